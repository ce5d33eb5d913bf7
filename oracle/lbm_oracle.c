/*
 * ORACLE -- test infrastructure only.  CPU restatement of the reference's lattice algorithms
 * (LBDEM-Taichi, src/lbm3d, src/lbdem, src/unlbdem, src/hybridlbdem) used by tests/ as the parity checker and
 * by bench.py as the timed CPU baseline.  The product path never links or calls this file.
 *
 * It keeps the reference's structure on purpose (that is what it checks against):
 *   - every lattice field is a dense [i][j][k] array, vectors stored AoS (19 / 3 contiguous);
 *   - collide writes feq and fpc, stream pulls fpc with per-neighbour CellType dispatch and
 *     refreshes rho, vel of FLUID cells (two passes per step, separate feq / fpc arrays);
 *   - Smagorinsky uses the previous step's feq and omega (lagged), as the reference does.
 * REAL is double (the FP64 oracle, suffix _f64) or float (suffix _f32, used for CPU timing).
 * Every function cites the reference file:line it restates.
 *
 * How each part is pinned (the reference cannot run here: every module imports Taichi):
 *   collide / stream / initialize: D3Q19 moment identities, mass conservation, Poiseuille and
 *     shear-wave closed forms, lagged-vs-fixed-point Smagorinsky steady state (tests/);
 *   psc_map: per-cell counts of a node-centred sphere equal integer lattice-point counts, the
 *     mean mapped volume over random offsets equals 4/3 pi (R - dx/2)^3; psc_forces: a spinning
 *     grain in a box of its own cubic symmetry feels a torque antiparallel to its spin and no
 *     force, a translating one a drag against its motion and no torque, zero force in fluid at
 *     rest (tests/test_oracle_psc_pins.py);
 *   threedelta, ibm_spread, ibm_drag, ibm_weight, vans_spread, vans_drag: closed-form kernel
 *     values, partition of unity, conservation of spread volume / momentum / force, dilute
 *     (Schiller-Naumann) drag limits (tests/test_oracle_unresolved.py);
 *   vans_collide: exact D3Q19 moments of the forcing term F_q and of the fluid-fraction
 *     equilibrium; tenneti_cd with vans_drag and the settling loop: the force balance and the
 *     terminal velocity of the reference's recorded one-particle runs (Re 31.9 and 4.1, solid
 *     fraction 0.054, within 0.5 %) (tests/test_oracle_vans_pins.py);
 *   boundaries: prescribed moments of the rebuilt Zou-He populations, exit equilibrium, channel
 *     profile (tests/test_oracle_boundaries.py);
 *   hybrid_map_coarse / _map_fine / _fine_weights / collide_hybrid / _fine_drag: equal to the PSC
 *     mapping of the coarse grains, conservation of the fine volume and momentum with the
 *     free-fraction correction, superposition of the single-solid weighted collisions and the
 *     momentum moved toward each solid, interpolation and dilute drag limit
 *     (tests/test_oracle_hybrid.py);
 *   stress / plane_stress: zero at rest, force balance of steady Poiseuille flow
 *     (tests/test_oracle_stress.py).
 * Pinned weakly: the phi^3 Re term of tenneti_cd adds 0.1 % to the drag at the recorded runs'
 * solid fraction and no other vector reaches it; the VANS effective viscosity of vans_collide is
 * pinned at zero solid fraction only.
 */
#include <math.h>
#include <stdint.h>
#include <string.h>

#ifndef REAL
#define REAL double
#define SFX(name) name##_f64
#endif

#define Q 19

/* CellType bit flags, src/lbm3d/lbmutils.py:69-97 */
#define CT_FLUID 1
#define CT_OBSTACLE 8
#define CT_FREE_SLIP 16
#define CT_VEL_LADD 32
#define CT_VEL_ZOUHE 64
#define CT_VEL_EXIT 128
#define CT_PRE_ZOUHE 256
#define CT_VEL_INLET_LADD 262144

/* D3Q19 set, src/lbm3d/lbm_solver3d.py:46-91 */
static const int CX[Q] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0};
static const int CY[Q] = {0, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 0, 0, 1, -1, 1, -1};
static const int CZ[Q] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1, 1, -1, -1, 1};
static const int QINV[Q] = {0, 2, 1, 4, 3, 6, 5, 8, 7, 10, 9, 12, 11, 14, 13, 16, 15, 18, 17};
static const REAL W[Q] = {1. / 3.,  1. / 18., 1. / 18., 1. / 18., 1. / 18., 1. / 18., 1. / 18.,
                          1. / 36., 1. / 36., 1. / 36., 1. / 36., 1. / 36., 1. / 36., 1. / 36.,
                          1. / 36., 1. / 36., 1. / 36., 1. / 36., 1. / 36.};

#define IDX(i, j, k) ((((long)(i)) * ny + (j)) * nz + (k))

/* compute_feq, src/lbm3d/lbm_solver3d.py:255-271 */
static void feq_of(REAL rho, const REAL *u, REAL *feq) {
  const REAL uv = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
  for (int q = 0; q < Q; ++q) {
    const REAL cu = CX[q] * u[0] + CY[q] * u[1] + CZ[q] * u[2];
    feq[q] = W[q] * rho * (1.0 + 3.0 * cu + 4.5 * cu * cu - 1.5 * uv);
  }
}

/* initialize, src/lbm3d/lbm_solver3d.py:157-171 (skip mask differs per class, passed in) */
void SFX(lbo_initialize)(int nx, int ny, int nz, const int32_t *CT, int32_t skip, const REAL *rho,
                         const REAL *vel, REAL *feq, REAL *f) {
  const long n = (long)nx * ny * nz;
#pragma omp parallel for schedule(static)
  for (long c = 0; c < n; ++c) {
    if (CT[c] & skip) continue;
    feq_of(rho[c], vel + 3 * c, feq + Q * c);
    for (int q = 0; q < Q; ++q) f[Q * c + q] = feq[Q * c + q];
  }
}

/* compute_strain_rate + computeOmega, src/lbm3d/lbm_solver3d.py:460-513 (uses the feq and
 * omega fields as they are, i.e. from the previous collide) */
static void smagorinsky(const REAL *f, const REAL *feq, REAL rho, REAL omega_old, REAL omega0, REAL cs,
                        REAL *omega_new, REAL *rate) {
  REAL xx = 0, yy = 0, zz = 0, xy = 0, xz = 0, yz = 0;
  for (int q = 0; q < Q; ++q) {
    const REAL d = f[q] - feq[q];
    xx += d * CX[q] * CX[q];
    yy += d * CY[q] * CY[q];
    zz += d * CZ[q] * CZ[q];
    xy += d * CX[q] * CY[q];
    xz += d * CX[q] * CZ[q];
    yz += d * CY[q] * CZ[q];
  }
  const REAL Pi = sqrt(2.0 * (xx * xx + yy * yy + zz * zz + 2.0 * (xy * xy + xz * xz + yz * yz)));
  *rate = omega_old / (2.0 * (1.0 / 3.0) * rho) * Pi;
  const REAL nuTurb = (cs * 1.0) * (cs * 1.0) * (*rate);
  const REAL tauEff = 1.0 / omega0 + nuTurb / (1.0 / 3.0);
  *omega_new = 1.0 / tauEff;
}

/* BasicLattice3D.collide, src/lbm3d/lbm_solver3d.py:176-192 (smag != 0), or
 * ForcedLattice3D.collide, src/lbm3d/lbm_foecelover.py:37-68 (smag == 0, F) */
void SFX(lbo_collide_basic)(int nx, int ny, int nz, const int32_t *CT, int32_t skip, int32_t smag, double omega0,
                            double cs, const REAL *F, const REAL *f, REAL *feq, const REAL *rho, const REAL *vel,
                            REAL *omega, REAL *rate, REAL *fpc) {
  const long n = (long)nx * ny * nz;
#pragma omp parallel for schedule(static)
  for (long c = 0; c < n; ++c) {
    if (CT[c] & skip) continue;
    if (smag) smagorinsky(f + Q * c, feq + Q * c, rho[c], omega[c], omega0, cs, &omega[c], &rate[c]);
    feq_of(rho[c], vel + 3 * c, feq + Q * c);
    const REAL om = omega[c];
    for (int q = 0; q < Q; ++q) {
      REAL v = (1.0 - om) * f[Q * c + q] + om * feq[Q * c + q];
      if (F) v += rho[c] * W[q] * (CX[q] * F[0] + CY[q] * F[1] + CZ[q] * F[2]) / (1.0 / 3.0);
      fpc[Q * c + q] = v;
    }
  }
}

/* Weighted (PSC-form) collision, PSCLattice3D.collide_solid (src/lbdem/psclattice3d.py:240-261)
 * and EqIMBlattice3D.collide_solid (src/unlbdem/eqlattice.py:272-292); BGK elsewhere.
 * smag: lagged Smagorinsky first (EqIBM, eqlattice.py:249).  hydroforce may be NULL. */
void SFX(lbo_collide_weighted)(int nx, int ny, int nz, const int32_t *CT, int32_t skip, int32_t smag,
                               double omega0, double cs, const REAL *F, const REAL *f, REAL *feq, const REAL *rho,
                               const REAL *vel, REAL *omega, REAL *rate, const REAL *volfrac,
                               const REAL *weight, const REAL *velsolid, REAL *hydroforce, REAL *fpc) {
  const long n = (long)nx * ny * nz;
  if (hydroforce) memset(hydroforce, 0, sizeof(REAL) * 3 * n);
#pragma omp parallel for schedule(static)
  for (long c = 0; c < n; ++c) {
    if (CT[c] & skip) continue;
    if (smag) smagorinsky(f + Q * c, feq + Q * c, rho[c], omega[c], omega0, cs, &omega[c], &rate[c]);
    feq_of(rho[c], vel + 3 * c, feq + Q * c);
    const REAL om = omega[c];
    const REAL *fc = f + Q * c, *fe = feq + Q * c;
    if (volfrac[c] > 0) {
      REAL fs[Q];
      feq_of(rho[c], velsolid + 3 * c, fs);
      const REAL B = weight[c];
      for (int q = 0; q < Q; ++q) {
        const REAL Os = fc[QINV[q]] - fe[QINV[q]] + fs[q] - fc[q];
        const REAL Of = -om * (fc[q] - fe[q]);
        REAL v = fc[q] + B * Os + (1.0 - B) * Of;
        if (F) v += (1.0 - B) * rho[c] * W[q] * (CX[q] * F[0] + CY[q] * F[1] + CZ[q] * F[2]) / (1.0 / 3.0);
        fpc[Q * c + q] = v;
        if (hydroforce) {
          hydroforce[3 * c] -= B * Os * CX[q];
          hydroforce[3 * c + 1] -= B * Os * CY[q];
          hydroforce[3 * c + 2] -= B * Os * CZ[q];
        }
      }
    } else {
      for (int q = 0; q < Q; ++q) {
        REAL v = (1.0 - om) * fc[q] + om * fe[q];
        if (F) v += rho[c] * W[q] * (CX[q] * F[0] + CY[q] * F[1] + CZ[q] * F[2]) / (1.0 / 3.0);
        fpc[Q * c + q] = v;
      }
    }
  }
}

#define CT_RIGHT 4096
#define CT_LEFT 8192
#define CT_TOP 16384
#define CT_BOTTOM 32768
#define CT_FRONT 65536
#define CT_BACK 131072

/* mirror maps qsyx / qsyy / qsyz, src/lbm3d/lbm_solver3d.py:50-57: the direction with the x / y /
 * z component of c reversed */
static const int QSY[3][Q] = {{0, 2, 1, 3, 4, 5, 6, 10, 9, 8, 7, 14, 13, 12, 11, 15, 16, 17, 18},
                              {0, 1, 2, 4, 3, 5, 6, 9, 10, 7, 8, 11, 12, 13, 14, 18, 17, 16, 15},
                              {0, 1, 2, 3, 4, 6, 5, 7, 8, 9, 10, 13, 14, 11, 12, 17, 18, 15, 16}};

/* compute_neigh_index, src/lbm3d/lbm_solver3d.py:308-342 */
static void neigh(int nx, int ny, int nz, int q, int i, int j, int k, int *a, int *b, int *d) {
  int in = i - CX[q], jn = j - CY[q], kn = k - CZ[q];
  if (in >= nx) in = 0; else if (in < 0) in = nx - 1;
  if (jn >= ny) jn = 0; else if (jn < 0) jn = ny - 1;
  if (kn >= nz) kn = 0; else if (kn < 0) kn = nz - 1;
  *a = in; *b = jn; *d = kn;
}

/* BasicLattice3D.stream task 1, src/lbm3d/lbm_solver3d.py:203-239, with compute_rho_vel
 * (:276-303).  vans != 0: Unresolvedlattice3D.stream (src/unlbdem/vanslattice.py:366-397):
 * rho, u refreshed at every streamed cell with compute_rhof_vel (:412-440). */
void SFX(lbo_stream)(int nx, int ny, int nz, const int32_t *CT, int32_t skip, const REAL *fpc, REAL *f,
                     REAL *rho, REAL *vel, int32_t vans, const REAL *volfrac, const REAL *hydroforce) {
#pragma omp parallel for schedule(static)
  for (int i = 0; i < nx; ++i)
    for (int j = 0; j < ny; ++j)
      for (int k = 0; k < nz; ++k) {
        const long c = IDX(i, j, k);
        if (CT[c] & skip) continue;
        f[Q * c] = fpc[Q * c];
        for (int q = 1; q < Q; ++q) {
          int a, b, d;
          neigh(nx, ny, nz, q, i, j, k, &a, &b, &d);
          const long m = IDX(a, b, d);
          const int32_t t = CT[m];
          if (t & (CT_FLUID | CT_VEL_ZOUHE | CT_VEL_EXIT | CT_PRE_ZOUHE)) {
            f[Q * c + q] = fpc[Q * m + q];
          } else if (t & CT_OBSTACLE) {
            f[Q * c + q] = fpc[Q * c + QINV[q]];
          } else if (t & CT_VEL_LADD) {
            const REAL cu = CX[q] * vel[3 * m] + CY[q] * vel[3 * m + 1] + CZ[q] * vel[3 * m + 2];
            f[Q * c + q] = fpc[Q * c + QINV[q]] + 2.0 * W[q] * rho[c] * cu / (1.0 / 3.0);
          } else if (t & CT_VEL_INLET_LADD) {  /* the inlet cell's density (:225-229) */
            const REAL cu = CX[q] * vel[3 * m] + CY[q] * vel[3 * m + 1] + CZ[q] * vel[3 * m + 2];
            f[Q * c + q] = fpc[Q * c + QINV[q]] + 2.0 * W[q] * rho[m] * cu / (1.0 / 3.0);
          } else if (t & CT_FREE_SLIP) {  /* specular reflection (:230-236) */
            /* the neighbour index with the wall-normal component kept at this cell's: a, b, d
             * with one of them replaced by i, j or k; fpc of a cell that never collides is 0 */
            if (t & (CT_LEFT | CT_RIGHT)) f[Q * c + q] = fpc[Q * IDX(i, b, d) + QSY[0][q]];
            else if (t & (CT_BOTTOM | CT_TOP)) f[Q * c + q] = fpc[Q * IDX(a, j, d) + QSY[1][q]];
            else if (t & (CT_BACK | CT_FRONT)) f[Q * c + q] = fpc[Q * IDX(a, b, k) + QSY[2][q]];
          }
          /* other neighbour types leave f[q] unchanged (lbm_solver3d.py:216-236) */
        }
        if (vans || (CT[c] & CT_FLUID)) {
          REAL r = 0, u[3] = {0, 0, 0};
          for (int q = 0; q < Q; ++q) {
            r += f[Q * c + q];
            u[0] += CX[q] * f[Q * c + q];
            u[1] += CY[q] * f[Q * c + q];
            u[2] += CZ[q] * f[Q * c + q];
          }
          rho[c] = r;
          for (int a = 0; a < 3; ++a) u[a] /= r;
          if (vans) {
            const REAL fvf = 1.0 - volfrac[c];
            for (int a = 0; a < 3; ++a) u[a] = (0.5 * hydroforce[3 * c + a] + u[a]) / (fvf * r);
          }
          vel[3 * c] = u[0];
          vel[3 * c + 1] = u[1];
          vel[3 * c + 2] = u[2];
        }
      }
}

/* ------------------------------------------------------------------------------------- */
/* PSC grain mapping, PSCLattice3D.grains2lattice (src/lbdem/psclattice3d.py:84-182).
 * grains: per grain [x, y, z, vx, vy, vz, wx, wy, wz, R] (physical units).  Outputs id (-1 =
 * none), volfrac, weight, velsolid (lattice units).  omega: per-cell omega field.          */
/* Wet-node boundaries: the second loop of BasicLattice3D.stream (src/lbm3d/lbm_solver3d.py:241-250),
 * after every cell has streamed: vel_zouHe (:671-820), bc_vel_exit (:645-667), pre_zouHe
 * (:823-957), bc_ladd_inlet (:516-527).  The face of a boundary cell is its first set bit among
 * LEFT, RIGHT, BOTTOM, TOP, BACK, FRONT (the reference's if-chain).  Zou-He for the face with
 * normal axis a: the populations with c_a = -s (s = +1 on the low face) are the known incoming
 * ones, those with c_a = s are rebuilt from their opposites. */

/* face index 0..5 = LEFT, RIGHT, BOTTOM, TOP, BACK, FRONT; -1 none */
static int bc_face(int32_t ct) {
  if (ct & CT_LEFT) return 0;
  if (ct & CT_RIGHT) return 1;
  if (ct & CT_BOTTOM) return 2;
  if (ct & CT_TOP) return 3;
  if (ct & CT_BACK) return 4;
  if (ct & CT_FRONT) return 5;
  return -1;
}

/* The terms of each face in the reference's order: the c_a = 0 populations, the c_a = -s
 * ones (doubled in the density sum), then per transverse axis t the c_t = +1 and c_t = -1
 * populations with c_a = 0 of its N_t term. */
static const int ZH_ZERO[3][9] = {{0, 3, 4, 5, 6, 15, 16, 17, 18},
                                  {0, 1, 2, 5, 6, 11, 12, 13, 14},
                                  {0, 1, 2, 3, 4, 7, 8, 9, 10}};
static const int ZH_IN[6][5] = {{2, 8, 10, 12, 14}, {1, 7, 9, 11, 13}, {4, 8, 9, 16, 18},
                                {3, 7, 10, 15, 17}, {6, 12, 13, 16, 17}, {5, 11, 14, 15, 18}};
/* [axis][transverse 0/1][plus/minus][3], transverse axes in (x, y, z) order skipping a */
static const int ZH_N[3][2][2][3] = {
    {{{3, 15, 17}, {4, 18, 16}}, {{5, 18, 15}, {6, 17, 16}}},
    {{{1, 11, 13}, {2, 14, 12}}, {{5, 11, 14}, {6, 13, 12}}},
    {{{1, 7, 9}, {2, 10, 8}}, {{3, 7, 10}, {4, 9, 8}}}};

static REAL zh_sum_known(const REAL *fc, int face) {
  const int a = face / 2;
  REAL s0 = 0, s1 = 0;
  for (int m = 0; m < 9; ++m) s0 += fc[ZH_ZERO[a][m]];
  for (int m = 0; m < 5; ++m) s1 += fc[ZH_IN[face][m]];
  return s0 + 2.0 * s1;
}

/* rebuilds the 5 unknown populations of face `face` for density rho and velocity u (u_t = 0 and
 * no rho u_t term in N_t for the pressure condition, `pressure` != 0) */
static void zh_rebuild(REAL *fc, int face, REAL rho, const REAL *u, int pressure) {
  const int a = face / 2, s = (face % 2 == 0) ? 1 : -1;
  const int ts[2] = {a == 0 ? 1 : 0, a == 2 ? 1 : 2};
  REAL N[2];
  for (int t = 0; t < 2; ++t) {
    const int *pp = ZH_N[a][t][0], *mm = ZH_N[a][t][1];
    N[t] = 0.5 * (fc[pp[0]] + fc[pp[1]] + fc[pp[2]] - (fc[mm[0]] + fc[mm[1]] + fc[mm[2]]));
    if (!pressure) N[t] = N[t] - 1.0 / 3.0 * rho * u[ts[t]];
  }
  for (int q = 1; q < Q; ++q) {
    const int c[3] = {CX[q], CY[q], CZ[q]};
    if (c[a] != s) continue;
    REAL cu = 0.0;
    int diag = -1;
    for (int b = 0; b < 3; ++b) {
      if (c[b] == 0) continue;
      if (pressure && b != a) { diag = b; continue; }
      cu += c[b] * u[b];
      if (b != a) diag = b;
    }
    if (diag < 0) {
      fc[q] = fc[QINV[q]] + (1.0 / 3.0) * rho * cu;
    } else {
      const int t = (diag == ts[0]) ? 0 : 1;
      fc[q] = fc[QINV[q]] + 1.0 / 6.0 * rho * cu - c[diag] * N[t];
    }
  }
}

void SFX(lbo_boundaries)(int nx, int ny, int nz, const int32_t *CT, REAL *f, REAL *rho, REAL *vel) {
  /* neighbour values are read from cells the loop does not change (FLUID interior cells) */
  for (int i = 0; i < nx; ++i)
    for (int j = 0; j < ny; ++j)
      for (int k = 0; k < nz; ++k) {
        const long c = IDX(i, j, k);
        const int32_t ct = CT[c];
        const int face = bc_face(ct);
        REAL *fc = f + Q * c;
        if (ct & CT_VEL_ZOUHE) {
          if (face < 0) continue;
          const int a = face / 2;
          const REAL ua = vel[3 * c + a];
          const REAL coeff = (face % 2 == 0) ? 1 / (1 - ua) : 1 / (1 + ua);
          rho[c] = coeff * zh_sum_known(fc, face);
          zh_rebuild(fc, face, rho[c], vel + 3 * c, 0);
        } else if (ct & CT_VEL_EXIT) {
          if (face >= 0) {
            static const int DI[6] = {1, -1, 0, 0, 0, 0}, DJ[6] = {0, 0, 1, -1, 0, 0}, DK[6] = {0, 0, 0, 0, 1, -1};
            const long m = IDX(i + DI[face], j + DJ[face], k + DK[face]);
            for (int b = 0; b < 3; ++b) vel[3 * c + b] = vel[3 * m + b];
          }
          rho[c] = 1.0;
          feq_of(rho[c], vel + 3 * c, fc);
        } else if (ct & CT_PRE_ZOUHE) {
          if (face < 0) continue;
          const int a = face / 2;
          const REAL r = rho[c];
          if (face % 2 == 0) vel[3 * c + a] = 1.0 - zh_sum_known(fc, face) / r;
          else vel[3 * c + a] = zh_sum_known(fc, face) / r - 1.0;
          REAL u[3] = {0, 0, 0};
          u[a] = vel[3 * c + a];
          zh_rebuild(fc, face, r, u, 1);
        } else if (ct & CT_VEL_INLET_LADD) {
          if (face < 0) continue;
          static const int DI[6] = {1, -1, 0, 0, 0, 0}, DJ[6] = {0, 0, 1, -1, 0, 0}, DK[6] = {0, 0, 0, 0, 1, -1};
          rho[c] = rho[IDX(i + DI[face], j + DJ[face], k + DK[face])];
        }
      }
}

void SFX(lbo_psc_map)(int nx, int ny, int nz, const int32_t *CT, int32_t ngrain, const double *grains,
                      const double *dmin, double dx, double dt, const REAL *omega, int32_t *id, REAL *volfrac,
                      REAL *weight, REAL *velsolid) {
  const int32_t skip = CT_OBSTACLE | CT_VEL_LADD | CT_FREE_SLIP;
  const long n = (long)nx * ny * nz;
#pragma omp parallel for schedule(dynamic, 64)
  for (long c = 0; c < n; ++c) {
    id[c] = -1;
    volfrac[c] = 0;
    weight[c] = 0;
    velsolid[3 * c] = velsolid[3 * c + 1] = velsolid[3 * c + 2] = 0;
    if (CT[c] & skip) continue;
    const int i = (int)(c / ((long)ny * nz)), j = (int)((c / nz) % ny), k = (int)(c % nz);
    for (int g = 0; g < ngrain; ++g) {
      const double *G = grains + 10 * g;
      const REAL xc = (REAL)((G[0] - dmin[0] + 0.5 * dx) / dx);
      const REAL yc = (REAL)((G[1] - dmin[1] + 0.5 * dx) / dx);
      const REAL zc = (REAL)((G[2] - dmin[2] + 0.5 * dx) / dx);
      const REAL r = (REAL)((G[9] - 0.5 * dx) / dx);
      const REAL dist = sqrt((xc - i) * (xc - i) + (yc - j) * (yc - j) + (zc - k) * (zc - k));
      int full = 0;
      REAL eps;
      if (dist >= r + 0.5 * sqrt(3.0)) continue;
      if (dist <= r - 0.5 * sqrt(3.0)) {
        full = 1;
        eps = 1.0;
      } else {
        int cnt = 0;
        for (int ii = 0; ii < 5; ++ii)
          for (int jj = 0; jj < 5; ++jj)
            for (int kk = 0; kk < 5; ++kk) {
              const REAL px = i - 0.4 + 0.2 * ii, py = j - 0.4 + 0.2 * jj, pz = k - 0.4 + 0.2 * kk;
              const REAL d2 = sqrt((xc - px) * (xc - px) + (yc - py) * (yc - py) + (zc - pz) * (zc - pz));
              if (d2 < r) ++cnt;
            }
        eps = cnt / 125.0;
        if (!(eps > volfrac[c])) continue;
      }
      id[c] = g;
      volfrac[c] = eps;
      if (full) weight[c] = 1.0;
      else {
        const REAL c1 = volfrac[c] * (1. / omega[c] - 0.5);
        const REAL c2 = (1. - volfrac[c]) + (1. / omega[c] - 0.5);
        weight[c] = c1 / c2;
      }
      /* compute_solid_vel, psclattice3d.py:160-182 */
      const REAL rx = (REAL)i - xc, ry = (REAL)j - yc, rz = (REAL)k - zc;
      const REAL wx = (REAL)G[6], wy = (REAL)G[7], wz = (REAL)G[8];
      const REAL ocx = wy * rz * dx - wz * ry * dx, ocy = wz * rx * dx - wx * rz * dx, ocz = wx * ry * dx - wy * rx * dx;
      velsolid[3 * c] = (G[3] + ocx) * dt / dx;
      velsolid[3 * c + 1] = (G[4] + ocy) * dt / dx;
      velsolid[3 * c + 2] = (G[5] + ocz) * dt / dx;
      if (full) break;
    }
  }
}

/* PSCLattice3D.lattice2grains, src/lbdem/psclattice3d.py:316-358 (serial per grain over its
 * bounding box); out: per grain [Fx, Fy, Fz, Tx, Ty, Tz] in physical units, FP64. */
void SFX(lbo_psc_forces)(int nx, int ny, int nz, const int32_t *CT, int32_t ngrain, const double *grains,
                         const double *dmin, double dx, double dt, double rho_phys, const int32_t *id,
                         const REAL *volfrac, const REAL *hydroforce, double *out) {
  const int32_t skip = CT_OBSTACLE | CT_VEL_LADD | CT_FREE_SLIP;
  const double scale = rho_phys * dx * dx * dx * dx / (dt * dt);
#pragma omp parallel for schedule(dynamic, 1)
  for (int g = 0; g < ngrain; ++g) {
    const double *G = grains + 10 * g;
    const double xc = (G[0] - dmin[0] + 0.5 * dx) / dx, yc = (G[1] - dmin[1] + 0.5 * dx) / dx,
                 zc = (G[2] - dmin[2] + 0.5 * dx) / dx, r = G[9] / dx;
    int xb = (int)(xc - r), xe = (int)(xc + r + 2), yb = (int)(yc - r), ye = (int)(yc + r + 2);
    int zb = (int)(zc - r), ze = (int)(zc + r + 2);
    if (xb < 0) xb = 0;
    if (yb < 0) yb = 0;
    if (zb < 0) zb = 0;
    if (xe > nx) xe = nx;
    if (ye > ny) ye = ny;
    if (ze > nz) ze = nz;
    double F[3] = {0, 0, 0}, T[3] = {0, 0, 0};
    for (int i = xb; i < xe; ++i)
      for (int j = yb; j < ye; ++j)
        for (int k = zb; k < ze; ++k) {
          const long c = IDX(i, j, k);
          if (CT[c] & skip) continue;
          if (volfrac[c] > 0.0 && id[c] == g) {
            const double Ff[3] = {hydroforce[3 * c] * scale, hydroforce[3 * c + 1] * scale,
                                  hydroforce[3 * c + 2] * scale};
            const double d[3] = {(xc - i) * dx, (yc - j) * dx, (zc - k) * dx};
            for (int a = 0; a < 3; ++a) F[a] += Ff[a];
            T[0] -= d[1] * Ff[2] - d[2] * Ff[1];
            T[1] -= d[2] * Ff[0] - d[0] * Ff[2];
            T[2] -= d[0] * Ff[1] - d[1] * Ff[0];
          }
        }
    for (int a = 0; a < 3; ++a) {
      out[6 * g + a] = F[a];
      out[6 * g + 3 + a] = T[a];
    }
  }
}

/* ------------------------------------------------------------------------------------- */
/* EqIMBlattice3D.threedelta, src/unlbdem/eqlattice.py:451-462 (also vanslattice.py:653-663) */
static double threedelta(double r) {
  double a = 0.0;
  if (r < 0.5) a = (1.0 + sqrt(-3.0 * r * r + 1.0)) / 3.0;
  else if (0.5 <= r && r <= 1.5) a = (5.0 - 3.0 * r - sqrt(-3.0 * (1.0 - r) * (1.0 - r) + 1.0)) / 6.0;
  return a;
}
double SFX(lbo_threedelta)(double r) { return threedelta(r); }

/* _kernel_with_mirror, src/unlbdem/eqlattice.py:468-528 */
static double kernel_mirror(int nx, int ny, int nz, double xc, double yc, double zc, int ii, int jj, int kk) {
  const double s = 1.5;
#define D3(a, b, c) sqrt(((a) - ii) * ((a) - ii) + ((b) - jj) * ((b) - jj) + ((c) - kk) * ((c) - kk))
  double w = threedelta(D3(xc, yc, zc));
  if (xc < s) w += threedelta(D3(-xc, yc, zc));
  if (xc > (double)(nx - 1) - s) w += threedelta(D3(2.0 * (nx - 1) - xc, yc, zc));
  if (yc < s) w += threedelta(D3(xc, -yc, zc));
  if (yc > (double)(ny - 1) - s) w += threedelta(D3(xc, 2.0 * (ny - 1) - yc, zc));
  if (zc < s) w += threedelta(D3(xc, yc, -zc));
  if (zc > (double)(nz - 1) - s) w += threedelta(D3(xc, yc, 2.0 * (nz - 1) - zc));
#undef D3
  return w;
}
double SFX(lbo_kernel_mirror)(int nx, int ny, int nz, double xc, double yc, double zc, int ii, int jj, int kk) {
  return kernel_mirror(nx, ny, nz, xc, yc, zc, ii, jj, kk);
}

/* EqIMBlattice3D.grains2lattice, src/unlbdem/eqlattice.py:104-194 (serial in grain order).
 * grains: [x, y, z, vx, vy, vz, R] per particle, physical.  Outputs volfrac (clamped),
 * velsolid, velsum, weight_sum (lattice units). */
void SFX(lbo_ibm_spread)(int nx, int ny, int nz, const int32_t *CT, int32_t np, const double *grains,
                         const double *dmin, double dx, double dt, REAL *volfrac, REAL *velsolid, REAL *velsum,
                         REAL *weight_sum) {
  const long n = (long)nx * ny * nz;
  const int32_t skip = CT_OBSTACLE | CT_VEL_LADD | CT_FREE_SLIP | CT_VEL_INLET_LADD;
  memset(volfrac, 0, sizeof(REAL) * n);
  memset(velsolid, 0, sizeof(REAL) * 3 * n);
  memset(velsum, 0, sizeof(REAL) * 3 * n);
  memset(weight_sum, 0, sizeof(REAL) * n);
  const double Vl = dx * dx * dx, s = 1.5, pi = 3.14159265358979323846;
  for (int p = 0; p < np; ++p) {
    const double *G = grains + 7 * p;
    const double xc = (G[0] - dmin[0] + 0.5 * dx) / dx, yc = (G[1] - dmin[1] + 0.5 * dx) / dx,
                 zc = (G[2] - dmin[2] + 0.5 * dx) / dx;
    const double Vp = 4.0 / 3.0 * pi * G[6] * G[6] * G[6];
    const double v[3] = {G[3] * dt / dx, G[4] * dt / dx, G[5] * dt / dx};
    int i0 = (int)(xc - s), i1 = (int)(xc + s + 1), j0 = (int)(yc - s), j1 = (int)(yc + s + 1);
    int k0 = (int)(zc - s), k1 = (int)(zc + s + 1);
    if (i0 < 0) i0 = 0;
    if (j0 < 0) j0 = 0;
    if (k0 < 0) k0 = 0;
    if (i1 > nx) i1 = nx;
    if (j1 > ny) j1 = ny;
    if (k1 > nz) k1 = nz;
    double denom = 0.0;
    for (int i = i0; i < i1; ++i)
      for (int j = j0; j < j1; ++j)
        for (int k = k0; k < k1; ++k) {
          if (CT[IDX(i, j, k)] & skip) continue;
          const double wb = kernel_mirror(nx, ny, nz, xc, yc, zc, i, j, k);
          if (wb < 0.0) continue;
          denom += wb;
        }
    if (denom < 1e-30) continue;
    for (int i = i0; i < i1; ++i)
      for (int j = j0; j < j1; ++j)
        for (int k = k0; k < k1; ++k) {
          const long c = IDX(i, j, k);
          if (CT[c] & skip) continue;
          const double wb = kernel_mirror(nx, ny, nz, xc, yc, zc, i, j, k);
          if (wb < 0.0) continue;
          const double wn = wb / denom, e = wn * Vp / Vl;
          volfrac[c] += e;
          for (int a = 0; a < 3; ++a) velsum[3 * c + a] += v[a] * e;
          weight_sum[c] += wn;
        }
  }
  for (long c = 0; c < n; ++c) {
    if (volfrac[c] >= 1.0) volfrac[c] = 0.99;
    if (volfrac[c] > 1e-15)
      for (int a = 0; a < 3; ++a) velsolid[3 * c + a] = velsum[3 * c + a] / volfrac[c];
  }
}

/* Tenneti drag coefficient, EqIMBlattice3D.weight_coefficient / compute_drag_force
 * (src/unlbdem/eqlattice.py:215-238, 392-444); lim: 1e-9 (EqIBM) or 1e-8 (VANS) */
static double tenneti_cd(double Re, double svf, double lim) {
  double Cd = 0.0;
  if (1.0 - svf > lim) {
    const double Cd0 = 1.0 + 0.15 * pow(Re, 0.687);
    const double A = 5.81 * svf / pow(1.0 - svf, 3) + 0.48 * pow(svf, 1.0 / 3.0) / pow(1.0 - svf, 4);
    const double s3 = svf * svf * svf;
    const double B = s3 * Re * (0.95 + 0.61 * s3 / pow(1.0 - svf, 2));
    Cd = (1.0 - svf) * (Cd0 / pow(1.0 - svf, 3) + A + B);
  }
  return Cd;
}
double SFX(lbo_tenneti_cd)(double Re, double svf, double lim) { return tenneti_cd(Re, svf, lim); }

/* EqIMBlattice3D.compute_weight, src/unlbdem/eqlattice.py:200-238 (physical units as written) */
void SFX(lbo_ibm_weight)(int nx, int ny, int nz, const int32_t *CT, double dx, double dt, double rho_phys,
                         double omega0, const REAL *volfrac, const REAL *velsolid, const REAL *vel, REAL *weight) {
  const long n = (long)nx * ny * nz;
  const double nuLu = (1.0 / omega0 - 0.5) / 3.0, nu = nuLu * dx * dx / dt, mu = rho_phys * nu;
  const double pi = 3.14159265358979323846;
#pragma omp parallel for schedule(static)
  for (long c = 0; c < n; ++c) {
    if (CT[c] & (CT_OBSTACLE | CT_VEL_LADD | CT_FREE_SLIP)) continue;
    if (!(volfrac[c] > 0)) continue;
    const double e = volfrac[c];
    const double R = pow(3.0 * dx * dx * dx * e / (4.0 * pi), 1.0 / 3.0), dp = 2.0 * R;
    double s2 = 0;
    for (int a = 0; a < 3; ++a) {
      const double d = (velsolid[3 * c + a] - vel[3 * c + a]) * dx / dt;
      s2 += d * d;
    }
    const double Re = (1.0 - e) * rho_phys * dp * sqrt(s2) / mu;
    const double Cd = tenneti_cd(Re, e, 1e-9);
    weight[c] = 3.0 * pi * (dp / dx) * nuLu * (1.0 - e) * Cd;
  }
}

/* EqIMBlattice3D.lattice2grains, src/unlbdem/eqlattice.py:309-386; out: force_fluid per
 * particle (physical units) */
void SFX(lbo_ibm_drag)(int nx, int ny, int nz, const int32_t *CT, int32_t np, const double *grains,
                       const double *dmin, double dx, double dt, double rho_phys, double omega0, const REAL *vel,
                       const REAL *volfrac, double *out) {
  const int32_t skip = CT_OBSTACLE | CT_VEL_LADD | CT_FREE_SLIP | CT_VEL_INLET_LADD;
  const double nuLu = (1.0 / omega0 - 0.5) / 3.0, nu = nuLu * dx * dx / dt, mu = rho_phys * nu;
  const double s = 1.5, pi = 3.14159265358979323846;
#pragma omp parallel for schedule(static)
  for (int p = 0; p < np; ++p) {
    const double *G = grains + 7 * p;
    const double xc = (G[0] - dmin[0] + 0.5 * dx) / dx, yc = (G[1] - dmin[1] + 0.5 * dx) / dx,
                 zc = (G[2] - dmin[2] + 0.5 * dx) / dx;
    int i0 = (int)(xc - s), i1 = (int)(xc + s + 1), j0 = (int)(yc - s), j1 = (int)(yc + s + 1);
    int k0 = (int)(zc - s), k1 = (int)(zc + s + 1);
    if (i0 < 0) i0 = 0;
    if (j0 < 0) j0 = 0;
    if (k0 < 0) k0 = 0;
    if (i1 > nx) i1 = nx;
    if (j1 > ny) j1 = ny;
    if (k1 > nz) k1 = nz;
    double vw[3] = {0, 0, 0}, es = 0, wt = 0;
    int nl = 0;
    for (int i = i0; i < i1; ++i)
      for (int j = j0; j < j1; ++j)
        for (int k = k0; k < k1; ++k) {
          const long c = IDX(i, j, k);
          if (CT[c] & skip) continue;
          const double w = kernel_mirror(nx, ny, nz, xc, yc, zc, i, j, k);
          if (w < 0.0) continue;
          for (int a = 0; a < 3; ++a) vw[a] += vel[3 * c + a] * w;
          es += 1.0 - volfrac[c];
          wt += w;
          nl += 1;
        }
    double uf[3] = {0, 0, 0}, ef = 0.0;
    if (wt > 1e-15)
      for (int a = 0; a < 3; ++a) uf[a] = vw[a] / wt * dx / dt;
    if (nl > 0) ef = es / (double)nl;
    double F[3] = {0, 0, 0};
    if (1.0 - ef > 1e-15) {
      const double ep = 1.0 - ef, dp = 2.0 * G[6];
      double us[3], m2 = 0;
      for (int a = 0; a < 3; ++a) {
        us[a] = G[3 + a] - uf[a];
        m2 += us[a] * us[a];
      }
      const double Re = (1.0 - ep) * rho_phys * dp * sqrt(m2) / mu;
      const double Cd = tenneti_cd(Re, ep, 1e-9);
      for (int a = 0; a < 3; ++a) F[a] = -3.0 * pi * dp * mu * (1.0 - ep) * Cd * us[a];
    }
    for (int a = 0; a < 3; ++a) out[3 * p + a] = F[a];
  }
}

/* ------------------------------------------------------------------------------------- */
/* VANS, src/unlbdem/vanslattice.py */

/* box [floor(x-1.5), ceil(x+1.5)] clamped to [0, N-1] (vanslattice.py:152-158, 511-516) */
static void vans_box(double c, int N, int *lo, int *hi) {
  int a = (int)floor(c - 1.5), b = (int)ceil(c + 1.5);
  if (a < 0) a = 0;
  if (b > N - 1) b = N - 1;
  *lo = a;
  *hi = b;
}

/* force2lattice (vanslattice.py:489-557), or grains2lattice (:135-196) when with_force == 0,
 * each grain's weights computed afresh (the reference's shared temp_weights buffer is a race
 * whose result depends on scheduling; this is its intended single-grain semantics).
 * grains: [x, y, z, R, Fx, Fy, Fz] per particle (physical). */
void SFX(lbo_vans_spread)(int nx, int ny, int nz, const int32_t *CT, int32_t np, const double *grains,
                          const double *dmin, double dx, double dt, double rho_phys, int32_t with_force,
                          REAL *volfrac, REAL *hydroforce) {
  const long n = (long)nx * ny * nz;
  const int32_t skip = CT_OBSTACLE | CT_VEL_LADD | CT_FREE_SLIP;
  memset(volfrac, 0, sizeof(REAL) * n);
  if (with_force) memset(hydroforce, 0, sizeof(REAL) * 3 * n);
  const double pi = 3.14159265358979323846, fs = dt * dt / (rho_phys * dx * dx * dx * dx);
  for (int p = 0; p < np; ++p) {
    const double *G = grains + 7 * p;
    const double xc = (G[0] - dmin[0] + 0.5 * dx) / dx, yc = (G[1] - dmin[1] + 0.5 * dx) / dx,
                 zc = (G[2] - dmin[2] + 0.5 * dx) / dx;
    int i0, i1, j0, j1, k0, k1;
    vans_box(xc, nx, &i0, &i1);
    vans_box(yc, ny, &j0, &j1);
    vans_box(zc, nz, &k0, &k1);
    double ws = 0.0;
    for (int i = i0; i <= i1; ++i)
      for (int j = j0; j <= j1; ++j)
        for (int k = k0; k <= k1; ++k) {
          if (CT[IDX(i, j, k)] & skip) continue;
          const double d = sqrt((xc - i) * (xc - i) + (yc - j) * (yc - j) + (zc - k) * (zc - k));
          if (d <= 1.5) ws += threedelta(d);
        }
    const double inv = ws > 1e-10 ? 1.0 / ws : 0.0;
    if (!with_force && !(ws > 1e-10)) continue; /* grains2lattice: correction 1, weights kept */
    const double Vg = 4.0 / 3.0 * pi * G[3] * G[3] * G[3], Vc = dx * dx * dx;
    for (int i = i0; i <= i1; ++i)
      for (int j = j0; j <= j1; ++j)
        for (int k = k0; k <= k1; ++k) {
          const long c = IDX(i, j, k);
          if (CT[c] & skip) continue;
          const double d = sqrt((xc - i) * (xc - i) + (yc - j) * (yc - j) + (zc - k) * (zc - k));
          if (!(d <= 1.5)) continue;
          const double w = threedelta(d);
          if (!(w > 1e-10) || !(inv > 0)) continue;
          const double wn = w * inv;
          volfrac[c] += wn * Vg / Vc;
          if (with_force)
            for (int a = 0; a < 3; ++a) hydroforce[3 * c + a] += wn * G[4 + a] * fs;
        }
  }
}

/* Unresolvedlattice3D.compute_hydrodynamic_force + compute_drag_force
 * (src/unlbdem/vanslattice.py:446-482, 563-651).  grains: [x, y, z, vx, vy, vz, R]. */
void SFX(lbo_vans_drag)(int nx, int ny, int nz, const int32_t *CT, int32_t np, const double *grains,
                        const double *dmin, double dx, double dt, double rho_phys, double omega0,
                        const REAL *volfrac, const REAL *vel, double *out) {
  const int32_t skip = CT_OBSTACLE | CT_VEL_LADD | CT_FREE_SLIP;
  const double nuLu = (1.0 / omega0 - 0.5) / 3.0, mu = rho_phys * nuLu * dx * dx / dt;
  const double pi = 3.14159265358979323846;
#pragma omp parallel for schedule(static)
  for (int p = 0; p < np; ++p) {
    const double *G = grains + 7 * p;
    const double xc = (G[0] - dmin[0] + 0.5 * dx) / dx, yc = (G[1] - dmin[1] + 0.5 * dx) / dx,
                 zc = (G[2] - dmin[2] + 0.5 * dx) / dx, r = G[6] / dx;
    int xb = (int)(xc - r - 2), xe = (int)(xc + r + 2), yb = (int)(yc - r - 2), ye = (int)(yc + r + 2);
    int zb = (int)(zc - r - 2), ze = (int)(zc + r + 2);
    if (xb < 0) xb = 0;
    if (yb < 0) yb = 0;
    if (zb < 0) zb = 0;
    if (xe > nx) xe = nx;
    if (ye > ny) ye = ny;
    if (ze > nz) ze = nz;
    double svf = 0, uf[3] = {0, 0, 0};
    for (int i = xb; i < xe; ++i)
      for (int j = yb; j < ye; ++j)
        for (int k = zb; k < ze; ++k) {
          const long c = IDX(i, j, k);
          if (CT[c] & skip) continue;
          if (volfrac[c] > 0.0) {
            const double d = sqrt((xc - i) * (xc - i) + (yc - j) * (yc - j) + (zc - k) * (zc - k));
            const double w = threedelta(d);
            svf += volfrac[c] * w;
            for (int a = 0; a < 3; ++a) uf[a] += vel[3 * c + a] * w;
          }
        }
    const double dp = 2.0 * G[6];
    double ur[3], m2 = 0;
    for (int a = 0; a < 3; ++a) {
      ur[a] = dx * uf[a] / dt - G[3 + a];
      m2 += ur[a] * ur[a];
    }
    const double Re = (1.0 - svf) * rho_phys * dp * sqrt(m2) / mu;
    const double Cd = tenneti_cd(Re, svf, 1e-8);
    for (int a = 0; a < 3; ++a) out[3 * p + a] = 3.0 * pi * dp * mu * (1.0 - svf) * Cd * ur[a];
  }
}

/* Unresolvedlattice3D.collide, src/unlbdem/vanslattice.py:203-358 (compute_omega,
 * compute_feq with fvf, compute_Fq); vel, rho are the stream-time fields. */
void SFX(lbo_vans_collide)(int nx, int ny, int nz, const int32_t *CT, double omega0, const REAL *f, REAL *feq,
                           const REAL *rho, const REAL *vel, REAL *omega, const REAL *volfrac,
                           const REAL *hydroforce, REAL *fpc) {
  const long n = (long)nx * ny * nz;
  const double nuLu = (1.0 / omega0 - 0.5) / 3.0;
#pragma omp parallel for schedule(static)
  for (long c = 0; c < n; ++c) {
    if (CT[c] & (CT_OBSTACLE | CT_VEL_LADD | CT_FREE_SLIP)) continue;
    const REAL svf = volfrac[c];
    const REAL g = 1.0 + 1.25 * svf * 0.64 / (1.0 - svf);
    omega[c] = 1.0 / (3.0 * nuLu * g * g + 0.5);
    const REAL *u = vel + 3 * c;
    const REAL fvf = 1.0 - svf, usq = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
    const REAL *Fh = hydroforce + 3 * c;
    for (int q = 0; q < Q; ++q) {
      const REAL cu = CX[q] * u[0] + CY[q] * u[1] + CZ[q] * u[2];
      feq[Q * c + q] = W[q] * rho[c] * (1.0 + 3.0 * cu * fvf + 4.5 * cu * cu * fvf - 1.5 * usq * fvf);
      const REAL t0 = 3.0 * CX[q] - 3.0 * u[0] + 9.0 * cu * CX[q];
      const REAL t1 = 3.0 * CY[q] - 3.0 * u[1] + 9.0 * cu * CY[q];
      const REAL t2 = 3.0 * CZ[q] - 3.0 * u[2] + 9.0 * cu * CZ[q];
      const REAL Fq = W[q] * (1.0 - 0.5 * omega[c]) * (t0 * Fh[0] + t1 * Fh[1] + t2 * Fh[2]);
      fpc[Q * c + q] = (1.0 - omega[c]) * f[Q * c + q] + omega[c] * feq[Q * c + q] + Fq;
    }
  }
}

/* ------------------------------------------------------------------------------------- */
/* HybridLattice3D, src/hybridlbdem/hybridlattice.py: coarse grains (2R > dx) as PSC, fine
 * grains (2R <= dx) as particle-equivalence IBM, one collision with both weights. */
#define HYB_SKIP (CT_OBSTACLE | CT_VEL_LADD | CT_FREE_SLIP | CT_VEL_INLET_LADD)

/* map_coarse_grains, hybridlattice.py:173-234 and _psc_set_velsolid :246-257 (cell loop over
 * the grains in index order).  grains: [x, y, z, vx, vy, vz, wx, wy, wz, R] per grain. */
void SFX(lbo_hybrid_map_coarse)(int nx, int ny, int nz, const int32_t *CT, int32_t ngrain, const double *grains,
                                const double *dmin, double dx, double dt, int32_t *id, REAL *volfrac,
                                REAL *velsolid) {
  const long n = (long)nx * ny * nz;
#pragma omp parallel for schedule(dynamic, 64)
  for (long c = 0; c < n; ++c) {
    id[c] = -1;
    volfrac[c] = 0;
    velsolid[3 * c] = velsolid[3 * c + 1] = velsolid[3 * c + 2] = 0;
    if (CT[c] & HYB_SKIP) continue;
    const int i = (int)(c / ((long)ny * nz)), j = (int)((c / nz) % ny), k = (int)(c % nz);
    for (int g = 0; g < ngrain; ++g) {
      const double *G = grains + 10 * g;
      if (2.0 * G[9] <= dx) continue;
      const double xc = (G[0] - dmin[0] + 0.5 * dx) / dx, yc = (G[1] - dmin[1] + 0.5 * dx) / dx,
                   zc = (G[2] - dmin[2] + 0.5 * dx) / dx;
      const double r = (G[9] - 0.5 * dx) / dx;
      const double dist = sqrt((xc - i) * (xc - i) + (yc - j) * (yc - j) + (zc - k) * (zc - k));
      const double hd = 0.5 * sqrt(3.0);
      int full = 0;
      double eps = 1.0;
      if (dist >= r + hd) continue;
      if (dist <= r - hd) {
        full = 1;
      } else {
        int cnt = 0;
        for (int si = 0; si < 5; ++si)
          for (int sj = 0; sj < 5; ++sj)
            for (int sk = 0; sk < 5; ++sk) {
              const double sx = i - 0.4 + 0.2 * si, sy = j - 0.4 + 0.2 * sj, sz = k - 0.4 + 0.2 * sk;
              if (sqrt((xc - sx) * (xc - sx) + (yc - sy) * (yc - sy) + (zc - sz) * (zc - sz)) < r) ++cnt;
            }
        eps = cnt / 125.0;
        if (!(eps > volfrac[c])) continue;
      }
      id[c] = g;
      volfrac[c] = (REAL)eps;
      const double rv[3] = {(i - xc) * dx, (j - yc) * dx, (k - zc) * dx};
      const double oc[3] = {G[7] * rv[2] - G[8] * rv[1], G[8] * rv[0] - G[6] * rv[2], G[6] * rv[1] - G[7] * rv[0]};
      for (int a = 0; a < 3; ++a) velsolid[3 * c + a] = (REAL)((G[3 + a] + oc[a]) * dt / dx);
      if (full) break;
    }
  }
}

/* map_fine_grains, hybridlattice.py:260-376 (serial in grain order): three passes per fine
 * grain (kernel sum, free-fraction correction, spread), then the clamp to the free fraction
 * and the solid velocity.  grains: [x, y, z, vx, vy, vz, R].  fine_velsolid keeps its value
 * where the fraction is at most 1e-10; *nwarn counts the overflow warnings. */
void SFX(lbo_hybrid_map_fine)(int nx, int ny, int nz, const int32_t *CT, int32_t np, const double *grains,
                              const double *dmin, double dx, double dt, const REAL *coarse_volfrac,
                              REAL *fine_volfrac, REAL *fine_velsolid, REAL *velsum, REAL *wsum, int32_t *nwarn) {
  const long n = (long)nx * ny * nz;
  memset(fine_volfrac, 0, sizeof(REAL) * n);
  memset(velsum, 0, sizeof(REAL) * 3 * n);
  memset(wsum, 0, sizeof(REAL) * n);
  const double Vl = dx * dx * dx, s = 1.5, pi = 3.14159265358979323846;
  for (int p = 0; p < np; ++p) {
    const double *G = grains + 7 * p;
    if (2.0 * G[6] > dx) continue;
    const double xc = (G[0] - dmin[0] + 0.5 * dx) / dx, yc = (G[1] - dmin[1] + 0.5 * dx) / dx,
                 zc = (G[2] - dmin[2] + 0.5 * dx) / dx;
    const double Vp = 4.0 / 3.0 * pi * G[6] * G[6] * G[6];
    const double v[3] = {G[3] * dt / dx, G[4] * dt / dx, G[5] * dt / dx};
    int i0 = (int)(xc - s), i1 = (int)(xc + s + 1), j0 = (int)(yc - s), j1 = (int)(yc + s + 1);
    int k0 = (int)(zc - s), k1 = (int)(zc + s + 1);
    if (i0 < 0) i0 = 0;
    if (j0 < 0) j0 = 0;
    if (k0 < 0) k0 = 0;
    if (i1 > nx) i1 = nx;
    if (j1 > ny) j1 = ny;
    if (k1 > nz) k1 = nz;
    double denom = 0.0, dcorr = 0.0;
    for (int pass = 0; pass < 3; ++pass) {
      for (int i = i0; i < i1; ++i)
        for (int j = j0; j < j1; ++j)
          for (int k = k0; k < k1; ++k) {
            const long c = IDX(i, j, k);
            if (CT[c] & HYB_SKIP) continue;
            if (coarse_volfrac[c] >= 1.0) continue;
            const double wb = kernel_mirror(nx, ny, nz, xc, yc, zc, i, j, k);
            if (!(wb > 0.0)) continue;
            if (pass == 0) {
              denom += wb;
            } else if (pass == 1) {
              dcorr += (wb / denom) * (1.0 - coarse_volfrac[c]);
            } else {
              const double wn = wb / denom, avail = 1.0 - coarse_volfrac[c];
              const double e = wn * avail / dcorr * Vp / Vl;
              fine_volfrac[c] += e;
              for (int a = 0; a < 3; ++a) velsum[3 * c + a] += v[a] * e;
              wsum[c] += wn;
            }
          }
      if (pass == 0 && denom < 1e-30) break;
      if (pass == 1 && dcorr < 1e-30) break;
    }
  }
  for (long c = 0; c < n; ++c) {
    const double avail = 1.0 - coarse_volfrac[c];
    if (fine_volfrac[c] > avail) {
      if (fine_volfrac[c] - avail > 1e-6) ++*nwarn;
      fine_volfrac[c] = (REAL)fmax(0.0, avail - 0.01);
    }
    if (fine_volfrac[c] > 1e-10)
      for (int a = 0; a < 3; ++a) fine_velsolid[3 * c + a] = velsum[3 * c + a] / fine_volfrac[c];
  }
}

/* compute_fine_weights + _weight_coefficient, hybridlattice.py:378-407 (physical units as
 * written); cells where the weight is not defined keep their value */
void SFX(lbo_hybrid_fine_weights)(int nx, int ny, int nz, const int32_t *CT, double dx, double dt,
                                  double rho_phys, double omega0, const REAL *coarse_volfrac,
                                  const REAL *fine_volfrac, const REAL *fine_velsolid, const REAL *vel,
                                  REAL *fine_weight) {
  const long n = (long)nx * ny * nz;
  const double nuLu = (1.0 / omega0 - 0.5) / 3.0, nu = nuLu * dx * dx / dt, mu = rho_phys * nu;
  const double pi = 3.14159265358979323846;
#pragma omp parallel for schedule(static)
  for (long c = 0; c < n; ++c) {
    if (CT[c] & HYB_SKIP) continue;
    const double svf = fine_volfrac[c], fvf = 1.0 - svf - coarse_volfrac[c];
    if (!(svf > 0.0 && fvf > 1e-10)) continue;
    const double Reff = pow(3.0 * dx * dx * dx * svf / (4.0 * pi), 1.0 / 3.0), dp = 2.0 * Reff;
    double s2 = 0;
    for (int a = 0; a < 3; ++a) {
      const double d = (fine_velsolid[3 * c + a] - vel[3 * c + a]) * dx / dt;
      s2 += d * d;
    }
    const double Re = (1.0 - svf) * rho_phys * dp * sqrt(s2) / mu;
    const double Cd = tenneti_cd(Re, svf, 1e-9);
    fine_weight[c] = (REAL)(3.0 * pi * (dp / dx) * nuLu * (1.0 - svf) * Cd);
  }
}

/* HybridLattice3D.collide, hybridlattice.py:433-593: computeOmega (lagged Smagorinsky of
 * BasicLattice3D), compute_feq, then the four-way dispatch collide_fluid / _coarse / _fine /
 * _hybrid; coarse_weight = B of _psc_compute_weight (:236-244) at coarse cells. */
void SFX(lbo_collide_hybrid)(int nx, int ny, int nz, const int32_t *CT, double omega0, double cs, const REAL *f,
                             REAL *feq, const REAL *rho, const REAL *vel, REAL *omega, REAL *rate,
                             const REAL *coarse_volfrac, const REAL *coarse_velsolid, REAL *coarse_weight,
                             const REAL *fine_volfrac, const REAL *fine_weight, const REAL *fine_velsolid,
                             REAL *hydroforce, REAL *fpc) {
  const long n = (long)nx * ny * nz;
  memset(hydroforce, 0, sizeof(REAL) * 3 * n);
#pragma omp parallel for schedule(static)
  for (long c = 0; c < n; ++c) {
    if (CT[c] & HYB_SKIP) continue;
    smagorinsky(f + Q * c, feq + Q * c, rho[c], omega[c], omega0, cs, &omega[c], &rate[c]);
    feq_of(rho[c], vel + 3 * c, feq + Q * c);
    const REAL om = omega[c];
    const REAL *fc = f + Q * c, *fe = feq + Q * c;
    const int has_c = coarse_volfrac[c] > 0.0, has_f = fine_volfrac[c] > 0.0;
    if (!has_c && !has_f) {
      for (int q = 0; q < Q; ++q) fpc[Q * c + q] = (1.0 - om) * fc[q] + om * fe[q];
      continue;
    }
    REAL B = 0, beta = 0, fs1[Q], fs2[Q];
    if (has_c) {
      const REAL eps = coarse_volfrac[c], tm = 1.0 / om - 0.5;
      B = (eps * tm) / ((1.0 - eps) + tm);
      coarse_weight[c] = B;
      feq_of(rho[c], coarse_velsolid + 3 * c, fs2);
    }
    if (has_f) {
      beta = fine_weight[c];
      feq_of(rho[c], fine_velsolid + 3 * c, fs1);
    }
    if (has_c && has_f && B + beta > 1.0) beta = 1.0 - B;
    for (int q = 0; q < Q; ++q) {
      const REAL Of = -om * (fc[q] - fe[q]);
      REAL v = fc[q];
      if (has_f) v += beta * (fc[QINV[q]] - fe[QINV[q]] + fs1[q] - fc[q]);
      if (has_c) {
        const REAL Os2 = fc[QINV[q]] - fe[QINV[q]] + fs2[q] - fc[q];
        v += B * Os2;
        hydroforce[3 * c] -= B * Os2 * CX[q];
        hydroforce[3 * c + 1] -= B * Os2 * CY[q];
        hydroforce[3 * c + 2] -= B * Os2 * CZ[q];
      }
      v += (1.0 - B - beta) * Of;
      fpc[Q * c + q] = v;
    }
  }
}

/* _transfer_fine_force + _compute_drag_force, hybridlattice.py:661-740; out: force_fluid of
 * the fine grains (rows of the coarse grains untouched).  grains: [x, y, z, vx, vy, vz, R]. */
void SFX(lbo_hybrid_fine_drag)(int nx, int ny, int nz, const int32_t *CT, int32_t np, const double *grains,
                               const double *dmin, double dx, double dt, double rho_phys, double omega0,
                               const REAL *vel, const REAL *fine_volfrac, const REAL *coarse_volfrac, double *out) {
  const double nuLu = (1.0 / omega0 - 0.5) / 3.0, nu = nuLu * dx * dx / dt, mu = rho_phys * nu;
  const double s = 1.5, pi = 3.14159265358979323846;
#pragma omp parallel for schedule(static)
  for (int p = 0; p < np; ++p) {
    const double *G = grains + 7 * p;
    if (2.0 * G[6] > dx) continue;
    const double xc = (G[0] - dmin[0] + 0.5 * dx) / dx, yc = (G[1] - dmin[1] + 0.5 * dx) / dx,
                 zc = (G[2] - dmin[2] + 0.5 * dx) / dx;
    int i0 = (int)(xc - s), i1 = (int)(xc + s + 1), j0 = (int)(yc - s), j1 = (int)(yc + s + 1);
    int k0 = (int)(zc - s), k1 = (int)(zc + s + 1);
    if (i0 < 0) i0 = 0;
    if (j0 < 0) j0 = 0;
    if (k0 < 0) k0 = 0;
    if (i1 > nx) i1 = nx;
    if (j1 > ny) j1 = ny;
    if (k1 > nz) k1 = nz;
    double vw[3] = {0, 0, 0}, es = 0, wt = 0;
    int nl = 0;
    for (int i = i0; i < i1; ++i)
      for (int j = j0; j < j1; ++j)
        for (int k = k0; k < k1; ++k) {
          const long c = IDX(i, j, k);
          if (CT[c] & HYB_SKIP) continue;
          const double w = kernel_mirror(nx, ny, nz, xc, yc, zc, i, j, k);
          if (w <= 0.0) continue;
          for (int a = 0; a < 3; ++a) vw[a] += vel[3 * c + a] * w;
          es += (1.0 - fine_volfrac[c] - coarse_volfrac[c]);
          wt += w;
          nl += 1;
        }
    double uf[3] = {0, 0, 0}, ef = 0.0;
    if (wt > 1e-15)
      for (int a = 0; a < 3; ++a) uf[a] = vw[a] / wt * dx / dt;
    if (nl > 0) ef = es / (double)nl;
    const double ep = 1.0 - ef;
    double F[3] = {0, 0, 0};
    if (ep > 1e-15) {
      const double dp = 2.0 * G[6];
      double us[3], m2 = 0;
      for (int a = 0; a < 3; ++a) {
        us[a] = G[3 + a] - uf[a];
        m2 += us[a] * us[a];
      }
      const double Re = (1.0 - ep) * rho_phys * dp * sqrt(m2) / mu;
      const double Cd = tenneti_cd(Re, ep, 1e-9);
      for (int a = 0; a < 3; ++a) F[a] = -3.0 * pi * dp * mu * (1.0 - ep) * Cd * us[a];
    }
    for (int a = 0; a < 3; ++a) out[3 * p + a] = F[a];
  }
}

/* ------------------------------------------------------------------------------------- */
/* BasicLattice3D.compute_stress, src/lbm3d/lbm_solver3d.py:347-405: f the post-stream
 * populations, feq and omega those the last collision left; out: 7 planes (xx, yy, zz, xy, xz,
 * yz, shear magnitude); skipped cells untouched. */
void SFX(lbo_stress)(int nx, int ny, int nz, const int32_t *CT, const REAL *f, const REAL *feq, const REAL *omega,
                     REAL *out) {
  const long n = (long)nx * ny * nz;
  const int32_t skip = CT_OBSTACLE | CT_VEL_LADD | CT_FREE_SLIP | CT_VEL_INLET_LADD;
#pragma omp parallel for schedule(static)
  for (long c = 0; c < n; ++c) {
    if (CT[c] & skip) continue;
    const REAL rf = 1.0 - 0.5 * omega[c];
    REAL s[6] = {0, 0, 0, 0, 0, 0};
    for (int q = 0; q < Q; ++q) {
      const REAL fneq = f[Q * c + q] - feq[Q * c + q];
      s[0] += CX[q] * CX[q] * fneq;
      s[1] += CY[q] * CY[q] * fneq;
      s[2] += CZ[q] * CZ[q] * fneq;
      s[3] += CX[q] * CY[q] * fneq;
      s[4] += CX[q] * CZ[q] * fneq;
      s[5] += CY[q] * CZ[q] * fneq;
    }
    for (int m = 0; m < 6; ++m) out[m * n + c] = -rf * s[m];
    out[6 * n + c] = sqrt(0.5 * (out[3 * n + c] * out[3 * n + c] + out[4 * n + c] * out[4 * n + c] +
                                 out[5 * n + c] * out[5 * n + c]));
  }
}

/* compute_plane_average_stress, src/lbm3d/lbm_solver3d.py:407-456: means of the xy, xz, yz
 * components over the cells of the plane whose CellType has the FLUID bit */
void SFX(lbo_plane_stress)(int nx, int ny, int nz, const int32_t *CT, const REAL *st, int axis, int pos,
                           double *out) {
  const long n = (long)nx * ny * nz;
  double s[3] = {0, 0, 0};
  long cnt = 0;
  const int n1 = axis == 0 ? ny : nx, n2 = axis == 2 ? ny : nz;
  for (int a = 0; a < n1; ++a)
    for (int b = 0; b < n2; ++b) {
      const int i = axis == 0 ? pos : a, j = axis == 0 ? a : (axis == 1 ? pos : b), k = axis == 2 ? pos : b;
      const long c = IDX(i, j, k);
      if (!(CT[c] & CT_FLUID)) continue;
      for (int m = 0; m < 3; ++m) s[m] += st[(3 + m) * n + c];
      ++cnt;
    }
  for (int m = 0; m < 3; ++m) out[m] = cnt > 0 ? s[m] / (double)cnt : 0.0;
}
