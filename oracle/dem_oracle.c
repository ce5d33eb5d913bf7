/*
 * ORACLE -- test infrastructure only.  Serial CPU restatement of the reference's DEM
 * (LBDEM-Taichi src/dem3d): Morton-hash broad phase with counting sort and 8-octant probe,
 * per-particle contact-history slots, Hertz-Mindlin and linear spring-dashpot contacts,
 * planar walls, and the integrator.  Used by tests/ as the parity checker and by bench.py as
 * the timed CPU baseline.  The product never links it.
 *
 * Deviation kept on purpose: candidate pairs go to a compact (i, j) list instead of the
 * reference's n x n bit table (demsolver.py:298-301), which overflows int32 indexing above
 * n = 46,340; the set of pairs is the same.  Slot overflow is counted instead of writing
 * cf[-1] (demsolver.py:666-680).
 * Every function cites the reference file:line it restates.  Always FP64 (REAL unused).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>



/* ---------------------------------------------------------------------------------- */
/* broad phase, src/dem3d/bpcd/bpcd.py                                                  */
/* ---------------------------------------------------------------------------------- */

/* morton3d32, bpcd.py:154-179 */
static uint32_t spread10(int32_t v) {
  uint32_t x = (uint32_t)v & 0x3ffu;
  x = (x | x << 16) & 0x30000ffu;
  x = (x | x << 8) & 0x300f00fu;
  x = (x | x << 4) & 0x30c30c3u;
  x = (x | x << 2) & 0x9249249u;
  return x;
}
static uint32_t morton(int32_t x, int32_t y, int32_t z) { return spread10(x) | (spread10(y) << 1) | (spread10(z) << 2); }

/* hash_code, bpcd.py:186-189: morton % T, on the int32 Morton value (non-negative: < 2^30) */
static int32_t hash_code(int32_t x, int32_t y, int32_t z, int32_t T) { return (int32_t)(morton(x, y, z) % (uint32_t)T); }

/* cell, bpcd.py:191-194: floor((x - min) / cs) */
static void cell_of(const double *x, const double *dmin, double cs, int32_t *ijk) {
  for (int a = 0; a < 3; ++a) ijk[a] = (int32_t)floor((x[a] - dmin[a]) / cs);
}

uint32_t lbo_morton3d32(int32_t x, int32_t y, int32_t z) { return morton(x, y, z); }

static int cmp_i32(const void *a, const void *b) {
  const int32_t x = *(const int32_t *)a, y = *(const int32_t *)b;
  return (x > y) - (x < y);
}

/* BPCD.create table size, bpcd.py:48-65, with next_pow2 (bpcd/utils.py:12-19) */
static int64_t next_pow2(int64_t x) {
  x -= 1;
  x |= x >> 1;
  x |= x >> 2;
  x |= x >> 4;
  x |= x >> 8;
  x |= x >> 16;
  return x + 1;
}
int64_t lbo_hash_table_size(const double *dmin, const double *dmax, double cell) {
  const double v0 = (dmax[0] - dmin[0]) / cell, v1 = (dmax[1] - dmin[1]) / cell, v2 = (dmax[2] - dmin[2]) / cell;
  int64_t size = (int64_t)(v0 * v1 * v2);
  size = next_pow2(size);
  size = next_pow2(size);
  if (size < (1 << 20)) size = 1 << 20;
  if (size > (1 << 22)) size = 1 << 22;
  return size;
}

/* detect_collision (bpcd.py:67-78): count (:80-84, :105-112), exclusive scan
 * (prefixsum.py:62-84), scatter (:86-95), probe (:114-152).  probe27 != 0 probes the 27
 * surrounding cells instead of the 8-cell octant.  Pairs (i < j) are appended to pairs
 * (capacity cap pairs); returns the number of distinct pairs (may exceed cap). */
int64_t lbo_broadphase(int32_t n, const double *pos, const double *dmin, double cs, int32_t T, int32_t probe27,
                       int32_t *pairs, int64_t cap) {
  /* the three table arrays persist between calls (a table of 2^20..2^22 entries is zeroed
   * once; afterwards only the buckets this call touched are cleared again).  Not reentrant. */
  static int32_t *count = NULL, *offset = NULL, *current = NULL;
  static int32_t cap_T = 0;
  if (T > cap_T) {
    free(count);
    free(offset);
    free(current);
    count = calloc((size_t)T, sizeof(int32_t));
    offset = calloc((size_t)T, sizeof(int32_t));
    current = calloc((size_t)T, sizeof(int32_t));
    cap_T = T;
  }
  int32_t *pid = malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
  int32_t *hsh = malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
  for (int32_t i = 0; i < n; ++i) {
    int32_t ijk[3];
    cell_of(pos + 3 * i, dmin, cs, ijk);
    hsh[i] = hash_code(ijk[0], ijk[1], ijk[2], T);
    count[hsh[i]] += 1;
  }
  /* exclusive scan (the reference uses a Blelloch scan; any exact scan gives the same).  Only
   * the occupied buckets are visited, in increasing hash order: their offsets equal those of
   * the dense scan, and an empty bucket's offset is never read (its count is 0). */
  int32_t *occ = malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
  memcpy(occ, hsh, sizeof(int32_t) * (size_t)n);
  qsort(occ, (size_t)n, sizeof(int32_t), cmp_i32);
  int32_t run = 0;
  for (int32_t u = 0; u < n; ++u) {
    if (u > 0 && occ[u] == occ[u - 1]) continue;
    offset[occ[u]] = run;
    run += count[occ[u]];
  }
  free(occ);
  for (int32_t i = 0; i < n; ++i) pid[offset[hsh[i]] + current[hsh[i]]++] = i;
  int64_t np = 0;
  for (int32_t i = 0; i < n; ++i) {
    const double *o = pos + 3 * i;
    int32_t ijk[3];
    cell_of(o, dmin, cs, ijk);
    int32_t cells[27][3];
    int nc = 0;
    if (!probe27) {
      int32_t d[3];
      for (int a = 0; a < 3; ++a) {
        const double center = ((double)ijk[a] + 0.5) * cs + dmin[a]; /* cell_center, bpcd.py:196-202 */
        d[a] = (o[a] > center) ? 1 : -1;
      }
      const int32_t off[8][3] = {{0, 0, 0}, {d[0], 0, 0}, {0, d[1], 0}, {0, 0, d[2]},
                                 {0, d[1], d[2]}, {d[0], 0, d[2]}, {d[0], d[1], 0}, {d[0], d[1], d[2]}};
      for (int c = 0; c < 8; ++c) {
        for (int a = 0; a < 3; ++a) cells[c][a] = ijk[a] + off[c][a];
      }
      nc = 8;
    } else {
      for (int dx = -1; dx <= 1; ++dx)
        for (int dy = -1; dy <= 1; ++dy)
          for (int dz = -1; dz <= 1; ++dz) {
            cells[nc][0] = ijk[0] + dx;
            cells[nc][1] = ijk[1] + dy;
            cells[nc][2] = ijk[2] + dz;
            ++nc;
          }
    }
    int32_t seen[27];
    for (int c = 0; c < nc; ++c) {
      const int32_t h = hash_code(cells[c][0], cells[c][1], cells[c][2], T);
      int dup = 0;
      for (int e = 0; e < c; ++e) dup |= (seen[e] == h);
      seen[c] = h;
      if (dup) continue; /* the bit table makes a revisited bucket idempotent */
      for (int32_t idx = offset[h]; idx < offset[h] + count[h]; ++idx) {
        const int32_t p = pid[idx];
        if (p > i) {
          if (np < cap) {
            pairs[2 * np] = i;
            pairs[2 * np + 1] = p;
          }
          ++np;
        }
      }
    }
  }
  for (int32_t i = 0; i < n; ++i) count[hsh[i]] = current[hsh[i]] = offset[hsh[i]] = 0;
  free(pid);
  free(hsh);
  return np;
}

/* ---------------------------------------------------------------------------------- */
/* DEM state and step                                                                   */
/* ---------------------------------------------------------------------------------- */
typedef struct {
  int32_t n, K;              /* particles, max_coordinate_number */
  int32_t model;             /* 0 = Hertz-Mindlin, 1 = linear */
  double dt, g[3];
  double dmin[3], dmax[3];   /* config.domain bounds */
  int32_t periodic[3];
  double cell;               /* hash cell size */
  int32_t T, probe27;
  /* materials [0] particle, [1] wall: density, E, nu, mu, e, kn, kt, dpn, dps */
  double mat[2][9];
  double wall_n[6][3], wall_d[6];
  int32_t wall_type[6];      /* 0 wall, 1 periodic */
} dem_params;

typedef struct {
  double *mass, *radius, *pos, *vel, *acc, *force, *quat, *omega, *omega_dot, *inertia, *moment;
  double *force_fluid, *moment_fluid;
  int32_t *freeze, *fixvel;
  /* contact slots, n*K */
  int32_t *cf_j, *cf_active, *cfn;
  double *cf_pos, *cf_force, *cf_shear;
  /* wall contacts, n*6 */
  int32_t *wcf_active;
  double *wcf_pos, *wcf_force, *wcf_shear;
  int32_t *status; /* [0] slot overflow, [1] dt warnings, [2] pp overlap warnings, [3] pw overlap warnings */
} dem_state;

enum { MAT_RHO, MAT_E, MAT_NU, MAT_MU, MAT_REST, MAT_KN, MAT_KT, MAT_DPN, MAT_DPS };

static void cross3(const double *a, const double *b, double *c) {
  c[0] = a[1] * b[2] - a[2] * b[1];
  c[1] = a[2] * b[0] - a[0] * b[2];
  c[2] = a[0] * b[1] - a[1] * b[0];
}
static double norm3(const double *a) { return sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]); }
static void matvec(const double *M, const double *v, double *o) {
  for (int r = 0; r < 3; ++r) o[r] = M[3 * r] * v[0] + M[3 * r + 1] * v[1] + M[3 * r + 2] * v[2];
}
/* 3x3 inverse by the adjugate (what ti.Matrix.inverse evaluates for n = 3) */
static void inv3(const double *m, double *o) {
  const double det = m[0] * (m[4] * m[8] - m[5] * m[7]) - m[1] * (m[3] * m[8] - m[5] * m[6]) +
                     m[2] * (m[3] * m[7] - m[4] * m[6]);
  const double id = 1.0 / det;
  o[0] = (m[4] * m[8] - m[5] * m[7]) * id;
  o[1] = (m[2] * m[7] - m[1] * m[8]) * id;
  o[2] = (m[1] * m[5] - m[2] * m[4]) * id;
  o[3] = (m[5] * m[6] - m[3] * m[8]) * id;
  o[4] = (m[0] * m[8] - m[2] * m[6]) * id;
  o[5] = (m[2] * m[3] - m[0] * m[5]) * id;
  o[6] = (m[3] * m[7] - m[4] * m[6]) * id;
  o[7] = (m[1] * m[6] - m[0] * m[7]) * id;
  o[8] = (m[0] * m[4] - m[1] * m[3]) * id;
}

/* rotation taking unit vector a to +x (demsolver.py:686-701, 757-773) */
static void rot_to_x(const double *a, double *R) {
  const double b[3] = {1.0, 0.0, 0.0};
  double v[3];
  cross3(a, b, v);
  const double s = norm3(v), c = a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
  if (s < 1e-12) {
    for (int k = 0; k < 9; ++k) R[k] = 0.0;
    if (c > 0.0) R[0] = R[4] = R[8] = 1.0;
    else {
      R[0] = -1.0;
      R[4] = 1.0;
      R[8] = -1.0;
    }
    return;
  }
  const double vx[9] = {0.0, -v[2], v[1], v[2], 0.0, -v[0], -v[1], v[0], 0.0};
  double vv[9];
  for (int r = 0; r < 3; ++r)
    for (int q = 0; q < 3; ++q) vv[3 * r + q] = vx[3 * r] * vx[q] + vx[3 * r + 1] * vx[3 + q] + vx[3 * r + 2] * vx[6 + q];
  const double k = (1.0 - c) / (s * s);
  for (int e = 0; e < 9; ++e) R[e] = ((e % 4 == 0) ? 1.0 : 0.0) + vx[e] + k * vv[e];
}

/* get_periodic_distance_vector, demsolver.py:483-512 */
static void periodic_delta(const dem_params *P, const double *xi, const double *xj, double *d) {
  for (int a = 0; a < 3; ++a) {
    d[a] = xj[a] - xi[a];
    if (P->periodic[a]) {
      const double L = P->dmax[a] - P->dmin[a];
      if (d[a] > 0.5 * L) d[a] -= L;
      else if (d[a] < -0.5 * L) d[a] += L;
    }
  }
}

/* contact law in the local frame (hertz.py:45-97, linear.py:43-101; walls :99-153, :103-152).
 * wall: particle-wall variant.  shear: history, updated in place.  Returns F (local). */
static void contact_law(const dem_params *P, dem_state *S, int i, int j, int wall, double delta_n, const double *vc,
                        double *shear, double *F) {
  const double *mi = P->mat[0], *mj = P->mat[wall ? 1 : 0];
  const double dt = P->dt;
  const double gap = -1.0 * delta_n;
  double kn, kt, gn, gt, mu;
  if (P->model == 0) {
    const double Ys = 1.0 / ((1.0 - mi[MAT_NU] * mi[MAT_NU]) / mi[MAT_E] + (1.0 - mj[MAT_NU] * mj[MAT_NU]) / mj[MAT_E]);
    const double Gs = 1.0 / (2.0 * (2.0 - mi[MAT_NU]) * (1.0 + mi[MAT_NU]) / mi[MAT_E] +
                             2.0 * (2.0 - mj[MAT_NU]) * (1.0 + mj[MAT_NU]) / mj[MAT_E]);
    double Rs, ms, e;
    if (wall) {
      Rs = S->radius[i];
      ms = S->mass[i];
      e = mj[MAT_REST];
      mu = mj[MAT_MU];
    } else {
      Rs = 1.0 / (1.0 / S->radius[i] + 1.0 / S->radius[j]);
      /* min_time, hertz.py:156-162: rho from i's mass and radius */
      const double rho = S->mass[i] * 3 / (4 * M_PI * S->radius[i] * S->radius[i] * S->radius[i]);
      const double tmin = M_PI * Rs * sqrt(rho / Gs) / (0.1631 * mi[MAT_NU] + 0.8766);
      if (dt > tmin) S->status[1] += 1;
      ms = 1.0 / (1.0 / S->mass[i] + 1.0 / S->mass[j]);
      e = 0.5 * (mi[MAT_REST] + mj[MAT_REST]);
      mu = 0.5 * (mi[MAT_MU] + mj[MAT_MU]);
    }
    const double le = log(e);
    const double beta = le / sqrt(le * le + M_PI * M_PI);
    const double Sn = 2.0 * Ys * sqrt(Rs * delta_n);
    const double St = 8.0 * Gs * sqrt(Rs * delta_n);
    kn = 4.0 / 3.0 * Ys * sqrt(Rs * delta_n);
    gn = -2.0 * beta * sqrt(5.0 / 6.0 * Sn * ms);
    kt = 8.0 * Gs * sqrt(Rs * delta_n);
    gt = -2.0 * beta * sqrt(5.0 / 6.0 * St * ms);
  } else {
    double ms;
    double dpn, dps;
    if (wall) {
      kn = mi[MAT_KN];
      kt = mi[MAT_KT];
      dpn = mi[MAT_DPN];
      dps = mi[MAT_DPS];
      ms = S->mass[i];
      mu = mj[MAT_MU];
    } else {
      kn = mi[MAT_KN] * mj[MAT_KN] / (mi[MAT_KN] + mj[MAT_KN]);
      kt = mi[MAT_KT] * mj[MAT_KT] / (mi[MAT_KT] + mj[MAT_KT]);
      dpn = 0.5 * (mi[MAT_DPN] + mj[MAT_DPN]);
      dps = 0.5 * (mi[MAT_DPS] + mj[MAT_DPS]);
      if (dt > sqrt(S->mass[i] / mi[MAT_KN])) S->status[1] += 1; /* linear.py:155-160 */
      ms = 1.0 / (1.0 / S->mass[i] + 1.0 / S->mass[j]);
      mu = 0.5 * (mi[MAT_MU] + mj[MAT_MU]);
    }
    gn = 2. * dpn * sqrt(kn * ms);
    gt = 2. * dps * sqrt(kn * ms);
  }
  shear[1] += vc[1] * dt;
  shear[2] += vc[2] * dt;
  F[0] = -kn * gap - gn * vc[0];
  const double tr[3] = {-kt * shear[0], -kt * shear[1], -kt * shear[2]};
  const double tl = norm3(tr);
  if (tl >= mu * F[0] && tl > 0.0) {
    const double ratio = mu * F[0] / tl;
    F[1] = tr[1] * ratio;
    F[2] = tr[2] * ratio;
    shear[1] = -F[1] / kt;
    shear[2] = -F[2] / kt;
  } else {
    F[1] = tr[1] - gt * vc[1];
    F[2] = tr[2] - gt * vc[2];
  }
}

/* resolve, demsolver.py:638-740 (history lookup :361-369, append :352-359) */
static void resolve(const dem_params *P, dem_state *S, int i, int j) {
  const int K = P->K;
  int off = -1;
  const int ncf = S->cfn[i] < K ? S->cfn[i] : K;
  for (int s = 0; s < ncf; ++s)
    if (S->cf_j[i * K + s] == j && S->cf_active[i * K + s]) {
      off = i * K + s;
      break;
    }
  double d[3];
  periodic_delta(P, S->pos + 3 * i, S->pos + 3 * j, d);
  const double dist = norm3(d);
  const double gap = dist - S->radius[i] - S->radius[j];
  int eval = 0;
  if (off >= 0) {
    if (gap < 0) eval = 1;
    else S->cf_active[off] = 0;
  } else if (gap < 0) {
    const int slot = S->cfn[i];
    if (slot >= K) {
      S->status[0] += 1;
      return;
    }
    S->cfn[i] = slot + 1;
    off = i * K + slot;
    S->cf_j[off] = j;
    S->cf_active[off] = 1;
    for (int a = 0; a < 3; ++a) S->cf_force[3 * off + a] = S->cf_shear[3 * off + a] = 0.0;
    eval = 1;
  }
  if (!eval) return;
  double a[3] = {d[0] / dist, d[1] / dist, d[2] / dist}, R[9], Ri[9];
  rot_to_x(a, R);
  const double delta_n = -gap;
  const double Rmin = S->radius[i] < S->radius[j] ? S->radius[i] : S->radius[j];
  if (delta_n > 0.05 * Rmin) S->status[2] += 1;
  double *cp = S->cf_pos + 3 * off;
  for (int k = 0; k < 3; ++k) cp[k] = S->pos[3 * i + k] + a[k] * (S->radius[i] - delta_n);
  double ri[3], rj[3], vci[3], vcj[3], dv[3], vc[3], F[3];
  for (int k = 0; k < 3; ++k) {
    ri[k] = cp[k] - S->pos[3 * i + k];
    rj[k] = cp[k] - (S->pos[3 * i + k] + d[k]);
  }
  cross3(S->omega + 3 * i, ri, vci);
  cross3(S->omega + 3 * j, rj, vcj);
  for (int k = 0; k < 3; ++k) dv[k] = (vcj[k] + S->vel[3 * j + k]) - (vci[k] + S->vel[3 * i + k]);
  matvec(R, dv, vc);
  contact_law(P, S, i, j, 0, delta_n, vc, S->cf_shear + 3 * off, F);
  for (int k = 0; k < 3; ++k) S->cf_force[3 * off + k] = F[k];
  inv3(R, Ri);
  const double mF[3] = {-F[0], -F[1], -F[2]};
  double Fi[3], Fj[3], Mi[3], Mj[3];
  matvec(Ri, mF, Fi);
  matvec(Ri, F, Fj);
  cross3(ri, Fi, Mi);
  cross3(rj, Fj, Mj);
  for (int k = 0; k < 3; ++k) {
    S->force[3 * i + k] += Fi[k];
    S->force[3 * j + k] += Fj[k];
    S->moment[3 * i + k] += Mi[k];
    S->moment[3 * j + k] += Mj[k];
  }
}

/* resolve_wall + evaluate_wall, demsolver.py:742-838 */
static void resolve_walls(const dem_params *P, dem_state *S) {
  for (int i = 0; i < P->n; ++i)
    for (int w = 0; w < 6; ++w) {
      if (P->wall_type[w] == 1) continue;
      const double *nrm = P->wall_n[w];
      const double *x = S->pos + 3 * i;
      const double pd = x[0] * nrm[0] + x[1] * nrm[1] + x[2] * nrm[2] - P->wall_d[w];
      const int c = 6 * i + w;
      if (S->wcf_active[c]) {
        if (-pd > S->radius[i]) {
          S->wcf_active[c] = 0;
          continue;
        }
      } else {
        if (!(-pd <= S->radius[i])) continue;
        S->wcf_active[c] = 1;
        for (int k = 0; k < 3; ++k) S->wcf_shear[3 * c + k] = 0.0;
      }
      double R[9], Ri[9];
      rot_to_x(nrm, R);
      const double delta_n = pd + S->radius[i];
      const double distance = S->radius[i] - delta_n;
      if (delta_n > 0.05 * S->radius[i]) S->status[3] += 1;
      double ri[3], vci[3], mv[3], vc[3], F[3];
      for (int k = 0; k < 3; ++k) {
        ri[k] = distance * nrm[k];
        S->wcf_pos[3 * c + k] = x[k] + ri[k];
      }
      cross3(S->omega + 3 * i, ri, vci);
      for (int k = 0; k < 3; ++k) mv[k] = -(vci[k] + S->vel[3 * i + k]);
      matvec(R, mv, vc);
      contact_law(P, S, i, w, 1, delta_n, vc, S->wcf_shear + 3 * c, F);
      for (int k = 0; k < 3; ++k) S->wcf_force[3 * c + k] = F[k];
      inv3(R, Ri);
      const double mF[3] = {-F[0], -F[1], -F[2]};
      double Fi[3], Mi[3];
      matvec(Ri, mF, Fi);
      cross3(ri, Fi, Mi);
      for (int k = 0; k < 3; ++k) {
        S->force[3 * i + k] += Fi[k];
        S->moment[3 * i + k] += Mi[k];
      }
    }
}

/* quat2RotMatrix, demsolver/utils.py:34-61 */
static void quat_rot(const double *q, double *R) {
  const double w2 = q[0] * q[0], i2 = q[1] * q[1], j2 = q[2] * q[2], k2 = q[3] * q[3];
  const double ij = 2.0 * q[1] * q[2], ik = 2.0 * q[1] * q[3], jk = 2.0 * q[2] * q[3];
  const double iw = 2.0 * q[1] * q[0], jw = 2.0 * q[2] * q[0], kw = 2.0 * q[3] * q[0];
  R[0] = w2 + i2 - j2 - k2;
  R[1] = ij - kw;
  R[2] = jw + ik;
  R[3] = ij + kw;
  R[4] = w2 - i2 + j2 - k2;
  R[5] = jk - iw;
  R[6] = ik - jw;
  R[7] = jk + iw;
  R[8] = w2 - i2 - j2 + k2;
}

/* apply_periodic_boundary, demsolver.py:451-480 */
static void wrap_position(const dem_params *P, double *x) {
  for (int a = 0; a < 3; ++a)
    if (P->periodic[a]) {
      const double L = P->dmax[a] - P->dmin[a];
      if (x[a] > P->dmax[a]) x[a] -= L;
      else if (x[a] < P->dmin[a]) x[a] += L;
    }
}

/* rotation part of update, demsolver.py:535-572 / 596-632 */
static void rotate(const dem_params *P, dem_state *S, int i) {
  const double dt = P->dt;
  double R[9], Ri[9], Ii[9], Ml[3], wl[3], Iw[3], cr[3], t[3], wdl[3], alpha[3];
  quat_rot(S->quat + 4 * i, R);
  matvec(R, S->moment + 3 * i, Ml);
  matvec(R, S->omega + 3 * i, wl);
  matvec(S->inertia + 9 * i, wl, Iw);
  cross3(wl, Iw, cr);
  for (int k = 0; k < 3; ++k) t[k] = Ml[k] - cr[k];
  inv3(S->inertia + 9 * i, Ii);
  matvec(Ii, t, wdl);
  inv3(R, Ri);
  matvec(Ri, wdl, alpha);
  double *w = S->omega + 3 * i, *q = S->quat + 4 * i;
  for (int k = 0; k < 3; ++k) w[k] += 0.5 * (S->omega_dot[3 * i + k] + alpha[k]) * dt;
  const double dq0 = -0.5 * (q[1] * w[0] + q[2] * w[1] + q[3] * w[2]);
  const double dq1 = +0.5 * (q[0] * w[0] - q[3] * w[1] + q[2] * w[2]);
  const double dq2 = +0.5 * (q[3] * w[0] + q[0] * w[1] + q[1] * w[2]);
  const double dq3 = +0.5 * (-q[2] * w[0] + q[1] * w[1] + q[0] * w[2]);
  q[0] += dq0 * dt;
  q[1] += dq1 * dt;
  q[2] += dq2 * dt;
  q[3] += dq3 * dt;
  const double qn = sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
  for (int k = 0; k < 4; ++k) q[k] /= qn;
  for (int k = 0; k < 3; ++k) S->omega_dot[3 * i + k] = alpha[k];
}

static int cmp_i64(const void *a, const void *b) {
  const int64_t x = *(const int64_t *)a, y = *(const int64_t *)b;
  return (x > y) - (x < y);
}

/* run_simulation, demsolver.py:848-862 (clear_state :422-429, contact :840-846,
 * resolve_wall :809-838, apply_body_force :437-448, update :514-636,
 * late_clear_state :431-436 with remove_inactive_contact :371-386).
 * pairs: scratch of capacity cap pairs; returns the candidate count of the step, or minus
 * that count, with nothing but the force / moment clearing done, when it exceeds cap (the
 * caller grows the scratch and repeats the step). */
int64_t lbo_dem_step(const dem_params *P, dem_state *S, int32_t *pairs, int64_t cap) {
  const int n = P->n, K = P->K;
  for (int k = 0; k < 3 * n; ++k) S->force[k] = S->moment[k] = 0.0;
  const int64_t np = lbo_broadphase(n, S->pos, P->dmin, P->cell, P->T, P->probe27, pairs, cap);
  if (np > cap) return -np;
  const int64_t use = np;
  /* resolve in the bit table's (i, j) order (demsolver.py:412-417) */
  int64_t *keys = malloc(sizeof(int64_t) * (size_t)(use > 0 ? use : 1));
  for (int64_t p = 0; p < use; ++p) keys[p] = (int64_t)pairs[2 * p] * n + pairs[2 * p + 1];
  qsort(keys, (size_t)use, sizeof(int64_t), cmp_i64);
  for (int64_t p = 0; p < use; ++p) resolve(P, S, (int)(keys[p] / n), (int)(keys[p] % n));
  free(keys);
  resolve_walls(P, S);
  const double dt = P->dt;
  for (int i = 0; i < n; ++i) {
    for (int k = 0; k < 3; ++k) {
      S->force[3 * i + k] += S->mass[i] * P->g[k];
      S->force[3 * i + k] += S->force_fluid[3 * i + k];
      S->moment[3 * i + k] += S->moment_fluid[3 * i + k];
    }
    if (S->freeze[i]) continue;
    double *x = S->pos + 3 * i, *v = S->vel + 3 * i;
    if (S->fixvel[i]) {
      for (int k = 0; k < 3; ++k) x[k] += v[k] * dt;
      wrap_position(P, x);
      rotate(P, S, i);
      continue;
    }
    double a[3];
    for (int k = 0; k < 3; ++k) a[k] = S->force[3 * i + k] / S->mass[i];
    for (int k = 0; k < 3; ++k) v[k] += 0.5 * (S->acc[3 * i + k] + a[k]) * dt;
    for (int k = 0; k < 3; ++k) x[k] += v[k] * dt + 0.5 * a[k] * dt * dt;
    wrap_position(P, x);
    for (int k = 0; k < 3; ++k) S->acc[3 * i + k] = a[k];
    rotate(P, S, i);
  }
  for (int i = 0; i < n; ++i) {
    const int m = S->cfn[i] < K ? S->cfn[i] : K;
    int act = 0;
    for (int s = 0; s < m; ++s) act += S->cf_active[i * K + s] ? 1 : 0;
    int o = 0;
    for (int s = 0; s < m && o < act; ++s) {
      const int src = i * K + s;
      if (!S->cf_active[src]) continue;
      const int dst = i * K + o;
      if (dst != src) {
        S->cf_j[dst] = S->cf_j[src];
        S->cf_active[dst] = S->cf_active[src];
        for (int k = 0; k < 3; ++k) {
          S->cf_pos[3 * dst + k] = S->cf_pos[3 * src + k];
          S->cf_force[3 * dst + k] = S->cf_force[3 * src + k];
          S->cf_shear[3 * dst + k] = S->cf_shear[3 * src + k];
        }
      }
      ++o;
    }
    for (int s = act; s < m; ++s) S->cf_active[i * K + s] = 0;
    S->cfn[i] = act;
  }
  return np;
}

