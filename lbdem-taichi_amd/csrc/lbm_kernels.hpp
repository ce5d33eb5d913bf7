// Fused D3Q19 collide-stream kernels (AA pattern, SoA, shifted populations) for gfx950.
//
// One thread per cell, 256-thread workgroups (four wave64).  Even steps read and write the
// cell's own 19 slots (writing to the opposite slot); odd steps gather from the neighbours'
// opposite slots and push to the neighbours' slots.  Each slot is touched by exactly one
// thread per step, so the update is in place and race-free: 19 loads + 19 stores per cell.
// Cells flagged NEAR (a wall neighbour) and cells on the domain edge take a slow path that
// bounces populations back at walls (halfway bounce-back, with the Ladd momentum term at moving
// walls; the wall directions come from the cell's neighbour mask) and wraps periodically at the
// domain edge.  The row-tiled kernels wrap by construction and take the slow path only for NEAR
// cells.
// Bounce-back in the AA pattern, chosen so that every row-tiled store writes whole rows (a
// store with some lanes masked off leaves partly written cache lines, which cost as much as the
// rest of the near-wall handling together):
//  - even step: a population f*_q leaving x towards a wall goes, like every other, into x's own
//    slot OPP[q], with the Ladd term at moving walls; the odd gather reads it back from there.
//    Wall cells store back the values they loaded, unchanged (they hold populations that the
//    near-wall cells read in this step);
//  - odd step: every cell, walls included, pushes all 19 populations; the one heading into a
//    wall w = x + c_q lands in w's slot q (Ladd term applied by x), and the next even step of x
//    reads direction OPP[q] from there.  A wall pushes back the values its target slots already
//    hold (loaded in its gather): a near-wall target reads its bounced population from that slot
//    in the same step.
#pragma once
#include <type_traits>
#include <utility>
#include "common.hpp"

namespace lbdem {

template <int I, int N, typename F>
__device__ __forceinline__ void static_for(F &&fn) {
  if constexpr (I < N) {
    fn(std::integral_constant<int, I>{});
    static_for<I + 1, N>(fn);
  }
}

// accumulate +v / -v / nothing according to a lattice velocity component
template <int a, typename T>
__device__ __forceinline__ void acc(T &s, T v) {
  if constexpr (a == 1) s += v;
  else if constexpr (a == -1) s -= v;
}
template <int a, typename T>
__device__ __forceinline__ T sgn(T v) {
  if constexpr (a == 1) return v;
  else return -v;
}
// c_q . (x, y, z) without multiplications by 0 (D3Q19 has at most two non-zero components)
template <int q, typename T>
__device__ __forceinline__ T cdot(T x, T y, T z) {
  constexpr int a = kCX[q], b = kCY[q], c = kCZ[q];
  if constexpr (a != 0 && b != 0) return sgn<a>(x) + sgn<b>(y);
  else if constexpr (a != 0 && c != 0) return sgn<a>(x) + sgn<c>(z);
  else if constexpr (b != 0 && c != 0) return sgn<b>(y) + sgn<c>(z);
  else if constexpr (a != 0) return sgn<a>(x);
  else if constexpr (b != 0) return sgn<b>(y);
  else if constexpr (c != 0) return sgn<c>(z);
  else return T(0);
}

constexpr int kMaskSolid = LBDEM_FLAG_WALL | LBDEM_FLAG_MOVING;

// all 18 neighbours lie inside the lattice without periodic wrap
__device__ __forceinline__ bool interior(const KParams &p, int i, int j, int k) {
  return i > 0 && i < p.nx - 1 && j > 0 && j < p.ny - 1 && k > 0 && k < p.nz - 1;
}

// Post-stream (shifted) populations of cell c after t steps, t parity `odd`.  Direction q of a
// NEAR cell whose upstream neighbour x - c_q is a wall is a bounced population: after an odd step
// it sits in the wall's slot OPP[q], after an even step in the cell's own slot q.
template <typename S>
__device__ __forceinline__ void gather(const KParams &p, const S *__restrict__ F, int c, uint8_t fl,
                                       bool odd, typename Store<S>::T (&f)[Q]) {
  const size_t N = (size_t)p.ncell;
  int i, j, k;
  cell_ijk(p, c, i, j, k);
  if (!odd) {
    if (!(fl & LBDEM_FLAG_NEAR)) {
      static_for<0, Q>([&](auto qc) {
        constexpr int q = qc;
        f[q] = Store<S>::ld(F + q * N + c);
      });
      return;
    }
    const uint32_t nm = p.nmask[c];
    static_for<0, Q>([&](auto qc) {
      constexpr int q = qc;
      if (q != 0 && ((nm >> kOPP[q]) & 1u))
        f[q] = Store<S>::ld(F + kOPP[q] * N + nbr_index(p, i, j, k, -kCX[q], -kCY[q], -kCZ[q]));
      else
        f[q] = Store<S>::ld(F + q * N + c);
    });
    return;
  }
  if (!(fl & LBDEM_FLAG_NEAR) && interior(p, i, j, k)) {
    const int sx = p.ny * p.nz, sy = p.nz;
    static_for<0, Q>([&](auto qc) {
      constexpr int q = qc;
      const int off = kCX[q] * sx + kCY[q] * sy + kCZ[q];
      f[q] = Store<S>::ld(F + kOPP[q] * N + (c - off));
    });
    return;
  }
  const uint32_t nm = (fl & LBDEM_FLAG_NEAR) ? p.nmask[c] : 0u;
  static_for<0, Q>([&](auto qc) {
    constexpr int q = qc;
    if (q != 0 && ((nm >> kOPP[q]) & 1u))
      f[q] = Store<S>::ld(F + q * N + c);
    else
      f[q] = Store<S>::ld(F + kOPP[q] * N + nbr_index(p, i, j, k, -kCX[q], -kCY[q], -kCZ[q]));
  });
}

// Stores the post-stream populations f of active cell c as the state after t steps (t parity
// `odd`): the inverse of gather.  Every slot written here is gather's source of exactly one value.
template <typename S>
__device__ __forceinline__ void scatter_state(const KParams &p, S *__restrict__ F, int c, uint8_t fl, bool odd,
                                              const typename Store<S>::T (&f)[Q]) {
  const size_t N = (size_t)p.ncell;
  int i, j, k;
  cell_ijk(p, c, i, j, k);
  const uint32_t nm = (fl & LBDEM_FLAG_NEAR) ? p.nmask[c] : 0u;
  static_for<0, Q>([&](auto qc) {
    constexpr int q = qc;
    const bool bounced = q != 0 && ((nm >> kOPP[q]) & 1u);
    if (!odd) {
      if (bounced) Store<S>::st(F + kOPP[q] * N + nbr_index(p, i, j, k, -kCX[q], -kCY[q], -kCZ[q]), f[q]);
      else Store<S>::st(F + q * N + c, f[q]);
    } else {
      if (q == 0 || bounced) Store<S>::st(F + q * N + c, f[q]);
      else Store<S>::st(F + kOPP[q] * N + nbr_index(p, i, j, k, -kCX[q], -kCY[q], -kCZ[q]), f[q]);
    }
  });
}

// 1/x in the collision: the hardware reciprocal (1 ulp) for FP32, a division for FP64
template <typename T>
__device__ __forceinline__ T recip(T x) {
  if constexpr (std::is_same_v<T, float>) {
    return __builtin_amdgcn_rcpf(x);
  } else {
    return T(1) / x;
  }
}

// square root in the collision: the hardware instruction (1 ulp) for FP32, IEEE for FP64
template <typename T>
__device__ __forceinline__ T root(T x) {
  if constexpr (std::is_same_v<T, float>) {
    return __builtin_amdgcn_sqrtf(x);
  } else {
    return sqrt(x);
  }
}

// Tenneti et al. (2011) drag weight of the particle-equivalence IBM, in lattice units.
// Follows EqIMBlattice3D.compute_weight / weight_coefficient (src/unlbdem/eqlattice.py:200-238):
// R = (3 eps / 4 pi)^(1/3), d = 2R, Re = (1-eps) d |u_s - u| / nu, W = 3 pi d nu (1-eps) C_d.
template <typename T>
__device__ __forceinline__ T tenneti_weight(T eps, T slip, T nu) {
  const T pi = T(3.14159265358979323846);
  const T R = cbrt(T(3) * eps / (T(4) * pi));
  const T d = T(2) * R;
  const T om = T(1) - eps;
  const T Re = om * d * slip / nu;
  T Cd = T(0);
  if (om > T(1e-9)) {
    const T om2 = om * om, om3 = om2 * om;
    const T Cd0 = T(1) + T(0.15) * pow(Re, T(0.687));
    const T A = T(5.81) * eps / om3 + T(0.48) * cbrt(eps) / (om3 * om);
    const T e3 = eps * eps * eps;
    const T B = e3 * Re * (T(0.95) + T(0.61) * e3 / om2);
    Cd = om * (Cd0 / om3 + A + B);
  }
  return T(3) * pi * d * nu * om * Cd;
}

// BGK / Smagorinsky / PSC / EqIBM / VANS collision of one cell, in shifted variables.
// f: pre-collision (in), post-collision (out).  Returns rho (for the moving-wall term).
template <int COLL, int COUP, typename T>
__device__ __forceinline__ T collide(const KParams &p, int c, int i, int j, int k, T (&f)[Q]) {
  T dr = T(0), jx = T(0), jy = T(0), jz = T(0);
  static_for<0, Q>([&](auto qc) {
    constexpr int q = qc;
    dr += f[q];
    acc<kCX[q]>(jx, f[q]);
    acc<kCY[q]>(jy, f[q]);
    acc<kCZ[q]>(jz, f[q]);
  });
  const T rho = T(p.rho0) + dr;
  const T inv = recip(rho);
  T ux = jx * inv, uy = jy * inv, uz = jz * inv;

  T fvf = T(1);  // fluid volume fraction in the VANS equilibrium
  T Fx = T(0), Fy = T(0), Fz = T(0);
  if constexpr (COUP == LBDEM_COUPLE_VANS) {
    // macroscopic velocity with the previous force / volume fraction
    // (compute_rhof_vel, src/unlbdem/vanslattice.py:412-440)
    const T fvfp = T(1) - (T)p.volfrac_prev[c];
    const T hx = (T)p.hforce_prev[3 * c], hy = (T)p.hforce_prev[3 * c + 1], hz = (T)p.hforce_prev[3 * c + 2];
    const T s = T(1) / (fvfp * rho);
    ux = (ux + T(0.5) * hx) * s;
    uy = (uy + T(0.5) * hy) * s;
    uz = (uz + T(0.5) * hz) * s;
    fvf = T(1) - (T)p.volfrac[c];
    Fx = (T)p.hforce[3 * c];
    Fy = (T)p.hforce[3 * c + 1];
    Fz = (T)p.hforce[3 * c + 2];
  }

  // relaxation frequency
  T omega;
  if constexpr (COUP == LBDEM_COUPLE_VANS) {
    // compute_omega (src/unlbdem/vanslattice.py:323-358)
    const T svf = T(1) - fvf;
    const T g = T(1) + T(1.25) * svf * T(0.64) / (T(1) - svf);
    const T nu_eff = T(p.nu_lu) * g * g;
    omega = T(1) / (T(3) * nu_eff + T(0.5));
  } else if constexpr (COLL == LBDEM_COLL_BGK) {
    omega = T(p.omega0);
  } else {
    // non-equilibrium momentum flux; sum_q c_a c_b feq = rho/3 delta_ab + rho u_a u_b holds
    // exactly for D3Q19, so Pi^neq = sum c c f - that closed form
    // (compute_strain_rate, src/lbm3d/lbm_solver3d.py:460-498)
    T pxx = T(0), pyy = T(0), pzz = T(0), pxy = T(0), pxz = T(0), pyz = T(0);
    static_for<1, Q>([&](auto qc) {
      constexpr int q = qc;
      if constexpr (kCX[q] != 0) pxx += f[q];
      if constexpr (kCY[q] != 0) pyy += f[q];
      if constexpr (kCZ[q] != 0) pzz += f[q];
      if constexpr (kCX[q] * kCY[q] == 1) pxy += f[q];
      if constexpr (kCX[q] * kCY[q] == -1) pxy -= f[q];
      if constexpr (kCX[q] * kCZ[q] == 1) pxz += f[q];
      if constexpr (kCX[q] * kCZ[q] == -1) pxz -= f[q];
      if constexpr (kCY[q] * kCZ[q] == 1) pyz += f[q];
      if constexpr (kCY[q] * kCZ[q] == -1) pyz -= f[q];
    });
    T er, eux, euy, euz;  // state whose equilibrium defines Pi^neq
    T omega_old = T(0);
    if constexpr (COLL == LBDEM_COLL_SMAG_LAGGED) {
      const size_t N = (size_t)p.ncell;
      const T *lag = reinterpret_cast<const T *>(p.lag);
      er = lag[c];
      eux = lag[N + c];
      euy = lag[2 * N + c];
      euz = lag[3 * N + c];
      omega_old = lag[4 * N + c];
    } else {
      er = rho;
      eux = ux;
      euy = uy;
      euz = uz;
    }
    const T edr3 = (er - T(p.rho0)) * T(1.0 / 3.0);
    const T nxx = pxx - (edr3 + er * eux * eux);
    const T nyy = pyy - (edr3 + er * euy * euy);
    const T nzz = pzz - (edr3 + er * euz * euz);
    const T nxy = pxy - er * eux * euy;
    const T nxz = pxz - er * eux * euz;
    const T nyz = pyz - er * euy * euz;
    const T Pi = root(T(2) * (nxx * nxx + nyy * nyy + nzz * nzz + T(2) * (nxy * nxy + nxz * nxz + nyz * nyz)));
    if constexpr (COLL == LBDEM_COLL_SMAG_LAGGED) {
      // rate = omega_old Pi / (2 cs^2 rho); tau = tau0 + Cs^2 rate / cs^2
      // (computeOmega, src/lbm3d/lbm_solver3d.py:500-513)
      const T rate = omega_old * Pi * T(1.5) * inv;
      omega = recip(T(p.tau0) + T(3) * T(p.cs2_smag) * rate);
      const size_t N = (size_t)p.ncell;
      T *lag = reinterpret_cast<T *>(p.lag);
      lag[c] = rho;
      lag[N + c] = ux;
      lag[2 * N + c] = uy;
      lag[3 * N + c] = uz;
      lag[4 * N + c] = omega;
    } else {
      // tau = tau0 + 9/2 Cs^2 Pi / (rho tau)  =>  tau = (tau0 + sqrt(tau0^2 + 18 Cs^2 Pi/rho)) / 2
      const T tau0 = T(p.tau0);
      omega = T(2) * recip(tau0 + root(tau0 * tau0 + T(18) * T(p.cs2_smag) * Pi * inv));
    }
  }

  const T usq = ux * ux + uy * uy + uz * uz;
  const T rf = rho * fvf;

  // solid coupling weight B and solid velocity (hybrid: B and us of the coarse grain, beta and
  // uf of the fine grains)
  T B = T(0), usx = T(0), usy = T(0), usz = T(0);
  T beta = T(0), ufx = T(0), ufy = T(0), ufz = T(0);
  bool solid = false;
  if constexpr (COUP == LBDEM_COUPLE_PSC) {
    const uint32_t key = p.psc_key[c];
    if (key != 0u) {
      solid = true;
      const int g = 0x7FFFFF - (int)(key & 0x7FFFFFu);
      const float *G = p.psc_grain + 12 * g;
      const T rx = T(i) - (T)G[0], ry = T(j) - (T)G[1], rz = T(k) - (T)G[2];
      const T wx = (T)G[6], wy = (T)G[7], wz = (T)G[8];
      usx = (T)G[3] + (wy * rz - wz * ry);
      usy = (T)G[4] + (wz * rx - wx * rz);
      usz = (T)G[5] + (wx * ry - wy * rx);
      if (key & (1u << 30)) {
        B = T(1);
      } else {
        // B = eps (tau - 1/2) / ((1 - eps) + (tau - 1/2))  (src/lbdem/psclattice3d.py:144-155)
        const T eps = T((key >> 23) & 0x7Fu) * T(1.0 / 125.0);
        const T th = T(p.tau0) - T(0.5);
        B = eps * th / ((T(1) - eps) + th);
      }
    }
  } else if constexpr (COUP == LBDEM_COUPLE_EQIBM) {
    const T eps = (T)p.volfrac[c];
    if (eps > T(0)) {
      solid = true;
      usx = (T)p.velsolid[3 * c];
      usy = (T)p.velsolid[3 * c + 1];
      usz = (T)p.velsolid[3 * c + 2];
      const T dx = usx - ux, dy = usy - uy, dz = usz - uz;
      B = tenneti_weight<T>(eps, sqrt(dx * dx + dy * dy + dz * dz), T(p.nu_lu));
    }
  } else if constexpr (COUP == LBDEM_COUPLE_HYBRID) {
    // HybridLattice3D.collide (src/hybridlbdem/hybridlattice.py:433-565): coarse grain from the
    // key, B = eps (tau - 1/2) / ((1 - eps) + (tau - 1/2)) with this step's omega
    // (_psc_compute_weight :236-244); fine grains from the fractions, Tenneti weight beta where
    // the cell keeps a fluid fraction 1 - eps_f - eps_c > 1e-10 (compute_fine_weights
    // :378-398); beta <= 1 - B where both are present (collide_hybrid :535-537).
    const uint32_t key = p.psc_key[c];
    T ec = T(0);
    if (key != 0u) {
      solid = true;
      const int g = 0x7FFFFF - (int)(key & 0x7FFFFFu);
      const float *G = p.psc_grain + 12 * g;
      const T rx = T(i) - (T)G[0], ry = T(j) - (T)G[1], rz = T(k) - (T)G[2];
      const T wx = (T)G[6], wy = (T)G[7], wz = (T)G[8];
      usx = (T)G[3] + (wy * rz - wz * ry);
      usy = (T)G[4] + (wz * rx - wx * rz);
      usz = (T)G[5] + (wx * ry - wy * rx);
      ec = (key & (1u << 30)) ? T(1) : T((key >> 23) & 0x7Fu) / T(125);
      const T th = recip(omega) - T(0.5);
      B = ec * th / ((T(1) - ec) + th);
    }
    const T ef = (T)p.volfrac[c];
    if (ef > T(0)) {
      solid = true;
      ufx = (T)p.velsolid[3 * c];
      ufy = (T)p.velsolid[3 * c + 1];
      ufz = (T)p.velsolid[3 * c + 2];
      if (T(1) - ef - ec > T(1e-10)) {
        const T dx = ufx - ux, dy = ufy - uy, dz = ufz - uz;
        beta = tenneti_weight<T>(ef, sqrt(dx * dx + dy * dy + dz * dz), T(p.nu_lu));
      }
      if (key != 0u && B + beta > T(1)) beta = T(1) - B;
    }
  }

  const bool forced = (COUP != LBDEM_COUPLE_VANS) && p.has_force;
  const T F3x = T(3) * rho * T(p.fx), F3y = T(3) * rho * T(p.fy), F3z = T(3) * rho * T(p.fz);
  const T vans_pref = T(1) - T(0.5) * omega;

  if (!solid && COUP == LBDEM_COUPLE_VANS) {
    static_for<0, Q>([&](auto qc) {
      constexpr int q = qc;
      const T w = T(kW[q]);
      const T cu = cdot<q>(ux, uy, uz);
      const T feq = w * (dr + rf * (T(3) * cu + T(4.5) * cu * cu - T(1.5) * usq));
      T nf = f[q] - omega * (f[q] - feq);
      // F_q = w (1 - omega/2) (3c - 3u + 9 (c.u) c) . F  (src/unlbdem/vanslattice.py:278-318)
      const T cF = cdot<q>(Fx, Fy, Fz);
      nf += w * vans_pref * (T(3) * cF - T(3) * (ux * Fx + uy * Fy + uz * Fz) + T(9) * cu * cF);
      f[q] = nf;
    });
  } else if (!solid) {
    // BGK relaxation f <- (1 - omega) f + omega feq (+ w c.F3), by opposite pairs:
    // omega feq_q = S_q + A_q with the even part S_q = omega w (A + 4.5 rf (c.u)^2) and the odd
    // part A_q = omega w 3 rf (c.u), so the pair (q, qbar) shares S_q and A_qbar = -A_q.
    const T omc = T(1) - omega;
    const T oA = omega * (dr - T(1.5) * rf * usq);
    const T k45 = omega * T(4.5) * rf, k3 = omega * T(3) * rf;
    // per weight class (axis 1/18, diagonal 1/36): w oA, w k45, w k3
    const T a_oA = T(kW[1]) * oA, a_k45 = T(kW[1]) * k45, a_k3 = T(kW[1]) * k3;
    const T d_oA = T(kW[7]) * oA, d_k45 = T(kW[7]) * k45, d_k3 = T(kW[7]) * k3;
    f[0] = omc * f[0] + T(kW[0]) * oA;
    static_for<1, Q>([&](auto qc) {
      constexpr int q = qc;
      if constexpr (kOPP[q] > q) {
        constexpr int r = kOPP[q];
        constexpr bool axis = kW[q] == kW[1];
        static_assert(axis || kW[q] == kW[7], "D3Q19 has two non-rest weight classes");
        const T cu = cdot<q>(ux, uy, uz);
        const T sym = (axis ? a_k45 : d_k45) * cu * cu + (axis ? a_oA : d_oA);
        T anti = (axis ? a_k3 : d_k3) * cu;
        if (forced) anti += T(kW[q]) * cdot<q>(F3x, F3y, F3z);
        f[q] = omc * f[q] + (sym + anti);
        f[r] = omc * f[r] + (sym - anti);
      }
    });
  } else {
    // weighted collision fpc = f + B Omega_s + (1-B) Omega_f with
    // Omega_s = f_qbar - feq_qbar + feq_s,q - f_q,  Omega_f = -omega (f_q - feq_q)
    // (src/lbdem/psclattice3d.py:240-261, src/unlbdem/eqlattice.py:272-292)
    // Hybrid: f + beta Omega_s1 + B Omega_s2 + (1 - B - beta) Omega_f, Omega_s1 with the fine
    // solid velocity, Omega_s2 with the coarse one (collide_hybrid, hybridlattice.py:523-565;
    // beta = 0 in coarse-only and B = 0 in fine-only cells gives collide_coarse / collide_fine)
    const T ussq = usx * usx + usy * usy + usz * usz;
    const T ufsq = ufx * ufx + ufy * ufy + ufz * ufz;
    T hx = T(0), hy = T(0), hz = T(0);
    const T oB = COUP == LBDEM_COUPLE_HYBRID ? T(1) - B - beta : T(1) - B;
    {
      // rest population: its own opposite
      const T feq0 = T(kW[0]) * (dr - rho * T(1.5) * usq);
      const T feqs0 = T(kW[0]) * (dr - rho * T(1.5) * ussq);
      const T Os = feqs0 - feq0;
      T n0 = f[0] + B * Os - oB * omega * (f[0] - feq0);
      if constexpr (COUP == LBDEM_COUPLE_HYBRID) n0 += beta * (T(kW[0]) * (dr - rho * T(1.5) * ufsq) - feq0);
      f[0] = n0;
    }
    static_for<1, Q>([&](auto qc) {
      constexpr int q = qc;
      if constexpr (kOPP[q] > q) {
        constexpr int r = kOPP[q];
        const T w = T(kW[q]);
        const T cu = cdot<q>(ux, uy, uz);
        const T cus = cdot<q>(usx, usy, usz);
        const T sym = dr + rho * (T(4.5) * cu * cu - T(1.5) * usq);
        const T syms = dr + rho * (T(4.5) * cus * cus - T(1.5) * ussq);
        const T feq_q = w * (sym + T(3) * rho * cu), feq_r = w * (sym - T(3) * rho * cu);
        const T fs_q = w * (syms + T(3) * rho * cus), fs_r = w * (syms - T(3) * rho * cus);
        const T fq = f[q], fr = f[r];
        const T Os_q = fr - feq_r + fs_q - fq;
        const T Os_r = fq - feq_q + fs_r - fr;
        T nq = fq + B * Os_q - oB * omega * (fq - feq_q);
        T nr = fr + B * Os_r - oB * omega * (fr - feq_r);
        if constexpr (COUP == LBDEM_COUPLE_HYBRID) {
          const T cuf = cdot<q>(ufx, ufy, ufz);
          const T symf = dr + rho * (T(4.5) * cuf * cuf - T(1.5) * ufsq);
          const T ff_q = w * (symf + T(3) * rho * cuf), ff_r = w * (symf - T(3) * rho * cuf);
          nq += beta * (fr - feq_r + ff_q - fq);
          nr += beta * (fq - feq_q + ff_r - fr);
        }
        if (forced) {
          const T cF = cdot<q>(F3x, F3y, F3z);
          nq += oB * w * cF;
          nr -= oB * w * cF;
        }
        f[q] = nq;
        f[r] = nr;
        // hydroforce -= B Omega_s c (c_r = -c_q)
        const T d = B * (Os_q - Os_r);
        acc<kCX[q]>(hx, -d);
        acc<kCY[q]>(hy, -d);
        acc<kCZ[q]>(hz, -d);
      }
    });
    if constexpr (COUP == LBDEM_COUPLE_PSC || COUP == LBDEM_COUPLE_HYBRID) {
      p.psc_force[3 * c] = (float)hx;
      p.psc_force[3 * c + 1] = (float)hy;
      p.psc_force[3 * c + 2] = (float)hz;
    }
  }
  return rho;
}

// Moving-wall (Ladd) momentum term for a population leaving along c_q into wall cell m:
// the bounced value is f*_q - 6 w_q rho (c_q . u_w)  (src/lbm3d/lbm_solver3d.py:221-224).
// density of the Ladd term: the bouncing cell's own (VEL_LADD, lbm_solver3d.py:221-224), or the
// inlet cell's for a Ladd velocity inlet (VEL_INLET_LADD, :225-229)
template <typename T>
__device__ __forceinline__ T wall_rho(const KParams &p, uint8_t fm, int m, T rho) {
  return (fm & LBDEM_FLAG_INLET) ? (T)p.rho_field[m] : rho;
}

template <int q, typename T>
__device__ __forceinline__ T ladd_term(const KParams &p, int m, T rho) {
  const T uwx = (T)p.vel[3 * m], uwy = (T)p.vel[3 * m + 1], uwz = (T)p.vel[3 * m + 2];
  return T(6) * T(kW[q]) * rho * cdot<q>(uwx, uwy, uwz);
}

template <typename S, int COLL, int COUP, bool ODD>
__global__ __launch_bounds__(256) void k_step(const KParams p) {
  using T = typename Store<S>::T;
  const int c = blockIdx.x * 256 + threadIdx.x;
  if (c >= p.ncell) return;
  const uint8_t fl = p.flags[c];
  if (!(fl & LBDEM_FLAG_ACTIVE)) return;
  S *__restrict__ F = reinterpret_cast<S *>(p.f);
  const size_t N = (size_t)p.ncell;
  int i, j, k;
  cell_ijk(p, c, i, j, k);
  const bool near = (fl & LBDEM_FLAG_NEAR) || !interior(p, i, j, k);  // bounce-back or periodic wrap

  T f[Q];
  gather<S>(p, F, c, fl, ODD, f);

  const T rho = collide<COLL, COUP, T>(p, c, i, j, k, f);

  if (!near) {
    if constexpr (!ODD) {
      static_for<0, Q>([&](auto qc) {
        constexpr int q = qc;
        Store<S>::st(F + kOPP[q] * N + c, f[q]);
      });
    } else {
      const int sx = p.ny * p.nz, sy = p.nz;
      static_for<0, Q>([&](auto qc) {
        constexpr int q = qc;
        const int off = kCX[q] * sx + kCY[q] * sy + kCZ[q];
        Store<S>::st(F + q * N + (c + off), f[q]);
      });
    }
  } else {
    const uint32_t nm = (fl & LBDEM_FLAG_NEAR) ? p.nmask[c] : 0u;
    static_for<0, Q>([&](auto qc) {
      constexpr int q = qc;
      if constexpr (q == 0) {
        Store<S>::st(F + c, f[0]);
      } else {
        const int m = nbr_index(p, i, j, k, kCX[q], kCY[q], kCZ[q]);
        T v = f[q];
        if (((nm >> q) & 1u) && (nm & kNbrMoving)) {
          const uint8_t fm = p.flags[m];
          if (fm & LBDEM_FLAG_MOVING) v -= ladd_term<q, T>(p, m, wall_rho(p, fm, m, rho));
        }
        if constexpr (!ODD) Store<S>::st(F + kOPP[q] * N + c, v);
        else Store<S>::st(F + q * N + m, v);  // into a wall cell too: read there by the next even step
      }
    });
  }
}

using StepKernel = void (*)(const KParams);

// f <- feq(rho, u) at active cells as the state t = 0; lag state initialised
template <typename S, int COUP>
__global__ __launch_bounds__(256) void k_init(const KParams p) {
  using T = typename Store<S>::T;
  const int c = blockIdx.x * 256 + threadIdx.x;
  if (c >= p.ncell) return;
  const uint8_t fl = p.flags[c];
  if (!(fl & LBDEM_FLAG_ACTIVE)) return;
  S *__restrict__ F = reinterpret_cast<S *>(p.f);
  const size_t N = (size_t)p.ncell;
  const T rho = (T)p.rho_out[c];
  const T ux = (T)p.vel[3 * c], uy = (T)p.vel[3 * c + 1], uz = (T)p.vel[3 * c + 2];
  const T dr = rho - T(p.rho0);
  const T usq = ux * ux + uy * uy + uz * uz;
  T fvf = T(1);
  if (COUP == LBDEM_COUPLE_VANS && p.volfrac) fvf = T(1) - (T)p.volfrac[c];
  T fe[Q];
  static_for<0, Q>([&](auto qc) {
    constexpr int q = qc;
    const T cu = cdot<q>(ux, uy, uz);
    fe[q] = T(kW[q]) * (dr + rho * fvf * (T(3) * cu + T(4.5) * cu * cu - T(1.5) * usq));
  });
  scatter_state<S>(p, F, c, fl, false, fe);
  if (p.lag) {
    T *lag = reinterpret_cast<T *>(p.lag);
    lag[c] = rho;
    lag[N + c] = ux;
    lag[2 * N + c] = uy;
    lag[3 * N + c] = uz;
    lag[4 * N + c] = (T)p.omega0;
  }
}

// rho, u of the post-stream state (compute_rho_vel / compute_rhof_vel)
template <typename S, int COUP>
__global__ __launch_bounds__(256) void k_macros(const KParams p) {
  using T = typename Store<S>::T;
  const int c = blockIdx.x * 256 + threadIdx.x;
  if (c >= p.ncell) return;
  const uint8_t fl = p.flags[c];
  if (!(fl & LBDEM_FLAG_ACTIVE)) return;
  if (COUP != LBDEM_COUPLE_VANS && !(fl & LBDEM_FLAG_FLUID)) return;
  if (p.mask && !(p.mask[c] > 0.0f)) return;
  const S *F = reinterpret_cast<const S *>(p.f);
  T f[Q];
  gather<S>(p, F, c, fl, p.t_odd != 0, f);
  T dr = T(0), jx = T(0), jy = T(0), jz = T(0);
  static_for<0, Q>([&](auto qc) {
    constexpr int q = qc;
    dr += f[q];
    acc<kCX[q]>(jx, f[q]);
    acc<kCY[q]>(jy, f[q]);
    acc<kCZ[q]>(jz, f[q]);
  });
  const T rho = T(p.rho0) + dr;
  T ux = jx / rho, uy = jy / rho, uz = jz / rho;
  if constexpr (COUP == LBDEM_COUPLE_VANS) {
    // the force / volume fraction of the step that produced this state
    const T fvf = T(1) - (T)p.volfrac[c];
    const T s = T(1) / (fvf * rho);
    ux = (ux + T(0.5) * (T)p.hforce[3 * c]) * s;
    uy = (uy + T(0.5) * (T)p.hforce[3 * c + 1]) * s;
    uz = (uz + T(0.5) * (T)p.hforce[3 * c + 2]) * s;
  }
  p.rho_out[c] = (float)rho;
  p.vel_out[3 * c] = (float)ux;
  p.vel_out[3 * c + 1] = (float)uy;
  p.vel_out[3 * c + 2] = (float)uz;
}

template <typename S>
__global__ __launch_bounds__(256) void k_read_f(const KParams p) {
  using T = typename Store<S>::T;
  const int c = blockIdx.x * 256 + threadIdx.x;
  if (c >= p.ncell) return;
  const size_t N = (size_t)p.ncell;
  const uint8_t fl = p.flags[c];
  T f[Q];
  const bool act = fl & LBDEM_FLAG_ACTIVE;
  if (act) gather<S>(p, reinterpret_cast<const S *>(p.f), c, fl, p.t_odd != 0, f);
  static_for<0, Q>([&](auto qc) {
    constexpr int q = qc;
    const double v = act ? (double)f[q] + kW[q] * p.rho0 : 0.0;
    if (p.out_f64) reinterpret_cast<double *>(p.out)[q * N + c] = v;
    else reinterpret_cast<float *>(p.out)[q * N + c] = (float)v;
  });
}

template <typename S>
__global__ __launch_bounds__(256) void k_write_f(const KParams p) {
  using T = typename Store<S>::T;
  const int c = blockIdx.x * 256 + threadIdx.x;
  if (c >= p.ncell) return;
  const uint8_t fl = p.flags[c];
  if (!(fl & LBDEM_FLAG_ACTIVE)) return;
  const size_t N = (size_t)p.ncell;
  S *F = reinterpret_cast<S *>(p.f);
  T f[Q];
  static_for<0, Q>([&](auto qc) {
    constexpr int q = qc;
    const double v = p.out_f64 ? reinterpret_cast<const double *>(p.in)[q * N + c]
                               : (double)reinterpret_cast<const float *>(p.in)[q * N + c];
    f[q] = (T)(v - kW[q] * p.rho0);
  });
  scatter_state<S>(p, F, c, fl, p.t_odd != 0, f);
}

// ---------------------------------------------------------------------------------------
// Wet-node boundaries after a step (the second loop of BasicLattice3D.stream,
// src/lbm3d/lbm_solver3d.py:241-250), one thread per boundary cell.  The face with normal axis a
// and inward sign s: the populations with c_a = -s arrived from inside; those with c_a = s are
// rebuilt from their opposites (vel_zouHe :671-820, pre_zouHe :823-957), in the reference's order
// of terms.  A boundary cell's populations are read and written at the state's parity through
// gather / scatter_state, which touch only that cell's own populations.
// ---------------------------------------------------------------------------------------
__constant__ int kZhZero[3][9] = {{0, 3, 4, 5, 6, 15, 16, 17, 18},
                                  {0, 1, 2, 5, 6, 11, 12, 13, 14},
                                  {0, 1, 2, 3, 4, 7, 8, 9, 10}};
__constant__ int kZhIn[6][5] = {{2, 8, 10, 12, 14}, {1, 7, 9, 11, 13}, {4, 8, 9, 16, 18},
                                {3, 7, 10, 15, 17}, {6, 12, 13, 16, 17}, {5, 11, 14, 15, 18}};
// [axis][transverse axis 0/1][c_t = +1 / -1][3], transverse axes in x, y, z order
__constant__ int kZhN[3][2][2][3] = {{{{3, 15, 17}, {4, 18, 16}}, {{5, 18, 15}, {6, 17, 16}}},
                                     {{{1, 11, 13}, {2, 14, 12}}, {{5, 11, 14}, {6, 13, 12}}},
                                     {{{1, 7, 9}, {2, 10, 8}}, {{3, 7, 10}, {4, 9, 8}}}};
__constant__ int kBcCx[Q] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0};
__constant__ int kBcCy[Q] = {0, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 0, 0, 1, -1, 1, -1};
__constant__ int kBcCz[Q] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1, 1, -1, -1, 1};
__constant__ int kBcOpp[Q] = {0, 2, 1, 4, 3, 6, 5, 8, 7, 10, 9, 12, 11, 14, 13, 16, 15, 18, 17};

// the known populations' sum of the density formula: c_a = 0 ones plus twice the incoming ones
template <typename T>
__device__ T zh_sum_known(const T *f, int face) {
  const int a = face >> 1;
  T s0 = T(0), s1 = T(0);
  for (int m = 0; m < 9; ++m) s0 += f[kZhZero[a][m]];
  for (int m = 0; m < 5; ++m) s1 += f[kZhIn[face][m]];
  return s0 + T(2) * s1;
}

// the five unknown populations for density rho and velocity u; the pressure condition uses only
// the normal velocity and no rho u_t term in N_t
template <typename T>
__device__ void zh_rebuild(T *f, int face, T rho, const T *u, bool pressure) {
  const int a = face >> 1, s = (face & 1) ? -1 : 1;
  const int ts[2] = {a == 0 ? 1 : 0, a == 2 ? 1 : 2};
  T N[2];
  for (int t = 0; t < 2; ++t) {
    const int *pp = kZhN[a][t][0], *mm = kZhN[a][t][1];
    N[t] = T(0.5) * (f[pp[0]] + f[pp[1]] + f[pp[2]] - (f[mm[0]] + f[mm[1]] + f[mm[2]]));
    if (!pressure) N[t] = N[t] - T(1.0 / 3.0) * rho * u[ts[t]];
  }
  for (int q = 1; q < Q; ++q) {
    const int c[3] = {kBcCx[q], kBcCy[q], kBcCz[q]};
    if (c[a] != s) continue;
    T cu = T(0);
    int diag = -1;
    for (int b = 0; b < 3; ++b) {
      if (c[b] == 0) continue;
      if (b != a) diag = b;
      if (pressure && b != a) continue;
      cu += T(c[b]) * u[b];
    }
    if (diag < 0) f[q] = f[kBcOpp[q]] + T(1.0 / 3.0) * rho * cu;
    else f[q] = f[kBcOpp[q]] + T(1.0 / 6.0) * rho * cu - T(c[diag]) * N[diag == ts[0] ? 0 : 1];
  }
}

template <typename S>
__global__ __launch_bounds__(256) void k_boundaries(const KParams p, const int32_t *cells, const uint8_t *kinds,
                                                    int ncells, float *rho_f, float *vel_f) {
  using T = typename Store<S>::T;
  const int e = blockIdx.x * 256 + threadIdx.x;
  if (e >= ncells) return;
  const int c = cells[e];
  const int type = kinds[e] & 15, face = kinds[e] >> 4;
  S *F = reinterpret_cast<S *>(p.f);
  const bool odd = p.t_odd != 0;
  int i, j, k;
  cell_ijk(p, c, i, j, k);
  const bool has_face = face <= 5;
  const int a = face >> 1, s = (face & 1) ? -1 : 1;
  const int m = has_face ? nbr_index(p, i, j, k, a == 0 ? s : 0, a == 1 ? s : 0, a == 2 ? s : 0) : c;
  // rho and u of the inner neighbour m: from its populations where it is a fluid cell (the
  // reference refreshed them in the first loop), else from the fields
  auto inner = [&](T &r, T (&u)[3]) {
    const uint8_t fm = p.flags[m];
    if ((fm & LBDEM_FLAG_FLUID) && (fm & LBDEM_FLAG_ACTIVE)) {
      T g[Q];
      gather<S>(p, F, m, fm, odd, g);
      T dr = T(0), jx = T(0), jy = T(0), jz = T(0);
      static_for<0, Q>([&](auto qc) {
        constexpr int q = qc;
        dr += g[q];
        acc<kCX[q]>(jx, g[q]);
        acc<kCY[q]>(jy, g[q]);
        acc<kCZ[q]>(jz, g[q]);
      });
      r = T(p.rho0) + dr;
      u[0] = jx / r;
      u[1] = jy / r;
      u[2] = jz / r;
    } else {
      r = (T)rho_f[m];
      for (int b = 0; b < 3; ++b) u[b] = (T)vel_f[3 * m + b];
    }
  };
  if (type == LBDEM_BC_LADD_INLET) {  // bc_ladd_inlet: density from the inner neighbour
    if (!has_face) return;
    T r, u[3];
    inner(r, u);
    rho_f[c] = (float)r;
    return;
  }
  const uint8_t fl = p.flags[c];
  if (!(fl & LBDEM_FLAG_ACTIVE)) return;
  T f[Q];
  if (type == LBDEM_BC_EXIT) {  // bc_vel_exit: the neighbour's velocity, rho = 1, f = feq
    T u[3];
    if (has_face) {
      T r;
      inner(r, u);
      for (int b = 0; b < 3; ++b) vel_f[3 * c + b] = (float)u[b];
    }
    for (int b = 0; b < 3; ++b) u[b] = (T)vel_f[3 * c + b];
    rho_f[c] = 1.0f;
    const T usq = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
    static_for<0, Q>([&](auto qc) {
      constexpr int q = qc;
      const T cu = cdot<q>(u[0], u[1], u[2]);
      f[q] = T(kW[q]) * (T(1) + T(3) * cu + T(4.5) * cu * cu - T(1.5) * usq) - T(kW[q] * p.rho0);
    });
    scatter_state<S>(p, F, c, fl, odd, f);
    return;
  }
  if (!has_face) return;  // Zou-He without a face: the reference's if-chain does nothing
  gather<S>(p, F, c, fl, odd, f);
  T g[Q];  // unshifted
  for (int q = 0; q < Q; ++q) g[q] = f[q] + T(kW[q] * p.rho0);
  T u[3];
  if (type == LBDEM_BC_VEL_ZOUHE) {
    for (int b = 0; b < 3; ++b) u[b] = (T)vel_f[3 * c + b];
    const T coeff = s > 0 ? T(1) / (T(1) - u[a]) : T(1) / (T(1) + u[a]);
    const T r = coeff * zh_sum_known(g, face);
    rho_f[c] = (float)r;
    zh_rebuild(g, face, r, u, false);
  } else if (type == LBDEM_BC_PRE_ZOUHE) {
    const T r = (T)rho_f[c];
    const T sk = zh_sum_known(g, face);
    const T un = s > 0 ? T(1) - sk / r : sk / r - T(1);
    vel_f[3 * c + a] = (float)un;
    u[0] = u[1] = u[2] = T(0);
    u[a] = un;
    zh_rebuild(g, face, r, u, true);
  } else {
    return;
  }
  for (int q = 0; q < Q; ++q) f[q] = g[q] - T(kW[q] * p.rho0);
  scatter_state<S>(p, F, c, fl, odd, f);
}

// ---------------------------------------------------------------------------------------
// Free-slip walls after a step (specular reflection, the FREE_SLIP branch of
// BasicLattice3D.stream, src/lbm3d/lbm_solver3d.py:230-236, mirror maps :50-57).  The fused step
// treats a free-slip cell as a resting wall (halfway bounce-back); this pass then replaces each
// population q of an active cell y whose upstream neighbour y - c_q is a free-slip cell by the
// post-collision population qsy[q] of the cell x' = y - c_q with the wall-normal component of the
// offset dropped (the reference pulls fpc[x'][qsy[q]]), or by 0 where x' does not collide (the
// reference's never-written fpc).  Every such post-collision value is a post-stream population
// of this step (the host finds which: the bounced one of x' where x' + c_qsy[q] is a wall, else
// the one streamed into x' + c_qsy[q]), and sources can be targets of other pairs, so all
// sources are read (k_free_slip_read) before any target is written (k_free_slip_write).
// pairs: four int32 per pair -- target cell, target direction, source cell (-1: zero), source
// direction.  Mirrored directions have the weight of the original, so the shifted stored values
// copy unchanged.
// ---------------------------------------------------------------------------------------
// index into the population array of the post-stream population s of active cell z at the
// state's parity (the address gather uses)
__device__ __forceinline__ size_t post_stream_slot(const KParams &p, int z, int s, bool odd) {
  if (s == 0) return (size_t)z;
  const int o = kBcOpp[s];
  const bool bounced = (p.flags[z] & LBDEM_FLAG_NEAR) && ((p.nmask[z] >> o) & 1u);
  if (bounced == odd) return (size_t)s * p.ncell + z;  // bounced after an odd step, streamed after an even one
  int i, j, k;
  cell_ijk(p, z, i, j, k);
  return (size_t)o * p.ncell + nbr_index(p, i, j, k, -kBcCx[s], -kBcCy[s], -kBcCz[s]);
}

template <typename S>
__global__ __launch_bounds__(256) void k_free_slip_read(const KParams p, const int32_t *pairs, int npairs,
                                                        double *val) {
  const int e = blockIdx.x * 256 + threadIdx.x;
  if (e >= npairs) return;
  const int q = pairs[4 * e + 1], z = pairs[4 * e + 2], s = pairs[4 * e + 3];
  if (z < 0) {  // a population of 0, shifted by w_q rho0
    const double w = q == 0 ? 1.0 / 3.0 : (q <= 6 ? 1.0 / 18.0 : 1.0 / 36.0);
    val[e] = (double)(typename Store<S>::T)(-w * p.rho0);
    return;
  }
  val[e] = (double)Store<S>::ld(reinterpret_cast<const S *>(p.f) + post_stream_slot(p, z, s, p.t_odd != 0));
}

template <typename S>
__global__ __launch_bounds__(256) void k_free_slip_write(const KParams p, const int32_t *pairs, int npairs,
                                                         const double *val) {
  const int e = blockIdx.x * 256 + threadIdx.x;
  if (e >= npairs) return;
  const size_t a = post_stream_slot(p, pairs[4 * e], pairs[4 * e + 1], p.t_odd != 0);
  Store<S>::st(reinterpret_cast<S *>(p.f) + a, (typename Store<S>::T)val[e]);
}

// Stress from the non-equilibrium populations, BasicLattice3D.compute_stress
// (src/lbm3d/lbm_solver3d.py:347-405): sigma_ab = -(1 - omega/2) sum_q c_a c_b (f_q - feq_q), f
// the populations after t steps, feq and omega those of the last collision.  With lagged
// Smagorinsky the lag state holds them (exact); otherwise the equilibrium of the current
// populations, omega0 (BGK) or the fixed-point omega of the current populations.  Out: seven
// planes of ncell floats (xx, yy, zz, xy, xz, yz, sqrt((xy^2 + xz^2 + yz^2) / 2)); cells the
// collision skips are left untouched.
template <typename S>
__global__ __launch_bounds__(256) void k_stress(const KParams p) {
  using T = typename Store<S>::T;
  const int c = blockIdx.x * 256 + threadIdx.x;
  if (c >= p.ncell) return;
  const uint8_t fl = p.flags[c];
  if (!(fl & LBDEM_FLAG_ACTIVE)) return;
  T f[Q];
  gather<S>(p, reinterpret_cast<const S *>(p.f), c, fl, p.t_odd != 0, f);
  T dr = T(0), jx = T(0), jy = T(0), jz = T(0);
  T pxx = T(0), pyy = T(0), pzz = T(0), pxy = T(0), pxz = T(0), pyz = T(0);
  static_for<0, Q>([&](auto qc) {
    constexpr int q = qc;
    dr += f[q];
    acc<kCX[q]>(jx, f[q]);
    acc<kCY[q]>(jy, f[q]);
    acc<kCZ[q]>(jz, f[q]);
    if constexpr (kCX[q] != 0) pxx += f[q];
    if constexpr (kCY[q] != 0) pyy += f[q];
    if constexpr (kCZ[q] != 0) pzz += f[q];
    if constexpr (kCX[q] * kCY[q] == 1) pxy += f[q];
    if constexpr (kCX[q] * kCY[q] == -1) pxy -= f[q];
    if constexpr (kCX[q] * kCZ[q] == 1) pxz += f[q];
    if constexpr (kCX[q] * kCZ[q] == -1) pxz -= f[q];
    if constexpr (kCY[q] * kCZ[q] == 1) pyz += f[q];
    if constexpr (kCY[q] * kCZ[q] == -1) pyz -= f[q];
  });
  const T rho = T(p.rho0) + dr;
  T er = rho, ux = jx / rho, uy = jy / rho, uz = jz / rho, omega = T(p.omega0);
  const size_t N = (size_t)p.ncell;
  if (p.collision == LBDEM_COLL_SMAG_LAGGED && p.coupling != LBDEM_COUPLE_VANS) {
    const T *lag = reinterpret_cast<const T *>(p.lag);
    er = lag[c];
    ux = lag[N + c];
    uy = lag[2 * N + c];
    uz = lag[3 * N + c];
    omega = lag[4 * N + c];
  }
  // sum_q c_a c_b feq_q = rho/3 delta_ab + rho u_a u_b (D3Q19), in shifted variables
  const T edr3 = (er - T(p.rho0)) * T(1.0 / 3.0);
  const T nxx = pxx - (edr3 + er * ux * ux), nyy = pyy - (edr3 + er * uy * uy), nzz = pzz - (edr3 + er * uz * uz);
  const T nxy = pxy - er * ux * uy, nxz = pxz - er * ux * uz, nyz = pyz - er * uy * uz;
  if (p.coupling == LBDEM_COUPLE_VANS) {
    // compute_omega (src/unlbdem/vanslattice.py:323-358)
    const T svf = (T)p.volfrac[c];
    const T g = T(1) + T(1.25) * svf * T(0.64) / (T(1) - svf);
    omega = T(1) / (T(3) * T(p.nu_lu) * g * g + T(0.5));
  } else if (p.collision == LBDEM_COLL_SMAG_FIXED) {
    const T Pi = sqrt(T(2) * (nxx * nxx + nyy * nyy + nzz * nzz + T(2) * (nxy * nxy + nxz * nxz + nyz * nyz)));
    const T tau0 = T(p.tau0);
    omega = T(2) / (tau0 + sqrt(tau0 * tau0 + T(18) * T(p.cs2_smag) * Pi / rho));
  }
  const T r = -(T(1) - T(0.5) * omega);
  float *out = reinterpret_cast<float *>(p.out);
  const T sxy = r * nxy, sxz = r * nxz, syz = r * nyz;
  out[c] = (float)(r * nxx);
  out[N + c] = (float)(r * nyy);
  out[2 * N + c] = (float)(r * nzz);
  out[3 * N + c] = (float)sxy;
  out[4 * N + c] = (float)sxz;
  out[5 * N + c] = (float)syz;
  out[6 * N + c] = (float)sqrt(T(0.5) * (sxy * sxy + sxz * sxz + syz * syz));
}

// per-workgroup FP64 partial sums: rho, rho u (3), rho |u|^2 / 2
template <typename S>
__global__ __launch_bounds__(256) void k_reduce_partial(const KParams p) {
  using T = typename Store<S>::T;
  const int c = blockIdx.x * 256 + threadIdx.x;
  double s[5] = {0, 0, 0, 0, 0};
  if (c < p.ncell) {
    const uint8_t fl = p.flags[c];
    if (fl & LBDEM_FLAG_ACTIVE) {
      T f[Q];
      gather<S>(p, reinterpret_cast<const S *>(p.f), c, fl, p.t_odd != 0, f);
      double dr = 0, jx = 0, jy = 0, jz = 0;
      static_for<0, Q>([&](auto qc) {
        constexpr int q = qc;
        const double v = (double)f[q];
        dr += v;
        acc<kCX[q]>(jx, v);
        acc<kCY[q]>(jy, v);
        acc<kCZ[q]>(jz, v);
      });
      const double rho = p.rho0 + dr;
      s[0] = rho;
      s[1] = jx;
      s[2] = jy;
      s[3] = jz;
      s[4] = 0.5 * (jx * jx + jy * jy + jz * jz) / rho;
    }
  }
  __shared__ double sh[4][5];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
  for (int m = 0; m < 5; ++m) {
    const double v = wave_sum(s[m]);
    if (lane == 0) sh[wv][m] = v;
  }
  __syncthreads();
  if (threadIdx.x < 5) {
    const int m = threadIdx.x;
    p.work[(size_t)blockIdx.x * 5 + m] = ((sh[0][m] + sh[1][m]) + (sh[2][m] + sh[3][m]));
  }
}

}  // namespace lbdem

// ---------------------------------------------------------------------------------------
// Pencil-tiled variant: one workgroup per (i, j) row of nz cells, V consecutive cells per
// thread.  Every global load and store of a step covers whole aligned row segments: the
// x/y neighbours of a row are other rows (same k range), and the +-1 shift along k of the
// odd step is applied by staging the ten z-moving populations through LDS, with periodic
// wrap inside the row.  Used when nz <= 1024 V and nz % V == 0.
// ---------------------------------------------------------------------------------------
namespace lbdem {

// index of a z-moving direction among {5, 6, 11, ..., 18}
template <int q>
constexpr int zslot() {
  return q == 5 ? 0 : q == 6 ? 1 : q - 9;
}
constexpr int kNZ = 10;

// V consecutive elements (V = 1 or 2) as one access of V * sizeof(S) bytes
template <int BYTES> struct Raw;
template <> struct Raw<4> { using type = uint32_t; };
template <> struct Raw<8> { using type = uint64_t; };
template <> struct Raw<16> { using type = __attribute__((ext_vector_type(2))) uint64_t; };

template <typename S, int V>
struct Vec {
  using T = typename Store<S>::T;
  static constexpr int kBytes = V * (int)sizeof(S);
  static __device__ __forceinline__ void ld(const S *p, T (&v)[V]) {
    if constexpr (V == 1) {
      v[0] = Store<S>::ld(p);
    } else {
      static_assert(kBytes == 4 || kBytes == 8 || kBytes == 16, "4, 8 or 16 bytes per access");
      using R = typename Raw<kBytes>::type;
      const R w = lbdem_ld_raw(reinterpret_cast<const R *>(p));
      S e[V];
      __builtin_memcpy(e, &w, sizeof(e));
#pragma unroll
      for (int m = 0; m < V; ++m) v[m] = Store<S>::cvt(e[m]);
    }
  }
  static __device__ __forceinline__ void st(S *p, const T (&v)[V]) {
    if constexpr (V == 1) {
      Store<S>::st(p, v[0]);
    } else {
      using R = typename Raw<kBytes>::type;
      S e[V];
#pragma unroll
      for (int m = 0; m < V; ++m) e[m] = Store<S>::pack(v[m]);
      R w;
      __builtin_memcpy(&w, e, sizeof(e));
      lbdem_st_raw(reinterpret_cast<R *>(p), w);
    }
  }
};

template <typename S, int COLL, int COUP, bool ODD, int V>
__global__ __launch_bounds__(pencil_max_threads(V)) void k_step_pencil(const KParams p) {
  // a row without active cells would store every slot unchanged: walls store what they loaded
  // (even step: their own slots; odd step: their active targets' slots, and anything into the
  // slots of wall targets, which nothing reads)
  if (p.flags[(size_t)blockIdx.x * p.nz] & LBDEM_FLAG_ROW_IDLE) return;
  using T = typename Store<S>::T;
  extern __shared__ __align__(16) unsigned char smem[];
  T *buf = reinterpret_cast<T *>(smem);  // [kNZ][ncz]
  const int nz = p.nz;
  const int ncz = blockDim.x * V;         // row cells covered by the block (>= nz)
  const int row = blockIdx.x;
  const int i = row / p.ny, j = row - (row / p.ny) * p.ny;
  const int k0 = threadIdx.x * V;
  const unsigned uk0 = threadIdx.x * V;  // per-lane part of every row address (uniform row base)
  const bool lane_in = k0 < nz;
  S *__restrict__ F = reinterpret_cast<S *>(p.f);
  const size_t N = (size_t)p.ncell;
  const int base = row * nz;

  uint8_t fl[V];
  bool act[V];
#pragma unroll
  for (int v = 0; v < V; ++v) {
    fl[v] = lane_in ? p.flags[base + k0 + v] : (uint8_t)0;
    act[v] = fl[v] & LBDEM_FLAG_ACTIVE;
  }
  uint32_t nmv[V];  // wall-neighbour bits (NEAR cells only)
  uint32_t tmv[V];  // wall cells: the active neighbours (odd step: slots to push back unchanged)
  uint32_t nm_any = 0u;  // union over this thread's cells: the wall paths below branch on it
#pragma unroll
  for (int v = 0; v < V; ++v) {
    const uint32_t m = (fl[v] & (LBDEM_FLAG_NEAR | LBDEM_FLAG_WALL | LBDEM_FLAG_MOVING)) ? p.nmask[base + k0 + v]
                                                                                          : 0u;
    nmv[v] = (fl[v] & LBDEM_FLAG_NEAR) ? m : 0u;
    tmv[v] = act[v] ? 0u : m;
    nm_any |= nmv[v];
  }
  T f[V][Q];

  // ---------------------------------------------------------------- gather
  if constexpr (!ODD) {
    if (lane_in) {
      // all 19 row loads first: a branch between a 16-bit load and its conversion would make
      // each conversion wait for its load before the next load is issued
      static_for<0, Q>([&](auto qc) {
        constexpr int q = qc;
        T t[V];
        Vec<S, V>::ld(F + (q * N + (size_t)base) + uk0, t);
#pragma unroll
        for (int v = 0; v < V; ++v) f[v][q] = t[v];
      });
      if (nm_any & ~kNbrMoving) {  // bounced at the wall x - c_q last step: in the wall's slot
        static_for<1, Q>([&](auto qc) {
          constexpr int q = qc;
#pragma unroll
          for (int v = 0; v < V; ++v) {
            if ((nmv[v] >> kOPP[q]) & 1u)
              f[v][q] = Store<S>::ld(F + kOPP[q] * N + nbr_index(p, i, j, k0 + v, -kCX[q], -kCY[q], -kCZ[q]));
          }
        });
      }
    }
  } else {
    if (lane_in) {  // one basic block for all 19 loads, so they are all issued before any wait
      static_for<0, Q>([&](auto qc) {
        constexpr int q = qc;
        constexpr int cx = kCX[q], cy = kCY[q], cz = kCZ[q];
        const int srow = wrapi(i - cx, p.nx) * p.ny + wrapi(j - cy, p.ny);
        T t[V];
        Vec<S, V>::ld(F + (kOPP[q] * N + (size_t)srow * nz) + uk0, t);
        if constexpr (cz == 0) {
#pragma unroll
          for (int v = 0; v < V; ++v) f[v][q] = t[v];
        } else {
#pragma unroll
          for (int v = 0; v < V; ++v) buf[zslot<q>() * ncz + k0 + v] = t[v];
        }
      });
    }
    __syncthreads();
    if (lane_in) {
      static_for<1, Q>([&](auto qc) {
        constexpr int q = qc;
        constexpr int cz = kCZ[q];
        if constexpr (cz != 0) {
#pragma unroll
          for (int v = 0; v < V; ++v) {
            int ks = k0 + v - cz;
            ks = ks < 0 ? ks + nz : (ks >= nz ? ks - nz : ks);
            f[v][q] = buf[zslot<q>() * ncz + ks];
          }
        }
      });
#pragma unroll
      for (int v = 0; v < V; ++v) {
        const int k = k0 + v;
        if (fl[v] & LBDEM_FLAG_NEAR) {  // bounced at the wall x - c_q last step: in the own slot q
          const uint32_t nm = nmv[v];
          static_for<1, Q>([&](auto qc) {
            constexpr int q = qc;
            if ((nm >> kOPP[q]) & 1u) f[v][q] = Store<S>::ld(F + q * N + base + k);
          });
        } else if (!act[v]) {
          // A wall pushes back what its targets hold: a target x + c_q that is an active cell
          // reads that slot (its bounced population) in this step, and no other thread writes it.
          // A target that is a wall itself holds nothing that is read (its slot q would carry the
          // bounced population of x, a wall), so its store may carry any value.
          const uint32_t tm = tmv[v];
          static_for<1, Q>([&](auto qc) {
            constexpr int q = qc;
            if ((tm >> q) & 1u) f[v][q] = Store<S>::ld(F + q * N + nbr_index(p, i, j, k, kCX[q], kCY[q], kCZ[q]));
          });
        }
      }
    }
  }

  // ---------------------------------------------------------------- collide
  T rho[V];
#pragma unroll
  for (int v = 0; v < V; ++v) {
    rho[v] = T(0);
    if (act[v]) rho[v] = collide<COLL, COUP, T>(p, base + k0 + v, i, j, k0 + v, f[v]);
  }

  // Ladd term of the populations heading into a moving wall (wall neighbours are rare: one
  // branch per direction on the union of this thread's masks)
  auto moving_wall = [&](auto qc, int v, T &val) {
    constexpr int q = decltype(qc)::value;
    if (((nmv[v] >> q) & 1u) && (nmv[v] & kNbrMoving)) {
      const int m = nbr_index(p, i, j, k0 + v, kCX[q], kCY[q], kCZ[q]);
      const uint8_t fm = p.flags[m];
      if (fm & LBDEM_FLAG_MOVING) val -= ladd_term<q, T>(p, m, wall_rho(p, fm, m, rho[v]));
    }
  };

  // ---------------------------------------------------------------- scatter (whole rows)
  if constexpr (!ODD) {
    if (lane_in) {
      static_for<0, Q>([&](auto qc) {
        constexpr int q = qc;
        T t[V];
#pragma unroll
        for (int v = 0; v < V; ++v) t[v] = act[v] ? f[v][q] : f[v][kOPP[q]];  // walls: unchanged
        if (q != 0 && (nm_any & kNbrMoving)) {
#pragma unroll
          for (int v = 0; v < V; ++v) moving_wall(qc, v, t[v]);
        }
        Vec<S, V>::st(F + (kOPP[q] * N + (size_t)base) + uk0, t);
      });
    }
  } else {
    __syncthreads();  // the gather's LDS reads are done
    if (lane_in) {
      static_for<0, Q>([&](auto qc) {
        constexpr int q = qc;
        constexpr int cx = kCX[q], cy = kCY[q], cz = kCZ[q];
        const int trow = wrapi(i + cx, p.nx) * p.ny + wrapi(j + cy, p.ny);
        T t[V];
#pragma unroll
        for (int v = 0; v < V; ++v) t[v] = f[v][q];
        if (q != 0 && (nm_any & kNbrMoving)) {
#pragma unroll
          for (int v = 0; v < V; ++v) moving_wall(qc, v, t[v]);
        }
        if constexpr (cz == 0) {
          Vec<S, V>::st(F + (q * N + (size_t)trow * nz) + uk0, t);
        } else {
#pragma unroll
          for (int v = 0; v < V; ++v) buf[zslot<q>() * ncz + k0 + v] = t[v];
        }
      });
    }
    __syncthreads();
    if (lane_in) {
      static_for<1, Q>([&](auto qc) {
        constexpr int q = qc;
        constexpr int cx = kCX[q], cy = kCY[q], cz = kCZ[q];
        if constexpr (cz != 0) {
          const int trow = wrapi(i + cx, p.nx) * p.ny + wrapi(j + cy, p.ny);
          T t[V];
#pragma unroll
          for (int v = 0; v < V; ++v) {
            int ks = k0 + v - cz;  // the cell pushing into (row trow, k0 + v)
            ks = ks < 0 ? ks + nz : (ks >= nz ? ks - nz : ks);
            t[v] = buf[zslot<q>() * ncz + ks];
          }
          Vec<S, V>::st(F + (q * N + (size_t)trow * nz) + uk0, t);
        }
      });
    }
  }
}

// row-tiled kernel with V cells per thread, or nullptr where one access would exceed 16 bytes
template <typename S, int COLL, int COUP, bool ODD, int V>
constexpr void (*pencil_fn())(const KParams) {
  if constexpr (V * sizeof(S) <= 16) {
    return k_step_pencil<S, COLL, COUP, ODD, V>;
  } else {
    return nullptr;
  }
}

}  // namespace lbdem
