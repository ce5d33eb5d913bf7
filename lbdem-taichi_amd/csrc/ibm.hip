// Unresolved LBM-DEM coupling on gfx950.
//
// Particle-equivalence IBM (EqIMBlattice3D): particle-centric spreading of the solid volume
// fraction and solid momentum with the radial 3-point kernel plus wall-mirror images, a
// per-cell finalisation (clamp, solid velocity, fluid velocity at covered cells), and the
// per-particle kernel interpolation + Tenneti drag.  The Tenneti collision weight is
// evaluated inside the collide-stream kernel.
// VANS (Unresolvedlattice3D): normalised spreading of particle volume and fluid force, and
// the kernel-weighted drag.
// Spreading uses FP32 global atomics (order of accumulation varies between runs).
#include "common.hpp"
#include "lbm_kernels.hpp"

namespace lbdem {

struct IbmArgs {
  int nx, ny, nz, np;
  const uint8_t *flags;
  const double *pos, *vel, *rad, *Ffluid;
  double dx, dt, rho, dmin0, dmin1, dmin2;
  double nu_lu;
  float *volfrac, *velsum, *wsum, *velsolid, *fvel, *frho, *hforce;
  double *force_out, *moment_out;
  int32_t *status;
  const uint32_t *key;  // hybrid coupling: coarse-grain keys of the PSC mapping
  const int32_t *order;  // EqIBM: grains visited in this order (a permutation), or NULL
};

// EqIMBlattice3D.threedelta (src/unlbdem/eqlattice.py:451-462)
__device__ __forceinline__ double threedelta(double r) {
  double a = 0.0;
  if (r < 0.5) a = (1.0 + sqrt(-3.0 * r * r + 1.0)) / 3.0;
  else if (0.5 <= r && r <= 1.5) a = (5.0 - 3.0 * r - sqrt(-3.0 * (1.0 - r) * (1.0 - r) + 1.0)) / 6.0;
  return a;
}

// threedelta(sqrt(r2)) with the support tested on r2 first: for r2 > 2.25 the rounded square
// root is at least 1.5, where threedelta is 0 (at exactly 1.5 its second branch gives
// (0.5 - sqrt(0.25)) / 6 = 0), so the value is the same and the square roots are skipped
__device__ __forceinline__ double threedelta_sq(double r2) { return r2 > 2.25 ? 0.0 : threedelta(sqrt(r2)); }

// _kernel_with_mirror (src/unlbdem/eqlattice.py:468-528)
__device__ __forceinline__ double kernel_mirror(const IbmArgs &a, double xc, double yc, double zc, int i, int j, int k) {
  const double s = 1.5;
  auto d2 = [&](double x, double y, double z) { return (x - i) * (x - i) + (y - j) * (y - j) + (z - k) * (z - k); };
  double w = threedelta_sq(d2(xc, yc, zc));
  if (xc < s) w += threedelta_sq(d2(-xc, yc, zc));
  if (xc > (double)(a.nx - 1) - s) w += threedelta_sq(d2(2.0 * (a.nx - 1) - xc, yc, zc));
  if (yc < s) w += threedelta_sq(d2(xc, -yc, zc));
  if (yc > (double)(a.ny - 1) - s) w += threedelta_sq(d2(xc, 2.0 * (a.ny - 1) - yc, zc));
  if (zc < s) w += threedelta_sq(d2(xc, yc, -zc));
  if (zc > (double)(a.nz - 1) - s) w += threedelta_sq(d2(xc, yc, 2.0 * (a.nz - 1) - zc));
  return w;
}

__device__ __forceinline__ void lattice_pos(const IbmArgs &a, int p, double &xc, double &yc, double &zc) {
  xc = (a.pos[3 * p] - a.dmin0 + 0.5 * a.dx) / a.dx;
  yc = (a.pos[3 * p + 1] - a.dmin1 + 0.5 * a.dx) / a.dx;
  zc = (a.pos[3 * p + 2] - a.dmin2 + 0.5 * a.dx) / a.dx;
}

// support box [max(0, int(c - 1.5)), min(N, int(c + 2.5)))  (eqlattice.py:142-147)
__device__ __forceinline__ void eq_box(double c, int N, int &lo, int &hi) {
  lo = max(0, (int)(c - 1.5));
  hi = min(N, (int)(c + 1.5 + 1));
}

// one wavefront per particle: lane l takes cell l of the particle's box, which spans at most
// 4 cells per axis ([int(c - 1.5), int(c + 2.5)) clamped), so 64 lanes cover it
constexpr int kGrainsPerBlock = 4;  // 256 threads

struct BoxLane {
  bool in;   // the lane's cell lies in the box
  int c;     // its cell index
  int i, j, k;
};

__device__ __forceinline__ BoxLane box_lane(const IbmArgs &a, double xc, double yc, double zc) {
  int i0, i1, j0, j1, k0, k1;
  eq_box(xc, a.nx, i0, i1);
  eq_box(yc, a.ny, j0, j1);
  eq_box(zc, a.nz, k0, k1);
  const int l = threadIdx.x & 63;
  BoxLane b;
  b.i = i0 + (l >> 4);
  b.j = j0 + ((l >> 2) & 3);
  b.k = k0 + (l & 3);
  b.in = b.i < i1 && b.j < j1 && b.k < k1;
  b.c = b.in ? (b.i * a.ny + b.j) * a.nz + b.k : 0;
  return b;
}

__global__ __launch_bounds__(256) void k_fill_zero(float *a, size_t n) {
  for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) a[i] = 0.f;
}

// EqIMBlattice3D.grains2lattice passes 1-2 (eqlattice.py:129-186), one wavefront per particle
// (lane per box cell; the kernel sum is a wave reduction, a tree instead of the reference's
// serial sum: round-off only)
__global__ __launch_bounds__(256) void k_eq_spread(const IbmArgs a) {
  const int p = blockIdx.x * kGrainsPerBlock + (int)(threadIdx.x >> 6);
  if (p >= a.np) return;  // uniform over the wavefront
  double xc, yc, zc;
  lattice_pos(a, p, xc, yc, zc);
  const BoxLane b = box_lane(a, xc, yc, zc);
  double wb = -1.0;
  if (b.in && (a.flags[b.c] & LBDEM_FLAG_ACTIVE)) wb = kernel_mirror(a, xc, yc, zc, b.i, b.j, b.k);
  const double denom = wave_sum(wb >= 0.0 ? wb : 0.0);
  if (denom < 1e-30 || wb < 0.0) return;
  const double pi = 3.14159265358979323846;
  const double R = a.rad[p];
  const double Vp = 4.0 / 3.0 * pi * R * R * R, Vl = a.dx * a.dx * a.dx;
  const double s = a.dt / a.dx;
  const double wn = wb / denom, e = wn * Vp / Vl;
  if (e == 0.0) return;
  const int c = b.c;
  unsafeAtomicAdd(a.volfrac + c, (float)e);
  unsafeAtomicAdd(a.velsum + 3 * c, (float)(a.vel[3 * p] * s * e));
  unsafeAtomicAdd(a.velsum + 3 * c + 1, (float)(a.vel[3 * p + 1] * s * e));
  unsafeAtomicAdd(a.velsum + 3 * c + 2, (float)(a.vel[3 * p + 2] * s * e));
  unsafeAtomicAdd(a.wsum + c, (float)wn);
}

// The same passes over grains in a spatial order (IbmArgs::order), kSpreadGrains per workgroup:
// the workgroup accumulates its grains' contributions in LDS over the union of their boxes and
// adds each covered cell to the global fields once.  In particle order every grain adds ~14
// cells x 5 fields with global atomics (14 million per step for 200,000 grains, 0.46 ms); in a
// spatial order without the LDS stage, concurrent wavefronts hit the same cells and it is
// slower still (0.63 ms).  A workgroup whose union box exceeds kSpreadCells adds directly.
// Each wavefront takes the workgroup's grains w, w + 4, ... one after another; their lattice
// positions, radii and velocities are staged in LDS by the pass that finds the union box, and
// the next grain's box cells' flags are loaded while the current grain is evaluated.
constexpr int kSpreadGrains = 128;  // 256 measured slower (0.31 -> 0.41 ms per update_coupling)
constexpr int kSpreadCells = 1536;
static_assert(kSpreadGrains <= 256, "the union box is built with one thread per grain of a 256-thread workgroup");

__global__ __launch_bounds__(256) void k_eq_spread_ordered(const IbmArgs a) {
  __shared__ float s_acc[5][kSpreadCells];
  __shared__ double s_g[7][kSpreadGrains];  // per grain: lattice position, radius, velocity
  __shared__ int s_lo[3], s_hi[3];
  const int g0 = blockIdx.x * kSpreadGrains;
  const int ng = min(kSpreadGrains, a.np - g0);
  if (threadIdx.x < 3) {
    s_lo[threadIdx.x] = INT_MAX;
    s_hi[threadIdx.x] = INT_MIN;
  }
  __syncthreads();
  if ((int)threadIdx.x < ng) {
    const int p = a.order[g0 + threadIdx.x];
    double c[3];
    lattice_pos(a, p, c[0], c[1], c[2]);
    s_g[0][threadIdx.x] = c[0];
    s_g[1][threadIdx.x] = c[1];
    s_g[2][threadIdx.x] = c[2];
    s_g[3][threadIdx.x] = a.rad[p];
    s_g[4][threadIdx.x] = a.vel[3 * p];
    s_g[5][threadIdx.x] = a.vel[3 * p + 1];
    s_g[6][threadIdx.x] = a.vel[3 * p + 2];
    const int N[3] = {a.nx, a.ny, a.nz};
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      int lo, hi;
      eq_box(c[d], N[d], lo, hi);
      if (hi > lo) {
        atomicMin(&s_lo[d], lo);
        atomicMax(&s_hi[d], hi);
      }
    }
  }
  __syncthreads();
  const int lo0 = s_lo[0], lo1 = s_lo[1], lo2 = s_lo[2];
  const int64_t e0 = (int64_t)s_hi[0] - lo0, e1 = (int64_t)s_hi[1] - lo1, e2 = (int64_t)s_hi[2] - lo2;
  const bool any = e0 > 0 && e1 > 0 && e2 > 0;
  const bool lds = any && e0 * e1 * e2 <= kSpreadCells;
  const int vol = lds ? (int)(e0 * e1 * e2) : 0, ex1 = (int)e1, ex2 = (int)e2;
  for (int l = threadIdx.x; l < vol; l += 256)
#pragma unroll
    for (int m = 0; m < 5; ++m) s_acc[m][l] = 0.f;
  __syncthreads();
  const double pi = 3.14159265358979323846;
  const double Vl = a.dx * a.dx * a.dx, s = a.dt / a.dx;
  if (any) {
    const int w = (int)(threadIdx.x >> 6);
    BoxLane bn{};
    uint8_t fn = 0;
    if (w < ng) {
      bn = box_lane(a, s_g[0][w], s_g[1][w], s_g[2][w]);
      fn = bn.in ? a.flags[bn.c] : 0;
    }
    for (int q = w; q < ng; q += 4) {  // uniform over the wavefront
      const BoxLane b = bn;
      const uint8_t fl = fn;
      if (q + 4 < ng) {
        bn = box_lane(a, s_g[0][q + 4], s_g[1][q + 4], s_g[2][q + 4]);
        fn = bn.in ? a.flags[bn.c] : 0;
      }
      const double xc = s_g[0][q], yc = s_g[1][q], zc = s_g[2][q];
      double wb = -1.0;
      if (b.in && (fl & LBDEM_FLAG_ACTIVE)) wb = kernel_mirror(a, xc, yc, zc, b.i, b.j, b.k);
      const double denom = wave_sum(wb >= 0.0 ? wb : 0.0);
      if (denom < 1e-30 || wb < 0.0) continue;
      const double R = s_g[3][q];
      const double Vp = 4.0 / 3.0 * pi * R * R * R;
      const double wn = wb / denom, e = wn * Vp / Vl;
      if (e == 0.0) continue;
      const float v[5] = {(float)e, (float)(s_g[4][q] * s * e), (float)(s_g[5][q] * s * e),
                          (float)(s_g[6][q] * s * e), (float)wn};
      if (lds) {
        const int l = ((b.i - lo0) * ex1 + (b.j - lo1)) * ex2 + (b.k - lo2);
#pragma unroll
        for (int m = 0; m < 5; ++m) atomicAdd(&s_acc[m][l], v[m]);
      } else {
        const int c = b.c;
        unsafeAtomicAdd(a.volfrac + c, v[0]);
        unsafeAtomicAdd(a.velsum + 3 * c, v[1]);
        unsafeAtomicAdd(a.velsum + 3 * c + 1, v[2]);
        unsafeAtomicAdd(a.velsum + 3 * c + 2, v[3]);
        unsafeAtomicAdd(a.wsum + c, v[4]);
      }
    }
  }
  __syncthreads();
  for (int l = threadIdx.x; l < vol; l += 256) {
    if (s_acc[0][l] == 0.f) continue;  // e > 0 for every contribution: 0 means none
    const int i = lo0 + l / (ex1 * ex2), j = lo1 + (l / ex2) % ex1, k = lo2 + l % ex2;
    const int c = (i * a.ny + j) * a.nz + k;
    unsafeAtomicAdd(a.volfrac + c, s_acc[0][l]);
    unsafeAtomicAdd(a.velsum + 3 * c, s_acc[1][l]);
    unsafeAtomicAdd(a.velsum + 3 * c + 1, s_acc[2][l]);
    unsafeAtomicAdd(a.velsum + 3 * c + 2, s_acc[3][l]);
    unsafeAtomicAdd(a.wsum + c, s_acc[4][l]);
  }
}

// solid-fraction clamp and solid velocity (eqlattice.py:188-194); fluid rho, u of the
// current populations at covered cells (the `vel` that compute_weight and lattice2grains read)
template <typename S>
__global__ __launch_bounds__(256) void k_eq_finalize(const IbmArgs a, const KParams kp) {
  const int c = blockIdx.x * 256 + threadIdx.x;
  if (c >= kp.ncell) return;
  float e = a.volfrac[c];
  if (e >= 1.0f) {
    e = 0.99f;
    a.volfrac[c] = e;
    if (a.status) atomicAdd(a.status, 1);
  }
  float us[3] = {0.f, 0.f, 0.f};
  if (e > 1e-15f) {
#pragma unroll
    for (int m = 0; m < 3; ++m) us[m] = a.velsum[3 * c + m] / e;
  }
#pragma unroll
  for (int m = 0; m < 3; ++m) a.velsolid[3 * c + m] = us[m];
  const uint8_t fl = kp.flags[c];
  if (e > 0.f && (fl & LBDEM_FLAG_ACTIVE) && (fl & LBDEM_FLAG_FLUID)) {
    using T = typename Store<S>::T;
    T f[Q];
    gather<S>(kp, reinterpret_cast<const S *>(kp.f), c, fl, kp.t_odd != 0, f);
    T dr = T(0), jx = T(0), jy = T(0), jz = T(0);
    static_for<0, Q>([&](auto qc) {
      constexpr int q = qc;
      dr += f[q];
      acc<kCX[q]>(jx, f[q]);
      acc<kCY[q]>(jy, f[q]);
      acc<kCZ[q]>(jz, f[q]);
    });
    const T rho = T(kp.rho0) + dr;
    a.frho[c] = (float)rho;
    a.fvel[3 * c] = (float)(jx / rho);
    a.fvel[3 * c + 1] = (float)(jy / rho);
    a.fvel[3 * c + 2] = (float)(jz / rho);
  }
}

// Tenneti drag coefficient (eqlattice.py:423-438, vanslattice.py:625-644); lim 1e-9 / 1e-8
__device__ __forceinline__ double tenneti_cd(double Re, double svf, double lim) {
  double Cd = 0.0;
  if (1.0 - svf > lim) {
    const double om = 1.0 - svf;
    const double Cd0 = 1.0 + 0.15 * pow(Re, 0.687);
    const double A = 5.81 * svf / (om * om * om) + 0.48 * pow(svf, 1.0 / 3.0) / (om * om * om * om);
    const double s3 = svf * svf * svf;
    const double B = s3 * Re * (0.95 + 0.61 * s3 / (om * om));
    Cd = om * (Cd0 / (om * om * om) + A + B);
  }
  return Cd;
}

// EqIMBlattice3D.lattice2grains + compute_drag_force (eqlattice.py:309-444), one thread per
// particle (measured faster than a wavefront per particle with six FP64 wave reductions:
// 88 against 209 us for 200,000 particles)
__global__ __launch_bounds__(256) void k_eq_drag(const IbmArgs a) {
  const int t = blockIdx.x * 256 + threadIdx.x;
  if (t >= a.np) return;
  // in a spatial order neighbouring threads read neighbouring cells (68 against 128 us for
  // 200,000 grains)
  const int p = a.order ? a.order[t] : t;
  double xc, yc, zc;
  lattice_pos(a, p, xc, yc, zc);
  int i0, i1, j0, j1, k0, k1;
  eq_box(xc, a.nx, i0, i1);
  eq_box(yc, a.ny, j0, j1);
  eq_box(zc, a.nz, k0, k1);
  double vw[3] = {0, 0, 0}, es = 0.0, wt = 0.0;
  int nl = 0;
  for (int i = i0; i < i1; ++i)
    for (int j = j0; j < j1; ++j)
      for (int k = k0; k < k1; ++k) {
        const int c = (i * a.ny + j) * a.nz + k;
        if (!(a.flags[c] & LBDEM_FLAG_ACTIVE)) continue;
        const double w = kernel_mirror(a, xc, yc, zc, i, j, k);
        if (w < 0.0) continue;
        if (w != 0.0) {
#pragma unroll
          for (int m = 0; m < 3; ++m) vw[m] += (double)a.fvel[3 * c + m] * w;
        }
        es += 1.0 - (double)a.volfrac[c];
        wt += w;
        nl += 1;
      }
  double uf[3] = {0, 0, 0}, ef = 0.0;
  if (wt > 1e-15) {
#pragma unroll
    for (int m = 0; m < 3; ++m) uf[m] = vw[m] / wt * a.dx / a.dt;
  }
  if (nl > 0) ef = es / (double)nl;
  double F[3] = {0, 0, 0};
  if (1.0 - ef > 1e-15) {
    const double pi = 3.14159265358979323846;
    const double ep = 1.0 - ef, dp = 2.0 * a.rad[p];
    const double mu = a.rho * a.nu_lu * a.dx * a.dx / a.dt;
    double us[3], m2 = 0.0;
#pragma unroll
    for (int m = 0; m < 3; ++m) {
      us[m] = a.vel[3 * p + m] - uf[m];
      m2 += us[m] * us[m];
    }
    const double Re = (1.0 - ep) * a.rho * dp * sqrt(m2) / mu;
    const double Cd = tenneti_cd(Re, ep, 1e-9);
#pragma unroll
    for (int m = 0; m < 3; ++m) F[m] = -3.0 * pi * dp * mu * (1.0 - ep) * Cd * us[m];
  }
#pragma unroll
  for (int m = 0; m < 3; ++m) a.force_out[3 * p + m] = F[m];
}

// Tenneti weight field for inspection (compute_weight, eqlattice.py:200-213)
__global__ __launch_bounds__(256) void k_eq_weight(const IbmArgs a, const KParams kp, float *weight) {
  const int c = blockIdx.x * 256 + threadIdx.x;
  if (c >= kp.ncell) return;
  if (!(kp.flags[c] & LBDEM_FLAG_ACTIVE)) return;
  const float e = a.volfrac[c];
  if (!(e > 0.f)) return;
  float d2 = 0.f;
#pragma unroll
  for (int m = 0; m < 3; ++m) {
    const float d = a.velsolid[3 * c + m] - a.fvel[3 * c + m];
    d2 += d * d;
  }
  weight[c] = tenneti_weight<float>(e, sqrtf(d2), (float)a.nu_lu);
}

// ------------------------------------------------------------------------------------
// Hybrid coupling, fine grains (HybridLattice3D, src/hybridlbdem/hybridlattice.py)
// ------------------------------------------------------------------------------------
// coarse solid fraction of a cell from its PSC key: 1 (full cover) or count / 125
__device__ __forceinline__ double coarse_eps(uint32_t key) {
  if (key == 0u) return 0.0;
  if (key & (1u << 30)) return 1.0;
  return (double)((key >> 23) & 0x7Fu) / 125.0;
}

// map_fine_grains passes 1-3 (hybridlattice.py:284-361), one wavefront per grain with
// 2R <= dx: lane l takes cell l of the grain's box (at most 4 x 4 x 4 cells), the kernel sum and
// the free-fraction correction sum are wave reductions, then each lane spreads into its cell
// (FP32 atomics).  The sums are trees instead of the reference's serial loops (round-off only).
__global__ __launch_bounds__(256) void k_hyb_spread(const IbmArgs a) {
  const int p = blockIdx.x * kGrainsPerBlock + (int)(threadIdx.x >> 6);
  if (p >= a.np) return;  // uniform over the wavefront
  const double R = a.rad[p];
  if (2.0 * R > a.dx) return;  // coarse grain (PSC side)
  double xc, yc, zc;
  lattice_pos(a, p, xc, yc, zc);
  const BoxLane b = box_lane(a, xc, yc, zc);
  // the cells a fine grain may occupy: active and not fully covered by a coarse grain
  double ec = 0.0, wb = 0.0;
  if (b.in && (a.flags[b.c] & LBDEM_FLAG_ACTIVE)) {
    ec = coarse_eps(a.key[b.c]);
    if (ec < 1.0) wb = kernel_mirror(a, xc, yc, zc, b.i, b.j, b.k);
  }
  const bool w_pos = wb > 0.0;
  const double denom = wave_sum(w_pos ? wb : 0.0);
  if (denom < 1e-30) return;
  const double dcorr = wave_sum(w_pos ? (wb / denom) * (1.0 - ec) : 0.0);
  if (dcorr < 1e-30 || !w_pos) return;
  const double pi = 3.14159265358979323846;
  const double Vp = 4.0 / 3.0 * pi * R * R * R, Vl = a.dx * a.dx * a.dx;
  const double wn = wb / denom;
  const double e = wn * (1.0 - ec) / dcorr * Vp / Vl;
  const int c = b.c;
  unsafeAtomicAdd(a.volfrac + c, (float)e);
  unsafeAtomicAdd(a.velsum + 3 * c, (float)(a.vel[3 * p] * a.dt / a.dx * e));
  unsafeAtomicAdd(a.velsum + 3 * c + 1, (float)(a.vel[3 * p + 1] * a.dt / a.dx * e));
  unsafeAtomicAdd(a.velsum + 3 * c + 2, (float)(a.vel[3 * p + 2] * a.dt / a.dx * e));
  unsafeAtomicAdd(a.wsum + c, (float)wn);
}

// map_fine_grains, last loop (hybridlattice.py:363-376): clamp to the free fraction, then
// the solid velocity where the fraction exceeds 1e-10 (elsewhere it keeps its value)
__global__ __launch_bounds__(256) void k_hyb_finalize(const IbmArgs a, int ncell) {
  const int c = blockIdx.x * 256 + threadIdx.x;
  if (c >= ncell) return;
  const double avail = 1.0 - coarse_eps(a.key[c]);
  float e = a.volfrac[c];
  if ((double)e > avail) {
    if ((double)e - avail > 1e-6 && a.status) atomicAdd(a.status, 1);
    e = (float)fmax(0.0, avail - 0.01);
    a.volfrac[c] = e;
  }
  if (e > 1e-10f) {
#pragma unroll
    for (int m = 0; m < 3; ++m) a.velsolid[3 * c + m] = a.velsum[3 * c + m] / e;
  }
}

// _transfer_fine_force + _compute_drag_force (hybridlattice.py:661-740), one thread per grain
// with 2R <= dx; nodes with a positive kernel weight only; total solid fraction fine + coarse
__global__ __launch_bounds__(256) void k_hyb_drag(const IbmArgs a) {
  const int p = blockIdx.x * 256 + threadIdx.x;
  if (p >= a.np) return;
  if (2.0 * a.rad[p] > a.dx) return;
  double xc, yc, zc;
  lattice_pos(a, p, xc, yc, zc);
  int i0, i1, j0, j1, k0, k1;
  eq_box(xc, a.nx, i0, i1);
  eq_box(yc, a.ny, j0, j1);
  eq_box(zc, a.nz, k0, k1);
  double vw[3] = {0, 0, 0}, es = 0.0, wt = 0.0;
  int nl = 0;
  for (int i = i0; i < i1; ++i)
    for (int j = j0; j < j1; ++j)
      for (int k = k0; k < k1; ++k) {
        const int c = (i * a.ny + j) * a.nz + k;
        if (!(a.flags[c] & LBDEM_FLAG_ACTIVE)) continue;
        const double w = kernel_mirror(a, xc, yc, zc, i, j, k);
        if (w <= 0.0) continue;
#pragma unroll
        for (int m = 0; m < 3; ++m) vw[m] += (double)a.fvel[3 * c + m] * w;
        es += 1.0 - (double)a.volfrac[c] - coarse_eps(a.key[c]);
        wt += w;
        nl += 1;
      }
  double uf[3] = {0, 0, 0}, ef = 0.0;
  if (wt > 1e-15) {
#pragma unroll
    for (int m = 0; m < 3; ++m) uf[m] = vw[m] / wt * a.dx / a.dt;
  }
  if (nl > 0) ef = es / (double)nl;
  double F[3] = {0, 0, 0};
  const double ep = 1.0 - ef;
  if (ep > 1e-15) {
    const double pi = 3.14159265358979323846;
    const double dp = 2.0 * a.rad[p];
    const double mu = a.rho * a.nu_lu * a.dx * a.dx / a.dt;
    double us[3], m2 = 0.0;
#pragma unroll
    for (int m = 0; m < 3; ++m) {
      us[m] = a.vel[3 * p + m] - uf[m];
      m2 += us[m] * us[m];
    }
    const double Re = (1.0 - ep) * a.rho * dp * sqrt(m2) / mu;
    const double Cd = tenneti_cd(Re, ep, 1e-9);
#pragma unroll
    for (int m = 0; m < 3; ++m) F[m] = -3.0 * pi * dp * mu * (1.0 - ep) * Cd * us[m];
  }
#pragma unroll
  for (int m = 0; m < 3; ++m) {
    a.force_out[3 * p + m] = F[m];
    a.moment_out[3 * p + m] = 0.0;  // force_fluid / moment_fluid.fill(0.0), :609-610
  }
}

// compute_fine_weights (hybridlattice.py:378-398) for inspection
__global__ __launch_bounds__(256) void k_hyb_weight(const IbmArgs a, const KParams kp, float *weight) {
  const int c = blockIdx.x * 256 + threadIdx.x;
  if (c >= kp.ncell) return;
  if (!(kp.flags[c] & LBDEM_FLAG_ACTIVE)) return;
  const float e = a.volfrac[c];
  const double fvf = 1.0 - (double)e - coarse_eps(a.key[c]);
  if (!(e > 0.f) || !(fvf > 1e-10)) return;
  float d2 = 0.f;
#pragma unroll
  for (int m = 0; m < 3; ++m) {
    const float d = a.velsolid[3 * c + m] - a.fvel[3 * c + m];
    d2 += d * d;
  }
  weight[c] = tenneti_weight<float>(e, sqrtf(d2), (float)a.nu_lu);
}

// ------------------------------------------------------------------------------------
// VANS
// ------------------------------------------------------------------------------------
// box [floor(c - 1.5), ceil(c + 1.5)] clamped to [0, N-1], inclusive (vanslattice.py:152-158)
__device__ __forceinline__ void vans_box(double c, int N, int &lo, int &hi) {
  lo = max(0, (int)floor(c - 1.5));
  hi = min(N - 1, (int)ceil(c + 1.5));
}

// force2lattice (vanslattice.py:489-557) / grains2lattice (:135-196) when Ffluid == NULL
__global__ __launch_bounds__(256) void k_vans_spread(const IbmArgs a) {
  const int p = blockIdx.x * 256 + threadIdx.x;
  if (p >= a.np) return;
  double xc, yc, zc;
  lattice_pos(a, p, xc, yc, zc);
  int i0, i1, j0, j1, k0, k1;
  vans_box(xc, a.nx, i0, i1);
  vans_box(yc, a.ny, j0, j1);
  vans_box(zc, a.nz, k0, k1);
  double ws = 0.0;
  for (int i = i0; i <= i1; ++i)
    for (int j = j0; j <= j1; ++j)
      for (int k = k0; k <= k1; ++k) {
        if (!(a.flags[(i * a.ny + j) * a.nz + k] & LBDEM_FLAG_ACTIVE)) continue;
        const double d = sqrt((xc - i) * (xc - i) + (yc - j) * (yc - j) + (zc - k) * (zc - k));
        if (d <= 1.5) ws += threedelta(d);
      }
  if (!(ws > 1e-10)) return;
  const double inv = 1.0 / ws;
  const double pi = 3.14159265358979323846;
  const double R = a.rad[p];
  const double Vg = 4.0 / 3.0 * pi * R * R * R, Vc = a.dx * a.dx * a.dx;
  const double fs = a.dt * a.dt / (a.rho * a.dx * a.dx * a.dx * a.dx);
  for (int i = i0; i <= i1; ++i)
    for (int j = j0; j <= j1; ++j)
      for (int k = k0; k <= k1; ++k) {
        const int c = (i * a.ny + j) * a.nz + k;
        if (!(a.flags[c] & LBDEM_FLAG_ACTIVE)) continue;
        const double d = sqrt((xc - i) * (xc - i) + (yc - j) * (yc - j) + (zc - k) * (zc - k));
        if (!(d <= 1.5)) continue;
        const double w = threedelta(d);
        if (!(w > 1e-10)) continue;
        const double wn = w * inv;
        unsafeAtomicAdd(a.volfrac + c, (float)(wn * Vg / Vc));
        if (a.Ffluid) {
#pragma unroll
          for (int m = 0; m < 3; ++m) unsafeAtomicAdd(a.hforce + 3 * c + m, (float)(wn * a.Ffluid[3 * p + m] * fs));
        }
      }
}

// compute_hydrodynamic_force + compute_drag_force (vanslattice.py:446-482, 563-651)
__global__ __launch_bounds__(256) void k_vans_drag(const IbmArgs a) {
  const int p = blockIdx.x * 256 + threadIdx.x;
  if (p >= a.np) return;
  double xc, yc, zc;
  lattice_pos(a, p, xc, yc, zc);
  const double r = a.rad[p] / a.dx;
  const int xb = max((int)(xc - r - 2), 0), xe = min((int)(xc + r + 2), a.nx);
  const int yb = max((int)(yc - r - 2), 0), ye = min((int)(yc + r + 2), a.ny);
  const int zb = max((int)(zc - r - 2), 0), ze = min((int)(zc + r + 2), a.nz);
  double svf = 0.0, uf[3] = {0, 0, 0};
  for (int i = xb; i < xe; ++i)
    for (int j = yb; j < ye; ++j)
      for (int k = zb; k < ze; ++k) {
        const int c = (i * a.ny + j) * a.nz + k;
        if (!(a.flags[c] & LBDEM_FLAG_ACTIVE)) continue;
        const double vf = (double)a.volfrac[c];
        if (vf > 0.0) {
          const double w = threedelta(sqrt((xc - i) * (xc - i) + (yc - j) * (yc - j) + (zc - k) * (zc - k)));
          svf += vf * w;
#pragma unroll
          for (int m = 0; m < 3; ++m) uf[m] += (double)a.fvel[3 * c + m] * w;
        }
      }
  const double pi = 3.14159265358979323846;
  const double dp = 2.0 * a.rad[p];
  const double mu = a.rho * a.nu_lu * a.dx * a.dx / a.dt;
  double ur[3], m2 = 0.0;
#pragma unroll
  for (int m = 0; m < 3; ++m) {
    ur[m] = a.dx * uf[m] / a.dt - a.vel[3 * p + m];
    m2 += ur[m] * ur[m];
  }
  const double Re = (1.0 - svf) * a.rho * dp * sqrt(m2) / mu;
  const double Cd = tenneti_cd(Re, svf, 1e-8);
#pragma unroll
  for (int m = 0; m < 3; ++m) a.force_out[3 * p + m] = 3.0 * pi * dp * mu * (1.0 - svf) * Cd * ur[m];
}

template <typename S>
static const void *finalize_for() {
  return reinterpret_cast<const void *>(&k_eq_finalize<S>);
}

}  // namespace lbdem

using namespace lbdem;

namespace {
int fill_ibm(const lbdem_lattice *lat, const uint8_t *flags, const lbdem_grains *g, const lbdem_units *u,
             double nu_lu, IbmArgs &a, const char *who) {
  LBDEM_REQUIRE(lat && flags && g && u, "%s: null argument", who);
  LBDEM_REQUIRE(g->n >= 0, "%s: negative particle count", who);
  LBDEM_REQUIRE(u->dx > 0 && u->dt > 0, "%s: dx and dt must be positive", who);
  a = IbmArgs{};
  a.nx = lat->nx;
  a.ny = lat->ny;
  a.nz = lat->nz;
  a.np = g->n;
  a.flags = flags;
  a.pos = g->position;
  a.vel = g->velocity;
  a.rad = g->radius;
  a.dx = u->dx;
  a.dt = u->dt;
  a.rho = u->rho;
  a.dmin0 = u->dmin[0];
  a.dmin1 = u->dmin[1];
  a.dmin2 = u->dmin[2];
  a.nu_lu = nu_lu;
  a.order = g->order;
  return LBDEM_OK;
}
unsigned nblk(int64_t n) { return (unsigned)((n + 255) / 256); }
}  // namespace

extern "C" {

int lbdem_ibm_spread(const lbdem_lattice *lat, const lbdem_step_args *st, const lbdem_grains *g,
                     const lbdem_units *u, float *volfrac, float *velsum, float *wsum, float *velsolid, float *rho,
                     float *vel, int32_t *status, void *stream) {
  IbmArgs a;
  int rc = fill_ibm(lat, st ? st->flags : nullptr, g, u, 0.0, a, "lbdem_ibm_spread");
  if (rc) return rc;
  LBDEM_REQUIRE(volfrac && velsum && wsum && velsolid && rho && vel, "lbdem_ibm_spread: null output");
  LBDEM_REQUIRE(g->n == 0 || (g->position && g->velocity && g->radius), "lbdem_ibm_spread: grain arrays");
  const int64_t n = (int64_t)lat->nx * lat->ny * lat->nz;
  a.volfrac = volfrac;
  a.velsum = velsum;
  a.wsum = wsum;
  a.velsolid = velsolid;
  a.frho = rho;
  a.fvel = vel;
  a.status = status;
  hipStream_t s = as_stream(stream);
  hipLaunchKernelGGL(k_fill_zero, dim3(2048), dim3(256), 0, s, volfrac, (size_t)n);
  hipLaunchKernelGGL(k_fill_zero, dim3(2048), dim3(256), 0, s, velsum, (size_t)(3 * n));
  hipLaunchKernelGGL(k_fill_zero, dim3(2048), dim3(256), 0, s, wsum, (size_t)n);
  if (g->n > 0 && a.order)
    hipLaunchKernelGGL(k_eq_spread_ordered, dim3((g->n + kSpreadGrains - 1) / kSpreadGrains), dim3(256), 0, s, a);
  else if (g->n > 0)
    hipLaunchKernelGGL(k_eq_spread, dim3((g->n + kGrainsPerBlock - 1) / kGrainsPerBlock), dim3(256), 0, s, a);
  LBDEM_CHECK_LAUNCH("lbdem_ibm_spread");
  // finalize reads the populations for the fluid velocity at covered cells
  KParams kp{};
  kp.nx = lat->nx;
  kp.ny = lat->ny;
  kp.nz = lat->nz;
  kp.ncell = (int)n;
  kp.f = st->f;
  kp.flags = st->flags;
  kp.nmask = st->nmask;
  kp.rho0 = lat->rho0;
  kp.t_odd = (int)(st->t & 1);
  LBDEM_REQUIRE(st->f && st->nmask, "lbdem_ibm_spread: populations and nmask are required");
  void *args[] = {&a, &kp};
  const void *fn = nullptr;
  switch (lat->storage) {
    case LBDEM_STORE_F32: fn = finalize_for<float>(); break;
    case LBDEM_STORE_F16: fn = finalize_for<_Float16>(); break;
    case LBDEM_STORE_BF16: fn = finalize_for<__bf16>(); break;
    case LBDEM_STORE_F64: fn = finalize_for<double>(); break;
    default: LBDEM_REQUIRE(false, "lbdem_ibm_spread: bad storage %d", lat->storage);
  }
  hipError_t e = hipLaunchKernel(fn, dim3(nblk(n)), dim3(256), args, 0, s);
  if (e != hipSuccess) {
    set_error("lbdem_ibm_spread: launch failed: %s", hipGetErrorString(e));
    return LBDEM_E_HIP;
  }
  return LBDEM_OK;
}

int lbdem_ibm_drag(const lbdem_lattice *lat, const uint8_t *flags, const lbdem_grains *g, const lbdem_units *u,
                   const float *volfrac, const float *vel, void *stream) {
  IbmArgs a;
  int rc = fill_ibm(lat, flags, g, u, (1.0 / lat->omega0 - 0.5) / 3.0, a, "lbdem_ibm_drag");
  if (rc) return rc;
  LBDEM_REQUIRE(volfrac && vel && (g->n == 0 || g->force_fluid), "lbdem_ibm_drag: null argument");
  a.volfrac = const_cast<float *>(volfrac);
  a.fvel = const_cast<float *>(vel);
  a.force_out = g->force_fluid;
  if (g->n > 0) hipLaunchKernelGGL(k_eq_drag, dim3(nblk(g->n)), dim3(256), 0, as_stream(stream), a);
  LBDEM_CHECK_LAUNCH("lbdem_ibm_drag");
  return LBDEM_OK;
}

int lbdem_ibm_weight(const lbdem_lattice *lat, const uint8_t *flags, const float *volfrac, const float *velsolid,
                     const float *vel, float *weight, void *stream) {
  LBDEM_REQUIRE(lat && flags && volfrac && velsolid && vel && weight, "lbdem_ibm_weight: null argument");
  IbmArgs a{};
  a.volfrac = const_cast<float *>(volfrac);
  a.velsolid = const_cast<float *>(velsolid);
  a.fvel = const_cast<float *>(vel);
  a.nu_lu = (1.0 / lat->omega0 - 0.5) / 3.0;
  KParams kp{};
  kp.ncell = lat->nx * lat->ny * lat->nz;
  kp.flags = flags;
  hipLaunchKernelGGL(k_eq_weight, dim3(nblk(kp.ncell)), dim3(256), 0, as_stream(stream), a, kp, weight);
  LBDEM_CHECK_LAUNCH("lbdem_ibm_weight");
  return LBDEM_OK;
}

int lbdem_vans_spread(const lbdem_lattice *lat, const uint8_t *flags, const lbdem_grains *g, const lbdem_units *u,
                      float *volfrac, float *hforce, void *stream) {
  IbmArgs a;
  int rc = fill_ibm(lat, flags, g, u, 0.0, a, "lbdem_vans_spread");
  if (rc) return rc;
  LBDEM_REQUIRE(volfrac, "lbdem_vans_spread: volfrac is required");
  const int64_t n = (int64_t)lat->nx * lat->ny * lat->nz;
  a.volfrac = volfrac;
  a.hforce = hforce;
  a.Ffluid = hforce ? g->force_fluid : nullptr;
  LBDEM_REQUIRE(!hforce || g->n == 0 || g->force_fluid, "lbdem_vans_spread: force_fluid is required");
  hipStream_t s = as_stream(stream);
  hipLaunchKernelGGL(k_fill_zero, dim3(2048), dim3(256), 0, s, volfrac, (size_t)n);
  if (hforce) hipLaunchKernelGGL(k_fill_zero, dim3(2048), dim3(256), 0, s, hforce, (size_t)(3 * n));
  if (g->n > 0) hipLaunchKernelGGL(k_vans_spread, dim3(nblk(g->n)), dim3(256), 0, s, a);
  LBDEM_CHECK_LAUNCH("lbdem_vans_spread");
  return LBDEM_OK;
}

int lbdem_vans_drag(const lbdem_lattice *lat, const uint8_t *flags, const lbdem_grains *g, const lbdem_units *u,
                    const float *volfrac, const float *vel, void *stream) {
  IbmArgs a;
  int rc = fill_ibm(lat, flags, g, u, (1.0 / lat->omega0 - 0.5) / 3.0, a, "lbdem_vans_drag");
  if (rc) return rc;
  LBDEM_REQUIRE(volfrac && vel && (g->n == 0 || g->force_fluid), "lbdem_vans_drag: null argument");
  a.volfrac = const_cast<float *>(volfrac);
  a.fvel = const_cast<float *>(vel);
  a.force_out = g->force_fluid;
  if (g->n > 0) hipLaunchKernelGGL(k_vans_drag, dim3(nblk(g->n)), dim3(256), 0, as_stream(stream), a);
  LBDEM_CHECK_LAUNCH("lbdem_vans_drag");
  return LBDEM_OK;
}

int lbdem_hybrid_map_fine(const lbdem_lattice *lat, const uint8_t *flags, const lbdem_grains *g,
                          const lbdem_units *u, const uint32_t *key, float *fine_volfrac, float *velsum,
                          float *wsum, float *fine_velsolid, int32_t *status, void *stream) {
  IbmArgs a;
  int rc = fill_ibm(lat, flags, g, u, 0.0, a, "lbdem_hybrid_map_fine");
  if (rc) return rc;
  LBDEM_REQUIRE(key && fine_volfrac && velsum && wsum && fine_velsolid, "lbdem_hybrid_map_fine: null argument");
  LBDEM_REQUIRE(g->n == 0 || (g->position && g->velocity && g->radius), "lbdem_hybrid_map_fine: grain arrays");
  const int64_t n = (int64_t)lat->nx * lat->ny * lat->nz;
  a.key = key;
  a.volfrac = fine_volfrac;
  a.velsum = velsum;
  a.wsum = wsum;
  a.velsolid = fine_velsolid;
  a.status = status;
  hipStream_t s = as_stream(stream);
  hipLaunchKernelGGL(k_fill_zero, dim3(2048), dim3(256), 0, s, fine_volfrac, (size_t)n);
  hipLaunchKernelGGL(k_fill_zero, dim3(2048), dim3(256), 0, s, velsum, (size_t)(3 * n));
  hipLaunchKernelGGL(k_fill_zero, dim3(2048), dim3(256), 0, s, wsum, (size_t)n);
  if (g->n > 0)
    hipLaunchKernelGGL(k_hyb_spread, dim3((g->n + kGrainsPerBlock - 1) / kGrainsPerBlock), dim3(256), 0, s, a);
  hipLaunchKernelGGL(k_hyb_finalize, dim3(nblk(n)), dim3(256), 0, s, a, (int)n);
  LBDEM_CHECK_LAUNCH("lbdem_hybrid_map_fine");
  return LBDEM_OK;
}

int lbdem_hybrid_forces(const lbdem_lattice *lat, const uint8_t *flags, const lbdem_grains *g, const lbdem_units *u,
                        const int32_t *coarse, int32_t ncoarse, const uint32_t *key, const float *cell_force,
                        uint32_t *stale_key, const int32_t *stale_bbox, const float *fine_volfrac, const float *vel,
                        void *stream) {
  IbmArgs a;
  int rc = fill_ibm(lat, flags, g, u, (1.0 / lat->omega0 - 0.5) / 3.0, a, "lbdem_hybrid_forces");
  if (rc) return rc;
  LBDEM_REQUIRE(coarse || ncoarse == 0, "lbdem_hybrid_forces: coarse grain list is required");
  LBDEM_REQUIRE(key && cell_force && fine_volfrac && vel, "lbdem_hybrid_forces: null argument");
  LBDEM_REQUIRE(g->n == 0 || (g->position && g->velocity && g->radius && g->force_fluid && g->moment_fluid),
                "lbdem_hybrid_forces: grain arrays");
  rc = launch_psc_forces(lat, g, u, coarse, ncoarse, key, cell_force, stale_key, stale_bbox, stream,
                         "lbdem_hybrid_forces");
  if (rc) return rc;
  a.key = key;
  a.volfrac = const_cast<float *>(fine_volfrac);
  a.fvel = const_cast<float *>(vel);
  a.force_out = g->force_fluid;
  a.moment_out = g->moment_fluid;
  if (g->n > 0) hipLaunchKernelGGL(k_hyb_drag, dim3(nblk(g->n)), dim3(256), 0, as_stream(stream), a);
  LBDEM_CHECK_LAUNCH("lbdem_hybrid_forces");
  return LBDEM_OK;
}

int lbdem_hybrid_fine_weight(const lbdem_lattice *lat, const uint8_t *flags, const uint32_t *key,
                             const float *fine_volfrac, const float *fine_velsolid, const float *vel, float *weight,
                             void *stream) {
  LBDEM_REQUIRE(lat && flags && key && fine_volfrac && fine_velsolid && vel && weight,
                "lbdem_hybrid_fine_weight: null argument");
  IbmArgs a{};
  a.key = key;
  a.volfrac = const_cast<float *>(fine_volfrac);
  a.velsolid = const_cast<float *>(fine_velsolid);
  a.fvel = const_cast<float *>(vel);
  a.nu_lu = (1.0 / lat->omega0 - 0.5) / 3.0;
  KParams kp{};
  kp.ncell = lat->nx * lat->ny * lat->nz;
  kp.flags = flags;
  hipLaunchKernelGGL(k_hyb_weight, dim3(nblk(kp.ncell)), dim3(256), 0, as_stream(stream), a, kp, weight);
  LBDEM_CHECK_LAUNCH("lbdem_hybrid_fine_weight");
  return LBDEM_OK;
}

}  // extern "C"
