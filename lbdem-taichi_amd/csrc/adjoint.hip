// Differentiable PSC channel: forward step and its hand-written adjoint (gfx950).
//
// State f^n (19 x ncell, SoA, unshifted), two arrays.  One step f^{n+1} = S(C(f^n)):
//   C: per-cell PSC collision with fixed grains (solid velocity 0) and a body force
//      g = f + B Omega_s + (1 - B) Omega_f + (1 - B) 3 rho w (c.F),
//      Omega_s = f_qbar - feq_qbar + w rho - f_q, Omega_f = -omega (f - feq),
//      B = eps (tau - 1/2) / ((1 - eps) + (tau - 1/2)), omega = 1/tau
//      (PSCLattice3D.collide_solid, src/lbdem/psclattice3d.py:240-261, with the forcing term of
//      ForcedLattice3D.collide, src/lbm3d/lbm_foecelover.py:53-68, applied to the fluid part);
//   S: streaming with halfway bounce-back at wall cells and periodic wrap
//      (BasicLattice3D.stream, src/lbm3d/lbm_solver3d.py:203-224).
// Adjoint: lambda^n = C'(f^n)^T S^T lambda^{n+1}; S^T gathers lambda_p(x + c_p), or
// lambda_pbar(x) where x + c_p is a wall.  The step also accumulates dL/d eps (per cell),
// dL/d tau and dL/d F.
#include "common.hpp"
#include "lbm_kernels.hpp"

namespace lbdem {

struct DiffArgs {
  int nx, ny, nz, ncell;
  const uint8_t *flags;
  const void *f_in;
  void *f_out;
  const void *lam_in;
  void *lam_out;
  const void *eps;   // per-cell solid fraction (T)
  void *g_eps;       // per-cell dL/d eps accumulator (T)
  double *g_red;     // 4 doubles per workgroup: dL/dtau, dL/dF (3), accumulated over steps
  double tau, Fx, Fy, Fz;
  int plane;         // outlet plane index along x
  const void *g_u;   // dL/d u_x on the plane (ny*nz, T)
  void *u_plane;     // u_x on the plane (ny*nz, T)
};

// Periodic neighbour index x + c from three wrapped offsets per axis (computed once per cell;
// each direction then costs two additions instead of three wraps and a multiply-add).
struct NbrIdx {
  int x[3], y[3], z[3];
  __device__ __forceinline__ NbrIdx(const DiffArgs &a, int c) {
    const int nyz = a.ny * a.nz;
    const int i = c / nyz, j = (c - i * nyz) / a.nz, k = c - i * nyz - j * a.nz;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      x[d] = wrapi(i + d - 1, a.nx) * nyz;
      y[d] = wrapi(j + d - 1, a.ny) * a.nz;
      z[d] = wrapi(k + d - 1, a.nz);
    }
  }
  template <int cx, int cy, int cz>
  __device__ __forceinline__ int at() const {
    return x[cx + 1] + y[cy + 1] + z[cz + 1];
  }
};

template <typename T>
__device__ __forceinline__ void psc_collide_fixed(const T (&f)[Q], T eps, T tau, T Fx, T Fy, T Fz, T (&g)[Q],
                                                  T &rho, T &ux, T &uy, T &uz, T &B) {
  T dr = T(0), jx = T(0), jy = T(0), jz = T(0);
  static_for<0, Q>([&](auto qc) {
    constexpr int q = qc;
    dr += f[q];
    acc<kCX[q]>(jx, f[q]);
    acc<kCY[q]>(jy, f[q]);
    acc<kCZ[q]>(jz, f[q]);
  });
  rho = dr;
  ux = jx / rho;
  uy = jy / rho;
  uz = jz / rho;
  const T omega = T(1) / tau, h = tau - T(0.5);
  B = eps * h / ((T(1) - eps) + h);
  const T usq = ux * ux + uy * uy + uz * uz;
  T feq[Q];
  static_for<0, Q>([&](auto qc) {
    constexpr int q = qc;
    const T cu = cdot<q>(ux, uy, uz);
    feq[q] = T(kW[q]) * rho * (T(1) + T(3) * cu + T(4.5) * cu * cu - T(1.5) * usq);
  });
  static_for<0, Q>([&](auto qc) {
    constexpr int q = qc;
    constexpr int r = kOPP[q];
    const T w = T(kW[q]);
    const T Os = f[r] - feq[r] + w * rho - f[q];
    const T Of = -omega * (f[q] - feq[q]);
    g[q] = f[q] + B * Os + (T(1) - B) * Of + (T(1) - B) * T(3) * rho * w * cdot<q>(Fx, Fy, Fz);
  });
}

template <typename T>
__global__ __launch_bounds__(256) void k_diff_step(const DiffArgs a) {
  const int c = blockIdx.x * 256 + threadIdx.x;
  if (c >= a.ncell) return;
  const uint8_t fl = a.flags[c];
  if (!(fl & LBDEM_FLAG_ACTIVE)) return;
  const size_t N = (size_t)a.ncell;
  const T *F = reinterpret_cast<const T *>(a.f_in);
  T *G = reinterpret_cast<T *>(a.f_out);
  T f[Q], g[Q];
  static_for<0, Q>([&](auto qc) {
    constexpr int q = qc;
    f[q] = F[q * N + c];
  });
  T rho, ux, uy, uz, B;
  psc_collide_fixed<T>(f, reinterpret_cast<const T *>(a.eps)[c], T(a.tau), T(a.Fx), T(a.Fy), T(a.Fz), g, rho, ux,
                       uy, uz, B);
  const NbrIdx nb(a, c);
  const bool near = fl & LBDEM_FLAG_NEAR;  // a wall among the neighbours (bounce-back)
  static_for<0, Q>([&](auto qc) {
    constexpr int q = qc;
    const int m = nb.at<kCX[q], kCY[q], kCZ[q]>();
    if (q != 0 && near && (a.flags[m] & kMaskSolid)) G[kOPP[q] * N + c] = g[q];  // bounce back
    else G[q * N + m] = g[q];
  });
}

// Adjoint step.  Live state per cell: f and mu (2 x 19 values); feq and c.u are recomputed where
// they are used instead of being kept, so the kernel stays below 128 VGPRs in FP32.
template <typename T>
__global__ __launch_bounds__(256) void k_diff_adjoint(const DiffArgs a) {
  const int c = blockIdx.x * 256 + threadIdx.x;
  double gt = 0.0, gfx = 0.0, gfy = 0.0, gfz = 0.0;
  const size_t N = (size_t)a.ncell;
  const uint8_t fl = c < a.ncell ? a.flags[c] : (uint8_t)0;
  if (fl & LBDEM_FLAG_ACTIVE) {
    const T *F = reinterpret_cast<const T *>(a.f_in);
    const T *L = reinterpret_cast<const T *>(a.lam_in);
    T *Lo = reinterpret_cast<T *>(a.lam_out);
    const NbrIdx nb(a, c);
    const bool near = fl & LBDEM_FLAG_NEAR;  // a wall among the neighbours (bounce-back)
    // mu = S^T lambda^{n+1}
    T mu[Q], f[Q];
    static_for<0, Q>([&](auto qc) {
      constexpr int q = qc;
      f[q] = F[q * N + c];
      const int m = nb.at<kCX[q], kCY[q], kCZ[q]>();
      if (q != 0 && near && (a.flags[m] & kMaskSolid)) mu[q] = L[kOPP[q] * N + c];
      else mu[q] = L[q * N + m];
    });
    T rho = T(0), jx = T(0), jy = T(0), jz = T(0);
    static_for<0, Q>([&](auto qc) {
      constexpr int q = qc;
      rho += f[q];
      acc<kCX[q]>(jx, f[q]);
      acc<kCY[q]>(jy, f[q]);
      acc<kCZ[q]>(jz, f[q]);
    });
    const T inv = recip(rho);
    const T ux = jx * inv, uy = jy * inv, uz = jz * inv;
    const T eps = reinterpret_cast<const T *>(a.eps)[c];
    const T tau = T(a.tau), h = tau - T(0.5), omega = T(1) / tau;
    const T B = eps * h / ((T(1) - eps) + h);
    const T Fx = T(a.Fx), Fy = T(a.Fy), Fz = T(a.Fz);
    const T usq = ux * ux + uy * uy + uz * uz;
    const T oB = T(1) - B;
    const T ca = T(1) - B - oB * omega;
    const T oBw = oB * omega;
    // nu = (dg/df)^T mu with g = a f + B P f + [(1-B) omega I - B P] feq + B w rho + (1-B) 3 rho w (c.F);
    // J_feq^T xi = s0 + sum_a G_a (c_ka - u_a) with xi_q = (1-B) omega mu_q - B mu_qbar.
    // Opposite pairs (q, qbar) share the even part E of feq (feq_q = E + O, feq_qbar = E - O).
    T sw, sF = T(0), s0, X, Gx = T(0), Gy = T(0), Gz = T(0);  // X = sum_q w xi_q
    T dB, dOm, dFx = T(0), dFy = T(0), dFz = T(0);
    {
      const T w = T(kW[0]);
      const T fe = w * rho * (T(1) - T(1.5) * usq);
      const T xi = (oBw - B) * mu[0];
      sw = w * mu[0];
      s0 = xi * fe;
      X = w * xi;
      dB = mu[0] * ((w * rho - fe) + omega * (f[0] - fe));
      dOm = mu[0] * (-(oB) * (f[0] - fe));
    }
    static_for<1, Q>([&](auto qc) {
      constexpr int q = qc;
      if constexpr (kOPP[q] > q) {
        constexpr int r = kOPP[q];
        const T w = T(kW[q]);
        const T cu = cdot<q>(ux, uy, uz);
        const T cF = cdot<q>(Fx, Fy, Fz);
        const T E = w * rho * (T(1) + T(4.5) * cu * cu - T(1.5) * usq), O = T(3) * w * rho * cu;
        const T fq = E + O, fr = E - O;
        const T xq = oBw * mu[q] - B * mu[r], xr = oBw * mu[r] - B * mu[q];
        const T xs = xq + xr, xd = xq - xr;
        sw += w * (mu[q] + mu[r]);
        sF += w * cF * (mu[q] - mu[r]);
        s0 += xs * E + xd * O;
        X += w * xs;
        // G_a += w c_qa (3 (xi_q - xi_qbar) + 9 (c.u) (xi_q + xi_qbar))
        const T gq = w * (T(3) * xd + T(9) * cu * xs);
        acc<kCX[q]>(Gx, gq);
        acc<kCY[q]>(Gy, gq);
        acc<kCZ[q]>(Gz, gq);
        // dg/dB = (P f - P feq + w rho - f) + omega (f - feq) - 3 rho w (c.F), for q and qbar
        const T wr = w * rho, t3 = T(3) * wr * cF;
        dB += mu[q] * ((f[r] - fr + wr - f[q]) + omega * (f[q] - fq) - t3) +
              mu[r] * ((f[q] - fq + wr - f[r]) + omega * (f[r] - fr) + t3);
        dOm += -(oB) * (mu[q] * (f[q] - fq) + mu[r] * (f[r] - fr));
        const T sfd = oB * T(3) * wr * (mu[q] - mu[r]);
        acc<kCX[q]>(dFx, sfd);
        acc<kCY[q]>(dFy, sfd);
        acc<kCZ[q]>(dFz, sfd);
      }
    });
    Gx -= T(3) * ux * X;
    Gy -= T(3) * uy * X;
    Gz -= T(3) * uz * X;
    s0 *= inv;
    const T common = s0 - (Gx * ux + Gy * uy + Gz * uz) + B * sw + oB * T(3) * sF;
    static_for<0, Q>([&](auto qc) {
      constexpr int q = qc;
      T nu = ca * mu[q] + B * mu[kOPP[q]] + common;
      acc<kCX[q]>(nu, Gx);
      acc<kCY[q]>(nu, Gy);
      acc<kCZ[q]>(nu, Gz);
      Lo[q * N + c] = nu;
    });
    const T den = (T(1) - eps) + h;
    const T dBde = h * (T(1) + h) / (den * den);
    const T dBdt = eps * (T(1) - eps) / (den * den);
    reinterpret_cast<T *>(a.g_eps)[c] += dB * dBde;
    gt = (double)(dB * dBdt) + (double)dOm * (-1.0 / ((double)tau * (double)tau));
    gfx = (double)dFx;
    gfy = (double)dFy;
    gfz = (double)dFz;
  }
  // block sums in a fixed order (one barrier), added to this workgroup's own accumulator slot:
  // no atomics (same-address FP64 atomics from every workgroup cost a third of the step) and the
  // final sum, lbdem_diff_reduce, runs in a fixed order: the gradients are reproducible
  __shared__ double sh[4][4];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const double v0 = wave_sum(gt), v1 = wave_sum(gfx), v2 = wave_sum(gfy), v3 = wave_sum(gfz);
  if (lane == 0) {
    sh[0][wv] = v0;
    sh[1][wv] = v1;
    sh[2][wv] = v2;
    sh[3][wv] = v3;
  }
  __syncthreads();
  if (threadIdx.x < 4) {
    const double r = (sh[threadIdx.x][0] + sh[threadIdx.x][1]) + (sh[threadIdx.x][2] + sh[threadIdx.x][3]);
    a.g_red[4 * (size_t)blockIdx.x + threadIdx.x] += r;
  }
}

// out[m] = sum over workgroups b of part[4 b + m], fixed order (one workgroup)
__global__ __launch_bounds__(256) void k_diff_reduce(const double *part, int nblocks, double *out) {
  __shared__ double sh[256][4];
  double s[4] = {0, 0, 0, 0};
  for (int b = threadIdx.x; b < nblocks; b += 256)
#pragma unroll
    for (int m = 0; m < 4; ++m) s[m] += part[4 * (size_t)b + m];
#pragma unroll
  for (int m = 0; m < 4; ++m) sh[threadIdx.x][m] = s[m];
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w)
#pragma unroll
      for (int m = 0; m < 4; ++m) sh[threadIdx.x][m] += sh[threadIdx.x + w][m];
    __syncthreads();
  }
  if (threadIdx.x < 4) out[threadIdx.x] = sh[0][threadIdx.x];
}

// u_x on the plane i = plane
template <typename T>
__global__ __launch_bounds__(256) void k_diff_outlet(const DiffArgs a) {
  const int t = blockIdx.x * 256 + threadIdx.x;
  const int nyz = a.ny * a.nz;
  if (t >= nyz) return;
  const int c = a.plane * nyz + t;
  T u = T(0);
  if (a.flags[c] & LBDEM_FLAG_ACTIVE) {
    const T *F = reinterpret_cast<const T *>(a.f_in);
    const size_t N = (size_t)a.ncell;
    T rho = T(0), jx = T(0);
    static_for<0, Q>([&](auto qc) {
      constexpr int q = qc;
      const T v = F[q * N + c];
      rho += v;
      acc<kCX[q]>(jx, v);
    });
    u = jx / rho;
  }
  reinterpret_cast<T *>(a.u_plane)[t] = u;
}

// lambda^T = dL/df^T from dL/du_x on the plane: dL/df_k = g (c_kx - u_x) / rho; zero elsewhere
template <typename T>
__global__ __launch_bounds__(256) void k_diff_seed(const DiffArgs a) {
  const int c = blockIdx.x * 256 + threadIdx.x;
  if (c >= a.ncell) return;
  const size_t N = (size_t)a.ncell;
  T *Lo = reinterpret_cast<T *>(a.lam_out);
  const int nyz = a.ny * a.nz;
  const int i = c / nyz;
  T g = T(0), rho = T(1), ux = T(0);
  if (i == a.plane && (a.flags[c] & LBDEM_FLAG_ACTIVE)) {
    g = reinterpret_cast<const T *>(a.g_u)[c - i * nyz];
    const T *F = reinterpret_cast<const T *>(a.f_in);
    T jx = T(0);
    rho = T(0);
    static_for<0, Q>([&](auto qc) {
      constexpr int q = qc;
      const T v = F[q * N + c];
      rho += v;
      acc<kCX[q]>(jx, v);
    });
    ux = jx / rho;
  }
  static_for<0, Q>([&](auto qc) {
    constexpr int q = qc;
    Lo[q * N + c] = g * (T(kCX[q]) - ux) / rho;
  });
}

template <typename T>
__global__ __launch_bounds__(256) void k_diff_init(const DiffArgs a) {
  const int c = blockIdx.x * 256 + threadIdx.x;
  if (c >= a.ncell) return;
  T *G = reinterpret_cast<T *>(a.f_out);
  const size_t N = (size_t)a.ncell;
  const bool act = a.flags[c] & LBDEM_FLAG_ACTIVE;
  static_for<0, Q>([&](auto qc) {
    constexpr int q = qc;
    G[q * N + c] = act ? T(kW[q]) : T(0);  // rest state rho = 1, u = 0
  });
}

static int fill_diff(const lbdem_diff_args *d, DiffArgs &a, const char *who) {
  LBDEM_REQUIRE(d && d->flags, "%s: null argument", who);
  LBDEM_REQUIRE(d->nx > 0 && d->ny > 0 && d->nz > 0, "%s: bad extents", who);
  LBDEM_REQUIRE(d->precision == 32 || d->precision == 64, "%s: precision must be 32 or 64", who);
  LBDEM_REQUIRE(d->tau > 0.5, "%s: tau must exceed 1/2", who);
  a = DiffArgs{};
  a.nx = d->nx;
  a.ny = d->ny;
  a.nz = d->nz;
  a.ncell = d->nx * d->ny * d->nz;
  a.flags = d->flags;
  a.tau = d->tau;
  a.Fx = d->force[0];
  a.Fy = d->force[1];
  a.Fz = d->force[2];
  a.eps = d->eps;
  a.plane = d->plane;
  return LBDEM_OK;
}

static int launch_diff(const void *fn, unsigned blocks, DiffArgs &a, void *stream, const char *who) {
  void *args[] = {&a};
  hipError_t e = hipLaunchKernel(fn, dim3(blocks), dim3(256), args, 0, as_stream(stream));
  if (e != hipSuccess) {
    set_error("%s: launch failed: %s", who, hipGetErrorString(e));
    return LBDEM_E_HIP;
  }
  return LBDEM_OK;
}

}  // namespace lbdem

using namespace lbdem;

#define LBDEM_DIFF_FN(kernel, prec) \
  ((prec) == 64 ? reinterpret_cast<const void *>(&kernel<double>) : reinterpret_cast<const void *>(&kernel<float>))

extern "C" {

int lbdem_diff_init(const lbdem_diff_args *d, void *f, void *stream) {
  DiffArgs a;
  int rc = fill_diff(d, a, "lbdem_diff_init");
  if (rc) return rc;
  LBDEM_REQUIRE(f, "lbdem_diff_init: f is required");
  a.f_out = f;
  return launch_diff(LBDEM_DIFF_FN(k_diff_init, d->precision), (unsigned)((a.ncell + 255) / 256), a, stream,
                     "lbdem_diff_init");
}

int lbdem_diff_step(const lbdem_diff_args *d, const void *f_in, void *f_out, void *stream) {
  DiffArgs a;
  int rc = fill_diff(d, a, "lbdem_diff_step");
  if (rc) return rc;
  LBDEM_REQUIRE(f_in && f_out && f_in != f_out && d->eps, "lbdem_diff_step: f_in, f_out (distinct) and eps");
  a.f_in = f_in;
  a.f_out = f_out;
  return launch_diff(LBDEM_DIFF_FN(k_diff_step, d->precision), (unsigned)((a.ncell + 255) / 256), a, stream,
                     "lbdem_diff_step");
}

int lbdem_diff_adjoint_step(const lbdem_diff_args *d, const void *f_n, const void *lam_in, void *lam_out,
                            void *g_eps, double *g_red, void *stream) {
  DiffArgs a;
  int rc = fill_diff(d, a, "lbdem_diff_adjoint_step");
  if (rc) return rc;
  LBDEM_REQUIRE(f_n && lam_in && lam_out && lam_in != lam_out && g_eps && g_red && d->eps,
                "lbdem_diff_adjoint_step: null or aliased argument");
  a.f_in = f_n;
  a.lam_in = lam_in;
  a.lam_out = lam_out;
  a.g_eps = g_eps;
  a.g_red = g_red;
  return launch_diff(LBDEM_DIFF_FN(k_diff_adjoint, d->precision), (unsigned)((a.ncell + 255) / 256), a, stream,
                     "lbdem_diff_adjoint_step");
}

int64_t lbdem_diff_red_doubles(const lbdem_diff_args *d) {
  if (!d || d->nx <= 0 || d->ny <= 0 || d->nz <= 0) return 0;
  return 4 * (((int64_t)d->nx * d->ny * d->nz + 255) / 256);
}

int lbdem_diff_reduce(const lbdem_diff_args *d, const double *g_red, double *out, void *stream) {
  DiffArgs a;
  int rc = fill_diff(d, a, "lbdem_diff_reduce");
  if (rc) return rc;
  LBDEM_REQUIRE(g_red && out, "lbdem_diff_reduce: null argument");
  const int nblocks = (a.ncell + 255) / 256;
  hipLaunchKernelGGL(k_diff_reduce, dim3(1), dim3(256), 0, as_stream(stream), g_red, nblocks, out);
  LBDEM_CHECK_LAUNCH("lbdem_diff_reduce");
  return LBDEM_OK;
}

int lbdem_diff_outlet(const lbdem_diff_args *d, const void *f, void *u_plane, void *stream) {
  DiffArgs a;
  int rc = fill_diff(d, a, "lbdem_diff_outlet");
  if (rc) return rc;
  LBDEM_REQUIRE(f && u_plane && d->plane >= 0 && d->plane < d->nx, "lbdem_diff_outlet: bad argument");
  a.f_in = f;
  a.u_plane = u_plane;
  return launch_diff(LBDEM_DIFF_FN(k_diff_outlet, d->precision), (unsigned)((d->ny * d->nz + 255) / 256), a,
                     stream, "lbdem_diff_outlet");
}

int lbdem_diff_seed(const lbdem_diff_args *d, const void *f, const void *g_u, void *lam, void *stream) {
  DiffArgs a;
  int rc = fill_diff(d, a, "lbdem_diff_seed");
  if (rc) return rc;
  LBDEM_REQUIRE(f && g_u && lam && d->plane >= 0 && d->plane < d->nx, "lbdem_diff_seed: bad argument");
  a.f_in = f;
  a.g_u = g_u;
  a.lam_out = lam;
  return launch_diff(LBDEM_DIFF_FN(k_diff_seed, d->precision), (unsigned)((a.ncell + 255) / 256), a, stream,
                     "lbdem_diff_seed");
}

}  // extern "C"
