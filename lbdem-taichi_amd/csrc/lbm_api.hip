// extern "C" entry points of the lattice core (see include/lbdem_hip.h).
#include <stdarg.h>
#include <stdio.h>
#include <mutex>
#include <string>
#include "common.hpp"
#include "lbm_table.hpp"

namespace lbdem {

static thread_local std::string g_last_error;

void set_error(const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_last_error = buf;
}

// Scratch of lbdem_lbm_reduce_static: the per-workgroup partials of the reduction (five doubles
// per 256 cells), for lattices of at most kReduceStaticMaxCells cells.  One copy per GPU, held
// until the process ends.  Host code never uses the name `scratch` as an address: the device
// address of each GPU's copy comes from hipGetSymbolAddress, taken once per device.
constexpr int64_t kReduceStaticMaxCells = int64_t(1) << 25;
constexpr int64_t kReduceStaticDoubles = (kReduceStaticMaxCells / 256) * 5;
static __device__ __align__(256) double scratch[kReduceStaticDoubles];

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static double *g_scratch_ptr[kMaxDevices];

// device address of the current device's scratch (nullptr, with the error set, on failure)
static double *scratch_for_current_device(const char *who) {
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess || dev < 0 || dev >= kMaxDevices) {
    set_error("%s: no usable current device (%s)", who, e != hipSuccess ? hipGetErrorString(e) : "index");
    return nullptr;
  }
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void *ptr = nullptr;
    e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr) {
      set_error("%s: hipGetSymbolAddress(scratch) failed: %s", who, hipGetErrorString(e));
      return nullptr;
    }
    g_scratch_ptr[dev] = static_cast<double *>(ptr);
  }
  return g_scratch_ptr[dev];
}

static const LbmKernels *table_for(int storage) {
  switch (storage) {
    case LBDEM_STORE_F32: return &lbm_kernels_f32();
    case LBDEM_STORE_F16: return &lbm_kernels_f16();
    case LBDEM_STORE_BF16: return &lbm_kernels_bf16();
    case LBDEM_STORE_F64: return &lbm_kernels_f64();
    default: return nullptr;
  }
}

static int fill_params(const lbdem_lattice *lat, const lbdem_step_args *a, KParams &p, const char *who) {
  LBDEM_REQUIRE(lat && a, "%s: null lattice or args", who);
  LBDEM_REQUIRE(lat->nx > 0 && lat->ny > 0 && lat->nz > 0, "%s: bad extents %d x %d x %d", who, lat->nx,
                lat->ny, lat->nz);
  const int64_t n = (int64_t)lat->nx * lat->ny * lat->nz;
  LBDEM_REQUIRE(n < (int64_t(1) << 31), "%s: %lld cells exceed 2^31", who, (long long)n);
  LBDEM_REQUIRE(lat->coupling >= 0 && lat->coupling <= LBDEM_COUPLE_HYBRID, "%s: bad coupling %d", who, lat->coupling);
  LBDEM_REQUIRE(lat->collision >= 0 && lat->collision < 3, "%s: bad collision %d", who, lat->collision);
  LBDEM_REQUIRE(a->f && a->flags && a->nmask, "%s: f, flags and nmask are required", who);
  LBDEM_REQUIRE(lat->omega0 > 0.0, "%s: omega0 must be positive", who);
  p = KParams{};
  p.nx = lat->nx;
  p.ny = lat->ny;
  p.nz = lat->nz;
  p.ncell = (int)n;
  p.f = a->f;
  p.flags = a->flags;
  p.nmask = a->nmask;
  p.vel = a->vel;
  p.lag = a->lag;
  p.psc_key = a->psc_key;
  p.psc_grain = a->psc_grain;
  p.psc_force = a->psc_force;
  p.volfrac = a->volfrac;
  p.velsolid = a->velsolid;
  p.hforce = a->hforce;
  p.volfrac_prev = a->volfrac_prev;
  p.hforce_prev = a->hforce_prev;
  p.omega0 = lat->omega0;
  p.tau0 = 1.0 / lat->omega0;
  p.collision = lat->collision;
  p.coupling = lat->coupling;
  p.cs2_smag = lat->cs_smag * lat->cs_smag;
  p.rho0 = lat->rho0;
  p.nu_lu = a->nu_lu;
  p.fx = lat->force[0];
  p.fy = lat->force[1];
  p.fz = lat->force[2];
  p.has_force = (lat->force[0] != 0.0 || lat->force[1] != 0.0 || lat->force[2] != 0.0) ? 1 : 0;
  p.t_odd = (int)(a->t & 1);
  p.rho_field = a->rho;
  return LBDEM_OK;
}

static int launch(KernelFn k, int ncell, const KParams &p, void *stream, const char *who) {
  if (ncell == 0) return LBDEM_OK;
  const unsigned blocks = (unsigned)((ncell + 255) / 256);
  KParams pp = p;
  void *args[] = {&pp};
  hipError_t e = hipLaunchKernel(reinterpret_cast<const void *>(k), dim3(blocks), dim3(256), args, 0,
                                 as_stream(stream));
  if (e != hipSuccess) {
    set_error("%s: launch failed: %s", who, hipGetErrorString(e));
    return LBDEM_E_HIP;
  }
  return LBDEM_OK;
}

// deterministic fixed-order final sum of per-workgroup partials (one workgroup)
// compute_plane_average_stress (src/lbm3d/lbm_solver3d.py:407-456): mean of the shear
// components over the FLUID cells of one plane; one workgroup, fixed summation order
__global__ __launch_bounds__(1024) void k_plane_stress(int nx, int ny, int nz, const uint8_t *flags,
                                                       const float *stress, int axis, int pos, double *out) {
  const size_t N = (size_t)nx * ny * nz;
  const int n1 = axis == 0 ? ny : nx, n2 = axis == 2 ? ny : nz;
  double s[4] = {0.0, 0.0, 0.0, 0.0};
  for (int e = threadIdx.x; e < n1 * n2; e += 1024) {
    const int a = e / n2, b = e - a * n2;
    const int i = axis == 0 ? pos : a, j = axis == 0 ? a : (axis == 1 ? pos : b), k = axis == 2 ? pos : b;
    const size_t c = ((size_t)i * ny + j) * nz + k;
    if (!(flags[c] & LBDEM_FLAG_FLUID)) continue;
    s[0] += (double)stress[3 * N + c];
    s[1] += (double)stress[4 * N + c];
    s[2] += (double)stress[5 * N + c];
    s[3] += 1.0;
  }
  __shared__ double sh[4][16];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
  for (int m = 0; m < 4; ++m) {
    const double v = wave_sum(s[m]);
    if (lane == 0) sh[m][wv] = v;
  }
  __syncthreads();
  if (threadIdx.x < 4) {
    double v = 0.0;
    for (int w = 0; w < 16; ++w) v += sh[threadIdx.x][w];
    sh[threadIdx.x][0] = v;
  }
  __syncthreads();
  if (threadIdx.x < 3) {
    const double cnt = sh[3][0];
    out[threadIdx.x] = cnt > 0.0 ? sh[threadIdx.x][0] / cnt : 0.0;
  }
}

__global__ __launch_bounds__(256) void k_reduce_final(const double *work, int nblocks, double *out) {
  __shared__ double sh[256][5];
  double s[5] = {0, 0, 0, 0, 0};
  for (int b = threadIdx.x; b < nblocks; b += 256)
#pragma unroll
    for (int m = 0; m < 5; ++m) s[m] += work[(size_t)b * 5 + m];
#pragma unroll
  for (int m = 0; m < 5; ++m) sh[threadIdx.x][m] = s[m];
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w)
#pragma unroll
      for (int m = 0; m < 5; ++m) sh[threadIdx.x][m] += sh[threadIdx.x + w][m];
    __syncthreads();
  }
  if (threadIdx.x < 5) out[threadIdx.x] = sh[0][threadIdx.x];
}

// NEAR bit and neighbour mask of every cell.  Active cells: bit q set where x + c_q is a wall
// (kNbrMoving: some neighbour moves), NEAR where any is.  Wall and moving-wall cells: bit q set
// where x + c_q is active -- the targets whose slots a wall must push back unchanged in the odd
// row-tiled step (the other targets are walls, whose slots hold nothing that is read); no NEAR.
// Each thread rewrites only its own flag byte and reads only the neighbours' ACTIVE / WALL /
// MOVING bits, which this kernel does not change.
__global__ __launch_bounds__(256) void k_prepare_flags(const KParams p, uint8_t *flags, uint32_t *nmask) {
  const int c = blockIdx.x * 256 + threadIdx.x;
  if (c >= p.ncell) return;
  const uint8_t fl = flags[c];
  uint32_t m = 0u;
  const bool act = fl & LBDEM_FLAG_ACTIVE;
  if (act || (fl & (LBDEM_FLAG_WALL | LBDEM_FLAG_MOVING))) {
    int i, j, k;
    cell_ijk(p, c, i, j, k);
#pragma unroll
    for (int q = 1; q < Q; ++q) {
      const uint8_t fn = flags[nbr_index(p, i, j, k, kCX[q], kCY[q], kCZ[q])];
      if (act) {
        if (fn & (LBDEM_FLAG_WALL | LBDEM_FLAG_MOVING)) m |= 1u << q;
        if (fn & LBDEM_FLAG_MOVING) m |= kNbrMoving;
      } else if (fn & LBDEM_FLAG_ACTIVE) {
        m |= 1u << q;
      }
    }
  }
  const uint8_t near = (act && m) ? LBDEM_FLAG_NEAR : 0;
  if ((fl & LBDEM_FLAG_NEAR) != near) flags[c] = (uint8_t)((fl & ~LBDEM_FLAG_NEAR) | near);
  nmask[c] = m;
}

// LBDEM_FLAG_ROW_IDLE on the first cell of each row with no active cell (one thread per row;
// runs after k_prepare_flags, which keeps the bit as it is)
__global__ __launch_bounds__(256) void k_mark_idle_rows(const KParams p, uint8_t *flags) {
  const int row = blockIdx.x * 256 + threadIdx.x;
  if (row >= p.nx * p.ny) return;
  const size_t base = (size_t)row * p.nz;
  bool any = false;
  for (int k = 0; k < p.nz && !any; ++k) any = flags[base + k] & LBDEM_FLAG_ACTIVE;
  const uint8_t fl = flags[base];
  const uint8_t idle = any ? 0 : LBDEM_FLAG_ROW_IDLE;
  if ((fl & LBDEM_FLAG_ROW_IDLE) != idle) flags[base] = (uint8_t)((fl & ~LBDEM_FLAG_ROW_IDLE) | idle);
}

}  // namespace lbdem

using namespace lbdem;

extern "C" {

int lbdem_abi_version(void) { return LBDEM_ABI_VERSION; }

int lbdem_lbm_prepare_flags(const lbdem_lattice *lat, uint8_t *flags, uint32_t *nmask, void *stream) {
  LBDEM_REQUIRE(lat && flags && nmask, "lbdem_lbm_prepare_flags: null argument");
  LBDEM_REQUIRE(lat->nx > 0 && lat->ny > 0 && lat->nz > 0, "lbdem_lbm_prepare_flags: bad extents");
  const int64_t n = (int64_t)lat->nx * lat->ny * lat->nz;
  LBDEM_REQUIRE(n < (int64_t(1) << 31), "lbdem_lbm_prepare_flags: too many cells");
  KParams p{};
  p.nx = lat->nx;
  p.ny = lat->ny;
  p.nz = lat->nz;
  p.ncell = (int)n;
  hipLaunchKernelGGL(k_prepare_flags, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, as_stream(stream), p, flags,
                     nmask);
  hipLaunchKernelGGL(k_mark_idle_rows, dim3((unsigned)((lat->nx * lat->ny + 255) / 256)), dim3(256), 0,
                     as_stream(stream), p, flags);
  LBDEM_CHECK_LAUNCH("lbdem_lbm_prepare_flags");
  return LBDEM_OK;
}

const char *lbdem_last_error(void) { return g_last_error.c_str(); }

int lbdem_lbm_init(const lbdem_lattice *lat, const lbdem_step_args *a, const float *rho, void *stream) {
  KParams p;
  int rc = fill_params(lat, a, p, "lbdem_lbm_init");
  if (rc) return rc;
  LBDEM_REQUIRE(rho && a->vel, "lbdem_lbm_init: rho and vel are required");
  LBDEM_REQUIRE(lat->collision != LBDEM_COLL_SMAG_LAGGED || a->lag,
                "lbdem_lbm_init: lagged Smagorinsky needs the lag buffer");
  const LbmKernels *t = table_for(lat->storage);
  LBDEM_REQUIRE(t, "lbdem_lbm_init: bad storage %d", lat->storage);
  p.rho_out = const_cast<float *>(rho);
  return launch(t->init[lat->coupling], p.ncell, p, stream, "lbdem_lbm_init");
}

int lbdem_lbm_step(const lbdem_lattice *lat, const lbdem_step_args *a, void *stream) {
  KParams p;
  int rc = fill_params(lat, a, p, "lbdem_lbm_step");
  if (rc) return rc;
  const LbmKernels *t = table_for(lat->storage);
  LBDEM_REQUIRE(t, "lbdem_lbm_step: bad storage %d", lat->storage);
  int coll = lat->collision;
  if (lat->coupling == LBDEM_COUPLE_VANS) coll = LBDEM_COLL_BGK;  // VANS carries its own omega
  KernelFn k = t->step[coll][lat->coupling][p.t_odd];
  if (!k) {
    set_error("lbdem_lbm_step: collision %d with coupling %d is not supported", lat->collision, lat->coupling);
    return LBDEM_E_UNSUPPORTED;
  }
  LBDEM_REQUIRE(coll != LBDEM_COLL_SMAG_LAGGED || a->lag, "lbdem_lbm_step: lagged Smagorinsky needs lag");
  LBDEM_REQUIRE(lat->coupling != LBDEM_COUPLE_PSC || (a->psc_key && a->psc_grain && a->psc_force),
                "lbdem_lbm_step: PSC needs psc_key, psc_grain, psc_force");
  LBDEM_REQUIRE(lat->coupling != LBDEM_COUPLE_EQIBM || (a->volfrac && a->velsolid && a->nu_lu > 0),
                "lbdem_lbm_step: EqIBM needs volfrac, velsolid and nu_lu > 0");
  LBDEM_REQUIRE(lat->coupling != LBDEM_COUPLE_HYBRID ||
                    (a->psc_key && a->psc_grain && a->psc_force && a->volfrac && a->velsolid && a->nu_lu > 0),
                "lbdem_lbm_step: hybrid coupling needs psc_key, psc_grain, psc_force, volfrac, velsolid and nu_lu > 0");
  LBDEM_REQUIRE(lat->coupling != LBDEM_COUPLE_VANS ||
                    (a->volfrac && a->hforce && a->volfrac_prev && a->hforce_prev && a->nu_lu > 0),
                "lbdem_lbm_step: VANS needs volfrac, hforce, volfrac_prev, hforce_prev, nu_lu > 0");
  // row-tiled variants: whole rows per workgroup, LDS staging of the z shift (<= 64 KiB);
  // kernel 2 / 3 / 4: one / two / four cells per thread.  Automatic choice (kernel 0), the
  // fastest measured on MI355X for each parity (256^3 TGV; all variants give identical bits):
  // one cell per thread for FP32 and FP64; for 16-bit storage two cells in the even step and
  // four in the odd one (BF16 235 / 260 us against 245 / 260 with four in both), and with the
  // lagged Smagorinsky one and two (340 / 379 us against 445 / 441: with the lag state the
  // four-cell variant holds 138 VGPRs, three wavefronts per SIMD; one cell 57, eight).
  // Narrower where the row does not fit.
  int kern = lat->kernel;
  LBDEM_REQUIRE(kern >= 0 && kern <= 4, "lbdem_lbm_step: bad kernel variant %d", kern);
  if (kern == 0) {
    if (lat->storage == LBDEM_STORE_F32 || lat->storage == LBDEM_STORE_F64) kern = 2;
    else if (coll == LBDEM_COLL_SMAG_LAGGED) kern = p.t_odd ? 3 : 2;
    else kern = p.t_odd ? (lat->nz % 4 == 0 ? 4 : 3) : 3;
  }
  const int tsize = lat->storage == LBDEM_STORE_F64 ? 8 : 4;
  auto fits = [&](int v) {
    const int b = ((lat->nz / v + 63) / 64) * 64;
    return lat->nz % v == 0 && b <= pencil_max_threads(v) && (size_t)10 * b * v * tsize <= 65536;
  };
  if (lat->kernel == 0) {  // automatic: the preferred width, else a narrower one that fits the row
    while (kern > 2 && !fits(kern == 4 ? 4 : 2)) --kern;
  }
  const int V = kern == 4 ? 4 : kern == 3 ? 2 : 1;
  const int block = ((lat->nz / V + 63) / 64) * 64;
  const size_t lds = p.t_odd ? (size_t)10 * block * V * tsize : 0;
  const bool pencil_ok = fits(V);
  KernelFn kp = V == 4 ? t->pencil4[coll][lat->coupling][p.t_odd]
                : V == 2 ? t->pencil2[coll][lat->coupling][p.t_odd] : t->pencil[coll][lat->coupling][p.t_odd];
  LBDEM_REQUIRE(lat->kernel < 2 || (pencil_ok && kp), "lbdem_lbm_step: row-tiled kernel %d needs nz %% %d == 0, a "
                "row of at most %d cells and (4 cells per thread) 32-bit or 16-bit storage", kern, V,
                pencil_max_threads(V) * V);
  if (kern >= 2 && pencil_ok && kp) {
    void *args[] = {&p};
    hipError_t e = hipLaunchKernel(reinterpret_cast<const void *>(kp), dim3((unsigned)(lat->nx * lat->ny)),
                                   dim3((unsigned)block), args, lds, as_stream(stream));
    if (e != hipSuccess) {
      set_error("lbdem_lbm_step: launch failed: %s", hipGetErrorString(e));
      return LBDEM_E_HIP;
    }
    return LBDEM_OK;
  }
  return launch(k, p.ncell, p, stream, "lbdem_lbm_step");
}

int lbdem_lbm_macros(const lbdem_lattice *lat, const lbdem_step_args *a, float *rho, float *vel,
                     const float *mask, void *stream) {
  KParams p;
  int rc = fill_params(lat, a, p, "lbdem_lbm_macros");
  if (rc) return rc;
  LBDEM_REQUIRE(rho && vel, "lbdem_lbm_macros: rho and vel are required");
  LBDEM_REQUIRE(lat->coupling != LBDEM_COUPLE_VANS || (a->volfrac && a->hforce),
                "lbdem_lbm_macros: VANS needs volfrac and hforce");
  const LbmKernels *t = table_for(lat->storage);
  LBDEM_REQUIRE(t, "lbdem_lbm_macros: bad storage %d", lat->storage);
  p.rho_out = rho;
  p.vel_out = vel;
  p.mask = mask;
  return launch(t->macros[lat->coupling], p.ncell, p, stream, "lbdem_lbm_macros");
}

int lbdem_lbm_read_f(const lbdem_lattice *lat, const lbdem_step_args *a, void *out, int32_t out_f64,
                     void *stream) {
  KParams p;
  int rc = fill_params(lat, a, p, "lbdem_lbm_read_f");
  if (rc) return rc;
  LBDEM_REQUIRE(out, "lbdem_lbm_read_f: out is required");
  const LbmKernels *t = table_for(lat->storage);
  LBDEM_REQUIRE(t, "lbdem_lbm_read_f: bad storage %d", lat->storage);
  p.out = out;
  p.out_f64 = out_f64 ? 1 : 0;
  return launch(t->read_f, p.ncell, p, stream, "lbdem_lbm_read_f");
}

int lbdem_lbm_write_f(const lbdem_lattice *lat, const lbdem_step_args *a, const void *in, int32_t in_f64,
                      void *stream) {
  KParams p;
  int rc = fill_params(lat, a, p, "lbdem_lbm_write_f");
  if (rc) return rc;
  LBDEM_REQUIRE(in, "lbdem_lbm_write_f: in is required");
  const LbmKernels *t = table_for(lat->storage);
  LBDEM_REQUIRE(t, "lbdem_lbm_write_f: bad storage %d", lat->storage);
  p.in = in;
  p.out_f64 = in_f64 ? 1 : 0;
  return launch(t->write_f, p.ncell, p, stream, "lbdem_lbm_write_f");
}

int lbdem_lbm_boundaries(const lbdem_lattice *lat, const lbdem_step_args *a, const int32_t *cells,
                         const uint8_t *kinds, int32_t ncells, float *rho, float *vel, void *stream) {
  KParams p;
  int rc = fill_params(lat, a, p, "lbdem_lbm_boundaries");
  if (rc) return rc;
  LBDEM_REQUIRE(ncells >= 0, "lbdem_lbm_boundaries: negative cell count");
  if (ncells == 0) return LBDEM_OK;
  LBDEM_REQUIRE(cells && kinds && rho && vel, "lbdem_lbm_boundaries: cells, kinds, rho and vel are required");
  const LbmKernels *t = table_for(lat->storage);
  LBDEM_REQUIRE(t, "lbdem_lbm_boundaries: bad storage %d", lat->storage);
  void *args[] = {&p, &cells, &kinds, &ncells, &rho, &vel};
  hipError_t e = hipLaunchKernel(reinterpret_cast<const void *>(t->boundaries), dim3((unsigned)((ncells + 255) / 256)),
                                 dim3(256), args, 0, as_stream(stream));
  if (e != hipSuccess) {
    set_error("lbdem_lbm_boundaries: launch failed: %s", hipGetErrorString(e));
    return LBDEM_E_HIP;
  }
  return LBDEM_OK;
}

int lbdem_lbm_free_slip(const lbdem_lattice *lat, const lbdem_step_args *a, const int32_t *pairs, int32_t npairs,
                        double *work, void *stream) {
  KParams p;
  int rc = fill_params(lat, a, p, "lbdem_lbm_free_slip");
  if (rc) return rc;
  LBDEM_REQUIRE(npairs >= 0, "lbdem_lbm_free_slip: negative pair count");
  if (npairs == 0) return LBDEM_OK;
  LBDEM_REQUIRE(pairs && work, "lbdem_lbm_free_slip: pairs and work are required");
  const LbmKernels *t = table_for(lat->storage);
  LBDEM_REQUIRE(t, "lbdem_lbm_free_slip: bad storage %d", lat->storage);
  const unsigned blocks = (unsigned)((npairs + 255) / 256);
  hipLaunchKernelGGL(t->free_slip_read, dim3(blocks), dim3(256), 0, as_stream(stream), p, pairs, (int)npairs, work);
  LBDEM_CHECK_LAUNCH("lbdem_lbm_free_slip");
  hipLaunchKernelGGL(t->free_slip_write, dim3(blocks), dim3(256), 0, as_stream(stream), p, pairs, (int)npairs,
                     (const double *)work);
  LBDEM_CHECK_LAUNCH("lbdem_lbm_free_slip");
  return LBDEM_OK;
}

int lbdem_lbm_stress(const lbdem_lattice *lat, const lbdem_step_args *a, float *stress, void *stream) {
  KParams p;
  int rc = fill_params(lat, a, p, "lbdem_lbm_stress");
  if (rc) return rc;
  LBDEM_REQUIRE(stress, "lbdem_lbm_stress: stress is required");
  LBDEM_REQUIRE(lat->collision != LBDEM_COLL_SMAG_LAGGED || a->lag, "lbdem_lbm_stress: lagged Smagorinsky needs lag");
  LBDEM_REQUIRE(lat->coupling != LBDEM_COUPLE_VANS || a->volfrac, "lbdem_lbm_stress: VANS needs volfrac");
  const LbmKernels *t = table_for(lat->storage);
  LBDEM_REQUIRE(t, "lbdem_lbm_stress: bad storage %d", lat->storage);
  p.out = stress;
  return launch(t->stress, p.ncell, p, stream, "lbdem_lbm_stress");
}

int lbdem_lbm_plane_stress(const lbdem_lattice *lat, const uint8_t *flags, const float *stress, int32_t axis,
                           int32_t position, double *out, void *stream) {
  LBDEM_REQUIRE(lat && flags && stress && out, "lbdem_lbm_plane_stress: null argument");
  LBDEM_REQUIRE(axis >= 0 && axis <= 2, "lbdem_lbm_plane_stress: axis %d is not 0, 1 or 2", axis);
  const int ext = axis == 0 ? lat->nx : axis == 1 ? lat->ny : lat->nz;
  LBDEM_REQUIRE(position >= 0 && position < ext, "lbdem_lbm_plane_stress: position %d outside [0, %d)", position,
                ext);
  hipLaunchKernelGGL(k_plane_stress, dim3(1), dim3(1024), 0, as_stream(stream), lat->nx, lat->ny, lat->nz, flags,
                     stress, axis, position, out);
  LBDEM_CHECK_LAUNCH("lbdem_lbm_plane_stress");
  return LBDEM_OK;
}

int64_t lbdem_reduce_work_doubles(const lbdem_lattice *lat) {
  if (!lat) return 0;
  const int64_t n = (int64_t)lat->nx * lat->ny * lat->nz;
  return ((n + 255) / 256) * 5;
}

int lbdem_lbm_reduce(const lbdem_lattice *lat, const lbdem_step_args *a, double *work, double *out,
                     void *stream) {
  KParams p;
  int rc = fill_params(lat, a, p, "lbdem_lbm_reduce");
  if (rc) return rc;
  LBDEM_REQUIRE(work && out, "lbdem_lbm_reduce: work and out are required");
  const LbmKernels *t = table_for(lat->storage);
  LBDEM_REQUIRE(t, "lbdem_lbm_reduce: bad storage %d", lat->storage);
  p.work = work;
  rc = launch(t->reduce_partial, p.ncell, p, stream, "lbdem_lbm_reduce");
  if (rc) return rc;
  const int nblocks = (p.ncell + 255) / 256;
  hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(256), 0, as_stream(stream), work, nblocks, out);
  LBDEM_CHECK_LAUNCH("lbdem_lbm_reduce");
  return LBDEM_OK;
}

int lbdem_lbm_reduce_static(const lbdem_lattice *lat, const lbdem_step_args *a, double *out, void *stream) {
  KParams p;
  int rc = fill_params(lat, a, p, "lbdem_lbm_reduce_static");
  if (rc) return rc;
  LBDEM_REQUIRE(out, "lbdem_lbm_reduce_static: out is required");
  LBDEM_REQUIRE((int64_t)p.ncell <= kReduceStaticMaxCells,
                "lbdem_lbm_reduce_static: %d cells exceed the static scratch (%lld cells; use lbdem_lbm_reduce)",
                p.ncell, (long long)kReduceStaticMaxCells);
  const LbmKernels *t = table_for(lat->storage);
  LBDEM_REQUIRE(t, "lbdem_lbm_reduce_static: bad storage %d", lat->storage);
  double *work = scratch_for_current_device("lbdem_lbm_reduce_static");
  if (!work) return LBDEM_E_HIP;
  // every workgroup writes all five of its partials before the final pass reads them: nothing
  // in the scratch needs clearing
  p.work = work;
  rc = launch(t->reduce_partial, p.ncell, p, stream, "lbdem_lbm_reduce_static");
  if (rc) return rc;
  const int nblocks = (p.ncell + 255) / 256;
  hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(256), 0, as_stream(stream), work, nblocks, out);
  LBDEM_CHECK_LAUNCH("lbdem_lbm_reduce_static");
  return LBDEM_OK;
}

}  // extern "C"
