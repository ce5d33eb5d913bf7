// Shared device/host helpers for liblbdem_hip.so (gfx950).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stddef.h>
#include "../../include/lbdem_hip.h"

namespace lbdem {

// ---------------------------------------------------------------------------------------
// D3Q19 velocity set, in the reference's direction order (src/lbm3d/lbm_solver3d.py:60-91):
// 0 rest; 1..6 faces (+x,-x,+y,-y,+z,-z); 7..18 edges.
// ---------------------------------------------------------------------------------------
constexpr int Q = 19;
constexpr int kCX[Q] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0};
constexpr int kCY[Q] = {0, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 0, 0, 1, -1, 1, -1};
constexpr int kCZ[Q] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1, 1, -1, -1, 1};
constexpr int kOPP[Q] = {0, 2, 1, 4, 3, 6, 5, 8, 7, 10, 9, 12, 11, 14, 13, 16, 15, 18, 17};
constexpr double kW[Q] = {1.0 / 3.0,  1.0 / 18.0, 1.0 / 18.0, 1.0 / 18.0, 1.0 / 18.0,
                          1.0 / 18.0, 1.0 / 18.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0,
                          1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0,
                          1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0};

// ---------------------------------------------------------------------------------------
// error reporting (thread-local message, see lbdem_last_error)
// ---------------------------------------------------------------------------------------
void set_error(const char *fmt, ...);

#define LBDEM_CHECK_LAUNCH(name)                                              \
  do {                                                                        \
    hipError_t e_ = hipGetLastError();                                        \
    if (e_ != hipSuccess) {                                                   \
      ::lbdem::set_error("%s: launch failed: %s", name, hipGetErrorString(e_)); \
      return LBDEM_E_HIP;                                                     \
    }                                                                         \
  } while (0)

#define LBDEM_REQUIRE(cond, ...)        \
  do {                                  \
    if (!(cond)) {                      \
      ::lbdem::set_error(__VA_ARGS__);  \
      return LBDEM_E_ARG;               \
    }                                   \
  } while (0)

inline hipStream_t as_stream(void *s) { return reinterpret_cast<hipStream_t>(s); }

// Largest workgroup of the row-tiled collide-stream with V cells per thread.  The two- and
// four-cell variants are capped at 256 threads: under a 1024-thread bound the compiler limits
// them to 64 VGPRs, which the lagged Smagorinsky collision (and FP64) overflows into scratch.
constexpr int pencil_max_threads(int V) { return V == 1 ? 1024 : 256; }

// ---------------------------------------------------------------------------------------
// storage types: load/store with conversion to the compute type
// ---------------------------------------------------------------------------------------
// Streaming loads/stores of the population arrays: non-temporal unless built with
// -DLBDEM_TEMPORAL_POPULATIONS.
template <typename R>
__device__ __forceinline__ R lbdem_ld_raw(const R *p) {
#ifdef LBDEM_TEMPORAL_POPULATIONS
  return *p;
#else
  return __builtin_nontemporal_load(p);
#endif
}
template <typename R>
__device__ __forceinline__ void lbdem_st_raw(R *p, R v) {
#ifdef LBDEM_TEMPORAL_POPULATIONS
  *p = v;
#else
  __builtin_nontemporal_store(v, p);
#endif
}

// cvt: stored element -> compute type; pack: compute type -> stored element
template <typename S> struct Store;
template <> struct Store<float> {
  using T = float;
  static __device__ __forceinline__ float cvt(float e) { return e; }
  static __device__ __forceinline__ float pack(float v) { return v; }
  static __device__ __forceinline__ float ld(const float *p) { return lbdem_ld_raw(p); }
  static __device__ __forceinline__ void st(float *p, float v) { lbdem_st_raw(p, v); }
};
template <> struct Store<double> {
  using T = double;
  static __device__ __forceinline__ double cvt(double e) { return e; }
  static __device__ __forceinline__ double pack(double v) { return v; }
  static __device__ __forceinline__ double ld(const double *p) { return lbdem_ld_raw(p); }
  static __device__ __forceinline__ void st(double *p, double v) { lbdem_st_raw(p, v); }
};
template <> struct Store<_Float16> {
  using T = float;
  static __device__ __forceinline__ float cvt(_Float16 e) { return (float)e; }
  static __device__ __forceinline__ _Float16 pack(float v) { return (_Float16)v; }
  static __device__ __forceinline__ float ld(const _Float16 *p) { return (float)lbdem_ld_raw(p); }
  static __device__ __forceinline__ void st(_Float16 *p, float v) { lbdem_st_raw(p, (_Float16)v); }
};
template <> struct Store<__bf16> {
  using T = float;
  static __device__ __forceinline__ float cvt(__bf16 e) {
    return __uint_as_float(((uint32_t)__builtin_bit_cast(uint16_t, e)) << 16);
  }
  static __device__ __forceinline__ __bf16 pack(float v) { return (__bf16)v; }  // RNE (v_cvt_pk_bf16_f32)
  static __device__ __forceinline__ float ld(const __bf16 *p) {
    return cvt(__builtin_bit_cast(__bf16, lbdem_ld_raw(reinterpret_cast<const uint16_t *>(p))));
  }
  static __device__ __forceinline__ void st(__bf16 *p, float v) {
    lbdem_st_raw(reinterpret_cast<uint16_t *>(p), __builtin_bit_cast(uint16_t, pack(v)));
  }
};

// ---------------------------------------------------------------------------------------
// kernel parameter block (passed by value)
// ---------------------------------------------------------------------------------------
struct KParams {
  int nx, ny, nz;
  int ncell;
  void *f;
  const uint8_t *flags;
  const uint32_t *nmask;  // wall-neighbour bits of NEAR cells (bit q: x + c_q is a wall)
  const float *rho_field;  // cell densities (Ladd term at INLET walls, boundary pass)
  const float *vel;
  void *lag;  // float, or double with FP64 storage
  const uint32_t *psc_key;
  const float *psc_grain;
  float *psc_force;
  const float *volfrac;
  const float *velsolid;
  const float *hforce;
  const float *volfrac_prev;
  const float *hforce_prev;
  // physics constants
  double omega0, tau0, cs2_smag, rho0, nu_lu;
  double fx, fy, fz;
  int has_force;
  int collision, coupling;  // LBDEM_COLL_* / LBDEM_COUPLE_* of the lattice
  int t_odd;  // layout parity of the state being read (t % 2)
  // outputs for macros / readout kernels
  float *rho_out;
  float *vel_out;
  const float *mask;
  void *out;
  int out_f64;
  const void *in;
  double *work;
};

constexpr uint32_t kNbrMoving = 1u << 31;  // nmask: some neighbour is a moving wall

__device__ __forceinline__ int wrapi(int v, int n) { return v < 0 ? n - 1 : (v >= n ? 0 : v); }

// periodic neighbour x + (dx,dy,dz), as compute_neigh_index (src/lbm3d/lbm_solver3d.py:308-342)
__device__ __forceinline__ int nbr_index(const KParams &p, int i, int j, int k, int dx, int dy, int dz) {
  const int ii = wrapi(i + dx, p.nx), jj = wrapi(j + dy, p.ny), kk = wrapi(k + dz, p.nz);
  return (ii * p.ny + jj) * p.nz + kk;
}

__device__ __forceinline__ void cell_ijk(const KParams &p, int c, int &i, int &j, int &k) {
  const int nyz = p.ny * p.nz;
  i = c / nyz;
  const int r = c - i * nyz;
  j = r / p.nz;
  k = r - j * p.nz;
}

// ---------------------------------------------------------------------------------------
// wave64 / block reductions in double
// ---------------------------------------------------------------------------------------
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// per-grain PSC force / torque reduction over the grains of glist (all grains when null);
// shared by lbdem_psc_forces and the hybrid coupling (psc.hip)
int launch_psc_forces(const lbdem_lattice *lat, const lbdem_grains *g, const lbdem_units *u, const int32_t *glist,
                      int32_t nlist, const uint32_t *key, const float *cell_force, uint32_t *stale_key,
                      const int32_t *stale_bbox, void *stream, const char *who);

}  // namespace lbdem
