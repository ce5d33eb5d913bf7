// Discrete-element method on gfx950: hash broad phase, contact resolution with per-particle
// history slots, Hertz-Mindlin / linear spring-dashpot laws, planar walls, integration.
//
// Two execution paths share the per-contact device code:
//  - the general path: about ten grid-wide kernels per DEM substep (broad phase with the
//    reference's Morton hash and counting sort, two-phase contact resolution, walls,
//    integration, slot compaction);
//  - the small-system path (n <= kSmallN): one workgroup runs many substeps in a single
//    launch, with barriers between phases, and finds candidate pairs through bucket lists in
//    LDS keyed by the same cell hash, which yields exactly the set the hash-table probe yields.
//
// FP64 throughout.  The float -> cell arithmetic is compiled without contraction
// (-ffp-contract=off) so cell keys match the CPU restatement bit for bit.
#include "common.hpp"

namespace lbdem {

constexpr int kSmallN = 2048;

struct DemP {  // lbdem_dem_params, copied into the kernel argument block
  lbdem_dem_params p;
};

// --------------------------------------------------------------------------------------
// broad phase (src/dem3d/bpcd/bpcd.py)
// --------------------------------------------------------------------------------------
__device__ __forceinline__ uint32_t spread10(int32_t v) {
  uint32_t x = (uint32_t)v & 0x3ffu;
  x = (x | x << 16) & 0x30000ffu;
  x = (x | x << 8) & 0x300f00fu;
  x = (x | x << 4) & 0x30c30c3u;
  x = (x | x << 2) & 0x9249249u;
  return x;
}
// morton3d32 % T  (bpcd.py:154-189)
__device__ __forceinline__ int32_t hash3(int32_t x, int32_t y, int32_t z, int32_t T) {
  return (int32_t)((spread10(x) | (spread10(y) << 1) | (spread10(z) << 2)) & (uint32_t)(T - 1));
}
// cell = floor((x - min) / cs)  (bpcd.py:191-194)
__device__ __forceinline__ void cell_of(const lbdem_dem_params &P, const double *x, int32_t c[3]) {
#pragma unroll
  for (int a = 0; a < 3; ++a) c[a] = (int32_t)floor((x[a] - P.dmin[a]) / P.cell);
}

// hashes of the probed cells of a particle in cell c at x: the octant's 8 cells
// (bpcd.py:114-152) or all 27; with DEDUP, duplicates (hash aliasing) become -1 so every bucket
// is visited once (callers that only ask whether a hash is among them skip that step).  NP is a
// template argument so every index into h is a compile-time constant.
template <int NP, bool DEDUP = true>
__device__ __forceinline__ void probe_cells(const lbdem_dem_params &P, const double *x, const int32_t (&c)[3],
                                            int32_t (&h)[NP]) {
  if constexpr (NP == 8) {
    int32_t d[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      const double center = ((double)c[a] + 0.5) * P.cell + P.dmin[a];
      d[a] = (x[a] > center) ? 1 : -1;
    }
    h[0] = hash3(c[0], c[1], c[2], P.T);
    h[1] = hash3(c[0] + d[0], c[1], c[2], P.T);
    h[2] = hash3(c[0], c[1] + d[1], c[2], P.T);
    h[3] = hash3(c[0], c[1], c[2] + d[2], P.T);
    h[4] = hash3(c[0], c[1] + d[1], c[2] + d[2], P.T);
    h[5] = hash3(c[0] + d[0], c[1], c[2] + d[2], P.T);
    h[6] = hash3(c[0] + d[0], c[1] + d[1], c[2], P.T);
    h[7] = hash3(c[0] + d[0], c[1] + d[1], c[2] + d[2], P.T);
  } else {
#pragma unroll
    for (int t = 0; t < 27; ++t) h[t] = hash3(c[0] + t / 9 - 1, c[1] + (t / 3) % 3 - 1, c[2] + t % 3 - 1, P.T);
  }
  if constexpr (DEDUP) {
#pragma unroll
    for (int k = 1; k < NP; ++k)
#pragma unroll
      for (int e = 0; e < k; ++e)
        if (h[e] == h[k]) h[k] = -1;  // a later match compares against -1 only, which changes nothing
  }
}

// atomicAdd(ptr + h, 1) for each active lane; lanes of the wave sharing h are combined into one
// atomic, for up to four distinct values (a coarse grid puts most of a wave into one or two
// buckets, whose counters would otherwise serialise the whole launch).  Returns the old value
// plus the lane's rank among the lanes combined with it.  Call from converged control flow.
__device__ __forceinline__ int32_t agg_add_one(int32_t *ptr, int32_t h, bool active) {
  const int lane = threadIdx.x & 63;
  int32_t res = 0;
  bool pending = active;
  for (int r = 0; r < 4; ++r) {
    const unsigned long long m = __ballot(pending);
    if (m == 0ull) break;
    const int leader = __ffsll((unsigned long long)m) - 1;
    const int32_t hl = __shfl(h, leader, 64);
    const bool mine = pending && h == hl;
    const unsigned long long grp = __ballot(mine);
    int32_t old = 0;
    if (lane == leader) old = atomicAdd(ptr + hl, (int32_t)__popcll(grp));
    old = __shfl(old, leader, 64);
    if (mine) {
      res = old + (int32_t)__popcll(grp & ((1ull << lane) - 1ull));
      pending = false;
    }
  }
  if (pending) res = atomicAdd(ptr + h, 1);
  return res;
}

// Hashes and bucket counts.  With `slot`, particle i's rank among the particles counted into its
// bucket (what the counter's atomic returned) is kept, and the scatter needs no atomics.
__global__ __launch_bounds__(256) void k_bp_count(const lbdem_dem_params P, const double *pos, int32_t *hash,
                                                  int32_t *count, double *force, double *moment, int32_t *slot) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  const bool active = i < P.n;
  int32_t h = 0;
  if (active) {
    int32_t c[3];
    cell_of(P, pos + 3 * i, c);
    h = hash3(c[0], c[1], c[2], P.T);
    hash[i] = h;
  }
  if (count) {  // NULL: hashes only
    const int32_t r = agg_add_one(count, h, active);
    if (slot && active) slot[i] = r;
  }
  if (!active) return;
  if (force) {
#pragma unroll
    for (int a = 0; a < 3; ++a) force[3 * i + a] = moment[3 * i + a] = 0.0;  // clear_state
  }
}

// exclusive scan of count[T] -> offset[T] in two passes over tiles of kScanTile entries: the
// tiles' sums, then each tile's scan plus the sums of the tiles before it (read back from L2 by
// every workgroup; at most kScanInlineTiles of them, 2^22 entries; a larger table first scans
// the tile sums in one workgroup, k_scan_sums, so the work stays linear in T).  Eight entries
// per lane, as two 16-byte accesses when the arrays allow it (the branch is uniform: a lane's
// first entry is a multiple of 8 entries into the array).  The three-pass form (tile scans, a
// one-workgroup scan of the sums, an add pass) took 17 us per substep on the 200,000-grain
// bed's 2^21-entry table.
constexpr int kScanTile = 2048;

__device__ __forceinline__ void scan_load8(const int32_t *in, int base, int n, int32_t v[8]) {
  if (base + 8 <= n && ((uintptr_t)in & 15u) == 0) {
    const int4 a = *reinterpret_cast<const int4 *>(in + base);
    const int4 b = *reinterpret_cast<const int4 *>(in + base + 4);
    v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w;
    v[4] = b.x; v[5] = b.y; v[6] = b.z; v[7] = b.w;
  } else {
#pragma unroll
    for (int m = 0; m < 8; ++m) v[m] = (base + m < n) ? in[base + m] : 0;
  }
}

__device__ __forceinline__ int32_t wave_sum(int32_t v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

__global__ __launch_bounds__(256) void k_scan_reduce(const int32_t *in, int32_t *tile_sum, int n) {
  __shared__ int32_t sh[4];
  int32_t v[8];
  scan_load8(in, blockIdx.x * kScanTile + threadIdx.x * 8, n, v);
  const int32_t s = wave_sum(v[0] + v[1] + v[2] + v[3] + v[4] + v[5] + v[6] + v[7]);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) tile_sum[blockIdx.x] = sh[0] + sh[1] + sh[2] + sh[3];
}

constexpr int kScanInlineTiles = 2048;

// exclusive scan of tile_sum[0, nt) in place, one workgroup
__global__ __launch_bounds__(1024) void k_scan_sums(int32_t *tile_sum, int nt) {
  __shared__ int32_t sh[16];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  int32_t carry = 0;
  for (int base = 0; base < nt; base += 1024) {
    const int k = base + (int)threadIdx.x;
    const int32_t v = k < nt ? tile_sum[k] : 0;
    int32_t x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const int32_t y = __shfl_up(x, o, 64);
      if (lane >= o) x += y;
    }
    if (lane == 63) sh[wave] = x;
    __syncthreads();
    int32_t pre = carry;
    for (int w = 0; w < wave; ++w) pre += sh[w];
    int32_t tot = 0;
    for (int w = 0; w < 16; ++w) tot += sh[w];
    if (k < nt) tile_sum[k] = pre + x - v;
    carry += tot;
    __syncthreads();  // sh is rewritten by the next chunk
  }
}

// prescanned: tile_sum holds the exclusive prefix of the tile sums (k_scan_sums)
__global__ __launch_bounds__(256) void k_scan_down(const int32_t *in, int32_t *out, const int32_t *tile_sum,
                                                   int n, int prescanned) {
  __shared__ int32_t sh[8];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  int32_t c = 0;
  if (prescanned) {
    if (threadIdx.x == 0) c = tile_sum[blockIdx.x];
  } else {
    for (int k = threadIdx.x; k < (int)blockIdx.x; k += 256) c += tile_sum[k];
  }
  c = wave_sum(c);
  const int base = blockIdx.x * kScanTile + threadIdx.x * 8;
  int32_t v[8];
  scan_load8(in, base, n, v);
  const int32_t s = v[0] + v[1] + v[2] + v[3] + v[4] + v[5] + v[6] + v[7];
  int32_t x = s;  // inclusive scan of the lanes' sums over the wavefront
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int32_t y = __shfl_up(x, o, 64);
    if (lane >= o) x += y;
  }
  if (lane == 0) sh[wave] = c;
  if (lane == 63) sh[4 + wave] = x;
  __syncthreads();
  int32_t run = sh[0] + sh[1] + sh[2] + sh[3] + x - s;
  for (int w = 0; w < wave; ++w) run += sh[4 + w];
  int32_t o8[8];
#pragma unroll
  for (int m = 0; m < 8; ++m) {
    o8[m] = run;
    run += v[m];
  }
  if (base + 8 <= n && ((uintptr_t)out & 15u) == 0) {
    *reinterpret_cast<int4 *>(out + base) = make_int4(o8[0], o8[1], o8[2], o8[3]);
    *reinterpret_cast<int4 *>(out + base + 4) = make_int4(o8[4], o8[5], o8[6], o8[7]);
  } else {
#pragma unroll
    for (int m = 0; m < 8; ++m)
      if (base + m < n) out[base + m] = o8[m];
  }
}

static void scan_table(const lbdem_dem_work *w, int T, hipStream_t s) {
  const unsigned nt = (unsigned)((T + kScanTile - 1) / kScanTile);
  const int prescan = nt > (unsigned)kScanInlineTiles ? 1 : 0;
  hipLaunchKernelGGL(k_scan_reduce, dim3(nt), dim3(256), 0, s, w->count, w->tile_sum, T);
  if (prescan) hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(1024), 0, s, w->tile_sum, (int)nt);
  hipLaunchKernelGGL(k_scan_down, dim3(nt), dim3(256), 0, s, w->count, w->offset, w->tile_sum, T, prescan);
}

__global__ __launch_bounds__(256) void k_bp_scatter(const lbdem_dem_params P, const int32_t *hash,
                                                    const int32_t *offset, int32_t *current, int32_t *pid) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  const bool active = i < P.n;
  const int32_t h = active ? hash[i] : 0;
  const int32_t slot = agg_add_one(current, h, active);
  if (active) pid[offset[h] + slot] = i;
}
// the scatter with the slots k_bp_count kept (12.8 -> 5.2 us per substep on the 200,000-grain bed)
__global__ __launch_bounds__(256) void k_bp_scatter_slots(const lbdem_dem_params P, const int32_t *hash,
                                                          const int32_t *offset, const int32_t *slot, int32_t *pid) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i < P.n) pid[offset[hash[i]] + slot[i]] = i;
}

// positions and radii in bucket order, one 32-byte record per particle, for the touching probe
__global__ __launch_bounds__(256) void k_bp_sort_state(const lbdem_dem_params P, const double *pos,
                                                       const double *rad, const int32_t *pid, double *spos) {
  const int t = blockIdx.x * 256 + threadIdx.x;
  if (t >= P.n) return;
  const int32_t j = pid[t];
  reinterpret_cast<double4 *>(spos)[t] = make_double4(pos[3 * j], pos[3 * j + 1], pos[3 * j + 2], rad[j]);
}

// probe: candidate pairs (i, pid > i); one atomic per wavefront reserves the output range
template <int NP>
__global__ __launch_bounds__(256) void k_bp_probe(const lbdem_dem_params P, const double *pos,
                                                  const int32_t *count, const int32_t *offset, const int32_t *pid,
                                                  int32_t *pairs, int64_t cap, unsigned long long *npairs) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  int32_t h[NP];
  constexpr int nc = NP;
  int mine = 0;
  if (i < P.n) {
    int32_t c[3];
    cell_of(P, pos + 3 * i, c);
    probe_cells<NP>(P, pos + 3 * i, c, h);
#pragma unroll
    for (int k = 0; k < nc; ++k) {
      if (h[k] < 0) continue;
      const int32_t b = offset[h[k]], e = b + count[h[k]];
      for (int idx = b; idx < e; ++idx) mine += (pid[idx] > i) ? 1 : 0;
    }
  }
  // wave-level exclusive scan of `mine`
  const int lane = threadIdx.x & 63;
  int incl = mine;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int t = __shfl_up(incl, o, 64);
    if (lane >= o) incl += t;
  }
  const int total = __shfl(incl, 63, 64);
  unsigned long long base = 0;
  if (lane == 63 && total > 0) base = atomicAdd(npairs, (unsigned long long)total);
  base = __shfl(base, 63, 64);
  unsigned long long w = base + (unsigned long long)(incl - mine);
  if (i < P.n && mine > 0) {
#pragma unroll
    for (int k = 0; k < nc; ++k) {
      if (h[k] < 0) continue;
      const int32_t b = offset[h[k]], e = b + count[h[k]];
      for (int idx = b; idx < e; ++idx) {
        const int32_t p = pid[idx];
        if (p > i) {
          if (w < (unsigned long long)cap) {
            pairs[2 * w] = i;
            pairs[2 * w + 1] = p;
          }
          ++w;
        }
      }
    }
  }
}

__global__ __launch_bounds__(256) void k_bp_reset(const lbdem_dem_params P, const int32_t *hash, int32_t *count,
                                                  int32_t *current) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= P.n) return;
  count[hash[i]] = 0;
  current[hash[i]] = 0;
}

// --------------------------------------------------------------------------------------
// contact mechanics
// --------------------------------------------------------------------------------------
// status counters (overlap and time-step warnings, slot overflow): the lanes of a wavefront that
// count together take one atomic (a dense bed raises hundreds of thousands per substep, and one
// atomic each on a single address serialises them)
__device__ __forceinline__ void count_event(int32_t *ctr) {
  const unsigned long long m = __ballot(1);
  if ((int)(threadIdx.x & 63) == __ffsll((unsigned long long)m) - 1) atomicAdd(ctr, (int32_t)__popcll(m));
}

enum { MAT_RHO, MAT_E, MAT_NU, MAT_MU, MAT_REST, MAT_KN, MAT_KT, MAT_DPN, MAT_DPS };

__device__ __forceinline__ void cross3(const double *a, const double *b, double *c) {
  c[0] = a[1] * b[2] - a[2] * b[1];
  c[1] = a[2] * b[0] - a[0] * b[2];
  c[2] = a[0] * b[1] - a[1] * b[0];
}
__device__ __forceinline__ double norm3(const double *a) { return sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]); }
__device__ __forceinline__ void matvec(const double *M, const double *v, double *o) {
#pragma unroll
  for (int r = 0; r < 3; ++r) o[r] = M[3 * r] * v[0] + M[3 * r + 1] * v[1] + M[3 * r + 2] * v[2];
}
// 3x3 inverse by the adjugate (ti.Matrix.inverse for n = 3)
__device__ __forceinline__ void inv3(const double *m, double *o) {
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
// rotation taking the unit vector a to +x (Rodrigues; demsolver.py:686-701)
__device__ __forceinline__ void rot_to_x(const double *a, double *R) {
  const double v[3] = {a[1] * 0.0 - a[2] * 0.0, a[2] * 1.0 - a[0] * 0.0, a[0] * 0.0 - a[1] * 1.0};
  const double s = norm3(v), c = a[0] * 1.0 + a[1] * 0.0 + a[2] * 0.0;
  if (s < 1e-12) {
#pragma unroll
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
  const double k = (1.0 - c) / (s * s);
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      const double vv = vx[3 * r] * vx[q] + vx[3 * r + 1] * vx[3 + q] + vx[3 * r + 2] * vx[6 + q];
      R[3 * r + q] = ((r == q) ? 1.0 : 0.0) + vx[3 * r + q] + k * vv;
    }
}

__device__ __forceinline__ void periodic_delta(const lbdem_dem_params &P, const double *xi, const double *xj,
                                               double *d) {
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    d[a] = xj[a] - xi[a];
    if (P.periodic[a]) {
      const double L = P.dmax[a] - P.dmin[a];
      if (d[a] > 0.5 * L) d[a] -= L;
      else if (d[a] < -0.5 * L) d[a] += L;
    }
  }
}

// contact law in the local frame (hertz.py:45-153, linear.py:43-152)
__device__ __forceinline__ void contact_law(const lbdem_dem_params &P, const lbdem_dem_state &S, int i, int j, bool wall,
                            double delta_n, const double *vc, double *shear, double *F) {
  const double *mi = P.mat[0], *mj = P.mat[wall ? 1 : 0];
  const double dt = P.dt;
  const double gap = -1.0 * delta_n;
  double kn, kt, gn, gt, mu;
  const double pi = 3.14159265358979323846;
  if (P.model == 0) {
    const double Ys = 1.0 / ((1.0 - mi[MAT_NU] * mi[MAT_NU]) / mi[MAT_E] + (1.0 - mj[MAT_NU] * mj[MAT_NU]) / mj[MAT_E]);
    const double Gs = 1.0 / (2.0 * (2.0 - mi[MAT_NU]) * (1.0 + mi[MAT_NU]) / mi[MAT_E] +
                             2.0 * (2.0 - mj[MAT_NU]) * (1.0 + mj[MAT_NU]) / mj[MAT_E]);
    double Rs, ms, e;
    if (wall) {
      Rs = S.radius[i];
      ms = S.mass[i];
      e = mj[MAT_REST];
      mu = mj[MAT_MU];
    } else {
      Rs = 1.0 / (1.0 / S.radius[i] + 1.0 / S.radius[j]);
      const double ri = S.radius[i];
      const double rho = S.mass[i] * 3 / (4 * pi * ri * ri * ri);
      const double tmin = pi * Rs * sqrt(rho / Gs) / (0.1631 * mi[MAT_NU] + 0.8766);  // hertz.py:156-162
      if (dt > tmin) count_event(S.status + 1);
      ms = 1.0 / (1.0 / S.mass[i] + 1.0 / S.mass[j]);
      e = 0.5 * (mi[MAT_REST] + mj[MAT_REST]);
      mu = 0.5 * (mi[MAT_MU] + mj[MAT_MU]);
    }
    const double le = log(e);
    const double beta = le / sqrt(le * le + pi * pi);
    const double Sn = 2.0 * Ys * sqrt(Rs * delta_n);
    const double St = 8.0 * Gs * sqrt(Rs * delta_n);
    kn = 4.0 / 3.0 * Ys * sqrt(Rs * delta_n);
    gn = -2.0 * beta * sqrt(5.0 / 6.0 * Sn * ms);
    kt = 8.0 * Gs * sqrt(Rs * delta_n);
    gt = -2.0 * beta * sqrt(5.0 / 6.0 * St * ms);
  } else {
    double ms, dpn, dps;
    if (wall) {
      kn = mi[MAT_KN];
      kt = mi[MAT_KT];
      dpn = mi[MAT_DPN];
      dps = mi[MAT_DPS];
      ms = S.mass[i];
      mu = mj[MAT_MU];
    } else {
      kn = mi[MAT_KN] * mj[MAT_KN] / (mi[MAT_KN] + mj[MAT_KN]);
      kt = mi[MAT_KT] * mj[MAT_KT] / (mi[MAT_KT] + mj[MAT_KT]);
      dpn = 0.5 * (mi[MAT_DPN] + mj[MAT_DPN]);
      dps = 0.5 * (mi[MAT_DPS] + mj[MAT_DPS]);
      if (dt > sqrt(S.mass[i] / mi[MAT_KN])) count_event(S.status + 1);  // linear.py:155-160
      ms = 1.0 / (1.0 / S.mass[i] + 1.0 / S.mass[j]);
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

// evaluate an active particle-particle contact held in slot `off` (demsolver.py:683-740)
__device__ __forceinline__ void eval_pair(const lbdem_dem_params &P, const lbdem_dem_state &S, int i, int j, int off, const double *d,
                          double dist, double gap) {
  double a[3] = {d[0] / dist, d[1] / dist, d[2] / dist}, R[9], Ri[9];
  rot_to_x(a, R);
  const double delta_n = -gap;
  const double Rmin = fmin(S.radius[i], S.radius[j]);
  if (delta_n > 0.05 * Rmin) count_event(S.status + 2);
  const double *xi = S.pos + 3 * i;
  double cp[3], ri[3], rj[3], vci[3], vcj[3], dv[3], vc[3], F[3], sh[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) cp[k] = xi[k] + a[k] * (S.radius[i] - delta_n);
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    S.cf_pos[3 * off + k] = cp[k];
    ri[k] = cp[k] - xi[k];
    rj[k] = cp[k] - (xi[k] + d[k]);
    sh[k] = S.cf_shear[3 * off + k];
  }
  cross3(S.omega + 3 * i, ri, vci);
  cross3(S.omega + 3 * j, rj, vcj);
#pragma unroll
  for (int k = 0; k < 3; ++k) dv[k] = (vcj[k] + S.vel[3 * j + k]) - (vci[k] + S.vel[3 * i + k]);
  matvec(R, dv, vc);
  contact_law(P, S, i, j, false, delta_n, vc, sh, F);
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    S.cf_shear[3 * off + k] = sh[k];
    S.cf_force[3 * off + k] = F[k];
  }
  inv3(R, Ri);
  const double mF[3] = {-F[0], -F[1], -F[2]};
  double Fi[3], Fj[3], Mi[3], Mj[3];
  matvec(Ri, mF, Fi);
  matvec(Ri, F, Fj);
  cross3(ri, Fi, Mi);
  cross3(rj, Fj, Mj);
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    unsafeAtomicAdd(S.force + 3 * i + k, Fi[k]);
    unsafeAtomicAdd(S.force + 3 * j + k, Fj[k]);
    unsafeAtomicAdd(S.moment + 3 * i + k, Mi[k]);
    unsafeAtomicAdd(S.moment + 3 * j + k, Mj[k]);
  }
}

// gap < 0 exactly as resolve_existing computes it (|d| - r_i - r_j < 0), with the square root
// skipped for pairs whose squared distance exceeds (r_i + r_j)^2 (1 + 1e-9): the computed gap of
// those is at least 5e-10 (r_i + r_j), far above its rounding error, so the answer is the same.
__device__ __forceinline__ bool touching(const lbdem_dem_params &P, const double *xi, double ri, const double *xj,
                                         double rj) {
  double d[3];
  periodic_delta(P, xi, xj, d);
  const double x2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
  const double a = ri + rj;
  if (x2 > a * a * (1.0 + 1e-9)) return false;
  return sqrt(x2) - ri - rj < 0;
}

// Phase 1 of resolve (demsolver.py:638-668): pairs with an active history slot are evaluated
// (gap < 0) or released (gap >= 0).  Returns 1 when the pair needs a new slot.  Only reads
// slots of other pairs, so it can run for all pairs concurrently.
__device__ __forceinline__ int resolve_existing(const lbdem_dem_params &P, const lbdem_dem_state &S, int i, int j) {
  const int K = P.K;
  const int ncf = min(S.cfn[i], K);
  int off = -1;
  for (int s = 0; s < ncf; ++s)
    if (S.cf_j[i * K + s] == j && S.cf_active[i * K + s]) {
      off = i * K + s;
      break;
    }
  double d[3];
  periodic_delta(P, S.pos + 3 * i, S.pos + 3 * j, d);
  const double dist = norm3(d);
  const double gap = dist - S.radius[i] - S.radius[j];
  if (off >= 0) {
    if (gap < 0) eval_pair(P, S, i, j, off, d, dist, gap);
    else S.cf_active[off] = 0;
    return 0;
  }
  return gap < 0 ? 1 : 0;
}

// Phase 2: a new contact takes the next slot of i (append_contact_offset, demsolver.py:352-359)
__device__ __forceinline__ void resolve_new(const lbdem_dem_params &P, const lbdem_dem_state &S, int i, int j) {
  const int K = P.K;
  const int slot = atomicAdd(S.cfn + i, 1);
  if (slot >= K) {  // the reference prints and writes cf[-1]; here: counted, no write
    count_event(S.status + 0);
    return;
  }
  const int off = i * K + slot;
  S.cf_j[off] = j;
  S.cf_active[off] = 1;
#pragma unroll
  for (int k = 0; k < 3; ++k) S.cf_force[3 * off + k] = S.cf_shear[3 * off + k] = 0.0;
  double d[3];
  periodic_delta(P, S.pos + 3 * i, S.pos + 3 * j, d);
  const double dist = norm3(d);
  eval_pair(P, S, i, j, off, d, dist, dist - S.radius[i] - S.radius[j]);
}

// Candidate pairs of the multi-kernel path, filtered: only touching pairs are listed (the
// candidate rule is the full probe's: j > i in one of i's probed buckets).  A candidate that
// does not touch changes nothing unless it holds a history slot, which k_dem_release clears;
// so the outcome is that of resolving every candidate, while the list stays the size of the
// contact set (a strongly polydisperse packing, whose cells are 4 r_max wide, has O(n^2)
// candidates).  blockIdx.y splits each particle's candidates over gridDim.y workgroups.
//
// When the probed buckets of the workgroup's particles number at most kTileBuckets (a coarse
// grid puts whole workgroups into the same few buckets), the buckets are walked together:
// 256 candidates at a time are staged in LDS and every lane tests them from there (lanes test
// only buckets of their own probe set).  Otherwise each lane walks its own buckets.
constexpr int kTileBuckets = 64;
constexpr int kTileTab = 128;  // LDS hash set collecting the workgroup's buckets
constexpr int kTileProbe = 8;  // linear-probing steps per insert before the set counts as full

// Appends (i, j); the lanes of a wavefront that emit together take one atomic on the shared
// counter (a dense bed lists millions of touching pairs per substep, and one atomic per pair
// on a single address serialises them).
__device__ __forceinline__ void emit_pair(int i, int j, int32_t *pairs, int64_t cap, unsigned long long *npairs) {
  const unsigned long long m = __ballot(1);
  const int lane = threadIdx.x & 63;
  const int leader = __ffsll((unsigned long long)m) - 1;
  unsigned long long base = 0ull;
  if (lane == leader) base = atomicAdd(npairs, (unsigned long long)__popcll(m));
  base = __shfl(base, leader, 64);
  const unsigned long long w = base + (unsigned long long)__popcll(m & ((1ull << lane) - 1ull));
  if (w < (unsigned long long)cap) {
    pairs[2 * w] = i;
    pairs[2 * w + 1] = j;
  }
}

// Pairs of the touching probe are staged in LDS and written out by the workgroup at its end,
// behind one atomic on the shared counter: a dense bed's substep lists hundreds of thousands of
// touching pairs, each wavefront's emits fall at different iterations of its lanes' loops, and
// one device-scope atomic per emit on that one address serialised the probe.  Pairs beyond the
// LDS buffer go straight to the global list.  The 8-cell probe's workgroups also hold the
// staged candidates in LDS and run a VALU-bound loop over them, so their buffer is kept small
// enough not to cost occupancy (7 workgroups per CU with it).
template <int NP>
constexpr int pair_stage_size() { return NP == 27 ? 2048 : 512; }

template <int S>
__device__ __forceinline__ void stage_pair(int i, int j, int2 *s_pairs, int *s_np, int32_t *pairs, int64_t cap,
                                           unsigned long long *npairs) {
  const unsigned long long m = __ballot(1);
  const int lane = threadIdx.x & 63;
  const int leader = __ffsll((unsigned long long)m) - 1;
  int base = 0;
  if (lane == leader) base = atomicAdd(s_np, __popcll(m));
  base = __shfl(base, leader, 64);
  const int u = base + __popcll(m & ((1ull << lane) - 1ull));
  if (u < S)
    s_pairs[u] = make_int2(i, j);
  else
    emit_pair(i, j, pairs, cap, npairs);
}

template <int NP>
__global__ __launch_bounds__(256) void k_bp_probe_touch(const lbdem_dem_params P, const double *pos,
                                                        const double *rad, const int32_t *count,
                                                        const int32_t *offset, const int32_t *pid, int32_t *pairs,
                                                        int64_t cap, unsigned long long *npairs) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  const bool active = i < P.n;
  const int split = gridDim.y, part = blockIdx.y;
  int32_t h[NP];
  double xi[3] = {0.0, 0.0, 0.0}, ri = 0.0;
#pragma unroll
  for (int k = 0; k < NP; ++k) h[k] = -1;
  if (active) {
#pragma unroll
    for (int a = 0; a < 3; ++a) xi[a] = pos[3 * i + a];
    ri = rad[i];
    int32_t c[3];
    cell_of(P, xi, c);
    probe_cells<NP>(P, xi, c, h);
  }
  if constexpr (NP == 8) {
    // ascending bucket order: lanes whose probe sets share buckets walk them in step
#pragma unroll
    for (int a = 0; a < NP - 1; ++a)
#pragma unroll
      for (int b = 0; b < NP - 1 - a; ++b) {
        const int32_t lo = min(h[b], h[b + 1]), hi = max(h[b], h[b + 1]);
        h[b] = lo;
        h[b + 1] = hi;
      }
  }
  __shared__ int32_t s_tab[kTileTab];
  __shared__ int32_t s_list[kTileBuckets];
  __shared__ int s_nb, s_over;
  __shared__ double s_x[256], s_y[256], s_z[256], s_r[256];
  __shared__ float4 s_pf[256];  // FP32 x, y, z, r of the staged candidates (one LDS read each)
  __shared__ float s_mf[256];
  __shared__ int32_t s_j[256];
  constexpr int kStage = pair_stage_size<NP>();
  __shared__ int2 s_pairs[kStage];
  __shared__ int s_np;
  __shared__ unsigned long long s_base;
  if (threadIdx.x < kTileTab) s_tab[threadIdx.x] = -1;
  if (threadIdx.x == 0) s_nb = s_over = s_np = 0;
  __syncthreads();
  // the workgroup's distinct buckets (hashes are >= 0; -1 marks a free slot), for the 8-cell
  // probe only (the 27-cell probe serves fine grids, where the set would overflow anyway).  An
  // insert gives up after kTileProbe slots and marks the set as overflowed, and a lane stops
  // inserting once it sees the mark, so a sparse grid costs a few LDS atomics per lane.
  if constexpr (NP == 8) {
#pragma unroll
    for (int k = 0; k < NP; ++k) {
      if (h[k] < 0) continue;
      if (*(volatile int *)&s_over) break;
      int slot = h[k] & (kTileTab - 1);
      bool placed = false;
      for (int t = 0; t < kTileProbe && !placed; ++t) {
        const int32_t old = atomicCAS(&s_tab[slot], -1, h[k]);
        placed = (old == -1 || old == h[k]);
        slot = (slot + 1) & (kTileTab - 1);
      }
      if (!placed) s_over = 1;
    }
  } else {
    if (threadIdx.x == 0) s_over = 1;
  }
  __syncthreads();
  if (threadIdx.x < kTileTab && s_tab[threadIdx.x] >= 0) {
    const int u = atomicAdd(&s_nb, 1);
    if (u < kTileBuckets) s_list[u] = s_tab[threadIdx.x];
  }
  __syncthreads();
  const int nb = s_nb;
  if (s_over == 0 && nb <= kTileBuckets) {
    // FP32 pre-test (no periodic axis): a candidate is dropped only when its FP32 distance
    // exceeds (r_i + r_j)(1 + 1e-6) + 2^-20 (|x_i|_inf + |x_j|_inf), a bound on the FP32
    // conversion and rounding errors four times over, so every pair the FP64 test could call
    // touching reaches it
    const bool pre = !(P.periodic[0] || P.periodic[1] || P.periodic[2]);
    const float xif = (float)xi[0], yif = (float)xi[1], zif = (float)xi[2], rif = (float)ri;
    const float mif = fmaxf(fabsf(xif), fmaxf(fabsf(yif), fabsf(zif)));
    for (int u = 0; u < nb; ++u) {
      const int32_t bk = s_list[u];
      bool mine = false;
#pragma unroll
      for (int k = 0; k < NP; ++k) mine |= (h[k] == bk);
      const int32_t b = offset[bk], e = b + count[bk];
      for (int base = b + part * 256; base < e; base += split * 256) {
        __syncthreads();  // the previous chunk has been read
        const int t = base + (int)threadIdx.x;
        if (t < e) {
          const int32_t j = pid[t];
          const double x = pos[3 * j], y = pos[3 * j + 1], z = pos[3 * j + 2], r = rad[j];
          s_j[threadIdx.x] = j;
          s_x[threadIdx.x] = x;
          s_y[threadIdx.x] = y;
          s_z[threadIdx.x] = z;
          s_r[threadIdx.x] = r;
          const float xf = (float)x, yf = (float)y, zf = (float)z;
          s_pf[threadIdx.x] = make_float4(xf, yf, zf, (float)r);
          s_mf[threadIdx.x] = fmaxf(fabsf(xf), fmaxf(fabsf(yf), fabsf(zf)));
        }
        __syncthreads();
        if (active && mine) {
          const int m = min(256, e - base);
          for (int q = 0; q < m; ++q) {
            const int32_t j = s_j[q];
            if (j <= i) continue;
            if (pre) {
              const float4 pj = s_pf[q];
              const float dx = xif - pj.x, dy = yif - pj.y, dz = zif - pj.z;
              const float x2 = dx * dx + dy * dy + dz * dz;
              const float T = (rif + pj.w) * 1.000001f + 9.53674316e-7f * (mif + s_mf[q]);
              if (x2 > T * T * 1.000004f) continue;
            }
            const double xj[3] = {s_x[q], s_y[q], s_z[q]};
            if (touching(P, xi, ri, xj, s_r[q])) stage_pair<kStage>(i, j, s_pairs, &s_np, pairs, cap, npairs);
          }
        }
      }
    }
  } else if (active) {
#pragma unroll
    for (int k = 0; k < NP; ++k) {
      if (h[k] < 0) continue;
      const int32_t b = offset[h[k]], e = b + count[h[k]];
      for (int idx = b + part; idx < e; idx += split) {
        const int32_t j = pid[idx];
        if (j <= i || !touching(P, xi, ri, pos + 3 * j, rad[j])) continue;
        stage_pair<kStage>(i, j, s_pairs, &s_np, pairs, cap, npairs);
      }
    }
  }
  __syncthreads();
  const int ns = min(s_np, kStage);
  if (threadIdx.x == 0) s_base = ns > 0 ? atomicAdd(npairs, (unsigned long long)ns) : 0ull;
  __syncthreads();
  const unsigned long long base = s_base;
  for (int u = threadIdx.x; u < ns; u += 256) {
    const unsigned long long w = base + (unsigned long long)u;
    if (w < (unsigned long long)cap) reinterpret_cast<int2 *>(pairs)[w] = s_pairs[u];
  }
}

// The touching probe of the 27-cell rule over particles in bucket order (spos): lane t takes
// the t-th particle of the bucket order, so a wavefront's particles share most of their buckets
// and read their candidates from neighbouring 32-byte records.  The walk is latency-bound, so
// blockIdx.y (three workgroups) splits each particle's buckets, not their entries; a lane loads
// its share's bucket headers together, then walks those buckets' entries as one range, four
// loads at a time (eight measured no faster).  Each touching pair is listed once, tested with the
// lower index first (as the resolve kernels and the oracle evaluate it):
//  - HALF (T >= 256): a half shell.  hash3 keeps each cell coordinate modulo a power of two of at
//    least 4 then (from T = 64 on, in fact), so the 27 probed cells have distinct hashes and j's bucket is i's cell + d for
//    exactly one d, and i's is j's cell - d.  A particle walks the 13 cells with d > 0 (z first,
//    then y, then x) and its own bucket's entries after its own position t0: every pair is found
//    from exactly one side.  The three workgroups take 5, 5 and 4 of these 14 buckets.
//  - otherwise (hashes may alias within the 27): each workgroup takes a z slab of 9 deduplicated
//    buckets and walks only their entries after t0, so a pair is found by whichever of its two
//    particles comes first in bucket order (the aliasing is symmetric: j's cell aliases i's
//    cell + d exactly when i's aliases j's - d).
template <bool HALF>
__global__ __launch_bounds__(256) void k_bp_probe_touch_sorted(const lbdem_dem_params P, const int32_t *count,
                                                               const int32_t *offset, const int32_t *pid,
                                                               const double4 *spos, int32_t *pairs, int64_t cap,
                                                               unsigned long long *npairs) {
  constexpr int kStage = pair_stage_size<27>();
  constexpr int NB = HALF ? 5 : 9;  // buckets per workgroup
  __shared__ int2 s_pairs[kStage];
  __shared__ int s_np;
  __shared__ unsigned long long s_base;
  if (threadIdx.x == 0) s_np = 0;
  __syncthreads();
  const int t0 = blockIdx.x * 256 + threadIdx.x;
  const int part = blockIdx.y;  // gridDim.y == 3
  if (t0 < P.n) {
    const int i = pid[t0];
    const double4 q = spos[t0];
    const double xi[3] = {q.x, q.y, q.z};
    const double ri = q.w;
    int32_t c[3], hs[NB];
    int self = -1;  // the bucket among hs holding this particle (walked after t0 only)
    cell_of(P, xi, c);
    if constexpr (HALF) {
      const int32_t x = c[0], y = c[1], z = c[2];
      if (part == 0) {
        hs[0] = hash3(x - 1, y - 1, z + 1, P.T);
        hs[1] = hash3(x, y - 1, z + 1, P.T);
        hs[2] = hash3(x + 1, y - 1, z + 1, P.T);
        hs[3] = hash3(x - 1, y, z + 1, P.T);
        hs[4] = hash3(x, y, z, P.T);
        self = 4;
      } else if (part == 1) {
        hs[0] = hash3(x, y, z + 1, P.T);
        hs[1] = hash3(x + 1, y, z + 1, P.T);
        hs[2] = hash3(x - 1, y + 1, z + 1, P.T);
        hs[3] = hash3(x, y + 1, z + 1, P.T);
        hs[4] = hash3(x + 1, y, z, P.T);
      } else {
        hs[0] = hash3(x + 1, y + 1, z + 1, P.T);
        hs[1] = hash3(x - 1, y + 1, z, P.T);
        hs[2] = hash3(x, y + 1, z, P.T);
        hs[3] = hash3(x + 1, y + 1, z, P.T);
        hs[4] = -1;
      }
    } else {
      int32_t h[27];
      probe_cells<27>(P, xi, c, h);
#pragma unroll
      for (int m = 0; m < 9; ++m) hs[m] = part == 0 ? h[3 * m] : (part == 1 ? h[3 * m + 1] : h[3 * m + 2]);
    }
    // all headers first, then the entries of the buckets as one range
    int32_t beg[NB], pre[NB];
    int total = 0;
#pragma unroll
    for (int m = 0; m < NB; ++m) {
      int32_t b = 0, cn = 0;
      if (hs[m] >= 0) {
        b = offset[hs[m]];
        cn = count[hs[m]];
        if (!HALF || m == self) {  // entries after t0 only
          const int32_t e1 = b + cn;
          b = max(b, t0 + 1);
          cn = max(e1 - b, 0);
        }
      }
      pre[m] = total;
      beg[m] = b - total;  // entry e of the range, pre[m] <= e < pre[m] + cn, is at beg[m] + e
      total += cn;
    }
    for (int e0 = 0; e0 < total; e0 += 4) {
      int32_t jj[4];
      double4 qq[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int e = e0 + u;
        jj[u] = -1;
        if (e < total) {
          int32_t b = beg[0];  // the last bucket starting at or before e holds it
#pragma unroll
          for (int m = 1; m < NB; ++m)
            if (e >= pre[m]) b = beg[m];
          jj[u] = pid[b + e];
          qq[u] = spos[b + e];
        }
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int j = jj[u];
        if (j < 0) continue;
        const double xj[3] = {qq[u].x, qq[u].y, qq[u].z};
        const bool hit = i < j ? touching(P, xi, ri, xj, qq[u].w) : touching(P, xj, qq[u].w, xi, ri);
        if (hit) stage_pair<kStage>(min(i, j), max(i, j), s_pairs, &s_np, pairs, cap, npairs);
      }
    }
  }
  __syncthreads();
  const int ns = min(s_np, kStage);
  if (threadIdx.x == 0) s_base = ns > 0 ? atomicAdd(npairs, (unsigned long long)ns) : 0ull;
  __syncthreads();
  const unsigned long long base = s_base;
  for (int u = threadIdx.x; u < ns; u += 256) {
    const unsigned long long w = base + (unsigned long long)u;
    if (w < (unsigned long long)cap) reinterpret_cast<int2 *>(pairs)[w] = s_pairs[u];
  }
}

// The touching pairs of the 8-cell rule on strongly polydisperse packings (lbdem_dem_params::
// fine_cell > 0).  The grid of the rule (cell 4 r_max) then puts most particles into a few
// buckets and its probe tests O(n^2) candidates.  Every touching pair lies within one cell of
// that grid's octant block, so the touching pairs among the candidates are the touching pairs
// that pass the candidate rule; they are found here on a second grid (fine_cell > 2 large_r,
// 27-cell probe, particles in bucket order as in k_bp_probe_touch_sorted) for the particles of
// radius <= large_r, and by testing each larger particle against all particles
// (k_bp_touch_large), and a pair (i, j), i < j, is kept when j's bucket (hash of the rule's grid)
// is one of i's probed buckets.
__device__ __forceinline__ bool probed_by(const lbdem_dem_params &P, const double *xi, int32_t hj) {
  int32_t c[3], h[8];
  cell_of(P, xi, c);
  probe_cells<8, false>(P, xi, c, h);
  bool in = false;
#pragma unroll
  for (int k = 0; k < 8; ++k) in |= (h[k] == hj);
  return in;
}

__global__ __launch_bounds__(256) void k_bp_probe_touch_fine(const lbdem_dem_params P, const lbdem_dem_params Pf,
                                                             const int32_t *count, const int32_t *offset,
                                                             const int32_t *pid, const double4 *spos,
                                                             const int32_t *hash, int32_t *pairs, int64_t cap,
                                                             unsigned long long *npairs) {
  constexpr int kStage = pair_stage_size<8>();
  __shared__ int2 s_pairs[kStage];
  __shared__ int s_np;
  __shared__ unsigned long long s_base;
  if (threadIdx.x == 0) s_np = 0;
  __syncthreads();
  const int t0 = blockIdx.x * 256 + threadIdx.x;
  const int split = gridDim.y, part = blockIdx.y;
  if (t0 < P.n) {
    const int i = pid[t0];
    const double4 q = spos[t0];
    const double xi[3] = {q.x, q.y, q.z};
    const double ri = q.w;
    if (ri <= P.large_r) {
      int32_t c[3], h[27], hm[8];
      cell_of(Pf, xi, c);
      probe_cells<27>(Pf, xi, c, h);
      {
        int32_t cm[3];
        cell_of(P, xi, cm);
        probe_cells<8, false>(P, xi, cm, hm);
      }
#pragma unroll
      for (int k = 0; k < 27; ++k) {
        if (k % split != part || h[k] < 0) continue;
        const int32_t b = offset[h[k]], e = b + count[h[k]];
        for (int idx = b; idx < e; idx += 4) {
          int32_t jj[4];
          double4 qq[4];
#pragma unroll
          for (int u = 0; u < 4; ++u) {
            jj[u] = -1;
            if (idx + u < e) {
              jj[u] = pid[idx + u];
              qq[u] = spos[idx + u];
            }
          }
#pragma unroll
          for (int u = 0; u < 4; ++u) {
            if (jj[u] <= i || qq[u].w > P.large_r) continue;
            const double xj[3] = {qq[u].x, qq[u].y, qq[u].z};
            if (!touching(P, xi, ri, xj, qq[u].w)) continue;
            const int32_t hj = hash[jj[u]];
            bool in = false;
#pragma unroll
            for (int m = 0; m < 8; ++m) in |= (hm[m] == hj);
            if (in) stage_pair<kStage>(i, jj[u], s_pairs, &s_np, pairs, cap, npairs);
          }
        }
      }
    }
  }
  __syncthreads();
  const int ns = min(s_np, kStage);
  if (threadIdx.x == 0) s_base = ns > 0 ? atomicAdd(npairs, (unsigned long long)ns) : 0ull;
  __syncthreads();
  const unsigned long long base = s_base;
  for (int u = threadIdx.x; u < ns; u += 256) {
    const unsigned long long w = base + (unsigned long long)u;
    if (w < (unsigned long long)cap) reinterpret_cast<int2 *>(pairs)[w] = s_pairs[u];
  }
}

// pairs of the large particles: blockIdx.y picks large particle L, threads run over all j; a
// pair of two large particles is taken by the one with the smaller index
__global__ __launch_bounds__(256) void k_bp_touch_large(const lbdem_dem_params P, const double *pos,
                                                        const double *rad, const int32_t *hash, const int32_t *large,
                                                        int32_t *pairs, int64_t cap, unsigned long long *npairs) {
  constexpr int kStage = pair_stage_size<8>();
  __shared__ int2 s_pairs[kStage];
  __shared__ int s_np;
  __shared__ unsigned long long s_base;
  if (threadIdx.x == 0) s_np = 0;
  __syncthreads();
  const int L = large[blockIdx.y];
  const int j = blockIdx.x * 256 + threadIdx.x;
  if (j < P.n && j != L && !(rad[j] > P.large_r && j < L)) {
    if (touching(P, pos + 3 * L, rad[L], pos + 3 * j, rad[j])) {
      const int a = min(L, j), b = max(L, j);
      if (probed_by(P, pos + 3 * a, hash[b])) stage_pair<kStage>(a, b, s_pairs, &s_np, pairs, cap, npairs);
    }
  }
  __syncthreads();
  const int ns = min(s_np, kStage);
  if (threadIdx.x == 0) s_base = ns > 0 ? atomicAdd(npairs, (unsigned long long)ns) : 0ull;
  __syncthreads();
  const unsigned long long base = s_base;
  for (int u = threadIdx.x; u < ns; u += 256) {
    const unsigned long long w = base + (unsigned long long)u;
    if (w < (unsigned long long)cap) reinterpret_cast<int2 *>(pairs)[w] = s_pairs[u];
  }
}

// The history slots of candidate pairs that no longer touch are released, as resolve_existing
// does for every candidate with gap >= 0 (demsolver.py:638-668); contacts whose partner left
// i's probed buckets keep their slot, as in the reference.  Runs before the resolve kernels.
// It also clears the bucket counters the substep's broad phase filled (count and current at
// rhash[i], the hashes of the grid the table held), as k_bp_reset does for the broad phase
// alone: nothing reads them after the probe.
template <int NP>
__global__ __launch_bounds__(256) void k_dem_release(const lbdem_dem_params P, const lbdem_dem_state S,
                                                     const int32_t *hash, const int32_t *rhash, int32_t *count,
                                                     int32_t *current) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= P.n) return;
  const int32_t hr = rhash[i];
  count[hr] = 0;
  if (current) current[hr] = 0;
  const int K = P.K;
  const int m = min(S.cfn[i], K);
  if (m == 0) return;
  const double *xi = S.pos + 3 * i;
  int32_t c[3];
  cell_of(P, xi, c);
  int32_t h[NP];
  probe_cells<NP, false>(P, xi, c, h);
  const double ri = S.radius[i];
  for (int s = 0; s < m; ++s) {
    const int off = i * K + s;
    if (!S.cf_active[off]) continue;
    const int j = S.cf_j[off];
    if (j <= i) continue;
    const int32_t hj = hash[j];
    bool cand = false;
#pragma unroll
    for (int k = 0; k < NP; ++k) cand |= (h[k] == hj);
    if (cand && !touching(P, xi, ri, S.pos + 3 * j, S.radius[j])) S.cf_active[off] = 0;
  }
}

// particle i's six wall-contact flags (wcf_active) as a bit mask, and back (changed bits only)
__device__ __forceinline__ unsigned load_wall_mask(const lbdem_dem_state &S, int i) {
  unsigned m = 0u;
#pragma unroll
  for (int w = 0; w < 6; ++w) m |= (S.wcf_active[6 * i + w] != 0 ? 1u : 0u) << w;
  return m;
}
__device__ __forceinline__ void store_wall_mask(const lbdem_dem_state &S, int i, unsigned before, unsigned after) {
#pragma unroll
  for (int w = 0; w < 6; ++w)
    if (((before ^ after) >> w) & 1u) S.wcf_active[6 * i + w] = (int32_t)((after >> w) & 1u);
}

// resolve_wall + evaluate_wall (demsolver.py:742-838) for particle i; forces added in place.
// wact: bit w = contact with wall w active (wcf_active), updated here; the caller keeps it.
__device__ __forceinline__ void resolve_walls(const lbdem_dem_params &P, const lbdem_dem_state &S, int i, unsigned &wact) {
  // walls within reach, all six tested without branches; the loop below visits only those and
  // the walls with an active contact (usually none)
  unsigned reach = 0u;
  {
    const double *x = S.pos + 3 * i;
    const double ri = S.radius[i];
#pragma unroll
    for (int w = 0; w < 6; ++w) {
      const double *nrm = P.wall_n[w];
      const double pd = x[0] * nrm[0] + x[1] * nrm[1] + x[2] * nrm[2] - P.wall_d[w];
      reach |= (P.wall_type[w] != 1 && -pd <= ri) ? 1u << w : 0u;
    }
  }
  const unsigned visit = reach | wact;
  if (!visit) return;
  for (int w = 0; w < 6; ++w) {
    if (!((visit >> w) & 1u) || P.wall_type[w] == 1) continue;
    const double *nrm = P.wall_n[w];
    const double *x = S.pos + 3 * i;
    const double pd = x[0] * nrm[0] + x[1] * nrm[1] + x[2] * nrm[2] - P.wall_d[w];
    const int c = 6 * i + w;
    if ((wact >> w) & 1u) {
      if (-pd > S.radius[i]) {
        wact &= ~(1u << w);
        continue;
      }
    } else {
      if (!(-pd <= S.radius[i])) continue;
      wact |= 1u << w;
#pragma unroll
      for (int k = 0; k < 3; ++k) S.wcf_shear[3 * c + k] = 0.0;
    }
    double R[9], Ri[9];
    rot_to_x(nrm, R);
    const double delta_n = pd + S.radius[i];
    const double distance = S.radius[i] - delta_n;
    if (delta_n > 0.05 * S.radius[i]) count_event(S.status + 3);
    double ri[3], vci[3], mv[3], vc[3], F[3], sh[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      ri[k] = distance * nrm[k];
      S.wcf_pos[3 * c + k] = x[k] + ri[k];
      sh[k] = S.wcf_shear[3 * c + k];
    }
    cross3(S.omega + 3 * i, ri, vci);
#pragma unroll
    for (int k = 0; k < 3; ++k) mv[k] = -(vci[k] + S.vel[3 * i + k]);
    matvec(R, mv, vc);
    contact_law(P, S, i, w, true, delta_n, vc, sh, F);
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      S.wcf_shear[3 * c + k] = sh[k];
      S.wcf_force[3 * c + k] = F[k];
    }
    inv3(R, Ri);
    const double mF[3] = {-F[0], -F[1], -F[2]};
    double Fi[3], Mi[3];
    matvec(Ri, mF, Fi);
    cross3(ri, Fi, Mi);
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      S.force[3 * i + k] += Fi[k];
      S.moment[3 * i + k] += Mi[k];
    }
  }
}

// quat2RotMatrix (demsolver/utils.py:34-61)
__device__ __forceinline__ void quat_rot(const double *q, double *R) {
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

__device__ __forceinline__ void wrap_position(const lbdem_dem_params &P, double *x) {
#pragma unroll
  for (int a = 0; a < 3; ++a)
    if (P.periodic[a]) {
      const double L = P.dmax[a] - P.dmin[a];
      if (x[a] > P.dmax[a]) x[a] -= L;
      else if (x[a] < P.dmin[a]) x[a] += L;
    }
}

__device__ __forceinline__ void rotate(const lbdem_dem_params &P, const lbdem_dem_state &S, int i) {
  const double dt = P.dt;
  double alpha[3];
  const double *In = S.inertia + 9 * i;
  if (In[1] == 0.0 && In[2] == 0.0 && In[3] == 0.0 && In[5] == 0.0 && In[6] == 0.0 && In[7] == 0.0 &&
      In[0] == In[4] && In[0] == In[8]) {
    // isotropic inertia c I (spheres): R^-1 (cI)^-1 (R M - w_l x c w_l) = M / c; equal to the
    // general expression up to its rounding
    const double ic = 1.0 / In[0];
#pragma unroll
    for (int k = 0; k < 3; ++k) alpha[k] = S.moment[3 * i + k] * ic;
  } else {
    double R[9], Ri[9], Ii[9], Ml[3], wl[3], Iw[3], cr[3], t[3], wdl[3];
    quat_rot(S.quat + 4 * i, R);
    matvec(R, S.moment + 3 * i, Ml);
    matvec(R, S.omega + 3 * i, wl);
    matvec(In, wl, Iw);
    cross3(wl, Iw, cr);
#pragma unroll
    for (int k = 0; k < 3; ++k) t[k] = Ml[k] - cr[k];
    inv3(In, Ii);
    matvec(Ii, t, wdl);
    inv3(R, Ri);
    matvec(Ri, wdl, alpha);
  }
  double *w = S.omega + 3 * i, *q = S.quat + 4 * i;
#pragma unroll
  for (int k = 0; k < 3; ++k) w[k] += 0.5 * (S.omega_dot[3 * i + k] + alpha[k]) * dt;
  const double dq0 = -0.5 * (q[1] * w[0] + q[2] * w[1] + q[3] * w[2]);
  const double dq1 = +0.5 * (q[0] * w[0] - q[3] * w[1] + q[2] * w[2]);
  const double dq2 = +0.5 * (q[3] * w[0] + q[0] * w[1] + q[1] * w[2]);
  const double dq3 = +0.5 * (-q[2] * w[0] + q[1] * w[1] + q[0] * w[2]);
  q[0] += dq0 * dt;
  q[1] += dq1 * dt;
  q[2] += dq2 * dt;
  q[3] += dq3 * dt;
  const double qn = sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
#pragma unroll
  for (int k = 0; k < 4; ++k) q[k] /= qn;
#pragma unroll
  for (int k = 0; k < 3; ++k) S.omega_dot[3 * i + k] = alpha[k];
}

// walls, apply_body_force (demsolver.py:437-448) and update (:514-636) for particle i
__device__ __forceinline__ void walls_and_integrate(const lbdem_dem_params &P, const lbdem_dem_state &S, int i, unsigned &wact) {
  resolve_walls(P, S, i, wact);
  const double dt = P.dt;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    S.force[3 * i + k] += S.mass[i] * P.g[k];
    S.force[3 * i + k] += S.force_fluid[3 * i + k];
    S.moment[3 * i + k] += S.moment_fluid[3 * i + k];
  }
  if (S.freeze && S.freeze[i]) return;
  double *x = S.pos + 3 * i, *v = S.vel + 3 * i;
  if (S.fixvel && S.fixvel[i]) {
#pragma unroll
    for (int k = 0; k < 3; ++k) x[k] += v[k] * dt;
    wrap_position(P, x);
    rotate(P, S, i);
    return;
  }
  double a[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) a[k] = S.force[3 * i + k] / S.mass[i];
#pragma unroll
  for (int k = 0; k < 3; ++k) v[k] += 0.5 * (S.acc[3 * i + k] + a[k]) * dt;
#pragma unroll
  for (int k = 0; k < 3; ++k) x[k] += v[k] * dt + 0.5 * a[k] * dt * dt;
  wrap_position(P, x);
#pragma unroll
  for (int k = 0; k < 3; ++k) S.acc[3 * i + k] = a[k];
  rotate(P, S, i);
}

// late_clear_state / remove_inactive_contact (demsolver.py:371-386, 431-436)
__device__ __forceinline__ void compact_slots(const lbdem_dem_params &P, const lbdem_dem_state &S, int i) {
  const int K = P.K;
  const int m = min(S.cfn[i], K);
  int act = 0;
  for (int s = 0; s < m; ++s) act += S.cf_active[i * K + s] ? 1 : 0;
  int o = 0;
  for (int s = 0; s < m && o < act; ++s) {
    const int src = i * K + s;
    if (!S.cf_active[src]) continue;
    const int dst = i * K + o;
    if (dst != src) {
      S.cf_j[dst] = S.cf_j[src];
      S.cf_active[dst] = S.cf_active[src];
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        S.cf_pos[3 * dst + k] = S.cf_pos[3 * src + k];
        S.cf_force[3 * dst + k] = S.cf_force[3 * src + k];
        S.cf_shear[3 * dst + k] = S.cf_shear[3 * src + k];
      }
    }
    ++o;
  }
  for (int s = act; s < m; ++s) S.cf_active[i * K + s] = 0;
  S.cfn[i] = act;
}

// --------------------------------------------------------------------------------------
// general-path kernels
// --------------------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_dem_resolve_existing(const lbdem_dem_params P, const lbdem_dem_state S,
                                                              const int32_t *pairs, const unsigned long long *npairs,
                                                              int64_t cap, int32_t *need_new) {
  const int64_t np = min((int64_t)*npairs, cap);
  for (int64_t p = (int64_t)blockIdx.x * 256 + threadIdx.x; p < np; p += (int64_t)gridDim.x * 256)
    need_new[p] = resolve_existing(P, S, pairs[2 * p], pairs[2 * p + 1]);
}
__global__ __launch_bounds__(256) void k_dem_resolve_new(const lbdem_dem_params P, const lbdem_dem_state S,
                                                         const int32_t *pairs, const unsigned long long *npairs,
                                                         int64_t cap, const int32_t *need_new) {
  const int64_t np = min((int64_t)*npairs, cap);
  for (int64_t p = (int64_t)blockIdx.x * 256 + threadIdx.x; p < np; p += (int64_t)gridDim.x * 256)
    if (need_new[p]) resolve_new(P, S, pairs[2 * p], pairs[2 * p + 1]);
}
// walls + integration and the compaction of particle i's history slots (independent: the
// slots are not read by the integration), one launch; the pair list is emptied for the next
// substep after its length is checked against the capacity
__global__ __launch_bounds__(256) void k_dem_integrate(const lbdem_dem_params P, const lbdem_dem_state S,
                                                       unsigned long long *npairs, int64_t cap) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i == 0) {
    const unsigned long long m = *npairs;
    if (m > (unsigned long long)cap) atomicMax(S.status + 4, (int)min(m, 0x7fffffffull));
    *npairs = 0ull;
  }
  if (i >= P.n) return;
  const unsigned w0 = load_wall_mask(S, i);
  unsigned w = w0;
  walls_and_integrate(P, S, i, w);
  store_wall_mask(S, i, w0, w);
  compact_slots(P, S, i);
}

// --------------------------------------------------------------------------------------
// small-system path: one workgroup, `nsub` substeps per launch
// --------------------------------------------------------------------------------------
// The whole kernel is one workgroup, so __syncthreads() (a workgroup-scope fence plus barrier)
// orders the phases' global-memory accesses; no device-scope fence (L2 write-back) is needed.
// Candidate pairs come from bucket lists kept in LDS: particle j is pushed onto the list of
// slot (hash_j & (kSmallSlots - 1)); particle i walks the lists of its distinct probe hashes and
// keeps the j > i whose own hash equals the probed one.  That is the candidate rule of the hash
// table (j's bucket is one of i's probed buckets) without the 2^20..2^22-entry table.
// With RESIDENT (n <= kResidentN, one thread per particle) the per-particle kinematic state
// lives in LDS for the whole launch: the state view's pointers are redirected to LDS copies (flat addressing), so the
// contact, wall and integration code is shared with the other paths; contact history stays in
// global memory.
// The resident kernel also keeps a Verlet list: each particle's j > i within
// r_i + r_j + skin, found with the 27-cell probe (which covers every distance below one cell),
// skin = min(cell - 2 r_max, r_max).  The list is rebuilt when some particle has moved more than
// skin / 2 since the last build, so every pair that touches, and every pair holding a history
// slot, is in the list: the contacts found are those of the per-substep hash probe.
constexpr int kSmallSlots = 4096;
constexpr int kResidentN = 256;
constexpr int kNbr = 16;  // Verlet neighbours kept per particle; more -> per-substep probe instead

__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
  return v;
}

template <bool RESIDENT, int NT>
__global__ __launch_bounds__(NT) void k_dem_small(const lbdem_dem_params P, const lbdem_dem_state G, int nsub,
                                                    int32_t *newlist, int newcap, int32_t *vl_g, double *vl_xg) {
  __shared__ int32_t head[kSmallSlots];
  __shared__ int32_t next[kSmallN];
  __shared__ int32_t chash[kSmallN];
  __shared__ int32_t ccell[3 * (RESIDENT ? kResidentN : kSmallN)];  // cell coordinates of phase A
  __shared__ int nnew;
  const int n = P.n;
  lbdem_dem_state S = G;
  // copied to LDS: pos, vel, acc, force, omega, omega_dot, moment, quat(4), mass, radius, cfn
  // (read and written), inertia, force_fluid, moment_fluid, freeze, fixvel (read only)
  constexpr int kRes = RESIDENT ? kResidentN : 1;
  __shared__ double r_vec[7][3 * kRes];
  __shared__ double r_quat[4 * kRes];
  __shared__ double r_mass[kRes], r_rad[kRes];
  __shared__ int32_t r_cfn[kRes];
  __shared__ double r_inertia[9 * kRes];
  __shared__ double r_fluid[2][3 * kRes];
  __shared__ int32_t r_freeze[kRes], r_fixvel[kRes];
  __shared__ uint32_t r_wact[kRes];  // wall-contact bits (resident kernel)
  __shared__ int32_t nb_cnt[kRes], nb_list[kNbr * kRes];  // Verlet lists (resident kernel)
  __shared__ double x_ref[3 * kRes];                       // positions at the last list build
  __shared__ double vl_skin;
  // list state; each is written only in a phase where no thread reads it before the next barrier:
  // vl_rebuild (set in the integration phase, cleared by thread 0 after the pair phase),
  // vl_off (lists unusable for this launch: no skin or a list overflowed), vl_over (build phase)
  __shared__ int vl_rebuild, vl_off, vl_over, vl_load;
  __shared__ double rmax_w[NT / 64];
  static_assert(!RESIDENT || NT >= kResidentN, "the resident kernel runs one thread per particle");
  // Verlet lists of the previous launch, loaded with the state (used only if still valid)
  unsigned wact_in = 0u;
  int32_t vl_cnt_in = 0, vl_nb_in[RESIDENT ? kNbr : 1], vl_valid_in = 0, vl_n_in = 0;
  double vl_x_in[3], vl_skin_in = 0.0;
  if constexpr (RESIDENT) {
    // thread t < n copies particle t (n <= NT): every load is issued before the first LDS
    // store, so the whole copy costs one memory round trip
    const int t = threadIdx.x;
    if (t < n) {
      const double *gv[7] = {G.pos, G.vel, G.acc, G.force, G.omega, G.omega_dot, G.moment};
      double v[7][3], q4[4], in9[9], fl[2][3];
#pragma unroll
      for (int a = 0; a < 7; ++a)
#pragma unroll
        for (int k = 0; k < 3; ++k) v[a][k] = gv[a][3 * t + k];
#pragma unroll
      for (int k = 0; k < 4; ++k) q4[k] = G.quat[4 * t + k];
#pragma unroll
      for (int k = 0; k < 9; ++k) in9[k] = G.inertia[9 * t + k];
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        fl[0][k] = G.force_fluid[3 * t + k];
        fl[1][k] = G.moment_fluid[3 * t + k];
      }
      wact_in = load_wall_mask(G, t);
      const double mass = G.mass[t], rad = G.radius[t];
      const int32_t cfn = G.cfn[t], frz = G.freeze ? G.freeze[t] : 0, fxv = G.fixvel ? G.fixvel[t] : 0;
      if (vl_g) {
        if (t == 0) {
          vl_valid_in = vl_g[0];
          vl_n_in = vl_g[1];
          vl_skin_in = vl_xg[0];
        }
        vl_cnt_in = vl_g[2 + t];
#pragma unroll
        for (int m = 0; m < kNbr; ++m) vl_nb_in[m] = vl_g[2 + n + kNbr * t + m];
#pragma unroll
        for (int k = 0; k < 3; ++k) vl_x_in[k] = vl_xg[1 + 3 * t + k];
      }
#pragma unroll
      for (int a = 0; a < 7; ++a)
#pragma unroll
        for (int k = 0; k < 3; ++k) r_vec[a][3 * t + k] = v[a][k];
#pragma unroll
      for (int k = 0; k < 4; ++k) r_quat[4 * t + k] = q4[k];
#pragma unroll
      for (int k = 0; k < 9; ++k) r_inertia[9 * t + k] = in9[k];
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        r_fluid[0][3 * t + k] = fl[0][k];
        r_fluid[1][3 * t + k] = fl[1][k];
      }
      r_wact[t] = wact_in;
      r_mass[t] = mass;
      r_rad[t] = rad;
      r_cfn[t] = cfn;
      r_freeze[t] = frz;
      r_fixvel[t] = fxv;
    }
    S.pos = r_vec[0];
    S.vel = r_vec[1];
    S.acc = r_vec[2];
    S.force = r_vec[3];
    S.omega = r_vec[4];
    S.omega_dot = r_vec[5];
    S.moment = r_vec[6];
    S.quat = r_quat;
    S.mass = r_mass;
    S.radius = r_rad;
    S.cfn = r_cfn;
    S.inertia = r_inertia;
    S.force_fluid = r_fluid[0];
    S.moment_fluid = r_fluid[1];
    if (G.freeze) S.freeze = r_freeze;
    if (G.fixvel) S.fixvel = r_fixvel;
  }
  // The hot broad-phase loops below read these views directly: in the resident kernel the
  // compiler then knows they are LDS and issues LDS loads, instead of the generic (flat)
  // accesses that the shared state view needs.
  const double *posv = RESIDENT ? &r_vec[0][0] : G.pos;
  const double *radv = RESIDENT ? r_rad : G.radius;
  const int32_t *cfnv = RESIDENT ? r_cfn : G.cfn;
  double *forcev = RESIDENT ? &r_vec[3][0] : G.force;
  double *momv = RESIDENT ? &r_vec[6][0] : G.moment;
  for (int t = threadIdx.x; t < kSmallSlots; t += NT) head[t] = -1;
  if constexpr (RESIDENT) {
    double rm = 0.0;
    for (int t = threadIdx.x; t < n; t += NT) rm = fmax(rm, r_rad[t]);
    rm = wave_max(rm);
    if ((threadIdx.x & 63) == 0) rmax_w[threadIdx.x >> 6] = rm;
    __syncthreads();
    if (threadIdx.x == 0) {
      double r = 0.0;
      for (int w = 0; w < NT / 64; ++w) r = fmax(r, rmax_w[w]);
      vl_skin = fmin(P.cell - 2.0 * r, r);
      vl_off = vl_skin > 0.0 ? 0 : 1;
      vl_over = 0;
      // lists kept by the previous launch: valid flag, particle count and skin must match
      vl_load = (!vl_off && vl_g && vl_valid_in == 1 && vl_n_in == n && vl_skin_in == vl_skin) ? 1 : 0;
      vl_rebuild = vl_load ? 0 : 1;
    }
    __syncthreads();
    if (vl_load) {
      const int i = threadIdx.x;
      if (i < n) {
        nb_cnt[i] = vl_cnt_in;
#pragma unroll
        for (int m = 0; m < kNbr; ++m) nb_list[kNbr * i + m] = vl_nb_in[m];
#pragma unroll
        for (int k = 0; k < 3; ++k) x_ref[3 * i + k] = vl_x_in[k];
        // positions changed between launches (by the caller) by more than half the skin: rebuild
        double d[3];
        periodic_delta(P, vl_x_in, posv + 3 * i, d);
        const double hs = 0.5 * vl_skin;
        if (d[0] * d[0] + d[1] * d[1] + d[2] * d[2] > hs * hs) atomicOr(&vl_rebuild, 1);
      }
    }
  }
  // clear_state of the first substep (later ones are cleared at the end of the one before)
  for (int i = threadIdx.x; i < n; i += NT) {
#pragma unroll
    for (int a = 0; a < 3; ++a) forcev[3 * i + a] = momv[3 * i + a] = 0.0;
  }
  if (threadIdx.x == 0) nnew = 0;
  __syncthreads();
  for (int sub = 0; sub < nsub; ++sub) {
    // bucket lists are needed by a Verlet-list build and by the per-substep probe
    bool pushed = true;
    if constexpr (RESIDENT) pushed = vl_rebuild || vl_off;  // uniform: stable in this phase
    // Barriers: after the bucket push (only when the lists are needed), after the pair phase,
    // after the new contacts (only if there are any) and after the integration.  clear_state
    // (forces, moments) and the new-contact count of a substep are reset at the end of the
    // previous one, or before the loop.
    if (pushed) {  // uniform: own cell hash, push onto the bucket list
      for (int i = threadIdx.x; i < n; i += NT) {
        int32_t c[3];
        cell_of(P, posv + 3 * i, c);
        const int32_t h = hash3(c[0], c[1], c[2], P.T);
        chash[i] = h;
#pragma unroll
        for (int a = 0; a < 3; ++a) ccell[3 * i + a] = c[a];
        next[i] = atomicExch(&head[h & (kSmallSlots - 1)], i);
      }
      __syncthreads();
    }
    // candidates of i; contacts already in a history slot are resolved here, new ones listed
    // (the probe count is a template argument of the walk so the hash array stays in registers)
    auto walk = [&](auto np) {
      constexpr int NP = decltype(np)::value;
      for (int i = threadIdx.x; i < n; i += NT) {
        const double *xi = posv + 3 * i;
        const int32_t ci[3] = {ccell[3 * i], ccell[3 * i + 1], ccell[3 * i + 2]};
        int32_t h[NP];
        probe_cells<NP>(P, xi, ci, h);
        const double ri = radv[i];
        const bool has_slots = cfnv[i] > 0;
#pragma unroll
        for (int k = 0; k < NP; ++k) {
          if (h[k] < 0) continue;
          for (int j = head[h[k] & (kSmallSlots - 1)]; j >= 0; j = next[j]) {
            if (j <= i || chash[j] != h[k]) continue;
            if (!has_slots && !touching(P, xi, ri, posv + 3 * j, radv[j])) continue;  // nothing to update
            if (resolve_existing(P, S, i, j)) {
              const int slot = atomicAdd(&nnew, 1);
              if (slot < newcap) {
                newlist[2 * slot] = i;
                newlist[2 * slot + 1] = j;
              }
            }
          }
        }
      }
    };
    // one pair (i, j > i) of a candidate list: touching pairs and pairs with history are resolved
    auto visit = [&](int i, const double *xi, double ri, bool has_slots, int j) {
      if (!has_slots && !touching(P, xi, ri, posv + 3 * j, radv[j])) return;  // nothing to update
      if (resolve_existing(P, S, i, j)) {
        const int slot = atomicAdd(&nnew, 1);
        if (slot < newcap) {
          newlist[2 * slot] = i;
          newlist[2 * slot + 1] = j;
        }
      }
    };
    bool use_lists = false;
    bool rebuilt = false;
    if constexpr (RESIDENT) {
      rebuilt = vl_rebuild && !vl_off;  // uniform: neither changes during this phase
      if (rebuilt) {  // (re)build the Verlet lists with the 27-cell probe
        const double skin = vl_skin;
        for (int i = threadIdx.x; i < n; i += NT) {
          const double *xi = posv + 3 * i;
          const int32_t ci[3] = {ccell[3 * i], ccell[3 * i + 1], ccell[3 * i + 2]};
          const double ri = radv[i];
          int cnt = 0;
          // the 27 cells around i; j is taken only from the list walk of its own cell, so a
          // bucket shared by several cells (hash aliasing) never yields a pair twice
          for (int t = 0; t < 27; ++t) {
            const int32_t cx = ci[0] + t / 9 - 1, cy = ci[1] + (t / 3) % 3 - 1, cz = ci[2] + t % 3 - 1;
            const int32_t hc = hash3(cx, cy, cz, P.T);
            for (int j = head[hc & (kSmallSlots - 1)]; j >= 0; j = next[j]) {
              if (j <= i || ccell[3 * j] != cx || ccell[3 * j + 1] != cy || ccell[3 * j + 2] != cz) continue;
              double d[3];
              periodic_delta(P, xi, posv + 3 * j, d);
              if (norm3(d) - ri - radv[j] < skin) {
                if (cnt < kNbr) nb_list[kNbr * i + cnt] = j;
                ++cnt;
              }
            }
          }
          nb_cnt[i] = cnt;
#pragma unroll
          for (int a = 0; a < 3; ++a) x_ref[3 * i + a] = xi[a];
          if (cnt > kNbr) atomicOr(&vl_over, 1);
        }
        __syncthreads();
        // every thread has read vl_rebuild for this substep; the integration phase may set it
        // again, after the pair phase's barrier.  vl_off turning 1 here leaves use_lists false
        // either way (vl_over is 1).
        if (threadIdx.x == 0) {
          vl_rebuild = 0;  // lists are current
          if (vl_over) vl_off = 1;
        }
      }
      use_lists = !vl_off && !vl_over;  // vl_over is final after the build barrier
    }
    if (use_lists) {
      for (int i = threadIdx.x; i < n; i += NT) {
        const double *xi = posv + 3 * i;
        const double ri = radv[i];
        const bool has_slots = cfnv[i] > 0;
        const int cnt = nb_cnt[i];
        for (int m = 0; m < cnt; ++m) visit(i, xi, ri, has_slots, nb_list[kNbr * i + m]);
      }
    } else if (P.probe27) {
      walk(std::integral_constant<int, 27>{});
    } else {
      walk(std::integral_constant<int, 8>{});
    }
    __syncthreads();
    const int nn = min(nnew, newcap);  // uniform: nnew is reset after this phase's reads
    if (threadIdx.x == 0 && nnew > newcap) atomicAdd(S.status + 0, nnew - newcap);
    for (int t = threadIdx.x; t < nn; t += NT) resolve_new(P, S, newlist[2 * t], newlist[2 * t + 1]);
    if (pushed)
      for (int t = threadIdx.x; t < kSmallSlots; t += NT) head[t] = -1;  // lists are rebuilt next substep
    if (nn > 0) __syncthreads();  // new contacts add forces that the integration reads
    // integration and slot compaction touch particle i's own data only
    for (int i = threadIdx.x; i < n; i += NT) {
      if constexpr (RESIDENT) {
        walls_and_integrate(P, S, i, r_wact[i]);
      } else {
        const unsigned w0 = load_wall_mask(S, i);
        unsigned w = w0;
        walls_and_integrate(P, S, i, w);
        store_wall_mask(S, i, w0, w);
      }
      compact_slots(P, S, i);
      if constexpr (RESIDENT) {
        if (!vl_off) {  // moved more than half the skin since the build: rebuild next substep
          double d[3];
          periodic_delta(P, x_ref + 3 * i, posv + 3 * i, d);
          const double hs = 0.5 * vl_skin;
          if (d[0] * d[0] + d[1] * d[1] + d[2] * d[2] > hs * hs) atomicOr(&vl_rebuild, 1);
        }
      }
      if (sub + 1 < nsub) {  // clear_state of the next substep; the last one's forces are kept
#pragma unroll
        for (int a = 0; a < 3; ++a) forcev[3 * i + a] = momv[3 * i + a] = 0.0;
      }
    }
    // every thread read nnew before this phase (it is 0 unless the new-contact barrier ran)
    if (threadIdx.x == 0) nnew = 0;
    __syncthreads();
  }
  if constexpr (RESIDENT) {
    const int t = threadIdx.x;
    if (t < n) {  // particle t's state back to global memory (the caller's arrays)
      double *gv[7] = {G.pos, G.vel, G.acc, G.force, G.omega, G.omega_dot, G.moment};
#pragma unroll
      for (int a = 0; a < 7; ++a)
#pragma unroll
        for (int k = 0; k < 3; ++k) gv[a][3 * t + k] = r_vec[a][3 * t + k];
#pragma unroll
      for (int k = 0; k < 4; ++k) G.quat[4 * t + k] = r_quat[4 * t + k];
      G.cfn[t] = r_cfn[t];
      store_wall_mask(G, t, wact_in, r_wact[t]);
    }
    if (vl_g) {  // keep the Verlet lists for the next launch
      const bool keep = !vl_off && !vl_rebuild;
      if (keep && t < n) {
        vl_g[2 + t] = nb_cnt[t];
#pragma unroll
        for (int m = 0; m < kNbr; ++m) vl_g[2 + n + kNbr * t + m] = nb_list[kNbr * t + m];
#pragma unroll
        for (int k = 0; k < 3; ++k) vl_xg[1 + 3 * t + k] = x_ref[3 * t + k];
      }
      if (threadIdx.x == 0) {
        vl_g[0] = keep ? 1 : 0;
        vl_g[1] = n;
        vl_xg[0] = vl_skin;
      }
    }
  }
}

static int check_params(const lbdem_dem_params *P, const char *who) {
  LBDEM_REQUIRE(P, "%s: null params", who);
  LBDEM_REQUIRE(P->n >= 0, "%s: negative particle count", who);
  LBDEM_REQUIRE(P->T > 0 && (P->T & (P->T - 1)) == 0, "%s: hash table size %d is not a power of two", who, P->T);
  LBDEM_REQUIRE(P->cell > 0.0, "%s: cell size must be positive", who);
  LBDEM_REQUIRE(P->K >= 0, "%s: max_coordinate_number must not be negative", who);
  LBDEM_REQUIRE(P->model == 0 || P->model == 1, "%s: model must be 0 (Hertz) or 1 (linear)", who);
  return LBDEM_OK;
}

static int run_broadphase(const lbdem_dem_params *P, const double *pos, const lbdem_dem_work *w, double *force,
                          double *moment, hipStream_t s, const char *who, const double *rad_touch = nullptr) {
  const int n = P->n;
  const unsigned nb = (unsigned)((n + 255) / 256);
  const bool fine = rad_touch && P->fine_cell > 0.0;
  // the bucket slots live in spos until k_bp_sort_state writes the records there (without spos
  // the scatter counts again, in `current`)
  int32_t *slot = w->spos ? reinterpret_cast<int32_t *>(w->spos) : nullptr;
  // the touching search of lbdem_dem_run leaves the counters to k_dem_release
  const bool reset = rad_touch == nullptr;
  // with the second grid the main grid needs its hashes only (the release kernel and the
  // candidate filter read them); the second grid is built in the same table
  hipLaunchKernelGGL(k_bp_count, dim3(nb), dim3(256), 0, s, *P, pos, w->hash, fine ? nullptr : w->count, force,
                     moment, fine ? nullptr : slot);
  if (fine) {
    lbdem_dem_params Pf = *P;
    Pf.cell = P->fine_cell;
    hipLaunchKernelGGL(k_bp_count, dim3(nb), dim3(256), 0, s, Pf, pos, w->fhash, w->count, nullptr, nullptr, slot);
    scan_table(w, P->T, s);
    if (slot)
      hipLaunchKernelGGL(k_bp_scatter_slots, dim3(nb), dim3(256), 0, s, Pf, w->fhash, w->offset, slot, w->pid);
    else
      hipLaunchKernelGGL(k_bp_scatter, dim3(nb), dim3(256), 0, s, Pf, w->fhash, w->offset, w->current, w->pid);
    hipLaunchKernelGGL(k_bp_sort_state, dim3(nb), dim3(256), 0, s, Pf, pos, rad_touch, w->pid, w->spos);
    const unsigned split = (unsigned)std::min(16, std::max(1, (int)((2048 + nb - 1) / nb)));
    hipLaunchKernelGGL(k_bp_probe_touch_fine, dim3(nb, std::min(split, 27u)), dim3(256), 0, s, *P, Pf, w->count,
                       w->offset, w->pid, reinterpret_cast<const double4 *>(w->spos), w->hash, w->pairs,
                       w->pair_cap, w->npairs);
    if (P->nlarge > 0)
      hipLaunchKernelGGL(k_bp_touch_large, dim3(nb, (unsigned)P->nlarge), dim3(256), 0, s, *P, pos, rad_touch,
                         w->hash, w->large, w->pairs, w->pair_cap, w->npairs);
    if (reset) hipLaunchKernelGGL(k_bp_reset, dim3(nb), dim3(256), 0, s, Pf, w->fhash, w->count, w->current);
    LBDEM_CHECK_LAUNCH(who);
    return LBDEM_OK;
  }
  scan_table(w, P->T, s);
  if (slot)
    hipLaunchKernelGGL(k_bp_scatter_slots, dim3(nb), dim3(256), 0, s, *P, w->hash, w->offset, slot, w->pid);
  else
    hipLaunchKernelGGL(k_bp_scatter, dim3(nb), dim3(256), 0, s, *P, w->hash, w->offset, w->current, w->pid);
  if (rad_touch) {
    // touching candidates only; each particle's candidates are split over more lanes when
    // there are fewer than 2048 workgroups (more waves in flight; also measured faster on the
    // sparse 200,000-particle bed, 0.72 against 0.84 ms with no split)
    const unsigned split = (unsigned)std::min(16, std::max(1, (int)((2048 + nb - 1) / nb)));
    // the 27-cell probe (fine grids) walks particles in bucket order (200,000-grain bed: 0.102
    // ms per substep with z slabs and the headers first, 0.116 ms with the buckets split k % 3
    // and walked one after another, 0.20 ms with a per-entry split, 0.54 ms in particle order;
    // 9 and 27 workgroups per particle measured slower than 3; a walk of each
    // wavefront's box of cells together, one candidate per lane and readlane, measured 0.47 ms:
    // waves at jumps of the Morton order get boxes of hundreds of cells and set the launch's
    // tail).  The 8-cell probe serves coarse grids, whose workgroups mostly take its LDS-tiled
    // walk; bucket order measured slower there (0.70 -> 0.86 ms, 37,464 grains)
    const double4 *sp = P->probe27 ? reinterpret_cast<const double4 *>(w->spos) : nullptr;
    if (sp) {
      hipLaunchKernelGGL(k_bp_sort_state, dim3(nb), dim3(256), 0, s, *P, pos, rad_touch, w->pid, w->spos);
      if (P->T >= 256)
        hipLaunchKernelGGL(k_bp_probe_touch_sorted<true>, dim3(nb, 3u), dim3(256), 0, s, *P, w->count,
                           w->offset, w->pid, sp, w->pairs, w->pair_cap, w->npairs);
      else
        hipLaunchKernelGGL(k_bp_probe_touch_sorted<false>, dim3(nb, 3u), dim3(256), 0, s, *P, w->count,
                           w->offset, w->pid, sp, w->pairs, w->pair_cap, w->npairs);
    } else if (P->probe27)
      hipLaunchKernelGGL(k_bp_probe_touch<27>, dim3(nb, split), dim3(256), 0, s, *P, pos, rad_touch, w->count,
                         w->offset, w->pid, w->pairs, w->pair_cap, w->npairs);
    else
      hipLaunchKernelGGL(k_bp_probe_touch<8>, dim3(nb, split), dim3(256), 0, s, *P, pos, rad_touch, w->count,
                         w->offset, w->pid, w->pairs, w->pair_cap, w->npairs);
  } else if (P->probe27) {
    hipLaunchKernelGGL(k_bp_probe<27>, dim3(nb), dim3(256), 0, s, *P, pos, w->count, w->offset, w->pid, w->pairs,
                       w->pair_cap, w->npairs);
  } else {
    hipLaunchKernelGGL(k_bp_probe<8>, dim3(nb), dim3(256), 0, s, *P, pos, w->count, w->offset, w->pid, w->pairs,
                       w->pair_cap, w->npairs);
  }
  if (reset) hipLaunchKernelGGL(k_bp_reset, dim3(nb), dim3(256), 0, s, *P, w->hash, w->count, w->current);
  LBDEM_CHECK_LAUNCH(who);
  return LBDEM_OK;
}

}  // namespace lbdem

using namespace lbdem;

extern "C" {

int lbdem_dem_broadphase(const lbdem_dem_params *P, const double *pos, const lbdem_dem_work *w, void *stream) {
  int rc = check_params(P, "lbdem_dem_broadphase");
  if (rc) return rc;
  LBDEM_REQUIRE(w && w->count && w->offset && w->current && w->hash && w->pid && w->tile_sum && w->npairs &&
                    (w->pairs || w->pair_cap == 0),
                "lbdem_dem_broadphase: incomplete workspace");
  LBDEM_REQUIRE(P->n == 0 || pos, "lbdem_dem_broadphase: pos is required");
  if (P->n == 0) return LBDEM_OK;
  return run_broadphase(P, pos, w, nullptr, nullptr, as_stream(stream), "lbdem_dem_broadphase");
}

int lbdem_dem_run(const lbdem_dem_params *P, const lbdem_dem_state *S, const lbdem_dem_work *w, int32_t nsub,
                  int32_t path, void *stream) {
  int rc = check_params(P, "lbdem_dem_run");
  if (rc) return rc;
  LBDEM_REQUIRE(S && w, "lbdem_dem_run: null state or workspace");
  LBDEM_REQUIRE(nsub >= 0, "lbdem_dem_run: negative substep count");
  LBDEM_REQUIRE(S->status, "lbdem_dem_run: status buffer is required");
  const int n = P->n;
  if (n == 0 || nsub == 0) return LBDEM_OK;
  hipStream_t s = as_stream(stream);
  const bool small = (path == 1) || (path == 0 && n <= kSmallN);
  if (small) {
    LBDEM_REQUIRE(n <= kSmallN, "lbdem_dem_run: the single-workgroup path takes at most %d particles", kSmallN);
    LBDEM_REQUIRE(w->pairs && w->pair_cap > 0, "lbdem_dem_run: small path needs the pairs buffer");
    const int newcap = (int)min(w->pair_cap, (int64_t)0x7fffffff);
    if (n <= kResidentN)
      hipLaunchKernelGGL((k_dem_small<true, kResidentN>), dim3(1), dim3(kResidentN), 0, s, *P, *S, nsub, w->pairs,
                         newcap, w->verlet, w->verlet_ref);
    else
      hipLaunchKernelGGL((k_dem_small<false, 512>), dim3(1), dim3(512), 0, s, *P, *S, nsub, w->pairs, newcap,
                         nullptr, nullptr);
    LBDEM_CHECK_LAUNCH("lbdem_dem_run(small)");
    return LBDEM_OK;
  }
  LBDEM_REQUIRE(w->need_new && w->pairs && w->pair_cap > 0, "lbdem_dem_run: need_new and pairs are required");
  LBDEM_REQUIRE(((uintptr_t)w->spos & 31u) == 0, "lbdem_dem_run: spos must be 32-byte aligned");
  if (P->fine_cell > 0.0) {
    LBDEM_REQUIRE(!P->probe27 && !P->periodic[0] && !P->periodic[1] && !P->periodic[2],
                  "lbdem_dem_run: fine_cell needs the 8-cell probe and no periodic axis");
    LBDEM_REQUIRE(P->fine_cell > 2.0 * P->large_r && P->large_r > 0.0,
                  "lbdem_dem_run: fine_cell must exceed 2 large_r > 0");
    LBDEM_REQUIRE(w->fhash && w->spos && P->nlarge >= 0 && P->nlarge <= 65535 && (P->nlarge == 0 || w->large),
                  "lbdem_dem_run: fine_cell needs fhash, spos and the large-particle list (at most 65535)");
  }
  const unsigned nb = (unsigned)((n + 255) / 256);
  const unsigned pb = 2048;  // grid-stride over the candidate list
  // the table's counters, cleared by the release kernel: the second grid's when it is in use;
  // `current` only when the scatter counted in it (no spos)
  const int32_t *rhash = P->fine_cell > 0.0 ? w->fhash : w->hash;
  int32_t *cur = w->spos ? nullptr : w->current;
  for (int sub = 0; sub < nsub; ++sub) {
    rc = run_broadphase(P, S->pos, w, S->force, S->moment, s, "lbdem_dem_run", S->radius);
    if (rc) return rc;
    if (P->probe27)
      hipLaunchKernelGGL(k_dem_release<27>, dim3(nb), dim3(256), 0, s, *P, *S, w->hash, rhash, w->count, cur);
    else
      hipLaunchKernelGGL(k_dem_release<8>, dim3(nb), dim3(256), 0, s, *P, *S, w->hash, rhash, w->count, cur);
    hipLaunchKernelGGL(k_dem_resolve_existing, dim3(pb), dim3(256), 0, s, *P, *S, w->pairs, w->npairs, w->pair_cap,
                       w->need_new);
    hipLaunchKernelGGL(k_dem_resolve_new, dim3(pb), dim3(256), 0, s, *P, *S, w->pairs, w->npairs, w->pair_cap,
                       w->need_new);
    hipLaunchKernelGGL(k_dem_integrate, dim3(nb), dim3(256), 0, s, *P, *S, w->npairs, w->pair_cap);
    LBDEM_CHECK_LAUNCH("lbdem_dem_run");
  }
  return LBDEM_OK;
}

}  // extern "C"
