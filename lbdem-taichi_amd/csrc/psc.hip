// Partially saturated cells (resolved LBM-DEM coupling): grain -> lattice mapping and the
// per-grain FP64 hydrodynamic force / torque reduction.  The weighted collision itself is
// folded into the collide-stream kernel (coupling LBDEM_COUPLE_PSC).
//
// Mapping is grain-centric: the workgroups of a grain walk the grain's bounding box and
// publishes, per covered cell, a 32-bit key with atomicMax:
//     bit 30      full coverage (dist <= r - sqrt(3)/2)
//     bits 29..23 sub-cell count (1..125) of a partially covered cell
//     bits 22..0  0x7FFFFF - grain index
// so the maximum key is exactly the reference's serial rule (cell loop over grains in
// index order, psclattice3d.py:95-138): the lowest-index full cover wins; otherwise the
// largest count, ties to the lowest index.
#include "common.hpp"

namespace lbdem {

constexpr uint32_t kFullBit = 1u << 30;
constexpr int kMaxGrain = 0x7FFFFF;

struct PscArgs {
  int nx, ny, nz, ngrain;
  const uint8_t *flags;
  const double *pos, *vel, *omg, *rad;
  double dx, dt, rho, dmin0, dmin1, dmin2;
  double tau0;
  uint32_t *key;
  float *snap;   // 12 per grain
  int *bbox;     // 6 per grain: i0, i1, j0, j1, k0, k1 (half-open)
  const float *cell_force;
  double *force_fluid, *moment_fluid;
  uint32_t *stale_key;     // keys of an earlier mapping, zeroed by the force kernel (may be null)
  const int *stale_bbox;   // that mapping's boxes
  int coarse_only;         // hybrid coupling: grains with 2R <= dx are not mapped (empty box)
  const int32_t *glist;    // grain of each workgroup column (null: blockIdx.x itself)
};

__device__ __forceinline__ int grain_of_block(const PscArgs &a) {
  return a.glist ? (int)a.glist[blockIdx.x] : (int)blockIdx.x;
}

// bounding box of lattice2grains (psclattice3d.py:326-337): [int(xc - r), int(xc + r + 2))
// with r = R/dx, clamped; it contains every cell the mapping can touch.
__device__ __forceinline__ void grain_box(const PscArgs &a, int g, double &xc, double &yc, double &zc, int b[6]) {
  xc = (a.pos[3 * g] - a.dmin0 + 0.5 * a.dx) / a.dx;
  yc = (a.pos[3 * g + 1] - a.dmin1 + 0.5 * a.dx) / a.dx;
  zc = (a.pos[3 * g + 2] - a.dmin2 + 0.5 * a.dx) / a.dx;
  const double r = a.rad[g] / a.dx;
  b[0] = max((int)(xc - r), 0);
  b[1] = min((int)(xc + r + 2), a.nx);
  b[2] = max((int)(yc - r), 0);
  b[3] = min((int)(yc + r + 2), a.ny);
  b[4] = max((int)(zc - r), 0);
  b[5] = min((int)(zc + r + 2), a.nz);
}

// Each grain's box is shared by gridDim.y workgroups (blockIdx.y): kSplit, or more when there
// are fewer than 128 grains (a handful of large grains still fills the chip; with 200 grains of
// R = 8 the larger split measured slower, 17.6 against 16.4 us).
constexpr int kSplit = 8;
inline int split_for(int ngrain) { return ngrain >= 128 ? kSplit : std::min(64, std::max(kSplit, 2048 / ngrain)); }
constexpr int kMapChunk = 1024;  // cells classified per workgroup between two counting passes

// min { x >= 0 : sqrt(x) >= r } for the correctly rounded sqrt: sqrt(x) < r <=> x < bound,
// which replaces one FP64 square root per sub-cell sample with a compare.
__device__ __forceinline__ double sqrt_lt_bound(double r) {
  if (!(r > 0.0)) return 0.0;
  double t = r * r;
  while (t > 0.0 && sqrt(t) >= r) t = nextafter(t, 0.0);
  while (sqrt(t) < r) t = nextafter(t, HUGE_VAL);
  return t;
}

// zero the keys written by the previous mapping (its boxes are in bbox)
__global__ __launch_bounds__(256) void k_psc_clear(const PscArgs a) {
  const int g = grain_of_block(a);
  const int *b = a.bbox + 6 * g;
  const int ni = b[1] - b[0], nj = b[3] - b[2], nk = b[5] - b[4];
  if (ni <= 0 || nj <= 0 || nk <= 0) return;
  const int tot = ni * nj * nk;
  for (int t = blockIdx.y * 256 + threadIdx.x; t < tot; t += 256 * gridDim.y) {
    const int k = b[4] + t % nk, j = b[2] + (t / nk) % nj, i = b[0] + t / (nk * nj);
    a.key[(i * a.ny + j) * a.nz + k] = 0u;
  }
}

// grains2lattice (psclattice3d.py:84-182), grain-centric
__global__ __launch_bounds__(256) void k_psc_map(const PscArgs a) {
  const int g = grain_of_block(a);
  double xc, yc, zc;
  int b[6];
  grain_box(a, g, xc, yc, zc, b);
  // map_coarse_grains skips the fine grains (src/hybridlbdem/hybridlattice.py:197-199)
  if (a.coarse_only && 2.0 * a.rad[g] <= a.dx) b[1] = b[0];
  const double r = (a.rad[g] - 0.5 * a.dx) / a.dx;
  const double half_diag = 0.5 * sqrt(3.0);
  const double r2 = sqrt_lt_bound(r);
  if (threadIdx.x == 0 && blockIdx.y == 0) {
#pragma unroll
    for (int m = 0; m < 6; ++m) a.bbox[6 * g + m] = b[m];
    // snapshot for the collision: centre (lattice units), v*dt/dx, omega*dt
    float *s = a.snap + 12 * g;
    s[0] = (float)xc;
    s[1] = (float)yc;
    s[2] = (float)zc;
    const double vs = a.dt / a.dx;
#pragma unroll
    for (int m = 0; m < 3; ++m) {
      s[3 + m] = (float)(a.vel[3 * g + m] * vs);
      s[6 + m] = (float)(a.omg[3 * g + m] * a.dt);
      s[9 + m] = 0.f;
    }
  }
  const int ni = b[1] - b[0], nj = b[3] - b[2], nk = b[5] - b[4];
  if (ni <= 0 || nj <= 0 || nk <= 0) return;  // uniform over the workgroup
  const int tot = ni * nj * nk;
  const uint32_t gid = (uint32_t)(kMaxGrain - g);
  // Cells cut by the sphere surface need 125 sub-cell samples, the others one distance: in
  // chunks of kMapChunk cells, full and empty cells are settled at once and the cut cells are
  // queued in LDS, then counted one per lane with every lane busy (evaluated where they fall,
  // about three lanes in four would idle through each count).
  __shared__ int queue[kMapChunk];
  __shared__ int qn;
  const int split = gridDim.y;
  for (int m0 = 0; (m0 * split + (int)blockIdx.y) * 256 < tot; m0 += kMapChunk / 256) {
    if (threadIdx.x == 0) qn = 0;
    __syncthreads();
    for (int mm = 0; mm < kMapChunk / 256; ++mm) {
      const int t = ((m0 + mm) * split + (int)blockIdx.y) * 256 + (int)threadIdx.x;
      if (t >= tot) break;
      const int k = b[4] + t % nk, j = b[2] + (t / nk) % nj, i = b[0] + t / (nk * nj);
      const int c = (i * a.ny + j) * a.nz + k;
      if (!(a.flags[c] & LBDEM_FLAG_ACTIVE)) continue;
      const double dist = sqrt((xc - i) * (xc - i) + (yc - j) * (yc - j) + (zc - k) * (zc - k));
      if (dist >= r + half_diag) continue;
      if (dist <= r - half_diag) atomicMax(a.key + c, kFullBit | gid);
      else queue[atomicAdd(&qn, 1)] = t;
    }
    __syncthreads();
    const int nq = qn;
    for (int e = threadIdx.x; e < nq; e += 256) {
      const int t = queue[e];
      const int k = b[4] + t % nk, j = b[2] + (t / nk) % nj, i = b[0] + t / (nk * nj);
      int cnt = 0;
      for (int ii = 0; ii < 5; ++ii) {
        const double px = (double)i - 0.4 + 0.2 * ii;
        const double dxx = (xc - px) * (xc - px);
        for (int jj = 0; jj < 5; ++jj) {
          const double py = (double)j - 0.4 + 0.2 * jj;
          const double dyy = (yc - py) * (yc - py);
          for (int kk = 0; kk < 5; ++kk) {
            const double pz = (double)k - 0.4 + 0.2 * kk;
            cnt += (dxx + dyy + (zc - pz) * (zc - pz) < r2) ? 1 : 0;  // sqrt(.) < r
          }
        }
      }
      if (cnt > 0) atomicMax(a.key + (i * a.ny + j) * a.nz + k, ((uint32_t)cnt << 23) | gid);
    }
    __syncthreads();  // the queue is refilled by the next chunk
  }
}

constexpr int kForceThreads = 1024;

// lattice2grains (psclattice3d.py:316-358): one 1024-thread workgroup per grain, FP64, fixed
// reduction order (per-thread sums, wavefront shuffles, then the 16 wave sums pairwise in LDS),
// so the result is deterministic.
__global__ __launch_bounds__(kForceThreads) void k_psc_forces(const PscArgs a) {
  const int g = grain_of_block(a);
  double xc, yc, zc;
  int b[6];
  grain_box(a, g, xc, yc, zc, b);
  const double scale = a.rho * a.dx * a.dx * a.dx * a.dx / (a.dt * a.dt);
  double F[3] = {0, 0, 0}, T[3] = {0, 0, 0};
  const int ni = b[1] - b[0], nj = b[3] - b[2], nk = b[5] - b[4];
  const uint32_t gid = (uint32_t)(kMaxGrain - g);
  if (ni > 0 && nj > 0 && nk > 0) {
    const int tot = ni * nj * nk;
    for (int t = threadIdx.x; t < tot; t += kForceThreads) {
      const int k = b[4] + t % nk, j = b[2] + (t / nk) % nj, i = b[0] + t / (nk * nj);
      const int c = (i * a.ny + j) * a.nz + k;
      const uint32_t key = a.key[c];
      if (key == 0u || (key & 0x7FFFFFu) != gid) continue;
      const double f0 = (double)a.cell_force[3 * c] * scale, f1 = (double)a.cell_force[3 * c + 1] * scale,
                   f2 = (double)a.cell_force[3 * c + 2] * scale;
      const double d0 = (xc - i) * a.dx, d1 = (yc - j) * a.dx, d2 = (zc - k) * a.dx;
      F[0] += f0;
      F[1] += f1;
      F[2] += f2;
      T[0] -= d1 * f2 - d2 * f1;
      T[1] -= d2 * f0 - d0 * f2;
      T[2] -= d0 * f1 - d1 * f0;
    }
  }
  if (a.stale_key) {  // the other key buffer, for the next mapping: zero this grain's old box
    const int *sb = a.stale_bbox + 6 * g;
    const int si = sb[1] - sb[0], sj = sb[3] - sb[2], sk = sb[5] - sb[4];
    if (si > 0 && sj > 0 && sk > 0) {
      const int tot = si * sj * sk;
      for (int t = threadIdx.x; t < tot; t += kForceThreads) {
        const int k = sb[4] + t % sk, j = sb[2] + (t / sk) % sj, i = sb[0] + t / (sk * sj);
        a.stale_key[(i * a.ny + j) * a.nz + k] = 0u;
      }
    }
  }
  constexpr int kWaves = kForceThreads / 64;
  __shared__ double sh[6][kWaves];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
  for (int m = 0; m < 3; ++m) {
    const double fs = wave_sum(F[m]), ts = wave_sum(T[m]);
    if (lane == 0) {
      sh[m][wv] = fs;
      sh[3 + m][wv] = ts;
    }
  }
  __syncthreads();
  if (threadIdx.x < 6) {
    double v[kWaves];
#pragma unroll
    for (int w = 0; w < kWaves; ++w) v[w] = sh[threadIdx.x][w];
#pragma unroll
    for (int h = kWaves / 2; h > 0; h /= 2)
#pragma unroll
      for (int w = 0; w < h; ++w) v[w] += v[w + h];
    const int m = threadIdx.x;
    if (m < 3) a.force_fluid[3 * g + m] = v[0];
    else a.moment_fluid[3 * g + m - 3] = v[0];
  }
}

// id / volfrac / weight / velsolid fields decoded from the keys (for inspection)
__global__ __launch_bounds__(256) void k_psc_decode(const PscArgs a, int *id, float *volfrac, float *weight,
                                                    float *velsolid, int ncell) {
  const int c = blockIdx.x * 256 + threadIdx.x;
  if (c >= ncell) return;
  const uint32_t key = a.key[c];
  if (key == 0u) {
    id[c] = -1;
    volfrac[c] = 0.f;
    weight[c] = 0.f;
    velsolid[3 * c] = velsolid[3 * c + 1] = velsolid[3 * c + 2] = 0.f;
    return;
  }
  const int g = kMaxGrain - (int)(key & 0x7FFFFFu);
  const int nyz = a.ny * a.nz;
  const int i = c / nyz, j = (c - i * nyz) / a.nz, k = c - i * nyz - j * a.nz;
  id[c] = g;
  float eps, B;
  if (key & kFullBit) {
    eps = 1.f;
    B = 1.f;
  } else {
    eps = (float)((key >> 23) & 0x7Fu) * (1.0f / 125.0f);
    const float th = (float)a.tau0 - 0.5f;
    B = eps * th / ((1.f - eps) + th);
  }
  volfrac[c] = eps;
  weight[c] = B;
  const float *s = a.snap + 12 * g;
  const float rx = i - s[0], ry = j - s[1], rz = k - s[2];
  velsolid[3 * c] = s[3] + (s[7] * rz - s[8] * ry);
  velsolid[3 * c + 1] = s[4] + (s[8] * rx - s[6] * rz);
  velsolid[3 * c + 2] = s[5] + (s[6] * ry - s[7] * rx);
}

static int fill_psc(const lbdem_lattice *lat, const lbdem_grains *g, const lbdem_units *u, PscArgs &a,
                    const char *who) {
  LBDEM_REQUIRE(lat && g && u, "%s: null argument", who);
  LBDEM_REQUIRE(g->n >= 0 && g->n <= kMaxGrain + 1, "%s: %d grains (at most %d)", who, g->n, kMaxGrain + 1);
  LBDEM_REQUIRE(u->dx > 0 && u->dt > 0, "%s: dx and dt must be positive", who);
  a = PscArgs{};
  a.nx = lat->nx;
  a.ny = lat->ny;
  a.nz = lat->nz;
  a.ngrain = g->n;
  a.pos = g->position;
  a.vel = g->velocity;
  a.omg = g->omega;
  a.rad = g->radius;
  a.force_fluid = g->force_fluid;
  a.moment_fluid = g->moment_fluid;
  a.dx = u->dx;
  a.dt = u->dt;
  a.rho = u->rho;
  a.dmin0 = u->dmin[0];
  a.dmin1 = u->dmin[1];
  a.dmin2 = u->dmin[2];
  a.tau0 = 1.0 / lat->omega0;
  return LBDEM_OK;
}

int launch_psc_forces(const lbdem_lattice *lat, const lbdem_grains *g, const lbdem_units *u, const int32_t *glist,
                      int32_t nlist, const uint32_t *key, const float *cell_force, uint32_t *stale_key,
                      const int32_t *stale_bbox, void *stream, const char *who) {
  PscArgs a;
  int rc = fill_psc(lat, g, u, a, who);
  if (rc) return rc;
  LBDEM_REQUIRE(key && cell_force, "%s: key and cell_force are required", who);
  LBDEM_REQUIRE(g->n == 0 || (g->position && g->radius && g->force_fluid && g->moment_fluid),
                "%s: grain position, radius, force_fluid and moment_fluid are required", who);
  LBDEM_REQUIRE(!glist || (nlist >= 0 && nlist <= g->n), "%s: grain list of %d entries for %d grains", who, nlist,
                g->n);
  const int ncol = glist ? nlist : g->n;
  if (ncol == 0) return LBDEM_OK;
  LBDEM_REQUIRE(!stale_key || (stale_bbox && stale_key != key), "%s: stale_key needs stale_bbox and must differ "
                "from key", who);
  a.key = const_cast<uint32_t *>(key);
  a.cell_force = cell_force;
  a.stale_key = stale_key;
  a.stale_bbox = stale_bbox;
  a.glist = glist;
  hipLaunchKernelGGL(k_psc_forces, dim3(ncol), dim3(kForceThreads), 0, as_stream(stream), a);
  LBDEM_CHECK_LAUNCH(who);
  return LBDEM_OK;
}

}  // namespace lbdem

using namespace lbdem;

extern "C" {

static int psc_map(const lbdem_lattice *lat, const uint8_t *flags, const lbdem_grains *g, const lbdem_units *u,
                   uint32_t *key, float *snap, int32_t *bbox, int32_t clear_previous, void *stream, int coarse_only,
                   const int32_t *glist, int32_t nlist, const char *who) {
  PscArgs a;
  int rc = fill_psc(lat, g, u, a, who);
  if (rc) return rc;
  LBDEM_REQUIRE(flags && key && snap && bbox, "%s: flags, key, snap and bbox are required", who);
  LBDEM_REQUIRE(g->n == 0 || (g->position && g->velocity && g->omega && g->radius),
                "%s: grain position, velocity, omega and radius are required", who);
  LBDEM_REQUIRE(!glist || (nlist >= 0 && nlist <= g->n), "%s: grain list of %d entries for %d grains", who, nlist,
                g->n);
  a.coarse_only = coarse_only;
  a.glist = glist;
  const int ncol = glist ? nlist : g->n;
  a.flags = flags;
  a.key = key;
  a.snap = snap;
  a.bbox = bbox;
  if (ncol == 0) return LBDEM_OK;
  if (clear_previous) {
    hipLaunchKernelGGL(k_psc_clear, dim3(ncol, split_for(ncol)), dim3(256), 0, as_stream(stream), a);
    LBDEM_CHECK_LAUNCH(who);
  }
  hipLaunchKernelGGL(k_psc_map, dim3(ncol, split_for(ncol)), dim3(256), 0, as_stream(stream), a);
  LBDEM_CHECK_LAUNCH(who);
  return LBDEM_OK;
}

int lbdem_psc_map(const lbdem_lattice *lat, const uint8_t *flags, const lbdem_grains *g, const lbdem_units *u,
                  uint32_t *key, float *snap, int32_t *bbox, int32_t clear_previous, void *stream) {
  return psc_map(lat, flags, g, u, key, snap, bbox, clear_previous, stream, 0, nullptr, 0, "lbdem_psc_map");
}

int lbdem_hybrid_map_coarse(const lbdem_lattice *lat, const uint8_t *flags, const lbdem_grains *g,
                            const lbdem_units *u, const int32_t *coarse, int32_t ncoarse, uint32_t *key, float *snap,
                            int32_t *bbox, int32_t clear_previous, void *stream) {
  return psc_map(lat, flags, g, u, key, snap, bbox, clear_previous, stream, 1, coarse, ncoarse,
                 "lbdem_hybrid_map_coarse");
}

int lbdem_psc_forces(const lbdem_lattice *lat, const lbdem_grains *g, const lbdem_units *u, const uint32_t *key,
                     const float *cell_force, uint32_t *stale_key, const int32_t *stale_bbox, void *stream) {
  return launch_psc_forces(lat, g, u, nullptr, 0, key, cell_force, stale_key, stale_bbox, stream,
                           "lbdem_psc_forces");
}

int lbdem_psc_decode(const lbdem_lattice *lat, const uint32_t *key, const float *snap, int32_t *id,
                     float *volfrac, float *weight, float *velsolid, void *stream) {
  LBDEM_REQUIRE(lat && key && snap && id && volfrac && weight && velsolid, "lbdem_psc_decode: null argument");
  PscArgs a{};
  a.nx = lat->nx;
  a.ny = lat->ny;
  a.nz = lat->nz;
  a.key = const_cast<uint32_t *>(key);
  a.snap = const_cast<float *>(snap);
  a.tau0 = 1.0 / lat->omega0;
  const int n = lat->nx * lat->ny * lat->nz;
  hipLaunchKernelGGL(k_psc_decode, dim3((n + 255) / 256), dim3(256), 0, as_stream(stream), a, id, volfrac,
                     weight, velsolid, n);
  LBDEM_CHECK_LAUNCH("lbdem_psc_decode");
  return LBDEM_OK;
}

}  // extern "C"
