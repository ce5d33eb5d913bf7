/*
 * lbdem_hip.h -- C ABI of liblbdem_hip.so, the gfx950 (MI355X) implementation of the
 * coupled lattice-Boltzmann / discrete-element time step of LBDEM-Taichi.
 *
 * Every entry point:
 *   - takes plain device pointers (allocated and owned by the caller, e.g. torch tensors),
 *     sizes, a parameter struct passed by pointer, and a HIP stream handle (`void*`,
 *     a hipStream_t; NULL = the null stream);
 *   - enqueues its kernels on that stream and returns without synchronising;
 *   - returns 0 (LBDEM_OK) or a negative LBDEM_E_* code; lbdem_last_error() then returns a
 *     thread-local message.
 * The library allocates no device memory: scratch space is passed in by the caller.
 *
 * Lattice layout (all lattice arrays): cell index c = (i*ny + j)*nz + k (k fastest), the
 * same order as the reference's Taichi fields `[i, j, k]`.  Populations are stored
 * structure-of-arrays, f[q*ncell + c], as shifted values f_q - w_q*rho0 in the storage type
 * chosen by lbdem_lattice.storage, updated in place with the AA access pattern: after an
 * even number of steps each cell holds its own post-stream populations in natural slot
 * order; after an odd number of steps it holds its post-collision populations in the
 * opposite slots.  Every entry point that reads f takes the number of steps taken, `t`.
 *
 * Reference file:line citations are relative to the LBDEM-Taichi repository root.
 */
#ifndef LBDEM_HIP_H
#define LBDEM_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LBDEM_ABI_VERSION 1

/* return codes */
#define LBDEM_OK 0
#define LBDEM_E_ARG (-1)         /* invalid argument (size, pointer, enum) */
#define LBDEM_E_HIP (-2)         /* a HIP runtime call failed */
#define LBDEM_E_UNSUPPORTED (-3) /* combination not implemented */
#define LBDEM_E_CAPACITY (-4)    /* caller-provided buffer too small */

/* population storage (lbdem_lattice.storage) */
#define LBDEM_STORE_F32 0
#define LBDEM_STORE_F16 1  /* shifted IEEE half, FP32 collision */
#define LBDEM_STORE_BF16 2 /* shifted bfloat16, FP32 collision */
#define LBDEM_STORE_F64 3  /* FP64 storage and FP64 collision */

/* collision operator (lbdem_lattice.collision) */
#define LBDEM_COLL_BGK 0        /* constant omega0 (ForcedLattice3D, PSCLattice3D) */
#define LBDEM_COLL_SMAG_LAGGED 1 /* Smagorinsky exactly as the reference: previous step's feq and omega */
#define LBDEM_COLL_SMAG_FIXED 2  /* Smagorinsky, lag-free self-consistent omega (fixed point of the above) */

/* fluid-solid coupling folded into the collision (lbdem_lattice.coupling) */
#define LBDEM_COUPLE_NONE 0
#define LBDEM_COUPLE_PSC 1   /* partially saturated cells, resolved grains */
#define LBDEM_COUPLE_EQIBM 2 /* particle-equivalence IBM, Tenneti weight */
#define LBDEM_COUPLE_VANS 3  /* volume-averaged NS with Guo-type force */

/* per-cell flag byte */
#define LBDEM_FLAG_ACTIVE 0x01 /* collided and streamed (not OBSTACLE / VEL_LADD) */
#define LBDEM_FLAG_NEAR 0x02   /* has a wall neighbour or lies on the domain edge (slow path) */
#define LBDEM_FLAG_WALL 0x04   /* bounce-back wall (CellType.OBSTACLE) */
#define LBDEM_FLAG_MOVING 0x08 /* moving bounce-back wall (CellType.VEL_LADD), velocity in vel[] */
#define LBDEM_FLAG_FLUID 0x10  /* rho,u are refreshed after streaming (CellType.FLUID) */

typedef struct lbdem_lattice {
  int32_t nx, ny, nz;  /* lattice extents */
  int32_t storage;     /* LBDEM_STORE_* */
  int32_t collision;   /* LBDEM_COLL_* */
  int32_t coupling;    /* LBDEM_COUPLE_* */
  double omega0;       /* 1/tau0 */
  double cs_smag;      /* Smagorinsky constant */
  double rho0;         /* shift density of the stored populations */
  double force[3];     /* body force in lattice units: adds rho*w_q*(c_q.F)/cs^2 */
} lbdem_lattice;

/* Per-step device pointers.  Unused members may be NULL. */
typedef struct lbdem_step_args {
  void *f;                     /* 19*ncell populations, storage type */
  const uint8_t *flags;        /* ncell flag bytes */
  const float *vel;            /* 3*ncell (AoS): wall velocity read at MOVING cells */
  float *lag;                  /* 5*ncell (SoA rho,ux,uy,uz,omega) for SMAG_LAGGED */
  /* PSC */
  const uint32_t *psc_key;     /* ncell packed (full, count, grain) keys, 0 = no grain */
  const float *psc_grain;      /* 12*ngrain: xc,yc,zc, v_lu xyz, omega*dt xyz, unused x3 */
  float *psc_force;            /* 3*ncell (AoS): -sum_q B*Omega_s*c_q at covered cells */
  /* EqIBM / VANS */
  const float *volfrac;        /* ncell solid volume fraction (this step) */
  const float *velsolid;       /* 3*ncell (AoS) solid velocity (EqIBM) */
  const float *hforce;         /* 3*ncell (AoS) fluid force density (VANS, this step) */
  const float *volfrac_prev;   /* VANS: volume fraction used for the macroscopic velocity */
  const float *hforce_prev;    /* VANS: force used for the macroscopic velocity */
  int64_t t;                   /* number of steps already taken (AA parity) */
  double nu_lu;                /* base lattice viscosity (1/omega0 - 1/2)/3 (EqIBM, VANS) */
} lbdem_step_args;

int lbdem_abi_version(void);
const char *lbdem_last_error(void);

/* f <- feq(rho, vel) at active cells, natural layout (t = 0); lag state initialised.
 * Replaces BasicLattice3D.initialize (src/lbm3d/lbm_solver3d.py:157-171), also
 * PSCLattice3D.initialize (src/lbdem/psclattice3d.py:60-78, lattice part),
 * EqIMBlattice3D.initialize (src/unlbdem/eqlattice.py:90-98),
 * Unresolvedlattice3D.initialize (src/unlbdem/vanslattice.py:106-127). */
int lbdem_lbm_init(const lbdem_lattice *lat, const lbdem_step_args *a, const float *rho,
                   void *stream);

/* One fused collide+stream step (AA pattern), t -> t+1.
 * Replaces BasicLattice3D.collide + stream (src/lbm3d/lbm_solver3d.py:176-250, with
 * computeOmega/compute_strain_rate :460-513, compute_feq :255-271, compute_neigh_index
 * :308-342), ForcedLattice3D.collide (src/lbm3d/lbm_foecelover.py:37-68),
 * PSCLattice3D.collide (src/lbdem/psclattice3d.py:187-261, cell loop; the grain mapping
 * and reduction are lbdem_psc_map / lbdem_psc_forces),
 * EqIMBlattice3D.collide (src/unlbdem/eqlattice.py:243-303, with the Tenneti weight of
 * compute_weight :200-238 evaluated inline), Unresolvedlattice3D.collide + stream
 * (src/unlbdem/vanslattice.py:203-440). */
int lbdem_lbm_step(const lbdem_lattice *lat, const lbdem_step_args *a, void *stream);

/* rho, vel (AoS) of the post-stream state after t steps, at FLUID cells (all active cells
 * for VANS, with its force/volume-fraction correction).  If `mask` is non-NULL only cells
 * with mask[c] > 0 are written.  Replaces compute_rho_vel
 * (src/lbm3d/lbm_solver3d.py:276-303) and compute_rhof_vel (src/unlbdem/vanslattice.py:412-440). */
int lbdem_lbm_macros(const lbdem_lattice *lat, const lbdem_step_args *a, float *rho,
                     float *vel, const float *mask, void *stream);

/* Unshifted populations of the post-stream state after t steps, natural order, into
 * out[q*ncell + c] (double if out_f64, else float).  Inactive cells are written as 0.
 * Backs `lattice.f.to_numpy()` (the reference's f field, src/lbm3d/lbm_solver3d.py:123). */
int lbdem_lbm_read_f(const lbdem_lattice *lat, const lbdem_step_args *a, void *out,
                     int32_t out_f64, void *stream);

/* Store unshifted natural-order populations (double if in_f64) as the state t = 0.
 * Backs `lattice.f.from_numpy()`. */
int lbdem_lbm_write_f(const lbdem_lattice *lat, const lbdem_step_args *a, const void *in,
                      int32_t in_f64, void *stream);

/* FP64 global sums over active cells of the post-stream state after t steps:
 * out[0] = sum rho, out[1..3] = sum rho*u, out[4] = sum rho|u|^2/2.  Deterministic
 * (fixed-order two-pass).  `work` must hold lbdem_reduce_work_doubles(lat) doubles.
 * No counterpart is live in the reference: its get_total_mass / get_kinetic_energy sit in
 * a string literal (src/lbm3d/lbm_solver3d.py:959-1073). */
int64_t lbdem_reduce_work_doubles(const lbdem_lattice *lat);
int lbdem_lbm_reduce(const lbdem_lattice *lat, const lbdem_step_args *a, double *work,
                     double *out, void *stream);

#ifdef __cplusplus
}
#endif
#endif /* LBDEM_HIP_H */
