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

#define LBDEM_ABI_VERSION 9

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
#define LBDEM_COUPLE_HYBRID 4 /* coarse grains as PSC (keys) and fine grains as EqIBM (fractions)
                                 in one collision, f + beta Os1 + B Os2 + (1-B-beta) Of */

/* per-cell flag byte */
#define LBDEM_FLAG_ACTIVE 0x01 /* collided and streamed (not OBSTACLE / VEL_LADD); every cell
                                  without it must be WALL or MOVING: its slots hold the
                                  populations bounced at it, not populations of its own */
#define LBDEM_FLAG_NEAR 0x02   /* has a wall neighbour (set by lbdem_lbm_prepare_flags) */
#define LBDEM_FLAG_WALL 0x04   /* bounce-back wall (CellType.OBSTACLE; CellType.FREE_SLIP, whose
                                  reflection lbdem_lbm_free_slip applies after the step) */
#define LBDEM_FLAG_MOVING 0x08 /* moving bounce-back wall (CellType.VEL_LADD), velocity in vel[] */
#define LBDEM_FLAG_FLUID 0x10  /* rho,u are refreshed after streaming (CellType.FLUID) */
#define LBDEM_FLAG_INLET 0x20  /* with MOVING: Ladd velocity inlet (CellType.VEL_INLET_LADD); its
                                  Ladd term uses this cell's density (step_args.rho) */
#define LBDEM_FLAG_ROW_IDLE 0x40 /* on the first cell (k = 0) of a row (fixed i, j) that has no ACTIVE
                                    cell: the row-tiled step leaves such rows alone (their step
                                    would store every slot unchanged).  Set and cleared by
                                    lbdem_lbm_prepare_flags; derived state, not an input */

/* wet-node boundary cells (lbdem_lbm_boundaries): kind = type | face << 4 */
#define LBDEM_BC_VEL_ZOUHE 1   /* CellType.VEL_ZOUHE: Zou-He velocity (vel given, rho computed) */
#define LBDEM_BC_PRE_ZOUHE 2   /* CellType.Pre_ZOUHE: Zou-He pressure (rho given, normal u computed) */
#define LBDEM_BC_EXIT 3        /* CellType.VEL_EXIT: neighbour's velocity, rho = 1, f = feq */
#define LBDEM_BC_LADD_INLET 4  /* CellType.VEL_INLET_LADD: rho = inner neighbour's rho */
/* face: 0 LEFT (x = 0), 1 RIGHT, 2 BOTTOM (y = 0), 3 TOP, 4 BACK (z = 0), 5 FRONT, 7 none */

typedef struct lbdem_lattice {
  int32_t nx, ny, nz;  /* lattice extents */
  int32_t storage;     /* LBDEM_STORE_* */
  int32_t collision;   /* LBDEM_COLL_* */
  int32_t coupling;    /* LBDEM_COUPLE_* */
  double omega0;       /* 1/tau0 */
  double cs_smag;      /* Smagorinsky constant */
  double rho0;         /* shift density of the stored populations */
  double force[3];     /* body force in lattice units: adds rho*w_q*(c_q.F)/cs^2 */
  int32_t kernel;      /* collide-stream variant: 0 automatic (2 for FP32 storage, 3 for FP64,
                          4 for 16-bit storage, 3 if nz % 4 != 0; 1 where the row does not fit),
                          1 one cell per thread,
                          2 / 3 / 4 row-tiled (one workgroup per (i, j) row, aligned accesses,
                          z shift of odd steps through LDS) with 1 / 2 / 4 cells per thread;
                          nz must be a multiple of that and at most 1024 times it; 4 is not
                          built for FP64 storage */
  int32_t reserved;
} lbdem_lattice;

/* Per-step device pointers.  Unused members may be NULL. */
typedef struct lbdem_step_args {
  void *f;                     /* 19*ncell populations, storage type */
  const uint8_t *flags;        /* ncell flag bytes */
  const uint32_t *nmask;       /* ncell wall-neighbour masks from lbdem_lbm_prepare_flags */
  const float *vel;            /* 3*ncell (AoS): wall velocity read at MOVING cells */
  void *lag;                   /* 5*ncell (SoA rho,ux,uy,uz,omega) for SMAG_LAGGED; float,
                                  double with LBDEM_STORE_F64 */
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
  const float *rho;            /* cell density field; read at INLET walls (may be NULL without) */
} lbdem_step_args;

int lbdem_abi_version(void);
const char *lbdem_last_error(void);

/* Derived flag data, on the device, from the ACTIVE / WALL / MOVING bits the caller sets: the
 * NEAR bit of every cell (active with a wall among its 18 periodic neighbours) and the
 * wall-neighbour mask of the NEAR cells (bit q, 1..18: x + c_q is a wall; bit 31: some
 * neighbour is a moving wall; 0 elsewhere).  Call after every change of the flags.  It replaces
 * the per-direction cell-type tests of BasicLattice3D.stream (src/lbm3d/lbm_solver3d.py:216-236)
 * with one load per near-wall cell.  Populations bounced at a wall are kept in the wall cell's
 * slots, so the layout of f depends on the flags: to change walls on a lattice that holds
 * populations, read them with lbdem_lbm_read_f, change the flags and call this function, then
 * store them back with lbdem_lbm_write_f (same t).  The next step then starts from the same
 * post-stream state, as in the reference. */
int lbdem_lbm_prepare_flags(const lbdem_lattice *lat, uint8_t *flags, uint32_t *nmask, void *stream);

/* f <- feq(rho, vel) at active cells as the state t = 0; lag state initialised.
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
 * (src/unlbdem/vanslattice.py:203-440), HybridLattice3D.collide (src/hybridlbdem/
 * hybridlattice.py:433-593: coarse keys in psc_key / psc_grain, fine fractions in volfrac /
 * velsolid, the momentum exchange of the coarse part into psc_force). */
int lbdem_lbm_step(const lbdem_lattice *lat, const lbdem_step_args *a, void *stream);

/* Wet-node boundaries, after the step that produced the state after a->t steps: the second
 * loop of BasicLattice3D.stream (src/lbm3d/lbm_solver3d.py:241-250) over the ncells boundary
 * cells (cells[e] with kinds[e], LBDEM_BC_* | face << 4): vel_zouHe (:671-820), pre_zouHe
 * (:823-957), bc_vel_exit (:645-667), bc_ladd_inlet (:516-527).  The Zou-He and exit cells are
 * ACTIVE (collided and streamed like fluid, not FLUID); their populations are rebuilt in place
 * and rho (velocity condition, exit) or vel (pressure condition, exit) updated.  Neighbour values
 * come from the neighbour's populations when it is FLUID, else from rho / vel. */
int lbdem_lbm_boundaries(const lbdem_lattice *lat, const lbdem_step_args *a, const int32_t *cells,
                         const uint8_t *kinds, int32_t ncells, float *rho, float *vel, void *stream);

/* Free-slip walls, after the step that produced the state after a->t steps and before
 * lbdem_lbm_boundaries: the FREE_SLIP branch of BasicLattice3D.stream (src/lbm3d/lbm_solver3d.py:
 * 230-236, mirror maps :50-57; also Unresolvedlattice3D.stream, src/unlbdem/vanslattice.py:
 * 388-395).  Free-slip cells are WALL cells for the step (halfway bounce-back); this call then
 * sets each listed post-stream population to the reflected post-collision one.  pairs: npairs
 * records of four int32 -- target cell and direction (an active cell whose upstream neighbour is
 * the free-slip cell), source cell and direction (a post-stream population of the same step
 * holding the post-collision value fpc[x'][qsy[q]]; source cell -1: the value 0, the reference's
 * fpc of a cell that never collides).  work: npairs doubles of scratch.  All sources are read
 * before any target is written. */
int lbdem_lbm_free_slip(const lbdem_lattice *lat, const lbdem_step_args *a, const int32_t *pairs, int32_t npairs,
                        double *work, void *stream);

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

/* Store unshifted natural-order populations (double if in_f64) at active cells as the
 * post-stream state after a->t steps (the inverse of lbdem_lbm_read_f at the same t).
 * Backs `lattice.f.from_numpy()`. */
int lbdem_lbm_write_f(const lbdem_lattice *lat, const lbdem_step_args *a, const void *in,
                      int32_t in_f64, void *stream);

/* Stress from the non-equilibrium populations, BasicLattice3D.compute_stress
 * (src/lbm3d/lbm_solver3d.py:347-405): sigma_ab = -(1 - omega/2) sum_q c_a c_b (f_q - feq_q),
 * f the populations after a->t steps, feq and omega those of the last collision: exact with
 * lagged Smagorinsky (kept in a->lag); otherwise the equilibrium of the current populations
 * with omega0 (BGK), the fixed-point omega, or the VANS omega of a->volfrac.  stress: seven
 * planes of ncell floats, xx, yy, zz, xy, xz, yz and the shear magnitude
 * sqrt((xy^2 + xz^2 + yz^2) / 2); cells the collision skips are not written. */
int lbdem_lbm_stress(const lbdem_lattice *lat, const lbdem_step_args *a, float *stress, void *stream);

/* Mean of the shear components (xy, xz, yz) of `stress` over the FLUID cells of the plane
 * {x_axis = position}, compute_plane_average_stress (src/lbm3d/lbm_solver3d.py:407-456);
 * out: 3 doubles on the device, zeros for a plane without fluid cells.  One workgroup, fixed
 * summation order. */
int lbdem_lbm_plane_stress(const lbdem_lattice *lat, const uint8_t *flags, const float *stress, int32_t axis,
                           int32_t position, double *out, void *stream);

/* FP64 global sums over active cells of the post-stream state after t steps:
 * out[0] = sum rho, out[1..3] = sum rho*u, out[4] = sum rho|u|^2/2.  Deterministic
 * (fixed-order two-pass).  `work` must hold lbdem_reduce_work_doubles(lat) doubles.
 * No counterpart is live in the reference: its get_total_mass / get_kinetic_energy sit in
 * a string literal (src/lbm3d/lbm_solver3d.py:959-1073). */
int64_t lbdem_reduce_work_doubles(const lbdem_lattice *lat);
int lbdem_lbm_reduce(const lbdem_lattice *lat, const lbdem_step_args *a, double *work,
                     double *out, void *stream);

/* lbdem_lbm_reduce without the caller's work buffer, and without any allocation: the
 * per-workgroup partials live in a file-scope device array of the library (5 MiB, one per GPU,
 * kept until the process ends), for lattices of at most 2^25 cells (larger ones are rejected
 * with LBDEM_E_ARG).  No counterpart in the reference: it is for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Calls on one GPU share that GPU's array: calls that overlap on one GPU (other
 * streams, other host threads) race; calls in stream order do not. */
int lbdem_lbm_reduce_static(const lbdem_lattice *lat, const lbdem_step_args *a, double *out, void *stream);

/* ------------------------------------------------------------------------------------ */
/* grain / particle state shared by the couplings and the DEM (device pointers, FP64)     */
/* ------------------------------------------------------------------------------------ */
typedef struct lbdem_grains {
  int32_t n;                 /* number of grains */
  const double *position;    /* 3n, physical units */
  const double *velocity;    /* 3n */
  const double *omega;       /* 3n angular velocity */
  const double *radius;      /* n */
  double *force_fluid;       /* 3n, written by the couplings */
  double *moment_fluid;      /* 3n */
  const int32_t *order;      /* n, may be NULL: a permutation of 0..n-1 that the EqIBM spreading
                                and drag visit the grains in (grains close in space close in the
                                order, e.g. the DEM broad phase's bucket order); other couplings
                                ignore it */
} lbdem_grains;

/* lattice <-> physical units (UnitConverter, src/lbm3d/lbmutils.py:14-63) and the DEM
 * domain minimum: a grain at x sits at lattice coordinate (x - dmin)/dx + 1/2. */
typedef struct lbdem_units {
  double dx, dt, rho;
  double dmin[3];
} lbdem_units;

/* PSC grain -> lattice mapping.  Replaces PSCLattice3D.grains2lattice with compute_weight and
 * compute_solid_vel (src/lbdem/psclattice3d.py:84-182): writes per-cell packed keys
 * (full bit 30, sub-cell count bits 29..23, 0x7FFFFF - grain bits 22..0; 0 = uncovered),
 * a 12-float per-grain snapshot (centre in lattice units, v*dt/dx, omega*dt) read by the
 * collision, and each grain's 6-int box.  clear_previous zeroes the keys inside the boxes
 * of the previous call first.  Grain-centric; geometry in FP64. */
int lbdem_psc_map(const lbdem_lattice *lat, const uint8_t *flags, const lbdem_grains *g,
                  const lbdem_units *u, uint32_t *key, float *snap, int32_t *bbox,
                  int32_t clear_previous, void *stream);

/* Per-grain hydrodynamic force and torque (physical units, FP64, fixed-order reduction) from
 * the per-cell force the PSC collision wrote.  Replaces PSCLattice3D.lattice2grains
 * (src/lbdem/psclattice3d.py:316-358).  stale_key / stale_bbox (may be NULL): a second key
 * array holding an earlier mapping with its boxes; the same launch zeroes those keys, so the
 * next lbdem_psc_map into that array needs no clear pass (two key arrays used in turn). */
int lbdem_psc_forces(const lbdem_lattice *lat, const lbdem_grains *g, const lbdem_units *u,
                     const uint32_t *key, const float *cell_force, uint32_t *stale_key,
                     const int32_t *stale_bbox, void *stream);

/* id / volfrac / weight / velsolid fields (psclattice3d.py:45-50) decoded from the keys. */
int lbdem_psc_decode(const lbdem_lattice *lat, const uint32_t *key, const float *snap, int32_t *id,
                     float *volfrac, float *weight, float *velsolid, void *stream);

/* ------------------------------------------------------------------------------------ */
/* hybrid resolved / unresolved coupling (src/hybridlbdem)                                */
/* ------------------------------------------------------------------------------------ */
/* Coarse-grain mapping, HybridLattice3D.map_coarse_grains (src/hybridlbdem/hybridlattice.py:
 * 173-234): lbdem_psc_map over the grains with 2R > dx only.  coarse (device, ncoarse
 * entries): the indices of those grains, one workgroup column each; NULL walks every grain and
 * skips the fine ones.  The PSC weight B is not mapped: the hybrid collision evaluates it with
 * the step's own omega (_psc_compute_weight :236-244). */
int lbdem_hybrid_map_coarse(const lbdem_lattice *lat, const uint8_t *flags, const lbdem_grains *g,
                            const lbdem_units *u, const int32_t *coarse, int32_t ncoarse, uint32_t *key,
                            float *snap, int32_t *bbox, int32_t clear_previous, void *stream);

/* Fine-grain spreading, HybridLattice3D.map_fine_grains (:260-376): for each grain with
 * 2R <= dx, the mirror 3-point kernel over the active cells not fully covered by a coarse grain
 * (key), normalised, corrected by the free fraction 1 - eps_c, spread into fine_volfrac,
 * velsum and wsum (FP32 atomics); then per cell the clamp to the free fraction (status[0] +=
 * cells that overflowed by more than 1e-6, the reference's warning) and the solid velocity
 * velsum / fine_volfrac where fine_volfrac > 1e-10 (elsewhere fine_velsolid keeps its value). */
int lbdem_hybrid_map_fine(const lbdem_lattice *lat, const uint8_t *flags, const lbdem_grains *g,
                          const lbdem_units *u, const uint32_t *key, float *fine_volfrac, float *velsum,
                          float *wsum, float *fine_velsolid, int32_t *status, void *stream);

/* Hydrodynamic loads of all grains, HybridLattice3D.lattice2grains (:595-740): for the coarse
 * grains (the ncoarse indices in coarse, as for lbdem_hybrid_map_coarse) force and torque from
 * the momentum exchange the collision wrote per cell (_transfer_coarse_force :626-659, FP64,
 * fixed-order reduction, as lbdem_psc_forces, with its stale-key clearing); for the others the
 * Tenneti drag (_transfer_fine_force, _compute_drag_force :661-740) from the kernel-interpolated
 * fluid velocity (vel: the post-stream velocity field) and the total solid fraction
 * fine + coarse, and zero torque. */
int lbdem_hybrid_forces(const lbdem_lattice *lat, const uint8_t *flags, const lbdem_grains *g,
                        const lbdem_units *u, const int32_t *coarse, int32_t ncoarse, const uint32_t *key,
                        const float *cell_force, uint32_t *stale_key, const int32_t *stale_bbox,
                        const float *fine_volfrac, const float *vel, void *stream);

/* Tenneti weight of the fine grains per cell, HybridLattice3D.compute_fine_weights (:378-431),
 * for inspection (the collision evaluates it inline); cells where it is not defined keep the
 * weight they hold, as in the reference. */
int lbdem_hybrid_fine_weight(const lbdem_lattice *lat, const uint8_t *flags, const uint32_t *key,
                             const float *fine_volfrac, const float *fine_velsolid, const float *vel,
                             float *weight, void *stream);

/* ------------------------------------------------------------------------------------ */
/* unresolved coupling (src/unlbdem)                                                      */
/* ------------------------------------------------------------------------------------ */
/* EqIBM spreading: EqIMBlattice3D.grains2lattice (src/unlbdem/eqlattice.py:104-194) with the
 * mirror kernel (:468-528) and threedelta (:451-462).  Zeroes and accumulates volfrac,
 * velsum (3n AoS), wsum; clamps volfrac >= 1 to 0.99 (counted in status[0] if non-NULL);
 * velsolid = velsum / volfrac.  Also writes rho, vel (AoS) of the current post-stream state
 * (step args: f, flags, t) at covered FLUID cells: the fluid velocity that compute_weight
 * (:200-238, evaluated inside lbdem_lbm_step) and lattice2grains read. */
int lbdem_ibm_spread(const lbdem_lattice *lat, const lbdem_step_args *a, const lbdem_grains *g,
                     const lbdem_units *u, float *volfrac, float *velsum, float *wsum,
                     float *velsolid, float *rho, float *vel, int32_t *status, void *stream);

/* EqIBM interpolation + Tenneti drag into g->force_fluid: EqIMBlattice3D.lattice2grains and
 * compute_drag_force (src/unlbdem/eqlattice.py:309-444). */
int lbdem_ibm_drag(const lbdem_lattice *lat, const uint8_t *flags, const lbdem_grains *g,
                   const lbdem_units *u, const float *volfrac, const float *vel, void *stream);

/* Tenneti collision weight per covered cell (EqIMBlattice3D.compute_weight,
 * src/unlbdem/eqlattice.py:200-238) for inspection; the step kernel evaluates it inline. */
int lbdem_ibm_weight(const lbdem_lattice *lat, const uint8_t *flags, const float *volfrac,
                     const float *velsolid, const float *vel, float *weight, void *stream);

/* VANS spreading: Unresolvedlattice3D.force2lattice (src/unlbdem/vanslattice.py:489-557), or
 * grains2lattice (:135-196, volume only) when hforce is NULL. */
int lbdem_vans_spread(const lbdem_lattice *lat, const uint8_t *flags, const lbdem_grains *g,
                      const lbdem_units *u, float *volfrac, float *hforce, void *stream);

/* VANS drag: Unresolvedlattice3D.compute_hydrodynamic_force + compute_drag_force
 * (src/unlbdem/vanslattice.py:446-482, 563-651) into g->force_fluid. */
int lbdem_vans_drag(const lbdem_lattice *lat, const uint8_t *flags, const lbdem_grains *g,
                    const lbdem_units *u, const float *volfrac, const float *vel, void *stream);

/* ------------------------------------------------------------------------------------ */
/* DEM (src/dem3d)                                                                        */
/* ------------------------------------------------------------------------------------ */
/* material row layout: density, E, nu, friction, restitution, k_n, k_t, damping_n, damping_t */
typedef struct lbdem_dem_params {
  int32_t n;            /* particles */
  int32_t K;            /* history slots per particle (max_coordinate_number) */
  int32_t model;        /* 0 Hertz-Mindlin, 1 linear spring-dashpot */
  int32_t T;            /* hash table size, power of two (BPCD.create, bpcd.py:48-65) */
  int32_t probe27;      /* 0: 8-cell octant probe (reference); 1: 27-cell probe */
  int32_t periodic[3];  /* periodic axes */
  double dt;            /* DEM time step */
  double g[3];          /* gravity */
  double dmin[3], dmax[3]; /* DEMSolverConfig.domain bounds */
  double cell;          /* hash cell size (reference: 4 r_max) */
  double mat[2][9];     /* [0] particle, [1] wall material */
  double wall_n[6][3];  /* outward wall normals (DEMSolverConfig._setup_walls) */
  double wall_d[6];     /* n.x = d */
  int32_t wall_type[6]; /* 0 wall, 1 periodic (skipped) */
  /* No counterpart in the reference (a search strategy, not a rule): with fine_cell > 0 the
   * grid-wide path of lbdem_dem_run finds its touching pairs on a second grid of that cell size
   * (27-cell probe) for the particles of radius <= large_r, and tests the nlarge particles of
   * lbdem_dem_work.large (radius > large_r) against all particles; every pair found is kept
   * only if it is a candidate of the grid above (j's bucket among i's probed buckets), so the
   * pairs are those of the 8-cell probe.  Requires fine_cell > 2 large_r, probe27 = 0 and no
   * periodic axis.  For strongly polydisperse packings, whose cell (4 r_max) holds most of the
   * particles in a few buckets. */
  double fine_cell;
  double large_r;
  int32_t nlarge;
} lbdem_dem_params;

/* Particle and contact state, device pointers (FP64 unless noted).  The Grain / Contact
 * structs of the reference (src/dem3d/dateclass/grain.py:23-52, contact.py:20-33) as
 * structure-of-arrays: vectors AoS per particle (3 or 4 or 9 doubles). */
typedef struct lbdem_dem_state {
  double *mass, *radius;             /* n */
  double *pos, *vel, *acc, *force;   /* 3n */
  double *quat;                      /* 4n (w, x, y, z) */
  double *omega, *omega_dot, *moment;/* 3n */
  double *inertia;                   /* 9n */
  double *force_fluid, *moment_fluid;/* 3n, added every substep */
  const int32_t *freeze, *fixvel;    /* n, may be NULL */
  int32_t *cf_j, *cf_active;         /* n*K history slots of particle i: partner j > i */
  int32_t *cfn;                      /* n slots in use */
  double *cf_pos, *cf_force, *cf_shear; /* 3*n*K */
  int32_t *wcf_active;               /* 6n wall contacts */
  double *wcf_pos, *wcf_force, *wcf_shear; /* 18n */
  int32_t *status;                   /* 8 counters: [0] slot overflow, [1] dt above the
                                        contact-law critical step, [2] particle overlap > 5%,
                                        [3] wall overlap > 5%, [4] candidate-list overflow
                                        (required length) */
} lbdem_dem_state;

/* Caller-provided DEM scratch.  count/current must be zero on first use; the library leaves
 * them zero.  npairs must be zero before lbdem_dem_broadphase; lbdem_dem_run resets it. */
typedef struct lbdem_dem_work {
  int32_t *count, *offset, *current; /* T each */
  int32_t *tile_sum;                 /* ceil(T / 1024) */
  int32_t *hash, *pid;               /* n each */
  int32_t *pairs;                    /* 2 * pair_cap candidate pairs (i < j): all of them
                                        (lbdem_dem_broadphase), the touching ones (general path of
                                        lbdem_dem_run), or the new contacts of a substep
                                        (single-workgroup path) */
  int64_t pair_cap;
  unsigned long long *npairs;        /* 1: candidate count (may exceed pair_cap) */
  int32_t *need_new;                 /* pair_cap */
  int32_t *verlet;                   /* 2 + 17n, may be NULL: neighbour lists the single-workgroup
                                        path keeps between calls (n <= 256); zero it, or its first
                                        element, whenever radii or contact history change */
  double *verlet_ref;                /* 1 + 3n, may be NULL (with verlet) */
  double *spos;                      /* 4n, may be NULL: x, y, z and radius of each particle in
                                        bucket order, written by the general path of lbdem_dem_run
                                        so that its probe walks particles and candidates in space
                                        order (NULL: particle order).  Both broad-phase entry
                                        points also keep each particle's bucket slot in its first
                                        4n bytes between the count and the scatter (NULL: the
                                        scatter counts again, in current) */
  int32_t *fhash;                    /* n, with fine_cell > 0: cell hashes on the second grid */
  const int32_t *large;              /* nlarge, with fine_cell > 0: indices of the particles of
                                        radius > large_r */
} lbdem_dem_work;

/* Candidate pairs of the Morton-hash broad phase: BPCD.detect_collision with the count,
 * Blelloch scan, scatter and 8-octant probe (src/dem3d/bpcd/bpcd.py:67-152,
 * src/dem3d/bpcd/prefixsum.py:62-84).  Pairs (i, j), i < j, land unordered in w->pairs;
 * w->npairs receives their number (also when it exceeds pair_cap). */
int lbdem_dem_broadphase(const lbdem_dem_params *P, const double *pos, const lbdem_dem_work *w,
                         void *stream);

/* nsub DEM substeps, each DEMSolver.run_simulation (src/dem3d/demsolver/demsolver.py:848-862):
 * clear_state, contact (broad phase + resolve :638-740), resolve_wall (:809-838),
 * apply_body_force (:437-448), update (:514-636), late_clear_state (:431-436).
 * path: 0 automatic, 1 single-workgroup path (n <= 2048, one launch for all substeps),
 * 2 general grid-wide path.  The general path lists only the touching candidate pairs in
 * w->pairs (and clears the history slots of candidates that stopped touching), which resolves
 * every candidate as the reference does while the list stays the size of the contact set;
 * w->npairs counts them, pair_cap bounds them (overflow: status[4]). */
int lbdem_dem_run(const lbdem_dem_params *P, const lbdem_dem_state *S, const lbdem_dem_work *w,
                  int32_t nsub, int32_t path, void *stream);

/* ------------------------------------------------------------------------------------ */
/* differentiable PSC channel (forward + hand-written adjoint)                            */
/* ------------------------------------------------------------------------------------ */
/* No counterpart in the reference (it has no adjoint).  The forward model is the PSC
 * collision with fixed grains (src/lbdem/psclattice3d.py:240-261) plus the body-force term of
 * ForcedLattice3D (src/lbm3d/lbm_foecelover.py:53-68) on the fluid part, (1 - B) rho w (c.F)/cs^2,
 * and bounce-back / periodic streaming (src/lbm3d/lbm_solver3d.py:203-224).  Two-array layout,
 * unshifted populations f[q*ncell + c], FP32 or FP64. */
typedef struct lbdem_diff_args {
  int32_t nx, ny, nz;
  int32_t precision;        /* 32 or 64: element type of every state / adjoint / eps array */
  const uint8_t *flags;     /* ncell flag bytes (ACTIVE, WALL; NEAR from lbdem_lbm_prepare_flags) */
  const void *eps;          /* ncell solid fraction per cell */
  double tau;               /* relaxation time */
  double force[3];          /* body force, lattice units */
  int32_t plane;            /* outlet plane index along x */
  int32_t reserved;
} lbdem_diff_args;

/* f <- rest state (rho = 1, u = 0) at active cells */
int lbdem_diff_init(const lbdem_diff_args *d, void *f, void *stream);
/* f_out = S(C(f_in)) */
int lbdem_diff_step(const lbdem_diff_args *d, const void *f_in, void *f_out, void *stream);
/* lam_out = C'(f_n)^T S^T lam_in; g_eps[c] += dL/d eps_c; the FP64 contributions to
 * dL/dtau, dL/dF (3) are added to g_red, an accumulator of lbdem_diff_red_doubles(d) doubles
 * (four per workgroup, zeroed by the caller before the first step), whose total
 * lbdem_diff_reduce returns. */
int lbdem_diff_adjoint_step(const lbdem_diff_args *d, const void *f_n, const void *lam_in, void *lam_out,
                            void *g_eps, double *g_red, void *stream);
int64_t lbdem_diff_red_doubles(const lbdem_diff_args *d);
/* out[0] = dL/dtau, out[1..3] = dL/dF: the sum of the accumulator in a fixed order
 * (deterministic, like lbdem_lbm_reduce) */
int lbdem_diff_reduce(const lbdem_diff_args *d, const double *g_red, double *out, void *stream);
/* u_x on the plane i = plane (ny*nz) */
int lbdem_diff_outlet(const lbdem_diff_args *d, const void *f, void *u_plane, void *stream);
/* lam = dL/df of the state f given dL/du_x on the plane (g_u, ny*nz); zero off the plane */
int lbdem_diff_seed(const lbdem_diff_args *d, const void *f, const void *g_u, void *lam, void *stream);

#ifdef __cplusplus
}
#endif
#endif /* LBDEM_HIP_H */
