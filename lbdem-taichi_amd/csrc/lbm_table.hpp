// Per-storage-type kernel tables (one translation unit per storage type).
#pragma once
#include "common.hpp"

namespace lbdem {

using KernelFn = void (*)(const KParams);
using BcKernelFn = void (*)(const KParams, const int32_t *, const uint8_t *, int, float *, float *);
using FsReadFn = void (*)(const KParams, const int32_t *, int, double *);
using FsWriteFn = void (*)(const KParams, const int32_t *, int, const double *);

struct LbmKernels {
  KernelFn step[3][5][2];  // [collision][coupling][odd]; nullptr = unsupported combination
  KernelFn pencil[3][5][2];   // row-tiled variant of step, one cell per thread
  KernelFn pencil2[3][5][2];  // row-tiled variant, two cells per thread
  KernelFn pencil4[3][5][2];  // row-tiled variant, four cells per thread (not for FP64)
  KernelFn init[5];        // [coupling]
  KernelFn macros[5];      // [coupling]
  KernelFn read_f;
  KernelFn write_f;
  BcKernelFn boundaries;
  FsReadFn free_slip_read;
  FsWriteFn free_slip_write;
  KernelFn stress;
  KernelFn reduce_partial;
};

const LbmKernels &lbm_kernels_f32();
const LbmKernels &lbm_kernels_f16();
const LbmKernels &lbm_kernels_bf16();
const LbmKernels &lbm_kernels_f64();

// hybrid-coupling entries, instantiated in their own translation units (lbm_hyb_<type>.hip)
// so that the four storage types and the two coupling sets compile in parallel
void add_hybrid_f32(LbmKernels &k);
void add_hybrid_f16(LbmKernels &k);
void add_hybrid_bf16(LbmKernels &k);
void add_hybrid_f64(LbmKernels &k);

}  // namespace lbdem

// Builds the table for storage type S (expanded once in each lbm_<type>.hip).
#define LBDEM_DEFINE_LBM_TABLE(S, NAME, ADD_HYBRID)                                            \
  namespace lbdem {                                                                             \
  const LbmKernels &NAME() {                                                                    \
    static const LbmKernels t = [] {                                                            \
      LbmKernels k{};                                                                           \
      k.step[LBDEM_COLL_BGK][LBDEM_COUPLE_NONE][0] = k_step<S, LBDEM_COLL_BGK, LBDEM_COUPLE_NONE, false>;               \
      k.pencil[LBDEM_COLL_BGK][LBDEM_COUPLE_NONE][0] = k_step_pencil<S, LBDEM_COLL_BGK, LBDEM_COUPLE_NONE, false, 1>;               \
      k.pencil2[LBDEM_COLL_BGK][LBDEM_COUPLE_NONE][0] = k_step_pencil<S, LBDEM_COLL_BGK, LBDEM_COUPLE_NONE, false, 2>;               \
      k.pencil4[LBDEM_COLL_BGK][LBDEM_COUPLE_NONE][0] = pencil_fn<S, LBDEM_COLL_BGK, LBDEM_COUPLE_NONE, false, 4>();               \
      k.step[LBDEM_COLL_BGK][LBDEM_COUPLE_NONE][1] = k_step<S, LBDEM_COLL_BGK, LBDEM_COUPLE_NONE, true>;                \
      k.pencil[LBDEM_COLL_BGK][LBDEM_COUPLE_NONE][1] = k_step_pencil<S, LBDEM_COLL_BGK, LBDEM_COUPLE_NONE, true, 1>;                \
      k.pencil2[LBDEM_COLL_BGK][LBDEM_COUPLE_NONE][1] = k_step_pencil<S, LBDEM_COLL_BGK, LBDEM_COUPLE_NONE, true, 2>;                \
      k.pencil4[LBDEM_COLL_BGK][LBDEM_COUPLE_NONE][1] = pencil_fn<S, LBDEM_COLL_BGK, LBDEM_COUPLE_NONE, true, 4>();                \
      k.step[LBDEM_COLL_SMAG_LAGGED][LBDEM_COUPLE_NONE][0] = k_step<S, LBDEM_COLL_SMAG_LAGGED, LBDEM_COUPLE_NONE, false>; \
      k.pencil[LBDEM_COLL_SMAG_LAGGED][LBDEM_COUPLE_NONE][0] = k_step_pencil<S, LBDEM_COLL_SMAG_LAGGED, LBDEM_COUPLE_NONE, false, 1>; \
      k.pencil2[LBDEM_COLL_SMAG_LAGGED][LBDEM_COUPLE_NONE][0] = k_step_pencil<S, LBDEM_COLL_SMAG_LAGGED, LBDEM_COUPLE_NONE, false, 2>; \
      k.pencil4[LBDEM_COLL_SMAG_LAGGED][LBDEM_COUPLE_NONE][0] = pencil_fn<S, LBDEM_COLL_SMAG_LAGGED, LBDEM_COUPLE_NONE, false, 4>(); \
      k.step[LBDEM_COLL_SMAG_LAGGED][LBDEM_COUPLE_NONE][1] = k_step<S, LBDEM_COLL_SMAG_LAGGED, LBDEM_COUPLE_NONE, true>;  \
      k.pencil[LBDEM_COLL_SMAG_LAGGED][LBDEM_COUPLE_NONE][1] = k_step_pencil<S, LBDEM_COLL_SMAG_LAGGED, LBDEM_COUPLE_NONE, true, 1>;  \
      k.pencil2[LBDEM_COLL_SMAG_LAGGED][LBDEM_COUPLE_NONE][1] = k_step_pencil<S, LBDEM_COLL_SMAG_LAGGED, LBDEM_COUPLE_NONE, true, 2>;  \
      k.pencil4[LBDEM_COLL_SMAG_LAGGED][LBDEM_COUPLE_NONE][1] = pencil_fn<S, LBDEM_COLL_SMAG_LAGGED, LBDEM_COUPLE_NONE, true, 4>();  \
      k.step[LBDEM_COLL_SMAG_FIXED][LBDEM_COUPLE_NONE][0] = k_step<S, LBDEM_COLL_SMAG_FIXED, LBDEM_COUPLE_NONE, false>;   \
      k.pencil[LBDEM_COLL_SMAG_FIXED][LBDEM_COUPLE_NONE][0] = k_step_pencil<S, LBDEM_COLL_SMAG_FIXED, LBDEM_COUPLE_NONE, false, 1>;   \
      k.pencil2[LBDEM_COLL_SMAG_FIXED][LBDEM_COUPLE_NONE][0] = k_step_pencil<S, LBDEM_COLL_SMAG_FIXED, LBDEM_COUPLE_NONE, false, 2>;   \
      k.pencil4[LBDEM_COLL_SMAG_FIXED][LBDEM_COUPLE_NONE][0] = pencil_fn<S, LBDEM_COLL_SMAG_FIXED, LBDEM_COUPLE_NONE, false, 4>();   \
      k.step[LBDEM_COLL_SMAG_FIXED][LBDEM_COUPLE_NONE][1] = k_step<S, LBDEM_COLL_SMAG_FIXED, LBDEM_COUPLE_NONE, true>;    \
      k.pencil[LBDEM_COLL_SMAG_FIXED][LBDEM_COUPLE_NONE][1] = k_step_pencil<S, LBDEM_COLL_SMAG_FIXED, LBDEM_COUPLE_NONE, true, 1>;    \
      k.pencil2[LBDEM_COLL_SMAG_FIXED][LBDEM_COUPLE_NONE][1] = k_step_pencil<S, LBDEM_COLL_SMAG_FIXED, LBDEM_COUPLE_NONE, true, 2>;    \
      k.pencil4[LBDEM_COLL_SMAG_FIXED][LBDEM_COUPLE_NONE][1] = pencil_fn<S, LBDEM_COLL_SMAG_FIXED, LBDEM_COUPLE_NONE, true, 4>();    \
      k.step[LBDEM_COLL_BGK][LBDEM_COUPLE_PSC][0] = k_step<S, LBDEM_COLL_BGK, LBDEM_COUPLE_PSC, false>;                 \
      k.pencil[LBDEM_COLL_BGK][LBDEM_COUPLE_PSC][0] = k_step_pencil<S, LBDEM_COLL_BGK, LBDEM_COUPLE_PSC, false, 1>;                 \
      k.pencil2[LBDEM_COLL_BGK][LBDEM_COUPLE_PSC][0] = k_step_pencil<S, LBDEM_COLL_BGK, LBDEM_COUPLE_PSC, false, 2>;                 \
      k.pencil4[LBDEM_COLL_BGK][LBDEM_COUPLE_PSC][0] = pencil_fn<S, LBDEM_COLL_BGK, LBDEM_COUPLE_PSC, false, 4>();                 \
      k.step[LBDEM_COLL_BGK][LBDEM_COUPLE_PSC][1] = k_step<S, LBDEM_COLL_BGK, LBDEM_COUPLE_PSC, true>;                  \
      k.pencil[LBDEM_COLL_BGK][LBDEM_COUPLE_PSC][1] = k_step_pencil<S, LBDEM_COLL_BGK, LBDEM_COUPLE_PSC, true, 1>;                  \
      k.pencil2[LBDEM_COLL_BGK][LBDEM_COUPLE_PSC][1] = k_step_pencil<S, LBDEM_COLL_BGK, LBDEM_COUPLE_PSC, true, 2>;                  \
      k.pencil4[LBDEM_COLL_BGK][LBDEM_COUPLE_PSC][1] = pencil_fn<S, LBDEM_COLL_BGK, LBDEM_COUPLE_PSC, true, 4>();                  \
      k.step[LBDEM_COLL_BGK][LBDEM_COUPLE_EQIBM][0] = k_step<S, LBDEM_COLL_BGK, LBDEM_COUPLE_EQIBM, false>;             \
      k.pencil[LBDEM_COLL_BGK][LBDEM_COUPLE_EQIBM][0] = k_step_pencil<S, LBDEM_COLL_BGK, LBDEM_COUPLE_EQIBM, false, 1>;             \
      k.pencil2[LBDEM_COLL_BGK][LBDEM_COUPLE_EQIBM][0] = k_step_pencil<S, LBDEM_COLL_BGK, LBDEM_COUPLE_EQIBM, false, 2>;             \
      k.pencil4[LBDEM_COLL_BGK][LBDEM_COUPLE_EQIBM][0] = pencil_fn<S, LBDEM_COLL_BGK, LBDEM_COUPLE_EQIBM, false, 4>();             \
      k.step[LBDEM_COLL_BGK][LBDEM_COUPLE_EQIBM][1] = k_step<S, LBDEM_COLL_BGK, LBDEM_COUPLE_EQIBM, true>;              \
      k.pencil[LBDEM_COLL_BGK][LBDEM_COUPLE_EQIBM][1] = k_step_pencil<S, LBDEM_COLL_BGK, LBDEM_COUPLE_EQIBM, true, 1>;              \
      k.pencil2[LBDEM_COLL_BGK][LBDEM_COUPLE_EQIBM][1] = k_step_pencil<S, LBDEM_COLL_BGK, LBDEM_COUPLE_EQIBM, true, 2>;              \
      k.pencil4[LBDEM_COLL_BGK][LBDEM_COUPLE_EQIBM][1] = pencil_fn<S, LBDEM_COLL_BGK, LBDEM_COUPLE_EQIBM, true, 4>();              \
      k.step[LBDEM_COLL_SMAG_LAGGED][LBDEM_COUPLE_EQIBM][0] = k_step<S, LBDEM_COLL_SMAG_LAGGED, LBDEM_COUPLE_EQIBM, false>; \
      k.pencil[LBDEM_COLL_SMAG_LAGGED][LBDEM_COUPLE_EQIBM][0] = k_step_pencil<S, LBDEM_COLL_SMAG_LAGGED, LBDEM_COUPLE_EQIBM, false, 1>; \
      k.pencil2[LBDEM_COLL_SMAG_LAGGED][LBDEM_COUPLE_EQIBM][0] = k_step_pencil<S, LBDEM_COLL_SMAG_LAGGED, LBDEM_COUPLE_EQIBM, false, 2>; \
      k.pencil4[LBDEM_COLL_SMAG_LAGGED][LBDEM_COUPLE_EQIBM][0] = pencil_fn<S, LBDEM_COLL_SMAG_LAGGED, LBDEM_COUPLE_EQIBM, false, 4>(); \
      k.step[LBDEM_COLL_SMAG_LAGGED][LBDEM_COUPLE_EQIBM][1] = k_step<S, LBDEM_COLL_SMAG_LAGGED, LBDEM_COUPLE_EQIBM, true>;  \
      k.pencil[LBDEM_COLL_SMAG_LAGGED][LBDEM_COUPLE_EQIBM][1] = k_step_pencil<S, LBDEM_COLL_SMAG_LAGGED, LBDEM_COUPLE_EQIBM, true, 1>;  \
      k.pencil2[LBDEM_COLL_SMAG_LAGGED][LBDEM_COUPLE_EQIBM][1] = k_step_pencil<S, LBDEM_COLL_SMAG_LAGGED, LBDEM_COUPLE_EQIBM, true, 2>;  \
      k.pencil4[LBDEM_COLL_SMAG_LAGGED][LBDEM_COUPLE_EQIBM][1] = pencil_fn<S, LBDEM_COLL_SMAG_LAGGED, LBDEM_COUPLE_EQIBM, true, 4>();  \
      k.step[LBDEM_COLL_SMAG_FIXED][LBDEM_COUPLE_EQIBM][0] = k_step<S, LBDEM_COLL_SMAG_FIXED, LBDEM_COUPLE_EQIBM, false>;   \
      k.pencil[LBDEM_COLL_SMAG_FIXED][LBDEM_COUPLE_EQIBM][0] = k_step_pencil<S, LBDEM_COLL_SMAG_FIXED, LBDEM_COUPLE_EQIBM, false, 1>;   \
      k.pencil2[LBDEM_COLL_SMAG_FIXED][LBDEM_COUPLE_EQIBM][0] = k_step_pencil<S, LBDEM_COLL_SMAG_FIXED, LBDEM_COUPLE_EQIBM, false, 2>;   \
      k.pencil4[LBDEM_COLL_SMAG_FIXED][LBDEM_COUPLE_EQIBM][0] = pencil_fn<S, LBDEM_COLL_SMAG_FIXED, LBDEM_COUPLE_EQIBM, false, 4>();   \
      k.step[LBDEM_COLL_SMAG_FIXED][LBDEM_COUPLE_EQIBM][1] = k_step<S, LBDEM_COLL_SMAG_FIXED, LBDEM_COUPLE_EQIBM, true>;    \
      k.pencil[LBDEM_COLL_SMAG_FIXED][LBDEM_COUPLE_EQIBM][1] = k_step_pencil<S, LBDEM_COLL_SMAG_FIXED, LBDEM_COUPLE_EQIBM, true, 1>;    \
      k.pencil2[LBDEM_COLL_SMAG_FIXED][LBDEM_COUPLE_EQIBM][1] = k_step_pencil<S, LBDEM_COLL_SMAG_FIXED, LBDEM_COUPLE_EQIBM, true, 2>;    \
      k.pencil4[LBDEM_COLL_SMAG_FIXED][LBDEM_COUPLE_EQIBM][1] = pencil_fn<S, LBDEM_COLL_SMAG_FIXED, LBDEM_COUPLE_EQIBM, true, 4>();    \
      k.step[LBDEM_COLL_BGK][LBDEM_COUPLE_VANS][0] = k_step<S, LBDEM_COLL_BGK, LBDEM_COUPLE_VANS, false>;               \
      k.pencil[LBDEM_COLL_BGK][LBDEM_COUPLE_VANS][0] = k_step_pencil<S, LBDEM_COLL_BGK, LBDEM_COUPLE_VANS, false, 1>;               \
      k.pencil2[LBDEM_COLL_BGK][LBDEM_COUPLE_VANS][0] = k_step_pencil<S, LBDEM_COLL_BGK, LBDEM_COUPLE_VANS, false, 2>;               \
      k.pencil4[LBDEM_COLL_BGK][LBDEM_COUPLE_VANS][0] = pencil_fn<S, LBDEM_COLL_BGK, LBDEM_COUPLE_VANS, false, 4>();               \
      k.step[LBDEM_COLL_BGK][LBDEM_COUPLE_VANS][1] = k_step<S, LBDEM_COLL_BGK, LBDEM_COUPLE_VANS, true>;                \
      k.pencil[LBDEM_COLL_BGK][LBDEM_COUPLE_VANS][1] = k_step_pencil<S, LBDEM_COLL_BGK, LBDEM_COUPLE_VANS, true, 1>;                \
      k.pencil2[LBDEM_COLL_BGK][LBDEM_COUPLE_VANS][1] = k_step_pencil<S, LBDEM_COLL_BGK, LBDEM_COUPLE_VANS, true, 2>;                \
      k.pencil4[LBDEM_COLL_BGK][LBDEM_COUPLE_VANS][1] = pencil_fn<S, LBDEM_COLL_BGK, LBDEM_COUPLE_VANS, true, 4>();                \
      k.init[LBDEM_COUPLE_NONE] = k_init<S, LBDEM_COUPLE_NONE>;                                 \
      k.init[LBDEM_COUPLE_PSC] = k_init<S, LBDEM_COUPLE_NONE>;                                  \
      k.init[LBDEM_COUPLE_EQIBM] = k_init<S, LBDEM_COUPLE_NONE>;                                \
      k.init[LBDEM_COUPLE_VANS] = k_init<S, LBDEM_COUPLE_VANS>;                                 \
      k.macros[LBDEM_COUPLE_NONE] = k_macros<S, LBDEM_COUPLE_NONE>;                             \
      k.macros[LBDEM_COUPLE_PSC] = k_macros<S, LBDEM_COUPLE_NONE>;                              \
      k.macros[LBDEM_COUPLE_EQIBM] = k_macros<S, LBDEM_COUPLE_NONE>;                            \
      k.macros[LBDEM_COUPLE_VANS] = k_macros<S, LBDEM_COUPLE_VANS>;                             \
      k.init[LBDEM_COUPLE_HYBRID] = k_init<S, LBDEM_COUPLE_NONE>;                               \
      k.macros[LBDEM_COUPLE_HYBRID] = k_macros<S, LBDEM_COUPLE_NONE>;                           \
      ADD_HYBRID(k);                                                                            \
      k.read_f = k_read_f<S>;                                                                   \
      k.write_f = k_write_f<S>;                                                                 \
      k.boundaries = k_boundaries<S>;                                                           \
      k.free_slip_read = k_free_slip_read<S>;                                                   \
      k.free_slip_write = k_free_slip_write<S>;                                                 \
      k.stress = k_stress<S>;                                                                   \
      k.reduce_partial = k_reduce_partial<S>;                                                   \
      return k;                                                                                 \
    }();                                                                                        \
    return t;                                                                                   \
  }                                                                                             \
  }

// Hybrid coupling (lagged and fixed-point Smagorinsky, as BasicLattice3D.computeOmega).
#define LBDEM_HYB_ENTRIES(S, COLL, ODD)                                                          \
  k.step[COLL][LBDEM_COUPLE_HYBRID][ODD] = k_step<S, COLL, LBDEM_COUPLE_HYBRID, ODD>;            \
  k.pencil[COLL][LBDEM_COUPLE_HYBRID][ODD] = k_step_pencil<S, COLL, LBDEM_COUPLE_HYBRID, ODD, 1>; \
  k.pencil2[COLL][LBDEM_COUPLE_HYBRID][ODD] = k_step_pencil<S, COLL, LBDEM_COUPLE_HYBRID, ODD, 2>; \
  k.pencil4[COLL][LBDEM_COUPLE_HYBRID][ODD] = pencil_fn<S, COLL, LBDEM_COUPLE_HYBRID, ODD, 4>();

#define LBDEM_DEFINE_HYBRID_ENTRIES(S, NAME)                                                     \
  namespace lbdem {                                                                             \
  void NAME(LbmKernels &k) {                                                                    \
    LBDEM_HYB_ENTRIES(S, LBDEM_COLL_SMAG_LAGGED, false)                                         \
    LBDEM_HYB_ENTRIES(S, LBDEM_COLL_SMAG_LAGGED, true)                                          \
    LBDEM_HYB_ENTRIES(S, LBDEM_COLL_SMAG_FIXED, false)                                          \
    LBDEM_HYB_ENTRIES(S, LBDEM_COLL_SMAG_FIXED, true)                                           \
  }                                                                                             \
  }
