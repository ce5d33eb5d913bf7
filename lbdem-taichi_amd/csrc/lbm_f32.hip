// Collide-stream kernels instantiated for float population storage.
#include "lbm_kernels.hpp"
#include "lbm_table.hpp"
LBDEM_DEFINE_LBM_TABLE(float, lbm_kernels_f32, add_hybrid_f32)
