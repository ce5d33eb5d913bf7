// Collide-stream kernels instantiated for __bf16 population storage.
#include "lbm_kernels.hpp"
#include "lbm_table.hpp"
LBDEM_DEFINE_LBM_TABLE(__bf16, lbm_kernels_bf16, add_hybrid_bf16)
