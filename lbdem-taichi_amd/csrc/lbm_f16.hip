// Collide-stream kernels instantiated for _Float16 population storage.
#include "lbm_kernels.hpp"
#include "lbm_table.hpp"
LBDEM_DEFINE_LBM_TABLE(_Float16, lbm_kernels_f16, add_hybrid_f16)
