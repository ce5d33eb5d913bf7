// Collide-stream kernels instantiated for double population storage.
#include "lbm_kernels.hpp"
#include "lbm_table.hpp"
LBDEM_DEFINE_LBM_TABLE(double, lbm_kernels_f64, add_hybrid_f64)
