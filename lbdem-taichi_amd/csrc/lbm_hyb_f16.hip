// Collide-stream kernels of the hybrid coupling instantiated for _Float16 population storage.
#include "lbm_kernels.hpp"
#include "lbm_table.hpp"
LBDEM_DEFINE_HYBRID_ENTRIES(_Float16, add_hybrid_f16)
