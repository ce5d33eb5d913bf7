// Collide-stream kernels of the hybrid coupling instantiated for double population storage.
#include "lbm_kernels.hpp"
#include "lbm_table.hpp"
LBDEM_DEFINE_HYBRID_ENTRIES(double, add_hybrid_f64)
