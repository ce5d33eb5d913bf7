// Collide-stream kernels of the hybrid coupling instantiated for float population storage.
#include "lbm_kernels.hpp"
#include "lbm_table.hpp"
LBDEM_DEFINE_HYBRID_ENTRIES(float, add_hybrid_f32)
