// Collide-stream kernels of the hybrid coupling instantiated for __bf16 population storage.
#include "lbm_kernels.hpp"
#include "lbm_table.hpp"
LBDEM_DEFINE_HYBRID_ENTRIES(__bf16, add_hybrid_bf16)
