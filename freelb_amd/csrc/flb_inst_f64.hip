// Kernel instantiations and launchers for the f64 storage kind.
#include "flb_kernels.hpp"

FLB_DEFINE_LAUNCHERS(f64, StoreF64)
