// Kernel instantiations and launchers for the f16s storage kind.
#include "flb_kernels.hpp"

FLB_DEFINE_LAUNCHERS(f16s, StoreF16S)
