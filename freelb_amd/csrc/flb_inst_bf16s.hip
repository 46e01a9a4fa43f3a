// Kernel instantiations and launchers for the bf16s storage kind.
#include "flb_kernels.hpp"

FLB_DEFINE_LAUNCHERS(bf16s, StoreBF16S)
