// Kernels for the custom 16-bit shifted storage (StoreF16C).
#include "flb_kernels.hpp"

FLB_DEFINE_LAUNCHERS(f16c, StoreF16C)
