// Kernel instantiations and launchers for the f32 storage kind.
#include "flb_kernels.hpp"

FLB_DEFINE_LAUNCHERS(f32, StoreF32)
