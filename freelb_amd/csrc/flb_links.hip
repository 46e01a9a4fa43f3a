// Precomputed halfway wall links (storage independent).
#include "flb_kernels.hpp"

namespace flb {

hipError_t launch_links(int q, const KArgs& a, uint64_t* links, hipStream_t s) {
  const int64_t n = int64_t(a.ex) * a.ey * a.ez;
  const dim3 grid(unsigned((n + kMaxThreads - 1) / kMaxThreads)), block(kMaxThreads);
  const bool per = a.px || a.py || a.pz;
  if (q == 19) {
    if (per) hipLaunchKernelGGL((kern::k_links<19, true>), grid, block, 0, s, a, links);
    else hipLaunchKernelGGL((kern::k_links<19, false>), grid, block, 0, s, a, links);
  } else {
    if (per) hipLaunchKernelGGL((kern::k_links<9, true>), grid, block, 0, s, a, links);
    else hipLaunchKernelGGL((kern::k_links<9, false>), grid, block, 0, s, a, links);
  }
  return hipGetLastError();
}

}  // namespace flb
