// Precomputed halfway wall links (storage independent).
#include "flb_kernels.hpp"

namespace flb {

hipError_t launch_links(int q, const KArgs& a, uint64_t* links, hipStream_t s) {
  const int64_t n = int64_t(a.ex) * a.ey * a.ez;
  const dim3 grid(unsigned((n + kMaxThreads - 1) / kMaxThreads)), block(kMaxThreads);
  const bool per = a.px || a.py || a.pz;
  const bool ok = kern::with_q<StoreF64>(q, [&](auto QC) {
    constexpr int Q = decltype(QC)::value;
    if (per) hipLaunchKernelGGL((kern::k_links<Q, true>), grid, block, 0, s, a, links);
    else hipLaunchKernelGGL((kern::k_links<Q, false>), grid, block, 0, s, a, links);
  });
  return ok ? hipGetLastError() : hipErrorInvalidValue;
}

}  // namespace flb
