// Convergence residual of the RHO / VELOCITY fields (BlockLattice::getToleranceRho /
// getToleranceU / getTolRho(shift) / getTolU(shift), data_struct/block_lattice.hh:252-367):
// the maximum over the cells [shift, n - shift) per axis of |rho - rho_old|
// (of max_a |u_a - u_old_a|), after which the old field takes the current one.
// The maximum of non-negative doubles is order-free: each thread folds its cell
// into one 64-bit word with an atomic max on the value's bits (IEEE order of
// non-negative doubles is the integer order of their bits), so the result is
// the same bits in every run. A NaN difference never wins, as in the
// reference's `res > maxres ? res : maxres`.

#include <hip/hip_runtime.h>

#include "flb_internal.hpp"

namespace flb {
namespace {

template <class R>
__global__ __launch_bounds__(kMaxThreads) void k_tolerance(const R* __restrict__ field, R* __restrict__ old,
                                                           int64_t fstride, int ncomp, int nx, int ny, int nz,
                                                           int xlo, int pitch, int64_t plane, int shift,
                                                           unsigned long long* __restrict__ res) {
  const int64_t k = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  const int ex = nx - 2 * shift, ey = ny - 2 * shift, ez = nz == 1 ? 1 : nz - 2 * shift;
  if (ex <= 0 || ey <= 0 || ez <= 0 || k >= int64_t(ex) * ey * ez) return;
  const int x = shift + int(k % ex);
  const int64_t r = k / ex;
  const int y = shift + int(r % ey);
  const int z = (nz == 1 ? 0 : shift) + int(r / ey);
  const int64_t slot = int64_t(x - xlo) + int64_t(pitch) * y + plane * z;
  R m = R(0);
  bool nan = false;
  for (int c = 0; c < ncomp; ++c) {
    const R v = field[c * fstride + slot];
    const R d = v - old[c * fstride + slot];
    const R a = d < R(0) ? -d : d;
    nan = nan || !(a == a);
    m = a > m ? a : m;
    old[c * fstride + slot] = v;
  }
  if (!nan && m > R(0)) atomicMax(res, static_cast<unsigned long long>(__double_as_longlong(double(m))));
}

}  // namespace

hipError_t launch_tolerance(bool dbl, const void* field, void* old, int64_t fstride, int ncomp, int nx, int ny,
                            int nz, int xlo, int pitch, int64_t plane, int shift, unsigned long long* res,
                            hipStream_t s) {
  const int64_t n = int64_t(nx) * ny * nz;
  const unsigned nb = unsigned((n + kMaxThreads - 1) / kMaxThreads);
  if (dbl)
    hipLaunchKernelGGL(k_tolerance<double>, dim3(nb), dim3(kMaxThreads), 0, s, static_cast<const double*>(field),
                       static_cast<double*>(old), fstride, ncomp, nx, ny, nz, xlo, pitch, plane, shift, res);
  else
    hipLaunchKernelGGL(k_tolerance<float>, dim3(nb), dim3(kMaxThreads), 0, s, static_cast<const float*>(field),
                       static_cast<float*>(old), fstride, ncomp, nx, ny, nz, xlo, pitch, plane, shift, res);
  return hipGetLastError();
}

}  // namespace flb
