// Copy of a shape parameter block from global memory to LDS by a whole
// workgroup (gfx950).
#pragma once

#include <hip/hip_runtime.h>

namespace vga {

// dst (LDS) and src (global) 16-byte aligned.  16-byte loads, four in flight
// per thread before their stores, so a workgroup waits about one L2 round
// trip for a block of up to 4 * 2 * nthreads doubles instead of one per
// 8-byte element per thread.
__device__ __forceinline__ void StageBlock(double *__restrict__ dst, const double *__restrict__ src, int len,
                                           int nthreads) {
  const int n2 = len >> 1;
  const double2 *s2 = reinterpret_cast<const double2 *>(src);
  double2 *d2 = reinterpret_cast<double2 *>(dst);
  int k = threadIdx.x;
  for (; k + 3 * nthreads < n2; k += 4 * nthreads) {
    const double2 a = s2[k], b = s2[k + nthreads], c = s2[k + 2 * nthreads], e = s2[k + 3 * nthreads];
    d2[k] = a;
    d2[k + nthreads] = b;
    d2[k + 2 * nthreads] = c;
    d2[k + 3 * nthreads] = e;
  }
  for (; k < n2; k += nthreads) d2[k] = s2[k];
  if ((len & 1) && threadIdx.x == 0) dst[len - 1] = src[len - 1];
}

}  // namespace vga
