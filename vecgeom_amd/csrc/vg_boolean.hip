// Basket queries on Boolean volumes (shapes_boolean.hpp), one query per
// lane.  The Boolean's placement is applied as the reference's
// SpecializedBooleanVolume does (Transform<transCode, rotCode> for the point,
// TransformDirection for the direction); the operands' parameter blocks are
// read from global memory through the Boolean's block (they are small and
// stay in the caches).  Query semantics, including the Minimize forms and the
// DistanceToOut next-node output, are those of the other shapes' basket
// kernels (volumes/ShapeImplementationHelper.h:588-748).
#include <hip/hip_runtime.h>

#include "shapes_boolean.hpp"
#include "vg_internal.hpp"

namespace vga {

namespace {

constexpr int kBoolBlock = 128;

template <int Q>
__global__ __launch_bounds__(kBoolBlock) void BooleanQueryKernel(const double *__restrict__ blk, const Xf X,
                                                                 int xfkind, const QArgs a) {
  const uint64_t i = (uint64_t)blockIdx.x * kBoolBlock + threadIdx.x;
  if (i >= a.n) return;
  const BooleanP P{blk};
  const V3 p{a.x[i], a.y[i], a.z[i]};
  V3 d{0, 0, 0};
  if constexpr (Q == kQDistIn || Q == kQDistInMin || Q == kQDistOut || Q == kQPlacedDistOut)
    d = V3{a.dx[i], a.dy[i], a.dz[i]};
  double *od = reinterpret_cast<double *>(a.out);
  if constexpr (Q == kQContains) {
    reinterpret_cast<uint8_t *>(a.out)[i] = BooleanS::Contains(P, XfPointRt(X, xfkind, p)) ? 1 : 0;
  } else if constexpr (Q == kQInside) {
    reinterpret_cast<int32_t *>(a.out)[i] = BooleanS::Inside(P, XfPointRt(X, xfkind, p));
  } else if constexpr (Q == kQDistIn) {
    const double sm = a.sm ? a.sm[i] : kInfinity;
    od[i] = BooleanS::DistanceToIn(P, XfPointRt(X, xfkind, p), XfDirRt(X, xfkind, d), sm);
  } else if constexpr (Q == kQDistInMin) {
    const double cur = od[i];
    const double r = BooleanS::DistanceToIn(P, XfPointRt(X, xfkind, p), XfDirRt(X, xfkind, d), cur);
    if (r < cur) {
      od[i] = r;
      a.next[i] = a.did;
    }
  } else if constexpr (Q == kQDistOut) {
    const double sm = a.sm ? a.sm[i] : kInfinity;
    const double r = BooleanS::DistanceToOut(P, p, d, sm);
    od[i] = r;
    if (a.next) a.next[i] = (r < sm) ? -1 : -2;
  } else if constexpr (Q == kQPlacedDistOut) {
    const double sm = a.sm ? a.sm[i] : kInfinity;
    od[i] = BooleanS::DistanceToOut(P, XfPointRt(X, xfkind, p), XfDirRt(X, xfkind, d), sm);
  } else if constexpr (Q == kQSafeIn) {
    od[i] = BooleanS::SafetyToIn(P, XfPointRt(X, xfkind, p));
  } else if constexpr (Q == kQSafeInMin) {
    const double cur = od[i], r = BooleanS::SafetyToIn(P, XfPointRt(X, xfkind, p));
    od[i] = (r < cur) ? r : cur;
  } else if constexpr (Q == kQSafeOut) {
    od[i] = BooleanS::SafetyToOut(P, p);
  } else {  // kQSafeOutMin
    const double cur = od[i], r = BooleanS::SafetyToOut(P, p);
    od[i] = (r < cur) ? r : cur;
  }
}

template <int Q>
int Launch(const double *blk, const PlacedRec &pr, const QArgs &a, hipStream_t s) {
  const unsigned blocks = (unsigned)((a.n + kBoolBlock - 1) / kBoolBlock);
  hipLaunchKernelGGL((BooleanQueryKernel<Q>), dim3(blocks), dim3(kBoolBlock), 0, s, blk, pr.xf, (int)pr.xfkind, a);
  return CheckHip(hipGetLastError(), "Boolean query launch");
}

}  // namespace

int LaunchBooleanQuery(const vg_geometry *g, int32_t placed, int q, const QArgs &a, hipStream_t s) {
  if (a.n == 0) return VG_OK;
  const PlacedRec &pr = g->placed[placed];
  const double *blk = g->d_params + pr.param_off;
  switch (q) {
    case kQContains: return Launch<kQContains>(blk, pr, a, s);
    case kQInside: return Launch<kQInside>(blk, pr, a, s);
    case kQDistIn: return Launch<kQDistIn>(blk, pr, a, s);
    case kQDistInMin: return Launch<kQDistInMin>(blk, pr, a, s);
    case kQDistOut: return Launch<kQDistOut>(blk, pr, a, s);
    case kQPlacedDistOut: return Launch<kQPlacedDistOut>(blk, pr, a, s);
    case kQSafeIn: return Launch<kQSafeIn>(blk, pr, a, s);
    case kQSafeInMin: return Launch<kQSafeInMin>(blk, pr, a, s);
    case kQSafeOut: return Launch<kQSafeOut>(blk, pr, a, s);
    case kQSafeOutMin: return Launch<kQSafeOutMin>(blk, pr, a, s);
  }
  return SetError(VG_ERR_INVALID, "unknown query");
}

}  // namespace vga
