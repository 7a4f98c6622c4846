// The VPlacedVolume virtuals that return a local point or take one
// (volumes/PlacedVolume.h:153-160), in basket form: Contains(point,
// localPoint&) and UnplacedContains(localPoint), one point per lane.
//
// Contains(point, localPoint&) transforms the point with the placement
// (Transform<transCode, rotCode>, base/Transformation3D.h:537-562) and
// returns the unplaced shape's Contains of it together with the local point
// (ShapeImplementationHelper.h:146-153); UnplacedContains is the unplaced
// shape's Contains of a local point (:139-144).  The shape is dispatched by
// its type tag, as the navigator does.
#include <hip/hip_runtime.h>

#include "shapes_boolean.hpp"
#include "vg_nav_impl.hpp"

namespace vga {

namespace {

using namespace nav;

constexpr int kExtraBlock = 256;

struct ExtraArgs {
  const double *x, *y, *z;
  uint8_t *out;
  double *lx, *ly, *lz;  // local points (ContainsLocal), nullable each
  const double *blk;
  Xf xf;
  uint64_t n;
  int32_t type, xfkind;
  int32_t local_in;  // 1: the points are local (UnplacedContains)
  int32_t pad;
};

__global__ __launch_bounds__(kExtraBlock) void ContainsExtraKernel(const ExtraArgs a) {
  const uint64_t i = (uint64_t)blockIdx.x * kExtraBlock + threadIdx.x;
  if (i >= a.n) return;
  const V3 p{a.x[i], a.y[i], a.z[i]};
  const V3 l = a.local_in ? p : XfPointRt(a.xf, a.xfkind, p);
  a.out[i] = (a.type == VG_BOOLEAN ? BooleanS::Contains(BooleanP{a.blk}, l) : DContains(a.type, a.blk, l)) ? 1 : 0;
  if (a.lx) a.lx[i] = l.x;
  if (a.ly) a.ly[i] = l.y;
  if (a.lz) a.lz[i] = l.z;
}

}  // namespace

int LaunchContainsExtra(const vg_geometry *g, int32_t placed, const double *x, const double *y, const double *z,
                        uint64_t n, uint8_t *out, double *lx, double *ly, double *lz, bool local_in, hipStream_t s) {
  if (n == 0) return VG_OK;
  const PlacedRec &r = g->placed[placed];
  ExtraArgs a{x, y, z, out, lx, ly, lz, g->d_params + r.param_off, r.xf, n, r.type, r.xfkind, local_in ? 1 : 0, 0};
  const unsigned blocks = (unsigned)((n + kExtraBlock - 1) / kExtraBlock);
  hipLaunchKernelGGL(ContainsExtraKernel, dim3(blocks), dim3(kExtraBlock), 0, s, a);
  return CheckHip(hipGetLastError(), "contains launch");
}

}  // namespace vga
