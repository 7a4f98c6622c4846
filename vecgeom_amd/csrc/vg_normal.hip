// Basket forms of the two normal queries, one point per lane:
//
//  * Normal(localPoint, normal&) -- VPlacedVolume::Normal
//    (volumes/PlacedVolume.h:266-272 and the shapes' overrides): the normal
//    (shapes_normal.hpp) and its valid flag;
//  * the USolids DistanceToOut(point, direction, normal&, convex&, stepMax)
//    (volumes/ShapeImplementationHelper.h:277-295, the same in
//    ScalarShapeImplementationHelper.h:249-264): the shape's DistanceToOut d,
//    then the Normal of point + d * direction (the vector scaled, then added
//    component by component), convex always true.
//
// Both take local-frame inputs, as the reference's do.  The shape is a
// template parameter: one launch covers one placed volume, so its parameter
// block is the same for every lane.
#include <hip/hip_runtime.h>

#include "shapes_normal.hpp"
#include "vg_internal.hpp"

namespace vga {

namespace {

constexpr int kNormalBlock = 256;

struct NormalArgs {
  const double *x, *y, *z;
  const double *dx, *dy, *dz;  // DistanceToOut form only
  const double *sm;            // nullable: +inf
  double *dist;                // DistanceToOut form only
  double *nx, *ny, *nz;
  uint8_t *valid;
  uint8_t *convex;  // nullable
  uint64_t n;
};

template <class S, bool WithDistance>
__global__ __launch_bounds__(kNormalBlock) void NormalKernel(const double *__restrict__ blk, const NormalArgs a) {
  const uint64_t i = (uint64_t)blockIdx.x * kNormalBlock + threadIdx.x;
  if (i >= a.n) return;
  const auto &P = ParamView<S>::Get(blk);
  V3 p{a.x[i], a.y[i], a.z[i]};
  if constexpr (WithDistance) {
    const V3 d{a.dx[i], a.dy[i], a.dz[i]};
    const double dist = S::DistanceToOut(P, p, d, a.sm ? a.sm[i] : kInfinity);
    a.dist[i] = dist;
    p = V3{p.x + d.x * dist, p.y + d.y * dist, p.z + d.z * dist};
    if (a.convex) a.convex[i] = 1;
  }
  V3 nrm{0, 0, 0};
  const bool ok = ShapeNormal(P, p, nrm);
  a.nx[i] = nrm.x;
  a.ny[i] = nrm.y;
  a.nz[i] = nrm.z;
  a.valid[i] = ok ? 1 : 0;
}

template <class S>
int Launch(const double *blk, const NormalArgs &a, bool with_distance, hipStream_t s) {
  const unsigned blocks = (unsigned)((a.n + kNormalBlock - 1) / kNormalBlock);
  if (with_distance)
    hipLaunchKernelGGL((NormalKernel<S, true>), dim3(blocks), dim3(kNormalBlock), 0, s, blk, a);
  else
    hipLaunchKernelGGL((NormalKernel<S, false>), dim3(blocks), dim3(kNormalBlock), 0, s, blk, a);
  return CheckHip(hipGetLastError(), "normal launch");
}

}  // namespace

bool ShapeHasNormal(int type) {
  switch (type) {
    case VG_BOX: case VG_TUBE: case VG_ORB: case VG_SPHERE: case VG_CONE: case VG_TRD: case VG_TRAPEZOID:
    case VG_POLYCONE:
      return true;
  }
  return false;
}

int LaunchNormal(const vg_geometry *g, int32_t placed, const double *x, const double *y, const double *z,
                 const double *dx, const double *dy, const double *dz, const double *step_max, uint64_t n,
                 double *dist, double *nx, double *ny, double *nz, uint8_t *valid, uint8_t *convex,
                 hipStream_t s) {
  if (n == 0) return VG_OK;
  const PlacedRec &r = g->placed[placed];
  const double *blk = g->d_params + r.param_off;
  const NormalArgs a{x, y, z, dx, dy, dz, step_max, dist, nx, ny, nz, valid, convex, n};
  const bool wd = dist != nullptr;
  switch (r.type) {
    case VG_BOX: return Launch<BoxS>(blk, a, wd, s);
    case VG_TUBE: return Launch<TubeS>(blk, a, wd, s);
    case VG_ORB: return Launch<OrbS>(blk, a, wd, s);
    case VG_SPHERE: return Launch<SphereS>(blk, a, wd, s);
    case VG_CONE: return Launch<ConeS>(blk, a, wd, s);
    case VG_TRD: return Launch<TrdS>(blk, a, wd, s);
    case VG_TRAPEZOID: return Launch<TrapS>(blk, a, wd, s);
    case VG_POLYCONE: return Launch<PolyconeS>(blk, a, wd, s);
  }
  return SetError(VG_ERR_UNSUPPORTED, "the reference has no Normal for this shape type");
}

}  // namespace vga
