// C-ABI entry points of libvecgeom_amd: device set-up, the geometry
// uploader (the replacement of CudaManager), and the basket shape queries.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "vg_internal.hpp"

namespace vga {

namespace {
int g_device = -1;
}  // namespace

int EnsureDevice() {
  if (g_device < 0) return SetError(VG_ERR_STATE, "vg_init has not been called");
  int cur = -1;
  if (hipGetDevice(&cur) != hipSuccess) return SetError(VG_ERR_HIP, "hipGetDevice failed");
  if (cur != g_device) {
    if (hipSetDevice(g_device) != hipSuccess) return SetError(VG_ERR_HIP, "hipSetDevice failed");
  }
  return VG_OK;
}

namespace {

void FreeDevice(vg_geometry *g) {
  if (g->d_params) (void)hipFree(g->d_params);
  if (g->d_placed) (void)hipFree(g->d_placed);
  if (g->d_daughters) (void)hipFree(g->d_daughters);
  if (g->d_aabb) (void)hipFree(g->d_aabb);
  if (g->d_b32) (void)hipFree(g->d_b32);
  if (g->d_b32side) (void)hipFree(g->d_b32side);
  if (g->d_extent32) (void)hipFree(g->d_extent32);
  if (g->d_slot_aabb) (void)hipFree(g->d_slot_aabb);
  if (g->d_scalar) (void)hipFree(g->d_scalar);
  g->d_slot_aabb = nullptr;
  g->d_scalar = nullptr;
  g->d_b32 = nullptr;
  g->d_b32side = nullptr;
  g->d_extent32 = nullptr;
  g->d_params = nullptr;
  g->d_placed = nullptr;
  g->d_daughters = nullptr;
  g->d_aabb = nullptr;
}

template <class T>
int Upload(const std::vector<T> &h, T **d) {
  const size_t bytes = std::max<size_t>(h.size() * sizeof(T), 16);
  int rc = CheckHip(hipMalloc(reinterpret_cast<void **>(d), bytes), "hipMalloc(geometry)");
  if (rc) return rc;
  if (!h.empty()) rc = CheckHip(hipMemcpy(*d, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice), "hipMemcpy(geometry)");
  return rc;
}

bool ValidPlaced(const vg_geometry *g, int32_t placed) {
  return g && placed >= 0 && (size_t)placed < g->placed.size();
}

}  // namespace
}  // namespace vga

using namespace vga;

extern "C" {

const char *vg_last_error(void) { return LastError(); }

const char *vg_version(void) { return "vecgeom_amd 0.1 (gfx950)"; }

int vg_init(int device) {
  int n = 0;
  int rc = CheckHip(hipGetDeviceCount(&n), "hipGetDeviceCount");
  if (rc) return rc;
  if (device < 0 || device >= n) return SetError(VG_ERR_INVALID, "device index out of range");
  if (g_device >= 0 && g_device != device) return SetError(VG_ERR_STATE, "vg_init called with a different device");
  rc = CheckHip(hipSetDevice(device), "hipSetDevice");
  if (rc) return rc;
  g_device = device;
  return VG_OK;
}

int vg_shutdown(void) {
  g_device = -1;
  return VG_OK;
}

int vg_shape_supported(int32_t shape_type) { return ShapeHasKernels(shape_type) ? 1 : 0; }

int vg_set_kernel_variant(int variant) { return SetKernelVariant(variant); }

int vg_geometry_upload(const vg_flat_geometry *f, vg_geometry **out) {
  if (!f || !out) return SetError(VG_ERR_INVALID, "null argument");
  *out = nullptr;
  int rc = EnsureDevice();
  if (rc) return rc;
  auto g = new vg_geometry;
  g->device = g_device;
  rc = BuildGeometry(f, g);
  if (rc) {
    delete g;
    return rc;
  }
  // Device copies.
  std::vector<float> faabb(g->aabb.begin(), g->aabb.end());
  rc = Upload(g->params, &g->d_params);
  if (!rc) rc = Upload(g->placed, &g->d_placed);
  if (!rc) rc = Upload(g->daughters, &g->d_daughters);
  if (!rc) rc = Upload(faabb, &g->d_aabb);
  if (!rc) rc = Upload(g->b32, &g->d_b32);
  if (!rc) rc = Upload(g->b32side, &g->d_b32side);
  if (!rc) rc = Upload(g->extent32, &g->d_extent32);
  if (!rc) rc = Upload(g->slot_aabb, &g->d_slot_aabb);
  if (!rc) rc = Upload(std::vector<double>(16, 0.), &g->d_scalar);
  if (rc) {
    FreeDevice(g);
    delete g;
    return rc;
  }
  *out = g;
  return VG_OK;
}

int vg_geometry_free(vg_geometry *g) {
  if (!g) return SetError(VG_ERR_INVALID, "null geometry");
  FreeDevice(g);
  delete g;
  return VG_OK;
}

int vg_geometry_max_depth(const vg_geometry *g, int32_t *depth) {
  if (!g || !depth) return SetError(VG_ERR_INVALID, "null argument");
  *depth = g->max_depth;
  return VG_OK;
}

int vg_geometry_placed_extent(const vg_geometry *g, int32_t placed, double *lo, double *hi) {
  if (!ValidPlaced(g, placed) || !lo || !hi) return SetError(VG_ERR_INVALID, "bad placed id or null argument");
  const PlacedRec &r = g->placed[placed];
  if (!ShapeExtent(r.type, g->params.data() + r.param_off, lo, hi))
    return SetError(VG_ERR_UNSUPPORTED, "no extent for shape type");
  return VG_OK;
}

int vg_geometry_placed_aabb(const vg_geometry *g, int32_t placed, double *lo, double *hi) {
  if (!ValidPlaced(g, placed) || !lo || !hi) return SetError(VG_ERR_INVALID, "bad placed id or null argument");
  for (int i = 0; i < 3; ++i) {
    lo[i] = g->aabb[6 * (size_t)placed + i];
    hi[i] = g->aabb[6 * (size_t)placed + 3 + i];
  }
  return VG_OK;
}

static int RunQuery(const vg_geometry *g, int32_t placed, int q, QArgs a, void *stream) {
  if (!ValidPlaced(g, placed)) return SetError(VG_ERR_INVALID, "bad geometry handle or placed id");
  if (a.n == 0) return VG_OK;
  if (!a.x || !a.y || !a.z || !a.out) return SetError(VG_ERR_INVALID, "null point or output array");
  const bool dir = q == kQDistIn || q == kQDistInMin || q == kQDistOut || q == kQPlacedDistOut;
  if (dir && (!a.dx || !a.dy || !a.dz)) return SetError(VG_ERR_INVALID, "null direction array");
  if (q == kQDistInMin && !a.next) return SetError(VG_ERR_INVALID, "null next-daughter array");
  int rc = EnsureDevice();
  if (rc) return rc;
  return LaunchQuery(g, placed, q, a, static_cast<hipStream_t>(stream));
}

static QArgs Args(const double *x, const double *y, const double *z, const double *dx, const double *dy,
                  const double *dz, const double *sm, void *out, int32_t *next, int32_t did, size_t n) {
  QArgs a;
  a.x = x; a.y = y; a.z = z;
  a.dx = dx; a.dy = dy; a.dz = dz;
  a.sm = sm;
  a.out = out;
  a.next = next;
  a.did = did;
  a.pad = 0;
  a.n = n;
  return a;
}

int vg_contains_local(const vg_geometry *g, int32_t placed, const double *x, const double *y, const double *z,
                      size_t n, uint8_t *out, double *lx, double *ly, double *lz, void *stream) {
  if (!ValidPlaced(g, placed)) return SetError(VG_ERR_INVALID, "bad geometry handle or placed id");
  if (n == 0) return VG_OK;
  if (!x || !y || !z || !out) return SetError(VG_ERR_INVALID, "null point or output array");
  if (!ShapeHasKernels(g->placed[placed].type)) return SetError(VG_ERR_UNSUPPORTED, "no device kernels for the shape");
  int rc = EnsureDevice();
  if (rc) return rc;
  return LaunchContainsExtra(g, placed, x, y, z, n, out, lx, ly, lz, false, static_cast<hipStream_t>(stream));
}

int vg_unplaced_contains(const vg_geometry *g, int32_t placed, const double *x, const double *y, const double *z,
                         size_t n, uint8_t *out, void *stream) {
  if (!ValidPlaced(g, placed)) return SetError(VG_ERR_INVALID, "bad geometry handle or placed id");
  if (n == 0) return VG_OK;
  if (!x || !y || !z || !out) return SetError(VG_ERR_INVALID, "null point or output array");
  if (!ShapeHasKernels(g->placed[placed].type)) return SetError(VG_ERR_UNSUPPORTED, "no device kernels for the shape");
  int rc = EnsureDevice();
  if (rc) return rc;
  return LaunchContainsExtra(g, placed, x, y, z, n, out, nullptr, nullptr, nullptr, true,
                             static_cast<hipStream_t>(stream));
}

int vg_normal(const vg_geometry *g, int32_t placed, const double *x, const double *y, const double *z, size_t n,
              double *nx, double *ny, double *nz, uint8_t *valid, void *stream) {
  if (!ValidPlaced(g, placed)) return SetError(VG_ERR_INVALID, "bad geometry handle or placed id");
  if (n == 0) return VG_OK;
  if (!x || !y || !z || !nx || !ny || !nz || !valid) return SetError(VG_ERR_INVALID, "null point or output array");
  if (!ShapeHasNormal(g->placed[placed].type))
    return SetError(VG_ERR_UNSUPPORTED, "the reference has no Normal for this shape type");
  int rc = EnsureDevice();
  if (rc) return rc;
  return LaunchNormal(g, placed, x, y, z, nullptr, nullptr, nullptr, nullptr, n, nullptr, nx, ny, nz, valid, nullptr,
                      static_cast<hipStream_t>(stream));
}

int vg_distance_to_out_normal(const vg_geometry *g, int32_t placed, const double *x, const double *y,
                              const double *z, const double *dx, const double *dy, const double *dz,
                              const double *step_max, size_t n, double *out, double *nx, double *ny, double *nz,
                              uint8_t *valid, uint8_t *convex, void *stream) {
  if (!ValidPlaced(g, placed)) return SetError(VG_ERR_INVALID, "bad geometry handle or placed id");
  if (n == 0) return VG_OK;
  if (!x || !y || !z || !dx || !dy || !dz || !out || !nx || !ny || !nz || !valid)
    return SetError(VG_ERR_INVALID, "null point, direction or output array");
  if (!ShapeHasNormal(g->placed[placed].type))
    return SetError(VG_ERR_UNSUPPORTED, "the reference has no Normal for this shape type");
  int rc = EnsureDevice();
  if (rc) return rc;
  return LaunchNormal(g, placed, x, y, z, dx, dy, dz, step_max, n, out, nx, ny, nz, valid, convex,
                      static_cast<hipStream_t>(stream));
}

int vg_scalar_query(vg_geometry *g, int32_t placed, int32_t query, const double point[3], const double dir[3],
                    double step_max, double *result, double local[3]) {
  if (!ValidPlaced(g, placed)) return SetError(VG_ERR_INVALID, "bad geometry handle or placed id");
  if (!point || !result) return SetError(VG_ERR_INVALID, "null point or result");
  const bool directional =
      query == VG_Q_DISTANCE_TO_IN || query == VG_Q_DISTANCE_TO_OUT || query == VG_Q_PLACED_DISTANCE_TO_OUT;
  if (directional && !dir) return SetError(VG_ERR_INVALID, "null direction");
  if (query == VG_Q_CONTAINS_LOCAL && !local) return SetError(VG_ERR_INVALID, "null local point");
  if (query < VG_Q_CONTAINS || query > VG_Q_UNPLACED_CONTAINS) return SetError(VG_ERR_INVALID, "unknown query");
  int rc = EnsureDevice();
  if (rc) return rc;
  std::lock_guard<std::mutex> lock(g->scalar_mutex);
  double h[16] = {};
  for (int i = 0; i < 3; ++i) h[i] = point[i];
  if (directional)
    for (int i = 0; i < 3; ++i) h[3 + i] = dir[i];
  h[6] = step_max;
  double *d = g->d_scalar;
  if ((rc = CheckHip(hipMemcpy(d, h, 7 * sizeof(double), hipMemcpyHostToDevice), "scalar query upload"))) return rc;
  void *o = d + 8;
  switch (query) {
    case VG_Q_CONTAINS: rc = vg_contains(g, placed, d, d + 1, d + 2, 1, static_cast<uint8_t *>(o), nullptr); break;
    case VG_Q_INSIDE: rc = vg_inside(g, placed, d, d + 1, d + 2, 1, static_cast<int32_t *>(o), nullptr); break;
    case VG_Q_DISTANCE_TO_IN:
      rc = vg_distance_to_in(g, placed, d, d + 1, d + 2, d + 3, d + 4, d + 5, d + 6, 1, d + 8, nullptr);
      break;
    case VG_Q_DISTANCE_TO_OUT:
      rc = vg_distance_to_out(g, placed, d, d + 1, d + 2, d + 3, d + 4, d + 5, d + 6, 1, d + 8, nullptr, nullptr);
      break;
    case VG_Q_PLACED_DISTANCE_TO_OUT:
      rc = vg_placed_distance_to_out(g, placed, d, d + 1, d + 2, d + 3, d + 4, d + 5, d + 6, 1, d + 8, nullptr);
      break;
    case VG_Q_SAFETY_TO_IN: rc = vg_safety_to_in(g, placed, d, d + 1, d + 2, 1, d + 8, nullptr); break;
    case VG_Q_SAFETY_TO_OUT: rc = vg_safety_to_out(g, placed, d, d + 1, d + 2, 1, d + 8, nullptr); break;
    case VG_Q_CONTAINS_LOCAL:
      rc = vg_contains_local(g, placed, d, d + 1, d + 2, 1, static_cast<uint8_t *>(o), d + 9, d + 10, d + 11, nullptr);
      break;
    default:  // VG_Q_UNPLACED_CONTAINS
      rc = vg_unplaced_contains(g, placed, d, d + 1, d + 2, 1, static_cast<uint8_t *>(o), nullptr);
      break;
  }
  if (rc) return rc;
  if ((rc = CheckHip(hipMemcpy(h + 8, d + 8, 4 * sizeof(double), hipMemcpyDeviceToHost), "scalar query result")))
    return rc;
  if (query == VG_Q_CONTAINS || query == VG_Q_CONTAINS_LOCAL || query == VG_Q_UNPLACED_CONTAINS) {
    uint8_t c;
    std::memcpy(&c, h + 8, 1);
    *result = c ? 1. : 0.;
  } else if (query == VG_Q_INSIDE) {
    int32_t v;
    std::memcpy(&v, h + 8, 4);
    *result = (double)v;
  } else {
    *result = h[8];
  }
  if (query == VG_Q_CONTAINS_LOCAL)
    for (int i = 0; i < 3; ++i) local[i] = h[9 + i];
  return VG_OK;
}

int vg_contains(const vg_geometry *g, int32_t placed, const double *x, const double *y, const double *z, size_t n,
                uint8_t *out, void *stream) {
  return RunQuery(g, placed, kQContains, Args(x, y, z, 0, 0, 0, 0, out, 0, 0, n), stream);
}

int vg_inside(const vg_geometry *g, int32_t placed, const double *x, const double *y, const double *z, size_t n,
              int32_t *out, void *stream) {
  return RunQuery(g, placed, kQInside, Args(x, y, z, 0, 0, 0, 0, out, 0, 0, n), stream);
}

int vg_distance_to_in(const vg_geometry *g, int32_t placed, const double *x, const double *y, const double *z,
                      const double *dx, const double *dy, const double *dz, const double *step_max, size_t n,
                      double *out, void *stream) {
  return RunQuery(g, placed, kQDistIn, Args(x, y, z, dx, dy, dz, step_max, out, 0, 0, n), stream);
}

int vg_distance_to_in_minimize(const vg_geometry *g, int32_t placed, int32_t daughter_index, const double *x,
                               const double *y, const double *z, const double *dx, const double *dy, const double *dz,
                               size_t n, double *inout_dist, int32_t *inout_next, void *stream) {
  return RunQuery(g, placed, kQDistInMin, Args(x, y, z, dx, dy, dz, 0, inout_dist, inout_next, daughter_index, n),
               stream);
}

int vg_distance_to_out(const vg_geometry *g, int32_t placed, const double *x, const double *y, const double *z,
                       const double *dx, const double *dy, const double *dz, const double *step_max, size_t n,
                       double *out, int32_t *next_node, void *stream) {
  return RunQuery(g, placed, kQDistOut, Args(x, y, z, dx, dy, dz, step_max, out, next_node, 0, n), stream);
}

int vg_placed_distance_to_out(const vg_geometry *g, int32_t placed, const double *x, const double *y, const double *z,
                              const double *dx, const double *dy, const double *dz, const double *step_max, size_t n,
                              double *out, void *stream) {
  return RunQuery(g, placed, kQPlacedDistOut, Args(x, y, z, dx, dy, dz, step_max, out, 0, 0, n), stream);
}

int vg_safety_to_in(const vg_geometry *g, int32_t placed, const double *x, const double *y, const double *z, size_t n,
                    double *out, void *stream) {
  return RunQuery(g, placed, kQSafeIn, Args(x, y, z, 0, 0, 0, 0, out, 0, 0, n), stream);
}

int vg_safety_to_in_minimize(const vg_geometry *g, int32_t placed, const double *x, const double *y, const double *z,
                             size_t n, double *inout, void *stream) {
  return RunQuery(g, placed, kQSafeInMin, Args(x, y, z, 0, 0, 0, 0, inout, 0, 0, n), stream);
}

int vg_safety_to_out(const vg_geometry *g, int32_t placed, const double *x, const double *y, const double *z, size_t n,
                     double *out, void *stream) {
  return RunQuery(g, placed, kQSafeOut, Args(x, y, z, 0, 0, 0, 0, out, 0, 0, n), stream);
}

int vg_safety_to_out_minimize(const vg_geometry *g, int32_t placed, const double *x, const double *y, const double *z,
                              size_t n, double *inout, void *stream) {
  return RunQuery(g, placed, kQSafeOutMin, Args(x, y, z, 0, 0, 0, 0, inout, 0, 0, n), stream);
}

}  // extern "C"
