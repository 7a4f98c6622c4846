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
thread_local std::string g_last_error;
int g_device = -1;
}  // namespace

int SetError(int code, const std::string &msg) {
  g_last_error = msg;
  return code;
}

int CheckHip(hipError_t e, const char *what) {
  if (e == hipSuccess) return VG_OK;
  return SetError(VG_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

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

// Placement kind as Transformation3D::SetProperties decides it
// (source/Transformation3D.cpp:81-90, 139-150).
int PlacementKind(const double *xf12) {
  const bool hasT = std::fabs(xf12[0]) > kTolerance || std::fabs(xf12[1]) > kTolerance || std::fabs(xf12[2]) > kTolerance;
  int code = 0;
  for (int i = 0; i < 9; ++i) code |= (1 << i) * (std::fabs(xf12[3 + i]) > kTolerance);
  if (code == 0x111 && xf12[3] == 1. && xf12[7] == 1. && xf12[11] == 1.) code = 0x200;
  const bool hasR = code != 0x200;
  if (!hasT && !hasR) return kXfIdentity;
  if (hasT && !hasR) return kXfTranslation;
  if (!hasT && hasR) return kXfRotation;
  return kXfGeneral;
}

// ABBoxManager::ComputeABBox, source/ABBoxManager.cpp:128-161: the 8 corners
// of the local extent mapped by InverseTransform
// (base/Transformation3D.h:601-632), padded by 1e-3.
void ComputeAABB(const double lo[3], const double hi[3], const Xf &X, double *out6) {
  const double delta[3] = {hi[0] - lo[0], hi[1] - lo[1], hi[2] - lo[2]};
  double mn[3] = {kInfinity, kInfinity, kInfinity}, mx[3] = {-kInfinity, -kInfinity, -kInfinity};
  for (int ix = 0; ix <= 1; ++ix)
    for (int iy = 0; iy <= 1; ++iy)
      for (int iz = 0; iz <= 1; ++iz) {
        const double c[3] = {lo[0] + ix * delta[0], lo[1] + iy * delta[1], lo[2] + iz * delta[2]};
        double m[3];
        for (int r = 0; r < 3; ++r) {
          m[r] = X.t[r];
          m[r] += c[0] * X.r[3 * r + 0];
          m[r] += c[1] * X.r[3 * r + 1];
          m[r] += c[2] * X.r[3 * r + 2];
        }
        for (int r = 0; r < 3; ++r) {
          mn[r] = std::min(mn[r], m[r]);
          mx[r] = std::max(mx[r], m[r]);
        }
      }
  for (int r = 0; r < 3; ++r) {
    out6[r] = mn[r] - 1E-3;
    out6[3 + r] = mx[r] + 1E-3;
  }
}

int Depth(const vg_geometry *g, int32_t placed, int level, int limit) {
  if (level > limit) return -1;
  const PlacedRec &p = g->placed[placed];
  int best = level;
  for (int k = p.dbegin; k < p.dend; ++k) {
    int d = Depth(g, g->daughters[k], level + 1, limit);
    if (d < 0) return -1;
    best = std::max(best, d);
  }
  return best;
}

void FreeDevice(vg_geometry *g) {
  if (g->d_params) (void)hipFree(g->d_params);
  if (g->d_placed) (void)hipFree(g->d_placed);
  if (g->d_daughters) (void)hipFree(g->d_daughters);
  if (g->d_aabb) (void)hipFree(g->d_aabb);
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

const char *vg_last_error(void) { return g_last_error.c_str(); }

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
  if (f->n_shapes < 0 || f->n_logical < 0 || f->n_placed <= 0)
    return SetError(VG_ERR_INVALID, "geometry needs at least one placed volume");
  if (!f->shape_type || !f->shape_param_begin || !f->logical_shape || !f->logical_daughter_begin ||
      !f->placed_logical || !f->placed_transform)
    return SetError(VG_ERR_INVALID, "null geometry array");
  if (f->world < 0 || f->world >= f->n_placed) return SetError(VG_ERR_INVALID, "world id out of range");

  auto g = new vg_geometry;
  g->device = g_device;
  g->world = f->world;
  // Shapes: parameter blocks.
  g->shapes.resize(f->n_shapes);
  for (int s = 0; s < f->n_shapes; ++s) {
    const int b = f->shape_param_begin[s], e = f->shape_param_begin[s + 1];
    if (b < 0 || e < b || (e > b && !f->params)) {
      delete g;
      return SetError(VG_ERR_INVALID, "bad parameter range for shape " + std::to_string(s));
    }
    std::string err;
    std::vector<double> block = BuildParams(f->shape_type[s], f->params + b, e - b, &err);
    if (block.empty()) {
      delete g;
      return SetError(ShapeHasKernels(f->shape_type[s]) ? VG_ERR_INVALID : VG_ERR_UNSUPPORTED,
                      "shape " + std::to_string(s) + ": " + err);
    }
    if (block.size() > (size_t)kMaxParamDoubles) {
      delete g;
      return SetError(VG_ERR_CAPACITY, "shape " + std::to_string(s) + ": parameter block too large");
    }
    // Keep every block 16-byte aligned within the device array.
    if (g->params.size() % 2) g->params.push_back(0.);
    g->shapes[s] = ShapeRec{f->shape_type[s], (int32_t)g->params.size(), (int32_t)block.size(), 0};
    g->params.insert(g->params.end(), block.begin(), block.end());
  }
  // Logical volumes.
  g->logical_shape.assign(f->logical_shape, f->logical_shape + f->n_logical);
  g->logical_dbegin.assign(f->logical_daughter_begin, f->logical_daughter_begin + f->n_logical + 1);
  const int nd = f->n_logical > 0 ? g->logical_dbegin[f->n_logical] : 0;
  if (nd > 0 && !f->daughters) {
    delete g;
    return SetError(VG_ERR_INVALID, "null daughters array");
  }
  g->daughters.assign(f->daughters, f->daughters + nd);
  for (int l = 0; l < f->n_logical; ++l) {
    if (g->logical_shape[l] < 0 || g->logical_shape[l] >= f->n_shapes || g->logical_dbegin[l] < 0 ||
        g->logical_dbegin[l + 1] < g->logical_dbegin[l] || g->logical_dbegin[l + 1] > nd) {
      delete g;
      return SetError(VG_ERR_INVALID, "bad logical volume " + std::to_string(l));
    }
  }
  for (int k = 0; k < nd; ++k) {
    if (g->daughters[k] < 0 || g->daughters[k] >= f->n_placed) {
      delete g;
      return SetError(VG_ERR_INVALID, "daughter id out of range");
    }
  }
  // Placed volumes.
  g->placed.resize(f->n_placed);
  g->aabb.resize(6 * (size_t)f->n_placed);
  for (int p = 0; p < f->n_placed; ++p) {
    const int l = f->placed_logical[p];
    if (l < 0 || l >= f->n_logical) {
      delete g;
      return SetError(VG_ERR_INVALID, "placed volume " + std::to_string(p) + " has a bad logical id");
    }
    PlacedRec &r = g->placed[p];
    const double *x = f->placed_transform + 12 * (size_t)p;
    r.logical = l;
    r.shape = g->logical_shape[l];
    r.type = g->shapes[r.shape].type;
    r.xfkind = PlacementKind(x);
    r.param_off = g->shapes[r.shape].param_off;
    r.dbegin = g->logical_dbegin[l];
    r.dend = g->logical_dbegin[l + 1];
    r.pad = 0;
    for (int i = 0; i < 3; ++i) r.xf.t[i] = x[i];
    for (int i = 0; i < 9; ++i) r.xf.r[i] = x[3 + i];
    double lo[3], hi[3];
    if (!ShapeExtent(r.type, g->params.data() + r.param_off, lo, hi)) {
      delete g;
      return SetError(VG_ERR_UNSUPPORTED, "no extent for shape type " + std::to_string(r.type));
    }
    ComputeAABB(lo, hi, r.xf, g->aabb.data() + 6 * (size_t)p);
  }
  g->max_depth = Depth(g, g->world, 1, 64);
  if (g->max_depth < 0) {
    delete g;
    return SetError(VG_ERR_INVALID, "volume tree deeper than 64 levels or cyclic");
  }
  // Device copies.
  std::vector<float> faabb(g->aabb.begin(), g->aabb.end());
  rc = Upload(g->params, &g->d_params);
  if (!rc) rc = Upload(g->placed, &g->d_placed);
  if (!rc) rc = Upload(g->daughters, &g->d_daughters);
  if (!rc) rc = Upload(faabb, &g->d_aabb);
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
