// Host-side flattening of a vg_flat_geometry into the records the device
// kernels read (the work CudaManager::Synchronize does for the reference,
// source/CudaManager.cpp:40-160), plus the error state shared by the C-ABI.
// Nothing here touches a device, so the same code also builds geometries for
// host-only checks of the navigation code.
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
}  // namespace

int SetError(int code, const std::string &msg) {
  g_last_error = msg;
  return code;
}

const char *LastError() { return g_last_error.c_str(); }

int CheckHip(hipError_t e, const char *what) {
  if (e == hipSuccess) return VG_OK;
  return SetError(VG_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

bool ShapeHasKernels(int type) {
  switch (type) {
    case VG_BOX: case VG_TUBE: case VG_ORB: case VG_TRD: case VG_PARALLELEPIPED: case VG_PARABOLOID:
    case VG_TRAPEZOID: case VG_SPHERE: case VG_CONE: case VG_POLYCONE: case VG_TORUS: case VG_POLYHEDRON:
    case VG_BOOLEAN:
      return true;
  }
  return false;
}

namespace {

// Placement kind as Transformation3D::SetProperties decides it
// (source/Transformation3D.cpp:81-90, 139-150).
int KindOfPlacement(const double *xf12) {
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

}  // namespace

int BuildGeometry(const vg_flat_geometry *f, vg_geometry *g) {
  if (f->n_shapes < 0 || f->n_logical < 0 || f->n_placed <= 0)
    return SetError(VG_ERR_INVALID, "geometry needs at least one placed volume");
  if (!f->shape_type || !f->shape_param_begin || !f->logical_shape || !f->logical_daughter_begin ||
      !f->placed_logical || !f->placed_transform)
    return SetError(VG_ERR_INVALID, "null geometry array");
  if (f->world < 0 || f->world >= f->n_placed) return SetError(VG_ERR_INVALID, "world id out of range");

  g->world = f->world;
  // Shapes: parameter blocks.
  g->shapes.resize(f->n_shapes);
  for (int s = 0; s < f->n_shapes; ++s) {
    const int b = f->shape_param_begin[s], e = f->shape_param_begin[s + 1];
    if (b < 0 || e < b || (e > b && !f->params)) {
      return SetError(VG_ERR_INVALID, "bad parameter range for shape " + std::to_string(s));
    }
    std::string err;
    std::vector<double> block = BuildParams(f->shape_type[s], f->params + b, e - b, &err);
    if (block.empty()) {
      return SetError(ShapeHasKernels(f->shape_type[s]) ? VG_ERR_INVALID : VG_ERR_UNSUPPORTED,
                      "shape " + std::to_string(s) + ": " + err);
    }
    if (block.size() > (size_t)kMaxParamDoubles) {
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
    return SetError(VG_ERR_INVALID, "null daughters array");
  }
  g->daughters.assign(f->daughters, f->daughters + nd);
  for (int l = 0; l < f->n_logical; ++l) {
    if (g->logical_shape[l] < 0 || g->logical_shape[l] >= f->n_shapes || g->logical_dbegin[l] < 0 ||
        g->logical_dbegin[l + 1] < g->logical_dbegin[l] || g->logical_dbegin[l + 1] > nd) {
      return SetError(VG_ERR_INVALID, "bad logical volume " + std::to_string(l));
    }
  }
  for (int k = 0; k < nd; ++k) {
    if (g->daughters[k] < 0 || g->daughters[k] >= f->n_placed) {
      return SetError(VG_ERR_INVALID, "daughter id out of range");
    }
  }
  // Placed volumes.
  g->placed.resize(f->n_placed);
  g->aabb.resize(6 * (size_t)f->n_placed);
  g->b32.resize(f->n_placed);
  g->extent32.resize(6 * (size_t)f->n_placed);
  g->b32side.clear();
  for (int p = 0; p < f->n_placed; ++p) {
    const int l = f->placed_logical[p];
    if (l < 0 || l >= f->n_logical) {
      return SetError(VG_ERR_INVALID, "placed volume " + std::to_string(p) + " has a bad logical id");
    }
    PlacedRec &r = g->placed[p];
    const double *x = f->placed_transform + 12 * (size_t)p;
    r.logical = l;
    r.shape = g->logical_shape[l];
    r.type = g->shapes[r.shape].type;
    r.xfkind = KindOfPlacement(x);
    r.param_off = g->shapes[r.shape].param_off;
    r.dbegin = g->logical_dbegin[l];
    r.dend = g->logical_dbegin[l + 1];
    r.pad = 0;
    for (int i = 0; i < 3; ++i) r.xf.t[i] = x[i];
    for (int i = 0; i < 9; ++i) r.xf.r[i] = x[3 + i];
    double lo[3], hi[3];
    if (!ShapeExtent(r.type, g->params.data() + r.param_off, lo, hi)) {
      return SetError(VG_ERR_UNSUPPORTED, "no extent for shape type " + std::to_string(r.type));
    }
    ComputeAABB(lo, hi, r.xf, g->aabb.data() + 6 * (size_t)p);
    BuildBound32(r.type, g->params.data() + r.param_off, r.xf, lo, hi, g->b32[p], g->b32side);
    for (int i = 0; i < 3; ++i) {
      g->extent32[6 * (size_t)p + i] = (float)lo[i];
      g->extent32[6 * (size_t)p + 3 + i] = (float)hi[i];
    }
    // Navigation and serving dispatch primitive shapes only.
    if (!g->nav_bad_type && (!ShapeHasKernels(r.type) || r.type == VG_BOOLEAN)) g->nav_bad_type = r.type;
    if (r.type >= 0 && r.type < 32) g->type_mask |= 1u << r.type;
  }
  // The daughters' float boxes in daughter-slot order: the navigator's box
  // loop then reads slot k's box without first reading slot k's id.
  g->slot_aabb.assign(8 * g->daughters.size(), 0.f);
  for (size_t k = 0; k < g->daughters.size(); ++k)
    for (int i = 0; i < 6; ++i) g->slot_aabb[8 * k + i] = (float)g->aabb[6 * (size_t)g->daughters[k] + i];
  g->max_depth = Depth(g, g->world, 1, 64);
  if (g->max_depth < 0) {
    return SetError(VG_ERR_INVALID, "volume tree deeper than 64 levels or cyclic");
  }
  return VG_OK;
}

int PlacementKind(const double *xf12) { return KindOfPlacement(xf12); }

}  // namespace vga
