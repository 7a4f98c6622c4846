// Host-side internals of libvecgeom_amd: the uploaded geometry and the
// dispatch tables shared by the C-ABI translation units.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/vecgeom_amd.h"
#include "vg_bound32.hpp"
#include "vg_math.hpp"

namespace vga {

// Device record of one placed volume (also kept on the host).
struct PlacedRec {
  int32_t logical;
  int32_t shape;      // index into shape tables
  int32_t type;       // vg_shape_type
  int32_t xfkind;     // XfKind of its placement
  int32_t param_off;  // offset (doubles) of the shape's parameter block
  int32_t dbegin;     // daughters of its logical volume: [dbegin, dend)
  int32_t dend;
  int32_t pad;
  Xf xf;
};
static_assert(sizeof(PlacedRec) % 8 == 0, "PlacedRec must be 8-byte sized");

struct ShapeRec {
  int32_t type;
  int32_t param_off;
  int32_t param_len;
  int32_t pad;
};

// Upper bound on the size of one shape parameter block (doubles).  Fixed
// blocks are a few hundred bytes; variable ones (Polycone sections,
// Polyhedron quadrilaterals) grow with the shape's planes and sides.
constexpr int kMaxParamDoubles = 1 << 20;

}  // namespace vga

struct vg_geometry {
  int device = -1;
  int max_depth = 0;
  int32_t world = -1;
  std::vector<vga::ShapeRec> shapes;
  std::vector<double> params;  // concatenated parameter blocks
  std::vector<vga::PlacedRec> placed;
  std::vector<int32_t> logical_shape;
  std::vector<int32_t> logical_dbegin;  // n_logical + 1
  std::vector<int32_t> daughters;
  std::vector<double> aabb;             // 6 per placed: lo xyz, hi xyz (mother frame)
  std::vector<vga::B32> b32;            // FP32 bound record per placed volume
  std::vector<float> b32side;           // variable-length bound data
  std::vector<float> extent32;          // 6 per placed: local extent lo xyz, hi xyz
  std::vector<float> slot_aabb;         // 8 per daughter slot: its float AABB, lo xyz, hi xyz, 2 unused
  // Device copies.
  double *d_params = nullptr;
  vga::PlacedRec *d_placed = nullptr;
  int32_t *d_daughters = nullptr;
  float *d_aabb = nullptr;  // 6 floats per placed (ABBoxManager keeps floats)
  vga::B32 *d_b32 = nullptr;
  float *d_b32side = nullptr;
  float *d_extent32 = nullptr;
  float *d_slot_aabb = nullptr;
  double *d_scalar = nullptr;  // device staging of vg_scalar_query (16 doubles)
  std::mutex scalar_mutex;     // one vg_scalar_query at a time per geometry
  int32_t nav_bad_type = 0;  // first shape type in the tree without device kernels (0: none)
  uint32_t type_mask = 0;    // bit t set: some placed volume has shape type t
};

namespace vga {

// Error handling (vg_capi.cpp).
int SetError(int code, const std::string &msg);
int CheckHip(hipError_t e, const char *what);
const char *LastError();
int EnsureDevice();
// Validates and flattens a geometry on the host (vg_geometry.cpp); fills
// every field but the device copies.  Returns a status.
int BuildGeometry(const vg_flat_geometry *f, vg_geometry *g);

// Builds the parameter block of one shape from its constructor parameters.
// Returns an empty vector and sets *err on invalid input.
std::vector<double> BuildParams(int type, const double *ctor, int n, std::string *err);
// Local-frame axis-aligned extent of a shape (UnplacedX::Extent).
bool ShapeExtent(int type, const double *block, double lo[3], double hi[3]);
// FP32 bound record of one placed volume (vg_bound32.cpp); variable-length
// data is appended to `side`.
void BuildBound32(int type, const double *blk, const Xf &xf, const double lo[3], const double hi[3], B32 &b,
                  std::vector<float> &side);

// Query kinds dispatched by vg_shape_kernels.hip.
enum Query : int {
  kQContains = 0,
  kQInside,
  kQDistIn,
  kQDistInMin,
  kQDistOut,
  kQPlacedDistOut,
  kQSafeIn,
  kQSafeInMin,
  kQSafeOut,
  kQSafeOutMin,
};

struct QArgs {
  const double *x, *y, *z;
  const double *dx, *dy, *dz;
  const double *sm;  // step max (nullable)
  void *out;         // uint8 / int32 / double, or inout distances
  int32_t *next;     // node index / next daughter (nullable)
  int32_t did;       // daughter index for DistanceToInMinimize
  int32_t pad;
  uint64_t n;
};

// Contains(point, localPoint&) / UnplacedContains baskets (vg_extra.hip).
int LaunchContainsExtra(const vg_geometry *g, int32_t placed, const double *x, const double *y, const double *z,
                        uint64_t n, uint8_t *out, double *lx, double *ly, double *lz, bool local_in, hipStream_t s);

// Launches one basket query on placed volume `placed` (vg_shape_kernels.hip).
int LaunchQuery(const vg_geometry *g, int32_t placed, int q, const QArgs &a, hipStream_t s);
// Basket queries on a Boolean volume (vg_boolean.hip).
int LaunchBooleanQuery(const vg_geometry *g, int32_t placed, int q, const QArgs &a, hipStream_t s);
// Normal baskets (vg_normal.hip): with dist == nullptr the Normal of local
// points; otherwise the USolids DistanceToOut with the normal at the exit.
int LaunchNormal(const vg_geometry *g, int32_t placed, const double *x, const double *y, const double *z,
                 const double *dx, const double *dy, const double *dz, const double *step_max, uint64_t n,
                 double *dist, double *nx, double *ny, double *nz, uint8_t *valid, uint8_t *convex, hipStream_t s);
// Shape types whose reference implements VPlacedVolume::Normal.
bool ShapeHasNormal(int type);
// Placement kind (identity / translation / rotation / general) of 12
// placement doubles, as Transformation3D::SetProperties decides it.
int PlacementKind(const double *xf12);
bool ShapeHasKernels(int type);
int SetKernelVariant(int v);

}  // namespace vga

namespace vga {
// Navigation launches (vg_nav.hip); arguments validated by the caller.
int NavLocate(const vg_geometry *g, const double *x, const double *y, const double *z, uint64_t n,
              const vg_nav_states &s, hipStream_t st);
int NavStep(const vg_geometry *g, int kind, const double *x, const double *y, const double *z, const double *dx,
            const double *dy, const double *dz, const double *pstep, uint64_t n, const vg_nav_states &cur,
            const vg_nav_states &nxt, double *step, hipStream_t st);
int NavSafety(const vg_geometry *g, const double *x, const double *y, const double *z, uint64_t n,
              const vg_nav_states &s, double *out, hipStream_t st);
int NavTrace(const vg_geometry *g, int kind, const double *x, const double *y, const double *z, const double *dx,
             const double *dy, const double *dz, uint64_t n, double pstep, int max_steps, int32_t *nsteps, double *fx,
             double *fy, double *fz, double *travelled, uint64_t *hash, int32_t *exact, hipStream_t st);
}  // namespace vga

namespace vga {
// Zoo shapes beyond Box and Tube (vg_params_zoo.cpp).
std::vector<double> BuildZooParams(int type, const double *ctor, int n, std::string *err);
bool ZooExtent(int type, const double *block, double lo[3], double hi[3]);
}  // namespace vga
