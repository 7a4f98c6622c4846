// Test infrastructure: a C++ caller of the C-ABI (include/vecgeom_amd.h),
// linked against vecgeom_amd/lib/libvecgeom_amd.so with no Python in
// between, as a reference-side binding would call it.  It uploads a small
// geometry (a world box holding a box, a tube and an orb), runs basket and
// scalar shape queries, the no-allocation batch and the navigator on the
// GPU, and checks closed-form answers.  Prints one line per check and
// "capi_gpu: all N checks passed"; exits non-zero on the first failure.
#include <hip/hip_runtime_api.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "../../include/vecgeom_amd.h"

namespace {

int g_checks = 0;

void Fail(const char *what) {
  std::printf("capi_gpu: FAILED %s (%s)\n", what, vg_last_error());
  std::exit(1);
}

void Ok(int rc, const char *what) {
  if (rc != VG_OK) Fail(what);
}

void Expect(bool cond, const char *what) {
  if (!cond) Fail(what);
  ++g_checks;
  std::printf("capi_gpu: ok %s\n", what);
}

bool Near(double a, double b) { return std::fabs(a - b) <= 1e-12 * (1 + std::fabs(b)); }

template <class T>
T *Dev(const std::vector<T> &h) {
  T *d = nullptr;
  if (hipMalloc(&d, h.size() * sizeof(T)) != hipSuccess) Fail("hipMalloc");
  if (hipMemcpy(d, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice) != hipSuccess) Fail("hipMemcpy");
  return d;
}

template <class T>
std::vector<T> Host(const T *d, size_t n) {
  std::vector<T> h(n);
  if (hipMemcpy(h.data(), d, n * sizeof(T), hipMemcpyDeviceToHost) != hipSuccess) Fail("hipMemcpy back");
  return h;
}

}  // namespace

int main() {
  Ok(vg_init(0), "vg_init");
  // World box 100^3; box (10,15,20) at (30,0,0); tube (5,20,30, full phi)
  // at (-40,0,0); orb 8 at (0,50,0).  Shapes 0..3, logical 0..3, placed 0..3
  // (0 = world).
  const int32_t types[4] = {VG_BOX, VG_BOX, VG_TUBE, VG_ORB};
  const double params[] = {100, 100, 100, 10, 15, 20, 5, 20, 30, 0, 2 * M_PI, 8};
  const int32_t pbegin[5] = {0, 3, 6, 11, 12};
  const int32_t lshape[4] = {0, 1, 2, 3};
  const int32_t dbegin[5] = {0, 3, 3, 3, 3};
  const int32_t daughters[3] = {1, 2, 3};
  const int32_t plogical[4] = {0, 1, 2, 3};
  const double xf[48] = {0,   0, 0, 1, 0, 0, 0, 1, 0, 0, 0, 1,   // world
                         30,  0, 0, 1, 0, 0, 0, 1, 0, 0, 0, 1,   // box
                         -40, 0, 0, 1, 0, 0, 0, 1, 0, 0, 0, 1,   // tube
                         0,  50, 0, 1, 0, 0, 0, 1, 0, 0, 0, 1};  // orb
  vg_flat_geometry flat{4, types, pbegin, params, 4, lshape, dbegin, daughters, 4, plogical, xf, 0};
  vg_geometry *g = nullptr;
  Ok(vg_geometry_upload(&flat, &g), "vg_geometry_upload");
  int32_t depth = 0;
  Ok(vg_geometry_max_depth(g, &depth), "vg_geometry_max_depth");
  Expect(depth == 2, "max depth 2");

  // Basket queries on the box (placed 1): points in the world frame.
  const std::vector<double> px = {30, 0, 45}, py = {0, 0, 0}, pz = {0, 0, 0};
  const std::vector<double> dx = {1, 1, -1}, dy = {0, 0, 0}, dz = {0, 0, 0};
  double *x = Dev(px), *y = Dev(py), *z = Dev(pz), *ddx = Dev(dx), *ddy = Dev(dy), *ddz = Dev(dz);
  double *out = Dev(std::vector<double>(3, 0.));
  uint8_t *c = Dev(std::vector<uint8_t>(3, 0));
  int32_t *ins = Dev(std::vector<int32_t>(3, 0));
  Ok(vg_contains(g, 1, x, y, z, 3, c, nullptr), "vg_contains");
  auto hc = Host(c, 3);
  Expect(hc[0] == 1 && hc[1] == 0 && hc[2] == 0, "box Contains (inside, outside, outside)");
  Ok(vg_inside(g, 1, x, y, z, 3, ins, nullptr), "vg_inside");
  auto hi = Host(ins, 3);
  Expect(hi[0] == VG_INSIDE_INSIDE && hi[1] == VG_INSIDE_OUTSIDE, "box Inside");
  Ok(vg_distance_to_in(g, 1, x, y, z, ddx, ddy, ddz, nullptr, 3, out, nullptr), "vg_distance_to_in");
  auto hd = Host(out, 3);
  Expect(Near(hd[1], 20.) && Near(hd[2], 5.), "box DistanceToIn from (0,0,0) along +x = 20, from (45,0,0) along -x = 5");
  Ok(vg_safety_to_in(g, 1, x, y, z, 3, out, nullptr), "vg_safety_to_in");
  hd = Host(out, 3);
  Expect(Near(hd[1], 20.) && Near(hd[2], 5.), "box SafetyToIn 20 and 5");

  // Contains(point, localPoint&) and UnplacedContains.
  double *lx = Dev(std::vector<double>(3, 0.)), *ly = Dev(std::vector<double>(3, 0.)), *lz = Dev(std::vector<double>(3, 0.));
  Ok(vg_contains_local(g, 1, x, y, z, 3, c, lx, ly, lz, nullptr), "vg_contains_local");
  hc = Host(c, 3);
  auto hlx = Host(lx, 3);
  Expect(hc[0] == 1 && Near(hlx[0], 0.) && Near(hlx[1], -30.) && Near(hlx[2], 15.), "Contains(point, local&) basket");
  Ok(vg_unplaced_contains(g, 1, lx, ly, lz, 3, c, nullptr), "vg_unplaced_contains");
  hc = Host(c, 3);
  Expect(hc[0] == 1 && hc[1] == 0 && hc[2] == 0, "UnplacedContains of the local points");

  // Normal and the USolids DistanceToOut with the exit normal, on local points
  // of the box (half-lengths 10, 15, 20): the +x face, an edge, the centre.
  {
    const std::vector<double> nx0 = {10, -10, 0}, ny0 = {0, 15, 0}, nz0 = {3, 0, 0};
    double *qx = Dev(nx0), *qy = Dev(ny0), *qz = Dev(nz0);
    double *ox = Dev(std::vector<double>(3, 0.)), *oy = Dev(std::vector<double>(3, 0.)),
           *oz = Dev(std::vector<double>(3, 0.));
    uint8_t *ok = Dev(std::vector<uint8_t>(3, 9));
    Ok(vg_normal(g, 1, qx, qy, qz, 3, ox, oy, oz, ok, nullptr), "vg_normal");
    auto hv = Host(ok, 3);
    auto hx = Host(ox, 3), hy = Host(oy, 3);
    Expect(hv[0] == 1 && hv[1] == 1 && hv[2] == 0 && hx[0] == 1. && Near(hx[1], -M_SQRT1_2) && Near(hy[1], M_SQRT1_2),
           "box Normal: +x face, (-x, +y) edge, centre invalid");
    Ok(vg_distance_to_out_normal(g, 1, lx, ly, lz, ddx, ddy, ddz, nullptr, 1, out, ox, oy, oz, ok, c, nullptr),
       "vg_distance_to_out_normal");
    auto hdo = Host(out, 1);
    hx = Host(ox, 1);
    hv = Host(ok, 1);
    auto hcv = Host(c, 1);
    Expect(Near(hdo[0], 10.) && hx[0] == 1. && hv[0] == 1 && hcv[0] == 1,
           "box DistanceToOut from its centre along +x = 10, exit normal +x, convex");
    Expect(vg_normal(g, 3, qx, qy, qz, 3, ox, oy, oz, ok, nullptr) == VG_OK, "orb Normal");
  }

  // Scalar virtuals.
  const double p1[3] = {35, 1, 2}, up[3] = {0, 1, 0};
  double r = -1, loc[3] = {0, 0, 0};
  Ok(vg_scalar_query(g, 1, VG_Q_CONTAINS_LOCAL, p1, nullptr, 0., &r, loc), "scalar Contains(point, local&)");
  Expect(r == 1. && Near(loc[0], 5.) && Near(loc[1], 1.) && Near(loc[2], 2.), "scalar Contains(point, local&)");
  const double l2[3] = {11, 0, 0};
  Ok(vg_scalar_query(g, 1, VG_Q_UNPLACED_CONTAINS, l2, nullptr, 0., &r, nullptr), "scalar UnplacedContains");
  Expect(r == 0., "scalar UnplacedContains (11,0,0) outside half-length 10");
  const double tc[3] = {0, 0, 0};
  Ok(vg_scalar_query(g, 2, VG_Q_DISTANCE_TO_OUT, tc, up, 1e30, &r, nullptr), "scalar DistanceToOut");
  Expect(Near(r, 20.), "tube DistanceToOut from its axis along +y = rmax 20");
  Ok(vg_scalar_query(g, 3, VG_Q_SAFETY_TO_OUT, tc, nullptr, 0., &r, nullptr), "scalar SafetyToOut");
  Expect(Near(r, 8.), "orb SafetyToOut at its centre = 8");
  const double pw[3] = {0, 0, 0};
  Ok(vg_scalar_query(g, 3, VG_Q_INSIDE, pw, nullptr, 0., &r, nullptr), "scalar Inside");
  Expect(r == (double)VG_INSIDE_OUTSIDE, "orb Inside of the world origin: outside");

  // The no-allocation batch on the box queries.
  uint8_t *c2 = Dev(std::vector<uint8_t>(3, 9));
  vg_query_task t{};
  t.placed = 1;
  t.query = VG_Q_CONTAINS;
  t.input_type = VG_F64;
  t.n = 3;
  t.x = x;
  t.y = y;
  t.z = z;
  t.out = c2;
  Ok(vg_batch_run_noalloc(g, &t, 1, nullptr), "vg_batch_run_noalloc");
  hc = Host(c2, 3);
  Expect(hc[0] == 1 && hc[1] == 0 && hc[2] == 0, "no-allocation batch Contains");

  // Navigation: locate, one ABBox step from the origin along +x.
  vg_nav_states s0{}, s1{};
  Ok(vg_nav_states_create(g, 3, 0, &s0), "vg_nav_states_create");
  Ok(vg_nav_states_create(g, 3, 0, &s1), "vg_nav_states_create");
  Ok(vg_locate_points(g, x, y, z, 3, &s0, nullptr), "vg_locate_points");
  std::vector<int32_t> lvl(3), path(3 * (size_t)s0.depth);
  std::vector<uint8_t> bnd(3);
  Ok(vg_nav_states_to_host(&s0, 3, lvl.data(), path.data(), bnd.data()), "vg_nav_states_to_host");
  Expect(lvl[0] == 2 && path[3 + 0] == 1 && lvl[1] == 1, "LocatePoint: (30,0,0) in the box, (0,0,0) in the world");
  double *step = Dev(std::vector<double>(3, 0.));
  Ok(vg_find_next_boundary_and_step(g, VG_NAV_ABBOX, x, y, z, ddx, ddy, ddz, nullptr, 3, &s0, &s1, step, nullptr),
     "vg_find_next_boundary_and_step");
  auto hs = Host(step, 3);
  Ok(vg_nav_states_to_host(&s1, 3, lvl.data(), path.data(), bnd.data()), "vg_nav_states_to_host");
  Expect(Near(hs[1], 20.) && lvl[1] == 2 && path[3 + 1] == 1 && bnd[1] == 1,
         "FindNextBoundaryAndStep: origin along +x enters the box after 20");
  Ok(vg_nav_states_free(&s0), "vg_nav_states_free");
  Ok(vg_nav_states_free(&s1), "vg_nav_states_free");
  Ok(vg_geometry_free(g), "vg_geometry_free");
  for (void *p : {(void *)x, (void *)y, (void *)z, (void *)ddx, (void *)ddy, (void *)ddz, (void *)out, (void *)c,
                  (void *)ins, (void *)lx, (void *)ly, (void *)lz, (void *)c2, (void *)step})
    (void)hipFree(p);
  std::printf("capi_gpu: all %d checks passed\n", g_checks);
  return 0;
}
