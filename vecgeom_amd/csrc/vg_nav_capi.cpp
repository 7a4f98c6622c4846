// C-ABI of the device navigator: navigation state pools, LocatePoint,
// FindNextBoundaryAndStep, GetSafety and the track loop.  Arguments are
// validated here; vg_nav.hip launches the kernels.
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "vg_internal.hpp"

using namespace vga;

namespace {

int CheckNav(const vg_geometry *g, int kind, bool basket_ok = false) {
  if (!g) return SetError(VG_ERR_INVALID, "null geometry");
  if (kind != VG_NAV_ABBOX && kind != VG_NAV_SIMPLE && kind != VG_NAV_ABBOX_MIXED &&
      !(basket_ok && kind == VG_NAV_SIMPLE_BASKET))
    return SetError(VG_ERR_INVALID, kind == VG_NAV_SIMPLE_BASKET ? "the basket navigator steps baskets only "
                                                                   "(vg_find_next_boundary_and_step)"
                                                                 : "unknown navigator");
  if (g->max_depth > VG_NAV_MAX_DEPTH)
    return SetError(VG_ERR_UNSUPPORTED, "geometry depth " + std::to_string(g->max_depth) + " exceeds " +
                                            std::to_string(VG_NAV_MAX_DEPTH));
  if (g->nav_bad_type)
    return SetError(VG_ERR_UNSUPPORTED, "no device kernels for shape type " + std::to_string(g->nav_bad_type) +
                                            " used in the volume tree");
  return EnsureDevice();
}

int CheckStates(const vg_geometry *g, const vg_nav_states *s, uint64_t n, const char *what) {
  if (!s || !s->level || !s->boundary || (s->depth > 0 && !s->path))
    return SetError(VG_ERR_INVALID, std::string(what) + ": null state arrays");
  if (s->depth < g->max_depth)
    return SetError(VG_ERR_INVALID, std::string(what) + ": state depth " + std::to_string(s->depth) +
                                        " below geometry depth " + std::to_string(g->max_depth));
  if (n > s->capacity) return SetError(VG_ERR_CAPACITY, std::string(what) + ": more tracks than state capacity");
  return VG_OK;
}

bool AnyNull(std::initializer_list<const void *> ps) {
  for (auto p : ps)
    if (!p) return true;
  return false;
}

}  // namespace

extern "C" {

int vg_nav_states_create(const vg_geometry *g, uint64_t capacity, int32_t depth, vg_nav_states *out) {
  if (!g || !out) return SetError(VG_ERR_INVALID, "null argument");
  int rc = EnsureDevice();
  if (rc) return rc;
  if (depth <= 0) depth = g->max_depth;
  if (depth < g->max_depth) return SetError(VG_ERR_INVALID, "state depth below geometry depth");
  vg_nav_states s{};
  s.capacity = capacity;
  s.depth = depth;
  const size_t c = capacity ? capacity : 1;
  rc = CheckHip(hipMalloc(&s.level, c * sizeof(int32_t)), "hipMalloc level");
  if (!rc) rc = CheckHip(hipMalloc(&s.path, c * (size_t)depth * sizeof(int32_t)), "hipMalloc path");
  if (!rc) rc = CheckHip(hipMalloc(&s.boundary, c), "hipMalloc boundary");
  if (!rc) rc = CheckHip(hipMemset(s.level, 0, c * sizeof(int32_t)), "hipMemset level");
  if (!rc) rc = CheckHip(hipMemset(s.path, 0xff, c * (size_t)depth * sizeof(int32_t)), "hipMemset path");
  if (!rc) rc = CheckHip(hipMemset(s.boundary, 0, c), "hipMemset boundary");
  if (rc) {
    (void)hipFree(s.level);
    (void)hipFree(s.path);
    (void)hipFree(s.boundary);
    return rc;
  }
  *out = s;
  return VG_OK;
}

int vg_nav_states_free(vg_nav_states *s) {
  if (!s) return SetError(VG_ERR_INVALID, "null states");
  int rc = EnsureDevice();
  if (rc) return rc;
  rc = CheckHip(hipFree(s->level), "hipFree level");
  int rc2 = CheckHip(hipFree(s->path), "hipFree path");
  int rc3 = CheckHip(hipFree(s->boundary), "hipFree boundary");
  s->level = s->path = nullptr;
  s->boundary = nullptr;
  s->capacity = 0;
  return rc ? rc : (rc2 ? rc2 : rc3);
}

int vg_nav_states_to_host(const vg_nav_states *s, uint64_t n, int32_t *level, int32_t *path, uint8_t *boundary) {
  if (!s || AnyNull({level, boundary}) || (s->depth > 0 && !path)) return SetError(VG_ERR_INVALID, "null argument");
  if (n > s->capacity) return SetError(VG_ERR_CAPACITY, "n exceeds state capacity");
  int rc = EnsureDevice();
  if (rc || n == 0) return rc;
  rc = CheckHip(hipMemcpy(level, s->level, n * sizeof(int32_t), hipMemcpyDeviceToHost), "copy level");
  if (!rc) rc = CheckHip(hipMemcpy(boundary, s->boundary, n, hipMemcpyDeviceToHost), "copy boundary");
  for (int l = 0; !rc && l < s->depth; ++l)
    rc = CheckHip(hipMemcpy(path + (size_t)l * n, s->path + (size_t)l * s->capacity, n * sizeof(int32_t),
                            hipMemcpyDeviceToHost),
                  "copy path");
  return rc;
}

int vg_nav_states_from_host(vg_nav_states *s, uint64_t n, const int32_t *level, const int32_t *path,
                            const uint8_t *boundary) {
  if (!s || AnyNull({level, boundary}) || (s->depth > 0 && !path)) return SetError(VG_ERR_INVALID, "null argument");
  if (n > s->capacity) return SetError(VG_ERR_CAPACITY, "n exceeds state capacity");
  int rc = EnsureDevice();
  if (rc || n == 0) return rc;
  rc = CheckHip(hipMemcpy(s->level, level, n * sizeof(int32_t), hipMemcpyHostToDevice), "copy level");
  if (!rc) rc = CheckHip(hipMemcpy(s->boundary, boundary, n, hipMemcpyHostToDevice), "copy boundary");
  for (int l = 0; !rc && l < s->depth; ++l)
    rc = CheckHip(hipMemcpy(s->path + (size_t)l * s->capacity, path + (size_t)l * n, n * sizeof(int32_t),
                            hipMemcpyHostToDevice),
                  "copy path");
  return rc;
}

int vg_locate_points(const vg_geometry *g, const double *x, const double *y, const double *z, size_t n,
                     vg_nav_states *states, void *stream) {
  int rc = CheckNav(g, VG_NAV_ABBOX);
  if (!rc) rc = CheckStates(g, states, n, "vg_locate_points");
  if (rc) return rc;
  if (n && AnyNull({x, y, z})) return SetError(VG_ERR_INVALID, "null point array");
  return NavLocate(g, x, y, z, n, *states, (hipStream_t)stream);
}

int vg_find_next_boundary_and_step(const vg_geometry *g, int32_t navigator, const double *x, const double *y,
                                   const double *z, const double *dx, const double *dy, const double *dz,
                                   const double *pstep, size_t n, const vg_nav_states *cur, vg_nav_states *next,
                                   double *step, void *stream) {
  int rc = CheckNav(g, navigator, true);
  if (!rc) rc = CheckStates(g, cur, n, "current states");
  if (!rc) rc = CheckStates(g, next, n, "new states");
  if (rc) return rc;
  if (n && AnyNull({x, y, z, dx, dy, dz, step})) return SetError(VG_ERR_INVALID, "null array argument");
  if (cur->level == next->level || cur->path == next->path || cur->boundary == next->boundary)
    return SetError(VG_ERR_INVALID, "current and new states alias");
  return NavStep(g, navigator, x, y, z, dx, dy, dz, pstep, n, *cur, *next, step, (hipStream_t)stream);
}

int vg_get_safety(const vg_geometry *g, const double *x, const double *y, const double *z, size_t n,
                  const vg_nav_states *cur, double *safety, void *stream) {
  int rc = CheckNav(g, VG_NAV_ABBOX);
  if (!rc) rc = CheckStates(g, cur, n, "vg_get_safety");
  if (rc) return rc;
  if (n && AnyNull({x, y, z, safety})) return SetError(VG_ERR_INVALID, "null array argument");
  return NavSafety(g, x, y, z, n, *cur, safety, (hipStream_t)stream);
}

int vg_trace_tracks(const vg_geometry *g, int32_t navigator, const double *x, const double *y, const double *z,
                    const double *dx, const double *dy, const double *dz, size_t n, double pstep, int32_t max_steps,
                    int32_t *nsteps, double *fx, double *fy, double *fz, double *travelled, uint64_t *hash,
                    int32_t *exact_steps, void *stream) {
  int rc = CheckNav(g, navigator);
  if (rc) return rc;
  if (max_steps < 0) return SetError(VG_ERR_INVALID, "max_steps must be >= 0");
  if (n && AnyNull({x, y, z, dx, dy, dz, nsteps, fx, fy, fz, travelled, hash}))
    return SetError(VG_ERR_INVALID, "null array argument");
  return NavTrace(g, navigator, x, y, z, dx, dy, dz, n, pstep, max_steps, nsteps, fx, fy, fz, travelled, hash,
                  exact_steps, (hipStream_t)stream);
}

}  // extern "C"
