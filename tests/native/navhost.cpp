// Test infrastructure: a host-only build of the device navigation code
// (vecgeom_amd/csrc/vg_nav_impl.hpp) with every table access bounds-checked
// (VGA_NAV_HOST_CHECKS).  It runs the same LocatePoint /
// FindNextBoundaryAndStep / GetSafety / TraceTrack functions the kernels run,
// on the geometry records vg_geometry_upload builds (BuildGeometry), so the
// CPU tests can compare the navigator's logic with the oracle before any
// GPU run.  Exports the oracle's navigation entry points with the prefix
// navhost_ (oracle/vgo_nav.cpp has the argument meanings); navigator kind 9
// runs the ABBox navigator with only its exact sorted-order candidate walk.
#include <thread>
#include <vector>

#include "../../vecgeom_amd/csrc/vg_nav_impl.hpp"

using namespace vga;
using namespace vga::nav;

namespace {

// Host-only navigator code: the ABBox walk in exact sorted order alone.
constexpr int kExactOnlyKind = 9;

struct HostGeom {
  vg_geometry g;
  std::vector<float> aabbf;
  NavDev dev;
};

template <class F>
void ParallelFor(long n, int nthreads, F f) {
  if (nthreads <= 1 || n < 64) {
    f(0L, n);
    return;
  }
  std::vector<std::thread> th;
  const long chunk = (n + nthreads - 1) / nthreads;
  for (int t = 0; t < nthreads; ++t) {
    const long b = t * chunk, e = std::min(n, b + chunk);
    if (b >= e) break;
    th.emplace_back([=] { f(b, e); });
  }
  for (auto &x : th) x.join();
}

template <int D>
void Load(const int *level, const int *path, long n, long i, DPath<D> &p) {
  p.Clear();
  p.level = level[i];
  VGA_NAV_CHECK(p.level >= 0 && p.level <= D, "state level out of range");
  for (int l = 0; l < p.level; ++l) p.v[l] = path[l * n + i];
}

template <int D>
void Store(const DPath<D> &p, int depth, long n, long i, int *level, int *path) {
  level[i] = p.level;
  for (int l = 0; l < depth; ++l) path[l * n + i] = l < D ? p.Get(l) : -1;
}

template <int D>
void Locate(const HostGeom *h, long n, const double *x, const double *y, const double *z, int depth, int *level,
            int *path, int nthreads) {
  ParallelFor(n, nthreads, [&](long b, long e) {
    for (long i = b; i < e; ++i) {
      DPath<D> p;
      p.Clear();
      LocatePoint(h->dev, h->dev.world, V3{x[i], y[i], z[i]}, p, true);
      Store(p, depth, n, i, level, path);
    }
  });
}

template <int D>
void Step(const HostGeom *h, long n, const double *x, const double *y, const double *z, const double *dx,
          const double *dy, const double *dz, int depth, const int *level_in, const int *path_in, int *level_out,
          int *path_out, unsigned char *boundary_out, double *step_out, double pstep, int kind, int nthreads) {
  ParallelFor(n, nthreads, [&](long b, long e) {
    for (long i = b; i < e; ++i) {
      DPath<D> c, o;
      Load(level_in, path_in, n, i, c);
      double step = 0;
      if (kind == VG_NAV_SIMPLE_BASKET) {
        DPath<D> c0;
        Load(level_in, path_in, n, 0, c0);
        if (c.level > 0 && c0.level > 0)
          FindNextBoundaryAndStepBasket(h->dev, V3{x[i], y[i], z[i]}, V3{dx[i], dy[i], dz[i]}, c, o, pstep, step,
                                        c0.Top());
        else
          o = c;
      } else if (c.level > 0 && kind == kExactOnlyKind) {
        FindNextBoundaryAndStep<D, true>(h->dev, V3{x[i], y[i], z[i]}, V3{dx[i], dy[i], dz[i]}, c, o, pstep, step,
                                         VG_NAV_ABBOX);
      } else if (c.level > 0 && kind == VG_NAV_ABBOX_MIXED) {
        FindNextBoundaryAndStep<D, false, kAllShapes, true>(h->dev, V3{x[i], y[i], z[i]}, V3{dx[i], dy[i], dz[i]}, c,
                                                            o, pstep, step, kind);
      } else if (c.level > 0) {
        FindNextBoundaryAndStep(h->dev, V3{x[i], y[i], z[i]}, V3{dx[i], dy[i], dz[i]}, c, o, pstep, step, kind);
      } else {
        o = c;
      }
      Store(o, depth, n, i, level_out, path_out);
      boundary_out[i] = (unsigned char)o.boundary;
      step_out[i] = step;
    }
  });
}

template <int D>
void Safety(const HostGeom *h, long n, const double *x, const double *y, const double *z, const int *level_in,
            const int *path_in, double *out, int nthreads) {
  ParallelFor(n, nthreads, [&](long b, long e) {
    for (long i = b; i < e; ++i) {
      DPath<D> c;
      Load(level_in, path_in, n, i, c);
      out[i] = c.level > 0 ? GetSafety(h->dev, V3{x[i], y[i], z[i]}, c) : 0.;
    }
  });
}

template <int D>
void Trace(const HostGeom *h, long n, const double *x, const double *y, const double *z, const double *dx,
           const double *dy, const double *dz, double pstep, int kind, int max_steps, int *nsteps, double *fx,
           double *fy, double *fz, double *travelled, unsigned long long *hash, int nthreads) {
  ParallelFor(n, nthreads, [&](long b, long e) {
    for (long i = b; i < e; ++i) {
      V3 p{x[i], y[i], z[i]};
      double dist;
      uint64_t hh;
      int nexact;
      const V3 dd{dx[i], dy[i], dz[i]};
      nsteps[i] = kind == kExactOnlyKind
                      ? TraceOne<D, true>(h->dev, p, dd, pstep, VG_NAV_ABBOX, max_steps, dist, hh, nexact)
                  : kind == VG_NAV_ABBOX_MIXED
                      ? TraceOne<D, false, kAllShapes, true>(h->dev, p, dd, pstep, kind, max_steps, dist, hh, nexact)
                      : TraceOne<D>(h->dev, p, dd, pstep, kind, max_steps, dist, hh, nexact);
      fx[i] = p.x;
      fy[i] = p.y;
      fz[i] = p.z;
      travelled[i] = dist;
      hash[i] = hh;
    }
  });
}

}  // namespace

extern "C" {

void *navhost_geom_create(int n_shapes, const int *shape_type, const int *shape_param_begin, const double *params,
                          int n_logical, const int *logical_shape, const int *logical_dbegin, const int *daughters,
                          int n_placed, const int *placed_logical, const double *placed_xf, int world) {
  vg_flat_geometry f{n_shapes,      shape_type, shape_param_begin, params,    n_logical, logical_shape,
                     logical_dbegin, daughters,  n_placed,          placed_logical, placed_xf, world};
  auto *h = new HostGeom;
  if (BuildGeometry(&f, &h->g) != VG_OK || h->g.max_depth > 16) {
    std::fprintf(stderr, "navhost_geom_create: %s\n", LastError());
    delete h;
    return nullptr;
  }
  h->aabbf.assign(h->g.aabb.begin(), h->g.aabb.end());
  h->dev.placed = h->g.placed.data();
  h->dev.params = h->g.params.data();
  h->dev.daughters = h->g.daughters.data();
  h->dev.aabb = h->aabbf.data();
  h->dev.b32 = h->g.b32.data();
  h->dev.b32side = h->g.b32side.data();
  h->dev.slot_box = h->g.slot_aabb.data();
  h->dev.world = h->g.world;
  h->dev.n_placed = (int32_t)h->g.placed.size();
  h->dev.n_daughters = (int32_t)h->g.daughters.size();
  h->dev.pad = 0;
  return h;
}

void navhost_geom_destroy(void *h) { delete static_cast<HostGeom *>(h); }

int navhost_geom_max_depth(void *h) { return static_cast<HostGeom *>(h)->g.max_depth; }

void navhost_locate(void *hv, long n, const double *x, const double *y, const double *z, int depth, int *level,
                    int *path, int nthreads) {
  auto *h = static_cast<HostGeom *>(hv);
  if (h->g.max_depth <= 8)
    Locate<8>(h, n, x, y, z, depth, level, path, nthreads);
  else
    Locate<16>(h, n, x, y, z, depth, level, path, nthreads);
}

void navhost_nav_step(void *hv, long n, const double *x, const double *y, const double *z, const double *dx,
                      const double *dy, const double *dz, int depth, const int *level_in, const int *path_in,
                      int *level_out, int *path_out, unsigned char *boundary_out, double *step_out, double pstep,
                      int kind, int nthreads) {
  auto *h = static_cast<HostGeom *>(hv);
  if (h->g.max_depth <= 8)
    Step<8>(h, n, x, y, z, dx, dy, dz, depth, level_in, path_in, level_out, path_out, boundary_out, step_out, pstep,
            kind, nthreads);
  else
    Step<16>(h, n, x, y, z, dx, dy, dz, depth, level_in, path_in, level_out, path_out, boundary_out, step_out, pstep,
             kind, nthreads);
}

void navhost_safety(void *hv, long n, const double *x, const double *y, const double *z, int, const int *level_in,
                    const int *path_in, double *out, int nthreads) {
  auto *h = static_cast<HostGeom *>(hv);
  if (h->g.max_depth <= 8)
    Safety<8>(h, n, x, y, z, level_in, path_in, out, nthreads);
  else
    Safety<16>(h, n, x, y, z, level_in, path_in, out, nthreads);
}

// The six queries of placed volume `pv` through the device shape code
// (frames as the C-ABI: mother frame for Contains, Inside, DistanceToIn,
// SafetyToIn; local frame for DistanceToOut, SafetyToOut).  query 0..5 as the
// oracle's; out is uint8 / int32 / double.
int navhost_query(void *hv, int pv, int query, long n, const double *x, const double *y, const double *z,
                  const double *dx, const double *dy, const double *dz, const double *sm, void *out) {
  auto *h = static_cast<HostGeom *>(hv);
  const NavDev &g = h->dev;
  const PlacedRec &r = g.Pv(pv);
  Xf xf;
  LoadXf(r, xf);
  const double *blk = g.params + r.param_off;
  for (long i = 0; i < n; ++i) {
    const V3 p{x[i], y[i], z[i]};
    const double s = sm ? sm[i] : kInfinity;
    switch (query) {
      case 0: static_cast<uint8_t *>(out)[i] = DContains(r.type, blk, XfPointRt(xf, r.xfkind, p)) ? 1 : 0; break;
      case 1: static_cast<int32_t *>(out)[i] = DInside(r.type, blk, XfPointRt(xf, r.xfkind, p)); break;
      case 2: {
        const V3 d{dx[i], dy[i], dz[i]};
        static_cast<double *>(out)[i] =
            DDistIn(r.type, blk, XfPointRt(xf, r.xfkind, p), XfDirRt(xf, r.xfkind, d), s);
        break;
      }
      case 3: static_cast<double *>(out)[i] = DDistOut(r.type, blk, p, V3{dx[i], dy[i], dz[i]}, s); break;
      case 4: static_cast<double *>(out)[i] = DSafeIn(r.type, blk, XfPointRt(xf, r.xfkind, p)); break;
      case 5: static_cast<double *>(out)[i] = DSafeOut(r.type, blk, p); break;
      default: return -1;
    }
  }
  return 0;
}

void navhost_trace(void *hv, long n, const double *x, const double *y, const double *z, const double *dx,
                   const double *dy, const double *dz, double pstep, int kind, int max_steps, int *nsteps, double *fx,
                   double *fy, double *fz, double *travelled, unsigned long long *hash, int nthreads) {
  auto *h = static_cast<HostGeom *>(hv);
  if (h->g.max_depth <= 8)
    Trace<8>(h, n, x, y, z, dx, dy, dz, pstep, kind, max_steps, nsteps, fx, fy, fz, travelled, hash, nthreads);
  else
    Trace<16>(h, n, x, y, z, dx, dy, dz, pstep, kind, max_steps, nsteps, fx, fy, fz, travelled, hash, nthreads);
}

// The ABBoxManager box (mother frame, double, before rounding to float) and
// the local extent BuildGeometry computed for placed volume `placed`.
void navhost_aabb(void *hv, int placed, double *box6, double *lo, double *hi) {
  const auto *h = static_cast<HostGeom *>(hv);
  VGA_NAV_CHECK(placed >= 0 && placed < (int)h->g.placed.size(), "placed id out of range");
  for (int i = 0; i < 6; ++i) box6[i] = h->g.aabb[6 * (size_t)placed + i];
  const PlacedRec &r = h->g.placed[placed];
  ShapeExtent(r.type, h->g.params.data() + r.param_off, lo, hi);
}

// FP32 filter of the mixed-precision forms (vg_bound32.hpp) on placed
// volume `placed`: Classify32 per point (0 inside, 2 outside, -1
// undecided), and whether each ray provably misses the extent box.
void navhost_classify32(void *hv, int placed, long n, const float *x, const float *y, const float *z,
                        signed char *out) {
  const auto *h = static_cast<HostGeom *>(hv);
  VGA_NAV_CHECK(placed >= 0 && placed < (int)h->g.b32.size(), "placed id out of range");
  const B32 &b = h->g.b32[placed];
  for (long i = 0; i < n; ++i) out[i] = (signed char)Classify32(b, h->g.b32side.data(), F3{x[i], y[i], z[i]});
}

void navhost_raymiss32(void *hv, int placed, long n, const float *x, const float *y, const float *z, const float *dx,
                       const float *dy, const float *dz, unsigned char *out) {
  const auto *h = static_cast<HostGeom *>(hv);
  VGA_NAV_CHECK(placed >= 0 && placed < (int)h->g.b32.size(), "placed id out of range");
  const B32 &b = h->g.b32[placed];
  const float *e = h->g.extent32.data() + 6 * (size_t)placed;
  // 1: miss with DistanceToIn kInfinity; 2: miss with NaN (the Orb's
  // approaching rays); 0: undecided.  As the batch kernel's filter.
  for (long i = 0; i < n; ++i) {
    const F3 p{x[i], y[i], z[i]}, d{dx[i], dy[i], dz[i]};
    int r = RayMissesBox32(b, e, p, d) ? 1 : 0;
    if (r && b.kind == VG_ORB) {
      const int sg = RadialSign32(b, p, d);
      r = sg == 0 ? 0 : (sg < 0 ? 2 : 1);
    }
    out[i] = (unsigned char)r;
  }
}

// so, si and the margin of each point (for checks of the bounds themselves).
void navhost_bound32(void *hv, int placed, long n, const float *x, const float *y, const float *z, float *so,
                     float *si, float *m) {
  const auto *h = static_cast<HostGeom *>(hv);
  VGA_NAV_CHECK(placed >= 0 && placed < (int)h->g.b32.size(), "placed id out of range");
  const B32 &b = h->g.b32[placed];
  for (long i = 0; i < n; ++i) Bound32(b, h->g.b32side.data(), F3{x[i], y[i], z[i]}, so[i], si[i], m[i]);
}

}  // extern "C"
