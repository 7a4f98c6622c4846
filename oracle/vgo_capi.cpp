// ORACLE -- test infrastructure only (see vgo.h).
// C-ABI of the oracle.  The same entry points are exported by the shim over
// the real reference build (oracle/ref/refshim.cpp, prefix vgref_), so tests
// can swap one for the other.
#include <cstring>
#include <thread>
#include <vector>

#include "vgo.h"

namespace vgo {
Shape *MakeBox(const double *p, int n);
Shape *MakeTube(const double *p, int n);
Shape *MakeZooShape(int type, const double *p, int n);
Shape *MakeBoolean(const double *p, int n);

Shape *MakeShape(int type, const double *p, int n) {
  switch (type) {
    case kBox: return MakeBox(p, n);
    case kTube: return MakeTube(p, n);
    case kBoolean: return MakeBoolean(p, n);
    default: return MakeZooShape(type, p, n);
  }
}
}  // namespace vgo

using namespace vgo;

namespace {
struct PlacedOwner {
  Placed placed;
  Shape *shape;
};

template <class F>
void ParallelFor(long n, int nthreads, F f) {
  if (nthreads <= 1 || n < 1024) {
    f(0L, n);
    return;
  }
  std::vector<std::thread> th;
  long chunk = (n + nthreads - 1) / nthreads;
  for (int t = 0; t < nthreads; ++t) {
    long b = t * chunk, e = std::min(n, b + chunk);
    if (b >= e) break;
    th.emplace_back([=] { f(b, e); });
  }
  for (auto &x : th) x.join();
}
}  // namespace

extern "C" {

// Query ids shared with the product's Python wrapper.
enum { Q_CONTAINS = 0, Q_INSIDE = 1, Q_DIST_IN = 2, Q_DIST_OUT = 3, Q_SAFE_IN = 4, Q_SAFE_OUT = 5, Q_PLACED_DIST_OUT = 6 };

void *vgo_placed_create(int type, const double *params, int nparams, const double *xf12) {
  Shape *s = MakeShape(type, params, nparams);
  if (!s) return nullptr;
  auto *o = new PlacedOwner;
  o->shape = s;
  o->placed.shape = s;
  o->placed.xf = Transform::FromArray(xf12);
  return o;
}

void vgo_placed_destroy(void *h) {
  auto *o = static_cast<PlacedOwner *>(h);
  if (!o) return;
  delete o->shape;
  delete o;
}

// Extent of the unplaced shape: lo[3], hi[3].
void vgo_shape_extent(void *h, double *lo, double *hi) {
  auto *o = static_cast<PlacedOwner *>(h);
  V3 l, u;
  o->shape->Extent(l, u);
  lo[0] = l.x; lo[1] = l.y; lo[2] = l.z;
  hi[0] = u.x; hi[1] = u.y; hi[2] = u.z;
}

// Runs one basket query over n points; returns 0 on success.
int vgo_query(void *h, int query, long n, const double *px, const double *py, const double *pz, const double *dx,
              const double *dy, const double *dz, const double *step_max, void *out, int nthreads) {
  auto *o = static_cast<PlacedOwner *>(h);
  if (!o) return -1;
  const Placed &P = o->placed;
  switch (query) {
    case Q_CONTAINS:
      ParallelFor(n, nthreads, [&](long b, long e) {
        uint8_t *r = static_cast<uint8_t *>(out);
        for (long i = b; i < e; ++i) r[i] = P.Contains(V3{px[i], py[i], pz[i]}) ? 1 : 0;
      });
      return 0;
    case Q_INSIDE:
      ParallelFor(n, nthreads, [&](long b, long e) {
        int32_t *r = static_cast<int32_t *>(out);
        for (long i = b; i < e; ++i) r[i] = P.Inside(V3{px[i], py[i], pz[i]});
      });
      return 0;
    case Q_DIST_IN:
      ParallelFor(n, nthreads, [&](long b, long e) {
        double *r = static_cast<double *>(out);
        for (long i = b; i < e; ++i)
          r[i] = P.DistanceToIn(V3{px[i], py[i], pz[i]}, V3{dx[i], dy[i], dz[i]}, step_max ? step_max[i] : kInfinity);
      });
      return 0;
    case Q_DIST_OUT:
      ParallelFor(n, nthreads, [&](long b, long e) {
        double *r = static_cast<double *>(out);
        for (long i = b; i < e; ++i)
          r[i] = P.DistanceToOut(V3{px[i], py[i], pz[i]}, V3{dx[i], dy[i], dz[i]}, step_max ? step_max[i] : kInfinity);
      });
      return 0;
    case Q_PLACED_DIST_OUT:
      ParallelFor(n, nthreads, [&](long b, long e) {
        double *r = static_cast<double *>(out);
        for (long i = b; i < e; ++i)
          r[i] = P.PlacedDistanceToOut(V3{px[i], py[i], pz[i]}, V3{dx[i], dy[i], dz[i]},
                                       step_max ? step_max[i] : kInfinity);
      });
      return 0;
    case Q_SAFE_IN:
      ParallelFor(n, nthreads, [&](long b, long e) {
        double *r = static_cast<double *>(out);
        for (long i = b; i < e; ++i) r[i] = P.SafetyToIn(V3{px[i], py[i], pz[i]});
      });
      return 0;
    case Q_SAFE_OUT:
      ParallelFor(n, nthreads, [&](long b, long e) {
        double *r = static_cast<double *>(out);
        for (long i = b; i < e; ++i) r[i] = P.SafetyToOut(V3{px[i], py[i], pz[i]});
      });
      return 0;
  }
  return -2;
}

// VPlacedVolume::Normal on local points: nx, ny, nz and valid (1/0) per
// point.  Returns -3 when the shape has no Normal in the reference.
int vgo_normal(void *h, long n, const double *px, const double *py, const double *pz, double *nx, double *ny,
               double *nz, uint8_t *valid) {
  auto *o = static_cast<PlacedOwner *>(h);
  if (!o) return -1;
  V3 t;
  if (o->shape->Normal(V3{0, 0, 0}, t) < 0) return -3;
  for (long i = 0; i < n; ++i) {
    V3 v;
    valid[i] = (uint8_t)o->shape->Normal(V3{px[i], py[i], pz[i]}, v);
    nx[i] = v.x; ny[i] = v.y; nz[i] = v.z;
  }
  return 0;
}

// The USolids DistanceToOut(point, direction, normal&, convex&, stepMax) on
// local rays: the distance, the normal at the exit point, its valid flag and
// convex (always 1).  Returns -3 when the shape has no Normal.
int vgo_distance_to_out_normal(void *h, long n, const double *px, const double *py, const double *pz,
                               const double *dx, const double *dy, const double *dz, const double *step_max,
                               double *dist, double *nx, double *ny, double *nz, uint8_t *valid, uint8_t *convex) {
  auto *o = static_cast<PlacedOwner *>(h);
  if (!o) return -1;
  V3 t;
  if (o->shape->Normal(V3{0, 0, 0}, t) < 0) return -3;
  for (long i = 0; i < n; ++i) {
    V3 v;
    bool cv;
    int ok;
    dist[i] = DistanceToOutNormal(*o->shape, V3{px[i], py[i], pz[i]}, V3{dx[i], dy[i], dz[i]},
                                  step_max ? step_max[i] : kInfinity, v, cv, ok);
    nx[i] = v.x; ny[i] = v.y; nz[i] = v.z;
    valid[i] = (uint8_t)ok;
    convex[i] = cv ? 1 : 0;
  }
  return 0;
}

// Forward transform of points by a placement (Transform<t,r> with the
// placement's codes); used by tests to build local-frame inputs.
void vgo_transform_points(const double *xf12, long n, const double *x, const double *y, const double *z, double *ox,
                          double *oy, double *oz, int direction) {
  Transform T = Transform::FromArray(xf12);
  for (long i = 0; i < n; ++i) {
    V3 v = direction ? T.ApplyDirection(V3{x[i], y[i], z[i]}) : T.Apply(V3{x[i], y[i], z[i]});
    ox[i] = v.x; oy[i] = v.y; oz[i] = v.z;
  }
}

int vgo_hardware_threads() { return (int)std::thread::hardware_concurrency(); }

}  // extern "C"
