// ORACLE -- test infrastructure only (see vgo.h).
//
// Navigation restated from the reference: the flattened volume tree, the
// ABBox bounding boxes, NavigationState as a path of placed ids,
// SimpleNavigator::LocatePoint / LocatePointExclVolume /
// RelocatePointFromPath, and the FindNextBoundaryAndStep of SimpleNavigator
// and ABBoxNavigator, plus the track loop of the reference's TraceTrack
// benchmark.
//
//   NavigationState      navigation/NavigationState.h:355-408
//   LocatePoint          navigation/SimpleNavigator.h:204-241
//   LocatePointExclVolume navigation/SimpleNavigator.h:246-289
//   RelocatePointFromPath navigation/SimpleNavigator.h:292-321
//   Simple FNBAS         navigation/SimpleNavigator.h:340-482
//   ABBox FNBAS          navigation/ABBoxNavigator.h:200-424
//   ComputeABBox         source/ABBoxManager.cpp:128-161
//   GetHitCandidates_v   source/ABBoxManager.cpp:80-126 (scalar float path)
//   IntersectCachedKernel2 volumes/kernel/BoxImplementation.h:334-384
//   GetSafety            navigation/ABBoxNavigator.h:428-452
//   track loop           test/globalbenchmarks/TraceTrack.cpp:188-220
//
// The ABBox hit list is sorted by distance with ties kept in daughter order
// (a stable sort).  The reference sorts with std::sort only when the hit
// count differs from that of the previous call (ABBoxNavigator.h:286-293);
// the reference shim forces the sort so that both follow the sorted order.
#include <algorithm>
#include <cstring>
#include <thread>
#include <vector>

#include "vgo.h"

namespace vgo {

struct NavGeom {
  std::vector<Shape *> shapes;
  std::vector<int> shapeType;
  std::vector<int> placedLogical;
  std::vector<Transform> xf;
  std::vector<int> logicalShape;
  std::vector<int> dbegin;  // n_logical + 1
  std::vector<int> daughters;
  std::vector<float> aabb;  // 6 floats per placed volume
  int world = -1;
  int maxDepth = 0;
  ~NavGeom() {
    for (auto s : shapes) delete s;
  }
  const Shape *ShapeOf(int pv) const { return shapes[logicalShape[placedLogical[pv]]]; }
  int DBegin(int pv) const { return dbegin[placedLogical[pv]]; }
  int DEnd(int pv) const { return dbegin[placedLogical[pv] + 1]; }
  // VPlacedVolume::Contains(point, local) with the placement's codes.
  bool Contains(int pv, V3 p, V3 &local) const {
    local = xf[pv].Apply(p);
    return ShapeOf(pv)->Contains(local);
  }
  double DistanceToIn(int pv, V3 p, V3 d, double sm) const {
    const Shape *s = ShapeOf(pv);
    V3 ld = s->DirectionUsesPointTransform() ? xf[pv].Apply(d) : xf[pv].ApplyDirection(d);
    return s->DistanceToIn(xf[pv].Apply(p), ld, sm);
  }
};

struct NavState {
  int level = 0;
  bool boundary = false;
  int path[32];
  int Top() const { return level > 0 ? path[level - 1] : -1; }
  void Push(int v) { path[level++] = v; }
  void Pop() {
    if (level > 0) path[--level] = -1;
  }
};

// NavigationState::TopMatrix, NavigationState.h:398-408
static Transform TopMatrix(const NavGeom &g, const NavState &s) {
  Transform m;
  for (int i = 1; i < s.level; ++i) m.MultiplyFromRight(g.xf[s.path[i]]);
  return m;
}

// SimpleNavigator::LocatePoint
static int LocatePoint(const NavGeom &g, int vol, V3 point, NavState &path, bool top) {
  int cand = vol;
  V3 tmp = point;
  if (top) cand = g.ShapeOf(vol)->Contains(point) ? vol : -1;
  if (cand >= 0) {
    path.Push(cand);
    bool godeeper = true;
    while (godeeper && g.DEnd(cand) > g.DBegin(cand)) {
      godeeper = false;
      for (int k = g.DBegin(cand); k < g.DEnd(cand); ++k) {
        int next = g.daughters[k];
        V3 tp;
        if (g.Contains(next, tmp, tp)) {
          path.Push(next);
          tmp = tp;
          cand = next;
          godeeper = true;
          break;
        }
      }
    }
  }
  return cand;
}

// SimpleNavigator::LocatePointExclVolume (reverse daughter order)
static int LocatePointExcl(const NavGeom &g, int vol, int excluded, V3 point, NavState &path) {
  int cand = vol;
  V3 tmp = point;
  path.Push(cand);
  bool godeeper = true;
  while (godeeper && g.DEnd(cand) > g.DBegin(cand)) {
    godeeper = false;
    for (int k = g.DEnd(cand) - 1; k >= g.DBegin(cand); --k) {
      int next = g.daughters[k];
      V3 tp;
      if (next != excluded && g.Contains(next, tmp, tp)) {
        path.Push(next);
        tmp = tp;
        cand = next;
        godeeper = true;
        break;
      }
    }
  }
  return cand;
}

// SimpleNavigator::RelocatePointFromPath
static int Relocate(const NavGeom &g, V3 localpoint, NavState &path) {
  int mother = path.Top();
  if (mother >= 0) {
    V3 tmp = localpoint;
    while (mother >= 0 && !g.ShapeOf(mother)->Contains(tmp)) {
      path.Pop();
      tmp = g.xf[mother].Inverse(tmp);
      mother = path.Top();
    }
    if (mother >= 0) {
      path.Pop();
      return LocatePointExcl(g, mother, mother, tmp, path);
    }
  }
  return mother;
}

// BoxImplementation::IntersectCachedKernel2 with float boxes
// (BoxImplementation.h:334-384, scalar float backend: IsFull = the bool).
static float IntersectBox(const float *lo, const float *hi, const float *p, const float *inv, const int *sign, float t0,
                          float t1) {
  const float *c[2] = {lo, hi};
  float tmin = (c[sign[0]][0] - p[0]) * inv[0];
  float tymax = (c[1 - sign[1]][1] - p[1]) * inv[1];
  if (tmin > tymax) return std::numeric_limits<float>::infinity();
  float tmax = (c[1 - sign[0]][0] - p[0]) * inv[0];
  float tymin = (c[sign[1]][1] - p[1]) * inv[1];
  if (tymin > tmax) return std::numeric_limits<float>::infinity();
  tmin = std::max(tmin, tymin);
  tmax = std::min(tmax, tymax);
  float tzmin = (c[sign[2]][2] - p[2]) * inv[2];
  float tzmax = (c[1 - sign[2]][2] - p[2]) * inv[2];
  if ((tmin > tzmax) || (tzmin > tmax)) return std::numeric_limits<float>::infinity();
  tmin = std::max(tmin, tzmin);
  tmax = std::min(tmax, tzmax);
  if (!((tmin < t1) && (tmax > t0))) tmin = std::numeric_limits<float>::infinity();
  return tmin;
}

// ABBoxNavigator::FindNextBoundaryAndStep (kind 0) and
// SimpleNavigator::FindNextBoundaryAndStep (kind 1).
static void FindNextBoundaryAndStep(const NavGeom &g, V3 gp, V3 gd, const NavState &cur, NavState &nst, double pstep,
                                    double &step, int kind) {
  Transform m = TopMatrix(g, cur);
  V3 lp = m.ApplyGeneric(gp);
  V3 ld = m.ApplyDirectionGeneric(gd);
  int cv = cur.Top();
  int next = -1;
  step = g.ShapeOf(cv)->DistanceToOut(lp, ld, pstep);
  if (kind == 0) {
    if (step < 0.) step = kInfinity;
    int db = g.DBegin(cv), de = g.DEnd(cv);
    if (de > db) {
      // GetHitCandidates_v (float slab test against every daughter box)
      float inv[3] = {1.f / (float)ld.x, 1.f / (float)ld.y, 1.f / (float)ld.z};
      float pf[3] = {(float)lp.x, (float)lp.y, (float)lp.z};
      int sign[3] = {inv[0] < 0, inv[1] < 0, inv[2] < 0};
      std::vector<std::pair<int, double>> hits;
      for (int k = db; k < de; ++k) {
        const float *bx = &g.aabb[6 * (size_t)g.daughters[k]];
        float dist = IntersectBox(bx, bx + 3, pf, inv, sign, 0.f, std::numeric_limits<float>::infinity());
        if (dist < std::numeric_limits<float>::infinity()) hits.push_back({k - db, (double)dist});
      }
      std::stable_sort(hits.begin(), hits.end(),
                       [](const std::pair<int, double> &a, const std::pair<int, double> &b) { return a.second < b.second; });
      for (auto &h : hits) {
        int cand = g.daughters[db + h.first];
        if (!(step < h.second)) {
          if (h.second < 0) {
            V3 tp;
            if (g.Contains(cand, lp, tp)) {
              step = 0;
              next = h.first;
              break;
            }
          }
          double dd = g.DistanceToIn(cand, lp, ld, step);
          next = (dd < step) ? h.first : next;
          step = (dd < step) ? dd : step;
        } else {
          break;
        }
      }
    }
    nst = cur;
    if (step > pstep) {
      step = pstep;
      nst.boundary = false;
      return;
    }
    nst.boundary = true;
    if (step > 1E30) {
      step = 0;
      nst.Pop();
      return;
    }
    if (step < 0.) step = 0.;
  } else {
    if (step <= 0.) step = kInfinity;
    for (int k = g.DBegin(cv); k < g.DEnd(cv); ++k) {
      double dd = g.DistanceToIn(g.daughters[k], lp, ld, step);
      bool valid = (dd < step && !std::isinf(dd));
      next = valid ? k - g.DBegin(cv) : next;
      step = valid ? dd : step;
    }
    nst = cur;
    if (step == kInfinity && pstep > 0.) {
      step = 0;
      nst.Pop();
      nst.boundary = true;
      return;
    }
    if (step > pstep) {
      step = pstep;
      nst.boundary = false;
      return;
    }
    nst.boundary = true;
    if (step < 0.) step = 0.;
  }
  const double push = (kind == 0) ? 1e-6 : 1e-6;
  V3 np{ld.x * (step + push) + lp.x, ld.y * (step + push) + lp.y, ld.z * (step + push) + lp.z};
  if (next != -1) {
    int nv = g.daughters[g.DBegin(cv) + next];
    LocatePoint(g, nv, g.xf[nv].ApplyGeneric(np), nst, false);
    return;
  }
  Relocate(g, np, nst);
  if (kind == 0) {
    while (nst.Top() == cur.Top()) nst.Pop();
  }
}

// SimpleNavigator::FindNextBoundaryAndStep for baskets (kind 3),
// navigation/SimpleNavigator.h:638-743, for one track of a basket whose
// current volume is `vol` (the reference takes the top volume of the
// basket's first state for every track, :672): local point and direction
// from the track's own TopMatrix (:660-667); the volume's DistanceToOut with
// the physics step as step limit, no repair of values <= 0 (:678-679;
// ShapeImplementationHelper.h:677-692 fills -1 / -2, both "no daughter");
// DistanceToInMinimize over the daughters in order: DistanceToIn with the
// current distance as step limit, taken if strictly smaller
// (ShapeImplementationHelper.h:637-658); distance > pstep -> physics step,
// off boundary (:708-713); otherwise on boundary, push by 1e-9 (:721-723),
// then LocatePoint into the daughter or RelocatePointFromPath (:725-740).
static void FindNextBoundaryAndStepBasket(const NavGeom &g, V3 gp, V3 gd, const NavState &cur, NavState &nst,
                                          double pstep, double &step, int vol) {
  Transform m = TopMatrix(g, cur);
  V3 lp = m.ApplyGeneric(gp);
  V3 ld = m.ApplyDirectionGeneric(gd);
  int next = -1;
  step = g.ShapeOf(vol)->DistanceToOut(lp, ld, pstep);
  for (int k = g.DBegin(vol); k < g.DEnd(vol); ++k) {
    double dd = g.DistanceToIn(g.daughters[k], lp, ld, step);
    if (dd < step) {
      step = dd;
      next = k - g.DBegin(vol);
    }
  }
  nst = cur;
  if (step > pstep) {
    step = pstep;
    nst.boundary = false;
    return;
  }
  nst.boundary = true;
  V3 np{ld.x * (step + 1e-9) + lp.x, ld.y * (step + 1e-9) + lp.y, ld.z * (step + 1e-9) + lp.z};
  if (next > -1) {
    int nv = g.daughters[g.DBegin(vol) + next];
    LocatePoint(g, nv, g.xf[nv].ApplyGeneric(np), nst, false);
    return;
  }
  Relocate(g, np, nst);
}

// ABBoxNavigator::GetSafety, ABBoxNavigator.h:428-452
static double GetSafety(const NavGeom &g, V3 gp, const NavState &cur) {
  Transform m = TopMatrix(g, cur);
  V3 lp = m.ApplyGeneric(gp);
  int cv = cur.Top();
  double safety = g.ShapeOf(cv)->SafetyToOut(lp);
  for (int k = g.DBegin(cv); k < g.DEnd(cv); ++k) {
    int d = g.daughters[k];
    double t = g.ShapeOf(d)->SafetyToIn(g.xf[d].Apply(lp));
    safety = Min(safety, t);
  }
  return safety;
}

static void ComputeAABB(const NavGeom &g, int pv, double *out6) {
  V3 lo, hi;
  g.ShapeOf(pv)->Extent(lo, hi);
  V3 delta = hi - lo;
  double mn[3] = {kInfinity, kInfinity, kInfinity}, mx[3] = {-kInfinity, -kInfinity, -kInfinity};
  for (int x = 0; x <= 1; ++x)
    for (int y = 0; y <= 1; ++y)
      for (int z = 0; z <= 1; ++z) {
        V3 c{lo.x + x * delta.x, lo.y + y * delta.y, lo.z + z * delta.z};
        V3 t = g.xf[pv].Inverse(c);
        for (int i = 0; i < 3; ++i) {
          mn[i] = std::min(mn[i], t[i]);
          mx[i] = std::max(mx[i], t[i]);
        }
      }
  for (int i = 0; i < 3; ++i) {
    out6[i] = mn[i] - 1E-3;
    out6[3 + i] = mx[i] + 1E-3;
  }
}

static int Depth(const NavGeom &g, int pv, int level) {
  int best = level;
  for (int k = g.DBegin(pv); k < g.DEnd(pv); ++k) best = std::max(best, Depth(g, g.daughters[k], level + 1));
  return best;
}

template <class F>
static void ParallelFor(long n, int nthreads, F f) {
  if (nthreads <= 1 || n < 64) {
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

}  // namespace vgo

using namespace vgo;

extern "C" {

// Flat geometry in the layout of include/vecgeom_amd.h (vg_flat_geometry).
void *vgo_geom_create(int n_shapes, const int *shape_type, const int *shape_param_begin, const double *params,
                      int n_logical, const int *logical_shape, const int *logical_dbegin, const int *daughters,
                      int n_placed, const int *placed_logical, const double *placed_xf, int world) {
  auto *g = new NavGeom;
  for (int s = 0; s < n_shapes; ++s) {
    Shape *sh = MakeShape(shape_type[s], params + shape_param_begin[s], shape_param_begin[s + 1] - shape_param_begin[s]);
    if (!sh) {
      delete g;
      return nullptr;
    }
    g->shapes.push_back(sh);
    g->shapeType.push_back(shape_type[s]);
  }
  g->logicalShape.assign(logical_shape, logical_shape + n_logical);
  g->dbegin.assign(logical_dbegin, logical_dbegin + n_logical + 1);
  g->daughters.assign(daughters, daughters + g->dbegin[n_logical]);
  g->placedLogical.assign(placed_logical, placed_logical + n_placed);
  for (int p = 0; p < n_placed; ++p) g->xf.push_back(Transform::FromArray(placed_xf + 12 * (size_t)p));
  g->world = world;
  g->aabb.resize(6 * (size_t)n_placed);
  for (int p = 0; p < n_placed; ++p) {
    double b[6];
    ComputeAABB(*g, p, b);
    for (int i = 0; i < 6; ++i) g->aabb[6 * (size_t)p + i] = (float)b[i];
  }
  g->maxDepth = Depth(*g, world, 1);
  return g;
}

void vgo_geom_destroy(void *h) { delete static_cast<NavGeom *>(h); }

int vgo_geom_max_depth(void *h) { return static_cast<NavGeom *>(h)->maxDepth; }

// Placed-volume AABB (as floats, the form the navigator tests).
void vgo_geom_aabb(void *h, int pv, float *out6) {
  auto *g = static_cast<NavGeom *>(h);
  for (int i = 0; i < 6; ++i) out6[i] = g->aabb[6 * (size_t)pv + i];
}

// LocatePoint from the world (top = true).  path is depth-major:
// path[l * n + i] = placed id at level l of point i (-1 beyond its level).
void vgo_locate(void *h, long n, const double *x, const double *y, const double *z, int depth, int *level, int *path,
                int nthreads) {
  auto *g = static_cast<NavGeom *>(h);
  ParallelFor(n, nthreads, [&](long b, long e) {
    for (long i = b; i < e; ++i) {
      NavState s;
      LocatePoint(*g, g->world, V3{x[i], y[i], z[i]}, s, true);
      level[i] = s.level;
      for (int l = 0; l < depth; ++l) path[l * n + i] = l < s.level ? s.path[l] : -1;
    }
  });
}

// One FindNextBoundaryAndStep per track; states in/out depth-major.
void vgo_nav_step(void *h, long n, const double *x, const double *y, const double *z, const double *dx,
                  const double *dy, const double *dz, int depth, const int *level_in, const int *path_in,
                  int *level_out, int *path_out, unsigned char *boundary_out, double *step_out, double pstep, int kind,
                  int nthreads) {
  auto *g = static_cast<NavGeom *>(h);
  // Basket navigator: the current volume of the whole basket is the top of
  // the first track's state.
  const int basket_vol = (kind == 3 && n > 0 && level_in[0] > 0) ? path_in[(level_in[0] - 1) * n + 0] : -1;
  ParallelFor(n, nthreads, [&](long b, long e) {
    for (long i = b; i < e; ++i) {
      NavState cur, nst;
      cur.level = level_in[i];
      for (int l = 0; l < cur.level; ++l) cur.path[l] = path_in[l * n + i];
      double step;
      if (kind == 3)
        FindNextBoundaryAndStepBasket(*g, V3{x[i], y[i], z[i]}, V3{dx[i], dy[i], dz[i]}, cur, nst, pstep, step,
                                      basket_vol);
      else
        FindNextBoundaryAndStep(*g, V3{x[i], y[i], z[i]}, V3{dx[i], dy[i], dz[i]}, cur, nst, pstep, step, kind);
      level_out[i] = nst.level;
      for (int l = 0; l < depth; ++l) path_out[l * n + i] = l < nst.level ? nst.path[l] : -1;
      boundary_out[i] = nst.boundary ? 1 : 0;
      step_out[i] = step;
    }
  });
}

void vgo_safety(void *h, long n, const double *x, const double *y, const double *z, int depth, const int *level_in,
                const int *path_in, double *out, int nthreads) {
  auto *g = static_cast<NavGeom *>(h);
  ParallelFor(n, nthreads, [&](long b, long e) {
    for (long i = b; i < e; ++i) {
      NavState cur;
      cur.level = level_in[i];
      for (int l = 0; l < cur.level; ++l) cur.path[l] = path_in[l * n + i];
      out[i] = GetSafety(*g, V3{x[i], y[i], z[i]}, cur);
    }
  });
}

// The TraceTrack loop: locate, then FindNextBoundaryAndStep and
// p += d * (step + 1e-6) until the state is outside, at most max_steps
// steps.  Per track: number of steps, final point, the summed step length,
// and a 64-bit hash of the sequence of (new top id, boundary flag).
void vgo_trace(void *h, long n, const double *x, const double *y, const double *z, const double *dx,
               const double *dy, const double *dz, double pstep, int kind, int max_steps, int *nsteps, double *fx,
               double *fy, double *fz, double *travelled, unsigned long long *hash, int nthreads) {
  auto *g = static_cast<NavGeom *>(h);
  ParallelFor(n, nthreads, [&](long b, long e) {
    for (long i = b; i < e; ++i) {
      V3 p{x[i], y[i], z[i]}, d{dx[i], dy[i], dz[i]};
      NavState cur, nst;
      LocatePoint(*g, g->world, p, cur, true);
      int k = 0;
      double dist = 0;
      unsigned long long hsh = 1469598103934665603ULL;
      while (cur.level > 0 && k < max_steps) {
        double step;
        FindNextBoundaryAndStep(*g, p, d, cur, nst, pstep, step, kind);
        dist += step;
        p = V3{p.x + d.x * (step + 1E-6), p.y + d.y * (step + 1E-6), p.z + d.z * (step + 1E-6)};
        cur = nst;
        unsigned long long v = (unsigned long long)(unsigned)(cur.Top() + 1) * 2ULL + (cur.boundary ? 1ULL : 0ULL);
        hsh = (hsh ^ v) * 1099511628211ULL;
        ++k;
      }
      nsteps[i] = k;
      fx[i] = p.x;
      fy[i] = p.y;
      fz[i] = p.z;
      travelled[i] = dist;
      hash[i] = hsh;
    }
  });
}

}  // extern "C"
