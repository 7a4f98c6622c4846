// ORACLE -- test infrastructure only.
//
// Navigation shim over the real reference build: it builds the reference's
// LogicalVolume / PlaceDaughter tree from the flat arrays of
// include/vecgeom_amd.h, closes the geometry, initialises the ABBox boxes,
// and calls SimpleNavigator / ABBoxNavigator through their public API.  It
// exports the same entry points as oracle/vgo_nav.cpp with the prefix
// vgref_.  No reference code is contained here.
//
// One adjustment goes through the reference's public API: before each
// ABBoxNavigator::FindNextBoundaryAndStep call the shared hit container
// (ABBoxManager::GetAllocatedHitContainer) is filled to one more element
// than any volume has daughters, so that the navigator's "sort only if the
// hit count changed since the previous call" test (ABBoxNavigator.h:286-293)
// always sorts.  Results then no longer depend on which track was navigated
// before.
#include <cmath>
#include <map>
#include <vector>

#include "base/Global.h"
#include "base/SOA3D.h"
#include "base/Transformation3D.h"
#include "base/Vector3D.h"
#include "management/GeoManager.h"
#include "navigation/ABBoxNavigator.h"
#include "navigation/NavigationState.h"
#include "navigation/SimpleNavigator.h"
#include "volumes/LogicalVolume.h"
#include "volumes/PlacedVolume.h"
#include "volumes/UnplacedVolume.h"

using namespace vecgeom;

VUnplacedVolume *RefMakeUnplaced(int type, const double *p, int n);  // refshim.cpp

namespace {

struct RefGeom {
  std::vector<VUnplacedVolume *> shapes;
  std::vector<LogicalVolume *> logical;
  std::vector<Transformation3D *> xf;
  std::vector<VPlacedVolume const *> placed;
  std::map<VPlacedVolume const *, int> ids;
  int world = -1;
  int depth = 0;
  int maxDaughters = 0;
};

RefGeom *g_active = nullptr;

void Activate(RefGeom *g) {
  if (g_active == g) return;
  GeoManager::Instance().SetWorld(g->placed[g->world]);
  GeoManager::Instance().CloseGeometry();
  ABBoxManager::Instance().InitABBoxesForCompleteGeometry();
  g_active = g;
}

NavigationState *Load(RefGeom *g, int level, const int *path, long n, long i) {
  NavigationState *s = NavigationState::MakeInstance(g->depth);
  s->Clear();
  for (int l = 0; l < level; ++l) s->Push(g->placed[path[l * n + i]]);
  return s;
}

void Store(RefGeom *g, NavigationState const *s, int depth, long n, long i, int *level, int *path) {
  level[i] = s->GetCurrentLevel();
  for (int l = 0; l < depth; ++l) path[l * n + i] = l < s->GetCurrentLevel() ? g->ids[s->At(l)] : -1;
}

void ForceSort(RefGeom *g) {
  auto &hits = ABBoxManager::Instance().GetAllocatedHitContainer();
  hits.assign(g->maxDaughters + 1, ABBoxManager::BoxIdDistancePair_t(0, 0.));
}

}  // namespace

extern "C" {

void *vgref_geom_create(int n_shapes, const int *shape_type, const int *shape_param_begin, const double *params,
                        int n_logical, const int *logical_shape, const int *logical_dbegin, const int *daughters,
                        int n_placed, const int *placed_logical, const double *placed_xf, int world) {
  auto *g = new RefGeom;
  for (int s = 0; s < n_shapes; ++s) {
    VUnplacedVolume *u =
        RefMakeUnplaced(shape_type[s], params + shape_param_begin[s], shape_param_begin[s + 1] - shape_param_begin[s]);
    if (!u) return nullptr;
    g->shapes.push_back(u);
  }
  for (int l = 0; l < n_logical; ++l) {
    g->logical.push_back(new LogicalVolume("lv", g->shapes[logical_shape[l]]));
    g->maxDaughters = std::max(g->maxDaughters, logical_dbegin[l + 1] - logical_dbegin[l]);
  }
  for (int p = 0; p < n_placed; ++p) {
    const double *x = placed_xf + 12 * (size_t)p;
    g->xf.push_back(new Transformation3D(x[0], x[1], x[2], x[3], x[4], x[5], x[6], x[7], x[8], x[9], x[10], x[11]));
  }
  g->placed.assign(n_placed, nullptr);
  // Place daughters in each logical volume's order.
  for (int l = 0; l < n_logical; ++l) {
    for (int k = logical_dbegin[l]; k < logical_dbegin[l + 1]; ++k) {
      int pid = daughters[k];
      g->placed[pid] = g->logical[l]->PlaceDaughter("pv", g->logical[placed_logical[pid]], g->xf[pid]);
    }
  }
  g->placed[world] = g->logical[placed_logical[world]]->Place("world", g->xf[world]);
  for (int p = 0; p < n_placed; ++p) g->ids[g->placed[p]] = p;
  g->world = world;
  GeoManager::Instance().SetWorld(g->placed[world]);
  GeoManager::Instance().CloseGeometry();
  g->depth = GeoManager::Instance().getMaxDepth();
  g_active = nullptr;
  Activate(g);
  return g;
}

int vgref_geom_max_depth(void *h) { return static_cast<RefGeom *>(h)->depth; }

void vgref_geom_destroy(void *) {}

void vgref_locate(void *h, long n, const double *x, const double *y, const double *z, int depth, int *level, int *path,
                  int) {
  auto *g = static_cast<RefGeom *>(h);
  Activate(g);
  SimpleNavigator nav;
  NavigationState *s = NavigationState::MakeInstance(g->depth);
  for (long i = 0; i < n; ++i) {
    s->Clear();
    nav.LocatePoint(g->placed[g->world], Vector3D<Precision>(x[i], y[i], z[i]), *s, true);
    Store(g, s, depth, n, i, level, path);
  }
  NavigationState::ReleaseInstance(s);
}

void vgref_nav_step(void *h, long n, const double *x, const double *y, const double *z, const double *dx,
                    const double *dy, const double *dz, int depth, const int *level_in, const int *path_in,
                    int *level_out, int *path_out, unsigned char *boundary_out, double *step_out, double pstep,
                    int kind, int) {
  auto *g = static_cast<RefGeom *>(h);
  Activate(g);
  ABBoxNavigator abnav;
  SimpleNavigator snav;
  if (kind == 3) {
    // The basket interface (SimpleNavigator.h:638-743) over the whole basket.
    SOA3D<Precision> gp(n), gd(n), lp(n), ld(n);
    for (long i = 0; i < n; ++i) {
      gp.set(i, x[i], y[i], z[i]);
      gd.set(i, dx[i], dy[i], dz[i]);
    }
    std::vector<NavigationState *> cur(n), nxt(n);
    for (long i = 0; i < n; ++i) {
      cur[i] = Load(g, level_in[i], path_in, n, i);
      nxt[i] = NavigationState::MakeInstance(g->depth);
    }
    std::vector<Precision> ps(n, pstep), safe(n), dist(n);
    std::vector<int> work(n);
    if (n > 0)
      snav.FindNextBoundaryAndStep(gp, gd, lp, ld, cur.data(), nxt.data(), ps.data(), safe.data(), dist.data(),
                                   work.data());
    for (long i = 0; i < n; ++i) {
      Store(g, nxt[i], depth, n, i, level_out, path_out);
      boundary_out[i] = nxt[i]->IsOnBoundary() ? 1 : 0;
      step_out[i] = dist[i];
      NavigationState::ReleaseInstance(cur[i]);
      NavigationState::ReleaseInstance(nxt[i]);
    }
    return;
  }
  NavigationState *nst = NavigationState::MakeInstance(g->depth);
  for (long i = 0; i < n; ++i) {
    NavigationState *cur = Load(g, level_in[i], path_in, n, i);
    nst->Clear();
    double step = 0;
    Vector3D<Precision> p(x[i], y[i], z[i]), d(dx[i], dy[i], dz[i]);
    if (kind == 0) {
      ForceSort(g);
      abnav.FindNextBoundaryAndStep(p, d, *cur, *nst, pstep, step);
    } else {
      snav.FindNextBoundaryAndStep(p, d, *cur, *nst, pstep, step);
    }
    Store(g, nst, depth, n, i, level_out, path_out);
    boundary_out[i] = nst->IsOnBoundary() ? 1 : 0;
    step_out[i] = step;
    NavigationState::ReleaseInstance(cur);
  }
  NavigationState::ReleaseInstance(nst);
}

void vgref_safety(void *h, long n, const double *x, const double *y, const double *z, int, const int *level_in,
                  const int *path_in, double *out, int) {
  auto *g = static_cast<RefGeom *>(h);
  Activate(g);
  ABBoxNavigator abnav;
  for (long i = 0; i < n; ++i) {
    NavigationState *cur = Load(g, level_in[i], path_in, n, i);
    out[i] = abnav.GetSafety(Vector3D<Precision>(x[i], y[i], z[i]), *cur);
    NavigationState::ReleaseInstance(cur);
  }
}

void vgref_trace(void *h, long n, const double *x, const double *y, const double *z, const double *dx,
                 const double *dy, const double *dz, double pstep, int kind, int max_steps, int *nsteps, double *fx,
                 double *fy, double *fz, double *travelled, unsigned long long *hash, int) {
  auto *g = static_cast<RefGeom *>(h);
  Activate(g);
  ABBoxNavigator abnav;
  SimpleNavigator snav;
  NavigationState *cur = NavigationState::MakeInstance(g->depth);
  NavigationState *nst = NavigationState::MakeInstance(g->depth);
  for (long i = 0; i < n; ++i) {
    Vector3D<Precision> p(x[i], y[i], z[i]), d(dx[i], dy[i], dz[i]);
    cur->Clear();
    snav.LocatePoint(g->placed[g->world], p, *cur, true);
    int k = 0;
    double dist = 0;
    unsigned long long hsh = 1469598103934665603ULL;
    while (!cur->IsOutside() && k < max_steps) {
      double step = 0;
      nst->Clear();
      if (kind == 0) {
        ForceSort(g);
        abnav.FindNextBoundaryAndStep(p, d, *cur, *nst, pstep, step);
      } else {
        snav.FindNextBoundaryAndStep(p, d, *cur, *nst, pstep, step);
      }
      dist += step;
      p = p + d * (step + 1E-6);
      nst->CopyTo(cur);
      int top = cur->Top() ? g->ids[cur->Top()] : -1;
      unsigned long long v = (unsigned long long)(unsigned)(top + 1) * 2ULL + (cur->IsOnBoundary() ? 1ULL : 0ULL);
      hsh = (hsh ^ v) * 1099511628211ULL;
      ++k;
    }
    nsteps[i] = k;
    fx[i] = p.x();
    fy[i] = p.y();
    fz[i] = p.z();
    travelled[i] = dist;
    hash[i] = hsh;
  }
  NavigationState::ReleaseInstance(cur);
  NavigationState::ReleaseInstance(nst);
}

}  // extern "C"
