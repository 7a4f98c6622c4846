// ORACLE -- test infrastructure only.
//
// C-ABI shim over the real reference build (compiled from the reference's own
// sources by oracle/Makefile into oracle/_ref/libvgref.so).  It exports the
// same entry points as the oracle restatement (vgo_capi.cpp) with the prefix
// vgref_, so tests check the restatement against the reference and write
// golden vectors from it.  This file contains no reference code: it only
// constructs reference objects and calls their public virtual interface
// (volumes/PlacedVolume.h:144-250).
#include <cstdint>
#include <cstring>
#include <cmath>
#include <thread>
#include <vector>

#include "base/Global.h"
#include "base/Transformation3D.h"
#include "base/SOA3D.h"
#include "base/Vector3D.h"
#include "volumes/LogicalVolume.h"
#include "volumes/PlacedVolume.h"
#include "volumes/Box.h"
#include "volumes/Tube.h"
#include "volumes/Orb.h"
#include "volumes/Sphere.h"
#include "volumes/Cone.h"
#include "volumes/Trd.h"
#include "volumes/Trapezoid.h"
#include "volumes/Parallelepiped.h"
#include "volumes/Paraboloid.h"
#include "volumes/Torus.h"
#include "volumes/Polycone.h"
#include "volumes/Polyhedron.h"
#include "volumes/UnplacedBooleanVolume.h"

using namespace vecgeom;

namespace {

struct RefPlaced {
  VUnplacedVolume *unplaced = nullptr;
  LogicalVolume *logical = nullptr;
  Transformation3D *xf = nullptr;
  VPlacedVolume const *placed = nullptr;
  int type = 0;
  std::vector<double> params;
};

// The reference has no Extent for the parallelepiped
// (volumes/PlacedVolume.h:260-263 asserts); this bound of its 8 corners is
// the convention shared with the oracle and the product.
void ParallelepipedExtent(const std::vector<double> &p, double *lo, double *hi) {
  const double k = kDegToRad;  // angles are in degrees (source/UnplacedParallelepiped.cpp:48-70)
  const double ta = std::tan(k * p[3]), tt = std::tan(k * p[4]);
  const double ex = p[0] + std::fabs(p[1] * ta) + std::fabs(p[2] * tt * std::cos(k * p[5]));
  const double ey = p[1] + std::fabs(p[2] * tt * std::sin(k * p[5]));
  const double ez = p[2];
  lo[0] = -ex; lo[1] = -ey; lo[2] = -ez;
  hi[0] = ex; hi[1] = ey; hi[2] = ez;
}

VUnplacedVolume *MakeUnplaced(int type, const double *p, int n) {
  switch (type) {
    case 1: if (n == 3) return new UnplacedBox(p[0], p[1], p[2]); break;
    case 2: if (n == 5) return new UnplacedTube(p[0], p[1], p[2], p[3], p[4]); break;
    case 3:
      if (n == 1) {
        // UnplacedOrb's constructor reads frTolerance before initialising it
        // (volumes/UnplacedOrb.h:29-36, source/UnplacedOrb.cpp:35-43);
        // SetRadius recomputes the tolerances from initialised members.
        auto *o = new UnplacedOrb(p[0]);
        o->SetRadius(p[0]);
        return o;
      }
      break;
    case 4: if (n == 6) return new UnplacedSphere(p[0], p[1], p[2], p[3], p[4], p[5]); break;
    case 5: if (n == 7) return new UnplacedCone(p[0], p[1], p[2], p[3], p[4], p[5], p[6]); break;
    case 6: if (n == 5) return new UnplacedTrd(p[0], p[1], p[2], p[3], p[4]); break;
    case 7: if (n == 11) return new UnplacedTrapezoid(p[0], p[1], p[2], p[3], p[4], p[5], p[6], p[7], p[8], p[9], p[10]); break;
    case 8: if (n == 6) return new UnplacedParallelepiped(p[0], p[1], p[2], p[3], p[4], p[5]); break;
    case 9: if (n == 3) return new UnplacedParaboloid(p[0], p[1], p[2]); break;
    case 10: if (n == 5) return new UnplacedTorus(p[0], p[1], p[2], p[3], p[4]); break;
    case 11: {
      // sphi, dphi, nz, rmin[nz], rmax[nz], z[nz]
      if (n < 3) break;
      int nz = (int)p[2];
      if (n != 3 + 3 * nz) break;
      std::vector<double> rmin(p + 3, p + 3 + nz), rmax(p + 3 + nz, p + 3 + 2 * nz), z(p + 3 + 2 * nz, p + 3 + 3 * nz);
      return new UnplacedPolycone(p[0], p[1], nz, rmin.data(), rmax.data(), z.data());
    }
    case 12: {
      // phiStart(deg), phiDelta(deg), sides, nz, z[nz], rmin[nz], rmax[nz]
      if (n < 4) break;
      int nz = (int)p[3];
      if (n != 4 + 3 * nz) break;
      std::vector<double> z(p + 4, p + 4 + nz), rmin(p + 4 + nz, p + 4 + 2 * nz), rmax(p + 4 + 2 * nz, p + 4 + 3 * nz);
      return new UnplacedPolyhedron(p[0], p[1], (int)p[2], nz, z.data(), rmin.data(), rmax.data());
    }
    case 13: {
      // op (BooleanOperation), then per operand: type, nparams, params,
      // placement[12]; operands placed with the reference's own classes
      // (volumes/UnplacedBooleanVolume.h:39-50).
      if (n < 1) break;
      VPlacedVolume const *operand[2] = {nullptr, nullptr};
      int at = 1;
      for (int k = 0; k < 2; ++k) {
        if (at + 2 > n) return nullptr;
        const int t = (int)p[at], np = (int)p[at + 1];
        if (t == 13 || np < 0 || at + 2 + np + 12 > n) return nullptr;
        VUnplacedVolume *u = MakeUnplaced(t, p + at + 2, np);
        if (!u) return nullptr;
        const double *x = p + at + 2 + np;
        auto *lv = new LogicalVolume(k == 0 ? "left" : "right", u);
        operand[k] = lv->Place(new Transformation3D(x[0], x[1], x[2], x[3], x[4], x[5], x[6], x[7], x[8], x[9], x[10],
                                                    x[11]));
        at += 2 + np + 12;
      }
      if (at != n || p[0] < 0 || p[0] > 2) break;
      return new UnplacedBooleanVolume((BooleanOperation)(int)p[0], operand[0], operand[1]);
    }
  }
  return nullptr;
}

Vector3D<Precision> V(const double *x, const double *y, const double *z, long i) {
  return Vector3D<Precision>(x[i], y[i], z[i]);
}

}  // namespace

VUnplacedVolume *RefMakeUnplaced(int type, const double *p, int n) { return MakeUnplaced(type, p, n); }

extern "C" {

void *vgref_placed_create(int type, const double *params, int nparams, const double *xf12) {
  VUnplacedVolume *u = MakeUnplaced(type, params, nparams);
  if (!u) return nullptr;
  auto *r = new RefPlaced;
  r->unplaced = u;
  r->logical = new LogicalVolume("shape", u);
  r->xf = new Transformation3D(xf12[0], xf12[1], xf12[2], xf12[3], xf12[4], xf12[5], xf12[6], xf12[7], xf12[8],
                               xf12[9], xf12[10], xf12[11]);
  r->placed = r->logical->Place(r->xf);
  r->type = type;
  r->params.assign(params, params + nparams);
  return r;
}

void vgref_placed_destroy(void *) {
  // Reference objects register themselves in GeoManager; they are kept
  // alive for the life of the process.
}

int vgref_query(void *h, int query, long n, const double *px, const double *py, const double *pz, const double *dx,
                const double *dy, const double *dz, const double *step_max, void *out, int) {
  auto *r = static_cast<RefPlaced *>(h);
  if (!r) return -1;
  VPlacedVolume const *pv = r->placed;
  for (long i = 0; i < n; ++i) {
    double sm = step_max ? step_max[i] : kInfinity;
    switch (query) {
      case 0: static_cast<uint8_t *>(out)[i] = pv->Contains(V(px, py, pz, i)) ? 1 : 0; break;
      case 1: static_cast<int32_t *>(out)[i] = (int32_t)pv->Inside(V(px, py, pz, i)); break;
      case 2: static_cast<double *>(out)[i] = pv->DistanceToIn(V(px, py, pz, i), V(dx, dy, dz, i), sm); break;
      case 3: static_cast<double *>(out)[i] = pv->DistanceToOut(V(px, py, pz, i), V(dx, dy, dz, i), sm); break;
      case 4: static_cast<double *>(out)[i] = pv->SafetyToIn(V(px, py, pz, i)); break;
      case 5: static_cast<double *>(out)[i] = pv->SafetyToOut(V(px, py, pz, i)); break;
      case 6: static_cast<double *>(out)[i] = pv->PlacedDistanceToOut(V(px, py, pz, i), V(dx, dy, dz, i), sm); break;
      default: return -2;
    }
  }
  return 0;
}

// The shapes whose Normal the reference implements (PlacedBox.h:90,
// PlacedTube.h:94, PlacedOrb.h:95, PlacedSphere.h:223, PlacedCone.h:96,
// PlacedTrd.h:96, PlacedTrapezoid.h:111, PlacedPolycone.h:87); the others
// assert (PlacedVolume.h:266-272, PlacedParaboloid.h:122).
static bool HasNormal(int type) { return (type >= 1 && type <= 7) || type == 11; }

// VPlacedVolume::Normal on local points.  The Cone's kernel asserts for a
// point on no surface where NDEBUG is not defined (the strict build): such
// points go only to the Release build.
int vgref_normal(void *h, long n, const double *px, const double *py, const double *pz, double *nx, double *ny,
                 double *nz, uint8_t *valid) {
  auto *r = static_cast<RefPlaced *>(h);
  if (!r) return -1;
  if (!HasNormal(r->type)) return -3;
  for (long i = 0; i < n; ++i) {
    Vector3D<Precision> nv;
    valid[i] = r->placed->Normal(V(px, py, pz, i), nv) ? 1 : 0;
    nx[i] = nv.x(); ny[i] = nv.y(); nz[i] = nv.z();
  }
  return 0;
}

// The USolids DistanceToOut(point, direction, normal&, convex&, stepMax)
// (volumes/ShapeImplementationHelper.h:277-295, compiled only with
// VECGEOM_USOLIDS), evaluated with the reference's own virtuals and
// Vector3D arithmetic: the distance, the Normal of point + d * direction,
// convex = true.
int vgref_distance_to_out_normal(void *h, long n, const double *px, const double *py, const double *pz,
                                 const double *dx, const double *dy, const double *dz, const double *step_max,
                                 double *dist, double *nx, double *ny, double *nz, uint8_t *valid, uint8_t *convex) {
  auto *r = static_cast<RefPlaced *>(h);
  if (!r) return -1;
  if (!HasNormal(r->type)) return -3;
  for (long i = 0; i < n; ++i) {
    const Vector3D<Precision> point = V(px, py, pz, i), direction = V(dx, dy, dz, i);
    const double d = r->placed->DistanceToOut(point, direction, step_max ? step_max[i] : kInfinity);
    Vector3D<double> hitpoint = point + d * direction;
    Vector3D<Precision> nv;
    valid[i] = r->placed->Normal(hitpoint, nv) ? 1 : 0;
    dist[i] = d;
    nx[i] = nv.x(); ny[i] = nv.y(); nz[i] = nv.z();
    convex[i] = 1;
  }
  return 0;
}

// The reference's own SOA3D basket virtuals (volumes/PlacedVolume.h:147-246,
// basket loops volumes/ShapeImplementationHelper.h:588-748), each thread on a
// contiguous range through non-owning SOA3D views (base/SOA3D.h:173-176).
// Queries as in vgref_query (0..5); Contains writes uint8 through a bool
// buffer.  Used to time the reference's CPU path.
int vgref_basket(void *h, int query, long n, const double *px, const double *py, const double *pz, const double *dx,
                 const double *dy, const double *dz, const double *step_max, void *out, int nthreads) {
  auto *r = static_cast<RefPlaced *>(h);
  if (!r || query < 0 || query > 5) return -1;
  VPlacedVolume const *pv = r->placed;
  auto run = [&](long b, long e) {
    const size_t m = (size_t)(e - b);
    SOA3D<Precision> P(const_cast<double *>(px + b), const_cast<double *>(py + b), const_cast<double *>(pz + b), m);
    switch (query) {
      case 0: {
        std::vector<char> tmp(m);
        bool *o = reinterpret_cast<bool *>(tmp.data());
        pv->Contains(P, o);
        for (size_t i = 0; i < m; ++i) static_cast<uint8_t *>(out)[b + i] = o[i] ? 1 : 0;
        break;
      }
      case 1: pv->Inside(P, static_cast<Inside_t *>(out) + b); break;
      case 2: {
        SOA3D<Precision> D(const_cast<double *>(dx + b), const_cast<double *>(dy + b), const_cast<double *>(dz + b), m);
        pv->DistanceToIn(P, D, step_max + b, static_cast<double *>(out) + b);
        break;
      }
      case 3: {
        SOA3D<Precision> D(const_cast<double *>(dx + b), const_cast<double *>(dy + b), const_cast<double *>(dz + b), m);
        pv->DistanceToOut(P, D, step_max + b, static_cast<double *>(out) + b);
        break;
      }
      case 4: pv->SafetyToIn(P, static_cast<double *>(out) + b); break;
      case 5: pv->SafetyToOut(P, static_cast<double *>(out) + b); break;
    }
  };
  if (nthreads <= 1 || n < 1024) {
    run(0, n);
    return 0;
  }
  std::vector<std::thread> th;
  const long chunk = (n + nthreads - 1) / nthreads;
  for (int t = 0; t < nthreads; ++t) {
    const long b = t * chunk, e = std::min(n, b + chunk);
    if (b >= e) break;
    th.emplace_back(run, b, e);
  }
  for (auto &x : th) x.join();
  return 0;
}

void vgref_shape_extent(void *h, double *lo, double *hi) {
  auto *r = static_cast<RefPlaced *>(h);
  if (r->type == 8) {
    ParallelepipedExtent(r->params, lo, hi);
    return;
  }
  Vector3D<Precision> l, u;
  r->placed->Extent(l, u);
  for (int i = 0; i < 3; ++i) {
    lo[i] = l[i];
    hi[i] = u[i];
  }
}

}  // extern "C"
