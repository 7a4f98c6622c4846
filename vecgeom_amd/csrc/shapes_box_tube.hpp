// Box, Wedge and Tube: parameter blocks and device kernels.
//
// Parameter blocks are plain structs of doubles filled on the host by the
// geometry uploader (vg_geometry.cpp) with the reference's cached values,
// then passed to kernels by value (they live in SGPRs) or read from the
// device parameter array by the navigator.
//
// Reference semantics: volumes/kernel/BoxImplementation.h:593-837,
// volumes/kernel/TubeImplementation.h:21-735, volumes/Wedge.h:183-372,
// volumes/UnplacedTube.h:41-60.  Early returns follow the Scalar backend
// (kScalar::early_returns = true); on the GPU they are per-lane branches.
#pragma once
#include "vg_math.hpp"

namespace vga {

// ------------------------------------------------------------------ Box
struct BoxP {
  double d[3];  // half-lengths
};

struct BoxS {
  using P = BoxP;
  template <bool ForInside>
  static VGA_HD void Kernel(const P &b, V3 p, bool &cin, bool &cout) {
    cout = Abs(p.x) > PlusTol<ForInside>(b.d[0]);
    if (ForInside) cin = Abs(p.x) < MinusTol<ForInside>(b.d[0]);
    if (cout) return;
    cout |= Abs(p.y) > PlusTol<ForInside>(b.d[1]);
    if (ForInside) cin &= Abs(p.y) < MinusTol<ForInside>(b.d[1]);
    if (cout) return;
    cout |= Abs(p.z) > PlusTol<ForInside>(b.d[2]);
    if (ForInside) cin &= Abs(p.z) < MinusTol<ForInside>(b.d[2]);
  }
  static VGA_HD bool Contains(const P &b, V3 p) {
    // Branch-free form of the ForInside=false kernel: identical result.
    return !(int(Abs(p.x) > b.d[0]) | int(Abs(p.y) > b.d[1]) | int(Abs(p.z) > b.d[2]));
  }
  static VGA_HD int Inside(const P &b, V3 p) {
    bool cin = false, cout;
    Kernel<true>(b, p, cin, cout);
    int r = kSurface;
    if (cout) r = kOutside;
    if (cin) r = kInside;
    return r;
  }
  static VGA_HD double DistanceToIn(const P &b, V3 p, V3 dir, double stepMax) {
    const double s0 = Abs(p.x) - b.d[0], s1 = Abs(p.y) - b.d[1], s2 = Abs(p.z) - b.d[2];
    if (s0 >= stepMax || s1 >= stepMax || s2 >= stepMax) return kInfinity;
    if (s0 < 0 && s1 < 0 && s2 < 0) return -1.;
    const double mt = MinusTol<true>(0.);
    double next = s0 / Abs(dir.x + kMinimum);
    double c1 = p.y + next * dir.y, c2 = p.z + next * dir.z;
    if (s0 >= mt && p.x * dir.x < 0 && Abs(c1) <= b.d[1] && Abs(c2) <= b.d[2]) return next;
    next = s1 / Abs(dir.y + kMinimum);
    c1 = p.x + next * dir.x;
    c2 = p.z + next * dir.z;
    if (s1 >= mt && p.y * dir.y < 0 && Abs(c1) <= b.d[0] && Abs(c2) <= b.d[2]) return next;
    next = s2 / Abs(dir.z + kMinimum);
    c1 = p.x + next * dir.x;
    c2 = p.y + next * dir.y;
    if (s2 >= mt && p.z * dir.z < 0 && Abs(c1) <= b.d[0] && Abs(c2) <= b.d[1]) return next;
    return kInfinity;
  }
  static VGA_HD double DistanceToOut(const P &b, V3 p, V3 dir, double) {
    const double ix = 1. / (dir.x + kMinimum), iy = 1. / (dir.y + kMinimum), iz = 1. / (dir.z + kMinimum);
    const double dx = (dir.x < 0 ? (-b.d[0] - p.x) : (b.d[0] - p.x)) * ix;
    const double dy = (dir.y < 0 ? (-b.d[1] - p.y) : (b.d[1] - p.y)) * iy;
    const double dz = (dir.z < 0 ? (-b.d[2] - p.z) : (b.d[2] - p.z)) * iz;
    double dist = dx;
    if (dy < dist) dist = dy;
    if (dz < dist) dist = dz;
    return dist;
  }
  static VGA_HD double SafetyToIn(const P &b, V3 p) {
    double s = -b.d[0] + Abs(p.x);
    const double sy = -b.d[1] + Abs(p.y), sz = -b.d[2] + Abs(p.z);
    if (sy > s) s = sy;
    if (sz > s) s = sz;
    return s;
  }
  static VGA_HD double SafetyToOut(const P &b, V3 p) {
    double s = b.d[0] - Abs(p.x);
    const double sy = b.d[1] - Abs(p.y), sz = b.d[2] - Abs(p.z);
    if (sy < s) s = sy;
    if (sz < s) s = sz;
    return s;
  }
};

// ------------------------------------------------------------------ Wedge
// volumes/Wedge.h; cached vectors from source/Wedge.cpp:16-43.
struct WedgeP {
  double dphi;
  double a1x, a1y, a2x, a2y;  // along vectors of the two phi planes
  double n1x, n1y, n2x, n2y;  // inward normals
};

// Host-side construction of a wedge (vg_params.cpp).
WedgeP MakeWedge(double angle, double zero);

struct WedgeS {
  static VGA_HD bool OnSurface(double ax, double ay, double nx, double ny, V3 p) {
    if (!(ax * p.x + ay * p.y >= 0.)) return false;
    return Abs(nx * p.x + ny * p.y) < kTolerance;
  }
  template <bool ForInside>
  static VGA_HD void Kernel(const WedgeP &w, V3 p, bool &cin, bool &cout) {
    const double startCheck = (-p.x * w.a1y + p.y * w.a1x);
    const double endCheck = (-w.a2x * p.y + w.a2y * p.x);
    cout = startCheck < 0.;
    if (w.dphi < kPi)
      cout |= endCheck < 0.;
    else
      cout &= endCheck < 0.;
    if (ForInside) {
      const bool onSurface = OnSurface(w.a1x, w.a1y, w.n1x, w.n1y, p) || OnSurface(w.a2x, w.a2y, w.n2x, w.n2y, p);
      cout &= !onSurface;
      cin = !onSurface && !cout;
    }
  }
  static VGA_HD bool ContainsWithBoundary(const WedgeP &w, V3 p) {
    bool cin, cout;
    Kernel<true>(w, p, cin, cout);
    return !cout;
  }
  static VGA_HD double SafetyToOut(const WedgeP &w, V3 p) {
    const double d1 = p.x * w.n1x + p.y * w.n1y;
    const double d2 = p.x * w.n2x + p.y * w.n2y;
    return (w.dphi < kPi) ? Min(d1, d2) : Max(d1, d2);
  }
  static VGA_HD double SafetyToIn(const WedgeP &w, V3 p) {
    const double d1 = p.x * w.n1x + p.y * w.n1y;
    const double d2 = p.x * w.n2x + p.y * w.n2y;
    return (w.dphi < kPi) ? Max(-1 * d1, -1 * d2) : Min(-1 * d1, -1 * d2);
  }
  static VGA_HD void DistanceToIn(const WedgeP &w, V3 p, V3 dir, double &dw1, double &dw2) {
    dw1 = kInfinity;
    dw2 = kInfinity;
    const double comp1 = dir.x * w.n1x + dir.y * w.n1y;
    const double comp2 = dir.x * w.n2x + dir.y * w.n2y;
    if (comp1 > 0.) {
      const double t = -(p.x * w.n1x + p.y * w.n1y) / comp1;
      if (t > 0.) dw1 = t;
    }
    if (comp2 > 0.) {
      const double t = -(p.x * w.n2x + p.y * w.n2y) / comp2;
      if (t > 0.) dw2 = t;
    }
  }
  static VGA_HD void DistanceToOut(const WedgeP &w, V3 p, V3 dir, double &dw1, double &dw2) {
    const double comp1 = dir.x * w.n1x + dir.y * w.n1y;
    const double comp2 = dir.x * w.n2x + dir.y * w.n2y;
    dw1 = kInfinity;
    dw2 = kInfinity;
    if (comp1 < 0.) {
      const double t = -(p.x * w.n1x + p.y * w.n1y) / comp1;
      if (t > 0.) dw1 = t;
    }
    if (comp2 < 0.) {
      const double t = -(p.x * w.n2x + p.y * w.n2y) / comp2;
      if (t > 0.) dw2 = t;
    }
  }
};

// ------------------------------------------------------------------ Tube
// The reference always instantiates UniversalTube (source/UnplacedTube.cpp:
// 31-75), so rmin and phi handling are runtime flags here too.
struct TubeP {
  double rmin, rmax, z, sphi, dphi;
  double rmin2, rmax2, a1x, a1y, a2x, a2y;
  double tolIrmin2, tolOrmin2, tolIrmax2, tolOrmax2, tolIz, tolOz;
  double hasRmin, hasPhi, smallerThanPi;  // 0/1 flags
  WedgeP w;
};

// Host-side construction of a tube's parameter block (vg_params.cpp).
TubeP MakeTube(double rmin, double rmax, double z, double sphi, double dphi);

struct TubeS {
  using P = TubeP;
  static VGA_HD bool InSector(const P &t, double x, double y) {
    const double startCheck = (-x * t.a1y + y * t.a1x);
    const double endCheck = (-t.a2x * y + t.a2y * x);
    if (t.smallerThanPi != 0) return (startCheck >= -kTolerance) & (endCheck >= -kTolerance);
    return (startCheck >= -kTolerance) | (endCheck >= -kTolerance);
  }
  template <bool Largest, bool InsectorCheck>
  static VGA_HD void CircleHit(const P &t, double b, double c, V3 pos, V3 dir, double &dist, bool &ok) {
    double delta = b * b - c;
    const bool mask = Largest ? (delta >= 0.) : (delta > 0.);
    if (!mask) delta = 0.;
    delta = Sqrt(delta);
    if (!Largest) delta = -delta;
    dist = -b + delta;
    if (InsectorCheck) {
      const double hitz = pos.z + dist * dir.z;
      const double hity = pos.y + dist * dir.y;
      const double hitx = pos.x + dist * dir.x;
      bool insector = true;
      if (t.hasPhi != 0) insector = WedgeS::ContainsWithBoundary(t.w, V3{hitx, hity, hitz});
      ok = mask & (dist >= 0) & (Abs(hitz) <= t.z) & insector;
    } else {
      ok = mask;
    }
  }
  template <bool inside>
  static VGA_HD double PhiPlaneSafety(const P &t, V3 pos) {
    double phi1 = pos.x * t.a1y - pos.y * t.a1x;
    if (inside) phi1 *= -1;
    if (phi1 < 0 || pos.x * t.a1x + pos.y * t.a1y < 0) phi1 = kInfinity;
    double phi2 = pos.x * t.a2y - pos.y * t.a2x;
    if (inside) phi2 *= -1;
    phi2 *= -1;
    if (phi2 < 0 || pos.x * t.a2x + pos.y * t.a2y < 0) phi2 = kInfinity;
    double s = Min(phi1, phi2);
    if (s == kInfinity) s = -1;
    return s;
  }
  template <bool PositiveDir, bool InsectorCheck>
  static VGA_HD void PhiPlaneHit(const P &t, double ax, double ay, V3 pos, V3 dir, double &dist, bool &ok) {
    dist = (ay * pos.x - ax * pos.y) / (dir.y * ax - dir.x * ay);
    if (InsectorCheck) {
      const double hitx = pos.x + dist * dir.x;
      const double hity = pos.y + dist * dir.y;
      const double hitz = pos.z + dist * dir.z;
      const double r2 = hitx * hitx + hity * hity;
      ok = Abs(hitz) <= t.tolIz && (r2 >= t.tolIrmin2) && (r2 <= t.tolIrmax2) && dist > 0;
      if (PositiveDir) ok = ok && (hitx * ax + hity * ay) > 0.;
    } else if (PositiveDir) {
      const double hitx = pos.x + dist * dir.x;
      const double hity = pos.y + dist * dir.y;
      ok = (hitx * ax + hity * ay) >= 0.;
    }
  }
  template <bool ForInside>
  static VGA_HD void Kernel(const P &t, V3 p, bool &cin, bool &cout) {
    const double absz = Abs(p.z);
    cout = absz > PlusTol<ForInside>(t.z);
    if (ForInside) cin = absz < MinusTol<ForInside>(t.z);
    if (cout) return;
    const double r2 = p.x * p.x + p.y * p.y;
    cout |= r2 > PlusTolSq<ForInside>(t.rmax, t.rmax2);
    if (ForInside) cin &= r2 < MinusTolSq<ForInside>(t.rmax, t.rmax2);
    if (cout) return;
    if (t.hasRmin != 0) {
      cout |= r2 <= MinusTolSq<ForInside>(t.rmin, t.rmin2);
      if (ForInside) cin &= r2 > PlusTolSq<ForInside>(t.rmin, t.rmin2);
      if (cout) return;
    }
    if (t.hasPhi != 0) {
      bool cinphi = false, coutphi;
      WedgeS::Kernel<ForInside>(t.w, p, cinphi, coutphi);
      cout |= coutphi;
      if (ForInside) cin &= cinphi;
    }
  }
  static VGA_HD bool Contains(const P &t, V3 p) {
    bool cin = false, cout;
    Kernel<false>(t, p, cin, cout);
    return !cout;
  }
  static VGA_HD int Inside(const P &t, V3 p) {
    bool cin = false, cout;
    Kernel<true>(t, p, cin, cout);
    int r = kSurface;
    if (cout) r = kOutside;
    if (cin) r = kInside;
    return r;
  }
  static VGA_HD double DistanceToIn(const P &t, V3 pos, V3 dir, double) {
    double distance = kInfinity;
    const double rsq = pos.x * pos.x + pos.y * pos.y;
    const double pdx = pos.x * dir.x;
    const double pdy = pos.y * dir.y;
    const double abspz = Abs(pos.z);
    const double distz = (abspz - t.z) / Abs(dir.z);
    bool okz = (pos.z * dir.z) < 0;
    if (abspz > t.z && !okz) return distance;
    if (Abs(pos.x) > t.rmax && pdx >= 0) return distance;
    if (Abs(pos.y) > t.rmax && pdy >= 0) return distance;
    const double hitx = pos.x + distz * dir.x;
    const double hity = pos.y + distz * dir.y;
    const double r2 = hitx * hitx + hity * hity;
    okz = okz & (distz > 0) & (r2 <= t.tolIrmax2);
    if (t.hasRmin != 0) okz = okz && (t.tolIrmin2 <= r2);
    if (t.hasPhi != 0) okz = okz & InSector(t, hitx, hity);
    if (okz) return distz;
    const double invnsq = 1 / (1 - dir.z * dir.z);
    const double rdotn = pdx + pdy;
    const double b = invnsq * rdotn;
    const double crmax = invnsq * (rsq - t.rmax2);
    double dist_rmax;
    bool ok_rmax;
    CircleHit<false, true>(t, b, crmax, pos, dir, dist_rmax, ok_rmax);
    if (ok_rmax) return dist_rmax;
    if (t.hasRmin != 0) {
      const double crmin = invnsq * (rsq - t.rmin2);
      double dist_rmin;
      bool ok_rmin;
      CircleHit<true, true>(t, b, crmin, pos, dir, dist_rmin, ok_rmin);
      if (ok_rmin) distance = dist_rmin;
    }
    if (t.hasPhi != 0) {
      double dist_phi;
      bool ok_phi;
      PhiPlaneHit<true, true>(t, t.a1x, t.a1y, pos, dir, dist_phi, ok_phi);
      if (ok_phi && dist_phi < distance) distance = dist_phi;
      PhiPlaneHit<true, true>(t, t.a2x, t.a2y, pos, dir, dist_phi, ok_phi);
      if (ok_phi && dist_phi < distance) distance = dist_phi;
    }
    return distance;
  }
  static VGA_HD double DistanceToOut(const P &t, V3 p, V3 dir, double) {
    double distance = kInfinity;
    const double dirzsign = dir.z < 0 ? -1. : 1.;
    const double distz = (t.z * dirzsign - p.z) / dir.z;
    if (distz >= 0) distance = distz;
    const double invnsq = 1 / (dir.x * dir.x + dir.y * dir.y);
    const double rdotn = dir.x * p.x + dir.y * p.y;
    const double rsq = p.x * p.x + p.y * p.y;
    const double b = invnsq * rdotn;
    if (t.hasRmin != 0) {
      const double crmin = invnsq * (rsq - t.rmin2);
      double d;
      bool ok;
      CircleHit<false, false>(t, b, crmin, p, dir, d, ok);
      if (ok && d >= 0 && d < distance) distance = d;
    }
    {
      const double crmax = invnsq * (rsq - t.rmax2);
      double d;
      bool ok;
      CircleHit<true, false>(t, b, crmax, p, dir, d, ok);
      if (ok && d >= 0 && d < distance) distance = d;
    }
    if (t.hasPhi != 0) {
      double d;
      bool ok = false;
      PhiPlaneHit<true, false>(t, t.a1x, t.a1y, p, dir, d, ok);
      if (ok && d > 0 && d < distance) distance = d;
      PhiPlaneHit<true, false>(t, t.a2x, t.a2y, p, dir, d, ok);
      if (ok && d > 0 && d < distance) distance = d;
    }
    return distance;
  }
  static VGA_HD double SafetyToIn(const P &t, V3 p) {
    const double r = Sqrt(p.x * p.x + p.y * p.y);
    double safety = Max(Abs(p.z) - t.z, r - t.rmax);
    if (t.hasRmin != 0) safety = Max(safety, t.rmin - r);
    if (t.hasPhi != 0) {
      if (InSector(t, p.x, p.y)) return safety;
      safety = Max(safety, PhiPlaneSafety<false>(t, p));
    }
    return safety;
  }
  static VGA_HD double SafetyToOut(const P &t, V3 p) {
    const double r = Sqrt(p.x * p.x + p.y * p.y);
    double safety = Min(t.z - Abs(p.z), t.rmax - r);
    if (t.hasRmin != 0) safety = Min(safety, r - t.rmin);
    if (t.hasPhi != 0) {
      double sp = PhiPlaneSafety<true>(t, p);
      if (sp < 0) sp = kInfinity;
      safety = Min(safety, sp);
    }
    return safety;
  }
};

}  // namespace vga
