// ORACLE -- test infrastructure only (see vgo.h).
// Box and Tube restated from the reference Scalar kernels.
#include "vgo.h"
#include "vgo_wedge.h"

namespace vgo {

// ---------------------------------------------------------------- Box
// volumes/kernel/BoxImplementation.h:593-837, volumes/UnplacedBox.h:19-23
struct Box : Shape {
  double d[3];
  Box(double dx, double dy, double dz) : d{dx, dy, dz} {}

  // BoxImplementation::NormalKernel, BoxImplementation.h:842-898: every face
  // within 100 kTolerance of the point adds the sign of that coordinate to
  // its axis; two or three faces (an edge, a corner) are scaled by 1/sqrt(2)
  // or 1/sqrt(3); valid when at least one face is that close.
  int Normal(V3 p, V3 &n) const override {
    const double delta = 100. * kTolerance;
    const double kInvSqrt2 = 0.7071067811865475, kInvSqrt3 = 0.5773502691896258;
    n = V3{0, 0, 0};
    double nsurf = 0;
    for (int dim = 0; dim < 3; ++dim) {
      const double safe = Abs(Abs(p[dim]) - d[dim]);
      if (safe < delta) {
        nsurf = nsurf + 1.;
        n[dim] = n[dim] + (p[dim] < 0 ? -1. : 1.);
      }
    }
    if (nsurf == 3.) n = n * kInvSqrt3;
    if (nsurf == 2.) n = n * kInvSqrt2;
    return nsurf > 0 ? 1 : 0;
  }

  // GenericKernelForContainsAndInside, BoxImplementation.h:608-653
  template <bool ForInside>
  void Kernel(V3 p, bool &cin, bool &cout) const {
    cout = Abs(p.x) > MakePlusTolerant<ForInside>(d[0]);
    if (ForInside) cin = Abs(p.x) < MakeMinusTolerant<ForInside>(d[0]);
    if (cout) return;
    cout |= Abs(p.y) > MakePlusTolerant<ForInside>(d[1]);
    if (ForInside) cin &= Abs(p.y) < MakeMinusTolerant<ForInside>(d[1]);
    if (cout) return;
    cout |= Abs(p.z) > MakePlusTolerant<ForInside>(d[2]);
    if (ForInside) cin &= Abs(p.z) < MakeMinusTolerant<ForInside>(d[2]);
  }
  bool Contains(V3 p) const override {
    bool cin = false, cout;
    Kernel<false>(p, cin, cout);
    return !cout;
  }
  // InsideKernel, BoxImplementation.h:655-668
  int Inside(V3 p) const override {
    bool cin = false, cout;
    Kernel<true>(p, cin, cout);
    int r = kSurface;
    if (cout) r = kOutside;
    if (cin) r = kInside;
    return r;
  }
  // DistanceToInKernel, BoxImplementation.h:673-746
  double DistanceToIn(V3 p, V3 dir, double stepMax) const override {
    double distance = kInfinity;
    double s[3] = {Abs(p.x) - d[0], Abs(p.y) - d[1], Abs(p.z) - d[2]};
    if (s[0] >= stepMax || s[1] >= stepMax || s[2] >= stepMax) return distance;
    if (s[0] < 0 && s[1] < 0 && s[2] < 0) return -1.;
    for (int a = 0; a < 3; ++a) {
      int b = a == 0 ? 1 : 0;
      int c = a == 2 ? 1 : 2;
      double next = s[a] / Abs(dir[a] + kMinimum);
      double c1 = p[b] + next * dir[b];
      double c2 = p[c] + next * dir[c];
      bool hit = s[a] >= MakeMinusTolerant<true>(0.) && p[a] * dir[a] < 0 && Abs(c1) <= d[b] && Abs(c2) <= d[c];
      if (hit) return next;
    }
    return distance;
  }
  // DistanceToOutKernel, BoxImplementation.h:751-799
  double DistanceToOut(V3 p, V3 dir, double) const override {
    double inv[3] = {1. / (dir.x + kMinimum), 1. / (dir.y + kMinimum), 1. / (dir.z + kMinimum)};
    double dist[3];
    for (int a = 0; a < 3; ++a) {
      dist[a] = (d[a] - p[a]) * inv[a];
      if (dir[a] < 0) dist[a] = (-d[a] - p[a]) * inv[a];
    }
    double distance = dist[0];
    if (dist[1] < distance) distance = dist[1];
    if (dist[2] < distance) distance = dist[2];
    return distance;
  }
  // SafetyToInKernel, BoxImplementation.h:804-818
  double SafetyToIn(V3 p) const override {
    double s = -d[0] + Abs(p.x);
    double sy = -d[1] + Abs(p.y);
    double sz = -d[2] + Abs(p.z);
    if (sy > s) s = sy;
    if (sz > s) s = sz;
    return s;
  }
  // SafetyToOutKernel, BoxImplementation.h:823-837
  double SafetyToOut(V3 p) const override {
    double s = d[0] - Abs(p.x);
    double sy = d[1] - Abs(p.y);
    double sz = d[2] - Abs(p.z);
    if (sy < s) s = sy;
    if (sz < s) s = sz;
    return s;
  }
  void Extent(V3 &lo, V3 &hi) const override {
    lo = {-d[0], -d[1], -d[2]};
    hi = {d[0], d[1], d[2]};
  }
};

// ---------------------------------------------------------------- Wedge
// volumes/Wedge.h:42-376, source/Wedge.cpp:16-43

// ---------------------------------------------------------------- Tube
// volumes/kernel/TubeImplementation.h:21-735, volumes/UnplacedTube.h:41-60,
// shapetypes/TubeTypes.h:104-211.  source/UnplacedTube.cpp:31-75 always
// instantiates UniversalTube (GENERATE_TUBE_SPECIALIZATIONS is never
// defined), so rmin/phi treatment and the sector class are runtime checks:
// hasRmin = rmin > 0, hasPhi = dphi < 2*M_PI, "smaller than pi" = dphi <= M_PI,
// and SectorType is kUnknownAngle.
struct Tube : Shape {
  double rmin, rmax, z, sphi, dphi;
  double rmin2, rmax2, a1x, a1y, a2x, a2y;
  double tolIrmin2, tolOrmin2, tolIrmax2, tolOrmax2, tolIz, tolOz;
  bool hasRmin, hasPhi, smallerThanPi;
  Wedge wedge;

  Tube(double rmin_, double rmax_, double z_, double sphi_, double dphi_)
      : rmin(rmin_), rmax(rmax_), z(z_), sphi(sphi_), dphi(dphi_) {
    // UnplacedTube::calculateCached, UnplacedTube.h:41-60
    tolIz = z - kTolerance;
    tolOz = z + kTolerance;
    rmin2 = rmin * rmin;
    rmax2 = rmax * rmax;
    double tolOrmin = rmin - kTolerance, tolIrmin = rmin + kTolerance;
    tolOrmin2 = tolOrmin * tolOrmin;
    tolIrmin2 = tolIrmin * tolIrmin;
    double tolOrmax = rmax + kTolerance, tolIrmax = rmax - kTolerance;
    tolOrmax2 = tolOrmax * tolOrmax;
    tolIrmax2 = tolIrmax * tolIrmax;
    a1x = std::cos(sphi);
    a1y = std::sin(sphi);
    a2x = std::cos(sphi + dphi);
    a2y = std::sin(sphi + dphi);
    wedge.Init(dphi, sphi);
    hasRmin = rmin > 0;
    hasPhi = dphi < 2 * M_PI;
    smallerThanPi = dphi <= M_PI;
  }

  // TubeUtilities::PointInCyclicalSector<onSurface=false>, :56-91
  bool InSector(double x, double y) const {
    double startCheck = (-x * a1y + y * a1x);
    double endCheck = (-a2x * y + a2y * x);
    if (smallerThanPi) return (startCheck >= -kTolerance) & (endCheck >= -kTolerance);
    return (startCheck >= -kTolerance) | (endCheck >= -kTolerance);
  }

  // CircleTrajectoryIntersection, :93-130
  template <bool Largest, bool InsectorCheck>
  void CircleHit(double b, double c, V3 pos, V3 dir, double &dist, bool &ok) const {
    double delta = b * b - c;
    bool mask = Largest ? (delta >= 0.) : (delta > 0.);
    if (!mask) delta = 0.;
    delta = Sqrt(delta);
    if (!Largest) delta = -delta;
    dist = -b + delta;
    if (InsectorCheck) {
      double hitz = pos.z + dist * dir.z;
      double hity = pos.y + dist * dir.y;
      double hitx = pos.x + dist * dir.x;
      bool insector = true;
      if (hasPhi) insector = wedge.ContainsWithBoundary(V3{hitx, hity, hitz});
      ok = mask & (dist >= 0) & (Abs(hitz) <= z) & insector;
    } else {
      ok = mask;
    }
  }

  // PhiPlaneSafety<inside>, :167-198 (SectorType kUnknownAngle branch)
  template <bool inside>
  double PhiPlaneSafety(V3 pos) const {
    double phi1 = pos.x * a1y - pos.y * a1x;
    if (inside) phi1 *= -1;
    if (phi1 < 0 || pos.x * a1x + pos.y * a1y < 0) phi1 = kInfinity;
    double phi2 = pos.x * a2y - pos.y * a2x;
    if (inside) phi2 *= -1;
    phi2 *= -1;
    if (phi2 < 0 || pos.x * a2x + pos.y * a2y < 0) phi2 = kInfinity;
    double safety = Min(phi1, phi2);
    if (safety == kInfinity) safety = -1;
    return safety;
  }

  // PhiPlaneTrajectoryIntersection, :215-265
  template <bool PositiveDir, bool InsectorCheck>
  void PhiPlaneHit(double ax, double ay, V3 pos, V3 dir, double &dist, bool &ok) const {
    dist = (ay * pos.x - ax * pos.y) / (dir.y * ax - dir.x * ay);
    if (InsectorCheck) {
      double hitx = pos.x + dist * dir.x;
      double hity = pos.y + dist * dir.y;
      double hitz = pos.z + dist * dir.z;
      double r2 = hitx * hitx + hity * hity;
      ok = Abs(hitz) <= tolIz && (r2 >= tolIrmin2) && (r2 <= tolIrmax2) && dist > 0;
      if (PositiveDir) ok = ok && (hitx * ax + hity * ay) > 0.;
    } else {
      if (PositiveDir) {
        double hitx = pos.x + dist * dir.x;
        double hity = pos.y + dist * dir.y;
        ok = (hitx * ax + hity * ay) >= 0.;
      }
    }
  }

  // GenericKernelForContainsAndInside, :287-343
  template <bool ForInside>
  void Kernel(V3 p, bool &cin, bool &cout) const {
    double absz = Abs(p.z);
    cout = absz > MakePlusTolerant<ForInside>(z);
    if (ForInside) cin = absz < MakeMinusTolerant<ForInside>(z);
    if (cout) return;
    double r2 = p.x * p.x + p.y * p.y;
    cout |= r2 > MakePlusTolerantSquare<ForInside>(rmax, rmax2);
    if (ForInside) cin &= r2 < MakeMinusTolerantSquare<ForInside>(rmax, rmax2);
    if (cout) return;
    if (hasRmin) {
      cout |= r2 <= MakeMinusTolerantSquare<ForInside>(rmin, rmin2);
      if (ForInside) cin &= r2 > MakePlusTolerantSquare<ForInside>(rmin, rmin2);
      if (cout) return;
    }
    if (hasPhi) {
      bool cinphi = false, coutphi;
      wedge.Kernel<ForInside>(p, cinphi, coutphi);
      cout |= coutphi;
      if (ForInside) cin &= cinphi;
    }
  }

  // UnplacedTube::Normal, source/UnplacedTube.cpp:141-229.  Surfaces are
  // tested in the order top, bottom, inner wall, outer wall, the two phi
  // planes; the first one found sets the vector, later ones add to it, and
  // the sum is divided by sqrt(count) when there are several.  The z test
  // takes kTolerance; the radial range of the lids is the exact one.
  int Normal(V3 p, V3 &n) const override {
    n = V3{0, 0, 0};
    int ns = 0;
    const double x2y2 = Sqrt(p.x * p.x + p.y * p.y);
    const bool inZ = (p.z < z + kTolerance) && (p.z > -z - kTolerance);
    const bool inR = (x2y2 >= rmin) && (x2y2 <= rmax);
    if (inR && Abs(p.z - z) <= kTolerance) {
      n = V3{0., 0., 1.};
      ns++;
    }
    if (inR && Abs(p.z + z) <= kTolerance) {
      if (ns > 0)
        n.z += -1;
      else
        n = V3{0.0, 0.0, -1};
      ns++;
    }
    if (rmin > 0.) {
      if (inZ && Abs(x2y2 - rmin) <= kTolerance) {
        const double inv = 1. / x2y2;
        if (ns == 0) {
          n = V3{-p.x * inv, -p.y * inv, 0.0};
        } else {
          n.x += -p.x * inv;
          n.y += -p.y * inv;
        }
        ns++;
      }
    }
    if (inZ && Abs(x2y2 - rmax) <= kTolerance) {
      const double inv = 1. / x2y2;
      if (ns > 0) {
        n.x += p.x * inv;
        n.y += p.y * inv;
      } else {
        n = V3{p.x * inv, p.y * inv, 0.0};
      }
      ns++;
    }
    if (dphi < kTwoPi) {
      // -fPhiWedge.GetNormal1/2(): the wedge's inward normals negated (z -0).
      const V3 o1{-wedge.n1x, -wedge.n1y, -0.0}, o2{-wedge.n2x, -wedge.n2y, -0.0};
      if (inR && Wedge::OnSurface(wedge.a1x, wedge.a1y, wedge.n1x, wedge.n1y, p)) {
        n = (ns == 0) ? o1 : n + o1;
        ns++;
      }
      if (inR && Wedge::OnSurface(wedge.a2x, wedge.a2y, wedge.n2x, wedge.n2y, p)) {
        n = (ns == 0) ? o2 : n + o2;
        ns++;
      }
    }
    if (ns > 1) {
      const double s = std::sqrt((double)ns);
      n = V3{n.x / s, n.y / s, n.z / s};
    }
    return ns != 0 ? 1 : 0;
  }

  bool Contains(V3 p) const override {
    bool cin = false, cout;
    Kernel<false>(p, cin, cout);
    return !cout;
  }
  int Inside(V3 p) const override {
    bool cin = false, cout;
    Kernel<true>(p, cin, cout);
    int r = kSurface;
    if (cout) r = kOutside;
    if (cin) r = kInside;
    return r;
  }

  // DistanceToIn, :400-554 (early returns as in the Scalar backend)
  double DistanceToIn(V3 pos, V3 dir, double) const override {
    double distance = kInfinity;
    double rsq = pos.x * pos.x + pos.y * pos.y;
    double pdx = pos.x * dir.x;
    double pdy = pos.y * dir.y;
    double abspz = Abs(pos.z);
    double distz = (abspz - z) / Abs(dir.z);
    bool okz = (pos.z * dir.z) < 0;
    if (abspz > z && !okz) return distance;
    if (Abs(pos.x) > rmax && pdx >= 0) return distance;
    if (Abs(pos.y) > rmax && pdy >= 0) return distance;
    double hitx = pos.x + distz * dir.x;
    double hity = pos.y + distz * dir.y;
    double r2 = hitx * hitx + hity * hity;
    okz = okz & (distz > 0) & (r2 <= tolIrmax2);
    if (hasRmin) okz = okz && (tolIrmin2 <= r2);
    if (hasPhi) okz = okz & InSector(hitx, hity);
    if (okz) return distz;

    double invnsq = 1 / (1 - dir.z * dir.z);
    double rdotn = pdx + pdy;
    double b = invnsq * rdotn;
    double crmax = invnsq * (rsq - rmax2);
    double dist_rmax;
    bool ok_rmax;
    CircleHit<false, true>(b, crmax, pos, dir, dist_rmax, ok_rmax);
    if (ok_rmax) return dist_rmax;
    double dist_rmin = 0;
    bool ok_rmin = false;
    if (hasRmin) {
      double crmin = invnsq * (rsq - rmin2);
      CircleHit<true, true>(b, crmin, pos, dir, dist_rmin, ok_rmin);
      if (ok_rmin) distance = dist_rmin;
    }
    // ok_rmax and okz are false here (early returns above).
    if (hasPhi) {
      double dist_phi;
      bool ok_phi;
      PhiPlaneHit<true, true>(a1x, a1y, pos, dir, dist_phi, ok_phi);
      if (ok_phi && dist_phi < distance) distance = dist_phi;
      PhiPlaneHit<true, true>(a2x, a2y, pos, dir, dist_phi, ok_phi);
      if (ok_phi && dist_phi < distance) distance = dist_phi;
    }
    return distance;
  }

  // DistanceToOut, :559-655 (the "bigger than pi or unknown" phi branch)
  double DistanceToOut(V3 p, V3 dir, double) const override {
    double distance = kInfinity;
    double dirzsign = 1.;
    if (dir.z < 0) dirzsign = -1.;
    double distz = (z * dirzsign - p.z) / dir.z;
    if (distz >= 0) distance = distz;
    double invnsq = 1 / (dir.x * dir.x + dir.y * dir.y);
    double rdotn = dir.x * p.x + dir.y * p.y;
    double rsq = p.x * p.x + p.y * p.y;
    double b = invnsq * rdotn;
    if (hasRmin) {
      double crmin = invnsq * (rsq - rmin2);
      double dist_rmin;
      bool ok_rmin;
      CircleHit<false, false>(b, crmin, p, dir, dist_rmin, ok_rmin);
      if (ok_rmin && dist_rmin >= 0 && dist_rmin < distance) distance = dist_rmin;
    }
    double crmax = invnsq * (rsq - rmax2);
    double dist_rmax;
    bool ok_rmax;
    CircleHit<true, false>(b, crmax, p, dir, dist_rmax, ok_rmax);
    if (ok_rmax && dist_rmax >= 0 && dist_rmax < distance) distance = dist_rmax;
    if (hasPhi) {
      double dist_phi;
      bool ok_phi = false;
      PhiPlaneHit<true, false>(a1x, a1y, p, dir, dist_phi, ok_phi);
      if (ok_phi && dist_phi > 0 && dist_phi < distance) distance = dist_phi;
      PhiPlaneHit<true, false>(a2x, a2y, p, dir, dist_phi, ok_phi);
      if (ok_phi && dist_phi > 0 && dist_phi < distance) distance = dist_phi;
    }
    return distance;
  }

  // SafetyToIn, :660-700
  double SafetyToIn(V3 p) const override {
    double r = Sqrt(p.x * p.x + p.y * p.y);
    double safez = Abs(p.z) - z;
    double safermax = r - rmax;
    double safety = Max(safez, safermax);
    if (hasRmin) {
      double safermin = rmin - r;
      safety = Max(safety, safermin);
    }
    if (hasPhi) {
      if (InSector(p.x, p.y)) return safety;
      double safephi = PhiPlaneSafety<false>(p);
      safety = Max(safety, safephi);
    }
    return safety;
  }

  // SafetyToOut, :705-733
  double SafetyToOut(V3 p) const override {
    double r = Sqrt(p.x * p.x + p.y * p.y);
    double safez = z - Abs(p.z);
    double safermax = rmax - r;
    double safety = Min(safez, safermax);
    if (hasRmin) {
      double safermin = r - rmin;
      safety = Min(safety, safermin);
    }
    if (hasPhi) {
      double safephi = PhiPlaneSafety<true>(p);
      if (safephi < 0) safephi = kInfinity;
      safety = Min(safety, safephi);
    }
    return safety;
  }
  // UnplacedTube::Extent, source/UnplacedTube.cpp:244-296
  void Extent(V3 &lo, V3 &hi) const override {
    lo = {-rmax, -rmax, -z};
    hi = {rmax, rmax, z};
    if (dphi == kTwoPi) return;
    double rin = 0.5 * (rmax + rmin);
    auto contains = [&](double x, double y) {
      bool cin = false, cout;
      wedge.Kernel<false>(V3{x, y, 0}, cin, cout);
      return !cout;
    };
    bool o0 = !contains(rin, 0), o90 = !contains(0, rin), o180 = !contains(-rin, 0), o270 = !contains(0, -rin);
    if (!(o0 || o90 || o180 || o270)) return;
    double Cx = rmax * std::cos(sphi), Dx = rmax * std::cos(sphi + dphi);
    double Cy = rmax * std::sin(sphi), Dy = rmax * std::sin(sphi + dphi);
    if (o0) hi.x = Max(Cx, Dx);
    if (o90) hi.y = Max(Cy, Dy);
    if (o180) lo.x = Min(Cx, Dx);
    if (o270) lo.y = Min(Cy, Dy);
    if (dphi >= kPi) return;
    double Ax = rmin * std::cos(sphi), Bx = rmin * std::cos(sphi + dphi);
    double Ay = rmin * std::sin(sphi), By = rmin * std::sin(sphi + dphi);
    double t = Max(Ax, Bx);
    hi.x = t > hi.x ? t : hi.x;
    t = Max(Ay, By);
    hi.y = t > hi.y ? t : hi.y;
    t = Min(Ax, Bx);
    lo.x = t < lo.x ? t : lo.x;
    t = Min(Ay, By);
    lo.y = t < lo.y ? t : lo.y;
  }
};

Shape *MakeBox(const double *p, int n) { return n == 3 ? new Box(p[0], p[1], p[2]) : nullptr; }
Shape *MakeTube(const double *p, int n) { return n == 5 ? new Tube(p[0], p[1], p[2], p[3], p[4]) : nullptr; }

// A tube whose rmin / phi treatment is fixed by its type rather than by its
// parameters (TubeTypes::HollowTube, NonHollowTube: shapetypes/TubeTypes.h:84-160).
Shape *MakeTubeTreat(const double *p, bool treatRmin, bool treatPhi) {
  auto *t = new Tube(p[0], p[1], p[2], p[3], p[4]);
  t->hasRmin = treatRmin;
  t->hasPhi = treatPhi;
  return t;
}

}  // namespace vgo
