// ORACLE -- test infrastructure only (see vgo.h).
// Sphere, Cone, Trapezoid, Torus, Polycone, Polyhedron restatements.
#include "vgo.h"

namespace vgo {
Shape *MakeCone(const double *, int) { return nullptr; }
Shape *MakeSphere(const double *, int) { return nullptr; }
// ---------------------------------------------------------------- Trapezoid
// volumes/kernel/TrapezoidImplementation.h:140-470 with the PlaneShell
// variant (VECGEOM_PLANESHELL_DISABLE is not set: base/PlaneShell.h:95-300);
// planes from source/UnplacedTrapezoid.cpp:28-52, 209-345, 673-685.
struct Trapezoid : Shape {
  double dz, theta, phi, dy1, dx1, dx2, ta1, dy2, dx3, dx4, ta2, tcp, tsp;
  double A[4], B[4], C[4], D[4];
  bool good = true;
  Trapezoid(const double *p) {
    dz = p[0]; theta = p[1]; phi = p[2]; dy1 = p[3]; dx1 = p[4]; dx2 = p[5]; ta1 = p[6];
    dy2 = p[7]; dx3 = p[8]; dx4 = p[9]; ta2 = p[10];
    tsp = std::tan(theta) * std::sin(phi);
    tcp = std::tan(theta) * std::cos(phi);
    for (int i = 0; i < 4; ++i) A[i] = B[i] = C[i] = D[i] = 0;
    V3 pt[8];
    Corners(pt);
    good &= MakePlane(pt[0], pt[1], pt[5], pt[4], 0);
    good &= MakePlane(pt[2], pt[6], pt[7], pt[3], 1);
    good &= MakePlane(pt[0], pt[4], pt[6], pt[2], 2);
    good &= MakePlane(pt[1], pt[3], pt[7], pt[5], 3);
  }
  void Corners(V3 *pt) const {
    pt[0] = {-dz * tcp - dy1 * ta1 - dx1, -dz * tsp - dy1, -dz};
    pt[1] = {-dz * tcp - dy1 * ta1 + dx1, -dz * tsp - dy1, -dz};
    pt[2] = {-dz * tcp + dy1 * ta1 - dx2, -dz * tsp + dy1, -dz};
    pt[3] = {-dz * tcp + dy1 * ta1 + dx2, -dz * tsp + dy1, -dz};
    pt[4] = {+dz * tcp - dy2 * ta2 - dx3, +dz * tsp - dy2, +dz};
    pt[5] = {+dz * tcp - dy2 * ta2 + dx3, +dz * tsp - dy2, +dz};
    pt[6] = {+dz * tcp + dy2 * ta2 - dx4, +dz * tsp + dy2, +dz};
    pt[7] = {+dz * tcp + dy2 * ta2 + dx4, +dz * tsp + dy2, +dz};
  }
  static V3 Cross(V3 a, V3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
  bool MakePlane(V3 p1, V3 p2, V3 p3, V3 p4, int i) {
    V3 v12 = p2 - p1, v13 = p3 - p1, v14 = p4 - p1;
    V3 vc = Cross(v12, v13);
    if (std::fabs(Dot(v14, vc) / (vc.Mag() * v14.Mag())) > 1.0e-7) return false;
    double a = +(p3.y - p1.y) * (p4.z - p2.z) - (p4.y - p2.y) * (p3.z - p1.z);
    double b = -(p3.x - p1.x) * (p4.z - p2.z) + (p4.x - p2.x) * (p3.z - p1.z);
    double c = +(p3.x - p1.x) * (p4.y - p2.y) - (p4.x - p2.x) * (p3.y - p1.y);
    double norm = 1.0 / std::sqrt(a * a + b * b + c * c);
    a *= norm;
    b *= norm;
    c *= norm;
    double d = -(a * p1.x + b * p1.y + c * p1.z);
    A[i] = a; B[i] = b; C[i] = c; D[i] = d;
    return true;
  }
  double Dist(int i, V3 p) const { return A[i] * p.x + B[i] * p.y + C[i] * p.z + D[i]; }
  // TrapezoidImplementation::GenericKernelForContainsAndInside :134-160 +
  // PlaneShell::GenericKernelForContainsAndInside (PlaneShell.h:128-152)
  template <bool ForInside>
  void Kernel(V3 p, bool &cin, bool &cout) const {
    cout = Abs(p.z) > MakePlusTolerant<ForInside>(dz);
    if (cout) return;
    if (ForInside) cin = Abs(p.z) < MakeMinusTolerant<ForInside>(dz);
    double dist[4];
    for (int i = 0; i < 4; ++i) dist[i] = Dist(i, p);
    for (int i = 0; i < 4; ++i) {
      cout |= dist[i] > MakePlusTolerant<ForInside>(0.);
      if (cout) return;
      if (ForInside) cin &= dist[i] < MakeMinusTolerant<ForInside>(0.);
    }
  }
  bool Contains(V3 p) const override {
    bool cin = false, cout;
    Kernel<false>(p, cin, cout);
    return !cout;
  }
  // Inside :196-211 (outside assigned last)
  int Inside(V3 p) const override {
    bool cin = false, cout;
    Kernel<true>(p, cin, cout);
    int r = kSurface;
    if (cin && !cout) r = kInside;
    if (cout) r = kOutside;
    return r;
  }
  // PlaneShell::DistanceToIn, PlaneShell.h:163-214
  double ShellDistanceToIn(V3 point, double smin0, V3 dir) const {
    double smax = kInfinity, smin = smin0;
    bool done = false;
    double pdist[4], proj[4], vdist[4];
    for (int i = 0; i < 4; ++i) {
      pdist[i] = Dist(i, point);
      proj[i] = A[i] * dir.x + B[i] * dir.y + C[i] * dir.z;
      vdist[i] = -pdist[i] / proj[i];
    }
    for (int i = 0; i < 4; ++i) {
      bool posPoint = pdist[i] >= -kHalfTolerance;
      bool posDir = proj[i] > 0;
      done |= (posPoint && posDir);
      bool fromInside = (!posPoint && posDir);
      done |= (fromInside && vdist[i] < smin);
      bool fromOutside = (posPoint && !posDir);
      done |= (fromOutside && vdist[i] > smax);
      if (done) return kInfinity;
      bool valid = (smin < vdist[i] && vdist[i] < smax);
      if (fromOutside && valid) smin = vdist[i];
      if (fromInside && valid) smax = vdist[i];
    }
    V3 entry = point + dir * smin;
    bool valid = true;
    for (int i = 0; i < 4; ++i) valid &= (Dist(i, entry) <= MakePlusTolerant<true>(0.));
    return valid ? smin : kInfinity;
  }
  // DistanceToIn :214-262
  double DistanceToIn(V3 point, V3 dir, double) const override {
    double distance = kInfinity;
    bool posZ = dir.z > 0.0, negZ = dir.z < 0.0;
    double zsign = negZ ? -1. : 1.;
    double mx = zsign * dz - point.z;
    bool done = (posZ && mx < MakePlusTolerant<true>(0.));
    done |= (negZ && mx > MakeMinusTolerant<true>(0.));
    if (done) return distance;
    double dirFactor = 1. / dir.z;
    double smax = mx * dirFactor;
    double smin = (-zsign * dz - point.z) * dirFactor;
    bool test = (!posZ) && (!negZ);
    bool zrange = Abs(point.z) < MakeMinusTolerant<true>(dz);
    done |= (test && !zrange);
    if (done) return distance;
    if (test && zrange) {
      smin = 0.0;
      smax = kInfinity;
    }
    double dpl = ShellDistanceToIn(point, smin, dir);
    done |= (dpl == kInfinity);
    done |= (dpl > smax);
    if (done) return distance;
    distance = smin;
    if ((dpl > smin) && (dpl < smax)) distance = dpl;
    if (distance < 0) distance = 0.0;
    return distance;
  }
  // PlaneShell::DistanceToOut, PlaneShell.h:218-245
  double ShellDistanceToOut(V3 point, V3 dir) const {
    double distOut = kInfinity;
    double proj[4], vdist[4];
    for (int i = 0; i < 4; ++i) {
      double pdist = Dist(i, point);
      proj[i] = A[i] * dir.x + B[i] * dir.y + C[i] * dir.z;
      vdist[i] = -pdist / proj[i];
    }
    for (int i = 0; i < 4; ++i)
      if (vdist[i] >= 0. && proj[i] > 0 && vdist[i] < distOut) distOut = vdist[i];
    return distOut;
  }
  // DistanceToOut :294-320
  double DistanceToOut(V3 point, V3 dir, double) const override {
    bool posZ = dir.z > 0.0, negZ = dir.z < 0.0;
    double zsign = negZ ? -1. : 1.;
    double mx = zsign * dz - point.z;
    bool done = (posZ && mx <= MakePlusTolerant<true>(0.));
    done |= (negZ && mx >= MakeMinusTolerant<true>(0.));
    if (done) return 0.;
    double distance = mx / dir.z;
    double dpl = ShellDistanceToOut(point, dir);
    if (dpl < distance) distance = dpl;
    return distance;
  }
  // SafetyToIn :366-390 + PlaneShell::SafetyToIn (PlaneShell.h:249-262)
  double SafetyToIn(V3 p) const override {
    double safety = Abs(p.z) - dz;
    double dist[4];
    for (int i = 0; i < 4; ++i) dist[i] = Dist(i, p);
    for (int i = 0; i < 4; ++i)
      if (dist[i] > safety) safety = dist[i];
    return safety;
  }
  // SafetyToOut :392-420 + PlaneShell::SafetyToOut (PlaneShell.h:266-280)
  double SafetyToOut(V3 p) const override {
    double safety = dz - Abs(p.z);
    if (safety < 0.0) safety = 0.0;
    if (safety == 0.0) return safety;
    double dist[4];
    for (int i = 0; i < 4; ++i) dist[i] = -Dist(i, p);
    for (int i = 0; i < 4; ++i)
      if (dist[i] < safety) safety = dist[i];
    return safety;
  }
  // UnplacedTrapezoid::Extent, source/UnplacedTrapezoid.cpp:428-466
  void Extent(V3 &lo, V3 &hi) const override {
    V3 pt[8];
    Corners(pt);
    auto mx = [](double a, double b) { return std::max(a, b); };
    auto mn = [](double a, double b) { return std::min(a, b); };
    for (int ax = 0; ax < 2; ++ax) {
      double e01 = mx(pt[0][ax], pt[1][ax]), e23 = mx(pt[2][ax], pt[3][ax]);
      double e45 = mx(pt[4][ax], pt[5][ax]), e67 = mx(pt[6][ax], pt[7][ax]);
      double eA = e01 > e23 ? e01 : e23, eB = e45 > e67 ? e45 : e67;
      hi[ax] = (eA > eB) ? eA : eB;
      e01 = mn(pt[0][ax], pt[1][ax]); e23 = mn(pt[2][ax], pt[3][ax]);
      e45 = mn(pt[4][ax], pt[5][ax]); e67 = mn(pt[6][ax], pt[7][ax]);
      eA = e01 < e23 ? e01 : e23; eB = e45 < e67 ? e45 : e67;
      lo[ax] = (eA < eB) ? eA : eB;
    }
    lo.z = -dz;
    hi.z = dz;
  }
};

Shape *MakeTrapezoid(const double *p, int n) { return n == 11 ? new Trapezoid(p) : nullptr; }
Shape *MakeTorus(const double *, int) { return nullptr; }
Shape *MakePolycone(const double *, int) { return nullptr; }
Shape *MakePolyhedron(const double *, int) { return nullptr; }
}  // namespace vgo
