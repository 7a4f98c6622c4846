// ORACLE -- test infrastructure only (see vgo.h).
// Zoo shapes restated from the reference Scalar kernels.
#include "vgo.h"

namespace vgo {

Shape *MakeCone(const double *p, int n);
Shape *MakeSphere(const double *p, int n);
Shape *MakeTrapezoid(const double *p, int n);
Shape *MakeTorus(const double *p, int n);
Shape *MakePolycone(const double *p, int n);
Shape *MakePolyhedron(const double *p, int n);

// ---------------------------------------------------------------- Orb
// volumes/kernel/OrbImplementation.h:234-570; cached values as
// UnplacedOrb::SetRadius computes them (source/UnplacedOrb.cpp:48-56).  The
// constructor (:35-43) reads frTolerance before it is initialised (member
// order, volumes/UnplacedOrb.h:29-36); SetRadius gives the intended
// fRTolerance = Max(1e-9, 2e-11 * r).
struct Orb : Shape {
  double r, rTol, rTolI, rTolO;
  explicit Orb(double r_) : r(r_) {
    rTol = Max(1e-9, 2e-11 * r);
    rTolI = r - rTol;
    rTolO = r + rTol;
  }
  // OrbImplementation::NormalKernel, OrbImplementation.h:208-229: the radial
  // direction, valid within the radial tolerance band.
  int Normal(V3 p, V3 &n) const override {
    const double rad2 = p.Mag2();
    const double rad = Sqrt(rad2);
    n = V3{p.x / rad, p.y / rad, p.z / rad};
    return ((rad2 <= rTolO * rTolO) && (rad2 >= rTolI * rTolI)) ? 1 : 0;
  }
  template <bool ForInside>
  void Kernel(V3 p, bool &cin, bool &cout) const {
    double rad2 = p.Mag2();
    double tolRMax2 = rTolO * rTolO;
    double tolRMin2 = rTolI * rTolI;
    cout = rad2 > tolRMax2;
    if (ForInside) cin = rad2 < tolRMin2;
  }
  bool Contains(V3 p) const override {
    bool cin = false, cout;
    Kernel<false>(p, cin, cout);
    return !cout;
  }
  int Inside(V3 p) const override {
    bool cin = false, cout;
    Kernel<true>(p, cin, cout);
    int res = kSurface;
    if (cout) res = kOutside;
    if (cin) res = kInside;
    return res;
  }
  // DistanceToInKernel :374-423
  double DistanceToIn(V3 p, V3 d, double) const override {
    double rad2 = p.Mag2();
    double rad = Sqrt(rad2);
    double pDotV = Dot(p, d);
    double radius2 = r * r;
    double c = rad2 - radius2;
    double d2 = pDotV * pDotV - c;
    if (rad < rTolI) return kInfinity;
    if ((rad >= rTolI) && (rad <= rTolO) && (pDotV < 0)) return 0.;
    if ((rad >= rTolI) && (rad <= rTolO) && (pDotV >= 0)) return kInfinity;
    if ((rad > rTolO) && (pDotV < 0)) return -pDotV - Sqrt(d2);
    return kInfinity;
  }
  // DistanceToOutKernel :425-480
  double DistanceToOut(V3 p, V3 d, double) const override {
    double rad2 = p.Mag2();
    double rad = Sqrt(rad2);
    double pDotV = Dot(p, d);
    double radius2 = r * r;
    double c = rad2 - radius2;
    double d2 = pDotV * pDotV - c;
    double tolRMax = rTolO;
    double tolRMax2 = tolRMax * tolRMax;
    if (rad2 > tolRMax2) return 0.;
    bool within = (rad <= tolRMax) && (c < (kTolerance * r));
    bool onTolSurface = (c > (-2 * kTolerance * r)) && ((pDotV >= 0) || (d2 < 0));
    if (within && onTolSurface) return 0.;
    if (within && !onTolSurface) return -pDotV + Sqrt(d2);
    return 0.;
  }
  // SafetyToInKernel / SafetyToOutKernel :482-540
  double SafetyToIn(V3 p) const override {
    double safe = Sqrt(p.Mag2()) - r;
    return safe < 0. ? 0. : safe;
  }
  double SafetyToOut(V3 p) const override {
    double safe = r - Sqrt(p.Mag2());
    return safe < 0. ? 0. : safe;
  }
  void Extent(V3 &lo, V3 &hi) const override {
    lo = {-r, -r, -r};
    hi = {r, r, r};
  }
};

// ---------------------------------------------------------------- Trd
// volumes/kernel/TrdImplementation.h:20-458, volumes/UnplacedTrd.h:38-48.
// source/UnplacedTrd.cpp:215-240 always instantiates UniversalTrd
// (GENERATE_TRD_SPECIALIZATIONS is never defined): the inside test takes the
// varying-y branch, safeties check dy1 != dy2 at run time.
struct Trd : Shape {
  double dx1, dx2, dy1, dy2, dz;
  double x2minusx1, y2minusy1, halfx1plusx2, halfy1plusy2, dztimes2, fx, fy, calfx, calfy;
  Trd(double a, double b, double c, double d, double e) : dx1(a), dx2(b), dy1(c), dy2(d), dz(e) {
    x2minusx1 = dx2 - dx1;
    y2minusy1 = dy2 - dy1;
    halfx1plusx2 = 0.5 * (dx1 + dx2);
    halfy1plusy2 = 0.5 * (dy1 + dy2);
    dztimes2 = dz * 2;
    fx = 0.5 * (dx1 - dx2) / dz;
    fy = 0.5 * (dy1 - dy2) / dz;
    calfx = 1. / Sqrt(1.0 + fx * fx);
    calfy = 1. / Sqrt(1.0 + fy * fy);
  }
  // UnplacedTrd::Normal, source/UnplacedTrd.cpp:154-199, as it executes.
  // Each of its six "if (onSurf && FacX) norm[0] = ...; norm[1] = ...;
  // norm[2] = ...; nosurface++;" lines guards only the first assignment, and
  // each Fac is a double product converted to bool (true when nonzero or
  // NaN).  So x takes the x-component of the last face whose Fac is true, y
  // and z those of the -z face, and the result is always valid.
  int Normal(V3 p, V3 &n) const override {
    const double xp = p.x, yp = p.y, zp = p.z;
    const double xx1 = dx1, xx2 = dx2, yy1 = dy1, yy2 = dy2, zz = dz;
    const double XplusX = 2.0 * yy1 * 2.0 * zz, XminusX = XplusX;
    const double XplusY = 0.0, XminusY = 0.0;
    const double XplusZ = 2.0 * yy1 * (xx1 - xx2), XminusZ = 2.0 * yy1 * (-xx1 + xx2);
    const double YplusX = 0.0, YminusX = 0.0;
    const double YplusY = -2.0 * xx1 * 2.0 * zz, YminusY = YplusY;
    const double YplusZ = 2.0 * xx1 * (-yy1 + yy2), YminusZ = 2.0 * xx1 * (yy1 - yy2);
    const double ZplusX = 0.0, ZminusX = 0.0, ZplusY = 0.0, ZminusY = 0.0;
    const double ZplusZ = -2.0 * xx2 * 2.0 * yy2, ZminusZ = 2.0 * xx2 * 2.0 * yy1;
    const bool fpx = (XplusX * (xp - xx2) + XplusY * (yp - yy2) + XplusZ * (zp - zz)) != 0;
    const bool fmx = (XminusX * (xp + xx2) + XminusY * (yp - yy2) + XminusZ * (zp - zz)) != 0;
    const bool fpy = (YplusX * (xp - xx2) + YplusY * (yp - yy2) + YplusZ * (zp - zz)) != 0;
    const bool fmy = (YplusX * (xp - xx2) + YminusY * (yp + yy2) + YminusZ * (zp - zz)) != 0;
    const bool fpz = (ZplusX * (xp - xx2) + ZplusY * (yp - yy2) + ZplusZ * (zp - zz)) != 0;
    const bool fmz = (ZminusX * (xp - xx2) + ZminusY * (yp + yy2) + ZminusZ * (zp - zz)) != 0;
    const bool onSurf = fpx || fmx || fpy || fmy || fpz || fmz;
    n = V3{0, 0, 0};
    if (onSurf && fpx) n.x = XplusX;
    if (onSurf && fmx) n.x = XminusX;
    if (onSurf && fpy) n.x = YplusX;
    if (onSurf && fmy) n.x = YminusX;
    if (onSurf && fpz) n.x = ZplusX;
    if (onSurf && fmz) n.x = ZminusX;
    n.y = ZminusY;
    n.z = ZminusZ;
    return 1;
  }
  // TrdUtilities::PointLineOrientation :40-46
  static double Cross(double px, double py, double vx, double vy) { return vx * py - vy * px; }
  // TrdUtilities::UnplacedInside :173-209
  template <bool surfaceT>
  void Kernel(V3 p, bool &cin, bool &cout) const {
    double pzPlusDz = p.z + dz;
    cout = Abs(p.z) > MakePlusTolerant<surfaceT>(dz);
    if (surfaceT) cin = Abs(p.z) < MakeMinusTolerant<surfaceT>(dz);
    double cross = Cross(Abs(p.x) - dx1, pzPlusDz, x2minusx1, dztimes2);
    cout |= MakePlusTolerant<surfaceT>(cross) < 0;
    if (surfaceT) cin &= MakeMinusTolerant<surfaceT>(cross) > 0;
    cross = Cross(Abs(p.y) - dy1, pzPlusDz, y2minusy1, dztimes2);
    cout |= MakePlusTolerant<surfaceT>(cross) < 0;
    if (surfaceT) cin &= MakeMinusTolerant<surfaceT>(cross) > 0;
  }
  bool Contains(V3 p) const override {
    bool cin = false, cout;
    Kernel<false>(p, cin, cout);
    return !cout;
  }
  // TrdImplementation::Inside :262-277 (outside assigned last)
  int Inside(V3 p) const override {
    bool cin = false, cout;
    Kernel<true>(p, cin, cout);
    int r = kSurface;
    if (cin) r = kInside;
    if (cout) r = kOutside;
    return r;
  }
  // TrdUtilities::FaceTrajectoryIntersection :86-131
  template <bool forY, bool mirrored>
  void FaceHit(V3 pos, V3 dir, double &dist, bool &ok) const {
    double alongV, posV, dirV, posK, dirK, fK, halfKplus, v1;
    if (forY) {
      alongV = y2minusy1; v1 = dy1; posV = pos.y; posK = pos.x; dirV = dir.y; dirK = dir.x; fK = fx;
      halfKplus = halfx1plusx2;
    } else {
      alongV = x2minusx1; v1 = dx1; posV = pos.x; posK = pos.y; dirV = dir.x; dirK = dir.y; fK = fy;
      halfKplus = halfy1plusy2;
    }
    if (mirrored) {
      posV *= -1;
      dirV *= -1;
    }
    double alongZ = dztimes2;
    dist = (alongZ * (posV - v1) - alongV * (pos.z + dz)) / (dir.z * alongV - dirV * alongZ);
    ok = dist > 0;
    if (ok) {
      double hitz = pos.z + dist * dir.z;
      ok &= Abs(hitz) <= dz;
      double hitk = posK + dist * dirK;
      double dK = halfKplus - fK * hitz;
      ok &= Abs(hitk) <= dK;
    }
  }
  // DistanceToIn :280-348
  double DistanceToIn(V3 pos, V3 dir, double) const override {
    double distance = kInfinity;
    bool okz = pos.z * dir.z < 0;
    if (okz) {
      double distz = (Abs(pos.z) - dz) / Abs(dir.z);
      double hitx = Abs(pos.x + distz * dir.x);
      double hity = Abs(pos.y + distz * dir.y);
      bool okzt = pos.z > dz && hitx <= dx2 && hity <= dy2;
      bool okzb = pos.z < -dz && hitx <= dx1 && hity <= dy1;
      okz &= (okzt | okzb);
      if (okz) distance = distz;
    }
    double d;
    bool ok;
    FaceHit<false, false>(pos, dir, d, ok);
    if (ok && d < distance) distance = d;
    FaceHit<false, true>(pos, dir, d, ok);
    if (ok && d < distance) distance = d;
    if (dy1 != dy2) {
      FaceHit<true, false>(pos, dir, d, ok);
      if (ok && d < distance) distance = d;
      FaceHit<true, true>(pos, dir, d, ok);
      if (ok && d < distance) distance = d;
    } else {
      double disty = (Abs(pos.y) - dy1) / Abs(dir.y);
      double zhit = pos.z + disty * dir.z;
      double xhit = pos.x + disty * dir.x;
      double dxw = halfx1plusx2 - fx * zhit;
      bool oky = pos.y * dir.y < 0 && disty > 0 && Abs(xhit) < dxw && Abs(zhit) < dz;
      if (oky && disty < distance) distance = disty;
    }
    return distance;
  }
  // DistanceToOut :350-430 (early returns)
  double DistanceToOut(V3 p, V3 dir, double) const override {
    double distance = kInfinity;
    if (dir.z > 0) {
      double distz = (dz - p.z) / Abs(dir.z);
      double hitx = Abs(p.x + distz * dir.x);
      double hity = Abs(p.y + distz * dir.y);
      if (hitx <= dx2 && hity <= dy2) return distz;
    }
    if (dir.z < 0) {
      double distz = (p.z - (-dz)) / Abs(dir.z);
      double hitx = Abs(p.x + distz * dir.x);
      double hity = Abs(p.y + distz * dir.y);
      if (hitx <= dx1 && hity <= dy1) return distz;
    }
    double d;
    bool ok;
    FaceHit<false, false>(p, dir, d, ok);
    if (ok) return d;
    FaceHit<false, true>(p, dir, d, ok);
    if (ok) return d;
    if (dy1 != dy2) {
      FaceHit<true, false>(p, dir, d, ok);
      if (ok) return d;
      FaceHit<true, true>(p, dir, d, ok);
      if (ok) distance = d;
    } else {
      double plane = dy1;
      if (dir.y < 0) plane = -dy1;
      double disty = (plane - p.y) / dir.y;
      double zhit = p.z + disty * dir.z;
      double xhit = p.x + disty * dir.x;
      double dxw = halfx1plusx2 - fx * zhit;
      if (Abs(xhit) < dxw && Abs(zhit) < dz) distance = disty;
    }
    return distance;
  }
  // TrdUtilities::Safety :133-171
  template <bool inside>
  double Safety(V3 pos) const {
    double dist = dz - Abs(pos.z);
    double distx = halfx1plusx2 - fx * pos.z;
    bool okx = distx >= 0;
    double safx = (distx - Abs(pos.x)) * calfx;
    if (okx && safx < dist) dist = safx;
    if (dy1 != dy2) {
      double disty = halfy1plusy2 - fy * pos.z;
      bool oky = disty >= 0;
      double safy = (disty - Abs(pos.y)) * calfy;
      if (oky && safy < dist) dist = safy;
    } else {
      double safy = dy1 - Abs(pos.y);
      if (safy < dist) dist = safy;
    }
    if (!inside) dist = -dist;
    return dist;
  }
  double SafetyToIn(V3 p) const override { return Safety<false>(p); }
  double SafetyToOut(V3 p) const override { return Safety<true>(p); }
  // UnplacedTrd::Extent, volumes/UnplacedTrd.h:147-150 (max x uses fDY1
  // as in the reference).
  void Extent(V3 &lo, V3 &hi) const override {
    lo = {-Max(dx1, dx2), -Max(dy1, dy2), -dz};
    hi = {Max(dy1, dx2), Max(dy1, dy2), dz};
  }
};

// ---------------------------------------------------------------- Parallelepiped
// volumes/kernel/ParallelepipedImplementation.h:120-356;
// source/UnplacedParallelepiped.cpp:15-70 (angles in degrees).
struct Parallelepiped : Shape {
  double d[3], tanAlpha, tanThetaSinPhi, tanThetaCosPhi;
  double alpha, theta, phi;
  Parallelepiped(double x, double y, double z, double a, double t, double ph) : d{x, y, z} {
    alpha = a;
    theta = t;
    phi = ph;
    tanAlpha = std::tan(kDegToRad * a);
    tanThetaCosPhi = std::tan(kDegToRad * t) * std::cos(kDegToRad * ph);
    tanThetaSinPhi = std::tan(kDegToRad * t) * std::sin(kDegToRad * ph);
  }
  // ParallelepipedImplementation::Transform :120-127
  V3 Shear(V3 p) const {
    p.y -= tanThetaSinPhi * p.z;
    p.x -= tanThetaCosPhi * p.z + tanAlpha * p.y;
    return p;
  }
  // Box kernels on the sheared point (BoxImplementation.h:593-746)
  template <bool ForInside>
  void BoxKernel(V3 p, bool &cin, bool &cout) const {
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
    BoxKernel<false>(Shear(p), cin, cout);
    return !cout;
  }
  int Inside(V3 p) const override {
    bool cin = false, cout;
    BoxKernel<true>(Shear(p), cin, cout);
    int r = kSurface;
    if (cout) r = kOutside;
    if (cin) r = kInside;
    return r;
  }
  double DistanceToIn(V3 p0, V3 d0, double stepMax) const override {
    V3 p = Shear(p0), dir = Shear(d0);
    double s[3] = {Abs(p.x) - d[0], Abs(p.y) - d[1], Abs(p.z) - d[2]};
    if (s[0] >= stepMax || s[1] >= stepMax || s[2] >= stepMax) return kInfinity;
    if (s[0] < 0 && s[1] < 0 && s[2] < 0) return -1.;
    for (int a = 0; a < 3; ++a) {
      int b = a == 0 ? 1 : 0;
      int c = a == 2 ? 1 : 2;
      double next = s[a] / Abs(dir[a] + kMinimum);
      double c1 = p[b] + next * dir[b];
      double c2 = p[c] + next * dir[c];
      if (s[a] >= MakeMinusTolerant<true>(0.) && p[a] * dir[a] < 0 && Abs(c1) <= d[b] && Abs(c2) <= d[c]) return next;
    }
    return kInfinity;
  }
  // DistanceToOut :208-300
  double DistanceToOut(V3 p, V3 dir, double) const override {
    double distance = kInfinity;
    bool inDir = dir.z > 0;
    double mx = d[2] - p.z;
    bool inPoint = mx > kHalfTolerance;
    bool away = inDir && !inPoint;
    if (away) distance = 0.;
    if (inPoint && inDir) distance = mx / dir.z;
    if (away) return distance;
    inDir = dir.z < 0;
    mx = -d[2] - p.z;
    inPoint = mx < -kHalfTolerance;
    away = inDir && !inPoint;
    if (away) distance = 0.;
    if (inPoint && inDir) distance = mx / dir.z;
    if (away) return distance;
    double lpy = p.y - tanThetaSinPhi * p.z;
    double ldy = dir.y - tanThetaSinPhi * dir.z;
    inDir = ldy > 0;
    mx = d[1] - lpy;
    inPoint = mx > kHalfTolerance;
    away = inDir && !inPoint;
    mx /= ldy;
    if (away) distance = 0.;
    if (inPoint && inDir && mx < distance) distance = mx;
    if (away) return distance;
    inDir = ldy < 0;
    mx = -d[1] - lpy;
    inPoint = mx < -kHalfTolerance;
    away = inDir && !inPoint;
    mx /= ldy;
    if (away) distance = 0.;
    if (inPoint && inDir && mx < distance) distance = mx;
    if (away) return distance;
    double lpx = p.x - tanThetaCosPhi * p.z - tanAlpha * lpy;
    double ldx = dir.x - tanThetaCosPhi * dir.z - tanAlpha * ldy;
    inDir = ldx > 0;
    mx = d[0] - lpx;
    inPoint = mx > kHalfTolerance;
    away = inDir && !inPoint;
    mx /= ldx;
    if (away) distance = 0.;
    if (inPoint && inDir && mx < distance) distance = mx;
    if (away) return distance;
    inDir = ldx < 0;
    mx = -d[0] - lpx;
    inPoint = mx < -kHalfTolerance;
    away = inDir && !inPoint;
    mx /= ldx;
    if (away) distance = 0.;
    if (inPoint && inDir && mx < distance) distance = mx;
    return distance;
  }
  // SafetyToIn / SafetyToOut :302-356
  double SafetyToIn(V3 p0) const override {
    V3 p = Shear(p0);
    double sx = Abs(p.x) - d[0], sy = Abs(p.y) - d[1], sz = Abs(p.z) - d[2];
    double ctx = 1.0 / Sqrt(1. + tanAlpha * tanAlpha + tanThetaCosPhi * tanThetaCosPhi);
    double cty = 1.0 / Sqrt(1. + tanThetaSinPhi * tanThetaSinPhi);
    sx *= ctx;
    sy *= cty;
    double s = sx;
    if (sy > s) s = sy;
    if (sz > s) s = sz;
    return s;
  }
  double SafetyToOut(V3 p0) const override {
    V3 p = Shear(p0);
    double sx = d[0] - Abs(p.x), sy = d[1] - Abs(p.y), sz = d[2] - Abs(p.z);
    double ctx = 1.0 / Sqrt(1. + tanAlpha * tanAlpha + tanThetaCosPhi * tanThetaCosPhi);
    double cty = 1.0 / Sqrt(1. + tanThetaSinPhi * tanThetaSinPhi);
    sx *= ctx;
    sy *= cty;
    double s = sx;
    if (sy < s) s = sy;
    if (sz < s) s = sz;
    return s;
  }
  // The reference has no Extent here (PlacedVolume.h:260-263 asserts); the
  // bound of the 8 sheared corners is the convention shared with the shim.
  void Extent(V3 &lo, V3 &hi) const override {
    double ta = std::tan(kDegToRad * alpha), tt = std::tan(kDegToRad * theta);
    double ex = d[0] + std::fabs(d[1] * ta) + std::fabs(d[2] * tt * std::cos(kDegToRad * phi));
    double ey = d[1] + std::fabs(d[2] * tt * std::sin(kDegToRad * phi));
    lo = {-ex, -ey, -d[2]};
    hi = {ex, ey, d[2]};
  }
};

// ---------------------------------------------------------------- Paraboloid
// volumes/kernel/ParaboloidImplementation.h:70-580;
// source/UnplacedParaboloid.cpp:92-127, 242-245.
struct Paraboloid : Shape {
  double rlo, rhi, dz, a, b, ainv, binv, a2, b2, rlo2, rhi2, tolIz, tolOz, tolIrlo2, tolOrlo2, tolIrhi2, tolOrhi2, dx,
      dy;
  Paraboloid(double lo, double hi, double z) {
    rlo = lo;
    rhi = hi;
    rlo2 = rlo * rlo;
    rhi2 = rhi * rhi;
    dz = z;
    double dd = 1. / (rhi2 - rlo2);
    a = 2. * dz * dd;
    b = -dz * (rlo2 + rhi2) * dd;
    ainv = 1 / a;
    binv = 1 / b;
    a2 = a * a;
    b2 = b * b;
    tolIz = dz - kHalfTolerance;
    tolOz = dz + kHalfTolerance;
    tolIrlo2 = (rlo - kHalfTolerance) * (rlo - kHalfTolerance);
    tolOrlo2 = (rlo + kHalfTolerance) * (rlo + kHalfTolerance);
    tolIrhi2 = (rhi - kHalfTolerance) * (rhi - kHalfTolerance);
    tolOrhi2 = (rhi + kHalfTolerance) * (rhi + kHalfTolerance);
    dx = Max(rhi, rlo);
    dy = dx;
  }
  // UnplacedInside :85-133 (inner z test uses the outer tolerance, as in
  // the reference)
  int Inside(V3 p) const override {
    double rho2 = p.x * p.x + p.y * p.y;
    double absZ = Abs(p.z);
    if (absZ > tolOz) return kOutside;
    double value = a * rho2 + b - p.z;
    if (value > kHalfTolerance) return kOutside;
    bool isInside = (absZ < tolOz) && (value < -kHalfTolerance);
    if (isInside) return kInside;
    return kSurface;
  }
  // UnplacedContains :136-160
  bool Contains(V3 p) const override {
    if (Abs(p.z) > dz) return false;
    double aa = a * (p.z - b);
    double rho2 = p.x * p.x + p.y * p.y;
    bool out = aa < 0 || aa < a2 * rho2;
    return !out;
  }
  // DistanceToIn :205-297
  double DistanceToIn(V3 lp, V3 ld, double) const override {
    double distance = kInfinity;
    double absZ = Abs(lp.z);
    double absDirZ = Abs(ld.z);
    double rho2 = lp.x * lp.x + lp.y * lp.y;
    double pdx = lp.x * ld.x;
    double pdy = lp.y * ld.y;
    bool checkZ = lp.z * ld.z > 0;
    if (absZ > dz && checkZ) return distance;
    if ((rho2 > rhi2) && (pdx > 0 && pdy > 0)) return distance;
    if ((Abs(lp.x) > rhi) && (pdx > 0)) return distance;
    if ((Abs(lp.y) > rhi) && (pdy > 0)) return distance;
    double distZ = (absZ - dz) / absDirZ;
    double xHit = lp.x + distZ * ld.x;
    double yHit = lp.y + distZ * ld.y;
    double rhoHit2 = xHit * xHit + yHit * yHit;
    double ray2 = rhi2;
    if (lp.z < -dz && ld.z > 0) ray2 = rlo2;
    bool crossing = (absZ > dz) && (!checkZ) && (rhoHit2 <= ray2);
    if (crossing) return distZ;
    double dirRho2 = ld.x * ld.x + ld.y * ld.y;
    double qa = a * dirRho2;
    double qb = 2. * a * (pdx + pdy) - ld.z;
    double qc = a * rho2 + b - lp.z;
    double qainv = 1. / qa;
    double t = qb * 0.5;
    double prod = qc * qa;
    double delta = t * t - prod;
    if (delta < 0) return distance;
    delta = Sqrt(delta);
    double distParab = qainv * (-t - delta);
    double zHit = lp.z + distParab * ld.z;
    bool hits = (distParab > 1E20) || (Abs(zHit) <= dz);
    if (hits && distParab > 0) distance = distParab;
    return distance;
  }
  // DistanceToOut :299-346
  double DistanceToOut(V3 p, V3 dir, double) const override {
    double distZ = kInfinity;
    double dirZinv = 1 / dir.z;
    bool dirMask = dir.z < 0;
    if (dirMask) distZ = -(dz + p.z) * dirZinv;
    if (!dirMask) distZ = (dz - p.z) * dirZinv;
    double distParab = kInfinity;
    double rho2 = p.x * p.x + p.y * p.y;
    double dirRho2 = dir.x * dir.x + dir.y * dir.y;
    double qa = a * dirRho2;
    double qb = 2. * a * (p.x * dir.x + p.y * dir.y) - dir.z;
    double qc = a * rho2 + b - p.z;
    double qainv = 1. / qa;
    double t = qb * 0.5;
    double prod = qc * qa;
    double delta = t * t - prod;
    delta = Sqrt(delta);
    double sign = 1.;
    if (qainv < 0) sign = -1.;
    double d1 = qainv * (-t - sign * delta);
    double d2 = qainv * (-t + sign * delta);
    if (d1 > 0) distParab = d1;
    if (d1 < 0 && d2 > 0) distParab = d2;
    if (d1 < 0 && d2 < 0) distParab = kInfinity;
    return Min(distParab, distZ);
  }
  // SafetyToIn, default bounding-cylinder branch :400-425
  double SafetyToIn(V3 p) const override {
    double safety = 0.;
    double absZ = Abs(p.z);
    double safeZ = absZ - dz;
    double r = Sqrt(p.x * p.x + p.y * p.y);
    double safeRhi = r - rhi;
    bool maskBc = (safeZ > 0) || (safeRhi > 0);
    double st = Max(safeZ, safeRhi);
    if (maskBc) {
      safety = st;
      return safety;
    }
    double r0sq = (p.z - b) * ainv;
    double dr = r - Sqrt(r0sq);
    double talf = -2. * a * Sqrt(r0sq);
    double salf = talf / Sqrt(1. + talf * talf);
    double safR = Abs(dr * salf);
    return Max(safR, safeZ);
  }
  // SafetyToOut, fast branch :452-470
  double SafetyToOut(V3 p) const override {
    double absZ = Abs(p.z);
    double safZ = dz - absZ;
    double safety = 0.;
    double r = Sqrt(p.x * p.x + p.y * p.y);
    double safRlo = rlo - r;
    double st = Min(safZ, safRlo);
    if (st > 0) safety = st;
    return safety;
  }
  void Extent(V3 &lo, V3 &hi) const override {
    lo = {-dx, -dy, -dz};
    hi = {dx, dy, dz};
  }
};

Shape *MakeZooShape(int type, const double *p, int n) {
  switch (type) {
    case kOrb: return n == 1 ? new Orb(p[0]) : nullptr;
    case kTrd: return n == 5 ? new Trd(p[0], p[1], p[2], p[3], p[4]) : nullptr;
    case kParallelepiped: return n == 6 ? new Parallelepiped(p[0], p[1], p[2], p[3], p[4], p[5]) : nullptr;
    case kParaboloid: return n == 3 ? new Paraboloid(p[0], p[1], p[2]) : nullptr;
    case kCone: return MakeCone(p, n);
    case kSphere: return MakeSphere(p, n);
    case kTrapezoid: return MakeTrapezoid(p, n);
    case kTorus: return MakeTorus(p, n);
    case kPolycone: return MakePolycone(p, n);
    case kPolyhedron: return MakePolyhedron(p, n);
  }
  return nullptr;
}

}  // namespace vgo
