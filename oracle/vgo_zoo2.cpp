// ORACLE -- test infrastructure only (see vgo.h).
// Sphere, Cone, Trapezoid, Torus, Polycone, Polyhedron restatements.
#include "vgo.h"

namespace vgo {

// ---------------------------------------------------------------- Sphere
// volumes/kernel/SphereImplementation.h:257-1520, volumes/UnplacedSphere.h:
// 63-215 (angle checks, cached trigonometry), source/UnplacedSphere.cpp:25-95.

struct Sphere : Shape {
  double fRmin, fRmax, fSPhi = 0, fDPhi = 0, fSTheta = 0, fDTheta = 0;
  double fRminTolerance, mkTolerance, kAngTol = 1e-9;
  double sinCPhi, cosCPhi, cosHDPhiOT, cosHDPhiIT, sinSPhi, cosSPhi, sinEPhi, cosEPhi, hDPhi, cPhi, ePhi;
  double sinSTheta, cosSTheta, sinETheta, cosETheta, tanSTheta, tanSTheta2, tanETheta, tanETheta2, eTheta;
  bool fullPhi = true, fullTheta = true;
  Sphere(const double *p) {
    fRmin = p[0];
    fRmax = p[1];
    fRminTolerance = (fRmin) ? Max(1e-9, 2.e-11 * fRmin) : 0;
    mkTolerance = Max(1e-9, 2.e-11 * fRmax);
    // CheckPhiAngles
    fullPhi = true;
    if (p[3] >= 2 * kPi - kAngTol * 0.5) {
      fDPhi = 2 * kPi;
      fSPhi = 0;
    } else {
      fullPhi = false;
      if (p[3] > 0) fDPhi = p[3];
    }
    if (!fullPhi) {
      double sPhi = p[2];
      if (sPhi < 0)
        fSPhi = 2 * kPi - std::fmod(std::fabs(sPhi), 2 * kPi);
      else
        fSPhi = std::fmod(sPhi, 2 * kPi);
      if (fSPhi + fDPhi > 2 * kPi) fSPhi -= 2 * kPi;
    }
    hDPhi = 0.5 * fDPhi;
    cPhi = fSPhi + hDPhi;
    ePhi = fSPhi + fDPhi;
    sinCPhi = std::sin(cPhi);
    cosCPhi = std::cos(cPhi);
    cosHDPhiIT = std::cos(hDPhi - 0.5 * kAngTol);
    cosHDPhiOT = std::cos(hDPhi + 0.5 * kAngTol);
    sinSPhi = std::sin(fSPhi);
    cosSPhi = std::cos(fSPhi);
    sinEPhi = std::sin(ePhi);
    cosEPhi = std::cos(ePhi);
    // CheckThetaAngles
    double sTheta = p[4], dTheta = p[5];
    if (!((sTheta < 0) || (sTheta > kPi))) fSTheta = sTheta;
    if (dTheta + sTheta >= kPi)
      fDTheta = kPi - sTheta;
    else if (dTheta > 0)
      fDTheta = dTheta;
    fullTheta = !(fDTheta - fSTheta < kPi);
    eTheta = fSTheta + fDTheta;
    sinSTheta = std::sin(fSTheta);
    cosSTheta = std::cos(fSTheta);
    sinETheta = std::sin(eTheta);
    cosETheta = std::cos(eTheta);
    tanSTheta = std::tan(fSTheta);
    tanSTheta2 = tanSTheta * tanSTheta;
    tanETheta = std::tan(eTheta);
    tanETheta2 = tanETheta * tanETheta;
  }

  // The reference's accessors (volumes/UnplacedSphere.h:262-270) return
  // swapped flags: IsFullSphere() is the phi flag, IsFullPhiSphere() the
  // flag for both.  The normal kernels call them by these names.
  bool IsFullSphere() const { return fullPhi; }
  bool IsFullPhiSphere() const { return fullPhi && fullTheta; }
  bool IsFullThetaSphere() const { return fullTheta; }

  // SphereImplementation::ApproxSurfaceNormalKernel,
  // SphereImplementation.h:463-610: the normal of the nearest surface for a
  // point on none (r shells, then phi planes, then theta cones; each later
  // candidate must be strictly nearer).
  V3 ApproxNormal(V3 p) const {
    const double kNRMin = 0., kNRMax = 1., kNSPhi = 2., kNEPhi = 3., kNSTheta = 4., kNETheta = 5.;
    double side = 10.;
    const double radius = Sqrt(p.Mag2());
    const double rho = Sqrt(p.x * p.x + p.y * p.y);
    double pPhi = ATan2(p.y, p.x);
    const double pTheta = std::atan2(rho, p.z);
    double distRMax, distRMin, distSPhi = 0., distEPhi = 0., distSTheta = 0., distETheta = 0., distMin = 0.;
    double temp = radius - fRmax;
    if (temp < 0.) temp = -1. * temp;
    distRMax = temp;
    temp = radius - fRmin;
    if (temp < 0.) temp = -1. * temp;
    distRMin = temp;
    if (fRmin > 0. && distRMin < distRMax) {
      distMin = distRMin;
      side = kNRMin;
    }
    if (fRmin > 0. && !(distRMin < distRMax)) {
      distMin = distRMax;
      side = kNRMax;
    }
    if (!(fRmin > 0.)) {
      distMin = distRMax;
      side = kNRMax;
    }
    if (pPhi < 0.) pPhi = pPhi + (2 * kPi);
    const bool phiCut = !IsFullPhiSphere() && rho > 0.;
    temp = pPhi - (fSPhi + 2 * kPi);
    if (temp < 0.) temp = -1. * temp;
    if (phiCut && fSPhi < 0.) distSPhi = temp * rho;
    temp = pPhi - fSPhi;
    if (temp < 0.) temp = -1. * temp;
    if (phiCut && !(fSPhi < 0.)) distSPhi = temp * rho;
    temp = pPhi - fSPhi - fDPhi;
    if (temp < 0.) temp = -1. * temp;
    if (phiCut) distEPhi = temp * rho;
    double prev = distMin;
    if (phiCut && distSPhi < distEPhi && distSPhi < distMin) distMin = distSPhi;
    if (phiCut && distSPhi < distEPhi && distSPhi < prev) side = kNSPhi;
    prev = distMin;
    if (phiCut && !(distSPhi < distEPhi) && distEPhi < distMin) distMin = distEPhi;
    if (phiCut && !(distSPhi < distEPhi) && distEPhi < prev) side = kNEPhi;
    const bool thetaCut = !IsFullThetaSphere() && radius > 0.;
    temp = pTheta - fSTheta;
    if (temp < 0.) temp = -1. * temp;
    if (thetaCut) distSTheta = temp * radius;
    temp = pTheta - fSTheta - fDTheta;
    if (temp < 0.) temp = -1. * temp;
    if (thetaCut) distETheta = temp * radius;
    prev = distMin;
    if (thetaCut && distSTheta < distETheta && distSTheta < distMin) distMin = distSTheta;
    if (thetaCut && distSTheta < distETheta && distSTheta < prev) side = kNSTheta;
    prev = distMin;
    if (thetaCut && !(distSTheta < distETheta) && distETheta < distMin) distMin = distETheta;
    if (thetaCut && !(distSTheta < distETheta) && distETheta < prev) side = kNETheta;
    if (side == kNRMin) return V3{-p.x / radius, -p.y / radius, -p.z / radius};
    if (side == kNRMax) return V3{p.x / radius, p.y / radius, p.z / radius};
    if (side == kNSPhi) return V3{sinSPhi, -cosSPhi, 0};
    if (side == kNEPhi) return V3{-sinEPhi, cosEPhi, 0};
    if (side == kNSTheta) return V3{-cosSTheta * std::cos(pPhi), -cosSTheta * std::sin(pPhi), sinSTheta};
    if (side == kNETheta) return V3{cosETheta * std::cos(pPhi), cosETheta * std::sin(pPhi), sinETheta};
    return V3{0, 0, 0};
  }

  // SphereImplementation::NormalKernel, SphereImplementation.h:619-760: sum
  // of the normals of every surface within tolerance (r shells 0.5e-9 mm,
  // phi planes and theta cones half the angular tolerance), unit length when
  // there are several; the approximate normal and invalid when there is none.
  int Normal(V3 p, V3 &n) const override {
    const double radius = Sqrt(p.Mag2());
    const double rho = Sqrt(p.x * p.x + p.y * p.y);
    double pPhi = ATan2(p.y, p.x);
    const double pTheta = std::atan2(rho, p.z);
    const double halfAng = 0.5 * kAngTol;
    double distSPhi = kInfinity, distSTheta = kInfinity, distEPhi = kInfinity, distETheta = kInfinity;
    V3 nR{0, 0, 0}, nPs{0, 0, 0}, nPe{0, 0, 0}, nTs{0, 0, 0}, nTe{0, 0, 0};
    const V3 nZ{0., 0., 1.};
    V3 norm{0, 0, 0}, sumnorm{0., 0., 0.};
    const bool fullPhiSphere = IsFullPhiSphere();  // the local of that name in the reference
    double temp = radius - fRmax;
    if (temp < 0.) temp = -1. * temp;
    const double distRMax = temp;
    double distRMin = 0;
    temp = radius - fRmin;
    if (temp < 0.) temp = -1. * temp;
    if (fRmin > 0) distRMin = temp;
    if (rho > 0. && !IsFullSphere() && pPhi < fSPhi - halfAng) pPhi = pPhi + (2 * kPi);
    if (rho > 0. && !IsFullSphere() && pPhi > ePhi + halfAng) pPhi = pPhi - (2 * kPi);
    temp = pPhi - fSPhi;
    if (temp < 0.) temp = -1. * temp;
    if (!IsFullPhiSphere() && rho > 0.) distSPhi = temp;
    temp = pPhi - ePhi;
    if (temp < 0.) temp = -1. * temp;
    if (!IsFullPhiSphere() && rho > 0.) distEPhi = temp;
    if (!IsFullPhiSphere() && !fRmin) distSPhi = 0.;
    if (!IsFullPhiSphere() && !fRmin) distEPhi = 0.;
    if (!IsFullPhiSphere()) {
      nPs = V3{sinSPhi, -cosSPhi, 0};
      nPe = V3{-sinEPhi, cosEPhi, 0};
    }
    temp = pTheta - fSTheta;
    if (temp < 0.) temp = -1. * temp;
    if (!IsFullThetaSphere() && rho > 0.) distSTheta = temp;
    temp = pTheta - eTheta;
    if (temp < 0.) temp = -1. * temp;
    if (!IsFullThetaSphere() && rho > 0.) distETheta = temp;
    if (!IsFullThetaSphere() && rho > 0.) nTs = V3{-cosSTheta * p.x / rho, -cosSTheta * p.y / rho, sinSTheta};
    if (!IsFullThetaSphere() && rho != 0.) nTe = V3{cosETheta * p.x / rho, cosETheta * p.y / rho, -sinETheta};
    if (!IsFullThetaSphere() && !fRmin && fSTheta != 0.) {
      distSTheta = 0.;
      nTs = V3{0., 0., -1.};
    }
    if (!IsFullThetaSphere() && !fRmin && eTheta < kPi) {
      distETheta = 0.;
      nTe = V3{0., 0., 1.};
    }
    if (radius != 0.) nR = V3{p.x / radius, p.y / radius, p.z / radius};
    double noSurfaces = 0;
    const double halfCar = 0.5 * 1e-9;
    if (distRMax <= halfCar) {
      noSurfaces = noSurfaces + 1;
      sumnorm = sumnorm + nR;
    }
    if (fRmin != 0. && distRMin <= halfCar) {
      noSurfaces = noSurfaces + 1;
      sumnorm = sumnorm - nR;
    }
    if (!IsFullPhiSphere() && distSPhi <= halfAng) {
      noSurfaces = noSurfaces + 1;
      sumnorm = sumnorm + nPs;
    }
    if (!IsFullPhiSphere() && distEPhi <= halfAng) {
      noSurfaces = noSurfaces + 1;
      sumnorm = sumnorm + nPe;
    }
    if (!IsFullThetaSphere() && distSTheta <= halfAng && fSTheta > 0.) {
      noSurfaces = noSurfaces + 1;
      if (radius <= halfCar && fullPhiSphere)
        sumnorm = sumnorm + nZ;
      else
        sumnorm = sumnorm + nTs;
    }
    if (!IsFullThetaSphere() && distETheta <= halfAng && eTheta < kPi) {
      noSurfaces = noSurfaces + 1;
      if (radius <= halfCar && fullPhiSphere)
        sumnorm = sumnorm - nZ;
      else
        sumnorm = sumnorm + nTe;
      if (sumnorm.z == 0.) sumnorm = sumnorm + nZ;
    }
    if (noSurfaces == 0) norm = ApproxNormal(p);
    if (noSurfaces == 1) norm = sumnorm;
    if (!(noSurfaces == 1) && noSurfaces != 0) {
      const double mag = Sqrt(sumnorm.Mag2());
      norm = V3{sumnorm.x * 1. / mag, sumnorm.y * 1. / mag, sumnorm.z * 1. / mag};
    }
    n = norm;
    return noSurfaces > 0 ? 1 : 0;
  }

  // GenericKernelForContainsAndInside :801-893 (the body does not depend on
  // ForInside)
  void Kernel(V3 p, bool &cin, bool &cout) const {
    double half = 0.5 * kAngTol;
    double rad2 = p.Mag2();
    double tolRMin = fRmin + (0.5 * fRminTolerance);
    double tolRMax = fRmax - (0.5 * fRminTolerance);
    cin = (rad2 <= tolRMax * tolRMax) && (rad2 >= tolRMin * tolRMin);
    tolRMin = fRmin - (0.5 * fRminTolerance);
    tolRMax = fRmax + (0.5 * fRminTolerance);
    cout = (rad2 <= tolRMin * tolRMin) || (rad2 >= tolRMax * tolRMax);
    if (cout) return;
    double pPhi = ATan2(p.y, p.x);
    double eph = fSPhi + fDPhi;
    if (pPhi < (fSPhi - half)) pPhi = pPhi + (2. * kPi);
    if (pPhi > (eph + half)) pPhi = pPhi - (2. * kPi);
    double tolAngMin = fSPhi + half, tolAngMax = eph - half;
    if (!fullPhi) {
      cin &= (pPhi <= tolAngMax) && (pPhi >= tolAngMin);
      tolAngMin = fSPhi - half;
      tolAngMax = eph + half;
      cout |= (pPhi < tolAngMin) || (pPhi > tolAngMax);
      if (cout) return;
    }
    double pTheta = ATan2(Sqrt(p.x * p.x + p.y * p.y), p.z);
    double eth = fSTheta + fDTheta;
    tolAngMin = fSTheta + half;
    tolAngMax = eth - half;
    if (!fullTheta) {
      cin &= (pTheta <= tolAngMax) && (pTheta >= tolAngMin);
      tolAngMin = fSTheta - half;
      tolAngMax = eth + half;
      cout |= (pTheta < tolAngMin) || (pTheta > tolAngMax);
    }
  }
  bool Contains(V3 p) const override {
    bool cin, cout;
    Kernel(p, cin, cout);
    return !cout;
  }
  int Inside(V3 p) const override {
    bool cin, cout;
    Kernel(p, cin, cout);
    int r = kSurface;
    if (cout) r = kOutside;
    if (cin) r = kInside;
    return r;
  }
  // The reference's DistanceToIn is an empty body (:85-94): the helper's
  // initial kInfinity is returned.
  double DistanceToIn(V3, V3, double) const override { return kInfinity; }
  // SafetyToInKernel :945-1040
  double SafetyToIn(V3 p) const override {
    double safety = kInfinity;
    double rad = Sqrt(p.Mag2());
    double rho2 = p.x * p.x + p.y * p.y;
    double rho = std::sqrt(rho2);
    double safeRMin = fRmin - rad, safeRMax = rad - fRmax;
    if ((fRmin > 0) && (safeRMin > safeRMax)) safety = safeRMin;
    if ((fRmin > 0) && !(safeRMin > safeRMax)) safety = safeRMax;
    if (!(fRmin > 0)) safety = safeRMax;
    double safePhi = 0;
    double cosPsi = (p.x * cosCPhi + p.y * sinCPhi) / rho;
    if (!fullPhi && (rho > 0)) {
      double test1 = (p.x * sinSPhi - p.y * cosSPhi);
      if (test1 < 0) test1 = -1. * test1;
      double test2 = (p.x * sinEPhi - p.y * cosEPhi);
      if (test2 < 0) test2 = -1. * test2;
      bool c = cosPsi < std::cos(hDPhi);
      bool side = (p.y * cosCPhi - p.x * sinCPhi) <= 0;
      if (c && side) safePhi = test1;
      if (c && !side) safePhi = test2;
      if (c && (safePhi > safety)) safety = safePhi;
    }
    double rds = p.Mag();
    double pTheta = (kPi / 2) - std::asin(p.z / rds);
    if (pTheta < 0) pTheta = pTheta + kPi;
    double dTheta1 = fSTheta - pTheta, dTheta2 = pTheta - eTheta;
    double sinDTheta1 = std::sin(dTheta1), sinDTheta2 = std::sin(dTheta2);
    double safeTheta = 0;
    if (!fullTheta && (rds != 0)) {
      bool a = dTheta1 > dTheta2;
      if (a && (dTheta1 >= 0)) safeTheta = rds * sinDTheta1;
      if (a && (dTheta1 >= 0) && (safety <= safeTheta)) safety = safeTheta;
      if (!a && (dTheta2 >= 0)) safeTheta = rds * sinDTheta2;
      if (!a && (dTheta2 >= 0) && (safety <= safeTheta)) safety = safeTheta;
    }
    if (safety < 0) safety = 0;
    return safety;
  }
  // SafetyToOutKernel :1053-1120
  double SafetyToOut(V3 p) const override {
    double safety = kInfinity;
    double rad = Sqrt(p.Mag2());
    double rho = Sqrt(p.x * p.x + p.y * p.y);
    double safeRMin = rad - fRmin, safeRMax = fRmax - rad;
    if ((fRmin > 0) && (safeRMin < safeRMax)) safety = safeRMin;
    if ((fRmin > 0) && !(safeRMin < safeRMax)) safety = safeRMax;
    if (!(fRmin > 0)) safety = safeRMax;
    double safePhi = 0;
    if (!fullPhi && (rho > 0)) {
      double test1 = (p.y * cosCPhi - p.x * sinCPhi);
      if (test1 <= 0) safePhi = -1. * (p.x * sinSPhi - p.y * cosSPhi);
      if (!(test1 <= 0)) safePhi = (p.x * sinEPhi - p.y * cosEPhi);
      if (safePhi < safety) safety = safePhi;
    }
    double rds = p.Mag();
    double pTheta = (kPi / 2) - std::asin(p.z / rds);
    if (pTheta < 0) pTheta = pTheta + kPi;
    double eth = fSTheta + fDTheta;
    double dTheta1 = pTheta - fSTheta, dTheta2 = eth - pTheta;
    double sinDTheta1 = std::sin(dTheta1), sinDTheta2 = std::sin(dTheta2);
    double safeTheta = 0;
    if (!fullTheta && (rds > 0)) {
      if (dTheta1 < dTheta2) safeTheta = rds * sinDTheta1;
      if (!(dTheta1 < dTheta2)) safeTheta = rds * sinDTheta2;
      if (safety > safeTheta) safety = safeTheta;
    }
    if (safety < 0) safety = 0;
    return safety;
  }
  // DistanceToSphere :384-446
  double ToSphere(V3 p, V3 d, double radius, bool check, bool firstcross) const {
    double rad2 = p.Mag2();
    double pDotV = Dot(p, d);
    double c = rad2 - radius * radius;
    double d2 = pDotV * pDotV - c;
    double dd = Sqrt(d2);
    if (d2 <= 0) return kInfinity;
    double distance;
    bool in = (c <= 0);
    if (in) distance = -1. * pDotV + dd;
    else if (firstcross) distance = -1. * pDotV - dd;
    else distance = -1. * pDotV + dd;
    if (distance < 0) distance = kInfinity;
    V3 pt = p + distance * d;
    bool cin, cout;
    Kernel(pt, cin, cout);
    if (check && cout) distance = kInfinity;
    return distance;
  }
  // DistanceToCone :314-382
  void ToCone(V3 p, V3 d, double dzv, double r1v, double r2v, double &bv, double &deltav) const {
    deltav = -1.;
    if (dzv < 0) return;
    double ro0v = (0.5 * (r1v + r2v));
    double tzv = (0.5 * (r2v - r1v) / dzv);
    double rsqv = p.x * p.x + p.y * p.y;
    double rcv = ro0v + p.z * tzv;
    double av = d.x * d.x + d.y * d.y - tzv * tzv * d.z * d.z;
    bv = p.x * d.x + p.y * d.y - tzv * rcv * d.z;
    double cv = rsqv - rcv * rcv;
    double t1 = av, t2 = bv;
    if (t1 < 0) t1 = -1. * t1;
    if (t2 < 0) t2 = -1. * t2;
    const double toler = 1.e-10;
    if ((t1 < toler) && (t2 < toler)) return;
    if (t1 < toler) {
      bv = 0.5 * (cv / bv);
      deltav = 0;
      return;
    }
    av = (1. / av);
    bv *= av;
    cv *= av;
    deltav = bv * bv - cv;
    if (deltav > 0)
      deltav = Sqrt(deltav);
    else
      deltav = -1.;
  }
  // DistanceToPhiMin :257-312 (in = true)
  double ToPhiMin(V3 p, V3 d, double s1, double c1, double s2, double c2, double sm, double cm) const {
    double sfi1 = kInfinity, sfi2 = kInfinity, s = 0;
    double un = d.x * s1 - d.y * c1;
    if (un > 0) s = (-1. * p.x * s1 + p.y * c1);
    if ((un > 0) && (s >= 0)) s = (s / un);
    if ((un > 0) && (s >= 0) && (((p.x + s * d.x) * sm - (p.y + s * d.y) * cm) >= 0)) sfi1 = s;
    un = -1. * d.x * s2 + d.y * c2;
    if (un > 0) s = (p.x * s2 - p.y * c2);
    if ((un > 0) && (s >= 0)) s = (s / un);
    if ((un > 0) && (s >= 0) && ((-1. * (p.x + s * d.x) * sm + (p.y + s * d.y) * cm) >= 0)) sfi2 = s;
    return Min(sfi1, sfi2);
  }
  // One theta cone block of DistanceToOutKernel :1336-1478; returns false
  // when the kernel returns 0 early (on the cone surface, moving out).
  bool ThetaBlock(V3 p, V3 d, double rxy2, bool cond, double si, double ci, bool first, double fPhi1,
                  double fPhi2, double &b, double &delta, double &snxt, double &sn) const {
    const double toler = 1.E-10, toler2 = 1.E+10;
    double r1 = 0, r2 = 0, z1 = 0, z2 = 0;
    if (cond && (ci > 0)) { r1 = fRmin * si; z1 = fRmin * ci; r2 = fRmax * si; z2 = fRmax * ci; }
    if (cond && !(ci > 0)) { r1 = fRmax * si; z1 = fRmax * ci; r2 = fRmin * si; z2 = fRmin * ci; }
    double dz = 0.5 * (z2 - z1);
    V3 ptnew = p;
    ptnew.z = p.z - 0.5 * (z1 + z2);
    double zinv = 1. / dz;
    double rin = 0.5 * (r1 + r2 + (r2 - r1) * ptnew.z * zinv);
    double sigz = 1.;
    if (p.z < 0) sigz = -1. * sigz;
    bool bigequ = first ? ((sigz * ci > 0) && (sigz * rxy2 < sigz * rin * (rin + sigz * toler)))
                        : ((sigz * ci > 0) && (sigz * rxy2 > sigz * rin * (rin - sigz * toler)));
    double ddotn = 0;
    double ddotnTemp = ptnew.x * d.x + ptnew.y * d.y + 0.5 * (r1 - r2) * d.z * zinv * Sqrt(rxy2);
    if (cond && bigequ) ddotn = ddotnTemp;
    if (first) {
      if (cond && bigequ && (sigz * ddotn <= 0)) return false;
    } else {
      if (cond && bigequ && (sigz * ddotn >= 0)) return false;
    }
    double bv = 0, deltav = 0;
    ToCone(ptnew, d, dz, r1, r2, bv, deltav);
    bool ok = cond && !bigequ;
    if (ok) { delta = deltav; b = bv; }
    double xnew, ynew, znew, phi0, ddp, temp;
    if (ok && (delta > 0)) {
      snxt = -1. * b - delta;
      znew = ptnew.z + snxt * d.z;
      temp = znew < 0 ? -1. * znew : znew;
      if ((snxt > 0) && (temp < dz)) {
        if (fullPhi) {
          sn = snxt;
        } else {
          xnew = ptnew.x + snxt * d.x;
          ynew = ptnew.y + snxt * d.y;
          phi0 = ATan2(ynew, xnew);
          ddp = phi0 - fPhi1;
          if (ddp < 0) ddp = ddp + 2 * kPi;
          if (ddp <= (fPhi2 - fPhi1)) sn = snxt;
        }
      }
      // The second root: block one tests sn > toler2 for both the root and
      // its use; block two computes the root when sn > toler and uses it
      // when sn > toler2 (:1404-1406 and :1466-1468).
      bool compute = first ? (sn > toler2) : (sn > toler);
      if (compute) snxt = -1. * b + delta;
      if (sn > toler2) {
        znew = ptnew.z + snxt * d.z;
        temp = znew < 0 ? -1. * znew : znew;
        if ((snxt > 0) && (temp < dz)) {
          if (fullPhi) {
            sn = snxt;
          } else {
            xnew = ptnew.x + snxt * d.x;
            ynew = ptnew.y + snxt * d.y;
            phi0 = ATan2(ynew, xnew);
            ddp = phi0 - fPhi1;
            if (ddp < 0) ddp = ddp + 2 * kPi;
            if (ddp <= (fPhi2 - fPhi1)) sn = snxt;
          }
        }
      }
    }
    return true;
  }
  // DistanceToOutKernel :1179-1520
  double DistanceToOut(V3 p, V3 d, double) const override {
    const double toler = 1.E-10;
    double rad = Sqrt(p.Mag2());
    double pDotV = Dot(p, d);
    double rxy2 = p.x * p.x + p.y * p.y;
    double fPhi1 = fSPhi, fPhi2 = ePhi;
    double sn1 = kInfinity, rn1 = kInfinity, sn2 = kInfinity, rn2 = kInfinity;
    if ((fRmin > 0) && (rad <= (fRmin + toler)) && (pDotV < 0)) return kInfinity;
    double dist = ToSphere(p, d, fRmin, true, true);
    if ((fRmin > 0) && !(rad <= (fRmin + toler)) && (pDotV < 0)) rn1 = dist;
    if ((rad >= (fRmax - toler)) && (pDotV >= 0)) return kInfinity;
    rn2 = ToSphere(p, d, fRmax, false, false);
    double sr = Min(rn1, rn2);
    double fTheta1 = fSTheta, fTheta2 = eTheta;
    double theta2Plane = std::abs(fTheta2 - kPi / 2);
    double b = 0, delta = 0, snxt = kInfinity;
    if (!fullTheta) {
      double theta1Plane = std::abs(fTheta1 - kPi / 2);
      if (theta1Plane < 1.e-10) {
        if ((p.z * d.z) < 0) sn1 = -1. * p.z / d.z;
      } else {
        if (!ThetaBlock(p, d, rxy2, fTheta1 > 0, sinSTheta, cosSTheta, true, fPhi1, fPhi2, b, delta, snxt, sn1))
          return 0.;
      }
      snxt = kInfinity;
      if (theta2Plane < 1.e-10) {
        if ((p.z * d.z) < 0) sn1 = -1. * p.z / d.z;
      } else {
        if (!ThetaBlock(p, d, rxy2, fTheta2 < kPi, sinETheta, cosETheta, false, fPhi1, fPhi2, b, delta, snxt, sn2))
          return 0.;
      }
    }
    double st = Min(sn1, sn2);
    double sp = kInfinity;
    if (!fullPhi) {
      double phim = 0.5 * (fPhi1 + fPhi2);
      sp = ToPhiMin(p, d, sinSPhi, cosSPhi, sinEPhi, cosEPhi, std::sin(phim), std::cos(phim));
    }
    double srt = Min(sr, st);
    return Min(srt, sp);
  }
  // UnplacedSphere::Extent, source/UnplacedSphere.cpp (full cube of rmax)
  void Extent(V3 &lo, V3 &hi) const override {
    lo = {-fRmax, -fRmax, -fRmax};
    hi = {fRmax, fRmax, fRmax};
  }
};

Shape *MakeSphere(const double *p, int n) { return n == 6 ? new Sphere(p) : nullptr; }
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

  // UnplacedTrapezoid::ApproxSurfaceNormal, source/UnplacedTrapezoid.cpp:583-612
  // (PlaneShell form): the nearest side plane's normal unless the z lid is
  // at least as near.
  V3 ApproxNormal(V3 p) const {
    double safe = kInfinity;
    int imin = 0;
    for (int i = 0; i < 4; i++) {
      const double dist = std::fabs(A[i] * p.x + B[i] * p.y + C[i] * p.z + D[i]);
      if (dist < safe) {
        safe = dist;
        imin = i;
      }
    }
    const double safez = std::fabs(std::fabs(p.z) - dz);
    if (safe < safez) return V3{A[imin], B[imin], C[imin]};
    return p.z > 0 ? V3{0, 0, 1} : V3{0, 0, -1};
  }
  // UnplacedTrapezoid::Normal, source/UnplacedTrapezoid.cpp:352-425
  // (PlaneShell form, the reference's default build): sum of the side planes
  // within kHalfTolerance and the z lid, as is with one surface, Unit() with
  // several, the approximate normal (invalid) with none.
  int Normal(V3 p, V3 &n) const override {
    int noSurfaces = 0;
    V3 sumnorm{0., 0., 0.};
    for (int i = 0; i < 4; ++i) {
      const double dist = A[i] * p.x + B[i] * p.y + C[i] * p.z + D[i];
      if (std::fabs(dist) <= kHalfTolerance) {
        noSurfaces++;
        sumnorm = sumnorm + V3{A[i], B[i], C[i]};
      }
    }
    const double distz = std::fabs(p.z) - dz;
    if (std::fabs(distz) <= kHalfTolerance) {
      noSurfaces++;
      if (p.z >= 0.)
        sumnorm = sumnorm + V3{0., 0., 1.};
      else
        sumnorm = sumnorm - V3{0., 0., 1.};
    }
    if (noSurfaces == 0) {
      n = ApproxNormal(p);
    } else if (noSurfaces == 1) {
      n = sumnorm;
    } else {
      const double s = Sqrt(sumnorm.Mag2() + kMinimum);
      n = V3{sumnorm.x / s, sumnorm.y / s, sumnorm.z / s};
    }
    return noSurfaces != 0 ? 1 : 0;
  }
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

}  // namespace vgo
