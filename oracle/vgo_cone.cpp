// ORACLE -- test infrastructure only (see vgo.h).
//
// Cone restated from the reference:
//   cached values             volumes/UnplacedCone.h:131-197
//   Contains / Inside         volumes/kernel/ConeImplementation.h:84-235
//                             (UniversalCone: rmin treatment when rmin1 > 0 or
//                             rmin2 > 0, phi treatment when dphi < 2 pi,
//                             shapetypes/ConeTypes.h:107-162)
//   DistanceToIn              ConeImplementation.h:237-1060 (the USolids
//                             algorithm the reference dispatches to)
//   DistanceToOut             ConeImplementation.h:1221-1796
//   SafetyToIn / SafetyToOut  ConeImplementation.h:1908-1996
//   Extent                    UnplacedCone.h:324-328
// IsFullPhi is dphi == 2 pi exactly (UnplacedCone.h:258); the phi treatment
// of Contains/Inside uses dphi < 2 pi.  stepMax is not used by the
// reference's cone.
#include <vector>

#include "vgo.h"
#include "vgo_wedge.h"

namespace vgo {

namespace {

constexpr double kRadTol = 1e-9;         // base/Global.h:299
constexpr double kAngTol = 1e-9;         // base/Global.h:300
constexpr double kHalfCarTol = 0.5e-9;   // kTolerance * 0.5
constexpr double kHalfRadTol = 0.5e-9;   // kRadTolerance * 0.5
constexpr double kHalfAngTol = 0.5e-9;   // kAngTolerance * 0.5

struct Cone : Shape {
  double rmin1, rmax1, rmin2, rmax2, dz, sphi, dphi;
  double innerSlope, outerSlope, innerOffset, outerOffset;
  double secRMin, secRMax, invSecRMin, invSecRMax, tanRMin, tanRMax;
  double cosCPhi, sinCPhi, cosSPhi, sinSPhi, cosEPhi, sinEPhi, cosHDPhi, cosHDPhiIT, cosHDPhiOT;
  bool fullPhi, phiTreatment, rminTreatment;
  Wedge wedge;

  Cone(const double *a) {
    rmin1 = a[0];
    rmax1 = a[1];
    rmin2 = a[2];
    rmax2 = a[3];
    dz = a[4];
    sphi = a[5];
    dphi = a[6];
    wedge.Init(dphi, sphi);
    const double hDPhi = 0.5 * dphi;
    const double cPhi = sphi + hDPhi;
    const double ePhi = sphi + dphi;
    sinCPhi = std::sin(cPhi);
    cosCPhi = std::cos(cPhi);
    cosHDPhi = std::cos(hDPhi);
    cosHDPhiIT = std::cos(hDPhi - 0.5 * kAngTol);
    cosHDPhiOT = std::cos(hDPhi + 0.5 * kAngTol);
    sinSPhi = std::sin(sphi);
    cosSPhi = std::cos(sphi);
    sinEPhi = std::sin(ePhi);
    cosEPhi = std::cos(ePhi);
    tanRMin = (rmin2 - rmin1) * 0.5 / dz;
    secRMin = std::sqrt(1.0 + tanRMin * tanRMin);
    invSecRMin = 1. / secRMin;
    tanRMax = (rmax2 - rmax1) * 0.5 / dz;
    secRMax = std::sqrt(1.0 + tanRMax * tanRMax);
    invSecRMax = 1. / secRMax;
    innerSlope = -(rmin1 - rmin2) / (2. * dz);
    outerSlope = -(rmax1 - rmax2) / (2. * dz);
    innerOffset = rmin2 - innerSlope * dz;
    outerOffset = rmax2 - outerSlope * dz;
    fullPhi = dphi == kTwoPi;
    phiTreatment = dphi < 2. * M_PI;
    rminTreatment = rmin1 > 0 || rmin2 > 0;
  }

  // The two normals of the reference's cone: ConeImplementation::NormalKernel
  // (ConeImplementation.h:2002-2128, PlacedCone::Normal) and
  // UnplacedCone::Normal (source/UnplacedCone.cpp:31-141, what a Polycone
  // section answers).  They differ only in how they scale by sec(alpha):
  // the kernel multiplies by the cached inverse where the unplaced form
  // divides.  Surfaces within 0.5 kTolerance (the phi planes: half the
  // angular tolerance, the same number) are summed in the order outer cone,
  // inner cone, start phi, end phi, z lid; unit length when several.
  template <bool Kernel_>
  int NormalT(V3 p, V3 &n) const {
    int noSurfaces = 0;
    double distSPhi = kInfinity, distEPhi = kInfinity;
    const double delta = 0.5 * kTolerance, dAngle = 0.5 * kAngTol;
    V3 sumnorm{0., 0., 0.}, nR{0, 0, 0}, nr{0., 0., 0.}, nPs{0, 0, 0}, nPe{0, 0, 0};
    const V3 nZ{0., 0., 1.};
    const double distZ = Abs(Abs(p.z) - dz);
    const double rho = Sqrt(p.x * p.x + p.y * p.y);
    const double pRMin = rho - p.z * tanRMin;
    const double widRMin = rmin2 - dz * tanRMin;
    const double distRMin = Kernel_ ? Abs(pRMin - widRMin) * invSecRMin : Abs(pRMin - widRMin) / secRMin;
    const double pRMax = rho - p.z * tanRMax;
    const double widRMax = rmax2 - dz * tanRMax;
    const double distRMax = Kernel_ ? Abs(pRMax - widRMax) * invSecRMax : Abs(pRMax - widRMax) / secRMax;
    if (!fullPhi) {
      if (rho != 0.) {
        // The kernel calls the backend's ATan2, the unplaced form std::atan2.
        double pPhi = Kernel_ ? ATan2(p.y, p.x) : std::atan2(p.y, p.x);
        if (pPhi < sphi - delta)
          pPhi += kTwoPi;
        else if (pPhi > sphi + dphi + delta)
          pPhi -= kTwoPi;
        distSPhi = Abs(pPhi - sphi);
        distEPhi = Abs(pPhi - sphi - dphi);
      } else if (!(rmin1 != 0.) || !(rmin2 != 0.)) {
        distSPhi = 0.;
        distEPhi = 0.;
      }
      nPs = V3{sinSPhi, -cosSPhi, 0};
      nPe = V3{-sinEPhi, cosEPhi, 0};
    }
    if (rho > delta) {
      nR = V3{p.x / rho / secRMax, p.y / rho / secRMax, Kernel_ ? -tanRMax * invSecRMax : -tanRMax / secRMax};
      if (rmin1 != 0. || rmin2 != 0.)
        nr = V3{-p.x / rho / secRMin, -p.y / rho / secRMin, Kernel_ ? tanRMin * invSecRMin : tanRMin / secRMin};
    }
    if (distRMax <= delta) {
      noSurfaces++;
      sumnorm = sumnorm + nR;
    }
    if ((rmin1 != 0. || rmin2 != 0.) && (distRMin <= delta)) {
      noSurfaces++;
      sumnorm = sumnorm + nr;
    }
    if (!fullPhi) {
      if (distSPhi <= dAngle) {
        noSurfaces++;
        sumnorm = sumnorm + nPs;
      }
      if (distEPhi <= dAngle) {
        noSurfaces++;
        sumnorm = sumnorm + nPe;
      }
    }
    if (distZ <= delta) {
      noSurfaces++;
      if (p.z >= 0.)
        sumnorm = sumnorm + nZ;
      else
        sumnorm = sumnorm - nZ;
    }
    if (noSurfaces == 0) {
      n = V3{0, 0, 0};  // the kernel's default vector, the unplaced form's zero sum
    } else if (noSurfaces == 1) {
      n = sumnorm;
    } else {
      // Vector3D::Unit: divide by Sqrt(Mag2 + kMinimum).
      const double s = Sqrt(sumnorm.Mag2() + kMinimum);
      n = V3{sumnorm.x / s, sumnorm.y / s, sumnorm.z / s};
    }
    return noSurfaces != 0 ? 1 : 0;
  }
  int Normal(V3 p, V3 &n) const override { return NormalT<true>(p, n); }
  int UnplacedNormal(V3 p, V3 &n) const { return NormalT<false>(p, n); }

  // GenericKernelForContainsAndInside, early returns as the Scalar backend.
  template <bool ForInside>
  void Kernel(V3 p, bool &cin, bool &cout) const {
    const double absz = Abs(p.z);
    cout = absz > MakePlusTolerant<ForInside>(dz);
    if (ForInside) cin = absz < MakeMinusTolerant<ForInside>(dz);
    if (cout) return;
    const double r2 = p.x * p.x + p.y * p.y;
    const double rmax = outerSlope * p.z + outerOffset;
    const double rmax2 = rmax * rmax;
    cout |= r2 > MakePlusTolerantSquare<ForInside>(rmax, rmax2);
    if (ForInside) cin &= r2 < MakeMinusTolerantSquare<ForInside>(rmax, rmax2);
    if (cout) return;
    if (rminTreatment) {
      const double rmin = innerSlope * p.z + innerOffset;
      const double rmin2 = rmin * rmin;
      cout |= r2 <= MakeMinusTolerantSquare<ForInside>(rmin, rmin2);
      if (ForInside) cin &= r2 > MakePlusTolerantSquare<ForInside>(rmin, rmin2);
      if (cout) return;
    }
    if (phiTreatment) {
      bool wi, wo;
      wedge.Kernel<ForInside>(p, wi, wo);
      cout |= wo;
      if (ForInside) cin &= wi;
    }
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

  // cos of the angle between (x, y) and the central phi, times r, against a
  // threshold: the "correct half of phi" test of the USolids algorithm.
  bool PhiAccept(double xi, double yi, double r, double cosLimit) const {
    const double cosPsi = (xi * cosCPhi + yi * sinCPhi) / r;
    return cosPsi >= cosLimit;
  }
  // The inner-surface normal at (xi, yi) does not point along v.
  bool InnerNormalAgainst(double xi, double yi, V3 v) const {
    const double risec = std::sqrt(xi * xi + yi * yi) * secRMin;
    const V3 norm{-xi / risec, -yi / risec, tanRMin / secRMin};
    return Dot(norm, v) <= 0;
  }
  // Larger / smaller root of t^2 + 2bt + c = 0 in the reference's two forms.
  static double RootFar(double b, double c, double d) { return b > 0 ? c / (-b - std::sqrt(d)) : -b + std::sqrt(d); }
  static double RootNear(double b, double c, double d) { return b > 0 ? -b - std::sqrt(d) : c / (-b + std::sqrt(d)); }

  // Handling of an inner-cone candidate sd at (xi, yi): records it in snxt
  // (phi section) or returns it (full phi) -- returns true if the caller
  // must return sd.
  bool InnerCandidate(V3 p, V3 v, double sd, double ri, double cosLimit, double &snxt) const {
    if (!fullPhi) {
      const double xi = p.x + sd * v.x, yi = p.y + sd * v.y;
      if (PhiAccept(xi, yi, ri, cosLimit)) {
        if (sd > kHalfRadTol || InnerNormalAgainst(xi, yi, v)) snxt = sd;
      }
      return false;
    }
    if (sd > kHalfRadTol) return true;
    const double xi = p.x + sd * v.x, yi = p.y + sd * v.y;
    return InnerNormalAgainst(xi, yi, v);
  }

  double DistanceToIn(V3 p, V3 v, double) const override {
    double snxt = kInfinity;
    const double rMinAv = (rmin1 + rmin2) * 0.5;
    const double rMinOAv = rMinAv > kHalfRadTol ? rMinAv - kHalfRadTol : 0.0;
    const double rMaxAv = (rmax1 + rmax2) * 0.5;
    const double rMaxOAv = rMaxAv + kHalfRadTol;
    const double tolIDz = dz - kHalfCarTol;
    const double tolODz = dz + kHalfCarTol;
    double sd, xi, yi, zi, ri, b, c, d;

    // z faces
    if (std::fabs(p.z) >= tolIDz) {
      if (!(p.z * v.z < 0)) return snxt;
      sd = (std::fabs(p.z) - dz) / std::fabs(v.z);
      if (sd < 0.0) sd = 0.0;
      xi = p.x + sd * v.x;
      yi = p.y + sd * v.y;
      const double rhoi2 = xi * xi + yi * yi;
      const double rminE = v.z > 0 ? rmin1 : rmin2, rmaxE = v.z > 0 ? rmax1 : rmax2;
      const double tolORMin = rminE - kHalfRadTol * secRMin;
      const double tolIRMin = rminE + kHalfRadTol * secRMin;
      const double tolIRMax = rmaxE - kHalfRadTol * secRMin;
      const double tolIRMin2 = tolORMin > 0 ? tolIRMin * tolIRMin : 0.0;
      const double tolIRMax2 = tolIRMax > 0 ? tolIRMax * tolIRMax : 0.0;
      if ((tolIRMin2 <= rhoi2) && (rhoi2 <= tolIRMax2)) {
        if (!fullPhi && rhoi2 != 0) {
          if (PhiAccept(xi, yi, std::sqrt(rhoi2), cosHDPhiIT)) return sd;
        } else {
          return sd;
        }
      }
    }

    // outer cone
    const double t1 = 1.0 - v.z * v.z;
    const double t2 = p.x * v.x + p.y * v.y;
    const double t3 = p.x * p.x + p.y * p.y;
    const double rin = tanRMin * p.z + rMinAv;
    const double rout = tanRMax * p.z + rMaxAv;
    double nt1 = t1 - (tanRMax * v.z) * (tanRMax * v.z);
    double nt2 = t2 - tanRMax * v.z * rout;
    double nt3 = t3 - rout * rout;
    if (std::fabs(nt1) > kRadTol) {
      b = nt2 / nt1;
      c = nt3 / nt1;
      d = b * b - c;
      if ((nt3 > rout * rout * kRadTol * kRadTol * secRMax * secRMax) || (rout < 0)) {
        if (d >= 0) {
          if ((rout < 0) && (nt3 <= 0)) {
            sd = RootFar(b, c, d);
          } else if ((b <= 0) && (c >= 0)) {
            sd = c / (-b + std::sqrt(d));
          } else if (c <= 0) {
            sd = -b + std::sqrt(d);
          } else {
            return kInfinity;
          }
          if (sd > 0) {
            zi = p.z + sd * v.z;
            if (std::fabs(zi) <= tolODz) {
              if (fullPhi) return sd;
              xi = p.x + sd * v.x;
              yi = p.y + sd * v.y;
              ri = rMaxAv + zi * tanRMax;
              if (PhiAccept(xi, yi, ri, cosHDPhiIT)) return sd;
            }
          }
        }
      } else {
        // inside the outer cone: on the surface heading in -> 0
        const double rinT = rin + kHalfRadTol * secRMin;
        if ((t3 > rinT * rinT) && (nt2 < 0) && (d >= 0) && (std::fabs(p.z) <= tolIDz)) {
          const double risec = std::sqrt(p.x * p.x + p.y * p.y) * secRMax;
          const V3 norm{p.x / risec, p.y / risec, -tanRMax / secRMax};
          if (!fullPhi) {
            if (PhiAccept(p.x, p.y, std::sqrt(t3), cosHDPhiIT) && Dot(norm, v) <= 0) return 0.0;
          } else if (Dot(norm, v) <= 0) {
            return 0.0;
          }
        }
      }
    } else if (std::fabs(nt2) > kRadTol) {
      sd = -0.5 * nt3 / nt2;
      if (sd < 0) return kInfinity;
      zi = p.z + sd * v.z;
      if ((std::fabs(zi) <= tolODz) && (nt2 < 0)) {
        if (fullPhi) return sd;
        xi = p.x + sd * v.x;
        yi = p.y + sd * v.y;
        ri = rMaxAv + zi * tanRMax;
        if (PhiAccept(xi, yi, ri, cosHDPhiIT)) return sd;
      }
    }

    // inner cone
    if (rMinAv != 0) {
      nt1 = t1 - (tanRMin * v.z) * (tanRMin * v.z);
      nt2 = t2 - tanRMin * v.z * rin;
      nt3 = t3 - rin * rin;
      if (nt1 != 0) {
        if (nt3 > rin * kRadTol * secRMin) {
          // outside the real and the imaginary inner cone: far root
          b = nt2 / nt1;
          c = nt3 / nt1;
          d = b * b - c;
          if (d >= 0) {
            sd = RootFar(b, c, d);
            if (sd >= 0) {
              zi = p.z + sd * v.z;
              if (std::fabs(zi) <= tolODz) {
                ri = rMinAv + zi * tanRMin;
                if (InnerCandidate(p, v, sd, ri, cosHDPhiIT, snxt)) return sd;
              }
            }
          }
        } else if (nt3 < -rin * kRadTol * secRMin) {
          // within the inner cone (real or imaginary): far root on the real
          // cone, else the near root
          b = nt2 / nt1;
          c = nt3 / nt1;
          d = b * b - c;
          if (d >= 0) {
            sd = RootFar(b, c, d);
            zi = p.z + sd * v.z;
            ri = rMinAv + zi * tanRMin;
            if (ri > 0) {
              if ((sd >= 0) && (std::fabs(zi) <= tolODz)) {
                if (InnerCandidate(p, v, sd, ri, cosHDPhiOT, snxt)) return sd;
              }
            } else {
              sd = RootNear(b, c, d);
              zi = p.z + sd * v.z;
              ri = rMinAv + zi * tanRMin;
              if ((sd >= 0) && (ri > 0) && (std::fabs(zi) <= tolODz)) {
                if (InnerCandidate(p, v, sd, ri, cosHDPhiIT, snxt)) return sd;
              }
            }
          }
        } else if (std::fabs(p.z) <= tolODz) {
          // on the inner cone within tolerance, inside the z range
          if (nt2 > 0) {
            if (!fullPhi) {
              if (PhiAccept(p.x, p.y, std::sqrt(t3), cosHDPhiIT)) return 0.0;
            } else {
              return 0.0;
            }
          } else {
            b = nt2 / nt1;
            c = nt3 / nt1;
            d = b * b - c;
            if (d >= 0) {
              sd = RootNear(b, c, d);
              zi = p.z + sd * v.z;
              ri = rMinAv + zi * tanRMin;
              if (!(ri > 0)) return kInfinity;
              sd = RootFar(b, c, d);
              zi = p.z + sd * v.z;
              if ((sd >= 0) && (std::fabs(zi) <= tolODz)) {
                if (fullPhi) return sd;
                xi = p.x + sd * v.x;
                yi = p.y + sd * v.y;
                ri = rMinAv + zi * tanRMin;
                if (PhiAccept(xi, yi, ri, cosHDPhiIT)) snxt = sd;
              }
            }
          }
        } else {
          // on the inner cone within tolerance, outside the z range
          b = nt2 / nt1;
          c = nt3 / nt1;
          d = b * b - c;
          if (d > 0) {
            sd = RootFar(b, c, d);
            zi = p.z + sd * v.z;
            if ((sd >= 0) && (std::fabs(zi) <= tolODz)) {
              if (fullPhi) return sd;
              xi = p.x + sd * v.x;
              yi = p.y + sd * v.y;
              ri = rMinAv + zi * tanRMin;
              if (PhiAccept(xi, yi, ri, cosHDPhiIT)) snxt = sd;
            }
          }
        }
      }
    }

    // phi planes
    if (!fullPhi) {
      PhiPlaneIn(p, v, v.x * sinSPhi - v.y * cosSPhi, p.y * cosSPhi - p.x * sinSPhi, false, rMinOAv, rMaxOAv, tolODz,
                 snxt);
      PhiPlaneIn(p, v, -(v.x * sinEPhi - v.y * cosEPhi), -(p.y * cosEPhi - p.x * sinEPhi), true, rMinOAv, rMaxOAv,
                 tolODz, snxt);
    }
    if (snxt < kHalfCarTol) snxt = 0.;
    return snxt;
  }

  // One phi plane of DistanceToIn (ConeImplementation.h:980-1048).
  void PhiPlaneIn(V3 p, V3 v, double comp, double dist, bool endPlane, double rMinOAv, double rMaxOAv, double tolODz,
                  double &snxt) const {
    if (!(comp < 0)) return;
    if (!(dist < kHalfCarTol)) return;
    double sd = dist / comp;
    if (!(sd < snxt)) return;
    if (sd < 0) sd = 0.0;
    const double zi = p.z + sd * v.z;
    if (!(std::fabs(zi) <= tolODz)) return;
    const double xi = p.x + sd * v.x, yi = p.y + sd * v.y;
    const double rhoi2 = xi * xi + yi * yi;
    const double tolORMin2 = (rMinOAv + zi * tanRMin) * (rMinOAv + zi * tanRMin);
    const double tolORMax2 = (rMaxOAv + zi * tanRMax) * (rMaxOAv + zi * tanRMax);
    if ((rhoi2 >= tolORMin2) && (rhoi2 <= tolORMax2)) {
      const double side = yi * cosCPhi - xi * sinCPhi;
      if (endPlane ? (side >= 0.0) : (side <= 0)) snxt = sd;
    }
  }

  double DistanceToOut(V3 p, V3 v, double) const override {
    double snxt, srd, pdist;
    if (v.z > 0.0) {
      pdist = dz - p.z;
      if (!(pdist > kHalfCarTol)) return 0.0;
      snxt = pdist / v.z;
    } else if (v.z < 0.0) {
      pdist = dz + p.z;
      if (!(pdist > kHalfCarTol)) return 0.0;
      snxt = -pdist / v.z;
    } else {
      snxt = kInfinity;
    }

    // outer cone
    const double rMaxAv = (rmax1 + rmax2) * 0.5;
    const double t1 = 1.0 - v.z * v.z;
    const double t2 = p.x * v.x + p.y * v.y;
    const double t3 = p.x * p.x + p.y * p.y;
    const double rout = tanRMax * p.z + rMaxAv;
    double nt1 = t1 - (tanRMax * v.z) * (tanRMax * v.z);
    double nt2 = t2 - tanRMax * v.z * rout;
    double nt3 = t3 - rout * rout;
    double deltaRoi2;
    if (v.z > 0.0) {
      deltaRoi2 = snxt * snxt * t1 + 2 * snxt * t2 + t3 - rmax2 * (rmax2 + kRadTol * secRMax);
    } else if (v.z < 0.0) {
      deltaRoi2 = snxt * snxt * t1 + 2 * snxt * t2 + t3 - rmax1 * (rmax1 + kRadTol * secRMax);
    } else {
      deltaRoi2 = 1.0;
    }
    double b, c, d, zi, ri;
    if (nt1 != 0 && (deltaRoi2 > 0.0)) {
      b = nt2 / nt1;
      c = nt3 / nt1;
      d = b * b - c;
      if (!(d >= 0)) return 0.0;  // no intersection, already outside
      if (nt3 > -kHalfRadTol && nt2 >= 0) return 0.0;
      srd = b > 0 ? -b - std::sqrt(d) : c / (-b + std::sqrt(d));
      zi = p.z + srd * v.z;
      ri = tanRMax * zi + rMaxAv;
      if ((ri < 0) || (srd < kHalfRadTol)) {
        const double sr2 = b > 0 ? c / (-b - std::sqrt(d)) : -b + std::sqrt(d);
        zi = p.z + sr2 * v.z;
        ri = tanRMax * zi + rMaxAv;
        srd = ((ri >= 0) && (sr2 > kHalfRadTol)) ? sr2 : kInfinity;
      }
    } else if (nt2 != 0 && (deltaRoi2 > 0.0)) {
      return 0.0;  // linear case: outside the outer cone
    } else {
      srd = kInfinity;
    }

    // inner cone
    if (rmin1 != 0 || rmin2 != 0) {
      nt1 = t1 - (tanRMin * v.z) * (tanRMin * v.z);
      if (nt1 != 0) {
        const double rMinAv = (rmin1 + rmin2) * 0.5;
        const double rin = tanRMin * p.z + rMinAv;
        nt2 = t2 - tanRMin * v.z * rin;
        nt3 = t3 - rin * rin;
        b = nt2 / nt1;
        c = nt3 / nt1;
        d = b * b - c;
        if (d >= 0.0) {
          if (nt3 < kRadTol * (rin + kRadTol * 0.25)) {
            if (nt2 < 0.0) return 0.0;
          } else {
            const double sr2 = b > 0 ? -b - std::sqrt(d) : c / (-b + std::sqrt(d));
            zi = p.z + sr2 * v.z;
            ri = tanRMin * zi + rMinAv;
            if ((ri < 0) || (sr2 < kHalfRadTol)) {
              const double sr3 = b > 0 ? c / (-b - std::sqrt(d)) : -b + std::sqrt(d);
              if (sr3 > kHalfRadTol && sr3 < srd) {
                zi = p.z + sr3 * v.z;
                ri = tanRMin * zi + rMinAv;
                if (ri >= 0.0) srd = sr3;
              }
            } else if ((sr2 < srd) && (sr2 > kHalfCarTol)) {
              srd = sr2;
            }
          }
        }
      }
    }

    // phi planes
    if (!fullPhi) {
      const double sphi = PhiOut(p, v);
      if (sphi < snxt) snxt = sphi;
    }
    if (srd < snxt) snxt = srd;
    if (snxt < kHalfCarTol) snxt = 0.;
    return snxt;
  }

  // Phi part of DistanceToOut (ConeImplementation.h:1620-1776).
  double PhiOut(V3 p, V3 v) const {
    double vphi = ATan2(v.y, v.x);
    if (vphi < sphi - kHalfAngTol) {
      vphi += 2 * kPi;
    } else if (vphi > sphi + dphi + kHalfAngTol) {
      vphi -= 2 * kPi;
    }
    const bool dirInside = (sphi - kHalfAngTol <= vphi) && (sphi + dphi + kHalfAngTol >= vphi);
    if (!(p.x != 0 || p.y != 0)) {
      // on the z axis: limited by phi only if the direction leaves the section
      return ((sphi - kHalfAngTol <= vphi) && (vphi <= sphi + dphi + kHalfAngTol)) ? kInfinity : 0.0;
    }
    const double pDistS = p.x * sinSPhi - p.y * cosSPhi;
    const double pDistE = -p.x * sinEPhi + p.y * cosEPhi;
    const double compS = -sinSPhi * v.x + cosSPhi * v.y;
    const double compE = sinEPhi * v.x - cosEPhi * v.y;
    const bool insidePlanes = ((dphi <= kPi) && ((pDistS <= kHalfCarTol) && (pDistE <= kHalfCarTol))) ||
                              ((dphi > kPi) && !((pDistS > kHalfCarTol) && (pDistE > kHalfCarTol)));
    if (!insidePlanes) return kInfinity;
    double sphiD = kInfinity;
    if (compS < 0) {
      sphiD = pDistS / compS;
      if (sphiD >= -kHalfCarTol) {
        const double xi = p.x + sphiD * v.x, yi = p.y + sphiD * v.y;
        if ((Abs(xi) <= kTolerance) && (Abs(yi) <= kTolerance)) {
          if (dirInside) sphiD = kInfinity;
        } else if ((yi * cosCPhi - xi * sinCPhi) >= 0) {
          sphiD = kInfinity;
        } else if (pDistS > -kHalfCarTol) {
          sphiD = 0.0;
        }
      } else {
        sphiD = kInfinity;
      }
    }
    if (compE < 0) {
      const double sphi2 = pDistE / compE;
      if ((sphi2 > -kHalfCarTol) && (sphi2 < sphiD)) {
        const double xi = p.x + sphi2 * v.x, yi = p.y + sphi2 * v.y;
        bool leaves;
        if ((Abs(xi) <= kTolerance) && (Abs(yi) <= kTolerance)) {
          leaves = !dirInside;
        } else {
          leaves = yi * cosCPhi - xi * sinCPhi >= 0;
        }
        if (leaves) sphiD = (pDistE <= -kHalfCarTol) ? sphi2 : 0.0;
      }
    }
    return sphiD;
  }

  double SafetyToIn(V3 p) const override { return SafetyToInImpl(p, false); }

  // SafetyToInUSOLIDS<Backend, ForPolycone>: the polycone asks for the
  // radial and phi parts only (no z term) for the section holding the point.
  double SafetyToInImpl(V3 p, bool forPolycone) const {
    double safe;
    const double rho = std::sqrt(p.x * p.x + p.y * p.y);
    const double safeZ = Abs(p.z) - dz;
    if (rmin1 != 0 || rmin2 != 0) {
      const double pRMin = tanRMin * p.z + (rmin1 + rmin2) * 0.5;
      const double safeR1 = (pRMin - rho) * invSecRMin;
      const double pRMax = tanRMax * p.z + (rmax1 + rmax2) * 0.5;
      const double safeR2 = (rho - pRMax) * invSecRMax;
      safe = safeR1 > safeR2 ? safeR1 : safeR2;
    } else {
      const double pRMax = tanRMax * p.z + (rmax1 + rmax2) * 0.5;
      safe = (rho - pRMax) * invSecRMax;
    }
    if (!forPolycone && safeZ > safe) safe = safeZ;
    if (!fullPhi && rho != 0) {
      const double cosPsi = p.x * cosCPhi + p.y * sinCPhi;
      if (cosPsi < cosHDPhi * rho) {
        const double safePhi = ((p.y * cosCPhi - p.x * sinCPhi) <= 0.0) ? Abs(p.x * sinSPhi - p.y * cosSPhi)
                                                                          : Abs(p.x * sinEPhi - p.y * cosEPhi);
        if (safePhi > safe) safe = safePhi;
      }
    }
    if (safe < 0.0) safe = 0.0;
    return safe;
  }

  double SafetyToOut(V3 p) const override {
    const double rho = std::sqrt(p.x * p.x + p.y * p.y);
    const double safeZ = dz - Abs(p.z);
    double safeR1 = kInfinity;
    if (rmin1 != 0 || rmin2 != 0) {
      const double pRMin = tanRMin * p.z + (rmin1 + rmin2) * 0.5;
      safeR1 = (rho - pRMin) * invSecRMin;
    }
    const double pRMax = tanRMax * p.z + (rmax1 + rmax2) * 0.5;
    const double safeR2 = (pRMax - rho) * invSecRMax;
    double safe = safeR1 < safeR2 ? safeR1 : safeR2;
    if (safeZ < safe) safe = safeZ;
    if (!fullPhi) {
      const double safePhi = ((p.y * cosCPhi - p.x * sinCPhi) <= 0) ? -(p.x * sinSPhi - p.y * cosSPhi)
                                                                    : (p.x * sinEPhi - p.y * cosEPhi);
      if (safePhi < safe) safe = safePhi;
    }
    if (safe < 0) safe = 0;
    return safe;
  }

  void Extent(V3 &lo, V3 &hi) const override {
    const double m = rmax1 > rmax2 ? rmax1 : rmax2;
    lo = V3{-m, -m, -dz};
    hi = V3{m, m, dz};
  }
};

// ---------------------------------------------------------------- Polycone
// volumes/UnplacedPolycone.h:32-261, source/UnplacedPolycone.cpp:27-180 (Init)
// and :755-774 (Extent), volumes/kernel/PolyconeImplementation.h:58-494.
struct Polycone : Shape {
  std::vector<Cone> sec;
  std::vector<double> shift;
  std::vector<double> zs;  // nSections + 1

  // construction parameters: sphi dphi nz rmin[nz] rmax[nz] z[nz]
  bool Init(const double *a, int n) {
    if (n < 3) return false;
    const int nz = (int)a[2];
    if (nz < 2 || n != 3 + 3 * nz) return false;
    const double phiStart = a[0], phiTotal = a[1];
    const double *rmin = a + 3, *rmax = a + 3 + nz, *z = a + 3 + 2 * nz;
    double prevZ = z[0], prevRmax = 0, prevRmin = 0;
    const int dirZ = z[1] < z[0] ? -1 : 1;
    for (int i = 0; i < nz; ++i) {
      if (i == 0) {
        zs.push_back(z[i]);
      } else if (((z[i] > prevZ) && (dirZ > 0)) || ((z[i] < prevZ) && (dirZ < 0))) {
        const double c[7] = {prevRmin, prevRmax, rmin[i], rmax[i], (z[i] - prevZ) / 2, phiStart, phiTotal};
        sec.emplace_back(c);
        zs.push_back(z[i]);
        const int zi = (int)zs.size() - 1;
        shift.push_back(zs[zi - 1] + 0.5 * (zs[zi] - zs[zi - 1]));
      }
      prevZ = z[i];
      prevRmin = rmin[i];
      prevRmax = rmax[i];
    }
    return !sec.empty();
  }
  int NSec() const { return (int)sec.size(); }
  // GetSectionIndex, UnplacedPolycone.h:111-120
  int SectionIndex(double z) const {
    if (z < zs[0]) return -1;
    for (int i = 0; i < NSec(); ++i)
      if (z >= zs[i] && z <= zs[i + 1]) return i;
    return -2;
  }
  static V3 Shift(V3 p, double s) { return V3{p.x - 0., p.y - 0., p.z - s}; }

  // UnplacedPolycone::Normal, source/UnplacedPolycone.cpp:684-698: -z or +z
  // below or above the sections (valid), else the section cone's
  // UnplacedCone::Normal of the point as given -- the reference does not
  // shift it into the section's frame.
  int Normal(V3 p, V3 &n) const override {
    const int index = SectionIndex(p.z);
    n = V3{0, 0, 0};
    if (index < 0) {
      if (index == -1) n = V3{0., 0., -1.};
      if (index == -2) n = V3{0., 0., 1.};
      return 1;
    }
    return sec[index].UnplacedNormal(p, n);
  }

  bool Contains(V3 p) const override {
    if (p.z < zs[0] || p.z > zs[NSec()]) return false;
    const int i = SectionIndex(p.z);
    return sec[i].Contains(Shift(p, shift[i]));
  }
  // The reference's polycone Inside reports only inside or outside.
  int Inside(V3 p) const override { return Contains(p) ? kInside : kOutside; }

  double DistanceToIn(V3 p, V3 v, double stepMax) const override {
    double distance = kInfinity;
    int increment = (v.z > 0) ? 1 : -1;
    if (std::fabs(v.z) < kTolerance) increment = 0;
    int index = SectionIndex(p.z);
    if (index == -1) index = 0;
    if (index == -2) index = NSec() - 1;
    do {
      distance = sec[index].DistanceToIn(Shift(p, shift[index]), v, stepMax);
      if (distance < kInfinity || !increment) break;
      index += increment;
    } while (index >= 0 && index < NSec());
    return distance;
  }

  double DistanceToOut(V3 point, V3 dir, double stepMax) const override {
    if (NSec() == 1) return sec[0].DistanceToOut(Shift(point, shift[0]), dir, stepMax);
    const int indexLow = SectionIndex(point.z - kTolerance);
    const int indexHigh = SectionIndex(point.z + kTolerance);
    int index;
    if (indexLow != indexHigh && indexLow >= 0) {
      index = sec[indexLow].Contains(Shift(point, shift[indexLow])) ? indexLow : indexHigh;
    } else {
      index = indexLow;
      if (index < 0) index = SectionIndex(point.z);
    }
    if (index < 0) return 0.;
    double totalDistance = 0.;
    int increment = (dir.z > 0) ? 1 : -1;
    if (std::fabs(dir.z) < kTolerance) increment = 0;
    int istep = 0;
    V3 pn = point;
    do {
      if ((totalDistance != 0) || (istep < 2)) {
        pn = V3{point.x + totalDistance * dir.x, point.y + totalDistance * dir.y, point.z + totalDistance * dir.z};
        pn.z -= shift[index];
        if (sec[index].Inside(pn) == kOutside) break;
      } else {
        pn.z -= shift[index];
      }
      istep++;
      const double dist = sec[index].DistanceToOut(pn, dir, stepMax);
      if (std::fabs(dist) < 0.5 * kTolerance) {
        int index1 = index;
        if ((index > 0) && (index < NSec() - 1)) {
          index1 += increment;
        } else {
          if ((index == 0) && (increment > 0)) index1 += increment;
          if ((index == NSec() - 1) && (increment < 0)) index1 += increment;
        }
        const double s = totalDistance + dist;
        V3 pte{point.x + s * dir.x, point.y + s * dir.y, point.z + s * dir.z};
        pte.z -= shift[index1];
        if (!sec[index1].Contains(pte)) break;
      }
      totalDistance += dist;
      index += increment;
    } while (increment != 0 && index >= 0 && index < NSec());
    return totalDistance;
  }

  double SafetyToIn(V3 p) const override {
    int index = SectionIndex(p.z);
    bool needZ = false;
    if (index < 0) {
      needZ = true;
      if (index == -1) index = 0;
      if (index == -2) index = NSec() - 1;
    }
    if (needZ) return sec[index].SafetyToInImpl(Shift(p, shift[index]), false);
    double safety = sec[index].SafetyToInImpl(Shift(p, shift[index]), true);
    if (safety < kTolerance) return safety;
    double minSafety = safety;
    double zbase = zs[index + 1];
    for (int i = index + 1; i < NSec(); ++i) {
      const double dz = zs[i] - zbase;
      if (dz >= minSafety) break;
      safety = sec[i].SafetyToInImpl(Shift(p, shift[i]), false);
      if (safety < minSafety) minSafety = safety;
    }
    if (index > 0) {
      zbase = zs[index - 1];
      for (int i = index - 1; i >= 0; --i) {
        const double dz = zbase - zs[i];
        if (dz >= minSafety) break;
        safety = sec[i].SafetyToInImpl(Shift(p, shift[i]), false);
        if (safety < minSafety) minSafety = safety;
      }
    }
    return minSafety;
  }

  double SafetyToOut(V3 point) const override {
    const int index = SectionIndex(point.z);
    if (index < 0) return 0;
    double minSafety = sec[index].SafetyToOut(Shift(point, shift[index]));
    if (minSafety == kInfinity) return 0.;
    if (minSafety < kTolerance) return 0.;
    double zbase = zs[index + 1];
    for (int i = index + 1; i < NSec(); ++i) {
      const double dz = zs[i] - zbase;
      if (dz >= minSafety) break;
      const double safety = sec[i].SafetyToInImpl(Shift(point, shift[i]), false);
      if (safety < minSafety) minSafety = safety;
    }
    if (index > 0) {
      zbase = zs[index - 1];
      for (int i = index - 1; i >= 0; --i) {
        const double dz = zbase - zs[i];
        if (dz >= minSafety) break;
        const double safety = sec[i].SafetyToInImpl(Shift(point, shift[i]), false);
        if (safety < minSafety) minSafety = safety;
      }
    }
    return minSafety;
  }

  void Extent(V3 &lo, V3 &hi) const override {
    double maxR = 0;
    for (const Cone &c : sec) {
      if (maxR < c.rmax1) maxR = c.rmax1;
      if (maxR < c.rmax2) maxR = c.rmax2;
    }
    lo = V3{-maxR, -maxR, zs[0]};
    hi = V3{maxR, maxR, zs[NSec()]};
  }
};

}  // namespace

Shape *MakePolycone(const double *p, int n) {
  auto *s = new Polycone;
  if (!s->Init(p, n)) {
    delete s;
    return nullptr;
  }
  return s;
}

Shape *MakeCone(const double *p, int n) {
  if (n != 7) return nullptr;
  return new Cone(p);
}

}  // namespace vgo
