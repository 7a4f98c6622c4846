// Cone (frustum with inner and outer conical surfaces and a phi section):
// parameter block and device kernels with the semantics of
// volumes/kernel/ConeImplementation.h:84-1996 (the USolids algorithms the
// reference dispatches DistanceToIn/Out and the safeties to) and the cached
// values of volumes/UnplacedCone.h:131-197.
//
// DistanceToOut's phi part uses atan2 of the direction from the device math
// library; it only decides which side of the phi range a direction points to
// within a 5e-10 rad tolerance, so an ulp of difference from the host libm
// changes nothing but directions within an ulp of that tolerance.
#pragma once
#include "shapes_box_tube.hpp"
#include "vg_math.hpp"

namespace vga {

struct ConeP {
  double rmin1, rmax1, rmin2, rmax2, dz, sphi, dphi;
  double innerSlope, outerSlope, innerOffset, outerOffset;
  double secRMin, secRMax, invSecRMin, invSecRMax, tanRMin, tanRMax;
  double cosCPhi, sinCPhi, cosSPhi, sinSPhi, cosEPhi, sinEPhi, cosHDPhi, cosHDPhiIT, cosHDPhiOT;
  double fullPhi;        // 1: dphi == 2 pi exactly (IsFullPhi)
  double phiTreatment;   // 1: dphi < 2 pi (Contains/Inside wedge)
  double rminTreatment;  // 1: rmin1 > 0 or rmin2 > 0
  double pad;
  WedgeP w;
  double pad2;
};
static_assert(sizeof(ConeP) % 16 == 0, "ConeP keeps 16-byte multiples");

struct ConeS {
  using P = ConeP;
  static constexpr double kRadTol = 1e-9;
  static constexpr double kHalfCarTol = 0.5e-9;
  static constexpr double kHalfRadTol = 0.5e-9;
  static constexpr double kHalfAngTol = 0.5e-9;

  template <bool ForInside>
  static VGA_HD void Kernel(const P &c, V3 p, bool &cin, bool &cout) {
    const double absz = Abs(p.z);
    cout = absz > PlusTol<ForInside>(c.dz);
    if (ForInside) cin = absz < MinusTol<ForInside>(c.dz);
    if (cout) return;
    const double r2 = p.x * p.x + p.y * p.y;
    const double rmax = c.outerSlope * p.z + c.outerOffset;
    const double rmax2 = rmax * rmax;
    cout |= r2 > PlusTolSq<ForInside>(rmax, rmax2);
    if (ForInside) cin &= r2 < MinusTolSq<ForInside>(rmax, rmax2);
    if (cout) return;
    if (c.rminTreatment != 0) {
      const double rmin = c.innerSlope * p.z + c.innerOffset;
      const double rmin2 = rmin * rmin;
      cout |= r2 <= MinusTolSq<ForInside>(rmin, rmin2);
      if (ForInside) cin &= r2 > PlusTolSq<ForInside>(rmin, rmin2);
      if (cout) return;
    }
    if (c.phiTreatment != 0) {
      bool wi, wo;
      WedgeS::Kernel<ForInside>(c.w, p, wi, wo);
      cout |= wo;
      if (ForInside) cin &= wi;
    }
  }
  static VGA_HD bool Contains(const P &c, V3 p) {
    bool cin = false, cout;
    Kernel<false>(c, p, cin, cout);
    return !cout;
  }
  static VGA_HD int Inside(const P &c, V3 p) {
    bool cin = false, cout;
    Kernel<true>(c, p, cin, cout);
    int r = kSurface;
    if (cout) r = kOutside;
    if (cin) r = kInside;
    return r;
  }

  static VGA_HD bool PhiAccept(const P &c, double xi, double yi, double r, double cosLimit) {
    const double cosPsi = (xi * c.cosCPhi + yi * c.sinCPhi) / r;
    return cosPsi >= cosLimit;
  }
  static VGA_HD bool InnerNormalAgainst(const P &c, double xi, double yi, V3 v) {
    const double risec = Sqrt(xi * xi + yi * yi) * c.secRMin;
    const V3 n{-xi / risec, -yi / risec, c.tanRMin / c.secRMin};
    return Dot(n, v) <= 0;
  }
  static VGA_HD double RootFar(double b, double c, double d) { return b > 0 ? c / (-b - Sqrt(d)) : -b + Sqrt(d); }
  static VGA_HD double RootNear(double b, double c, double d) { return b > 0 ? -b - Sqrt(d) : c / (-b + Sqrt(d)); }

  // An inner-cone hit at distance sd: returns true if DistanceToIn returns
  // sd at once, or records it in snxt (phi-sectioned cones).
  static VGA_HD bool InnerCandidate(const P &c, V3 p, V3 v, double sd, double ri, double cosLimit, double &snxt) {
    const double xi = p.x + sd * v.x, yi = p.y + sd * v.y;
    if (c.fullPhi == 0) {
      if (PhiAccept(c, xi, yi, ri, cosLimit)) {
        if (sd > kHalfRadTol || InnerNormalAgainst(c, xi, yi, v)) snxt = sd;
      }
      return false;
    }
    return sd > kHalfRadTol || InnerNormalAgainst(c, xi, yi, v);
  }

  static VGA_HD void PhiPlaneIn(const P &c, V3 p, V3 v, double comp, double dist, bool endPlane, double rMinOAv,
                                double rMaxOAv, double tolODz, double &snxt) {
    if (!(comp < 0)) return;
    if (!(dist < kHalfCarTol)) return;
    double sd = dist / comp;
    if (!(sd < snxt)) return;
    if (sd < 0) sd = 0.0;
    const double zi = p.z + sd * v.z;
    if (!(Abs(zi) <= tolODz)) return;
    const double xi = p.x + sd * v.x, yi = p.y + sd * v.y;
    const double rhoi2 = xi * xi + yi * yi;
    const double tolORMin2 = (rMinOAv + zi * c.tanRMin) * (rMinOAv + zi * c.tanRMin);
    const double tolORMax2 = (rMaxOAv + zi * c.tanRMax) * (rMaxOAv + zi * c.tanRMax);
    if ((rhoi2 >= tolORMin2) && (rhoi2 <= tolORMax2)) {
      const double side = yi * c.cosCPhi - xi * c.sinCPhi;
      if (endPlane ? (side >= 0.0) : (side <= 0)) snxt = sd;
    }
  }

  static VGA_HD double DistanceToIn(const P &c, V3 p, V3 v, double) {
    double snxt = kInfinity;
    const bool fullPhi = c.fullPhi != 0;
    const double rMinAv = (c.rmin1 + c.rmin2) * 0.5;
    const double rMinOAv = rMinAv > kHalfRadTol ? rMinAv - kHalfRadTol : 0.0;
    const double rMaxAv = (c.rmax1 + c.rmax2) * 0.5;
    const double rMaxOAv = rMaxAv + kHalfRadTol;
    const double tolIDz = c.dz - kHalfCarTol;
    const double tolODz = c.dz + kHalfCarTol;
    double sd, xi, yi, zi, ri, b, cc, d;

    if (Abs(p.z) >= tolIDz) {
      if (!(p.z * v.z < 0)) return snxt;
      sd = (Abs(p.z) - c.dz) / Abs(v.z);
      if (sd < 0.0) sd = 0.0;
      xi = p.x + sd * v.x;
      yi = p.y + sd * v.y;
      const double rhoi2 = xi * xi + yi * yi;
      const double rminE = v.z > 0 ? c.rmin1 : c.rmin2, rmaxE = v.z > 0 ? c.rmax1 : c.rmax2;
      const double tolORMin = rminE - kHalfRadTol * c.secRMin;
      const double tolIRMin = rminE + kHalfRadTol * c.secRMin;
      const double tolIRMax = rmaxE - kHalfRadTol * c.secRMin;
      const double tolIRMin2 = tolORMin > 0 ? tolIRMin * tolIRMin : 0.0;
      const double tolIRMax2 = tolIRMax > 0 ? tolIRMax * tolIRMax : 0.0;
      if ((tolIRMin2 <= rhoi2) && (rhoi2 <= tolIRMax2)) {
        if (!fullPhi && rhoi2 != 0) {
          if (PhiAccept(c, xi, yi, Sqrt(rhoi2), c.cosHDPhiIT)) return sd;
        } else {
          return sd;
        }
      }
    }

    const double t1 = 1.0 - v.z * v.z;
    const double t2 = p.x * v.x + p.y * v.y;
    const double t3 = p.x * p.x + p.y * p.y;
    const double rin = c.tanRMin * p.z + rMinAv;
    const double rout = c.tanRMax * p.z + rMaxAv;
    double nt1 = t1 - (c.tanRMax * v.z) * (c.tanRMax * v.z);
    double nt2 = t2 - c.tanRMax * v.z * rout;
    double nt3 = t3 - rout * rout;
    if (Abs(nt1) > kRadTol) {
      b = nt2 / nt1;
      cc = nt3 / nt1;
      d = b * b - cc;
      if ((nt3 > rout * rout * kRadTol * kRadTol * c.secRMax * c.secRMax) || (rout < 0)) {
        if (d >= 0) {
          if ((rout < 0) && (nt3 <= 0)) {
            sd = RootFar(b, cc, d);
          } else if ((b <= 0) && (cc >= 0)) {
            sd = cc / (-b + Sqrt(d));
          } else if (cc <= 0) {
            sd = -b + Sqrt(d);
          } else {
            return kInfinity;
          }
          if (sd > 0) {
            zi = p.z + sd * v.z;
            if (Abs(zi) <= tolODz) {
              if (fullPhi) return sd;
              xi = p.x + sd * v.x;
              yi = p.y + sd * v.y;
              ri = rMaxAv + zi * c.tanRMax;
              if (PhiAccept(c, xi, yi, ri, c.cosHDPhiIT)) return sd;
            }
          }
        }
      } else {
        const double rinT = rin + kHalfRadTol * c.secRMin;
        if ((t3 > rinT * rinT) && (nt2 < 0) && (d >= 0) && (Abs(p.z) <= tolIDz)) {
          const double risec = Sqrt(p.x * p.x + p.y * p.y) * c.secRMax;
          const V3 n{p.x / risec, p.y / risec, -c.tanRMax / c.secRMax};
          if (!fullPhi) {
            if (PhiAccept(c, p.x, p.y, Sqrt(t3), c.cosHDPhiIT) && Dot(n, v) <= 0) return 0.0;
          } else if (Dot(n, v) <= 0) {
            return 0.0;
          }
        }
      }
    } else if (Abs(nt2) > kRadTol) {
      sd = -0.5 * nt3 / nt2;
      if (sd < 0) return kInfinity;
      zi = p.z + sd * v.z;
      if ((Abs(zi) <= tolODz) && (nt2 < 0)) {
        if (fullPhi) return sd;
        xi = p.x + sd * v.x;
        yi = p.y + sd * v.y;
        ri = rMaxAv + zi * c.tanRMax;
        if (PhiAccept(c, xi, yi, ri, c.cosHDPhiIT)) return sd;
      }
    }

    if (rMinAv != 0) {
      nt1 = t1 - (c.tanRMin * v.z) * (c.tanRMin * v.z);
      nt2 = t2 - c.tanRMin * v.z * rin;
      nt3 = t3 - rin * rin;
      if (nt1 != 0) {
        if (nt3 > rin * kRadTol * c.secRMin) {
          b = nt2 / nt1;
          cc = nt3 / nt1;
          d = b * b - cc;
          if (d >= 0) {
            sd = RootFar(b, cc, d);
            if (sd >= 0) {
              zi = p.z + sd * v.z;
              if (Abs(zi) <= tolODz) {
                ri = rMinAv + zi * c.tanRMin;
                if (InnerCandidate(c, p, v, sd, ri, c.cosHDPhiIT, snxt)) return sd;
              }
            }
          }
        } else if (nt3 < -rin * kRadTol * c.secRMin) {
          b = nt2 / nt1;
          cc = nt3 / nt1;
          d = b * b - cc;
          if (d >= 0) {
            sd = RootFar(b, cc, d);
            zi = p.z + sd * v.z;
            ri = rMinAv + zi * c.tanRMin;
            if (ri > 0) {
              if ((sd >= 0) && (Abs(zi) <= tolODz)) {
                if (InnerCandidate(c, p, v, sd, ri, c.cosHDPhiOT, snxt)) return sd;
              }
            } else {
              sd = RootNear(b, cc, d);
              zi = p.z + sd * v.z;
              ri = rMinAv + zi * c.tanRMin;
              if ((sd >= 0) && (ri > 0) && (Abs(zi) <= tolODz)) {
                if (InnerCandidate(c, p, v, sd, ri, c.cosHDPhiIT, snxt)) return sd;
              }
            }
          }
        } else if (Abs(p.z) <= tolODz) {
          if (nt2 > 0) {
            if (!fullPhi) {
              if (PhiAccept(c, p.x, p.y, Sqrt(t3), c.cosHDPhiIT)) return 0.0;
            } else {
              return 0.0;
            }
          } else {
            b = nt2 / nt1;
            cc = nt3 / nt1;
            d = b * b - cc;
            if (d >= 0) {
              sd = RootNear(b, cc, d);
              zi = p.z + sd * v.z;
              ri = rMinAv + zi * c.tanRMin;
              if (!(ri > 0)) return kInfinity;
              sd = RootFar(b, cc, d);
              zi = p.z + sd * v.z;
              if ((sd >= 0) && (Abs(zi) <= tolODz)) {
                if (fullPhi) return sd;
                xi = p.x + sd * v.x;
                yi = p.y + sd * v.y;
                ri = rMinAv + zi * c.tanRMin;
                if (PhiAccept(c, xi, yi, ri, c.cosHDPhiIT)) snxt = sd;
              }
            }
          }
        } else {
          b = nt2 / nt1;
          cc = nt3 / nt1;
          d = b * b - cc;
          if (d > 0) {
            sd = RootFar(b, cc, d);
            zi = p.z + sd * v.z;
            if ((sd >= 0) && (Abs(zi) <= tolODz)) {
              if (fullPhi) return sd;
              xi = p.x + sd * v.x;
              yi = p.y + sd * v.y;
              ri = rMinAv + zi * c.tanRMin;
              if (PhiAccept(c, xi, yi, ri, c.cosHDPhiIT)) snxt = sd;
            }
          }
        }
      }
    }

    if (!fullPhi) {
      PhiPlaneIn(c, p, v, v.x * c.sinSPhi - v.y * c.cosSPhi, p.y * c.cosSPhi - p.x * c.sinSPhi, false, rMinOAv,
                 rMaxOAv, tolODz, snxt);
      PhiPlaneIn(c, p, v, -(v.x * c.sinEPhi - v.y * c.cosEPhi), -(p.y * c.cosEPhi - p.x * c.sinEPhi), true, rMinOAv,
                 rMaxOAv, tolODz, snxt);
    }
    if (snxt < kHalfCarTol) snxt = 0.;
    return snxt;
  }

  static VGA_HD double ATan2(double y, double x) {
    if (x != 0) return atan2(y, x);
    if (y > 0) return kPi / 2;
    if (y < 0) return -kPi / 2;
    return 0;
  }

  static VGA_HD double PhiOut(const P &c, V3 p, V3 v) {
    double vphi = ATan2(v.y, v.x);
    if (vphi < c.sphi - kHalfAngTol) {
      vphi += 2 * kPi;
    } else if (vphi > c.sphi + c.dphi + kHalfAngTol) {
      vphi -= 2 * kPi;
    }
    const bool dirInside = (c.sphi - kHalfAngTol <= vphi) && (c.sphi + c.dphi + kHalfAngTol >= vphi);
    if (!(p.x != 0 || p.y != 0)) return dirInside ? kInfinity : 0.0;
    const double pDistS = p.x * c.sinSPhi - p.y * c.cosSPhi;
    const double pDistE = -p.x * c.sinEPhi + p.y * c.cosEPhi;
    const double compS = -c.sinSPhi * v.x + c.cosSPhi * v.y;
    const double compE = c.sinEPhi * v.x - c.cosEPhi * v.y;
    const bool insidePlanes = ((c.dphi <= kPi) && ((pDistS <= kHalfCarTol) && (pDistE <= kHalfCarTol))) ||
                              ((c.dphi > kPi) && !((pDistS > kHalfCarTol) && (pDistE > kHalfCarTol)));
    if (!insidePlanes) return kInfinity;
    double s = kInfinity;
    if (compS < 0) {
      s = pDistS / compS;
      if (s >= -kHalfCarTol) {
        const double xi = p.x + s * v.x, yi = p.y + s * v.y;
        if ((Abs(xi) <= kTolerance) && (Abs(yi) <= kTolerance)) {
          if (dirInside) s = kInfinity;
        } else if ((yi * c.cosCPhi - xi * c.sinCPhi) >= 0) {
          s = kInfinity;
        } else if (pDistS > -kHalfCarTol) {
          s = 0.0;
        }
      } else {
        s = kInfinity;
      }
    }
    if (compE < 0) {
      const double s2 = pDistE / compE;
      if ((s2 > -kHalfCarTol) && (s2 < s)) {
        const double xi = p.x + s2 * v.x, yi = p.y + s2 * v.y;
        const bool leaves = ((Abs(xi) <= kTolerance) && (Abs(yi) <= kTolerance))
                                ? !dirInside
                                : (yi * c.cosCPhi - xi * c.sinCPhi >= 0);
        if (leaves) s = (pDistE <= -kHalfCarTol) ? s2 : 0.0;
      }
    }
    return s;
  }

  static VGA_HD double DistanceToOut(const P &c, V3 p, V3 v, double) {
    double snxt, srd, pdist;
    if (v.z > 0.0) {
      pdist = c.dz - p.z;
      if (!(pdist > kHalfCarTol)) return 0.0;
      snxt = pdist / v.z;
    } else if (v.z < 0.0) {
      pdist = c.dz + p.z;
      if (!(pdist > kHalfCarTol)) return 0.0;
      snxt = -pdist / v.z;
    } else {
      snxt = kInfinity;
    }
    const double rMaxAv = (c.rmax1 + c.rmax2) * 0.5;
    const double t1 = 1.0 - v.z * v.z;
    const double t2 = p.x * v.x + p.y * v.y;
    const double t3 = p.x * p.x + p.y * p.y;
    const double rout = c.tanRMax * p.z + rMaxAv;
    double nt1 = t1 - (c.tanRMax * v.z) * (c.tanRMax * v.z);
    double nt2 = t2 - c.tanRMax * v.z * rout;
    double nt3 = t3 - rout * rout;
    double deltaRoi2;
    if (v.z > 0.0) {
      deltaRoi2 = snxt * snxt * t1 + 2 * snxt * t2 + t3 - c.rmax2 * (c.rmax2 + kRadTol * c.secRMax);
    } else if (v.z < 0.0) {
      deltaRoi2 = snxt * snxt * t1 + 2 * snxt * t2 + t3 - c.rmax1 * (c.rmax1 + kRadTol * c.secRMax);
    } else {
      deltaRoi2 = 1.0;
    }
    double b, cc, d, zi, ri;
    if (nt1 != 0 && (deltaRoi2 > 0.0)) {
      b = nt2 / nt1;
      cc = nt3 / nt1;
      d = b * b - cc;
      if (!(d >= 0)) return 0.0;
      if (nt3 > -kHalfRadTol && nt2 >= 0) return 0.0;
      srd = b > 0 ? -b - Sqrt(d) : cc / (-b + Sqrt(d));
      zi = p.z + srd * v.z;
      ri = c.tanRMax * zi + rMaxAv;
      if ((ri < 0) || (srd < kHalfRadTol)) {
        const double sr2 = b > 0 ? cc / (-b - Sqrt(d)) : -b + Sqrt(d);
        zi = p.z + sr2 * v.z;
        ri = c.tanRMax * zi + rMaxAv;
        srd = ((ri >= 0) && (sr2 > kHalfRadTol)) ? sr2 : kInfinity;
      }
    } else if (nt2 != 0 && (deltaRoi2 > 0.0)) {
      return 0.0;
    } else {
      srd = kInfinity;
    }
    if (c.rmin1 != 0 || c.rmin2 != 0) {
      nt1 = t1 - (c.tanRMin * v.z) * (c.tanRMin * v.z);
      if (nt1 != 0) {
        const double rMinAv = (c.rmin1 + c.rmin2) * 0.5;
        const double rin = c.tanRMin * p.z + rMinAv;
        nt2 = t2 - c.tanRMin * v.z * rin;
        nt3 = t3 - rin * rin;
        b = nt2 / nt1;
        cc = nt3 / nt1;
        d = b * b - cc;
        if (d >= 0.0) {
          if (nt3 < kRadTol * (rin + kRadTol * 0.25)) {
            if (nt2 < 0.0) return 0.0;
          } else {
            const double sr2 = b > 0 ? -b - Sqrt(d) : cc / (-b + Sqrt(d));
            zi = p.z + sr2 * v.z;
            ri = c.tanRMin * zi + rMinAv;
            if ((ri < 0) || (sr2 < kHalfRadTol)) {
              const double sr3 = b > 0 ? cc / (-b - Sqrt(d)) : -b + Sqrt(d);
              if (sr3 > kHalfRadTol && sr3 < srd) {
                zi = p.z + sr3 * v.z;
                ri = c.tanRMin * zi + rMinAv;
                if (ri >= 0.0) srd = sr3;
              }
            } else if ((sr2 < srd) && (sr2 > kHalfCarTol)) {
              srd = sr2;
            }
          }
        }
      }
    }
    if (c.fullPhi == 0) {
      const double sphi = PhiOut(c, p, v);
      if (sphi < snxt) snxt = sphi;
    }
    if (srd < snxt) snxt = srd;
    if (snxt < kHalfCarTol) snxt = 0.;
    return snxt;
  }

  static VGA_HD double SafetyToIn(const P &c, V3 p) { return SafetyToInImpl(c, p, false); }

  // SafetyToInUSOLIDS<Backend, ForPolycone> (ConeImplementation.h:1908-1996):
  // a polycone asks for the radial and phi parts only of the section that
  // holds the point.
  static VGA_HD double SafetyToInImpl(const P &c, V3 p, bool forPolycone) {
    double safe;
    const double rho = Sqrt(p.x * p.x + p.y * p.y);
    const double safeZ = Abs(p.z) - c.dz;
    if (c.rmin1 != 0 || c.rmin2 != 0) {
      const double pRMin = c.tanRMin * p.z + (c.rmin1 + c.rmin2) * 0.5;
      const double safeR1 = (pRMin - rho) * c.invSecRMin;
      const double pRMax = c.tanRMax * p.z + (c.rmax1 + c.rmax2) * 0.5;
      const double safeR2 = (rho - pRMax) * c.invSecRMax;
      safe = safeR1 > safeR2 ? safeR1 : safeR2;
    } else {
      const double pRMax = c.tanRMax * p.z + (c.rmax1 + c.rmax2) * 0.5;
      safe = (rho - pRMax) * c.invSecRMax;
    }
    if (!forPolycone && safeZ > safe) safe = safeZ;
    if (c.fullPhi == 0 && rho != 0) {
      const double cosPsi = p.x * c.cosCPhi + p.y * c.sinCPhi;
      if (cosPsi < c.cosHDPhi * rho) {
        const double safePhi = ((p.y * c.cosCPhi - p.x * c.sinCPhi) <= 0.0) ? Abs(p.x * c.sinSPhi - p.y * c.cosSPhi)
                                                                            : Abs(p.x * c.sinEPhi - p.y * c.cosEPhi);
        if (safePhi > safe) safe = safePhi;
      }
    }
    if (safe < 0.0) safe = 0.0;
    return safe;
  }

  static VGA_HD double SafetyToOut(const P &c, V3 p) {
    const double rho = Sqrt(p.x * p.x + p.y * p.y);
    const double safeZ = c.dz - Abs(p.z);
    double safeR1 = kInfinity;
    if (c.rmin1 != 0 || c.rmin2 != 0) {
      const double pRMin = c.tanRMin * p.z + (c.rmin1 + c.rmin2) * 0.5;
      safeR1 = (rho - pRMin) * c.invSecRMin;
    }
    const double pRMax = c.tanRMax * p.z + (c.rmax1 + c.rmax2) * 0.5;
    const double safeR2 = (pRMax - rho) * c.invSecRMax;
    double safe = safeR1 < safeR2 ? safeR1 : safeR2;
    if (safeZ < safe) safe = safeZ;
    if (c.fullPhi == 0) {
      const double safePhi = ((p.y * c.cosCPhi - p.x * c.sinCPhi) <= 0) ? -(p.x * c.sinSPhi - p.y * c.cosSPhi)
                                                                      : (p.x * c.sinEPhi - p.y * c.cosEPhi);
      if (safePhi < safe) safe = safePhi;
    }
    if (safe < 0) safe = 0;
    return safe;
  }
};

}  // namespace vga
