// Sphere (rmin/rmax shell with phi and theta sections): parameter block and
// device kernels with the semantics of
// volumes/kernel/SphereImplementation.h:257-1520 and the cached values of
// volumes/UnplacedSphere.h:63-215.
//
// Point-dependent transcendentals (atan2, asin, sin) use the device math
// library; their last-ulp differences from the host libm can move safeties
// by an ulp and can flip a classification only for points within an ulp of
// an angular tolerance boundary.  Transcendentals of constants are computed
// on the host.
#pragma once
#include "vg_math.hpp"

namespace vga {

struct SphereP {
  double fRmin, fRmax, fSPhi, fDPhi, fSTheta, fDTheta, fRminTolerance;
  double sinCPhi, cosCPhi, sinSPhi, cosSPhi, sinEPhi, cosEPhi, hDPhi, cPhi, ePhi;
  double sinSTheta, cosSTheta, sinETheta, cosETheta, eTheta;
  double cosHDPhi;      // std::cos(hDPhi) (SafetyToIn)
  double sinPhim, cosPhim;  // sin/cos of 0.5*(sphi+ephi) (DistanceToOut)
  double fullPhi, fullTheta;  // 0/1
};

struct SphereS {
  using P = SphereP;
  static constexpr double kAngTol = 1e-9;

  static VGA_HD double ATan2(double y, double x) {
    if (x != 0) return atan2(y, x);
    if (y > 0) return kPi / 2;
    if (y < 0) return -kPi / 2;
    return 0;
  }
  static VGA_HD void Kernel(const P &s, V3 p, bool &cin, bool &cout) {
    const double half = 0.5 * kAngTol;
    const double rad2 = p.x * p.x + p.y * p.y + p.z * p.z;
    double tolRMin = s.fRmin + (0.5 * s.fRminTolerance);
    double tolRMax = s.fRmax - (0.5 * s.fRminTolerance);
    cin = (rad2 <= tolRMax * tolRMax) && (rad2 >= tolRMin * tolRMin);
    tolRMin = s.fRmin - (0.5 * s.fRminTolerance);
    tolRMax = s.fRmax + (0.5 * s.fRminTolerance);
    cout = (rad2 <= tolRMin * tolRMin) || (rad2 >= tolRMax * tolRMax);
    if (cout) return;
    if (s.fullPhi == 0) {
      double pPhi = ATan2(p.y, p.x);
      const double eph = s.fSPhi + s.fDPhi;
      if (pPhi < (s.fSPhi - half)) pPhi = pPhi + (2. * kPi);
      if (pPhi > (eph + half)) pPhi = pPhi - (2. * kPi);
      cin &= (pPhi <= eph - half) && (pPhi >= s.fSPhi + half);
      cout |= (pPhi < s.fSPhi - half) || (pPhi > eph + half);
      if (cout) return;
    }
    if (s.fullTheta == 0) {
      const double pTheta = ATan2(Sqrt(p.x * p.x + p.y * p.y), p.z);
      const double eth = s.fSTheta + s.fDTheta;
      cin &= (pTheta <= eth - half) && (pTheta >= s.fSTheta + half);
      cout |= (pTheta < s.fSTheta - half) || (pTheta > eth + half);
    }
  }
  static VGA_HD bool Contains(const P &s, V3 p) {
    bool cin, cout;
    Kernel(s, p, cin, cout);
    return !cout;
  }
  static VGA_HD int Inside(const P &s, V3 p) {
    bool cin, cout;
    Kernel(s, p, cin, cout);
    int r = kSurface;
    if (cout) r = kOutside;
    if (cin) r = kInside;
    return r;
  }
  // The reference's DistanceToIn body is empty: the caller's kInfinity.
  static VGA_HD double DistanceToIn(const P &, V3, V3, double) { return kInfinity; }

  static VGA_HD double SafetyToIn(const P &s, V3 p) {
    double safety = kInfinity;
    const double rad = Sqrt(p.x * p.x + p.y * p.y + p.z * p.z);
    const double rho = Sqrt(p.x * p.x + p.y * p.y);
    const double safeRMin = s.fRmin - rad, safeRMax = rad - s.fRmax;
    if (s.fRmin > 0)
      safety = (safeRMin > safeRMax) ? safeRMin : safeRMax;
    else
      safety = safeRMax;
    if (s.fullPhi == 0 && (rho > 0)) {
      const double cosPsi = (p.x * s.cosCPhi + p.y * s.sinCPhi) / rho;
      if (cosPsi < s.cosHDPhi) {
        double test1 = (p.x * s.sinSPhi - p.y * s.cosSPhi);
        if (test1 < 0) test1 = -1. * test1;
        double test2 = (p.x * s.sinEPhi - p.y * s.cosEPhi);
        if (test2 < 0) test2 = -1. * test2;
        const double safePhi = ((p.y * s.cosCPhi - p.x * s.sinCPhi) <= 0) ? test1 : test2;
        if (safePhi > safety) safety = safePhi;
      }
    }
    const double rds = Sqrt(p.x * p.x + p.y * p.y + p.z * p.z);
    if (s.fullTheta == 0 && (rds != 0)) {
      double pTheta = (kPi / 2) - asin(p.z / rds);
      if (pTheta < 0) pTheta = pTheta + kPi;
      const double dTheta1 = s.fSTheta - pTheta, dTheta2 = pTheta - s.eTheta;
      if (dTheta1 > dTheta2) {
        if (dTheta1 >= 0) {
          const double st = rds * sin(dTheta1);
          if (safety <= st) safety = st;
        }
      } else if (dTheta2 >= 0) {
        const double st = rds * sin(dTheta2);
        if (safety <= st) safety = st;
      }
    }
    if (safety < 0) safety = 0;
    return safety;
  }
  static VGA_HD double SafetyToOut(const P &s, V3 p) {
    double safety;
    const double rad = Sqrt(p.x * p.x + p.y * p.y + p.z * p.z);
    const double rho = Sqrt(p.x * p.x + p.y * p.y);
    const double safeRMin = rad - s.fRmin, safeRMax = s.fRmax - rad;
    if (s.fRmin > 0)
      safety = (safeRMin < safeRMax) ? safeRMin : safeRMax;
    else
      safety = safeRMax;
    if (s.fullPhi == 0 && (rho > 0)) {
      const double test1 = (p.y * s.cosCPhi - p.x * s.sinCPhi);
      const double safePhi =
          (test1 <= 0) ? -1. * (p.x * s.sinSPhi - p.y * s.cosSPhi) : (p.x * s.sinEPhi - p.y * s.cosEPhi);
      if (safePhi < safety) safety = safePhi;
    }
    const double rds = Sqrt(p.x * p.x + p.y * p.y + p.z * p.z);
    if (s.fullTheta == 0 && (rds > 0)) {
      double pTheta = (kPi / 2) - asin(p.z / rds);
      if (pTheta < 0) pTheta = pTheta + kPi;
      const double eth = s.fSTheta + s.fDTheta;
      const double dTheta1 = pTheta - s.fSTheta, dTheta2 = eth - pTheta;
      const double safeTheta = (dTheta1 < dTheta2) ? rds * sin(dTheta1) : rds * sin(dTheta2);
      if (safety > safeTheta) safety = safeTheta;
    }
    if (safety < 0) safety = 0;
    return safety;
  }
  static VGA_HD double ToSphere(const P &s, V3 p, V3 d, double radius, bool check, bool firstcross) {
    const double rad2 = p.x * p.x + p.y * p.y + p.z * p.z;
    const double pDotV = p.x * d.x + p.y * d.y + p.z * d.z;
    const double c = rad2 - radius * radius;
    const double d2 = pDotV * pDotV - c;
    if (d2 <= 0) return kInfinity;
    const double dd = Sqrt(d2);
    double distance;
    if (c <= 0 || !firstcross)
      distance = -1. * pDotV + dd;
    else
      distance = -1. * pDotV - dd;
    if (distance < 0) distance = kInfinity;
    if (check) {
      const V3 pt{p.x + distance * d.x, p.y + distance * d.y, p.z + distance * d.z};
      bool cin, cout;
      Kernel(s, pt, cin, cout);
      if (cout) distance = kInfinity;
    }
    return distance;
  }
  static VGA_HD void ToCone(V3 p, V3 d, double dzv, double r1v, double r2v, double &bv, double &deltav) {
    deltav = -1.;
    if (dzv < 0) return;
    const double ro0v = (0.5 * (r1v + r2v));
    const double tzv = (0.5 * (r2v - r1v) / dzv);
    const double rsqv = p.x * p.x + p.y * p.y;
    const double rcv = ro0v + p.z * tzv;
    double av = d.x * d.x + d.y * d.y - tzv * tzv * d.z * d.z;
    bv = p.x * d.x + p.y * d.y - tzv * rcv * d.z;
    double cv = rsqv - rcv * rcv;
    const double t1 = Abs(av), t2 = Abs(bv);
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
    deltav = (deltav > 0) ? Sqrt(deltav) : -1.;
  }
  static VGA_HD bool PhiOk(const P &s, double x, double y) {
    double ddp = ATan2(y, x) - s.fSPhi;
    if (ddp < 0) ddp = ddp + 2 * kPi;
    return ddp <= (s.ePhi - s.fSPhi);
  }
  // One theta-cone block; false means "return 0" (on the cone, moving out).
  template <bool first>
  static VGA_HD bool ThetaBlock(const P &s, V3 p, V3 d, double rxy2, bool cond, double si, double ci, double &sn) {
    const double toler = 1.E-10, toler2 = 1.E+10;
    double r1 = 0, r2 = 0, z1 = 0, z2 = 0;
    if (cond) {
      if (ci > 0) {
        r1 = s.fRmin * si; z1 = s.fRmin * ci; r2 = s.fRmax * si; z2 = s.fRmax * ci;
      } else {
        r1 = s.fRmax * si; z1 = s.fRmax * ci; r2 = s.fRmin * si; z2 = s.fRmin * ci;
      }
    }
    const double dz = 0.5 * (z2 - z1);
    const V3 ptnew{p.x, p.y, p.z - 0.5 * (z1 + z2)};
    const double zinv = 1. / dz;
    const double rin = 0.5 * (r1 + r2 + (r2 - r1) * ptnew.z * zinv);
    const double sigz = p.z < 0 ? -1. : 1.;
    const bool bigequ = first ? ((sigz * ci > 0) && (sigz * rxy2 < sigz * rin * (rin + sigz * toler)))
                              : ((sigz * ci > 0) && (sigz * rxy2 > sigz * rin * (rin - sigz * toler)));
    if (cond && bigequ) {
      const double ddotn = ptnew.x * d.x + ptnew.y * d.y + 0.5 * (r1 - r2) * d.z * zinv * Sqrt(rxy2);
      if (first ? (sigz * ddotn <= 0) : (sigz * ddotn >= 0)) return false;
    }
    if (!(cond && !bigequ)) return true;
    double b, delta;
    ToCone(ptnew, d, dz, r1, r2, b, delta);
    if (!(delta > 0)) return true;
    double snxt = -1. * b - delta;
    double znew = ptnew.z + snxt * d.z;
    if ((snxt > 0) && (Abs(znew) < dz)) {
      if (s.fullPhi != 0 || PhiOk(s, ptnew.x + snxt * d.x, ptnew.y + snxt * d.y)) sn = snxt;
    }
    if (sn > toler2) {
      snxt = -1. * b + delta;
      znew = ptnew.z + snxt * d.z;
      if ((snxt > 0) && (Abs(znew) < dz)) {
        if (s.fullPhi != 0 || PhiOk(s, ptnew.x + snxt * d.x, ptnew.y + snxt * d.y)) sn = snxt;
      }
    }
    (void)toler;
    return true;
  }
  static VGA_HD double DistanceToOut(const P &s, V3 p, V3 d, double) {
    const double toler = 1.E-10;
    const double rad = Sqrt(p.x * p.x + p.y * p.y + p.z * p.z);
    const double pDotV = p.x * d.x + p.y * d.y + p.z * d.z;
    const double rxy2 = p.x * p.x + p.y * p.y;
    double sn1 = kInfinity, rn1 = kInfinity, sn2 = kInfinity;
    if (s.fRmin > 0 && pDotV < 0) {
      if (rad <= (s.fRmin + toler)) return kInfinity;
      rn1 = ToSphere(s, p, d, s.fRmin, true, true);
    }
    if ((rad >= (s.fRmax - toler)) && (pDotV >= 0)) return kInfinity;
    const double rn2 = ToSphere(s, p, d, s.fRmax, false, false);
    const double sr = Min(rn1, rn2);
    if (s.fullTheta == 0) {
      if (Abs(s.fSTheta - kPi / 2) < 1.e-10) {
        if ((p.z * d.z) < 0) sn1 = -1. * p.z / d.z;
      } else if (!ThetaBlock<true>(s, p, d, rxy2, s.fSTheta > 0, s.sinSTheta, s.cosSTheta, sn1)) {
        return 0.;
      }
      if (Abs(s.eTheta - kPi / 2) < 1.e-10) {
        if ((p.z * d.z) < 0) sn1 = -1. * p.z / d.z;
      } else if (!ThetaBlock<false>(s, p, d, rxy2, s.eTheta < kPi, s.sinETheta, s.cosETheta, sn2)) {
        return 0.;
      }
    }
    const double st = Min(sn1, sn2);
    double sp = kInfinity;
    if (s.fullPhi == 0) {
      // DistanceToPhiMin with in = true
      double sfi1 = kInfinity, sfi2 = kInfinity;
      double un = d.x * s.sinSPhi - d.y * s.cosSPhi;
      if (un > 0) {
        double t = (-1. * p.x * s.sinSPhi + p.y * s.cosSPhi);
        if (t >= 0) {
          t = t / un;
          if (((p.x + t * d.x) * s.sinPhim - (p.y + t * d.y) * s.cosPhim) >= 0) sfi1 = t;
        }
      }
      un = -1. * d.x * s.sinEPhi + d.y * s.cosEPhi;
      if (un > 0) {
        double t = (p.x * s.sinEPhi - p.y * s.cosEPhi);
        if (t >= 0) {
          t = t / un;
          if ((-1. * (p.x + t * d.x) * s.sinPhim + (p.y + t * d.y) * s.cosPhim) >= 0) sfi2 = t;
        }
      }
      sp = Min(sfi1, sfi2);
    }
    return Min(Min(sr, st), sp);
  }
};

}  // namespace vga
