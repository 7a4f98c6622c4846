// Orb, Trd, Parallelepiped and Paraboloid: parameter blocks and device
// kernels with the reference Scalar backend's semantics.
//
//   Orb             volumes/kernel/OrbImplementation.h:234-570,
//                   source/UnplacedOrb.cpp:48-56 (tolerances as SetRadius
//                   computes them)
//   Trd             volumes/kernel/TrdImplementation.h:20-458,
//                   volumes/UnplacedTrd.h:38-48 (UniversalTrd)
//   Parallelepiped  volumes/kernel/ParallelepipedImplementation.h:120-356,
//                   source/UnplacedParallelepiped.cpp:48-70
//   Paraboloid      volumes/kernel/ParaboloidImplementation.h:70-470,
//                   source/UnplacedParaboloid.cpp:92-127
#pragma once
#include "shapes_box_tube.hpp"

namespace vga {

// ------------------------------------------------------------------ Orb
struct OrbP {
  double r, rTolI, rTolO, pad;
};

struct OrbS {
  using P = OrbP;
  static VGA_HD bool Contains(const P &o, V3 p) {
    const double rad2 = p.x * p.x + p.y * p.y + p.z * p.z;
    return !(rad2 > o.rTolO * o.rTolO);
  }
  static VGA_HD int Inside(const P &o, V3 p) {
    const double rad2 = p.x * p.x + p.y * p.y + p.z * p.z;
    int r = kSurface;
    if (rad2 > o.rTolO * o.rTolO) r = kOutside;
    if (rad2 < o.rTolI * o.rTolI) r = kInside;
    return r;
  }
  static VGA_HD double DistanceToIn(const P &o, V3 p, V3 d, double) {
    const double rad2 = p.x * p.x + p.y * p.y + p.z * p.z;
    const double rad = Sqrt(rad2);
    const double pDotV = p.x * d.x + p.y * d.y + p.z * d.z;
    const double c = rad2 - o.r * o.r;
    const double d2 = pDotV * pDotV - c;
    if (rad < o.rTolI) return kInfinity;
    if ((rad >= o.rTolI) && (rad <= o.rTolO) && (pDotV < 0)) return 0.;
    if ((rad >= o.rTolI) && (rad <= o.rTolO) && (pDotV >= 0)) return kInfinity;
    if ((rad > o.rTolO) && (pDotV < 0)) return -pDotV - Sqrt(d2);
    return kInfinity;
  }
  static VGA_HD double DistanceToOut(const P &o, V3 p, V3 d, double) {
    const double rad2 = p.x * p.x + p.y * p.y + p.z * p.z;
    const double rad = Sqrt(rad2);
    const double pDotV = p.x * d.x + p.y * d.y + p.z * d.z;
    const double c = rad2 - o.r * o.r;
    const double d2 = pDotV * pDotV - c;
    if (rad2 > o.rTolO * o.rTolO) return 0.;
    const bool within = (rad <= o.rTolO) && (c < (kTolerance * o.r));
    const bool onTol = (c > (-2 * kTolerance * o.r)) && ((pDotV >= 0) || (d2 < 0));
    if (within && !onTol) return -pDotV + Sqrt(d2);
    return 0.;
  }
  static VGA_HD double SafetyToIn(const P &o, V3 p) {
    const double s = Sqrt(p.x * p.x + p.y * p.y + p.z * p.z) - o.r;
    return s < 0. ? 0. : s;
  }
  static VGA_HD double SafetyToOut(const P &o, V3 p) {
    const double s = o.r - Sqrt(p.x * p.x + p.y * p.y + p.z * p.z);
    return s < 0. ? 0. : s;
  }
};

// ------------------------------------------------------------------ Trd
struct TrdP {
  double dx1, dx2, dy1, dy2, dz;
  double x2minusx1, y2minusy1, halfx1plusx2, halfy1plusy2, dztimes2, fx, fy, calfx, calfy;
  double varyingY;  // dy1 != dy2
  double pad;
};

struct TrdS {
  using P = TrdP;
  template <bool surfaceT>
  static VGA_HD void Kernel(const P &t, V3 p, bool &cin, bool &cout) {
    const double pzPlusDz = p.z + t.dz;
    cout = Abs(p.z) > PlusTol<surfaceT>(t.dz);
    if (surfaceT) cin = Abs(p.z) < MinusTol<surfaceT>(t.dz);
    double cross = t.x2minusx1 * pzPlusDz - t.dztimes2 * (Abs(p.x) - t.dx1);
    cout |= PlusTol<surfaceT>(cross) < 0;
    if (surfaceT) cin &= MinusTol<surfaceT>(cross) > 0;
    cross = t.y2minusy1 * pzPlusDz - t.dztimes2 * (Abs(p.y) - t.dy1);
    cout |= PlusTol<surfaceT>(cross) < 0;
    if (surfaceT) cin &= MinusTol<surfaceT>(cross) > 0;
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
    if (cin) r = kInside;
    if (cout) r = kOutside;
    return r;
  }
  template <bool forY, bool mirrored>
  static VGA_HD void FaceHit(const P &t, V3 pos, V3 dir, double &dist, bool &ok) {
    double alongV, posV, dirV, posK, dirK, fK, halfKplus, v1;
    if (forY) {
      alongV = t.y2minusy1; v1 = t.dy1; posV = pos.y; posK = pos.x; dirV = dir.y; dirK = dir.x; fK = t.fx;
      halfKplus = t.halfx1plusx2;
    } else {
      alongV = t.x2minusx1; v1 = t.dx1; posV = pos.x; posK = pos.y; dirV = dir.x; dirK = dir.y; fK = t.fy;
      halfKplus = t.halfy1plusy2;
    }
    if (mirrored) {
      posV = -posV;
      dirV = -dirV;
    }
    const double alongZ = t.dztimes2;
    dist = (alongZ * (posV - v1) - alongV * (pos.z + t.dz)) / (dir.z * alongV - dirV * alongZ);
    ok = dist > 0;
    if (ok) {
      const double hitz = pos.z + dist * dir.z;
      ok &= Abs(hitz) <= t.dz;
      const double hitk = posK + dist * dirK;
      const double dK = halfKplus - fK * hitz;
      ok &= Abs(hitk) <= dK;
    }
  }
  static VGA_HD double DistanceToIn(const P &t, V3 pos, V3 dir, double) {
    double distance = kInfinity;
    bool okz = pos.z * dir.z < 0;
    if (okz) {
      const double distz = (Abs(pos.z) - t.dz) / Abs(dir.z);
      const double hitx = Abs(pos.x + distz * dir.x);
      const double hity = Abs(pos.y + distz * dir.y);
      const bool okzt = pos.z > t.dz && hitx <= t.dx2 && hity <= t.dy2;
      const bool okzb = pos.z < -t.dz && hitx <= t.dx1 && hity <= t.dy1;
      okz &= (okzt | okzb);
      if (okz) distance = distz;
    }
    double d;
    bool ok;
    FaceHit<false, false>(t, pos, dir, d, ok);
    if (ok && d < distance) distance = d;
    FaceHit<false, true>(t, pos, dir, d, ok);
    if (ok && d < distance) distance = d;
    if (t.varyingY != 0) {
      FaceHit<true, false>(t, pos, dir, d, ok);
      if (ok && d < distance) distance = d;
      FaceHit<true, true>(t, pos, dir, d, ok);
      if (ok && d < distance) distance = d;
    } else {
      const double disty = (Abs(pos.y) - t.dy1) / Abs(dir.y);
      const double zhit = pos.z + disty * dir.z;
      const double xhit = pos.x + disty * dir.x;
      const double dxw = t.halfx1plusx2 - t.fx * zhit;
      const bool oky = pos.y * dir.y < 0 && disty > 0 && Abs(xhit) < dxw && Abs(zhit) < t.dz;
      if (oky && disty < distance) distance = disty;
    }
    return distance;
  }
  static VGA_HD double DistanceToOut(const P &t, V3 p, V3 dir, double) {
    if (dir.z > 0) {
      const double distz = (t.dz - p.z) / Abs(dir.z);
      const double hitx = Abs(p.x + distz * dir.x);
      const double hity = Abs(p.y + distz * dir.y);
      if (hitx <= t.dx2 && hity <= t.dy2) return distz;
    }
    if (dir.z < 0) {
      const double distz = (p.z - (-t.dz)) / Abs(dir.z);
      const double hitx = Abs(p.x + distz * dir.x);
      const double hity = Abs(p.y + distz * dir.y);
      if (hitx <= t.dx1 && hity <= t.dy1) return distz;
    }
    double d;
    bool ok;
    FaceHit<false, false>(t, p, dir, d, ok);
    if (ok) return d;
    FaceHit<false, true>(t, p, dir, d, ok);
    if (ok) return d;
    double distance = kInfinity;
    if (t.varyingY != 0) {
      FaceHit<true, false>(t, p, dir, d, ok);
      if (ok) return d;
      FaceHit<true, true>(t, p, dir, d, ok);
      if (ok) distance = d;
    } else {
      const double plane = dir.y < 0 ? -t.dy1 : t.dy1;
      const double disty = (plane - p.y) / dir.y;
      const double zhit = p.z + disty * dir.z;
      const double xhit = p.x + disty * dir.x;
      const double dxw = t.halfx1plusx2 - t.fx * zhit;
      if (Abs(xhit) < dxw && Abs(zhit) < t.dz) distance = disty;
    }
    return distance;
  }
  template <bool inside>
  static VGA_HD double Safety(const P &t, V3 pos) {
    double dist = t.dz - Abs(pos.z);
    const double distx = t.halfx1plusx2 - t.fx * pos.z;
    const double safx = (distx - Abs(pos.x)) * t.calfx;
    if (distx >= 0 && safx < dist) dist = safx;
    if (t.varyingY != 0) {
      const double disty = t.halfy1plusy2 - t.fy * pos.z;
      const double safy = (disty - Abs(pos.y)) * t.calfy;
      if (disty >= 0 && safy < dist) dist = safy;
    } else {
      const double safy = t.dy1 - Abs(pos.y);
      if (safy < dist) dist = safy;
    }
    return inside ? dist : -dist;
  }
  static VGA_HD double SafetyToIn(const P &t, V3 p) { return Safety<false>(t, p); }
  static VGA_HD double SafetyToOut(const P &t, V3 p) { return Safety<true>(t, p); }
};

// ------------------------------------------------------------------ Parallelepiped
struct ParaP {
  double d[3];
  double tanAlpha, tanThetaSinPhi, tanThetaCosPhi;
  double ctx, cty;  // 1/sqrt(1+..) factors of the safeties
  double alpha, theta, phi, pad;
};

struct ParaS {
  using P = ParaP;
  static VGA_HD V3 Shear(const P &q, V3 p) {
    p.y -= q.tanThetaSinPhi * p.z;
    p.x -= q.tanThetaCosPhi * p.z + q.tanAlpha * p.y;
    return p;
  }
  static VGA_HD bool Contains(const P &q, V3 p) {
    const BoxP b{{q.d[0], q.d[1], q.d[2]}};
    return BoxS::Contains(b, Shear(q, p));
  }
  static VGA_HD int Inside(const P &q, V3 p) {
    const BoxP b{{q.d[0], q.d[1], q.d[2]}};
    return BoxS::Inside(b, Shear(q, p));
  }
  static VGA_HD double DistanceToIn(const P &q, V3 p, V3 d, double stepMax) {
    const BoxP b{{q.d[0], q.d[1], q.d[2]}};
    return BoxS::DistanceToIn(b, Shear(q, p), Shear(q, d), stepMax);
  }
  static VGA_HD double DistanceToOut(const P &q, V3 p, V3 dir, double) {
    double distance = kInfinity;
    bool inDir = dir.z > 0;
    double mx = q.d[2] - p.z;
    bool inPoint = mx > kHalfTolerance;
    bool away = inDir && !inPoint;
    if (away) return 0.;
    if (inPoint && inDir) distance = mx / dir.z;
    inDir = dir.z < 0;
    mx = -q.d[2] - p.z;
    inPoint = mx < -kHalfTolerance;
    away = inDir && !inPoint;
    if (away) return 0.;
    if (inPoint && inDir) distance = mx / dir.z;
    const double lpy = p.y - q.tanThetaSinPhi * p.z;
    const double ldy = dir.y - q.tanThetaSinPhi * dir.z;
    inDir = ldy > 0;
    mx = q.d[1] - lpy;
    inPoint = mx > kHalfTolerance;
    away = inDir && !inPoint;
    mx /= ldy;
    if (away) return 0.;
    if (inPoint && inDir && mx < distance) distance = mx;
    inDir = ldy < 0;
    mx = -q.d[1] - lpy;
    inPoint = mx < -kHalfTolerance;
    away = inDir && !inPoint;
    mx /= ldy;
    if (away) return 0.;
    if (inPoint && inDir && mx < distance) distance = mx;
    const double lpx = p.x - q.tanThetaCosPhi * p.z - q.tanAlpha * lpy;
    const double ldx = dir.x - q.tanThetaCosPhi * dir.z - q.tanAlpha * ldy;
    inDir = ldx > 0;
    mx = q.d[0] - lpx;
    inPoint = mx > kHalfTolerance;
    away = inDir && !inPoint;
    mx /= ldx;
    if (away) return 0.;
    if (inPoint && inDir && mx < distance) distance = mx;
    inDir = ldx < 0;
    mx = -q.d[0] - lpx;
    inPoint = mx < -kHalfTolerance;
    away = inDir && !inPoint;
    mx /= ldx;
    if (away) return 0.;
    if (inPoint && inDir && mx < distance) distance = mx;
    return distance;
  }
  static VGA_HD double SafetyToIn(const P &q, V3 p0) {
    const V3 p = Shear(q, p0);
    const double sx = (Abs(p.x) - q.d[0]) * q.ctx, sy = (Abs(p.y) - q.d[1]) * q.cty, sz = Abs(p.z) - q.d[2];
    double s = sx;
    if (sy > s) s = sy;
    if (sz > s) s = sz;
    return s;
  }
  static VGA_HD double SafetyToOut(const P &q, V3 p0) {
    const V3 p = Shear(q, p0);
    const double sx = (q.d[0] - Abs(p.x)) * q.ctx, sy = (q.d[1] - Abs(p.y)) * q.cty, sz = q.d[2] - Abs(p.z);
    double s = sx;
    if (sy < s) s = sy;
    if (sz < s) s = sz;
    return s;
  }
};

// ------------------------------------------------------------------ Paraboloid
struct ParabP {
  double rlo, rhi, dz, a, b, ainv, a2, rlo2, rhi2, tolOz, dx, dy;
};

struct ParabS {
  using P = ParabP;
  static VGA_HD int Inside(const P &q, V3 p) {
    const double rho2 = p.x * p.x + p.y * p.y;
    const double absZ = Abs(p.z);
    if (absZ > q.tolOz) return kOutside;
    const double value = q.a * rho2 + q.b - p.z;
    if (value > kHalfTolerance) return kOutside;
    if ((absZ < q.tolOz) && (value < -kHalfTolerance)) return kInside;
    return kSurface;
  }
  static VGA_HD bool Contains(const P &q, V3 p) {
    if (Abs(p.z) > q.dz) return false;
    const double aa = q.a * (p.z - q.b);
    const double rho2 = p.x * p.x + p.y * p.y;
    return !(aa < 0 || aa < q.a2 * rho2);
  }
  static VGA_HD double DistanceToIn(const P &q, V3 lp, V3 ld, double) {
    const double absZ = Abs(lp.z);
    const double absDirZ = Abs(ld.z);
    const double rho2 = lp.x * lp.x + lp.y * lp.y;
    const double pdx = lp.x * ld.x;
    const double pdy = lp.y * ld.y;
    const bool checkZ = lp.z * ld.z > 0;
    if (absZ > q.dz && checkZ) return kInfinity;
    if ((rho2 > q.rhi2) && (pdx > 0 && pdy > 0)) return kInfinity;
    if ((Abs(lp.x) > q.rhi) && (pdx > 0)) return kInfinity;
    if ((Abs(lp.y) > q.rhi) && (pdy > 0)) return kInfinity;
    const double distZ = (absZ - q.dz) / absDirZ;
    const double xHit = lp.x + distZ * ld.x;
    const double yHit = lp.y + distZ * ld.y;
    const double rhoHit2 = xHit * xHit + yHit * yHit;
    const double ray2 = (lp.z < -q.dz && ld.z > 0) ? q.rlo2 : q.rhi2;
    if ((absZ > q.dz) && (!checkZ) && (rhoHit2 <= ray2)) return distZ;
    const double dirRho2 = ld.x * ld.x + ld.y * ld.y;
    const double qa = q.a * dirRho2;
    const double qb = 2. * q.a * (pdx + pdy) - ld.z;
    const double qc = q.a * rho2 + q.b - lp.z;
    const double qainv = 1. / qa;
    const double t = qb * 0.5;
    double delta = t * t - qc * qa;
    if (delta < 0) return kInfinity;
    delta = Sqrt(delta);
    const double distParab = qainv * (-t - delta);
    const double zHit = lp.z + distParab * ld.z;
    if (((distParab > 1E20) || (Abs(zHit) <= q.dz)) && distParab > 0) return distParab;
    return kInfinity;
  }
  static VGA_HD double DistanceToOut(const P &q, V3 p, V3 dir, double) {
    const double dirZinv = 1 / dir.z;
    const double distZ = dir.z < 0 ? -(q.dz + p.z) * dirZinv : (q.dz - p.z) * dirZinv;
    double distParab = kInfinity;
    const double rho2 = p.x * p.x + p.y * p.y;
    const double dirRho2 = dir.x * dir.x + dir.y * dir.y;
    const double qa = q.a * dirRho2;
    const double qb = 2. * q.a * (p.x * dir.x + p.y * dir.y) - dir.z;
    const double qc = q.a * rho2 + q.b - p.z;
    const double qainv = 1. / qa;
    const double t = qb * 0.5;
    const double delta = Sqrt(t * t - qc * qa);
    const double sign = qainv < 0 ? -1. : 1.;
    const double d1 = qainv * (-t - sign * delta);
    const double d2 = qainv * (-t + sign * delta);
    if (d1 > 0) distParab = d1;
    if (d1 < 0 && d2 > 0) distParab = d2;
    if (d1 < 0 && d2 < 0) distParab = kInfinity;
    return Min(distParab, distZ);
  }
  static VGA_HD double SafetyToIn(const P &q, V3 p) {
    const double safeZ = Abs(p.z) - q.dz;
    const double r = Sqrt(p.x * p.x + p.y * p.y);
    const double safeRhi = r - q.rhi;
    if ((safeZ > 0) || (safeRhi > 0)) return Max(safeZ, safeRhi);
    const double r0sq = (p.z - q.b) * q.ainv;
    const double dr = r - Sqrt(r0sq);
    const double talf = -2. * q.a * Sqrt(r0sq);
    const double salf = talf / Sqrt(1. + talf * talf);
    return Max(Abs(dr * salf), safeZ);
  }
  static VGA_HD double SafetyToOut(const P &q, V3 p) {
    const double safZ = q.dz - Abs(p.z);
    const double r = Sqrt(p.x * p.x + p.y * p.y);
    const double st = Min(safZ, q.rlo - r);
    return st > 0 ? st : 0.;
  }
};

}  // namespace vga
