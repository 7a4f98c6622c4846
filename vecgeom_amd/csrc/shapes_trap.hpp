// Trapezoid: four side planes in SOA form (the reference's PlaneShell,
// base/PlaneShell.h:95-300) plus the z-slab, with the semantics of
// volumes/kernel/TrapezoidImplementation.h:134-420.  Planes are computed on
// the host from the corners (source/UnplacedTrapezoid.cpp:209-345, 673-685).
#pragma once
#include "vg_math.hpp"

namespace vga {

struct TrapP {
  double dz;
  double A[4], B[4], C[4], D[4];
  double prm[11];  // constructor parameters (for the extent)
  double pad[2];
};

struct TrapS {
  using P = TrapP;
  static VGA_HD double Dist(const P &t, int i, V3 p) { return t.A[i] * p.x + t.B[i] * p.y + t.C[i] * p.z + t.D[i]; }
  template <bool ForInside>
  static VGA_HD void Kernel(const P &t, V3 p, bool &cin, bool &cout) {
    cout = Abs(p.z) > PlusTol<ForInside>(t.dz);
    if (cout) return;
    if (ForInside) cin = Abs(p.z) < MinusTol<ForInside>(t.dz);
    double dist[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) dist[i] = Dist(t, i, p);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      cout |= dist[i] > PlusTol<ForInside>(0.);
      if (cout) return;
      if (ForInside) cin &= dist[i] < MinusTol<ForInside>(0.);
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
    if (cout) return kOutside;
    return cin ? kInside : kSurface;
  }
  static VGA_HD double ShellDistanceToIn(const P &t, V3 point, double smin, V3 dir) {
    double smax = kInfinity;
    double pdist[4], proj[4], vdist[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      pdist[i] = Dist(t, i, point);
      proj[i] = t.A[i] * dir.x + t.B[i] * dir.y + t.C[i] * dir.z;
      vdist[i] = -pdist[i] / proj[i];
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const bool posPoint = pdist[i] >= -kHalfTolerance;
      const bool posDir = proj[i] > 0;
      const bool fromInside = (!posPoint && posDir);
      const bool fromOutside = (posPoint && !posDir);
      if ((posPoint && posDir) || (fromInside && vdist[i] < smin) || (fromOutside && vdist[i] > smax))
        return kInfinity;
      const bool valid = (smin < vdist[i] && vdist[i] < smax);
      if (fromOutside && valid) smin = vdist[i];
      if (fromInside && valid) smax = vdist[i];
    }
    const V3 entry{point.x + dir.x * smin, point.y + dir.y * smin, point.z + dir.z * smin};
    bool valid = true;
#pragma unroll
    for (int i = 0; i < 4; ++i) valid &= (Dist(t, i, entry) <= PlusTol<true>(0.));
    return valid ? smin : kInfinity;
  }
  static VGA_HD double DistanceToIn(const P &t, V3 point, V3 dir, double) {
    const bool posZ = dir.z > 0.0, negZ = dir.z < 0.0;
    const double zsign = negZ ? -1. : 1.;
    const double mx = zsign * t.dz - point.z;
    if ((posZ && mx < PlusTol<true>(0.)) || (negZ && mx > MinusTol<true>(0.))) return kInfinity;
    const double dirFactor = 1. / dir.z;
    double smax = mx * dirFactor;
    double smin = (-zsign * t.dz - point.z) * dirFactor;
    const bool test = (!posZ) && (!negZ);
    const bool zrange = Abs(point.z) < MinusTol<true>(t.dz);
    if (test && !zrange) return kInfinity;
    if (test && zrange) {
      smin = 0.0;
      smax = kInfinity;
    }
    const double dpl = ShellDistanceToIn(t, point, smin, dir);
    if (dpl == kInfinity || dpl > smax) return kInfinity;
    double distance = smin;
    if ((dpl > smin) && (dpl < smax)) distance = dpl;
    if (distance < 0) distance = 0.0;
    return distance;
  }
  static VGA_HD double DistanceToOut(const P &t, V3 point, V3 dir, double) {
    const bool posZ = dir.z > 0.0, negZ = dir.z < 0.0;
    const double zsign = negZ ? -1. : 1.;
    const double mx = zsign * t.dz - point.z;
    if ((posZ && mx <= PlusTol<true>(0.)) || (negZ && mx >= MinusTol<true>(0.))) return 0.;
    double distance = mx / dir.z;
    double distOut = kInfinity;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const double pdist = Dist(t, i, point);
      const double proj = t.A[i] * dir.x + t.B[i] * dir.y + t.C[i] * dir.z;
      const double vdist = -pdist / proj;
      if (vdist >= 0. && proj > 0 && vdist < distOut) distOut = vdist;
    }
    if (distOut < distance) distance = distOut;
    return distance;
  }
  static VGA_HD double SafetyToIn(const P &t, V3 p) {
    double safety = Abs(p.z) - t.dz;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const double d = Dist(t, i, p);
      if (d > safety) safety = d;
    }
    return safety;
  }
  static VGA_HD double SafetyToOut(const P &t, V3 p) {
    double safety = t.dz - Abs(p.z);
    if (safety < 0.0) safety = 0.0;
    if (safety == 0.0) return safety;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const double d = -Dist(t, i, p);
      if (d < safety) safety = d;
    }
    return safety;
  }
};

}  // namespace vga
