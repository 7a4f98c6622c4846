// Polycone (a stack of cone sections along z): parameter block and device
// kernels with the semantics of volumes/kernel/PolyconeImplementation.h:58-494
// and the sections UnplacedPolycone::Init builds (source/UnplacedPolycone.cpp:
// 27-180).
//
// The parameter block has a variable length, so kernels receive a pointer to
// it in device memory (PolyconeP) instead of the block itself:
//   blk[0]                 number of sections n
//   blk[2 .. 2+n]          z planes of the sections (n + 1 values, padded to
//                          an even count)
//   then n records of 42 doubles: ConeP (40 doubles), shift, pad
#pragma once
#include "shapes_cone.hpp"

namespace vga {

struct PolyconeP {
  const double *blk;
};

struct PolyconeS {
  using P = PolyconeP;
  static constexpr int kSecDoubles = 42;

  static VGA_HD int NSec(const P &q) { return (int)q.blk[0]; }
  static VGA_HD int SecBase(int n) { return 2 + ((n + 2) & ~1); }
  static VGA_HD double Z(const P &q, int i) { return q.blk[2 + i]; }
  static VGA_HD const ConeP &Sec(const P &q, int i) {
    return *reinterpret_cast<const ConeP *>(q.blk + SecBase(NSec(q)) + kSecDoubles * i);
  }
  static VGA_HD double Shift(const P &q, int i) { return q.blk[SecBase(NSec(q)) + kSecDoubles * i + 40]; }
  static VGA_HD V3 Local(V3 p, double s) { return V3{p.x - 0., p.y - 0., p.z - s}; }

  // GetSectionIndex, volumes/UnplacedPolycone.h:111-120
  static VGA_HD int SectionIndex(const P &q, double z) {
    if (z < Z(q, 0)) return -1;
    const int n = NSec(q);
    for (int i = 0; i < n; ++i)
      if (z >= Z(q, i) && z <= Z(q, i + 1)) return i;
    return -2;
  }

  // False only when the cone of section i would return kInfinity from
  // DistanceToIn: every finite result of the cone is the distance to a point
  // of the ray within its z slab and between its inner and outer radii (hit
  // points on its surfaces, or the start point itself when on a surface), so
  // a ray that stays more than kPad away from the slab-bounded annulus
  // min(rmin) <= rho <= max(rmax) cannot get one.  kPad (1e-6 mm) exceeds
  // every tolerance of the cone and the rounding of this test.
  static VGA_HD bool RayNearSection(const P &q, int i, V3 p, V3 v) {
    constexpr double kPad = 1e-6;
    const ConeP &c = Sec(q, i);
    const double hz = c.dz + kPad;
    const double pz = p.z - Shift(q, i);
    double t0 = 0., t1 = kInfinity;
    if (v.z != 0) {
      const double inv = 1. / v.z;
      double ta = (-hz - pz) * inv, tb = (hz - pz) * inv;
      if (ta > tb) {
        const double t = ta;
        ta = tb;
        tb = t;
      }
      t0 = ta > t0 ? ta : t0;
      t1 = tb;
      if (!(t0 <= t1)) return false;
    } else if (Abs(pz) > hz) {
      return false;
    }
    const double a = v.x * v.x + v.y * v.y, b = p.x * v.x + p.y * v.y;
    double t = a > 0 ? -b / a : t0;
    t = t < t0 ? t0 : (t > t1 ? t1 : t);
    const double x = p.x + t * v.x, y = p.y + t * v.y;
    const double rOut = (c.rmax1 > c.rmax2 ? c.rmax1 : c.rmax2) + kPad;
    if (x * x + y * y > rOut * rOut) return false;
    const double rIn = (c.rmin1 < c.rmin2 ? c.rmin1 : c.rmin2) - kPad;
    if (rIn > 0 && t1 < kInfinity) {
      // rho^2 is convex along the ray: its maximum on [t0, t1] is at an end.
      const double x0 = p.x + t0 * v.x, y0 = p.y + t0 * v.y, x1 = p.x + t1 * v.x, y1 = p.y + t1 * v.y;
      const double r0 = x0 * x0 + y0 * y0, r1 = x1 * x1 + y1 * y1;
      if ((r0 > r1 ? r0 : r1) < rIn * rIn) return false;
    }
    return true;
  }

  // The z term of the section's cone SafetyToIn, evaluated as the cone does:
  // its safety is at least this, so a section whose z term already reaches
  // the running minimum cannot lower it.
  static VGA_HD double SectionSafeZ(const P &q, int i, V3 p) {
    return Abs(Local(p, Shift(q, i)).z) - Sec(q, i).dz;
  }

  static VGA_HD bool Contains(const P &q, V3 p) {
    const int n = NSec(q);
    if (p.z < Z(q, 0) || p.z > Z(q, n)) return false;
    const int i = SectionIndex(q, p.z);
    return ConeS::Contains(Sec(q, i), Local(p, Shift(q, i)));
  }
  // The reference's polycone Inside reports inside or outside only.
  static VGA_HD int Inside(const P &q, V3 p) { return Contains(q, p) ? kInside : kOutside; }

  // PolyconeImplementation.h:148-192.  (Run as a refilled walk, as the
  // Polyhedron's DistanceToIn is, or with each lane skipping ahead over the
  // sections it cannot hit, it measured slower: the cone code per section is
  // short next to the walk's bookkeeping.)
  static VGA_HD double DistanceToIn(const P &q, V3 p, V3 v, double stepMax) {
    const int n = NSec(q);
    double distance = kInfinity;
    int increment = (v.z > 0) ? 1 : -1;
    if (Abs(v.z) < kTolerance) increment = 0;
    int index = SectionIndex(q, p.z);
    if (index == -1) index = 0;
    if (index == -2) index = n - 1;
    do {
      distance = RayNearSection(q, index, p, v)
                     ? ConeS::DistanceToIn(Sec(q, index), Local(p, Shift(q, index)), v, stepMax)
                     : kInfinity;
      if (distance < kInfinity || !increment) break;
      index += increment;
    } while (index >= 0 && index < n);
    return distance;
  }

  static VGA_HD double DistanceToOut(const P &q, V3 point, V3 dir, double stepMax) {
    const int n = NSec(q);
    if (n == 1) return ConeS::DistanceToOut(Sec(q, 0), Local(point, Shift(q, 0)), dir, stepMax);
    const int indexLow = SectionIndex(q, point.z - kTolerance);
    const int indexHigh = SectionIndex(q, point.z + kTolerance);
    int index;
    if (indexLow != indexHigh && indexLow >= 0) {
      index = ConeS::Contains(Sec(q, indexLow), Local(point, Shift(q, indexLow))) ? indexLow : indexHigh;
    } else {
      index = indexLow;
      if (index < 0) index = SectionIndex(q, point.z);
    }
    if (index < 0) return 0.;
    double totalDistance = 0.;
    int increment = (dir.z > 0) ? 1 : -1;
    if (Abs(dir.z) < kTolerance) increment = 0;
    int istep = 0;
    V3 pn = point;
    do {
      const ConeP &c = Sec(q, index);
      if ((totalDistance != 0) || (istep < 2)) {
        pn = V3{point.x + totalDistance * dir.x, point.y + totalDistance * dir.y, point.z + totalDistance * dir.z};
        pn.z -= Shift(q, index);
        if (ConeS::Inside(c, pn) == kOutside) break;
      } else {
        pn.z -= Shift(q, index);
      }
      istep++;
      const double dist = ConeS::DistanceToOut(c, pn, dir, stepMax);
      if (Abs(dist) < 0.5 * kTolerance) {
        // On a section boundary: stop unless the next section holds the exit point.
        int index1 = index;
        if ((index > 0) && (index < n - 1)) {
          index1 += increment;
        } else {
          if ((index == 0) && (increment > 0)) index1 += increment;
          if ((index == n - 1) && (increment < 0)) index1 += increment;
        }
        const double s = totalDistance + dist;
        V3 pte{point.x + s * dir.x, point.y + s * dir.y, point.z + s * dir.z};
        pte.z -= Shift(q, index1);
        if (!ConeS::Contains(Sec(q, index1), pte)) break;
      }
      totalDistance += dist;
      index += increment;
    } while (increment != 0 && index >= 0 && index < n);
    return totalDistance;
  }

  static VGA_HD double SafetyToIn(const P &q, V3 p) {
    const int n = NSec(q);
    int index = SectionIndex(q, p.z);
    bool needZ = false;
    if (index < 0) {
      needZ = true;
      if (index == -1) index = 0;
      if (index == -2) index = n - 1;
    }
    if (needZ) return ConeS::SafetyToInImpl(Sec(q, index), Local(p, Shift(q, index)), false);
    double safety = ConeS::SafetyToInImpl(Sec(q, index), Local(p, Shift(q, index)), true);
    if (safety < kTolerance) return safety;
    double minSafety = safety;
    double zbase = Z(q, index + 1);
    for (int i = index + 1; i < n; ++i) {
      if (Z(q, i) - zbase >= minSafety) break;
      if (SectionSafeZ(q, i, p) >= minSafety) continue;
      safety = ConeS::SafetyToInImpl(Sec(q, i), Local(p, Shift(q, i)), false);
      if (safety < minSafety) minSafety = safety;
    }
    if (index > 0) {
      zbase = Z(q, index - 1);
      for (int i = index - 1; i >= 0; --i) {
        if (zbase - Z(q, i) >= minSafety) break;
        if (SectionSafeZ(q, i, p) >= minSafety) continue;
        safety = ConeS::SafetyToInImpl(Sec(q, i), Local(p, Shift(q, i)), false);
        if (safety < minSafety) minSafety = safety;
      }
    }
    return minSafety;
  }

  static VGA_HD double SafetyToOut(const P &q, V3 p) {
    const int n = NSec(q);
    const int index = SectionIndex(q, p.z);
    if (index < 0) return 0;
    double minSafety = ConeS::SafetyToOut(Sec(q, index), Local(p, Shift(q, index)));
    if (minSafety == kInfinity) return 0.;
    if (minSafety < kTolerance) return 0.;
    double zbase = Z(q, index + 1);
    for (int i = index + 1; i < n; ++i) {
      if (Z(q, i) - zbase >= minSafety) break;
      if (SectionSafeZ(q, i, p) >= minSafety) continue;
      const double safety = ConeS::SafetyToInImpl(Sec(q, i), Local(p, Shift(q, i)), false);
      if (safety < minSafety) minSafety = safety;
    }
    if (index > 0) {
      zbase = Z(q, index - 1);
      for (int i = index - 1; i >= 0; --i) {
        if (zbase - Z(q, i) >= minSafety) break;
        if (SectionSafeZ(q, i, p) >= minSafety) continue;
        const double safety = ConeS::SafetyToInImpl(Sec(q, i), Local(p, Shift(q, i)), false);
        if (safety < minSafety) minSafety = safety;
      }
    }
    return minSafety;
  }
};

// How a kernel obtains a shape's parameters from its block: by value for
// fixed-size blocks, through a pointer to the device copy for variable-size
// ones.
// Length() gives a variable-size block's length in doubles.
template <class S>
struct ParamView {
  static constexpr bool kIndirect = false;
  static VGA_HD const typename S::P &Get(const double *b) { return *reinterpret_cast<const typename S::P *>(b); }
};
template <>
struct ParamView<PolyconeS> {
  static constexpr bool kIndirect = true;
  static VGA_HD PolyconeP Get(const double *b) { return PolyconeP{b}; }
  static VGA_HD int Length(const double *b) {
    const int n = (int)b[0];
    return PolyconeS::SecBase(n) + PolyconeS::kSecDoubles * n;
  }
};

}  // namespace vga
