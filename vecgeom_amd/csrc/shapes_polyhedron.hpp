// Polyhedron (regular n-sided sections stacked along z): parameter block and
// device kernels with the semantics of the scalar kernels of
// volumes/kernel/PolyhedronImplementation.h:317-1010 over the quadrilateral
// sets of volumes/Quadrilaterals.h:94-572 and the planes of
// volumes/Planes.h:207-276, as UnplacedPolyhedron builds and specialises
// them (source/UnplacedPolyhedron.cpp:27-320).
//
// Where the reference indexes out of range, results are defined:
//  * Inside / Contains of a point inside the bounding tube but not strictly
//    between the end planes (the reference indexes its segment array with -1
//    or n): the end segment is used and kInside becomes kSurface; Contains is
//    false.
//  * the phi-wedge test of an end segment that is degenerate in z (no phi
//    quadrilaterals): skipped.
//
// The block has a variable length (every quadrilateral is precomputed on
// the host), so kernels receive a pointer to it in device memory:
//   [0] nseg  [1] sides  [2] hasPhi  [3] largePhi  [4] bounding-tube z offset
//   [5] offset of the z planes (nseg + 1)  [6] offset of the phi-section
//   vectors (3 per side boundary)  [7] offset of the segment records
//   [8] length of the whole block
//   [16 ..] TubeP of the bounding tube with its treatment flags fixed as the
//   reference's HollowTube / NonHollowTube kernels fix them
//   segment record (10): hasInner, nOuter, nPhi, nInner, offOuter, offPhi,
//   offInner, rLo, rHi, rInHi -- every quadrilateral of the segment lies in
//   the annular slab Z(k) <= z <= Z(k+1), rLo <= rho <= rHi (rLo: the
//   smaller inner side distance, rHi: the larger corner radius), and the
//   inner quadrilaterals within rho <= rInHi (the larger inner corner radius)
//   quadrilateral record (33): plane normal (3) and distance, four side
//   planes (normal 3, distance), four corners (3 each), unused
#pragma once
#include "shapes_polycone.hpp"

namespace vga {

struct PolyhedronP {
  const double *blk;
};

struct PolyhedronS {
  using P = PolyhedronP;
  static constexpr int kTubeOff = 16;
  static constexpr int kSegDoubles = 10;
  // 33, not 32: with a record stride of 64 dwords every quadrilateral's
  // field k would sit in the same LDS bank, and lanes reading different
  // quadrilaterals would conflict.
  static constexpr int kQuadDoubles = 33;

  struct Quads {
    const double *q;  // first record
    int n;
    VGA_HD const double *R(int i) const { return q + kQuadDoubles * i; }
    VGA_HD V3 N(int i) const { return V3{R(i)[0], R(i)[1], R(i)[2]}; }
    VGA_HD double D(int i) const { return R(i)[3]; }
    VGA_HD V3 SN(int i, int j) const { return V3{R(i)[4 + 4 * j], R(i)[5 + 4 * j], R(i)[6 + 4 * j]}; }
    VGA_HD double SD(int i, int j) const { return R(i)[7 + 4 * j]; }
    VGA_HD V3 C(int i, int j) const { return V3{R(i)[20 + 3 * j], R(i)[21 + 3 * j], R(i)[22 + 3 * j]}; }
    VGA_HD double Plane(int i, V3 p) const { return Dot(p, N(i)) + D(i); }
  };
  struct Seg {
    bool hasInner;
    Quads outer, phi, inner;
  };

  static VGA_HD int NSeg(const P &q) { return (int)q.blk[0]; }
  static VGA_HD int Sides(const P &q) { return (int)q.blk[1]; }
  static VGA_HD bool HasPhi(const P &q) { return q.blk[2] != 0; }
  static VGA_HD bool LargePhi(const P &q) { return q.blk[3] != 0; }
  static VGA_HD double Z(const P &q, int i) { return q.blk[(int)q.blk[5] + i]; }
  static VGA_HD V3 PhiSec(const P &q, int i) {
    const double *v = q.blk + (int)q.blk[6] + 3 * i;
    return V3{v[0], v[1], v[2]};
  }
  static VGA_HD const TubeP &Tube(const P &q) { return *reinterpret_cast<const TubeP *>(q.blk + kTubeOff); }
  static VGA_HD V3 BoundsPoint(const P &q, V3 p) { return V3{p.x, p.y, p.z - q.blk[4]}; }
  static VGA_HD Seg GetSeg(const P &q, int k) {
    const double *r = SegRec(q, k);
    Seg s;
    s.hasInner = r[0] != 0;
    s.outer = Quads{q.blk + (int)r[4], (int)r[1]};
    s.phi = Quads{q.blk + (int)r[5], (int)r[2]};
    s.inner = Quads{q.blk + (int)r[6], (int)r[3]};
    return s;
  }

  static VGA_HD const double *SegRec(const P &q, int k) { return q.blk + (int)q.blk[7] + kSegDoubles * k; }

  // Pruning of whole segments.  The reference visits every segment along the
  // ray (DistanceToIn) or around the point (safety); a segment is skipped
  // here only when its result provably cannot change the outcome, so the
  // results are those of the full walk.

  // 0 only when no quadrilateral of segment k can yield a hit: a hit is a
  // point of one of the segment's quadrilaterals (to within rounding), and
  // those lie in the segment's annular slab, so a ray (t >= 0) that stays
  // more than kPad (1e-6 mm, far above every rounding of the tests) away from
  // the slab gets kInfinity from every quadrilateral of the segment.  1 when
  // the ray may hit the outer or phi quadrilaterals but stays more than kPad
  // outside radius rInHi within the slab (no inner hit possible), 2 otherwise.
  // The ray's constants of RayNearSegment, computed once per query.
  struct RayPre {
    double invz, tc;  // 1 / v.z; the parameter of closest approach to the z axis
    bool radial;      // v has a component across the z axis
  };
  static VGA_HD RayPre MakeRayPre(V3 p, V3 v) {
    RayPre r;
    r.invz = v.z != 0 ? 1. / v.z : 0.;
    const double a = v.x * v.x + v.y * v.y, b = p.x * v.x + p.y * v.y;
    r.radial = a > 0;
    r.tc = r.radial ? -b / a : 0.;
    return r;
  }
  static VGA_HD int RayNearSegment(const P &q, int k, V3 p, V3 v, const RayPre &pre) {
    constexpr double kPad = 1e-6;
    const double *r = SegRec(q, k);
    const double zlo = Z(q, k) - kPad, zhi = Z(q, k + 1) + kPad;
    double t0 = 0., t1 = kInfinity;
    if (v.z != 0) {
      const double inv = pre.invz;
      double ta = (zlo - p.z) * inv, tb = (zhi - p.z) * inv;
      if (ta > tb) {
        const double t = ta;
        ta = tb;
        tb = t;
      }
      t0 = ta > t0 ? ta : t0;
      t1 = tb;
      if (!(t0 <= t1)) return false;
    } else if (p.z < zlo || p.z > zhi) {
      return false;
    }
    double t = pre.radial ? pre.tc : t0;
    t = t < t0 ? t0 : (t > t1 ? t1 : t);
    const double x = p.x + t * v.x, y = p.y + t * v.y;
    const double rho2 = x * x + y * y;  // smallest rho^2 within the slab
    const double rOut = r[8] + kPad;
    if (rho2 > rOut * rOut) return 0;
    const double rIn = r[7] - kPad;
    if (rIn > 0 && t1 < kInfinity) {
      // rho^2 is convex along the ray: its maximum on [t0, t1] is at an end.
      const double x0 = p.x + t0 * v.x, y0 = p.y + t0 * v.y, x1 = p.x + t1 * v.x, y1 = p.y + t1 * v.y;
      const double r0 = x0 * x0 + y0 * y0, r1 = x1 * x1 + y1 * y1;
      if ((r0 > r1 ? r0 : r1) < rIn * rIn) return 0;
    }
    const double rInHi = r[9] + kPad;
    return rho2 > rInHi * rInHi ? 1 : 2;
  }

  // True only when segment k's squared safety cannot be below `safety2`
  // (a squared distance): every quadrilateral lies in the segment's annular
  // slab, and the value the reference computes for a quadrilateral is the
  // squared distance to the quadrilateral (point inside its side planes) or
  // to one of its edges, never below the squared distance to the slab by
  // more than rounding; pad covers that rounding with a wide margin.
  // rho: Sqrt(p.x * p.x + p.y * p.y), computed once per query.
  static VGA_HD bool SegmentSafetyAtLeast(const P &q, int k, V3 p, double rho, double safety2) {
    const double *r = SegRec(q, k);
    const double zlo = Z(q, k), zhi = Z(q, k + 1);
    const double dz = p.z < zlo ? zlo - p.z : (p.z > zhi ? p.z - zhi : 0.);
    const double dr = rho < r[7] ? r[7] - rho : (rho > r[8] ? rho - r[8] : 0.);
    const double scale = Abs(p.x) + Abs(p.y) + Abs(p.z) + r[8] + Abs(zlo) + Abs(zhi);
    const double pad = 1e-6 + 1e-8 * scale;
    // The slab distance sqrt(d2) exceeds sqrt(safety2) by more than pad when
    // d2 >= safety2 + pad (safety2 + 1) + pad^2, since 2 sqrt(s) <= s + 1;
    // the last factor covers the rounding of the right-hand side.  (No
    // square root; it prunes a pad-wide sliver less than the exact bound.)
    const double d2 = dz * dz + dr * dr;
    return d2 >= (safety2 + pad * (safety2 + 1.) + pad * pad) * (1. + 1e-12);
  }

  // DistanceToLineSegmentSquared<kScalar>, GenericKernels.h:118-147
  static VGA_HD double SegmentDistanceSquared(V3 c0, V3 c1, V3 p) {
    const V3 line = c1 - c0;
    const V3 proj = p - c0;
    const double dot0 = Dot(proj, line);
    if (dot0 <= 0) return (p - c0).Mag2();
    const double dot1 = line.Mag2();
    if (dot1 <= dot0) return (p - c1).Mag2();
    return ((c0 + (dot0 / dot1) * line) - p).Mag2();
  }

  // Planes::Contains, Planes.h:207-219
  static VGA_HD bool QContains(const Quads &s, V3 p) {
    bool r = true;
    for (int i = 0; i < s.n; ++i) r &= s.Plane(i, p) < 0;
    return r;
  }
  // Planes::Inside, Planes.h:252-270
  static VGA_HD int QInside(const Quads &s, V3 p) {
    int r = kInside;
    for (int i = 0; i < s.n; ++i) {
      const double *R = s.R(i);
      const double v = R[0] * p.x + R[1] * p.y + R[2] * p.z + R[3];
      if (v > kTolerance) r = kOutside;
      if (r == kInside && v > -kTolerance) r = kSurface;
      if (r == kOutside) break;
    }
    return r;
  }
  static VGA_HD bool RayHits(const Quads &s, int i, V3 x) {
    bool valid = true;
    for (int j = 0; j < 4; ++j) {
      valid &= Dot(x, s.SN(i, j)) + s.SD(i, j) >= 0;
      if (!valid) break;
    }
    return valid;
  }
  // Quadrilaterals::DistanceToIn<kScalar, behind>, Quadrilaterals.h:302-350.
  // The same tests in the same order as the reference's scalar loop, run in
  // two passes over chunks of 32 quadrilaterals: the first evaluates the
  // cheap plane tests of every quadrilateral (a wave stays converged there)
  // and records the survivors in a bit mask; the second visits only each
  // lane's own survivors in index order for the division and the side-plane
  // tests.  A wave then pays for the largest survivor count among its lanes
  // instead of for every quadrilateral that any lane survives.
  template <bool Behind>
  static VGA_HD double QDistanceToIn(const Quads &s, V3 p, V3 d) {
    double best = kInfinity;
    for (int base = 0; base < s.n; base += 32) {
      const int m = s.n - base < 32 ? s.n - base : 32;
      uint32_t mask = 0;
      for (int k = 0; k < m; ++k) {
        const V3 nn = s.N(base + k);
        const double distance = Dot(p, nn) + s.D(base + k);
        const double proj = Dot(d, nn);
        const bool ok = (Behind ? -distance : distance) >= 0 && (Behind ? proj : -proj) >= 0;
        mask |= (uint32_t)ok << k;
      }
      while (mask) {
        const int i = base + __builtin_ctz(mask);
        mask &= mask - 1;
        const V3 nn = s.N(i);
        double distance = Dot(p, nn) + s.D(i);
        distance /= -Dot(d, nn);
        const V3 x = p + d * distance;
        if (RayHits(s, i, x)) best = distance;
        if (best < kInfinity) return best;
      }
    }
    return best;
  }
  // Quadrilaterals::DistanceToOut<kScalar>, Quadrilaterals.h:435-490 (a
  // two-pass form as in QDistanceToIn measured slower here: from inside, about
  // half of the quadrilaterals are candidates for every lane).
  static VGA_HD double QDistanceToOut(const Quads &s, V3 p, V3 d, double zMin, double zMax) {
    double best = kInfinity;
    for (int i = 0; i < s.n; ++i) {
      const V3 nn = s.N(i);
      double t = Dot(p, nn) + s.D(i);
      if (!(t < 0)) continue;
      const double proj = Dot(d, nn);
      if (!(proj >= 0)) continue;
      t /= -proj;
      if (!(t < best)) continue;
      bool valid;
      if (zMin == zMax) {
        valid = RayHits(s, i, p + t * d);
      } else {
        const double zp = p.z + t * d.z;
        valid = zp >= zMin && zp < zMax;
      }
      if (valid) best = t;
    }
    return best;
  }
  // Quadrilaterals::ScalarDistanceSquared, Quadrilaterals.h:500-568
  static VGA_HD double QDistanceSquared(const Quads &s, int i, V3 p) {
    const V3 nn = s.N(i);
    const double distance = Dot(p, nn) + s.D(i);
    const V3 x = p + distance * nn;
    bool w[4];
    for (int j = 0; j < 4; ++j) {
      const V3 sn = s.SN(i, j);
      w[j] = x.x * sn.x + x.y * sn.y + x.z * sn.z + s.SD(i, j) >= 0;
    }
    if (w[0] && w[1] && w[2] && w[3]) return distance * distance;
    const int a = !w[0] ? 0 : (!w[2] ? 2 : (!w[1] ? 1 : 3));
    return SegmentDistanceSquared(s.C(i, a), s.C(i, (a + 1) & 3), p);
  }

  // FindZSegmentKernel<kScalar>, PolyhedronImplementation.h:317-331
  static VGA_HD int FindZ(const P &q, double pz) {
    const int nz = NSeg(q) + 1;
    int index = -1;
    for (int k = 0; k < nz; ++k) {
      if (!(pz > Z(q, k))) break;
      ++index;
    }
    return index;
  }
  static VGA_HD int ClampZ(const P &q, int k) {
    const int n = NSeg(q);
    return k < 0 ? 0 : (k >= n ? n - 1 : k);
  }
  // FindPhiSegment, :357-385
  static VGA_HD int FindPhi(const P &q, V3 p) {
    int index = -1;
    V3 s = PhiSec(q, 0);
    double first = p.x * s.x + p.y * s.y + p.z * s.z;
    const int n = Sides(q) + 1;
    for (int i = 1; i < n; ++i) {
      s = PhiSec(q, i);
      const double second = p.x * s.x + p.y * s.y + p.z * s.z;
      if (first >= 0 && second < 0) index = i - 1;
      if (index >= 0) break;
      first = second;
    }
    return index;
  }
  // InPhiCutoutWedge, :679-700
  static VGA_HD bool InPhiWedge(const P &q, const Seg &s, V3 p) {
    if (s.phi.n != 2) return false;
    const bool first = s.phi.Plane(0, p) >= 0;
    const bool second = s.phi.Plane(1, p) >= 0;
    return LargePhi(q) ? (first || second) : (first && second);
  }
  // ScalarSegmentContainsKernel, :702-726
  static VGA_HD bool SegmentContains(const P &q, V3 p, int k) {
    const Seg s = GetSeg(q, k);
    if (!QContains(s.outer, p)) return false;
    if (s.hasInner && QContains(s.inner, p)) return false;
    if (HasPhi(q)) {
      bool cin, cout;
      WedgeS::Kernel<false>(Tube(q).w, p, cin, cout);
      if (cout) return false;
    }
    return true;
  }

  // ScalarContainsKernel, :728-770 (a form that evaluates every test on
  // every lane and combines the results measured slower: whole waves of
  // points outside the bounding tube skip the plane loops here).
  static VGA_HD bool Contains(const P &q, V3 p) {
    if (!TubeS::Contains(Tube(q), BoundsPoint(q, p))) return false;
    const int k = FindZ(q, p.z);
    if (!(k >= 0) || k >= NSeg(q)) return false;
    const Seg s = GetSeg(q, k);
    if (!QContains(s.outer, p)) return false;
    if (s.hasInner && QContains(s.inner, p)) return false;
    return true;
  }

  // ScalarInsideKernel, :772-830
  static VGA_HD int Inside(const P &q, V3 p) {
    if (!TubeS::Contains(Tube(q), BoundsPoint(q, p))) return kOutside;
    const int k0 = FindZ(q, p.z);
    const bool endBand = k0 < 0 || k0 >= NSeg(q);
    const Seg s = GetSeg(q, ClampZ(q, k0));
    int r = QInside(s.outer, p);
    if (r == kInside) {
      const int ii = s.hasInner ? QInside(s.inner, p) : kOutside;
      if (ii == kInside)
        r = kOutside;
      else if (ii == kSurface)
        r = kSurface;
      else if (HasPhi(q) && InPhiWedge(q, s, p))
        r = kOutside;
    }
    if (endBand && r == kInside) r = kSurface;
    return r;
  }

  // DistanceToInZSegment, :387-428
  // inner: false when RayNearSegment has ruled out inner hits (the inner
  // shell would return kInfinity).
  static VGA_HD double DistanceToInSegment(const P &q, int k, V3 p, V3 d, bool inner) {
    const Seg s = GetSeg(q, k);
    double distance = QDistanceToIn<false>(s.outer, p, d);
    if (distance < kInfinity) return distance;
    if (HasPhi(q)) distance = QDistanceToIn<false>(s.phi, p, d);
    if (distance < kInfinity) return distance;
    if (s.hasInner) distance = inner ? QDistanceToIn<true>(s.inner, p, d) : kInfinity;
    return distance;
  }
  // DistanceToOutZSegment, :430-476
  static VGA_HD double DistanceToOutSegment(const P &q, int k, V3 p, V3 d) {
    const Seg s = GetSeg(q, k);
    bool done = false;
    double distance = kInfinity;
    if (s.hasInner) {
      distance = QDistanceToIn<false>(s.inner, p, d);
      done = distance < kInfinity;
      if (done) return distance;
    }
    if (HasPhi(q)) {
      const double dphi = QDistanceToIn<true>(s.phi, p, d);
      if (!done && distance >= 0.) distance = dphi;
      done = distance >= 0. && distance < kInfinity;
      if (done) return distance;
    }
    const double dout = QDistanceToOut(s.outer, p, d, Z(q, k), Z(q, k + 1));
    if (!done) distance = dout;
    return distance;
  }
  // ScalarSafetyToZSegmentSquared, :478-521
  static VGA_HD double SafetySegmentSquared(const P &q, int k, int phiIndex, V3 p) {
    const Seg s = GetSeg(q, k);
    if (HasPhi(q) && s.phi.n == 2 && InPhiWedge(q, s, p))
      return Min(QDistanceSquared(s.phi, 0, p), QDistanceSquared(s.phi, 1, p));
    if (phiIndex < 0) return kInfinity;
    double so = kInfinity;
    if (s.outer.n > 0) so = QDistanceSquared(s.outer, phiIndex, p);
    double si = kInfinity;
    if (s.hasInner && s.inner.n > 0) si = QDistanceSquared(s.inner, phiIndex, p);
    return Min(si, so);
  }
  // The endcap tests of ScalarDistanceToEndcaps / ScalarSafetyToEndcapsSquared
  // on end segment k.  A point more than 1e-6 mm outside the segment's corner
  // radius is outside some outer plane, so the outer test is skipped there.
  static VGA_HD bool OnEndcap(const P &q, int k, V3 x) {
    const double rOut = SegRec(q, k)[8] + 1e-6;
    if (x.x * x.x + x.y * x.y > rOut * rOut) return false;
    const Seg s = GetSeg(q, k);
    if (!QContains(s.outer, x)) return false;
    if (s.hasInner && QContains(s.inner, x)) return false;
    if (HasPhi(q) && InPhiWedge(q, s, x)) return false;
    return true;
  }
  // ScalarDistanceToEndcaps<pointInside>, :523-602
  template <bool PointInside>
  static VGA_HD void DistanceToEndcaps(const P &q, V3 p, V3 d, double &distance) {
    const bool useFirst = PointInside ? (d.z < 0) : !(d.z < 0);
    const int n = NSeg(q);
    const double zPlane = useFirst ? Z(q, 0) : Z(q, n);
    const double t = (zPlane - p.z) / d.z;
    if (t < 0 || t >= distance) return;
    if (OnEndcap(q, useFirst ? 0 : n - 1, p + t * d)) distance = t;
  }
  // ScalarSafetyToEndcapsSquared, :604-645
  static VGA_HD void SafetyToEndcapsSquared(const P &q, V3 p, double &distSq) {
    const int n = NSeg(q);
    const double firstD = Z(q, 0) - p.z;
    const double lastD = Z(q, n) - p.z;
    const bool isFirst = Abs(firstD) < Abs(lastD);
    const double t = isFirst ? firstD : lastD;
    const double tSq = t * t;
    if (tSq >= distSq) return;
    if (OnEndcap(q, isFirst ? 0 : n - 1, V3{p.x, p.y, p.z + t})) distSq = tSq;
  }

  // ScalarDistanceToInKernel, :832-913 (local point and direction), split
  // into a walk: DistInBegin (bounding tube, start segment, containment),
  // DistInStep (one z segment's quadrilaterals), DistInEnd (endcaps, step
  // limit).  The reference walks the segments in the direction of d.z; a
  // segment the ray cannot come near yields kInfinity there, and the walk
  // here skips ahead over such segments.  DistanceToIn runs the walk on one
  // lane; the refill kernel interleaves the walks of a wave's lanes so that
  // every DistInStep runs on as many lanes as possible.
  static constexpr bool kWalkIn = true, kWalkOut = true;
  struct DistInWalk {
    V3 p, d;
    double stepMax, distance;
    RayPre pre;
    int k, step, near;
    bool final;
  };
  // Moves w.k to the next segment (from w.k on) that the ray can hit; false
  // when there is none.
  static VGA_HD bool DistInAdvance(const P &q, DistInWalk &w) {
    const int n = NSeg(q);
    int near = RayNearSegment(q, w.k, w.p, w.d, w.pre);
    while (!near) {
      w.k += w.step;
      if (w.k < 0 || w.k >= n) return false;
      near = RayNearSegment(q, w.k, w.p, w.d, w.pre);
    }
    w.near = near;
    return true;
  }
  // True when the query is resolved without a segment step.
  static VGA_HD bool DistInBegin(const P &q, V3 p, V3 d, double stepMax, DistInWalk &w) {
    w.p = p;
    w.d = d;
    w.stepMax = stepMax;
    w.final = true;
    const V3 bp = BoundsPoint(q, p);
    const bool inBounds = TubeS::Contains(Tube(q), bp);
    if (!inBounds && TubeS::DistanceToIn(Tube(q), bp, d, stepMax) == kInfinity) {
      w.distance = kInfinity;
      return true;
    }
    w.k = ClampZ(q, FindZ(q, p.z));
    if (inBounds && SegmentContains(q, p, w.k)) {
      w.distance = -1.;
      return true;
    }
    w.final = false;
    w.step = d.z >= 0 ? 1 : -1;
    w.distance = kInfinity;
    w.pre = MakeRayPre(p, d);
    return !DistInAdvance(q, w);
  }
  // True when the walk is over.
  static VGA_HD bool DistInStep(const P &q, DistInWalk &w) {
    w.distance = DistanceToInSegment(q, w.k, w.p, w.d, w.near == 2);
    if (w.distance >= 0 && w.distance < kInfinity) return true;
    w.k += w.step;
    if (w.k < 0 || w.k >= NSeg(q)) return true;
    if (DistInAdvance(q, w)) return false;
    w.distance = kInfinity;
    return true;
  }
  static VGA_HD double DistInEnd(const P &q, DistInWalk &w) {
    if (w.final) return w.distance;
    double distance = w.distance;
    DistanceToEndcaps<false>(q, w.p, w.d, distance);
    return distance < w.stepMax ? distance : w.stepMax;
  }
  static VGA_HD double DistanceToIn(const P &q, V3 p, V3 d, double stepMax) {
    DistInWalk w;
    if (!DistInBegin(q, p, d, stepMax, w))
      while (!DistInStep(q, w)) {
      }
    return DistInEnd(q, w);
  }
  struct DistInOps {
    using State = DistInWalk;
    static VGA_HD bool Begin(const P &q, V3 p, V3 d, double sm, State &w) { return DistInBegin(q, p, d, sm, w); }
    static VGA_HD bool Step(const P &q, State &w) { return DistInStep(q, w); }
    static VGA_HD double End(const P &q, State &w) { return DistInEnd(q, w); }
  };

  // ScalarDistanceToOutKernel, :946-1004 (local point and direction), as a
  // walk over the z segments in the direction of d.z like DistanceToIn:
  // DistOutBegin (start segment), DistOutStep (one segment and the
  // reference's two stopping tests), DistOutEnd (endcaps, step limit).
  struct DistOutWalk {
    V3 p, d;
    double stepMax, distance;
    int k, step;
  };
  static VGA_HD bool DistOutBegin(const P &q, V3 p, V3 d, double stepMax, DistOutWalk &w) {
    w.p = p;
    w.d = d;
    w.stepMax = stepMax;
    w.k = ClampZ(q, FindZ(q, p.z));
    w.step = d.z >= 0 ? 1 : -1;
    w.distance = kInfinity;
    return false;
  }
  static VGA_HD bool DistOutStep(const P &q, DistOutWalk &w) {
    w.distance = DistanceToOutSegment(q, w.k, w.p, w.d);
    if (w.distance >= 0 && w.distance < kInfinity) return true;
    if (w.step > 0 ? (Z(q, w.k) - w.p.z > w.distance) : (w.p.z - Z(q, w.k) > w.distance)) return true;
    w.k += w.step;
    return w.k < 0 || w.k >= NSeg(q);
  }
  static VGA_HD double DistOutEnd(const P &q, DistOutWalk &w) {
    double distance = w.distance;
    DistanceToEndcaps<true>(q, w.p, w.d, distance);
    return distance < w.stepMax ? distance : w.stepMax;
  }
  static VGA_HD double DistanceToOut(const P &q, V3 p, V3 d, double stepMax) {
    DistOutWalk w;
    if (!DistOutBegin(q, p, d, stepMax, w))
      while (!DistOutStep(q, w)) {
      }
    return DistOutEnd(q, w);
  }
  struct DistOutOps {
    using State = DistOutWalk;
    static VGA_HD bool Begin(const P &q, V3 p, V3 d, double sm, State &w) { return DistOutBegin(q, p, d, sm, w); }
    static VGA_HD bool Step(const P &q, State &w) { return DistOutStep(q, w); }
    static VGA_HD double End(const P &q, State &w) { return DistOutEnd(q, w); }
  };

  // ScalarSafetyKernel, :915-944 (both SafetyToIn and SafetyToOut)
  static VGA_HD double Safety(const P &q, V3 p) {
    const int n = NSeg(q);
    double safety = kInfinity;
    const int k = ClampZ(q, FindZ(q, p.z));
    const int phiIndex = FindPhi(q, p);
    const double rho = Sqrt(p.x * p.x + p.y * p.y);
    // (Skipping ahead over pruned segments per lane, as DistanceToIn does,
    // measured slower here: the segment computation is short.)
    for (int z = k; z < n;) {
      if (!SegmentSafetyAtLeast(q, z, p, rho, safety)) safety = Min(safety, SafetySegmentSquared(q, z, phiIndex, p));
      ++z;
      if (Z(q, z) - p.z > safety) break;
    }
    for (int z = k - 1; z >= 0; --z) {
      if (!SegmentSafetyAtLeast(q, z, p, rho, safety)) safety = Min(safety, SafetySegmentSquared(q, z, phiIndex, p));
      if (p.z - Z(q, z) > safety) break;
    }
    SafetyToEndcapsSquared(q, p, safety);
    return Sqrt(safety);
  }
  static VGA_HD double SafetyToIn(const P &q, V3 p) { return Safety(q, p); }
  static VGA_HD double SafetyToOut(const P &q, V3 p) { return Safety(q, p); }
};

template <>
struct ParamView<PolyhedronS> {
  static constexpr bool kIndirect = true;
  static VGA_HD PolyhedronP Get(const double *b) { return PolyhedronP{b}; }
  static VGA_HD int Length(const double *b) { return (int)b[8]; }
};

}  // namespace vga
