// ORACLE -- test infrastructure only (see vgo.h).
//
// Polyhedron restatement: construction as UnplacedPolyhedron
// (source/UnplacedPolyhedron.cpp:27-233), the quadrilateral sets of
// Quadrilaterals (volumes/Quadrilaterals.h:94-572, source/Quadrilaterals.cpp:
// 49-150) over Planes (volumes/Planes.h:207-276, source/Planes.cpp:38-61), and
// the scalar kernels of PolyhedronImplementation
// (volumes/kernel/PolyhedronImplementation.h:317-1010) as instantiated by
// UnplacedPolyhedron::SpecializedVolume (source/UnplacedPolyhedron.cpp:
// 268-320): inner-radius treatment iff any rMin > 0, phi treatment iff the
// phi range is below 360 degrees, "large cutout" iff it is below 180.
//
// Where the reference reads out of range the restatement defines a result
// (the product does the same):
//  * Inside and Contains for a point inside the bounding tube but not
//    strictly between the end planes (FindZSegment gives -1 or N;
//    PolyhedronImplementation.h:686, 731 index fZSegments with it): the end
//    segment is used and a kInside result becomes kSurface (the bounding tube
//    reaches only kTolerance beyond the end planes); Contains is false.
//  * Endcap tests on an end segment that is degenerate in z (no phi
//    quadrilaterals, PolyhedronImplementation.h:616, 652): the phi-wedge test
//    is skipped.
#include <vector>

#include "vgo.h"
#include "vgo_wedge.h"

namespace vgo {

Shape *MakeTube(const double *p, int n);
Shape *MakeTubeTreat(const double *p, bool treatRmin, bool treatPhi);

namespace {

V3 Cross(V3 l, V3 r) { return {l.y * r.z - l.z * r.y, l.z * r.x - l.x * r.z, l.x * r.y - l.y * r.x}; }
V3 Normalized(V3 v) { return v * (1. / v.Mag()); }
V3 FixZeroes(V3 v) {
  for (int i = 0; i < 3; ++i)
    if (Abs(v[i]) < kTolerance) v[i] = 0.;
  return v;
}
V3 FromCylindrical(double r, double phi, double z) { return {r * std::cos(phi), r * std::sin(phi), z}; }
// Vector3D operator== (base/Vector3D.h:751-757): equal within kTolerance.
bool Same(V3 a, V3 b) { return Abs(a.x - b.x) < kTolerance && Abs(a.y - b.y) < kTolerance && Abs(a.z - b.z) < kTolerance; }
// volumes/kernel/GenericKernels.h:108-110
double NormalizeAngle(double a) { return a + kTwoPi * ((a < 0) - int(a / kTwoPi)); }

// DistanceToLineSegmentSquared<kScalar>, GenericKernels.h:118-147
double SegmentDistanceSquared(V3 c0, V3 c1, V3 p) {
  const V3 line = c1 - c0;
  const V3 proj = p - c0;
  const double dot0 = Dot(proj, line);
  if (dot0 <= 0) return (p - c0).Mag2();
  const double dot1 = line.Mag2();
  if (dot1 <= dot0) return (p - c1).Mag2();
  return ((c0 + (dot0 / dot1) * line) - p).Mag2();
}

// Quadrilaterals: planes (normal, distance), four side planes and corners.
struct Quads {
  int n = 0;
  std::vector<V3> nrm, sn[4], cor[4];
  std::vector<double> dist, sd[4];

  void Resize(int k) {
    n = k;
    nrm.assign(k, V3{0, 0, 0});
    dist.assign(k, 0.);
    for (int j = 0; j < 4; ++j) {
      sn[j].assign(k, V3{0, 0, 0});
      sd[j].assign(k, 0.);
      cor[j].assign(k, V3{0, 0, 0});
    }
  }

  // Quadrilaterals::Set, source/Quadrilaterals.cpp:49-140
  void Set(int i, V3 c0, V3 c1, V3 c2, V3 c3) {
    cor[0][i] = c0;
    cor[1][i] = c1;
    cor[2][i] = c2;
    cor[3][i] = c3;
    V3 ch[4];
    ch[0] = cor[0][i];
    int assigned = 1, tested = 1;
    while (tested <= 4) {
      do {
        ch[assigned] = cor[tested][i];
        tested++;
      } while (Same(ch[assigned], ch[assigned - 1]) && tested < 4);
      assigned++;
      if (assigned == 3) break;
    }
    const double a = ch[0][1] * (ch[1][2] - ch[2][2]) + ch[1][1] * (ch[2][2] - ch[0][2]) +
                     ch[2][1] * (ch[0][2] - ch[1][2]);
    const double b = ch[0][2] * (ch[1][0] - ch[2][0]) + ch[1][2] * (ch[2][0] - ch[0][0]) +
                     ch[2][2] * (ch[0][0] - ch[1][0]);
    const double c = ch[0][0] * (ch[1][1] - ch[2][1]) + ch[1][0] * (ch[2][1] - ch[0][1]) +
                     ch[2][0] * (ch[0][1] - ch[1][1]);
    double d = -ch[0][0] * (ch[1][1] * ch[2][2] - ch[2][1] * ch[1][2]) -
               ch[1][0] * (ch[2][1] * ch[0][2] - ch[0][1] * ch[2][2]) -
               ch[2][0] * (ch[0][1] * ch[1][2] - ch[1][1] * ch[0][2]);
    V3 normal{a, b, c};
    const double inv = 1. / normal.Mag();
    normal = normal * inv;
    d *= inv;
    normal = FixZeroes(normal);
    nrm[i] = normal;
    dist[i] = d;
    auto side = [&](int j, V3 s0, V3 s1) {
      if (!Same(s1, s0)) {
        const V3 sv = Normalized(Cross(normal, s1 - s0));
        // Planes::Set(index, normal, origin), source/Planes.cpp:38-47
        const V3 f = FixZeroes(sv);
        const double il = 1. / f.Mag();
        sn[j][i] = il * f;
        sd[j][i] = il * -Dot(f, s0);
      } else {
        sn[j][i] = V3{0, 0, 0};
        sd[j][i] = 0;
      }
    };
    side(0, c0, c1);
    side(1, c1, c2);
    side(2, c2, c3);
    side(3, c3, c0);
  }
  void FlipSign(int i) {
    nrm[i] = V3{-nrm[i].x, -nrm[i].y, -nrm[i].z};
    dist[i] = -dist[i];
  }
  double Plane(int i, V3 p) const { return Dot(p, nrm[i]) + dist[i]; }

  // Planes::Contains, Planes.h:207-219
  bool Contains(V3 p) const {
    bool r = true;
    for (int i = 0; i < n; ++i) r &= Plane(i, p) < 0;
    return r;
  }
  // Planes::Inside, Planes.h:252-270
  int Inside(V3 p) const {
    int r = kInside;
    for (int i = 0; i < n; ++i) {
      const double v = nrm[i].x * p.x + nrm[i].y * p.y + nrm[i].z * p.z + dist[i];
      if (v > kTolerance) r = kOutside;
      if (r == kInside && v > -kTolerance) r = kSurface;
      if (r == kOutside) break;
    }
    return r;
  }
  // Quadrilaterals::RayHitsQuadrilateral, Quadrilaterals.h:94-103
  bool RayHits(int i, V3 x) const {
    bool valid = true;
    for (int j = 0; j < 4; ++j) {
      valid &= Dot(x, sn[j][i]) + sd[j][i] >= 0;
      if (!valid) break;
    }
    return valid;
  }
  // Quadrilaterals::DistanceToIn<kScalar, behind>, Quadrilaterals.h:302-350
  template <bool Behind>
  double DistanceToIn(V3 p, V3 d) const {
    double best = kInfinity;
    for (int i = 0; i < n; ++i) {
      double distance = Dot(p, nrm[i]) + dist[i];
      if (!((Behind ? -distance : distance) >= 0)) continue;
      const double proj = Dot(d, nrm[i]);
      if (!((Behind ? proj : -proj) >= 0)) continue;
      distance /= -proj;
      const V3 x = p + d * distance;
      bool valid = true;
      for (int j = 0; j < 4; ++j) {
        valid &= Dot(x, sn[j][i]) + sd[j][i] >= 0;
        if (!valid) break;
      }
      if (valid) best = distance;
      if (best < kInfinity) break;
    }
    return best;
  }
  // Quadrilaterals::DistanceToOut<kScalar>, Quadrilaterals.h:435-490
  double DistanceToOut(V3 p, V3 d, double zMin, double zMax) const {
    double best = kInfinity;
    for (int i = 0; i < n; ++i) {
      double t = Dot(p, nrm[i]) + dist[i];
      if (!(t < 0)) continue;
      const double proj = Dot(d, nrm[i]);
      if (!(proj >= 0)) continue;
      t /= -proj;
      if (!(t < best)) continue;
      bool valid;
      if (zMin == zMax) {
        valid = RayHits(i, p + t * d);
      } else {
        const double zp = p.z + t * d.z;
        valid = zp >= zMin && zp < zMax;
      }
      if (!valid) continue;
      best = t;
    }
    return best;
  }
  // Quadrilaterals::ScalarDistanceSquared, Quadrilaterals.h:500-568
  double DistanceSquared(int i, V3 p) const {
    const V3 nn = nrm[i];
    const double distance = Dot(p, nn) + dist[i];
    const V3 x = p + distance * nn;
    bool w[4];
    for (int j = 0; j < 4; ++j)
      w[j] = x.x * sn[j][i].x + x.y * sn[j][i].y + x.z * sn[j][i].z + sd[j][i] >= 0;
    if (w[0] && w[1] && w[2] && w[3]) return distance * distance;
    V3 c0, c1;
    if (!w[0]) {
      c0 = cor[0][i];
      c1 = cor[1][i];
    } else if (!w[2]) {
      c0 = cor[2][i];
      c1 = cor[3][i];
    } else if (!w[1]) {
      c0 = cor[1][i];
      c1 = cor[2][i];
    } else {
      c0 = cor[3][i];
      c1 = cor[0][i];
    }
    return SegmentDistanceSquared(c0, c1, p);
  }
};

struct Segment {
  Quads outer, phi, inner;
  bool hasInner = false;
};

struct Polyhedron : Shape {
  int sides = 0, nseg = 0;
  bool hasInnerRadii = false, hasPhi = false, largePhi = false;
  std::vector<Segment> seg;
  std::vector<double> zp;
  std::vector<V3> phiSec;
  Shape *btube = nullptr;   // bounding tube with its own parameters (Extent, wedge)
  Shape *btubeT = nullptr;  // the same tube, treated as HollowTube / NonHollowTube
  Wedge bwedge;             // the bounding tube's phi wedge
  double btubeOffset = 0;

  ~Polyhedron() override {
    delete btube;
    delete btubeT;
  }

  // UnplacedPolyhedron(phiStart, phiDelta, sideCount, zPlaneCount, zPlanes,
  // rMin, rMax), source/UnplacedPolyhedron.cpp:27-233
  bool Init(double phiStartDeg, double phiDeltaDeg, int sideCount, int nz, const double *z, const double *rmin,
            const double *rmax) {
    if (nz < 2 || sideCount < 1 || !(phiDeltaDeg > 0)) return false;
    sides = sideCount;
    nseg = nz - 1;
    hasPhi = phiDeltaDeg < 360;
    largePhi = phiDeltaDeg < 180;
    seg.resize(nseg);
    zp.assign(z, z + nz);
    for (int i = 0; i < nseg; ++i) {
      if (!(z[i] <= z[i + 1])) return false;
      seg[i].hasInner = rmin[i] > 0 || rmin[i + 1] > 0;
      int m = (z[i] == z[i + 1] && rmax[i] == rmax[i + 1]) ? 0 : 1;
      seg[i].outer.Resize(sides * m);
      if (hasPhi) seg[i].phi.Resize(z[i] == z[i + 1] ? 0 : 2);
      m = (z[i] == z[i + 1] && rmin[i] == rmin[i + 1]) ? 0 : 1;
      if (seg[i].hasInner) {
        seg[i].inner.Resize(sides * m);
        hasInnerRadii = true;
      }
    }
    double phiStart = NormalizeAngle(kDegToRad * phiStartDeg);
    double phiDelta = phiDeltaDeg * kDegToRad;
    if (phiDelta > kTwoPi) phiDelta = kTwoPi;
    const double sidePhi = phiDelta / sides;
    std::vector<double> vphi(sides + 1);
    phiSec.resize(sides + 1);
    for (int i = 0; i <= sides; ++i) {
      vphi[i] = NormalizeAngle(phiStart + i * sidePhi);
      const V3 cv = FixZeroes(Normalized(FromCylindrical(1., vphi[i], 0)));
      phiSec[i] = Cross(Normalized(cv), V3{0, 0, -1});
    }
    if (!hasPhi) vphi[sides] = vphi[0];
    const double cosHalf = std::cos(0.5 * sidePhi);
    double innerRadius = kInfinity, outerRadius = -kInfinity;
    for (int i = 0; i < nz; ++i) {
      if (rmin[i] < innerRadius) innerRadius = rmin[i];
      if (!(rmin[i] >= 0 && rmax[i] > 0)) return false;
      if (rmax[i] > outerRadius) outerRadius = rmax[i];
    }
    outerRadius /= cosHalf;
    const double boundingTubeZ = z[nz - 1] - z[0] + 2. * kTolerance;
    const double bPhiStart = !hasPhi ? 0 : phiStart;
    const double bPhiDelta = !hasPhi ? kTwoPi : phiDelta;
    const double tp[5] = {innerRadius - kTolerance, outerRadius + kTolerance, 0.5 * boundingTubeZ, bPhiStart,
                          bPhiDelta};
    btube = MakeTube(tp, 5);
    btubeT = MakeTubeTreat(tp, hasInnerRadii, false);
    bwedge.Init(bPhiDelta, bPhiStart);
    btubeOffset = z[0] + 0.5 * boundingTubeZ;

    auto vi = [&](int plane, int corner) { return plane * (sides + 1) + corner; };
    std::vector<V3> ov(nz * (sides + 1)), iv(nz * (sides + 1));
    for (int i = 0; i < nz; ++i) {
      for (int j = 0, jm = sides + (hasPhi ? 1 : 0); j < jm; ++j) {
        ov[vi(i, j)] = FixZeroes(FromCylindrical(rmax[i] / cosHalf, vphi[j], z[i]));
        iv[vi(i, j)] = FixZeroes(FromCylindrical(rmin[i] / cosHalf, vphi[j], z[i]));
      }
      if (!hasPhi) {
        ov[vi(i, sides)] = ov[vi(i, 0)];
        iv[vi(i, sides)] = iv[vi(i, 0)];
      }
    }
    auto wrong = [](V3 nn, V3 c) { return nn[0] * c[0] + nn[1] * c[1] < 0; };
    for (int ip = 0; ip < nseg; ++ip) {
      Segment &s = seg[ip];
      for (int k = 0; k < s.outer.n; ++k) {
        s.outer.Set(k, ov[vi(ip, k)], ov[vi(ip, k + 1)], ov[vi(ip + 1, k + 1)], ov[vi(ip + 1, k)]);
        if (wrong(s.outer.nrm[k], ov[vi(ip, k)])) s.outer.FlipSign(k);
      }
      for (int k = 0; k < s.inner.n; ++k) {
        if (s.hasInner) {
          s.inner.Set(k, iv[vi(ip, k)], iv[vi(ip, k + 1)], iv[vi(ip + 1, k + 1)], iv[vi(ip + 1, k)]);
          if (wrong(s.inner.nrm[k], iv[vi(ip, k)])) s.inner.FlipSign(k);
        }
      }
      if (hasPhi && s.phi.n == 2) {
        s.phi.Set(0, iv[vi(ip, 0)], iv[vi(ip + 1, 0)], ov[vi(ip + 1, 0)], ov[vi(ip, 0)]);
        if (Dot(s.phi.nrm[0], phiSec[0]) > 0) s.phi.FlipSign(0);
        s.phi.Set(1, ov[vi(ip, sides)], ov[vi(ip + 1, sides)], iv[vi(ip + 1, sides)], iv[vi(ip, sides)]);
        if (Dot(s.phi.nrm[1], phiSec[sides]) < 0) s.phi.FlipSign(1);
      }
    }
    return true;
  }

  // FindZSegmentKernel<kScalar>, PolyhedronImplementation.h:317-331
  int FindZ(double pz) const {
    int index = -1;
    for (double zk : zp) {
      if (!(pz > zk)) break;
      ++index;
    }
    return index;
  }
  int ClampZ(int k) const { return k < 0 ? 0 : (k >= nseg ? nseg - 1 : k); }
  // FindPhiSegment, :357-385
  int FindPhi(V3 p) const {
    int index = -1;
    double first = p.x * phiSec[0].x + p.y * phiSec[0].y + p.z * phiSec[0].z;
    for (int i = 1; i < sides + 1; ++i) {
      const double second = p.x * phiSec[i].x + p.y * phiSec[i].y + p.z * phiSec[i].z;
      if (first >= 0 && second < 0) index = i - 1;
      if (index >= 0) break;
      first = second;
    }
    return index;
  }
  // InPhiCutoutWedge, :679-700
  bool InPhiWedge(const Segment &s, V3 p) const {
    if (s.phi.n != 2) return false;  // degenerate segment (see header)
    const bool first = Dot(p, s.phi.nrm[0]) + s.phi.dist[0] >= 0;
    const bool second = Dot(p, s.phi.nrm[1]) + s.phi.dist[1] >= 0;
    return largePhi ? (first || second) : (first && second);
  }
  V3 BoundsPoint(V3 p) const { return V3{p.x, p.y, p.z - btubeOffset}; }

  // ScalarSegmentContainsKernel, :702-726
  bool SegmentContains(V3 p, int k) const {
    const Segment &s = seg[k];
    if (!s.outer.Contains(p)) return false;
    if (s.hasInner && s.inner.Contains(p)) return false;
    if (hasPhi) {
      bool cin, cout;
      bwedge.Kernel<false>(p, cin, cout);
      if (cout) return false;
    }
    return true;
  }

  // ScalarContainsKernel, :728-770
  bool Contains(V3 p) const override {
    if (!btubeT->Contains(BoundsPoint(p))) return false;
    const int k = FindZ(p.z);
    if (!(k >= 0) || k >= nseg) return false;
    const Segment &s = seg[k];
    if (!s.outer.Contains(p)) return false;
    if (s.hasInner && s.inner.Contains(p)) return false;
    return true;
  }

  // ScalarInsideKernel, :772-830
  int Inside(V3 p) const override {
    if (!btubeT->Contains(BoundsPoint(p))) return kOutside;
    const int k0 = FindZ(p.z);
    const bool endBand = k0 < 0 || k0 >= nseg;
    const Segment &s = seg[ClampZ(k0)];
    int r = kInside;
    const int io = s.outer.Inside(p);
    if (io != kInside) {
      r = io;
    } else {
      int ii = kOutside;
      if (s.hasInner) ii = s.inner.Inside(p);
      if (s.hasInner && ii == kInside) {
        r = kOutside;
      } else if (s.hasInner && ii == kSurface) {
        r = kSurface;
      } else if (hasPhi && InPhiWedge(s, p)) {
        r = kOutside;
      }
    }
    if (endBand && r == kInside) r = kSurface;
    return r;
  }

  // DistanceToInZSegment, :387-428
  double DistanceToInSegment(int k, V3 p, V3 d) const {
    const Segment &s = seg[k];
    double distance = s.outer.DistanceToIn<false>(p, d);
    if (distance < kInfinity) return distance;
    if (hasPhi) distance = s.phi.DistanceToIn<false>(p, d);
    if (distance < kInfinity) return distance;
    if (s.hasInner) distance = s.inner.DistanceToIn<true>(p, d);
    return distance;
  }
  // DistanceToOutZSegment, :430-476
  double DistanceToOutSegment(int k, double zMin, double zMax, V3 p, V3 d) const {
    const Segment &s = seg[k];
    bool done = false;
    double distance = kInfinity;
    if (s.hasInner) {
      distance = s.inner.DistanceToIn<false>(p, d);
      done = distance < kInfinity;
      if (done) return distance;
    }
    if (hasPhi) {
      const double dphi = s.phi.DistanceToIn<true>(p, d);
      if (!done && distance >= 0.) distance = dphi;
      done = distance >= 0. && distance < kInfinity;
      if (done) return distance;
    }
    const double dout = s.outer.DistanceToOut(p, d, zMin, zMax);
    if (!done) distance = dout;
    return distance;
  }
  // ScalarSafetyToZSegmentSquared, :478-521
  double SafetySegmentSquared(int k, int phiIndex, V3 p) const {
    const Segment &s = seg[k];
    if (hasPhi && s.phi.n == 2 && InPhiWedge(s, p))
      return Min(s.phi.DistanceSquared(0, p), s.phi.DistanceSquared(1, p));
    if (phiIndex < 0) return kInfinity;
    double so = kInfinity;
    if (s.outer.n > 0) so = s.outer.DistanceSquared(phiIndex, p);
    double si = kInfinity;
    if (s.hasInner && s.inner.n > 0) si = s.inner.DistanceSquared(phiIndex, p);
    return Min(si, so);
  }
  // ScalarDistanceToEndcaps<pointInside>, :523-602
  template <bool PointInside>
  void DistanceToEndcaps(V3 p, V3 d, double &distance) const {
    const bool useFirst = PointInside ? (d.z < 0) : !(d.z < 0);
    const Segment &s = useFirst ? seg[0] : seg[nseg - 1];
    const double zPlane = useFirst ? zp[0] : zp[nseg];
    const double t = (zPlane - p.z) / d.z;
    if (t < 0 || t >= distance) return;
    const V3 x = p + t * d;
    if (!s.outer.Contains(x)) return;
    if (s.hasInner && s.inner.Contains(x)) return;
    if (hasPhi && InPhiWedge(s, x)) return;
    distance = t;
  }
  // ScalarSafetyToEndcapsSquared, :604-645
  void SafetyToEndcapsSquared(V3 p, double &distSq) const {
    const double firstD = zp[0] - p.z;
    const double lastD = zp[nseg] - p.z;
    const bool isFirst = Abs(firstD) < Abs(lastD);
    const Segment &s = isFirst ? seg[0] : seg[nseg - 1];
    const double t = isFirst ? firstD : lastD;
    const double tSq = t * t;
    if (tSq >= distSq) return;
    const V3 x{p.x, p.y, p.z + t};
    if (!s.outer.Contains(x)) return;
    if (s.hasInner && s.inner.Contains(x)) return;
    if (hasPhi && InPhiWedge(s, x)) return;
    distSq = tSq;
  }

  // ScalarDistanceToInKernel, :832-913 (local point and direction)
  double DistanceToIn(V3 p, V3 d, double stepMax) const override {
    const V3 bp = BoundsPoint(p);
    const bool inBounds = btubeT->Contains(bp);
    if (!inBounds) {
      const double td = btubeT->DistanceToIn(bp, d, stepMax);
      if (td == kInfinity) return kInfinity;
    }
    int k = ClampZ(FindZ(p.z));
    if (inBounds && SegmentContains(p, k)) return -1.;
    const bool right = d.z >= 0;
    double distance = kInfinity;
    if (right) {
      for (; k < nseg; ++k) {
        distance = DistanceToInSegment(k, p, d);
        if (distance >= 0 && distance < kInfinity) break;
      }
    } else {
      for (; k >= 0; --k) {
        distance = DistanceToInSegment(k, p, d);
        if (distance >= 0 && distance < kInfinity) break;
      }
    }
    DistanceToEndcaps<false>(p, d, distance);
    return distance < stepMax ? distance : stepMax;
  }

  // ScalarDistanceToOutKernel, :946-1004
  double DistanceToOut(V3 p, V3 d, double stepMax) const override {
    int k = ClampZ(FindZ(p.z));
    const bool right = d.z >= 0;
    double distance = kInfinity;
    if (right) {
      for (; k < nseg; ++k) {
        distance = DistanceToOutSegment(k, zp[k], zp[k + 1], p, d);
        if (distance >= 0 && distance < kInfinity) break;
        if (zp[k] - p.z > distance) break;
      }
    } else {
      for (; k >= 0; --k) {
        distance = DistanceToOutSegment(k, zp[k], zp[k + 1], p, d);
        if (distance >= 0 && distance < kInfinity) break;
        if (p.z - zp[k] > distance) break;
      }
    }
    DistanceToEndcaps<true>(p, d, distance);
    return distance < stepMax ? distance : stepMax;
  }

  // ScalarSafetyKernel, :915-944 (both SafetyToIn and SafetyToOut)
  double Safety(V3 p) const {
    double safety = kInfinity;
    const int k = ClampZ(FindZ(p.z));
    const int phiIndex = FindPhi(p);
    for (int z = k; z < nseg;) {
      safety = Min(safety, SafetySegmentSquared(z, phiIndex, p));
      ++z;
      if (zp[z] - p.z > safety) break;
    }
    for (int z = k - 1; z >= 0; --z) {
      safety = Min(safety, SafetySegmentSquared(z, phiIndex, p));
      if (p.z - zp[z] > safety) break;
    }
    SafetyToEndcapsSquared(p, safety);
    return std::sqrt(safety);
  }
  double SafetyToIn(V3 p) const override { return Safety(p); }
  double SafetyToOut(V3 p) const override { return Safety(p); }

  // UnplacedPolyhedron::Extent, source/UnplacedPolyhedron.cpp:324-330
  void Extent(V3 &lo, V3 &hi) const override {
    btube->Extent(lo, hi);
    lo.z += btubeOffset;
    hi.z += btubeOffset;
  }
};

}  // namespace

// Parameters: phiStart phiDelta [deg], sides, nz, z[nz], rmin[nz], rmax[nz].
Shape *MakePolyhedron(const double *p, int n) {
  if (n < 4) return nullptr;
  const int sides = (int)p[2], nz = (int)p[3];
  if (sides < 1 || nz < 2 || n != 4 + 3 * nz) return nullptr;
  auto *s = new Polyhedron;
  if (!s->Init(p[0], p[1], sides, nz, p + 4, p + 4 + nz, p + 4 + 2 * nz)) {
    delete s;
    return nullptr;
  }
  return s;
}

}  // namespace vgo
