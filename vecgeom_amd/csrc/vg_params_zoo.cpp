// Host-side parameter blocks of the zoo shapes beyond Box and Tube: the
// constructor parameters plus the values each Unplaced* class caches,
// computed with the reference's expressions on the host.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "shapes.hpp"
#include "vg_internal.hpp"

namespace vga {

namespace {

template <class T>
std::vector<double> Pack(const T &p) {
  static_assert(sizeof(T) % sizeof(double) == 0, "parameter block must be whole doubles");
  std::vector<double> v(sizeof(T) / sizeof(double));
  std::memcpy(v.data(), &p, sizeof(T));
  return v;
}

std::vector<double> Fail(std::string *err, const std::string &m) {
  if (err) *err = m;
  return {};
}

constexpr double kDegToRad = kPi / 180.;

// Trapezoid corners and side planes, source/UnplacedTrapezoid.cpp:209-345,
// 673-685 (tan/sin/cos of theta, phi in radians).
struct TrapBuild {
  double dz, dy1, dx1, dx2, ta1, dy2, dx3, dx4, ta2, tcp, tsp;
  V3 pt[8];
  explicit TrapBuild(const double *p) {
    dz = p[0]; dy1 = p[3]; dx1 = p[4]; dx2 = p[5]; ta1 = p[6]; dy2 = p[7]; dx3 = p[8]; dx4 = p[9]; ta2 = p[10];
    tsp = std::tan(p[1]) * std::sin(p[2]);
    tcp = std::tan(p[1]) * std::cos(p[2]);
    pt[0] = {-dz * tcp - dy1 * ta1 - dx1, -dz * tsp - dy1, -dz};
    pt[1] = {-dz * tcp - dy1 * ta1 + dx1, -dz * tsp - dy1, -dz};
    pt[2] = {-dz * tcp + dy1 * ta1 - dx2, -dz * tsp + dy1, -dz};
    pt[3] = {-dz * tcp + dy1 * ta1 + dx2, -dz * tsp + dy1, -dz};
    pt[4] = {+dz * tcp - dy2 * ta2 - dx3, +dz * tsp - dy2, +dz};
    pt[5] = {+dz * tcp - dy2 * ta2 + dx3, +dz * tsp - dy2, +dz};
    pt[6] = {+dz * tcp + dy2 * ta2 - dx4, +dz * tsp + dy2, +dz};
    pt[7] = {+dz * tcp + dy2 * ta2 + dx4, +dz * tsp + dy2, +dz};
  }
  static bool Plane(V3 p1, V3 p2, V3 p3, V3 p4, double &a, double &b, double &c, double &d) {
    const V3 v12 = p2 - p1, v13 = p3 - p1, v14 = p4 - p1;
    const V3 vc{v12.y * v13.z - v12.z * v13.y, v12.z * v13.x - v12.x * v13.z, v12.x * v13.y - v12.y * v13.x};
    const double vcm = std::sqrt(vc.x * vc.x + vc.y * vc.y + vc.z * vc.z);
    const double v14m = std::sqrt(v14.x * v14.x + v14.y * v14.y + v14.z * v14.z);
    if (std::fabs((v14.x * vc.x + v14.y * vc.y + v14.z * vc.z) / (vcm * v14m)) > 1.0e-7) return false;
    a = +(p3.y - p1.y) * (p4.z - p2.z) - (p4.y - p2.y) * (p3.z - p1.z);
    b = -(p3.x - p1.x) * (p4.z - p2.z) + (p4.x - p2.x) * (p3.z - p1.z);
    c = +(p3.x - p1.x) * (p4.y - p2.y) - (p4.x - p2.x) * (p3.y - p1.y);
    const double norm = 1.0 / std::sqrt(a * a + b * b + c * c);
    a *= norm;
    b *= norm;
    c *= norm;
    d = -(a * p1.x + b * p1.y + c * p1.z);
    return true;
  }
};


// ------------------------------------------------------------ Polyhedron
// UnplacedPolyhedron's construction (source/UnplacedPolyhedron.cpp:27-233)
// with the quadrilateral set-up of source/Quadrilaterals.cpp:49-140 and
// Planes::Set (source/Planes.cpp:38-47), written into the block layout
// PolyhedronS reads.
V3 PhCross(V3 l, V3 r) { return {l.y * r.z - l.z * r.y, l.z * r.x - l.x * r.z, l.x * r.y - l.y * r.x}; }
double PhMag(V3 v) { return std::sqrt(v.Mag2()); }
V3 PhUnit(V3 v) { return v * (1. / PhMag(v)); }
V3 PhFixZeroes(V3 v) {
  double c[3] = {v.x, v.y, v.z};
  for (double &x : c)
    if (std::fabs(x) < kTolerance) x = 0.;
  return V3{c[0], c[1], c[2]};
}
// Vector3D equality is within kTolerance per component (base/Vector3D.h:751-757).
bool PhSame(V3 a, V3 b) {
  return std::fabs(a.x - b.x) < kTolerance && std::fabs(a.y - b.y) < kTolerance && std::fabs(a.z - b.z) < kTolerance;
}
double PhNormalizeAngle(double a) { return a + 2 * kPi * ((a < 0) - int(a / (2 * kPi))); }

// One quadrilateral record: plane, four side planes, four corners.
void PhQuad(double *r, V3 c0, V3 c1, V3 c2, V3 c3) {
  const V3 cs[4] = {c0, c1, c2, c3};
  // Three corners for the plane, skipping a corner equal to the previous one.
  V3 ch[4];
  ch[0] = cs[0];
  int assigned = 1, tested = 1;
  while (tested <= 4) {
    do {
      ch[assigned] = cs[tested];
      tested++;
    } while (PhSame(ch[assigned], ch[assigned - 1]) && tested < 4);
    assigned++;
    if (assigned == 3) break;
  }
  const double a = ch[0].y * (ch[1].z - ch[2].z) + ch[1].y * (ch[2].z - ch[0].z) + ch[2].y * (ch[0].z - ch[1].z);
  const double b = ch[0].z * (ch[1].x - ch[2].x) + ch[1].z * (ch[2].x - ch[0].x) + ch[2].z * (ch[0].x - ch[1].x);
  const double c = ch[0].x * (ch[1].y - ch[2].y) + ch[1].x * (ch[2].y - ch[0].y) + ch[2].x * (ch[0].y - ch[1].y);
  double d = -ch[0].x * (ch[1].y * ch[2].z - ch[2].y * ch[1].z) - ch[1].x * (ch[2].y * ch[0].z - ch[0].y * ch[2].z) -
             ch[2].x * (ch[0].y * ch[1].z - ch[1].y * ch[0].z);
  V3 nrm{a, b, c};
  const double inv = 1. / PhMag(nrm);
  nrm = nrm * inv;
  d *= inv;
  nrm = PhFixZeroes(nrm);
  r[0] = nrm.x; r[1] = nrm.y; r[2] = nrm.z; r[3] = d;
  for (int j = 0; j < 4; ++j) {
    const V3 s0 = cs[j], s1 = cs[(j + 1) & 3];
    double *sr = r + 4 + 4 * j;
    if (!PhSame(s1, s0)) {
      const V3 f = PhFixZeroes(PhUnit(PhCross(nrm, s1 - s0)));
      const double il = 1. / PhMag(f);
      sr[0] = il * f.x; sr[1] = il * f.y; sr[2] = il * f.z;
      sr[3] = il * -Dot(f, s0);
    } else {
      sr[0] = sr[1] = sr[2] = sr[3] = 0.;
    }
    r[20 + 3 * j] = cs[j].x; r[21 + 3 * j] = cs[j].y; r[22 + 3 * j] = cs[j].z;
  }
}
void PhFlip(double *r) {
  for (int k = 0; k < 4; ++k) r[k] = -r[k];
}

std::vector<double> BuildPolyhedron(const double *c, int n, std::string *err) {
  if (n < 4) return Fail(err, "polyhedron needs phiStart phiDelta sides nz z[nz] rmin[nz] rmax[nz]");
  const int sides = (int)c[2], nz = (int)c[3];
  if (sides < 1 || nz < 2 || n != 4 + 3 * nz) return Fail(err, "polyhedron needs sides >= 1, nz >= 2, 4 + 3 nz values");
  if (!(c[1] > 0)) return Fail(err, "polyhedron needs phiDelta > 0");
  const double *z = c + 4, *rmin = c + 4 + nz, *rmax = c + 4 + 2 * nz;
  const int nseg = nz - 1;
  for (int i = 0; i < nseg; ++i)
    if (!(z[i] <= z[i + 1])) return Fail(err, "polyhedron z planes must not decrease");
  for (int i = 0; i < nz; ++i)
    if (!(rmin[i] >= 0 && rmax[i] > 0)) return Fail(err, "polyhedron needs rmin >= 0 and rmax > 0");
  const bool hasPhi = c[1] < 360, largePhi = c[1] < 180;
  struct SegCount {
    bool inner;
    int nOuter, nPhi, nInner;
  };
  std::vector<SegCount> sc(nseg);
  bool hasInnerRadii = false;
  for (int i = 0; i < nseg; ++i) {
    sc[i].inner = rmin[i] > 0 || rmin[i + 1] > 0;
    sc[i].nOuter = (z[i] == z[i + 1] && rmax[i] == rmax[i + 1]) ? 0 : sides;
    sc[i].nPhi = hasPhi ? (z[i] == z[i + 1] ? 0 : 2) : 0;
    sc[i].nInner = sc[i].inner ? ((z[i] == z[i + 1] && rmin[i] == rmin[i + 1]) ? 0 : sides) : 0;
    hasInnerRadii |= sc[i].inner;
  }
  const double phiStart = PhNormalizeAngle(kDegToRad * c[0]);
  double phiDelta = c[1] * kDegToRad;
  if (phiDelta > 2 * kPi) phiDelta = 2 * kPi;
  const double sidePhi = phiDelta / sides;
  std::vector<double> vphi(sides + 1);
  std::vector<V3> phiSec(sides + 1);
  for (int i = 0; i <= sides; ++i) {
    vphi[i] = PhNormalizeAngle(phiStart + i * sidePhi);
    const V3 cv = PhFixZeroes(PhUnit(V3{1. * std::cos(vphi[i]), 1. * std::sin(vphi[i]), 0}));
    phiSec[i] = PhCross(PhUnit(cv), V3{0, 0, -1});
  }
  if (!hasPhi) vphi[sides] = vphi[0];
  const double cosHalf = std::cos(0.5 * sidePhi);
  double innerRadius = kInfinity, outerRadius = -kInfinity;
  for (int i = 0; i < nz; ++i) {
    if (rmin[i] < innerRadius) innerRadius = rmin[i];
    if (rmax[i] > outerRadius) outerRadius = rmax[i];
  }
  outerRadius /= cosHalf;
  const double tubeZ = z[nz - 1] - z[0] + 2. * kTolerance;
  TubeP tube = MakeTube(innerRadius - kTolerance, outerRadius + kTolerance, 0.5 * tubeZ, !hasPhi ? 0 : phiStart,
                        !hasPhi ? 2 * kPi : phiDelta);
  // The reference bounds the polyhedron with a HollowTube or NonHollowTube
  // kernel: rmin treated iff there are inner radii, phi never.
  tube.hasRmin = hasInnerRadii ? 1. : 0.;
  tube.hasPhi = 0.;

  // Layout.
  const int tubeDoubles = (int)((sizeof(TubeP) + sizeof(double) - 1) / sizeof(double));
  const int zOff = PolyhedronS::kTubeOff + ((tubeDoubles + 1) & ~1);
  const int phiOff = zOff + ((nz + 1) & ~1);
  const int segOff = phiOff + ((3 * (sides + 1) + 1) & ~1);
  // Segment k's quadrilaterals start at an offset of k * sides (mod 32)
  // doubles: with 33-double records, the lanes of a wave that read the same
  // quadrilateral slot of different segments (or different quadrilaterals of
  // up to 32 / sides segments) then read different LDS banks.
  const int quadBase = segOff + PolyhedronS::kSegDoubles * nseg;
  std::vector<int> segQuadOff(nseg);
  int end = quadBase;
  for (int ip = 0; ip < nseg; ++ip) {
    const int want = (ip * sides) & 31;
    end += (want - end) & 31;
    segQuadOff[ip] = end;
    end += PolyhedronS::kQuadDoubles * (sc[ip].nOuter + sc[ip].nPhi + sc[ip].nInner);
  }
  std::vector<double> blk(end, 0.);
  blk[0] = nseg; blk[1] = sides; blk[2] = hasPhi; blk[3] = largePhi; blk[4] = z[0] + 0.5 * tubeZ;
  blk[5] = zOff; blk[6] = phiOff; blk[7] = segOff; blk[8] = (double)blk.size();
  std::memcpy(blk.data() + PolyhedronS::kTubeOff, &tube, sizeof(TubeP));
  for (int i = 0; i < nz; ++i) blk[zOff + i] = z[i];
  for (int i = 0; i <= sides; ++i) {
    blk[phiOff + 3 * i] = phiSec[i].x; blk[phiOff + 3 * i + 1] = phiSec[i].y; blk[phiOff + 3 * i + 2] = phiSec[i].z;
  }
  // Corners: rMin / rMax are distances to the sides, corners lie at r / cos(half side angle).
  auto vi = [&](int plane, int corner) { return plane * (sides + 1) + corner; };
  std::vector<V3> ov(nz * (sides + 1)), iv(nz * (sides + 1));
  for (int i = 0; i < nz; ++i) {
    for (int j = 0, jm = sides + (hasPhi ? 1 : 0); j < jm; ++j) {
      const double ro = rmax[i] / cosHalf, ri = rmin[i] / cosHalf;
      ov[vi(i, j)] = PhFixZeroes(V3{ro * std::cos(vphi[j]), ro * std::sin(vphi[j]), z[i]});
      iv[vi(i, j)] = PhFixZeroes(V3{ri * std::cos(vphi[j]), ri * std::sin(vphi[j]), z[i]});
    }
    if (!hasPhi) {
      ov[vi(i, sides)] = ov[vi(i, 0)];
      iv[vi(i, sides)] = iv[vi(i, 0)];
    }
  }
  // Normals of the side faces point away from the z axis.
  auto wrong = [](const double *r, V3 corner) { return r[0] * corner.x + r[1] * corner.y < 0; };
  for (int ip = 0; ip < nseg; ++ip) {
    double *rec = blk.data() + segOff + PolyhedronS::kSegDoubles * ip;
    rec[0] = sc[ip].inner; rec[1] = sc[ip].nOuter; rec[2] = sc[ip].nPhi; rec[3] = sc[ip].nInner;
    // Annular slab holding the segment's quadrilaterals: inner side distance
    // and outer corner radius.
    rec[7] = sc[ip].inner ? std::min(rmin[ip], rmin[ip + 1]) : 0.;
    rec[8] = std::max(std::max(rmax[ip], rmax[ip + 1]), std::max(rmin[ip], rmin[ip + 1])) / cosHalf;
    rec[9] = std::max(rmin[ip], rmin[ip + 1]) / cosHalf;
    int quadOff = segQuadOff[ip];
    rec[4] = quadOff;
    for (int k = 0; k < sc[ip].nOuter; ++k) {
      double *q = blk.data() + quadOff + PolyhedronS::kQuadDoubles * k;
      PhQuad(q, ov[vi(ip, k)], ov[vi(ip, k + 1)], ov[vi(ip + 1, k + 1)], ov[vi(ip + 1, k)]);
      if (wrong(q, ov[vi(ip, k)])) PhFlip(q);
    }
    quadOff += PolyhedronS::kQuadDoubles * sc[ip].nOuter;
    rec[5] = quadOff;
    if (sc[ip].nPhi == 2) {
      double *q0 = blk.data() + quadOff, *q1 = q0 + PolyhedronS::kQuadDoubles;
      PhQuad(q0, iv[vi(ip, 0)], iv[vi(ip + 1, 0)], ov[vi(ip + 1, 0)], ov[vi(ip, 0)]);
      if (q0[0] * phiSec[0].x + q0[1] * phiSec[0].y + q0[2] * phiSec[0].z > 0) PhFlip(q0);
      PhQuad(q1, ov[vi(ip, sides)], ov[vi(ip + 1, sides)], iv[vi(ip + 1, sides)], iv[vi(ip, sides)]);
      if (q1[0] * phiSec[sides].x + q1[1] * phiSec[sides].y + q1[2] * phiSec[sides].z < 0) PhFlip(q1);
    }
    quadOff += PolyhedronS::kQuadDoubles * sc[ip].nPhi;
    rec[6] = quadOff;
    for (int k = 0; k < sc[ip].nInner; ++k) {
      double *q = blk.data() + quadOff + PolyhedronS::kQuadDoubles * k;
      PhQuad(q, iv[vi(ip, k)], iv[vi(ip, k + 1)], iv[vi(ip + 1, k + 1)], iv[vi(ip + 1, k)]);
      if (wrong(q, iv[vi(ip, k)])) PhFlip(q);
    }
    quadOff += PolyhedronS::kQuadDoubles * sc[ip].nInner;
  }
  return blk;
}

}  // namespace

// UnplacedCone constructor (volumes/UnplacedCone.h:131-197): the cached
// values of a cone with parameters c = rmin1 rmax1 rmin2 rmax2 dz sphi dphi.
ConeP MakeConeP(const double *c) {
  ConeP q;
  std::memset(&q, 0, sizeof(q));
  const double kAngTol = 1e-9;
  q.rmin1 = c[0];
  q.rmax1 = c[1];
  q.rmin2 = c[2];
  q.rmax2 = c[3];
  q.dz = c[4];
  q.sphi = c[5];
  q.dphi = c[6];
  q.w = MakeWedge(q.dphi, q.sphi);
  const double hDPhi = 0.5 * q.dphi;
  const double cPhi = q.sphi + hDPhi;
  const double ePhi = q.sphi + q.dphi;
  q.sinCPhi = std::sin(cPhi);
  q.cosCPhi = std::cos(cPhi);
  q.cosHDPhi = std::cos(hDPhi);
  q.cosHDPhiIT = std::cos(hDPhi - 0.5 * kAngTol);
  q.cosHDPhiOT = std::cos(hDPhi + 0.5 * kAngTol);
  q.sinSPhi = std::sin(q.sphi);
  q.cosSPhi = std::cos(q.sphi);
  q.sinEPhi = std::sin(ePhi);
  q.cosEPhi = std::cos(ePhi);
  q.tanRMin = (q.rmin2 - q.rmin1) * 0.5 / q.dz;
  q.secRMin = std::sqrt(1.0 + q.tanRMin * q.tanRMin);
  q.invSecRMin = 1. / q.secRMin;
  q.tanRMax = (q.rmax2 - q.rmax1) * 0.5 / q.dz;
  q.secRMax = std::sqrt(1.0 + q.tanRMax * q.tanRMax);
  q.invSecRMax = 1. / q.secRMax;
  q.innerSlope = -(q.rmin1 - q.rmin2) / (2. * q.dz);
  q.outerSlope = -(q.rmax1 - q.rmax2) / (2. * q.dz);
  q.innerOffset = q.rmin2 - q.innerSlope * q.dz;
  q.outerOffset = q.rmax2 - q.outerSlope * q.dz;
  q.fullPhi = q.dphi == kTwoPi ? 1. : 0.;
  q.phiTreatment = q.dphi < 2. * M_PI ? 1. : 0.;
  q.rminTreatment = (q.rmin1 > 0 || q.rmin2 > 0) ? 1. : 0.;
  return q;
}

std::vector<double> BuildZooParams(int type, const double *c, int n, std::string *err) {
  switch (type) {
    case VG_ORB: {
      if (n != 1 || !(c[0] > 0)) return Fail(err, "orb needs 1 parameter r > 0");
      // UnplacedOrb::SetRadius, source/UnplacedOrb.cpp:48-56
      OrbP o;
      o.r = c[0];
      const double tol = Max(1e-9, 2e-11 * o.r);
      o.rTolI = o.r - tol;
      o.rTolO = o.r + tol;
      o.pad = 0;
      return Pack(o);
    }
    case VG_TRD: {
      if (n != 5 || !(c[0] >= 0 && c[1] >= 0 && c[2] >= 0 && c[3] >= 0 && c[4] > 0))
        return Fail(err, "trd needs 5 parameters dx1 dx2 dy1 dy2 dz (dz > 0)");
      // UnplacedTrd::calculateCached, volumes/UnplacedTrd.h:38-48
      TrdP t;
      t.dx1 = c[0]; t.dx2 = c[1]; t.dy1 = c[2]; t.dy2 = c[3]; t.dz = c[4];
      t.x2minusx1 = t.dx2 - t.dx1;
      t.y2minusy1 = t.dy2 - t.dy1;
      t.halfx1plusx2 = 0.5 * (t.dx1 + t.dx2);
      t.halfy1plusy2 = 0.5 * (t.dy1 + t.dy2);
      t.dztimes2 = t.dz * 2;
      t.fx = 0.5 * (t.dx1 - t.dx2) / t.dz;
      t.fy = 0.5 * (t.dy1 - t.dy2) / t.dz;
      t.calfx = 1. / std::sqrt(1.0 + t.fx * t.fx);
      t.calfy = 1. / std::sqrt(1.0 + t.fy * t.fy);
      t.varyingY = t.dy1 != t.dy2 ? 1. : 0.;
      t.pad = 0;
      return Pack(t);
    }
    case VG_PARALLELEPIPED: {
      if (n != 6 || !(c[0] > 0 && c[1] > 0 && c[2] > 0))
        return Fail(err, "parallelepiped needs x y z (> 0) alpha theta phi [deg]");
      // source/UnplacedParallelepiped.cpp:48-70; safety factors of
      // ParallelepipedImplementation.h:318-352
      ParaP q;
      q.d[0] = c[0]; q.d[1] = c[1]; q.d[2] = c[2];
      q.alpha = c[3]; q.theta = c[4]; q.phi = c[5];
      q.tanAlpha = std::tan(kDegToRad * c[3]);
      q.tanThetaCosPhi = std::tan(kDegToRad * c[4]) * std::cos(kDegToRad * c[5]);
      q.tanThetaSinPhi = std::tan(kDegToRad * c[4]) * std::sin(kDegToRad * c[5]);
      q.ctx = 1.0 / std::sqrt(1. + q.tanAlpha * q.tanAlpha + q.tanThetaCosPhi * q.tanThetaCosPhi);
      q.cty = 1.0 / std::sqrt(1. + q.tanThetaSinPhi * q.tanThetaSinPhi);
      q.pad = 0;
      return Pack(q);
    }
    case VG_PARABOLOID: {
      if (n != 3 || !(c[0] >= 0 && c[1] >= 0 && c[2] > 0 && c[1] != c[0]))
        return Fail(err, "paraboloid needs rlo >= 0, rhi >= 0, rlo != rhi, dz > 0");
      // UnplacedParaboloid::SetRloAndRhiAndDz, source/UnplacedParaboloid.cpp:92-127
      ParabP q;
      q.rlo = c[0]; q.rhi = c[1]; q.dz = c[2];
      q.rlo2 = q.rlo * q.rlo;
      q.rhi2 = q.rhi * q.rhi;
      const double dd = 1. / (q.rhi2 - q.rlo2);
      q.a = 2. * q.dz * dd;
      q.b = -q.dz * (q.rlo2 + q.rhi2) * dd;
      q.ainv = 1 / q.a;
      q.a2 = q.a * q.a;
      q.tolOz = q.dz + kHalfTolerance;
      q.dx = Max(q.rhi, q.rlo);
      q.dy = q.dx;
      return Pack(q);
    }
    case VG_SPHERE: {
      if (n != 6 || !(c[0] >= 0 && c[1] > c[0] && c[3] > 0 && c[5] > 0))
        return Fail(err, "sphere needs 0 <= rmin < rmax, dphi > 0, dtheta > 0");
      // source/UnplacedSphere.cpp:25-95, volumes/UnplacedSphere.h:63-215
      SphereP q;
      std::memset(&q, 0, sizeof(q));
      const double kAngTol = 1e-9;
      q.fRmin = c[0];
      q.fRmax = c[1];
      q.fRminTolerance = (q.fRmin) ? Max(1e-9, 2.e-11 * q.fRmin) : 0;
      bool fullPhi = true;
      if (c[3] >= 2 * kPi - kAngTol * 0.5) {
        q.fDPhi = 2 * kPi;
        q.fSPhi = 0;
      } else {
        fullPhi = false;
        q.fDPhi = c[3];
        const double sPhi = c[2];
        q.fSPhi = sPhi < 0 ? 2 * kPi - std::fmod(std::fabs(sPhi), 2 * kPi) : std::fmod(sPhi, 2 * kPi);
        if (q.fSPhi + q.fDPhi > 2 * kPi) q.fSPhi -= 2 * kPi;
      }
      q.hDPhi = 0.5 * q.fDPhi;
      q.cPhi = q.fSPhi + q.hDPhi;
      q.ePhi = q.fSPhi + q.fDPhi;
      q.sinCPhi = std::sin(q.cPhi);
      q.cosCPhi = std::cos(q.cPhi);
      q.sinSPhi = std::sin(q.fSPhi);
      q.cosSPhi = std::cos(q.fSPhi);
      q.sinEPhi = std::sin(q.ePhi);
      q.cosEPhi = std::cos(q.ePhi);
      const double sTheta = c[4], dTheta = c[5];
      if (!((sTheta < 0) || (sTheta > kPi))) q.fSTheta = sTheta;
      if (dTheta + sTheta >= kPi)
        q.fDTheta = kPi - sTheta;
      else
        q.fDTheta = dTheta;
      const bool fullTheta = !(q.fDTheta - q.fSTheta < kPi);
      q.eTheta = q.fSTheta + q.fDTheta;
      q.sinSTheta = std::sin(q.fSTheta);
      q.cosSTheta = std::cos(q.fSTheta);
      q.sinETheta = std::sin(q.eTheta);
      q.cosETheta = std::cos(q.eTheta);
      q.cosHDPhi = std::cos(q.hDPhi);
      const double phim = 0.5 * (q.fSPhi + q.ePhi);
      q.sinPhim = std::sin(phim);
      q.cosPhim = std::cos(phim);
      q.fullPhi = fullPhi ? 1. : 0.;
      q.fullTheta = fullTheta ? 1. : 0.;
      return Pack(q);
    }
    case VG_CONE: {
      if (n != 7 || !(c[4] > 0 && c[0] >= 0 && c[2] >= 0 && c[1] >= c[0] && c[3] >= c[2] && c[6] > 0))
        return Fail(err, "cone needs dz > 0, 0 <= rmin <= rmax at both ends, dphi > 0");
      return Pack(MakeConeP(c));
    }
    case VG_TORUS: {
      // volumes/UnplacedTorus.h:43-69; a ring torus (rtor > rmax) keeps the
      // bounding tube's inner radius positive.
      if (n != 5 || !(c[0] >= 0 && c[1] > c[0] && c[2] > c[1] + kTolerance && c[4] > 0))
        return Fail(err, "torus needs 0 <= rmin < rmax < rtor and dphi > 0");
      TorusP q;
      std::memset(&q, 0, sizeof(q));
      q.rmin = c[0];
      q.rmax = c[1];
      q.rtor = c[2];
      q.sphi = c[3];
      q.dphi = c[4];
      q.rmin2 = q.rmin * q.rmin;
      q.rmax2 = q.rmax * q.rmax;
      q.rtor2 = q.rtor * q.rtor;
      q.w = MakeWedge(q.dphi, q.sphi);
      // The reference's DistanceToIn uses the HollowTube specialisation of the
      // bounding tube: rmin treated, phi never.
      q.btube = MakeTube(q.rtor - q.rmax - kTolerance, q.rtor + q.rmax + kTolerance, q.rmax, 0., kTwoPi);
      return Pack(q);
    }
    case VG_POLYCONE: {
      // UnplacedPolycone::Init, source/UnplacedPolycone.cpp:27-180: one cone
      // section per rise in z, shifted to the middle of its z range.
      if (n < 3) return Fail(err, "polycone needs sphi dphi nz rmin[nz] rmax[nz] z[nz]");
      const int nz = (int)c[2];
      if (nz < 2 || n != 3 + 3 * nz) return Fail(err, "polycone needs nz >= 2 and 3 + 3 nz parameters");
      const double *rmin = c + 3, *rmax = c + 3 + nz, *z = c + 3 + 2 * nz;
      for (int i = 0; i < nz; ++i)
        if (!(rmin[i] >= 0 && rmax[i] >= rmin[i])) return Fail(err, "polycone needs 0 <= rmin <= rmax per plane");
      if (!(c[1] > 0)) return Fail(err, "polycone needs dphi > 0");
      std::vector<double> zs;
      std::vector<ConeP> cones;
      std::vector<double> shifts;
      double prevZ = z[0], prevRmax = 0, prevRmin = 0;
      const int dirZ = z[1] < z[0] ? -1 : 1;
      for (int i = 0; i < nz; ++i) {
        if (i == 0) {
          zs.push_back(z[i]);
        } else if (((z[i] > prevZ) && (dirZ > 0)) || ((z[i] < prevZ) && (dirZ < 0))) {
          const double cc[7] = {prevRmin, prevRmax, rmin[i], rmax[i], (z[i] - prevZ) / 2, c[0], c[1]};
          cones.push_back(MakeConeP(cc));
          zs.push_back(z[i]);
          const size_t zi = zs.size() - 1;
          shifts.push_back(zs[zi - 1] + 0.5 * (zs[zi] - zs[zi - 1]));
        }
        prevZ = z[i];
        prevRmin = rmin[i];
        prevRmax = rmax[i];
      }
      const int ns = (int)cones.size();
      if (ns == 0) return Fail(err, "polycone has no section");
      std::vector<double> blk(PolyconeS::SecBase(ns) + PolyconeS::kSecDoubles * (size_t)ns, 0.);
      blk[0] = ns;
      for (int i = 0; i <= ns; ++i) blk[2 + i] = zs[i];
      for (int i = 0; i < ns; ++i) {
        double *r = blk.data() + PolyconeS::SecBase(ns) + PolyconeS::kSecDoubles * (size_t)i;
        std::memcpy(r, &cones[i], sizeof(ConeP));
        r[40] = shifts[i];
      }
      return blk;
    }
    case VG_POLYHEDRON:
      return BuildPolyhedron(c, n, err);
    case VG_TRAPEZOID: {
      if (n != 11 || !(c[0] > 0 && c[3] > 0 && c[4] > 0 && c[5] > 0 && c[7] > 0 && c[8] > 0 && c[9] > 0))
        return Fail(err, "trapezoid needs 11 parameters with positive lengths");
      TrapBuild tb(c);
      TrapP t;
      std::memset(&t, 0, sizeof(t));
      t.dz = c[0];
      const int corners[4][4] = {{0, 1, 5, 4}, {2, 6, 7, 3}, {0, 4, 6, 2}, {1, 3, 7, 5}};
      for (int i = 0; i < 4; ++i) {
        if (!TrapBuild::Plane(tb.pt[corners[i][0]], tb.pt[corners[i][1]], tb.pt[corners[i][2]],
                              tb.pt[corners[i][3]], t.A[i], t.B[i], t.C[i], t.D[i]))
          return Fail(err, "trapezoid side face " + std::to_string(i) + " is not planar");
      }
      for (int i = 0; i < 11; ++i) t.prm[i] = c[i];
      return Pack(t);
    }
  }
  if (err) *err = "shape type " + std::to_string(type) + " has no device kernels";
  return {};
}

bool ZooExtent(int type, const double *b, double lo[3], double hi[3]) {
  switch (type) {
    case VG_ORB: {
      const OrbP *o = reinterpret_cast<const OrbP *>(b);
      for (int i = 0; i < 3; ++i) {
        lo[i] = -o->r;
        hi[i] = o->r;
      }
      return true;
    }
    case VG_TRD: {
      // UnplacedTrd::Extent, volumes/UnplacedTrd.h:147-150 (max x uses dy1,
      // as the reference does).
      const TrdP *t = reinterpret_cast<const TrdP *>(b);
      lo[0] = -Max(t->dx1, t->dx2); lo[1] = -Max(t->dy1, t->dy2); lo[2] = -t->dz;
      hi[0] = Max(t->dy1, t->dx2); hi[1] = Max(t->dy1, t->dy2); hi[2] = t->dz;
      return true;
    }
    case VG_PARALLELEPIPED: {
      // No Extent in the reference (volumes/PlacedVolume.h:260-263); bound of
      // the sheared corners.
      const ParaP *q = reinterpret_cast<const ParaP *>(b);
      const double ta = std::tan(kDegToRad * q->alpha), tt = std::tan(kDegToRad * q->theta);
      const double ex = q->d[0] + std::fabs(q->d[1] * ta) + std::fabs(q->d[2] * tt * std::cos(kDegToRad * q->phi));
      const double ey = q->d[1] + std::fabs(q->d[2] * tt * std::sin(kDegToRad * q->phi));
      lo[0] = -ex; lo[1] = -ey; lo[2] = -q->d[2];
      hi[0] = ex; hi[1] = ey; hi[2] = q->d[2];
      return true;
    }
    case VG_PARABOLOID: {
      // UnplacedParaboloid::Extent, source/UnplacedParaboloid.cpp:169-176
      const ParabP *q = reinterpret_cast<const ParabP *>(b);
      lo[0] = -q->dx; lo[1] = -q->dy; lo[2] = -q->dz;
      hi[0] = q->dx; hi[1] = q->dy; hi[2] = q->dz;
      return true;
    }
    case VG_CONE: {
      // UnplacedCone::Extent, volumes/UnplacedCone.h:324-328
      const ConeP *q = reinterpret_cast<const ConeP *>(b);
      const double m = q->rmax1 > q->rmax2 ? q->rmax1 : q->rmax2;
      lo[0] = -m; lo[1] = -m; lo[2] = -q->dz;
      hi[0] = m; hi[1] = m; hi[2] = q->dz;
      return true;
    }
    case VG_TORUS: {
      // UnplacedTorus::Extent = the bounding tube's (with its phi section)
      const TorusP *q = reinterpret_cast<const TorusP *>(b);
      const TubeP t = MakeTube(q->rtor - q->rmax - kTolerance, q->rtor + q->rmax + kTolerance, q->rmax, q->sphi,
                               q->dphi);
      return ShapeExtent(VG_TUBE, reinterpret_cast<const double *>(&t), lo, hi);
    }
    case VG_POLYCONE: {
      // UnplacedPolycone::Extent, source/UnplacedPolycone.cpp:755-774
      const PolyconeP q{b};
      const int ns = PolyconeS::NSec(q);
      double maxR = 0;
      for (int i = 0; i < ns; ++i) {
        const ConeP &c = PolyconeS::Sec(q, i);
        if (maxR < c.rmax1) maxR = c.rmax1;
        if (maxR < c.rmax2) maxR = c.rmax2;
      }
      lo[0] = -maxR; lo[1] = -maxR; lo[2] = PolyconeS::Z(q, 0);
      hi[0] = maxR; hi[1] = maxR; hi[2] = PolyconeS::Z(q, ns);
      return true;
    }
    case VG_POLYHEDRON: {
      // UnplacedPolyhedron::Extent: the bounding tube's, shifted by its z
      // offset (source/UnplacedPolyhedron.cpp:324-330).
      const PolyhedronP q{b};
      const TubeP &t = PolyhedronS::Tube(q);
      if (!ShapeExtent(VG_TUBE, reinterpret_cast<const double *>(&t), lo, hi)) return false;
      lo[2] += b[4];
      hi[2] += b[4];
      return true;
    }
    case VG_SPHERE: {
      const SphereP *q = reinterpret_cast<const SphereP *>(b);
      for (int i = 0; i < 3; ++i) {
        lo[i] = -q->fRmax;
        hi[i] = q->fRmax;
      }
      return true;
    }
    case VG_TRAPEZOID: {
      // UnplacedTrapezoid::Extent, source/UnplacedTrapezoid.cpp:428-466
      const TrapP *t = reinterpret_cast<const TrapP *>(b);
      TrapBuild tb(t->prm);
      for (int ax = 0; ax < 2; ++ax) {
        auto g = [&](int k) { return ax == 0 ? tb.pt[k].x : tb.pt[k].y; };
        double e01 = std::max(g(0), g(1)), e23 = std::max(g(2), g(3));
        double e45 = std::max(g(4), g(5)), e67 = std::max(g(6), g(7));
        double eA = e01 > e23 ? e01 : e23, eB = e45 > e67 ? e45 : e67;
        hi[ax] = (eA > eB) ? eA : eB;
        e01 = std::min(g(0), g(1)); e23 = std::min(g(2), g(3));
        e45 = std::min(g(4), g(5)); e67 = std::min(g(6), g(7));
        eA = e01 < e23 ? e01 : e23; eB = e45 < e67 ? e45 : e67;
        lo[ax] = (eA < eB) ? eA : eB;
      }
      lo[2] = -t->dz;
      hi[2] = t->dz;
      return true;
    }
  }
  return false;
}

}  // namespace vga
