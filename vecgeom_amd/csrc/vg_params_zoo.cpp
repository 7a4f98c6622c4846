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
