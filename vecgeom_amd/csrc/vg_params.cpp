// Host-side construction of shape parameter blocks: the constructor
// parameters of each Unplaced* class plus the values the reference caches at
// construction, computed with the same expressions on the host (glibc
// cos/sin, IEEE double) so that device kernels see identical constants.
#include <cmath>
#include <algorithm>
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

bool Fail(std::string *err, const std::string &m) {
  if (err) *err = m;
  return false;
}



// UnplacedTube::Extent, source/UnplacedTube.cpp:244-296.
void TubeExtent(const TubeP &t, double lo[3], double hi[3]) {
  lo[0] = -t.rmax; lo[1] = -t.rmax; lo[2] = -t.z;
  hi[0] = t.rmax; hi[1] = t.rmax; hi[2] = t.z;
  if (t.dphi == kTwoPi) return;
  const double rin = 0.5 * (t.rmax + t.rmin);
  auto contains = [&](double x, double y) {
    bool cin = false, cout;
    WedgeS::Kernel<false>(t.w, V3{x, y, 0}, cin, cout);
    return !cout;
  };
  const bool phi0out = !contains(rin, 0), phi90out = !contains(0, rin);
  const bool phi180out = !contains(-rin, 0), phi270out = !contains(0, -rin);
  if (!(phi0out || phi90out || phi180out || phi270out)) return;
  const double cx = t.rmax * std::cos(t.sphi), dx = t.rmax * std::cos(t.sphi + t.dphi);
  const double cy = t.rmax * std::sin(t.sphi), dy = t.rmax * std::sin(t.sphi + t.dphi);
  if (phi0out) hi[0] = Max(cx, dx);
  if (phi90out) hi[1] = Max(cy, dy);
  if (phi180out) lo[0] = Min(cx, dx);
  if (phi270out) lo[1] = Min(cy, dy);
  if (t.dphi >= kPi) return;
  const double ax = t.rmin * std::cos(t.sphi), bx = t.rmin * std::cos(t.sphi + t.dphi);
  const double ay = t.rmin * std::sin(t.sphi), by = t.rmin * std::sin(t.sphi + t.dphi);
  double tmp = Max(ax, bx);
  hi[0] = tmp > hi[0] ? tmp : hi[0];
  tmp = Max(ay, by);
  hi[1] = tmp > hi[1] ? tmp : hi[1];
  tmp = Min(ax, bx);
  lo[0] = tmp < lo[0] ? tmp : lo[0];
  tmp = Min(ay, by);
  lo[1] = tmp < lo[1] ? tmp : lo[1];
}

}  // namespace

// volumes/UnplacedTube.h:41-60 (calculateCached) and constructor :64-88.
TubeP MakeTube(double rmin, double rmax, double z, double sphi, double dphi) {
  TubeP t;
  t.rmin = rmin;
  t.rmax = rmax;
  t.z = z;
  t.sphi = sphi;
  t.dphi = dphi;
  t.tolIz = z - kTolerance;
  t.tolOz = z + kTolerance;
  t.rmin2 = rmin * rmin;
  t.rmax2 = rmax * rmax;
  const double tolOrmin = rmin - kTolerance, tolIrmin = rmin + kTolerance;
  t.tolOrmin2 = tolOrmin * tolOrmin;
  t.tolIrmin2 = tolIrmin * tolIrmin;
  const double tolOrmax = rmax + kTolerance, tolIrmax = rmax - kTolerance;
  t.tolOrmax2 = tolOrmax * tolOrmax;
  t.tolIrmax2 = tolIrmax * tolIrmax;
  t.a1x = std::cos(sphi);
  t.a1y = std::sin(sphi);
  t.a2x = std::cos(sphi + dphi);
  t.a2y = std::sin(sphi + dphi);
  t.w = MakeWedge(dphi, sphi);
  // shapetypes/TubeTypes.h:154-160, 196-201 and TubeImplementation.h:71-74
  t.hasRmin = rmin > 0 ? 1. : 0.;
  t.hasPhi = dphi < 2 * M_PI ? 1. : 0.;
  t.smallerThanPi = dphi <= M_PI ? 1. : 0.;
  return t;
}

// source/Wedge.cpp:16-43
WedgeP MakeWedge(double angle, double zero) {
  WedgeP w;
  w.dphi = angle;
  w.a1x = std::cos(zero);
  w.a1y = std::sin(zero);
  w.a2x = std::cos(zero + angle);
  w.a2y = std::sin(zero + angle);
  w.n1x = -std::sin(zero);
  w.n1y = std::cos(zero);
  w.n2x = std::sin(zero + angle);
  w.n2y = -std::cos(zero + angle);
  return w;
}

// Boolean block (shapes_boolean.hpp): op, per operand type / block offset /
// placement kind, total length, the two placements, then the operands'
// blocks, each starting at an even offset (16-byte aligned).
std::vector<double> BuildBooleanParams(const double *c, int n, std::string *err) {
  std::vector<double> none;
  if (n < 1 || !(c[0] == 0 || c[0] == 1 || c[0] == 2))
    return Fail(err, "Boolean needs op 0 (union), 1 (intersection) or 2 (subtraction)"), none;
  std::vector<double> blk(32, 0.);
  blk[0] = c[0];
  int at = 1;
  for (int k = 0; k < 2; ++k) {
    if (at + 2 > n) return Fail(err, "Boolean operand record truncated"), none;
    const int t = (int)c[at], np = (int)c[at + 1];
    if (t == VG_BOOLEAN) return Fail(err, "Boolean operands must be primitive shapes"), none;
    if (np < 0 || at + 2 + np + 12 > n) return Fail(err, "Boolean operand record truncated"), none;
    std::vector<double> ob = BuildParams(t, c + at + 2, np, err);
    if (ob.empty()) return none;
    const double *x = c + at + 2 + np;
    blk.resize((blk.size() + 1) & ~size_t(1), 0.);
    blk[1 + 3 * k] = t;
    blk[2 + 3 * k] = (double)blk.size();
    blk[3 + 3 * k] = PlacementKind(x);
    for (int i = 0; i < 12; ++i) blk[8 + 12 * k + i] = x[i];
    blk.insert(blk.end(), ob.begin(), ob.end());
    at += 2 + np + 12;
  }
  if (at != n) return Fail(err, "Boolean parameter count does not match its operands"), none;
  blk.resize((blk.size() + 1) & ~size_t(1), 0.);
  blk[7] = (double)blk.size();
  return blk;
}

std::vector<double> BuildParams(int type, const double *c, int n, std::string *err) {
  std::vector<double> none;
  switch (type) {
    case VG_BOX: {
      if (n != 3) return Fail(err, "box needs 3 parameters (dx dy dz)"), none;
      if (!(c[0] > 0 && c[1] > 0 && c[2] > 0)) return Fail(err, "box half-lengths must be > 0"), none;
      BoxP b{{c[0], c[1], c[2]}};
      return Pack(b);
    }
    case VG_TUBE: {
      if (n != 5) return Fail(err, "tube needs 5 parameters (rmin rmax dz sphi dphi)"), none;
      if (!(c[0] >= 0 && c[1] > c[0] && c[2] > 0 && c[4] > 0))
        return Fail(err, "tube needs 0 <= rmin < rmax, dz > 0, dphi > 0"), none;
      return Pack(MakeTube(c[0], c[1], c[2], c[3], c[4]));
    }
    case VG_BOOLEAN:
      return BuildBooleanParams(c, n, err);
    default:
      return BuildZooParams(type, c, n, err);
  }
}

bool ShapeExtent(int type, const double *b, double lo[3], double hi[3]) {
  switch (type) {
    case VG_BOX: {
      const BoxP *p = reinterpret_cast<const BoxP *>(b);
      for (int i = 0; i < 3; ++i) {
        lo[i] = -p->d[i];
        hi[i] = p->d[i];
      }
      return true;
    }
    case VG_TUBE: {
      const TubeP *p = reinterpret_cast<const TubeP *>(b);
      TubeExtent(*p, lo, hi);
      return true;
    }
    case VG_BOOLEAN: {
      // source/UnplacedBooleanVolume.cpp:94-122: the operands' unplaced
      // extents (their placements not applied), combined per operation.
      double la[3], ha[3], lb[3], hb[3];
      if (!ShapeExtent((int)b[1], b + (int)b[2], la, ha) || !ShapeExtent((int)b[4], b + (int)b[5], lb, hb)) return false;
      for (int i = 0; i < 3; ++i) {
        if (b[0] == 0) {
          lo[i] = std::min(la[i], lb[i]);
          hi[i] = std::max(ha[i], hb[i]);
        } else if (b[0] == 1) {
          lo[i] = std::max(la[i], lb[i]);
          hi[i] = std::min(ha[i], hb[i]);
        } else {
          lo[i] = la[i];
          hi[i] = ha[i];
        }
      }
      return true;
    }
    default:
      return ZooExtent(type, b, lo, hi);
  }
}

}  // namespace vga
