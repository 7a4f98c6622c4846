// Host-side construction of the FP32 bound records (vg_bound32.hpp) of every
// placed volume, built with the geometry at upload.
//
// Data layout of B32::s per shape kind (all in the shape's local frame):
//   Box            dx, dy, dz
//   Orb            r
//   Tube           rmin, rmax, dz, hasRmin, wedge[5]
//   Cone           aI, bI, cI, aO, bO, cO, hasRmin, dz, wedge[5]
//                  (rmin(z) = aI z + bI, rmax(z) = aO z + bO, c = 1/sqrt(1+a^2))
//   Sphere         rmin, rmax, hasTheta1, sin1, cos1, hasTheta2, sin2, cos2, wedge[5]
//   Torus          rmin, rmax, rtor, wedge[5]
//   Trd            hx0, sx, cx, hy0, sy, cy, dz  (half-widths hx0 + sx z, hy0 + sy z)
//   Trapezoid      dz, then 4 side planes (unit outward normal, offset)
//   Parallelepiped dx, dy, dz, tanAlpha, tanThetaSinPhi, tanThetaCosPhi, 1/|ny|, 1/|nx|
//   Paraboloid     a, b, dz  (inside: z >= a rho^2 + b, |z| <= dz)
// and in the side array (B32::off):
//   Polycone       nsec, wedge[5], then per section zlo, zhi, cone[7]
//   Polyhedron     nseg, the offset of each segment record (relative to the
//                  start), then per segment zlo, zhi, nout, nin, nout + nin
//                  planes (unit normal, offset), oriented with the z axis at
//                  the segment's mid-plane on the negative side (outer planes:
//                  the solid side; inner planes: the hole's side)
// wedge = mode (0 none, 1 opening <= pi: intersection, 2: union), inward
// unit normals n1 = (-sin s, cos s), n2 = (sin e, -cos e) of the start and
// end half-planes (s, e the start and end angles).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "shapes.hpp"
#include "vg_bound32.hpp"
#include "vg_internal.hpp"

namespace vga {

namespace {

void PutWedge(float *w, bool treated, double sphi, double dphi) {
  for (int i = 0; i < 5; ++i) w[i] = 0.f;
  if (!treated || dphi >= kTwoPi) return;
  const double e = sphi + dphi;
  w[0] = dphi <= kPi ? kPhiConvex : kPhiUnion;
  w[1] = (float)-std::sin(sphi);
  w[2] = (float)std::cos(sphi);
  w[3] = (float)std::sin(e);
  w[4] = (float)-std::cos(e);
}

// Cone lines of a ConeP whose local z is shifted by `shift` from the frame
// the bound is evaluated in.
void PutCone(float *c, const ConeP &q, double shift) {
  c[0] = (float)q.innerSlope;
  c[1] = (float)(q.innerOffset - q.innerSlope * shift);
  c[2] = (float)(1. / std::sqrt(1. + q.innerSlope * q.innerSlope));
  c[3] = (float)q.outerSlope;
  c[4] = (float)(q.outerOffset - q.outerSlope * shift);
  c[5] = (float)(1. / std::sqrt(1. + q.outerSlope * q.outerSlope));
  c[6] = q.rminTreatment != 0 ? 1.f : 0.f;
}

// A plane n.p + d scaled to a unit normal, oriented so that `ref` lies on
// its negative side.
void PutPlane(float *o, V3 n, double d, V3 ref, double *dmax) {
  const double len = std::sqrt(Dot(n, n));
  double s = len > 0 ? 1. / len : 0.;
  if (Dot(n, ref) + d > 0) s = -s;
  o[0] = (float)(n.x * s);
  o[1] = (float)(n.y * s);
  o[2] = (float)(n.z * s);
  o[3] = (float)(d * s);
  *dmax = std::max(*dmax, std::fabs(d * s));
}

}  // namespace

void BuildBound32(int type, const double *blk, const Xf &xf, const double lo[3], const double hi[3], B32 &b,
                  std::vector<float> &side) {
  std::memset(&b, 0, sizeof(b));
  for (int i = 0; i < 9; ++i) b.r[i] = (float)xf.r[i];
  for (int i = 0; i < 3; ++i) b.t[i] = (float)xf.t[i];
  b.tsum = (float)((std::fabs(xf.t[0]) + std::fabs(xf.t[1]) + std::fabs(xf.t[2])) * (1. + 1e-6));
  double ext = 0;
  for (int i = 0; i < 3; ++i) ext = std::max(ext, std::max(std::fabs(lo[i]), std::fabs(hi[i])));
  double extra = 0;  // parameters larger than the extent
  b.kind = type;
  b.off = 0;
  float *s = b.s;
  switch (type) {
    case VG_BOX: {
      const BoxP &q = *reinterpret_cast<const BoxP *>(blk);
      for (int i = 0; i < 3; ++i) s[i] = (float)q.d[i];
      break;
    }
    case VG_ORB: {
      s[0] = (float)reinterpret_cast<const OrbP *>(blk)->r;
      break;
    }
    case VG_TUBE: {
      const TubeP &q = *reinterpret_cast<const TubeP *>(blk);
      s[0] = (float)q.rmin;
      s[1] = (float)q.rmax;
      s[2] = (float)q.z;
      s[3] = q.hasRmin != 0 ? 1.f : 0.f;
      PutWedge(s + 4, q.hasPhi != 0, q.sphi, q.dphi);
      break;
    }
    case VG_CONE: {
      const ConeP &q = *reinterpret_cast<const ConeP *>(blk);
      PutCone(s, q, 0.);
      s[7] = (float)q.dz;
      PutWedge(s + 8, q.phiTreatment != 0, q.sphi, q.dphi);
      extra = std::max(std::fabs(q.innerOffset), std::fabs(q.outerOffset));
      break;
    }
    case VG_SPHERE: {
      const SphereP &q = *reinterpret_cast<const SphereP *>(blk);
      s[0] = (float)q.fRmin;
      s[1] = (float)q.fRmax;
      const double t1 = q.fSTheta, t2 = q.fSTheta + q.fDTheta;
      const bool theta = q.fullTheta == 0;
      s[2] = theta && t1 > 0 ? 1.f : 0.f;
      s[3] = (float)std::sin(t1);
      s[4] = (float)std::cos(t1);
      s[5] = theta && t2 < kPi ? 1.f : 0.f;
      s[6] = (float)std::sin(t2);
      s[7] = (float)std::cos(t2);
      PutWedge(s + 8, q.fullPhi == 0, q.fSPhi, q.fDPhi);
      break;
    }
    case VG_TORUS: {
      const TorusP &q = *reinterpret_cast<const TorusP *>(blk);
      s[0] = (float)q.rmin;
      s[1] = (float)q.rmax;
      s[2] = (float)q.rtor;
      PutWedge(s + 3, q.dphi < kTwoPi, q.sphi, q.dphi);
      break;
    }
    case VG_TRD: {
      const TrdP &q = *reinterpret_cast<const TrdP *>(blk);
      const double sx = (q.dx2 - q.dx1) / (2. * q.dz), sy = (q.dy2 - q.dy1) / (2. * q.dz);
      s[0] = (float)(0.5 * (q.dx1 + q.dx2));
      s[1] = (float)sx;
      s[2] = (float)(1. / std::sqrt(1. + sx * sx));
      s[3] = (float)(0.5 * (q.dy1 + q.dy2));
      s[4] = (float)sy;
      s[5] = (float)(1. / std::sqrt(1. + sy * sy));
      s[6] = (float)q.dz;
      break;
    }
    case VG_TRAPEZOID: {
      const TrapP &q = *reinterpret_cast<const TrapP *>(blk);
      s[0] = (float)q.dz;
      for (int i = 0; i < 4; ++i) {
        const double len = std::sqrt(q.A[i] * q.A[i] + q.B[i] * q.B[i] + q.C[i] * q.C[i]);
        const double k = len > 0 ? 1. / len : 0.;
        s[1 + 4 * i] = (float)(q.A[i] * k);
        s[2 + 4 * i] = (float)(q.B[i] * k);
        s[3 + 4 * i] = (float)(q.C[i] * k);
        s[4 + 4 * i] = (float)(q.D[i] * k);
        extra = std::max(extra, std::fabs(q.D[i] * k));
      }
      break;
    }
    case VG_PARALLELEPIPED: {
      const ParaP &q = *reinterpret_cast<const ParaP *>(blk);
      for (int i = 0; i < 3; ++i) s[i] = (float)q.d[i];
      s[3] = (float)q.tanAlpha;
      s[4] = (float)q.tanThetaSinPhi;
      s[5] = (float)q.tanThetaCosPhi;
      // y' = y - tts z: normal (0, 1, -tts); x' = x - tα y + (tα tts - ttc) z.
      const double ny = std::sqrt(1. + q.tanThetaSinPhi * q.tanThetaSinPhi);
      const double c = q.tanAlpha * q.tanThetaSinPhi - q.tanThetaCosPhi;
      const double nx = std::sqrt(1. + q.tanAlpha * q.tanAlpha + c * c);
      s[6] = (float)(1. / ny);
      s[7] = (float)(1. / nx);
      extra = ext * (1. + std::fabs(q.tanAlpha) + std::fabs(q.tanThetaSinPhi) + std::fabs(q.tanThetaCosPhi));
      break;
    }
    case VG_PARABOLOID: {
      const ParabP &q = *reinterpret_cast<const ParabP *>(blk);
      s[0] = (float)q.a;
      s[1] = (float)q.b;
      s[2] = (float)q.dz;
      extra = std::fabs(q.b) + q.a * q.rhi * q.rhi;
      break;
    }
    case VG_POLYCONE: {
      const PolyconeP q{blk};
      const int n = PolyconeS::NSec(q);
      b.off = (int32_t)side.size();
      side.push_back((float)n);
      float w[5];
      const ConeP &c0 = PolyconeS::Sec(q, 0);
      PutWedge(w, c0.phiTreatment != 0, c0.sphi, c0.dphi);
      side.insert(side.end(), w, w + 5);
      for (int k = 0; k < n; ++k) {
        const ConeP &c = PolyconeS::Sec(q, k);
        const double sh = PolyconeS::Shift(q, k);
        float r[9];
        r[0] = (float)(sh - c.dz);
        r[1] = (float)(sh + c.dz);
        PutCone(r + 2, c, sh);
        side.insert(side.end(), r, r + 9);
        extra = std::max(extra, std::max(std::fabs((double)r[3]), std::fabs((double)r[6])));
      }
      break;
    }
    case VG_POLYHEDRON: {
      const PolyhedronP q{blk};
      if (PolyhedronS::HasPhi(q)) {
        // Contains of the reference ignores the phi cutout; no bound.
        b.kind = kBoundNone;
        break;
      }
      const int n = PolyhedronS::NSeg(q);
      b.off = (int32_t)side.size();
      side.push_back((float)n);
      const size_t offs = side.size();
      side.resize(offs + n, 0.f);  // record offsets, relative to b.off
      double dmax = 0;
      for (int k = 0; k < n; ++k) {
        side[offs + k] = (float)(side.size() - b.off);
        const auto sg = PolyhedronS::GetSeg(q, k);
        const double zlo = PolyhedronS::Z(q, k), zhi = PolyhedronS::Z(q, k + 1);
        const V3 ref{0., 0., 0.5 * (zlo + zhi)};
        const int ni = sg.hasInner ? sg.inner.n : 0;
        side.push_back((float)zlo);
        side.push_back((float)zhi);
        side.push_back((float)sg.outer.n);
        side.push_back((float)ni);
        float pl[4];
        for (int i = 0; i < sg.outer.n; ++i) {
          PutPlane(pl, sg.outer.N(i), sg.outer.D(i), ref, &dmax);
          side.insert(side.end(), pl, pl + 4);
        }
        for (int i = 0; i < ni; ++i) {
          PutPlane(pl, sg.inner.N(i), sg.inner.D(i), ref, &dmax);
          side.insert(side.end(), pl, pl + 4);
        }
      }
      extra = dmax;
      break;
    }
    default:
      b.kind = kBoundNone;
      break;
  }
  b.pscale = (float)((2. * ext + extra) * (1. + 1e-6));
}

}  // namespace vga
