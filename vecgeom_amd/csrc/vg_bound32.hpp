// FP32 distance bounds: the single-precision filter of the mixed-precision
// forms (shape queries with float inputs, and the navigator's daughter
// walk).
//
// For a placed volume and a point p in its mother frame, Bound32 computes in
// FP32 two numbers:
//   so  a lower bound on the Euclidean distance from p to the shape when p
//       is outside it (so > 0 proves p is outside);
//   si  a lower bound on the distance from p to the shape's boundary when p
//       is inside it (si > 0 proves p is inside);
// and a margin M that bounds, rigorously, how far the FP32 values can lie
// from the same functions evaluated exactly on the exact local point.  A
// lane with so > M is outside the shape by more than the rounding of the
// whole FP32 computation; with si > M it is inside by as much.  The margin
// also holds the reference's tolerance band (every classification of the
// reference differs from the geometric one only within a few kTolerance of
// a surface, far below the 1e-6 mm added to M), so such a lane's Contains
// and Inside equal the FP64 kernels' results; every other lane is evaluated
// by the FP64 kernel.
//
// The bounds are distance functions of simple sets, combined so that they
// stay lower bounds:
//   * for an intersection A n B:  dist(p, A n B) >= max(dist(p, A), dist(p, B))
//     and depth(p, A n B) >= min(depth(p, A), depth(p, B));
//   * for a union A u B:          dist(p, A u B) = min(dist(p, A), dist(p, B))
//     and depth(p, A u B) >= max(depth(p, A), depth(p, B));
//   * half-spaces n.p + d <= 0 with a unit normal: signed distance n.p + d
//     (exact); convex polytopes (Box, Trd, Trapezoid, Parallelepiped,
//     Polyhedron prisms): max / min over their face planes;
//   * z slabs, cylinders rho <= R and spheres r <= R: |z| - dz, rho - R,
//     r - R (exact);
//   * cones rho <= a z + b (Cone, Polycone sections): the distance to the
//     generating line in the meridian half-plane, (rho - a z - b)/sqrt(1+a^2)
//     (exact for the solid cone, a lower bound beyond its apex);
//   * theta cones of a Sphere (half-lines from the origin at angle theta):
//     r sin(theta_p - theta) = rho cos(theta) - z sin(theta) and its mirror;
//   * the solid torus: sqrt((rho - R)^2 + z^2) - rmax (exact);
//   * the paraboloid z >= a rho^2 + b: outside, the distance to its tangent
//     plane at the point's own radius, (a rho^2 + b - z)/sqrt(1 + 4a^2 rho^2);
//     inside, depth >= min(1, f/sqrt(1 + 4a^2 (rho+1)^2)) with f = z - a rho^2 - b
//     (f is Lipschitz with that constant within 1 mm of p);
//   * phi wedges (half-planes through the z axis with inward unit normals
//     n1, n2): the intersection of the two half-planes when the opening
//     is at most pi, their union when it is larger.
// Every one of these functions is 1-Lipschitz in p (the paraboloid's outside
// bound up to a factor that tends to 1 near its surface).
//
// Error bound.  The local point is l = R (p - t) evaluated in FP32 from the
// float point, the float rounding of t and of R: each component is within
// 8u S of the exact value (u = 2^-24, S = |p|_1 + |t|_1; one rounding of p,
// of t, of each subtraction and of the three-term dot product with rounded
// R, whose entries are at most 1).  A 1-Lipschitz function then moves by at
// most sqrt(3) * 8u S, and its own FP32 evaluation (a few additions,
// multiplications and at most two square roots on values of magnitude at
// most |l|_1 + P, P the largest parameter of the shape) adds at most
// 16u (|l|_1 + P) <= 16u (sqrt(3) S + P).  M = 64u (S + P) + 1e-6 mm covers
// the sum with a factor of about 1.5 to spare.  tests/test_bound32.py checks
// the bounds against the oracle on points in the tolerance band.
#pragma once
#include "../../include/vecgeom_amd.h"
#include "vg_math.hpp"

namespace vga {

constexpr float kU32 = 5.9604644775390625e-8f;  // 2^-24
constexpr float kBoundMul = 64.f;
constexpr float kBoundAbs = 1e-6f;

// Bound kinds (the shape type of the placed volume, or none).
constexpr int kBoundNone = -1;

// Phi wedge modes.
constexpr float kPhiNone = 0.f, kPhiConvex = 1.f, kPhiUnion = 2.f;

// FP32 record of one placed volume (144 bytes).
struct alignas(16) B32 {
  float r[9];    // local = Rotate(p - t), l_j = sum_i (p - t)_i r[3i + j]
  float t[3];
  float tsum;    // |t|_1 (rounded up)
  float pscale;  // largest parameter magnitude of the shape (mm)
  int32_t kind;  // shape type with a bound, or kBoundNone
  int32_t off;   // offset of the variable-length data in the side array
  float s[20];   // shape data (layout per kind in vg_bound32.cpp)
};
static_assert(sizeof(B32) == 144, "B32 is 144 bytes");

struct F3 {
  float x, y, z;
};

VGA_HD float FAbs(float a) { return __builtin_fabsf(a); }
VGA_HD float FSqrt(float a) { return __builtin_sqrtf(a); }
VGA_HD float FMax(float a, float b) { return a < b ? b : a; }
VGA_HD float FMin(float a, float b) { return b < a ? b : a; }

VGA_HD F3 Local32(const B32 &b, F3 p) {
  const float ax = p.x - b.t[0], ay = p.y - b.t[1], az = p.z - b.t[2];
  return F3{ax * b.r[0] + ay * b.r[3] + az * b.r[6], ax * b.r[1] + ay * b.r[4] + az * b.r[7],
            ax * b.r[2] + ay * b.r[5] + az * b.r[8]};
}
VGA_HD F3 LocalDir32(const B32 &b, F3 d) {
  return F3{d.x * b.r[0] + d.y * b.r[3] + d.z * b.r[6], d.x * b.r[1] + d.y * b.r[4] + d.z * b.r[7],
            d.x * b.r[2] + d.y * b.r[5] + d.z * b.r[8]};
}

// The margin M for a mother-frame point p.
VGA_HD float Margin32(const B32 &b, F3 p) {
  const float S = FAbs(p.x) + FAbs(p.y) + FAbs(p.z) + b.tsum;
  return kBoundMul * kU32 * (S + b.pscale) + kBoundAbs;
}

// Wedge (mode, n1x, n1y, n2x, n2y at w[0..4]): (distance outside, depth inside).
VGA_HD void Wedge32(const float *w, F3 l, float &so, float &si) {
  if (w[0] == kPhiNone) return;
  const float d1 = w[1] * l.x + w[2] * l.y, d2 = w[3] * l.x + w[4] * l.y;
  if (w[0] == kPhiConvex) {
    so = FMax(so, FMax(-d1, -d2));
    si = FMin(si, FMin(d1, d2));
  } else {
    so = FMax(so, FMin(-d1, -d2));
    si = FMin(si, FMax(d1, d2));
  }
}

// Cone surfaces: inner rho >= aI z + bI (if hasRmin), outer rho <= aO z + bO,
// with cI, cO = 1/sqrt(1 + a^2).  c = {aI, bI, cI, aO, bO, cO, hasRmin}.
VGA_HD void Cone32(const float *c, float rho, float z, float &so, float &si) {
  const float out = (rho - (c[3] * z + c[4])) * c[5];
  so = FMax(so, out);
  si = FMin(si, -out);
  if (c[6] != 0.f) {
    const float in = ((c[0] * z + c[1]) - rho) * c[2];
    so = FMax(so, in);
    si = FMin(si, -in);
  }
}

// Bounds of a shape of kind K at local point l.  side: variable-length data.
template <int K>
VGA_HD void BoundOf(const B32 &b, const float *side, F3 l, float &so, float &si) {
  const float *s = b.s;
  so = -__builtin_huge_valf();
  si = __builtin_huge_valf();
  switch (K) {
    case VG_BOX: {
      const float ox = FAbs(l.x) - s[0], oy = FAbs(l.y) - s[1], oz = FAbs(l.z) - s[2];
      so = FMax(ox, FMax(oy, oz));
      si = -so;
      break;
    }
    case VG_ORB: {
      const float r = FSqrt(l.x * l.x + l.y * l.y + l.z * l.z);
      so = r - s[0];
      si = -so;
      break;
    }
    case VG_TUBE: {  // rmin, rmax, dz, hasRmin, wedge[5]
      const float rho = FSqrt(l.x * l.x + l.y * l.y);
      const float oz = FAbs(l.z) - s[2], orr = rho - s[1];
      so = FMax(oz, orr);
      si = FMin(-oz, -orr);
      if (s[3] != 0.f) {
        so = FMax(so, s[0] - rho);
        si = FMin(si, rho - s[0]);
      }
      Wedge32(s + 4, l, so, si);
      break;
    }
    case VG_CONE: {  // cone[7], dz, wedge[5]
      const float rho = FSqrt(l.x * l.x + l.y * l.y);
      const float oz = FAbs(l.z) - s[7];
      so = oz;
      si = -oz;
      Cone32(s, rho, l.z, so, si);
      Wedge32(s + 8, l, so, si);
      break;
    }
    case VG_SPHERE: {  // rmin, rmax, hasTheta1, sin1, cos1, hasTheta2, sin2, cos2, wedge[5]
      const float rho = FSqrt(l.x * l.x + l.y * l.y);
      const float r = FSqrt(l.x * l.x + l.y * l.y + l.z * l.z);
      so = r - s[1];
      si = -so;
      if (s[0] > 0.f) {
        so = FMax(so, s[0] - r);
        si = FMin(si, r - s[0]);
      }
      if (s[2] != 0.f) {  // theta >= theta1
        const float o = l.z * s[3] - rho * s[4];
        so = FMax(so, o);
        si = FMin(si, -o);
      }
      if (s[5] != 0.f) {  // theta <= theta2
        const float o = rho * s[7] - l.z * s[6];
        so = FMax(so, o);
        si = FMin(si, -o);
      }
      Wedge32(s + 8, l, so, si);
      break;
    }
    case VG_TORUS: {  // rmin, rmax, rtor, wedge[5]
      const float rho = FSqrt(l.x * l.x + l.y * l.y);
      const float a = rho - s[2];
      const float dt = FSqrt(a * a + l.z * l.z);
      so = dt - s[1];
      si = -so;
      if (s[0] > 0.f) {
        so = FMax(so, s[0] - dt);
        si = FMin(si, dt - s[0]);
      }
      Wedge32(s + 3, l, so, si);
      break;
    }
    case VG_TRD: {  // hx0, sx, cx, hy0, sy, cy, dz
      const float oz = FAbs(l.z) - s[6];
      const float ox = (FAbs(l.x) - (s[0] + s[1] * l.z)) * s[2];
      const float oy = (FAbs(l.y) - (s[3] + s[4] * l.z)) * s[5];
      so = FMax(oz, FMax(ox, oy));
      si = -so;
      break;
    }
    case VG_TRAPEZOID: {  // dz, then 4 planes (a, b, c, d) with unit outward normals
      so = FAbs(l.z) - s[0];
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const float *q = s + 1 + 4 * i;
        so = FMax(so, q[0] * l.x + q[1] * l.y + q[2] * l.z + q[3]);
      }
      si = -so;
      break;
    }
    case VG_PARALLELEPIPED: {  // dx, dy, dz, tanAlpha, tanThetaSinPhi, tanThetaCosPhi, 1/|ny|, 1/|nx|
      const float yp = l.y - s[4] * l.z;
      const float xp = l.x - s[5] * l.z - s[3] * yp;
      const float oz = FAbs(l.z) - s[2];
      const float oy = (FAbs(yp) - s[1]) * s[6];
      const float ox = (FAbs(xp) - s[0]) * s[7];
      so = FMax(oz, FMax(oy, ox));
      si = -so;
      break;
    }
    case VG_PARABOLOID: {  // a, b, dz
      const float rho2 = l.x * l.x + l.y * l.y;
      const float f = (s[0] * rho2 + s[1]) - l.z;  // > 0 outside the paraboloid
      const float oz = FAbs(l.z) - s[2];
      const float g2 = 4.f * s[0] * s[0];
      const float out = f / FSqrt(1.f + g2 * rho2);
      const float rp = FSqrt(rho2) + 1.f;
      const float in = FMin(1.f, -f / FSqrt(1.f + g2 * rp * rp));
      so = FMax(oz, out);
      si = FMin(-oz, in);
      break;
    }
    case VG_POLYCONE: {  // side: nsec, wedge[5], then per section zlo, zhi, cone[7]
      // Union over the sections.  Sections are ordered along z, and a
      // section's bounds are at least its slab distance max(zlo - z, z - zhi)
      // away (so) / at most minus it (si): starting from the section nearest
      // in z, sections are visited outwards while their slab distance can
      // still lower so; only sections whose slab holds z can raise si.
      const float *q = side + b.off;
      const int n = (int)q[0];
      const float rho = FSqrt(l.x * l.x + l.y * l.y);
      auto sec = [&](int k, float &o, float &in) {
        const float *c = q + 6 + 9 * k;
        o = FMax(c[0] - l.z, l.z - c[1]);
        in = -o;
        Cone32(c + 2, rho, l.z, o, in);
      };
      int k0 = 0;
      while (k0 + 1 < n && l.z > q[6 + 9 * k0 + 1]) ++k0;
      sec(k0, so, si);
      for (int k = k0 - 1; k >= 0; --k) {  // below: slab distance z - zhi
        if (l.z - q[6 + 9 * k + 1] >= so) break;
        float o, in;
        sec(k, o, in);
        so = FMin(so, o);
        si = FMax(si, in);
      }
      for (int k = k0 + 1; k < n; ++k) {  // above: slab distance zlo - z
        if (q[6 + 9 * k] - l.z >= so) break;
        float o, in;
        sec(k, o, in);
        so = FMin(so, o);
        si = FMax(si, in);
      }
      Wedge32(q + 1, l, so, si);
      break;
    }
    case VG_POLYHEDRON: {  // side: nseg, then offsets of the segment records, then the records
      // Segment record: zlo, zhi, nout, nin, nout + nin planes.  Visited as
      // the Polycone's sections are.
      const float *q = side + b.off;
      const int n = (int)q[0];
      auto seg = [&](int k, float &o, float &in) {
        const float *g = q + (int)q[1 + k];
        const int no = (int)g[2], ni = (int)g[3];
        o = FMax(g[0] - l.z, l.z - g[1]);
        const float *pl = g + 4;
        for (int i = 0; i < no; ++i, pl += 4) o = FMax(o, pl[0] * l.x + pl[1] * l.y + pl[2] * l.z + pl[3]);
        in = -o;
        if (ni > 0) {
          // Inner planes are oriented with the hole on their negative side:
          // outside the hole when some plane is positive (depth of the
          // material >= that value); inside the hole at depth min(-plane).
          float hmax = -__builtin_huge_valf();
          for (int i = 0; i < ni; ++i, pl += 4) hmax = FMax(hmax, pl[0] * l.x + pl[1] * l.y + pl[2] * l.z + pl[3]);
          o = FMax(o, -hmax);
          in = FMin(in, hmax);
        }
      };
      auto zlo = [&](int k) { return q[(int)q[1 + k]]; };
      auto zhi = [&](int k) { return q[(int)q[1 + k] + 1]; };
      int k0 = 0;
      while (k0 + 1 < n && l.z > zhi(k0)) ++k0;
      seg(k0, so, si);
      for (int k = k0 - 1; k >= 0; --k) {
        if (l.z - zhi(k) >= so) break;
        float o, in;
        seg(k, o, in);
        so = FMin(so, o);
        si = FMax(si, in);
      }
      for (int k = k0 + 1; k < n; ++k) {
        if (zlo(k) - l.z >= so) break;
        float o, in;
        seg(k, o, in);
        so = FMin(so, o);
        si = FMax(si, in);
      }
      break;
    }
    default:
      break;
  }
}

// Shapes present in a mask of kinds (bit k: kind k may occur): a kernel that
// knows its shapes compiles only their bounds.
constexpr uint32_t kAllBoundKinds = 0xFFFFFFFFu;

#define VGA_BOUND_CASE(K)                                 \
  case K:                                                 \
    if constexpr ((MASK >> K) & 1u) BoundOf<K>(b, side, l, so, si); \
    break;

template <uint32_t MASK = kAllBoundKinds>
VGA_HD void ShapeBound32(const B32 &b, const float *side, F3 l, float &so, float &si) {
  so = -__builtin_huge_valf();
  si = __builtin_huge_valf();
  switch (b.kind) {
    VGA_BOUND_CASE(VG_BOX)
    VGA_BOUND_CASE(VG_ORB)
    VGA_BOUND_CASE(VG_TUBE)
    VGA_BOUND_CASE(VG_CONE)
    VGA_BOUND_CASE(VG_SPHERE)
    VGA_BOUND_CASE(VG_TORUS)
    VGA_BOUND_CASE(VG_TRD)
    VGA_BOUND_CASE(VG_TRAPEZOID)
    VGA_BOUND_CASE(VG_PARALLELEPIPED)
    VGA_BOUND_CASE(VG_PARABOLOID)
    VGA_BOUND_CASE(VG_POLYCONE)
    VGA_BOUND_CASE(VG_POLYHEDRON)
    default:
      break;
  }
}
#undef VGA_BOUND_CASE

// so, si and the margin of a mother-frame point.  For kBoundNone every
// point is undecided (so = si = -inf).
template <uint32_t MASK = kAllBoundKinds>
VGA_HD void Bound32(const B32 &b, const float *side, F3 p, float &so, float &si, float &m) {
  m = Margin32(b, p);
  if (b.kind == kBoundNone) {
    so = si = -__builtin_huge_valf();
    return;
  }
  ShapeBound32<MASK>(b, side, Local32(b, p), so, si);
}

// Decision of the FP32 filter: 0 inside, 2 outside, -1 undecided.
template <uint32_t MASK = kAllBoundKinds>
VGA_HD int Classify32(const B32 &b, const float *side, F3 p) {
  float so, si, m;
  Bound32<MASK>(b, side, p, so, si, m);
  if (so > m) return kOutside;
  if (si > m) return kInside;
  return -1;
}

// Sign of p.v in the local frame (p the point, v the direction): +1 or -1
// when |p.v| exceeds the margin M (its FP32 error is below
// |dl| |v| + |l| |dv| + 3u |l| <= M), 0 otherwise.  The reference Orb's
// DistanceToIn of a missing ray depends on it (NaN when the ray approaches
// the centre: the square root of a negative discriminant; kInfinity when it
// recedes).
VGA_HD int RadialSign32(const B32 &b, F3 p, F3 d) {
  const float m = Margin32(b, p);
  const F3 l = Local32(b, p);
  const F3 v = LocalDir32(b, d);
  const float pv = l.x * v.x + l.y * v.y + l.z * v.z;
  return pv > m ? 1 : (pv < -m ? -1 : 0);
}

// True when the ray p + t d (t >= 0; d a unit vector to float rounding)
// provably misses the box [lo - pad, hi + pad] of the local frame, with
// pad = 2M + 1e-3.  Axes along which the direction component is below 1e-2
// in magnitude impose no constraint (which only weakens the test); for the
// others the slab times carry relative errors below 1e-3 and an absolute
// error below M / |d_i|, covered by the padding and the 1e-3 relative
// slack.  The extent box is e[0..5] = lo xyz, hi xyz.
VGA_HD bool RayMissesBox32(const B32 &b, const float *e, F3 p, F3 d) {
  const float m = Margin32(b, p);
  const F3 l = Local32(b, p);
  const F3 v = LocalDir32(b, d);
  const float pad = 2.f * m + 1e-3f;
  const float lp[3] = {l.x, l.y, l.z}, vv[3] = {v.x, v.y, v.z};
  float t0 = 0.f, t1 = __builtin_huge_valf();
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    if (FAbs(vv[i]) < 1e-2f) continue;
    const float inv = 1.f / vv[i];
    float ta = (e[i] - pad - lp[i]) * inv, tb = (e[3 + i] + pad - lp[i]) * inv;
    if (ta > tb) {
      const float x = ta;
      ta = tb;
      tb = x;
    }
    t0 = FMax(t0, ta);
    t1 = FMin(t1, tb);
  }
  return t0 > t1 * (1.f + 1e-3f) + 1e-3f;
}

}  // namespace vga
