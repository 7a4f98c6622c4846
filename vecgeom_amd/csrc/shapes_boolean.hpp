// Boolean volumes: union, intersection and subtraction of two placed
// primitive shapes, with the semantics of volumes/kernel/BooleanImplementation.h
// (subtraction), BooleanUnionImplementation.h and
// BooleanIntersectionImplementation.h.  The operands are placed in the
// Boolean's own frame; each is evaluated through its shape's kernels the way
// the reference's kernels call the operand's VPlacedVolume virtuals: Contains,
// Inside, DistanceToIn and SafetyToIn transform the point (and direction),
// DistanceToOut and SafetyToOut take the operand's local point,
// PlacedDistanceToOut transforms both.
//
// Device parameter block (doubles, built by BuildParams in vg_params.cpp):
//   [0] op (0 union, 1 intersection, 2 subtraction), [1..3] left operand:
//   shape type, offset of its parameter block, placement kind; [4..6] the
//   same for the right operand; [7] the block's total length; [8..19] left
//   placement, [20..31] right placement (translation, then rotation); then
//   the two operands' own parameter blocks at their offsets (16-byte
//   aligned).  Operands are primitive shapes.
//
// The reference's loops (subtraction DistanceToIn, union DistanceToOut,
// intersection DistanceToIn) have no iteration limit; each stops here after
// kMaxBooleanIterations passes and returns the distance reached, as the
// oracle does.
#pragma once
#include "../../include/vecgeom_amd.h"
#include "shapes.hpp"

namespace vga {

struct BooleanP {
  const double *b;
};

struct BooleanS {
  using P = BooleanP;
  static constexpr int kMaxBooleanIterations = 1000;
  enum { kUnionOp = 0, kIntersectionOp = 1, kSubtractionOp = 2 };

  struct Operand {
    int type, kind;
    const double *blk;
    const Xf *xf;
  };
  static VGA_HD int Op(const P &q) { return (int)q.b[0]; }
  static VGA_HD Operand Get(const P &q, int k) {
    const double *h = q.b + 1 + 3 * k;
    return Operand{(int)h[0], (int)h[2], q.b + (int)h[1], reinterpret_cast<const Xf *>(q.b + 8 + 12 * k)};
  }

  // Calls f(S{}, params) for the operand's shape type.
  template <class F>
  static VGA_HD auto With(int type, const double *blk, F f) -> decltype(f(BoxS{}, ParamView<BoxS>::Get(blk))) {
    switch (type) {
      case VG_TUBE: return f(TubeS{}, ParamView<TubeS>::Get(blk));
      case VG_ORB: return f(OrbS{}, ParamView<OrbS>::Get(blk));
      case VG_SPHERE: return f(SphereS{}, ParamView<SphereS>::Get(blk));
      case VG_CONE: return f(ConeS{}, ParamView<ConeS>::Get(blk));
      case VG_TRD: return f(TrdS{}, ParamView<TrdS>::Get(blk));
      case VG_TRAPEZOID: return f(TrapS{}, ParamView<TrapS>::Get(blk));
      case VG_PARALLELEPIPED: return f(ParaS{}, ParamView<ParaS>::Get(blk));
      case VG_PARABOLOID: return f(ParabS{}, ParamView<ParabS>::Get(blk));
      case VG_TORUS: return f(TorusS{}, ParamView<TorusS>::Get(blk));
      case VG_POLYCONE: return f(PolyconeS{}, ParamView<PolyconeS>::Get(blk));
      case VG_POLYHEDRON: return f(PolyhedronS{}, ParamView<PolyhedronS>::Get(blk));
      default: return f(BoxS{}, ParamView<BoxS>::Get(blk));
    }
  }

  // The operand's VPlacedVolume virtuals (volumes/PlacedVolume.h:147-250).
  static VGA_HD V3 Local(const Operand &o, V3 p) { return XfPointRt(*o.xf, o.kind, p); }
  static VGA_HD bool OContains(const Operand &o, V3 p) {
    const V3 l = Local(o, p);
    return With(o.type, o.blk, [&](auto s, const auto &Q) { return decltype(s)::Contains(Q, l); });
  }
  static VGA_HD int OInside(const Operand &o, V3 p) {
    const V3 l = Local(o, p);
    return With(o.type, o.blk, [&](auto s, const auto &Q) { return decltype(s)::Inside(Q, l); });
  }
  static VGA_HD double ODistIn(const Operand &o, V3 p, V3 v, double sm) {
    const V3 l = Local(o, p), ld = XfDirRt(*o.xf, o.kind, v);
    return With(o.type, o.blk, [&](auto s, const auto &Q) { return decltype(s)::DistanceToIn(Q, l, ld, sm); });
  }
  static VGA_HD double ODistOut(const Operand &o, V3 p, V3 v, double sm) {
    return With(o.type, o.blk, [&](auto s, const auto &Q) { return decltype(s)::DistanceToOut(Q, p, v, sm); });
  }
  static VGA_HD double OPlacedDistOut(const Operand &o, V3 p, V3 v, double sm) {
    const V3 l = Local(o, p), ld = XfDirRt(*o.xf, o.kind, v);
    return With(o.type, o.blk, [&](auto s, const auto &Q) { return decltype(s)::DistanceToOut(Q, l, ld, sm); });
  }
  static VGA_HD double OSafeIn(const Operand &o, V3 p) {
    const V3 l = Local(o, p);
    return With(o.type, o.blk, [&](auto s, const auto &Q) { return decltype(s)::SafetyToIn(Q, l); });
  }
  static VGA_HD double OSafeOut(const Operand &o, V3 p) {
    return With(o.type, o.blk, [&](auto s, const auto &Q) { return decltype(s)::SafetyToOut(Q, p); });
  }
  // Transformation3D::Transform (generic) of the operand's placement.
  static VGA_HD V3 Generic(const Operand &o, V3 p) { return XfPointRt(*o.xf, kXfGeneral, p); }

  // ContainsKernel of each operation.
  static VGA_HD bool Contains(const P &q, V3 p) {
    const Operand A = Get(q, 0), B = Get(q, 1);
    const int op = Op(q);
    if (op == kUnionOp) {
      if (OContains(A, p)) return true;
      return OContains(B, p);
    }
    if (op == kIntersectionOp) {
      const bool a = OContains(A, p);
      const bool b = OContains(B, p);
      return a && b;
    }
    if (!OContains(A, p)) return false;
    return !OContains(B, p);
  }

  // InsideKernel of each operation (Geant4's algorithm).
  static VGA_HD int Inside(const P &q, V3 p) {
    const Operand A = Get(q, 0), B = Get(q, 1);
    const int op = Op(q);
    const int a = OInside(A, p);
    if (op == kUnionOp) {
      if (a == kInside) return kInside;
      const int b = OInside(B, p);
      if (b == kInside) return kInside;
      if (b == kSurface || a == kSurface) return kSurface;
      return kOutside;
    }
    if (a == kOutside) return kOutside;
    const int b = OInside(B, p);
    if (op == kIntersectionOp) {
      if (a == kInside && b == kInside) return kInside;
      if ((a == kInside && b == kSurface) || (b == kInside && a == kSurface) || (a == kSurface && b == kSurface))
        return kSurface;
      return kOutside;
    }
    if (a == kInside && b == kOutside) return kInside;
    if ((a == kInside && b == kSurface) || (b == kOutside && a == kSurface)) return kSurface;
    return kOutside;
  }

  static VGA_HD double DistanceToIn(const P &q, V3 p, V3 v, double stepMax) {
    const Operand A = Get(q, 0), B = Get(q, 1);
    const int op = Op(q);
    if (op == kUnionOp) {
      const double d1 = ODistIn(A, p, v, stepMax);
      const double d2 = ODistIn(B, p, v, stepMax);
      return Min(d1, d2);
    }
    if (op == kIntersectionOp) {
      // BooleanIntersectionImplementation.h DistanceToInKernel.
      V3 hit = p;
      bool inleft = OContains(A, hit);
      bool inright = OContains(B, hit);
      double d1 = 0., d2 = 0., snext = 0.;
      if (inleft && inright) {
        d1 = OPlacedDistOut(A, hit, v, stepMax);
        d2 = OPlacedDistOut(B, hit, v, stepMax);
        if (d1 < 2 * kTolerance) inleft = false;
        if (d2 < 2 * kTolerance) inright = false;
        if (inleft && inright) return 0.0;
      }
      for (int it = 0; it < kMaxBooleanIterations; ++it) {
        d1 = d2 = 0;
        if (!inleft) {
          d1 = ODistIn(A, hit, v, kInfinity);
          d1 = Max(d1, kTolerance);
          if (d1 > 1E20) return kInfinity;
        }
        if (!inright) {
          d2 = ODistIn(B, hit, v, kInfinity);
          d2 = Max(d2, kTolerance);
          if (d2 > 1E20) return kInfinity;
        }
        if (d1 > d2) {
          snext += d1;
          inleft = true;
          hit = hit + d1 * v;
          inright = OContains(B, hit + kTolerance * v);
          if (inright) return snext;
        } else {
          snext += d2;
          inright = true;
          hit = hit + d2 * v;
          inleft = OContains(A, hit + kTolerance * v);
          if (inleft) return snext;
        }
      }
      return snext;
    }
    // BooleanImplementation.h:456-500 (subtraction): leave B, enter A.
    double d1, d2, snxt = 0.;
    V3 hit = p;
    bool insideRight = OContains(B, p);
    double epsil = 0.;
    for (int it = 0; it < kMaxBooleanIterations; ++it) {
      if (insideRight) {
        d1 = OPlacedDistOut(B, hit, v, stepMax);
        snxt += (d1 > 0. && d1 < kInfinity) ? d1 + epsil : 0.;
        hit = hit + ((d1 > 0. && d1 < kInfinity) ? (d1 + 1E-8) * v : 1E-8 * v);
        epsil = 1.E-8;
        if (OContains(A, hit)) return snxt;
      }
      d2 = ODistIn(A, hit, v, stepMax);
      if (d2 == kInfinity) return kInfinity;
      d1 = ODistIn(B, hit, v, stepMax);
      if (d2 < d1 - kTolerance) return snxt + (d2 + epsil);
      snxt += (d1 > 0. && d1 < kInfinity) ? d1 + epsil : 0.;
      hit = hit + ((d1 > 0. && d1 < kInfinity) ? (d1 + epsil) * v : epsil * v);
      insideRight = true;
    }
    return snxt;
  }

  static VGA_HD double DistanceToOut(const P &q, V3 p, V3 v, double stepMax) {
    const Operand A = Get(q, 0), B = Get(q, 1);
    const int op = Op(q);
    if (op == kIntersectionOp) return Min(ODistOut(A, p, v, kInfinity), OPlacedDistOut(B, p, v, kInfinity));
    if (op == kSubtractionOp) {
      const double d = ODistOut(A, p, v, stepMax);
      const double dinright = ODistIn(B, p, v, stepMax);
      return Min(d, dinright);
    }
    // BooleanUnionImplementation.h DistanceToOutKernel.
    double dist = 0., disTmp = 0.;
    int push = 0;
    if (OContains(A, p)) {
      for (int it = 0; it < kMaxBooleanIterations; ++it) {
        disTmp = ODistOut(A, p + dist * v, v, kInfinity);
        dist += (disTmp >= 0. && disTmp < kInfinity) ? disTmp : 0;
        dist += kTolerance;
        push++;
        if (OContains(B, p + dist * v)) {
          disTmp = OPlacedDistOut(B, p + dist * v, v, kInfinity);
          dist += (disTmp >= 0. && disTmp < kInfinity) ? disTmp : 0;
        }
        if (!OContains(A, p + dist * v)) break;
      }
    } else {
      for (int it = 0; it < kMaxBooleanIterations; ++it) {
        disTmp = OPlacedDistOut(B, p + dist * v, v, kInfinity);
        dist += (disTmp >= 0. && disTmp < kInfinity) ? disTmp : 0.;
        dist += kTolerance;
        push++;
        if (OContains(A, p + dist * v)) {
          disTmp = ODistOut(A, p + dist * v, v, kInfinity);
          dist += disTmp;
        }
        if (!OContains(B, p + dist * v)) break;
      }
    }
    return dist - push * kTolerance;
  }

  static VGA_HD double SafetyToIn(const P &q, V3 p) {
    const Operand A = Get(q, 0), B = Get(q, 1);
    const int op = Op(q);
    if (op == kUnionOp) {
      const double distA = OSafeIn(A, p);
      const double distB = OSafeIn(B, p);
      double safety = Min(distA, distB);
      if (safety < 0) safety = 0.;
      return safety;
    }
    if (op == kIntersectionOp) {
      const bool insideA = OContains(A, p);
      const bool insideB = OContains(B, p);
      if (!insideA && insideB) return OSafeIn(A, p);
      if (!insideB && insideA) return OSafeIn(B, p);
      return Min(OSafeIn(A, p), OSafeIn(B, p));
    }
    // BooleanImplementation.h:529-546.
    if (OContains(A, p) && OContains(B, p)) return OSafeOut(B, Generic(B, p));
    return OSafeIn(A, p);
  }

  static VGA_HD double SafetyToOut(const P &q, V3 p) {
    const Operand A = Get(q, 0), B = Get(q, 1);
    const int op = Op(q);
    if (op == kUnionOp) {
      const bool containedA = OContains(A, p);
      const bool containedB = OContains(B, p);
      if (containedA && containedB) return Max(OSafeOut(A, p), OSafeOut(B, Generic(B, p)));
      if (containedB) return OSafeOut(B, Generic(B, p));
      return OSafeOut(A, p);
    }
    if (op == kIntersectionOp) {
      double safety = Min(OSafeOut(A, p), OSafeOut(B, Generic(B, p)));
      if (safety < 0) safety = 0.;
      return safety;
    }
    const double safety = OSafeOut(A, p);
    const double safetyright = OSafeIn(B, p);
    return Min(safety, safetyright);
  }
};

template <>
struct ParamView<BooleanS> {
  static constexpr bool kIndirect = true;
  static VGA_HD BooleanP Get(const double *b) { return BooleanP{b}; }
  static VGA_HD int Length(const double *b) { return (int)b[7]; }
};

}  // namespace vga
