// ORACLE -- test infrastructure only (see vgo.h).
//
// Boolean volumes (union, intersection, subtraction of two placed primitive
// shapes), restating volumes/kernel/BooleanImplementation.h (subtraction),
// BooleanUnionImplementation.h, BooleanIntersectionImplementation.h and
// source/UnplacedBooleanVolume.cpp (Extent).  The operands are placed volumes
// whose placements are relative to the Boolean's own frame; their virtuals
// are used exactly as the reference's kernels call them (VPlacedVolume
// Contains/Inside/DistanceToIn/SafetyToIn transform the point, DistanceToOut
// and SafetyToOut take the operand's local point, PlacedDistanceToOut
// transforms both).
//
// Parameter block (shared with the device code and the reference shim):
//   op (0 union, 1 intersection, 2 subtraction: the BooleanOperation enum of
//   volumes/UnplacedBooleanVolume.h:10-14),
//   then for the left and the right operand: type, nparams, params[nparams],
//   placement[12] (translation, then the 9 rotation entries).
// Operands must be primitive shapes (not Boolean).
//
// The reference's loops (subtraction DistanceToIn, union DistanceToOut,
// intersection DistanceToIn) have no iteration limit; here each stops after
// kMaxBooleanIterations passes and returns the distance reached.
#include "vgo.h"

namespace vgo {

Shape *MakeShape(int type, const double *params, int nparams);

namespace {

constexpr int kMaxBooleanIterations = 1000;
enum { kUnion = 0, kIntersection = 1, kSubtraction = 2 };
enum { kInsideIn = 0, kSurfaceIn = 1, kOutsideIn = 2 };  // EInside

struct Boolean : Shape {
  int op = 0;
  Placed A, B;  // left, right
  Shape *sa = nullptr, *sb = nullptr;
  ~Boolean() override {
    delete sa;
    delete sb;
  }

  // volumes/kernel/BooleanImplementation.h:296-313 (subtraction),
  // BooleanUnionImplementation.h ContainsKernel, BooleanIntersectionImplementation.h ContainsKernel.
  bool Contains(V3 p) const override {
    if (op == kUnion) {
      if (A.Contains(p)) return true;
      return B.Contains(p);
    }
    if (op == kIntersection) {
      const bool a = A.Contains(p);
      const bool b = B.Contains(p);
      return a && b;
    }
    if (!A.Contains(p)) return false;
    return !B.Contains(p);
  }

  // The InsideKernels (Geant4's algorithm, surface states as written).
  int Inside(V3 p) const override {
    const int a = A.Inside(p);
    if (op == kUnion) {
      if (a == kInsideIn) return kInsideIn;
      const int b = B.Inside(p);
      if (b == kInsideIn) return kInsideIn;
      if (b == kSurfaceIn || a == kSurfaceIn) return kSurfaceIn;
      return kOutsideIn;
    }
    if (a == kOutsideIn) return kOutsideIn;
    const int b = B.Inside(p);
    if (op == kIntersection) {
      if (a == kInsideIn && b == kInsideIn) return kInsideIn;
      if ((a == kInsideIn && b == kSurfaceIn) || (b == kInsideIn && a == kSurfaceIn) ||
          (a == kSurfaceIn && b == kSurfaceIn))
        return kSurfaceIn;
      return kOutsideIn;
    }
    if (a == kInsideIn && b == kOutsideIn) return kInsideIn;
    if ((a == kInsideIn && b == kSurfaceIn) || (b == kOutsideIn && a == kSurfaceIn)) return kSurfaceIn;
    return kOutsideIn;
  }

  double DistanceToIn(V3 p, V3 v, double stepMax) const override {
    if (op == kUnion) {
      // BooleanUnionImplementation.h DistanceToInKernel: the nearer operand.
      const double d1 = A.DistanceToIn(p, v, stepMax);
      const double d2 = B.DistanceToIn(p, v, stepMax);
      return Min(d1, d2);
    }
    if (op == kIntersection) return IntersectionDistanceToIn(p, v, stepMax);
    return SubtractionDistanceToIn(p, v, stepMax);
  }

  // BooleanImplementation.h:456-500: leave B, then enter A, alternating.
  double SubtractionDistanceToIn(V3 p, V3 v, double stepMax) const {
    double d1, d2, snxt = 0.;
    V3 hitpoint = p;
    bool insideRight = B.Contains(p);
    double epsil = 0.;
    for (int it = 0; it < kMaxBooleanIterations; ++it) {
      if (insideRight) {
        d1 = B.PlacedDistanceToOut(hitpoint, v, stepMax);
        snxt += (d1 > 0. && d1 < kInfinity) ? d1 + epsil : 0.;
        hitpoint = hitpoint + ((d1 > 0. && d1 < kInfinity) ? (d1 + 1E-8) * v : 1E-8 * v);
        epsil = 1.E-8;
        if (A.Contains(hitpoint)) return snxt;
      }
      d2 = A.DistanceToIn(hitpoint, v, stepMax);
      if (d2 == kInfinity) return kInfinity;
      d1 = B.DistanceToIn(hitpoint, v, stepMax);
      if (d2 < d1 - kTolerance) {
        snxt += d2 + epsil;
        return snxt;
      }
      snxt += (d1 > 0. && d1 < kInfinity) ? d1 + epsil : 0.;
      hitpoint = hitpoint + ((d1 > 0. && d1 < kInfinity) ? (d1 + epsil) * v : epsil * v);
      insideRight = true;
    }
    return snxt;
  }

  // BooleanIntersectionImplementation.h DistanceToInKernel.
  double IntersectionDistanceToIn(V3 p, V3 v, double stepMax) const {
    V3 hitpoint = p;
    bool inleft = A.Contains(hitpoint);
    bool inright = B.Contains(hitpoint);
    double d1 = 0., d2 = 0., snext = 0.;
    if (inleft && inright) {
      d1 = A.PlacedDistanceToOut(hitpoint, v, stepMax);
      d2 = B.PlacedDistanceToOut(hitpoint, v, stepMax);
      if (d1 < 2 * kTolerance) inleft = false;
      if (d2 < 2 * kTolerance) inright = false;
      if (inleft && inright) return 0.0;
    }
    for (int it = 0; it < kMaxBooleanIterations; ++it) {
      d1 = d2 = 0;
      if (!inleft) {
        d1 = A.DistanceToIn(hitpoint, v, kInfinity);
        d1 = Max(d1, kTolerance);
        if (d1 > 1E20) return kInfinity;
      }
      if (!inright) {
        d2 = B.DistanceToIn(hitpoint, v, kInfinity);
        d2 = Max(d2, kTolerance);
        if (d2 > 1E20) return kInfinity;
      }
      if (d1 > d2) {
        snext += d1;
        inleft = true;
        hitpoint = hitpoint + d1 * v;
        inright = B.Contains(hitpoint + kTolerance * v);
        if (inright) return snext;
      } else {
        snext += d2;
        inright = true;
        hitpoint = hitpoint + d2 * v;
        inleft = A.Contains(hitpoint + kTolerance * v);
        if (inleft) return snext;
      }
    }
    return snext;
  }

  double DistanceToOut(V3 p, V3 v, double stepMax) const override {
    if (op == kIntersection) {
      // BooleanIntersectionImplementation.h DistanceToOutKernel (default step limits).
      return Min(A.DistanceToOut(p, v, kInfinity), B.PlacedDistanceToOut(p, v, kInfinity));
    }
    if (op == kSubtraction) {
      // BooleanImplementation.h:510-520.
      const double d = A.DistanceToOut(p, v, stepMax);
      const double dinright = B.DistanceToIn(p, v, stepMax);
      return Min(d, dinright);
    }
    // BooleanUnionImplementation.h DistanceToOutKernel: step through the
    // operand the point is in, then through the other while it overlaps.
    double dist = 0., disTmp = 0.;
    int push = 0;
    if (A.Contains(p)) {
      for (int it = 0; it < kMaxBooleanIterations; ++it) {
        disTmp = A.DistanceToOut(p + dist * v, v, kInfinity);
        dist += (disTmp >= 0. && disTmp < kInfinity) ? disTmp : 0;
        dist += kTolerance;
        push++;
        if (B.Contains(p + dist * v)) {
          disTmp = B.PlacedDistanceToOut(p + dist * v, v, kInfinity);
          dist += (disTmp >= 0. && disTmp < kInfinity) ? disTmp : 0;
        }
        if (!A.Contains(p + dist * v)) break;
      }
    } else {
      for (int it = 0; it < kMaxBooleanIterations; ++it) {
        disTmp = B.PlacedDistanceToOut(p + dist * v, v, kInfinity);
        dist += (disTmp >= 0. && disTmp < kInfinity) ? disTmp : 0.;
        dist += kTolerance;
        push++;
        if (A.Contains(p + dist * v)) {
          disTmp = A.DistanceToOut(p + dist * v, v, kInfinity);
          dist += disTmp;
        }
        if (!B.Contains(p + dist * v)) break;
      }
    }
    return dist - push * kTolerance;
  }

  double SafetyToIn(V3 p) const override {
    if (op == kUnion) {
      // BooleanUnionImplementation.h SafetyToInKernel.
      const double distA = A.SafetyToIn(p);
      const double distB = B.SafetyToIn(p);
      double safety = Min(distA, distB);
      if (safety < 0) safety = 0.;
      return safety;
    }
    if (op == kIntersection) {
      // BooleanIntersectionImplementation.h SafetyToInKernel.
      const bool insideA = A.Contains(p);
      const bool insideB = B.Contains(p);
      if (!insideA && insideB) return A.SafetyToIn(p);
      if (!insideB && insideA) return B.SafetyToIn(p);
      return Min(A.SafetyToIn(p), B.SafetyToIn(p));
    }
    // BooleanImplementation.h:529-546.
    if (A.Contains(p) && B.Contains(p)) return B.SafetyToOut(B.xf.Apply(p));
    return A.SafetyToIn(p);
  }

  double SafetyToOut(V3 p) const override {
    if (op == kUnion) {
      // BooleanUnionImplementation.h SafetyToOutKernel.
      const bool containedA = A.Contains(p);
      const bool containedB = B.Contains(p);
      if (containedA && containedB) return Max(A.SafetyToOut(p), B.SafetyToOut(B.xf.Apply(p)));
      if (containedB) return B.SafetyToOut(B.xf.Apply(p));
      return A.SafetyToOut(p);
    }
    if (op == kIntersection) {
      // BooleanIntersectionImplementation.h SafetyToOutKernel.
      double safety = Min(A.SafetyToOut(p), B.SafetyToOut(B.xf.Apply(p)));
      if (safety < 0) safety = 0.;
      return safety;
    }
    // BooleanImplementation.h:560-567.
    const double safety = A.SafetyToOut(p);
    const double safetyright = B.SafetyToIn(p);
    return Min(safety, safetyright);
  }

  // source/UnplacedBooleanVolume.cpp:94-122: the operands' unplaced extents
  // (their placements are not applied), combined per operation.
  void Extent(V3 &lo, V3 &hi) const override {
    V3 la, ha, lb, hb;
    sa->Extent(la, ha);
    sb->Extent(lb, hb);
    if (op == kUnion) {
      lo = V3{Min(la.x, lb.x), Min(la.y, lb.y), Min(la.z, lb.z)};
      hi = V3{Max(ha.x, hb.x), Max(ha.y, hb.y), Max(ha.z, hb.z)};
    } else if (op == kIntersection) {
      lo = V3{Max(la.x, lb.x), Max(la.y, lb.y), Max(la.z, lb.z)};
      hi = V3{Min(ha.x, hb.x), Min(ha.y, hb.y), Min(ha.z, hb.z)};
    } else {
      lo = la;
      hi = ha;
    }
  }
};

// Parses one operand; returns the number of doubles read, or -1.
int ReadOperand(const double *p, int n, Placed &pl, Shape *&s) {
  if (n < 2) return -1;
  const int type = (int)p[0], np = (int)p[1];
  if (type == kBoolean || np < 0 || n < 2 + np + 12) return -1;
  s = MakeShape(type, p + 2, np);
  if (!s) return -1;
  pl.shape = s;
  pl.xf = Transform::FromArray(p + 2 + np);
  return 2 + np + 12;
}

}  // namespace

Shape *MakeBoolean(const double *p, int n) {
  if (n < 1) return nullptr;
  const int op = (int)p[0];
  if (op < kUnion || op > kSubtraction) return nullptr;
  auto *b = new Boolean;
  b->op = op;
  const int ra = ReadOperand(p + 1, n - 1, b->A, b->sa);
  const int rb = ra < 0 ? -1 : ReadOperand(p + 1 + ra, n - 1 - ra, b->B, b->sb);
  if (ra < 0 || rb < 0 || 1 + ra + rb != n) {
    delete b;
    return nullptr;
  }
  return b;
}

}  // namespace vgo
