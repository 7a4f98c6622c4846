// ORACLE -- test infrastructure only.
//
// CPU restatement of VecGeom's Scalar backend for the shape-query and
// navigation hot path.  It exists to check the HIP product against the
// reference algorithm on identical inputs, and to time the reference-style
// CPU path beside the GPU.  Only tests/, __graft_entry__.smoke() and the
// cpu_baseline leg of bench.py may load it.  Nothing in vecgeom_amd/ links
// against it.
//
// Semantics followed (file:line relative to the reference tree):
//   constants            base/Global.h:193-309
//   Min/Max/Abs/Sqrt     backend/scalar/Backend.h:60-170 (std::min/std::max)
//   Transformation3D     base/Transformation3D.h:469-478 (general DoRotation),
//                        :518-525 (DoTranslation), :537-562 (Transform<t,r>),
//                        :601-632 (InverseTransform), :674-743
//                        (MultiplyFromRight), :754-767 (TransformDirection),
//                        source/Transformation3D.cpp:81-161 (SetProperties,
//                        GenerateRotationCode)
//   early returns        kScalar::early_returns = true
//                        (backend/scalar/Backend.h:20)
//   scalar output init   volumes/ShapeImplementationHelper.h:135-318
//                        (Inside -> kOutside, Contains -> false, distances and
//                        safeties -> kInfinity before the kernel runs)
//
// The oracle is compiled with -ffp-contract=off and without -ffast-math (the
// strict build, oracle of record).  A second copy compiled with the
// reference's Release flags (-O3 -ffast-math -ftree-vectorize) is built only
// to count lanes where the two disagree.
#pragma once
#include <cmath>
#include <cstdint>
#include <limits>
#include <algorithm>
#include <vector>

namespace vgo {

// base/Global.h:262-309
constexpr double kPi = 3.14159265358979323846;
constexpr double kTwoPi = 2. * kPi;
constexpr double kDegToRad = kPi / 180.;
constexpr double kInfinity = std::numeric_limits<double>::infinity();
constexpr double kMinimum = std::numeric_limits<double>::min();
constexpr double kTolerance = 1e-9;
constexpr double kHalfTolerance = 0.5 * kTolerance;
constexpr double kAngTolerance = 1e-9;
constexpr double kRadTolerance = 1e-9;
// base/Global.h:198-205
constexpr int kInside = 0;
constexpr int kSurface = 1;
constexpr int kOutside = 2;

// backend/scalar/Backend.h:145-165: std::min / std::max argument order.
inline double Min(double a, double b) { return (b < a) ? b : a; }
inline double Max(double a, double b) { return (a < b) ? b : a; }
inline double Abs(double a) { return std::fabs(a); }
inline double Sqrt(double a) { return std::sqrt(a); }

// volumes/kernel/GenericKernels.h:23-53
template <bool tol> inline double MakePlusTolerant(double x) { return tol ? x + kHalfTolerance : x; }
template <bool tol> inline double MakeMinusTolerant(double x) { return tol ? x - kHalfTolerance : x; }
template <bool tol> inline double MakePlusTolerantSquare(double x, double xsq) { return tol ? xsq + kTolerance * x : xsq; }
template <bool tol> inline double MakeMinusTolerantSquare(double x, double xsq) { return tol ? xsq - kTolerance * x : xsq; }

struct V3 {
  double x, y, z;
  double operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
  double &operator[](int i) { return i == 0 ? x : (i == 1 ? y : z); }
  double Mag2() const { return x * x + y * y + z * z; }  // base/Vector3D.h Mag2
  double Mag() const { return std::sqrt(Mag2()); }
  double Perp2() const { return x * x + y * y; }
};
inline V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 operator*(double s, V3 a) { return {s * a.x, s * a.y, s * a.z}; }
inline V3 operator*(V3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
// base/Vector3D.h Dot: x*x' + y*y' + z*z' evaluated left to right.
inline double Dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }

// ATan2 follows backend/scalar/Backend.h:130-137.
inline double ATan2(double y, double x) {
  if (x != 0) return std::atan2(y, x);
  if (y > 0) return kPi / 2;
  if (y < 0) return -kPi / 2;
  return 0;
}

// Transformation3D restated with the reference's operation order.
struct Transform {
  double t[3] = {0, 0, 0};
  double r[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
  bool hasTranslation = false;
  bool hasRotation = false;
  bool identity = true;

  // source/Transformation3D.cpp:81-90 (SetProperties) and :139-150.
  void SetProperties() {
    hasTranslation = std::fabs(t[0]) > kTolerance || std::fabs(t[1]) > kTolerance || std::fabs(t[2]) > kTolerance;
    hasRotation = RotationCode() != 0x200;
    identity = !hasTranslation && !hasRotation;
  }
  int RotationCode() const {
    int code = 0;
    for (int i = 0; i < 9; ++i) code |= (1 << i) * (std::fabs(r[i]) > kTolerance);
    if (code == 0x111 && (r[0] == 1. && r[4] == 1. && r[8] == 1.)) code = 0x200;
    return code;
  }
  static Transform FromArray(const double *a12) {
    Transform T;
    for (int i = 0; i < 3; ++i) T.t[i] = a12[i];
    for (int i = 0; i < 9; ++i) T.r[i] = a12[3 + i];
    T.SetProperties();
    return T;
  }
  // base/Transformation3D.h:469-478 general rotation (local = R^T-indexed).
  V3 Rotate(V3 m) const {
    V3 l;
    l.x = m.x * r[0];
    l.y = m.x * r[1];
    l.z = m.x * r[2];
    l.x += m.y * r[3];
    l.y += m.y * r[4];
    l.z += m.y * r[5];
    l.x += m.z * r[6];
    l.y += m.z * r[7];
    l.z += m.z * r[8];
    return l;
  }
  // base/Transformation3D.h:537-562 with the placement's codes: translation is
  // skipped when SetProperties found none, rotation when the code is identity.
  V3 Apply(V3 m) const {
    if (!hasTranslation && !hasRotation) return m;
    if (hasTranslation && !hasRotation) return {m.x - t[0], m.y - t[1], m.z - t[2]};
    if (!hasTranslation && hasRotation) return Rotate(m);
    V3 tmp{m.x - t[0], m.y - t[1], m.z - t[2]};
    return Rotate(tmp);
  }
  // base/Transformation3D.h:754-767
  V3 ApplyDirection(V3 m) const {
    if (!hasRotation) return m;
    return Rotate(m);
  }
  // base/Transformation3D.h:601-632 (InverseTransformKernel<false>)
  V3 Inverse(V3 l) const {
    V3 m;
    m.x = t[0];
    m.x += l.x * r[0];
    m.x += l.y * r[1];
    m.x += l.z * r[2];
    m.y = t[1];
    m.y += l.x * r[3];
    m.y += l.y * r[4];
    m.y += l.z * r[5];
    m.z = t[2];
    m.z += l.x * r[6];
    m.z += l.y * r[7];
    m.z += l.z * r[8];
    return m;
  }
  // InverseTransformKernel<true>
  V3 InverseDirection(V3 l) const {
    V3 m;
    m.x = l.x * r[0];
    m.x += l.y * r[1];
    m.x += l.z * r[2];
    m.y = l.x * r[3];
    m.y += l.y * r[4];
    m.y += l.z * r[5];
    m.z = l.x * r[6];
    m.z += l.y * r[7];
    m.z += l.z * r[8];
    return m;
  }
  // base/Transformation3D.h:674-743
  void MultiplyFromRight(const Transform &rhs) {
    if (rhs.identity) return;
    if (rhs.hasTranslation) {
      t[0] += r[0] * rhs.t[0];
      t[0] += r[1] * rhs.t[1];
      t[0] += r[2] * rhs.t[2];
      t[1] += r[3] * rhs.t[0];
      t[1] += r[4] * rhs.t[1];
      t[1] += r[5] * rhs.t[2];
      t[2] += r[6] * rhs.t[0];
      t[2] += r[7] * rhs.t[1];
      t[2] += r[8] * rhs.t[2];
    }
    if (rhs.hasRotation) {
      for (int row = 0; row < 3; ++row) {
        double tx = r[3 * row], ty = r[3 * row + 1], tz = r[3 * row + 2];
        r[3 * row + 0] = tx * rhs.r[0];
        r[3 * row + 1] = tx * rhs.r[1];
        r[3 * row + 2] = tx * rhs.r[2];
        r[3 * row + 0] += ty * rhs.r[3];
        r[3 * row + 1] += ty * rhs.r[4];
        r[3 * row + 2] += ty * rhs.r[5];
        r[3 * row + 0] += tz * rhs.r[6];
        r[3 * row + 1] += tz * rhs.r[7];
        r[3 * row + 2] += tz * rhs.r[8];
      }
    }
    // The reference leaves fIdentity/fHasRotation/fHasTranslation as they
    // were on the accumulating matrix (a fresh Transformation3D starts as the
    // identity whose flags say identity); TopMatrix callers then use the
    // generic Transform, so flags of the product do not matter there.
  }
  // Generic (unspecialised) Transform used by the navigators on TopMatrix.
  V3 ApplyGeneric(V3 m) const {
    V3 tmp{m.x - t[0], m.y - t[1], m.z - t[2]};
    return Rotate(tmp);
  }
  V3 ApplyDirectionGeneric(V3 m) const { return Rotate(m); }
};

// Shape interface in the local frame, one method per reference kernel entry.
struct Shape {
  virtual ~Shape() {}
  virtual bool Contains(V3 p) const = 0;                              // UnplacedContains
  virtual int Inside(V3 p) const = 0;                                 // Inside kernel on local point
  virtual double DistanceToIn(V3 p, V3 d, double stepMax) const = 0;  // local point/dir
  virtual double DistanceToOut(V3 p, V3 d, double stepMax) const = 0;
  virtual double SafetyToIn(V3 p) const = 0;
  virtual double SafetyToOut(V3 p) const = 0;
  // Axis-aligned extent in the local frame (UnplacedX::Extent).
  virtual void Extent(V3 &lo, V3 &hi) const = 0;
  // Shapes whose placed DistanceToIn does not use TransformDirection (the
  // reference's Torus) override this; default is the correct form.
  virtual bool DirectionUsesPointTransform() const { return false; }
  // VPlacedVolume::Normal(localPoint, normal&) (volumes/PlacedVolume.h:266-272
  // and the shapes' overrides).  n starts as a zero vector (a default
  // Vector3D) and is set as the reference sets it; returns the reference's
  // valid flag (1 or 0), or -1 where the reference has no Normal for the
  // shape (the base class asserts).
  virtual int Normal(V3 p, V3 &n) const {
    (void)p;
    n = V3{0, 0, 0};
    return -1;
  }
};

// The USolids DistanceToOut(point, direction, normal&, convex&, stepMax)
// (volumes/ShapeImplementationHelper.h:277-295, identical in
// ScalarShapeImplementationHelper.h:249-264): DistanceToOut, then the Normal
// of point + d * direction, convex always true.  Returns the distance; valid
// as Normal returns it.
inline double DistanceToOutNormal(const Shape &s, V3 p, V3 d, double stepMax, V3 &n, bool &convex, int &valid) {
  const double dist = s.DistanceToOut(p, d, stepMax);
  // Vector3D operator*(scalar, vector) scales a copy of the vector; then
  // point + that, component by component.
  const V3 hit{p.x + d.x * dist, p.y + d.y * dist, p.z + d.z * dist};
  valid = s.Normal(hit, n);
  convex = true;
  return dist;
}

// Shape type tags shared with the product C-ABI (include/vecgeom_amd.h).
enum ShapeType {
  kBox = 1,
  kTube = 2,
  kOrb = 3,
  kSphere = 4,
  kCone = 5,
  kTrd = 6,
  kTrapezoid = 7,
  kParallelepiped = 8,
  kParaboloid = 9,
  kTorus = 10,
  kPolycone = 11,
  kPolyhedron = 12,
  kBoolean = 13,  // union / intersection / subtraction of two placed primitives (vgo_boolean.cpp)
};

// Factory: returns nullptr for an unknown type or a bad parameter count.
Shape *MakeShape(int type, const double *params, int nparams);

// A placed shape: placement transform + unplaced shape.
// volumes/ShapeImplementationHelper.h:135-318 (scalar virtuals).
struct Placed {
  const Shape *shape = nullptr;
  Transform xf;
  bool Contains(V3 p) const { return shape->Contains(xf.Apply(p)); }
  int Inside(V3 p) const { return shape->Inside(xf.Apply(p)); }
  double DistanceToIn(V3 p, V3 d, double stepMax) const {
    V3 ld = shape->DirectionUsesPointTransform() ? xf.Apply(d) : xf.ApplyDirection(d);
    return shape->DistanceToIn(xf.Apply(p), ld, stepMax);
  }
  double DistanceToOut(V3 p, V3 d, double stepMax) const { return shape->DistanceToOut(p, d, stepMax); }
  double PlacedDistanceToOut(V3 p, V3 d, double stepMax) const {
    return shape->DistanceToOut(xf.Apply(p), xf.ApplyDirection(d), stepMax);
  }
  double SafetyToIn(V3 p) const { return shape->SafetyToIn(xf.Apply(p)); }
  double SafetyToOut(V3 p) const { return shape->SafetyToOut(p); }
};

}  // namespace vgo
