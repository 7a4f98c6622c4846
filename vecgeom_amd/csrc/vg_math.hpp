// Device-side scalar math shared by every shape kernel.
//
// All shape code in this directory is compiled with -ffp-contract=off and
// without fast-math: classification and distances must agree bit for bit
// with VecGeom's Scalar backend, which evaluates every expression in IEEE
// double with no fused multiply-add.  The helpers below fix operation order
// and NaN behaviour where the reference relies on it:
//   Min/Max        backend/scalar/Backend.h:145-165 (std::min / std::max)
//   tolerances     base/Global.h:262-309, volumes/kernel/GenericKernels.h:23-53
#pragma once

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define VGA_HD __host__ __device__ __forceinline__
#else
#include <cmath>
#define VGA_HD inline
#endif

#include <cstdint>

namespace vga {

constexpr double kPi = 3.14159265358979323846;
constexpr double kTwoPi = 2. * kPi;
constexpr double kInfinity = __builtin_huge_val();
constexpr double kMinimum = 2.2250738585072014e-308;  // DBL_MIN
constexpr double kTolerance = 1e-9;
constexpr double kHalfTolerance = 0.5 * kTolerance;
constexpr int kInside = 0;
constexpr int kSurface = 1;
constexpr int kOutside = 2;

VGA_HD double Min(double a, double b) { return (b < a) ? b : a; }
VGA_HD double Max(double a, double b) { return (a < b) ? b : a; }
VGA_HD double Abs(double a) { return __builtin_fabs(a); }
VGA_HD double Sqrt(double a) { return __builtin_sqrt(a); }

template <bool tol> VGA_HD double PlusTol(double x) { return tol ? x + kHalfTolerance : x; }
template <bool tol> VGA_HD double MinusTol(double x) { return tol ? x - kHalfTolerance : x; }
template <bool tol> VGA_HD double PlusTolSq(double x, double xsq) { return tol ? xsq + kTolerance * x : xsq; }
template <bool tol> VGA_HD double MinusTolSq(double x, double xsq) { return tol ? xsq - kTolerance * x : xsq; }

struct V3 {
  double x, y, z;
  VGA_HD double operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
  VGA_HD double Mag2() const { return x * x + y * y + z * z; }
};
VGA_HD V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
VGA_HD V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
VGA_HD V3 operator*(double s, V3 a) { return {s * a.x, s * a.y, s * a.z}; }
VGA_HD V3 operator*(V3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
VGA_HD double Dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }

// Placement transform.  Kinds select the reference's specialised
// Transform<transCode, rotCode> paths (base/Transformation3D.h:537-562):
// identity, translation only, rotation only, general.  Rotations always use
// the 9-term general form of DoRotation (base/Transformation3D.h:469-478).
enum XfKind : int { kXfIdentity = 0, kXfTranslation = 1, kXfRotation = 2, kXfGeneral = 3 };

struct Xf {
  double t[3];
  double r[9];
};

VGA_HD V3 Rotate(const Xf &X, V3 m) {
  double lx = m.x * X.r[0];
  double ly = m.x * X.r[1];
  double lz = m.x * X.r[2];
  lx += m.y * X.r[3];
  ly += m.y * X.r[4];
  lz += m.y * X.r[5];
  lx += m.z * X.r[6];
  ly += m.z * X.r[7];
  lz += m.z * X.r[8];
  return {lx, ly, lz};
}

template <int K>
VGA_HD V3 XfPoint(const Xf &X, V3 m) {
  if (K == kXfIdentity) return m;
  if (K == kXfTranslation) return {m.x - X.t[0], m.y - X.t[1], m.z - X.t[2]};
  if (K == kXfRotation) return Rotate(X, m);
  return Rotate(X, V3{m.x - X.t[0], m.y - X.t[1], m.z - X.t[2]});
}

template <int K>
VGA_HD V3 XfDir(const Xf &X, V3 m) {
  if (K == kXfIdentity || K == kXfTranslation) return m;
  return Rotate(X, m);
}

// Runtime-kind variants used by the navigator (kind stored per placement).
VGA_HD V3 XfPointRt(const Xf &X, int k, V3 m) {
  switch (k) {
    case kXfIdentity: return m;
    case kXfTranslation: return {m.x - X.t[0], m.y - X.t[1], m.z - X.t[2]};
    case kXfRotation: return Rotate(X, m);
    default: return Rotate(X, V3{m.x - X.t[0], m.y - X.t[1], m.z - X.t[2]});
  }
}
VGA_HD V3 XfDirRt(const Xf &X, int k, V3 m) { return (k == kXfIdentity || k == kXfTranslation) ? m : Rotate(X, m); }

}  // namespace vga
