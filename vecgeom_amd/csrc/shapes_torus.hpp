// Torus (ring of circular cross-section with optional inner radius and phi
// section): parameter block and device kernels with the semantics of
// volumes/kernel/TorusImplementation.h:36-1170 and the cached values of
// volumes/UnplacedTorus.h:43-69.
//
// Distances come from the reference's complex-arithmetic Ferrari solution of
// the quartic (solveQuartic2) followed by two Newton steps.  The solver calls
// cbrt, atan2, sin and cos; the device math library may differ from the host
// libm in the last ulp, so distances agree with the host to a relative
// 1e-12-ish level after the Newton steps rather than bit for bit, and a root
// whose imaginary part sits exactly at the 1e-9 acceptance threshold may be
// classified differently.  Contains, Inside and the safeties use no
// transcendental and agree exactly.
#pragma once
#include "shapes_box_tube.hpp"
#include "vg_math.hpp"

namespace vga {

struct TorusP {
  double rmin, rmax, rtor, sphi, dphi, rmin2, rmax2, rtor2;
  WedgeP w;
  TubeP btube;  // bounding tube, phi never treated (HollowTube)
};
static_assert(sizeof(TorusP) % 16 == 0, "TorusP keeps 16-byte multiples");

struct TorusS {
  using P = TorusP;

  struct Cx {
    double re, im;
  };
  static VGA_HD Cx Add(Cx x, Cx y) { return {x.re + y.re, x.im + y.im}; }
  static VGA_HD Cx Add(double s, Cx x) { return {x.re + s, x.im}; }
  static VGA_HD Cx Sub(Cx x, Cx y) { return {x.re - y.re, x.im - y.im}; }
  static VGA_HD Cx Sub(double s, Cx x) { return {s - x.re, -x.im}; }
  static VGA_HD Cx Neg(Cx x) { return {-x.re, -x.im}; }
  static VGA_HD Cx Mul(Cx x, Cx y) { return {x.re * y.re - x.im * y.im, x.re * y.im + x.im * y.re}; }
  static VGA_HD Cx Mul(Cx x, double s) { return {x.re * s, x.im * s}; }
  static VGA_HD Cx Conj(Cx x) { return {x.re, -x.im}; }
  static VGA_HD Cx DivReal(Cx x, double y) {
    const double inv = 1. / y;
    return {x.re * inv, x.im * inv};
  }
  static VGA_HD Cx Div(double s, Cx y) {
    const Cx t = Mul(y, Conj(y));
    return DivReal(Mul(Conj(y), s), t.re);
  }
  static VGA_HD Cx CSqrt(Cx x) {
    const double l = Sqrt(x.re * x.re + x.im * x.im);
    return {Sqrt(0.5 * (l + x.re)), Sqrt(0.5 * (l - x.re))};
  }
  static VGA_HD Cx CSqrtReal(double x) {
    const double imcoef = (x >= 0.) ? 0. : 1.;
    const double realcoef = (x >= 0) ? 1. : 0.;
    const double l = Sqrt(Abs(x));
    return {realcoef * l, imcoef * l};
  }
  static VGA_HD double ATan2(double y, double x) {
    if (x != 0) return atan2(y, x);
    if (y > 0) return kPi / 2;
    if (y < 0) return -kPi / 2;
    return 0;
  }
  static VGA_HD Cx CCbrt(Cx x) {
    double r, s, c;
    if (x.im != 0.0) {
      r = Sqrt(x.re * x.re + x.im * x.im);
      const double na = ATan2(x.im, x.re) / 3.;
      s = sin(na);
      c = cos(na);
    } else {
      r = x.re;
      s = 0.;
      c = 1.;
    }
    const double rc = cbrt(r);
    return {rc * c, rc * s};
  }

  // solveQuartic2 for the monic x^4 + b x^3 + c x^2 + d x + e.
  static VGA_HD void SolveQuartic(double b, double c, double d, double e, Cx *roots) {
    constexpr double inv256 = 1. / 256, inv16 = 1. / 16, inv3 = 1. / 3, inv27 = 1. / 27, inv6 = 1. / 6,
                     inv108 = 1. / 108, inv12 = 1. / 12;
    const double bb = b * b;
    const double bbb = bb * b;
    const double alpha = -3.0 * 0.125 * bb + c;
    const double beta = 0.125 * bbb - 0.5 * b * c + d;
    const double alpha2 = alpha * alpha;
    const double gamma = -3.0 * bbb * b * inv256 + c * bb * inv16 - 0.25 * b * d + e;
    const double P = -alpha2 * inv12 - gamma;
    const double Q = -alpha2 * alpha * inv108 + alpha * gamma * inv3 - 0.125 * beta * beta;
    const double tmp = 0.25 * Q * Q + P * P * P * inv27;
    const Cx R = Add(Q * 0.5, CSqrtReal(tmp));
    const Cx U = CCbrt(R);
    Cx y = Sub(-5.0 * alpha * inv6, U);
    y = Add(y, Div(P, Mul(U, 3.)));
    const Cx W = CSqrt(Add(Add(alpha, y), y));
    const double firstPart = -0.25 * b;
    const Cx second = Sub(-3.0 * alpha, Mul(y, 2.0));
    const Cx third = Div(2.0 * beta, W);
    roots[0] = Add(firstPart, Mul(Sub(Neg(W), CSqrt(Add(second, third))), 0.5));
    roots[1] = Add(firstPart, Mul(Add(Neg(W), CSqrt(Add(second, third))), 0.5));
    roots[2] = Add(firstPart, Mul(Sub(W, CSqrt(Sub(second, third))), 0.5));
    roots[3] = Add(firstPart, Mul(Add(W, CSqrt(Sub(second, third))), 0.5));
  }

  template <bool ForInside, bool Phi>
  static VGA_HD void Kernel(const P &t, V3 p, bool &cin, bool &cout) {
    const double rxy = Sqrt(p.x * p.x + p.y * p.y);
    const double radsq = (rxy - t.rtor) * (rxy - t.rtor) + p.z * p.z;
    cout = radsq > PlusTolSq<ForInside>(t.rmax, t.rmax2);
    if (ForInside) cin = radsq < MinusTolSq<ForInside>(t.rmax, t.rmax2);
    if (cout) return;
    cout |= radsq < MinusTolSq<ForInside>(t.rmin, t.rmin2);
    if (ForInside) cin &= radsq > PlusTolSq<ForInside>(t.rmin, t.rmin2);
    if (cout) return;
    if (t.dphi < kTwoPi && Phi) {
      bool wi, wo;
      WedgeS::Kernel<ForInside>(t.w, p, wi, wo);
      cout |= wo;
      if (ForInside) cin &= wi;
    }
  }
  static VGA_HD bool Contains(const P &t, V3 p) {
    bool cin = false, cout;
    Kernel<false, true>(t, p, cin, cout);
    return !cout;
  }
  static VGA_HD bool ContainsDisk(const P &t, V3 p) {
    bool cin = false, cout;
    Kernel<false, false>(t, p, cin, cout);
    return !cout;
  }
  static VGA_HD int Inside(const P &t, V3 p) {
    bool cin = false, cout;
    Kernel<true, true>(t, p, cin, cout);
    int r = kSurface;
    if (cout) r = kOutside;
    if (cin) r = kInside;
    return r;
  }

  static VGA_HD double CheckZero(double b, double c, double d, double e, double x) {
    const double x2 = x * x;
    return x2 * x2 + b * x2 * x + c * x2 + d * x + e;
  }
  static VGA_HD double NewtonIter(double b, double c, double d, double x, double fold) {
    const double x2 = x * x;
    const double fprime = 4 * x2 * x + 3 * b * x2 + 2 * c * x + d;
    return x - fold / fprime;
  }

  template <bool ForRmin>
  static VGA_HD double ToBoundary(const P &t, V3 p, V3 dir, double radius, bool out) {
    const double r0sq = p.x * p.x + p.y * p.y + p.z * p.z;
    const double rdotn = p.x * dir.x + p.y * dir.y + p.z * dir.z;
    const double rsumsq = t.rtor2 + radius * radius;
    const double a = 4. * rdotn;
    const double b = 2. * (r0sq + 2. * rdotn * rdotn - rsumsq + 2. * t.rtor2 * dir.z * dir.z);
    const double c = 4. * (r0sq * rdotn - rsumsq * rdotn + 2. * t.rtor2 * p.z * dir.z);
    const double d = r0sq * r0sq - 2. * r0sq * rsumsq + 4. * t.rtor2 * p.z * p.z +
                     (t.rtor2 - radius * radius) * (t.rtor2 - radius * radius);
    const bool needphi = t.dphi < kTwoPi;
    Cx roots[4];
    SolveQuartic(a, b, c, d, roots);
    double valid = kInfinity, second = kInfinity;
    bool have = false;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      have = Abs(roots[k].im) < 1E-9 && roots[k].re > -1E-9;
      if (needphi && have) {
        const double s = roots[k].re;
        have = WedgeS::ContainsWithBoundary(t.w, V3{p.x + s * dir.x, p.y + s * dir.y, p.z + s * dir.z});
      }
      if (k == 0) {
        if (have) valid = roots[0].re;
        if (out && have) second = valid;
      } else {
        if (have) valid = Min(roots[k].re, valid);
        if (out && have) second = Max(roots[k].re, valid);
      }
    }
    // Newton refinement only when the last root examined was valid.
    have = have && (valid < kInfinity);
    if (have) {
      valid = NewtonIter(a, b, c, valid, CheckZero(a, b, c, d, valid));
      valid = NewtonIter(a, b, c, valid, CheckZero(a, b, c, d, valid));
    }
    if (valid < 1E-9) {
      valid = 0.;
      const double rho = Sqrt(p.x * p.x + p.y * p.y);
      const double red = (rho - t.rtor) / rho;
      V3 n{p.x * red, p.y * red, p.z};
      if (ForRmin) n = V3{-n.x, -n.y, -n.z};
      const double dot = Dot(dir, n);
      if (!out && (dot > 0.)) valid = second;
      if (out && (dot < 0.)) valid = second;
    }
    return valid;
  }

  static VGA_HD double PhiLimit(const P &t, V3 p, V3 v, double d1, double d2, double current) {
    if (d1 < current) {
      const V3 ip{p.x + v.x * d1, p.y + v.y * d1, p.z + v.z * d1};
      if (ContainsDisk(t, ip) && (ip.x * t.w.a1x + ip.y * t.w.a1y >= 0)) current = d1;
    }
    if (d2 < current) {
      const V3 ip{p.x + v.x * d2, p.y + v.y * d2, p.z + v.z * d2};
      if (ContainsDisk(t, ip) && (ip.x * t.w.a2x + ip.y * t.w.a2y >= 0)) current = d2;
    }
    return current;
  }

  // The bounding-tube rejection at the head of DistanceToIn: true (and
  // out = kInfinity) exactly when DistanceToIn returns kInfinity there.
  static constexpr bool kHasQuick = true;
  static VGA_HD bool DistanceToInQuick(const P &t, V3 p, V3 v, double stepMax, double &out) {
    if (TubeS::Contains(t.btube, p) || TubeS::DistanceToIn(t.btube, p, v, stepMax) != kInfinity) return false;
    out = kInfinity;
    return true;
  }

  static VGA_HD double DistanceToIn(const P &t, V3 p, V3 v, double stepMax) {
    const bool inBounds = TubeS::Contains(t.btube, p);
    const double tubeDistance = TubeS::DistanceToIn(t.btube, p, v, stepMax);
    if (!inBounds && tubeDistance == kInfinity) return kInfinity;
    V3 lp = p;
    if (!inBounds) lp = V3{p.x + v.x * tubeDistance, p.y + v.y * tubeDistance, p.z + v.z * tubeDistance};
    const double dout = ToBoundary<false>(t, lp, v, t.rmax, false);
    double din = kInfinity;
    if (t.rmin > 0) din = ToBoundary<true>(t, lp, v, t.rmin, false);
    double minDist = Min(dout, din);
    if (t.dphi < kTwoPi) {
      double d1, d2;
      WedgeS::DistanceToIn(t.w, lp, v, d1, d2);
      minDist = PhiLimit(t, lp, v, d1, d2, minDist);
    }
    return inBounds ? minDist : minDist + tubeDistance;
  }

  static VGA_HD double DistanceToOut(const P &t, V3 p, V3 v, double) {
    const double dout = ToBoundary<false>(t, p, v, t.rmax, true);
    double din = kInfinity;
    if (t.rmin > 0) din = ToBoundary<true>(t, p, v, t.rmin, true);
    double distance = Min(dout, din);
    if (t.dphi < kTwoPi) {
      double d1, d2;
      WedgeS::DistanceToOut(t.w, p, v, d1, d2);
      distance = PhiLimit(t, p, v, d1, d2, distance);
    }
    return distance;
  }

  static VGA_HD double SafetyToIn(const P &t, V3 p) {
    const double rxy = Sqrt(p.x * p.x + p.y * p.y);
    const double rad = Sqrt((rxy - t.rtor) * (rxy - t.rtor) + p.z * p.z);
    double safety = rad - t.rmax;
    if (t.rmin != 0) safety = Max(t.rmin - rad, rad - t.rmax);
    if (t.dphi < kTwoPi && (rxy != 0.)) safety = Max(WedgeS::SafetyToIn(t.w, p), safety);
    return safety;
  }
  static VGA_HD double SafetyToOut(const P &t, V3 p) {
    const double rxy = Sqrt(p.x * p.x + p.y * p.y);
    const double rad = Sqrt((rxy - t.rtor) * (rxy - t.rtor) + p.z * p.z);
    double safety = t.rmax - rad;
    if (t.rmin != 0) safety = Min(rad - t.rmin, t.rmax - rad);
    if (t.dphi < kTwoPi) safety = Min(WedgeS::SafetyToOut(t.w, p), safety);
    return safety;
  }
};

}  // namespace vga
