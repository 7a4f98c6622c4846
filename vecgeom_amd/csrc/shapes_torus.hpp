// Torus (ring of circular cross-section with optional inner radius and phi
// section): parameter block and device kernels with the semantics of
// volumes/kernel/TorusImplementation.h:36-1170 and the cached values of
// volumes/UnplacedTorus.h:43-69.
//
// Distances: the reference intersects the ray with each tube surface by
// solving the quartic in closed form (Ferrari, complex arithmetic with
// cbrt/atan2/sin/cos), keeps the roots whose imaginary part is below 1e-9
// and polishes the smallest with two Newton steps
// (TorusImplementation.h:669-887).  That acceptance test is ill-conditioned:
// on near-tangent rays a real pair comes out with an imaginary part above
// 1e-9 and is dropped, or a root is off by 1e-8..1e-3 mm, so the reference's
// answer can miss the first crossing altogether (measured against the exact
// crossings of the same double inputs).  Here the same quartic, with the
// same coefficients, is solved for its real roots by bracketing
// (QuarticRoots below), which finds every root with a sign change to a few
// ulps, and the first root above -1e-9 that passes the phi test is the
// distance, as in the reference.  Starts within ~1e-8 mm of the surface,
// where the reference's surface rule (distances below 1e-9) takes over, run
// the reference's computation unchanged.  Distances
// therefore agree with the reference to the FP64 bar wherever its closed
// form is accurate, and lie within the surface tolerance of the true first
// crossing where it is not.  Contains, Inside and the safeties are the
// reference's and agree exactly.
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

  // The reference's closed form (solveQuartic2, TorusImplementation.h:669-744)
  // for the monic x^4 + b x^3 + c x^2 + d x + e, in the reference's order of
  // operations.  Used only by ClosedFormToBoundary (starts on the surface).
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

  // Root of a function with a sign change on [xl, xh] (fl = f(xl),
  // fh = f(xh)), from the secant point by Newton steps kept inside the
  // bracket, bisecting where a step would leave it or shrink it too slowly;
  // stops when the step falls below rtol |x|.  fn(x, f, df) gives f and f'.
  template <class Fn>
  static VGA_HD double Bracketed(Fn fn, double xl, double xh, double fl, double fh, double rtol) {
    if (fl > 0.) {
      double t = xl;
      xl = xh;
      xh = t;
      t = fl;
      fl = fh;
      fh = t;
    }
    double x = xl - fl * (xh - xl) / (fh - fl), dxold = Abs(xh - xl), dx = dxold, f, df;
    fn(x, f, df);
    for (int it = 0; it < 100 && f != 0.; ++it) {
      if (f < 0.) {
        xl = x;
      } else {
        xh = x;
      }
      if (((x - xh) * df - f) * ((x - xl) * df - f) > 0. || Abs(2. * f) > Abs(dxold * df)) {
        dxold = dx;
        dx = 0.5 * (xh - xl);
        x = xl + dx;
      } else {
        dxold = dx;
        dx = f / df;
        x -= dx;
      }
      if (Abs(dx) <= rtol * Abs(x) + 1e-20) break;
      fn(x, f, df);
    }
    return x;
  }

  static VGA_HD int SignOf(double v) { return (v > 0.) - (v < 0.); }

  // Principal square root (branch cut on the negative real axis), in the
  // cancellation-free form.
  static VGA_HD Cx PSqrt(Cx x) {
    const double l = Sqrt(x.re * x.re + x.im * x.im);
    if (x.re >= 0.) {
      const double s = Sqrt(0.5 * (l + x.re));
      return {s, s > 0. ? 0.5 * x.im / s : 0.};
    }
    const double s = Sqrt(0.5 * (l - x.re));
    return {0.5 * Abs(x.im) / s, x.im < 0. ? -s : s};
  }

  // All four roots of the monic x^4 + a x^3 + b x^2 + c x + d by Ferrari's
  // method through the resolvent cubic, with principal square and cube
  // roots.  No claim of accuracy: CertifiedFirstRoot checks every root.
  static VGA_HD void Ferrari(double a, double b, double c, double d, Cx *z) {
    const double aa = a * a;
    const double alpha = -0.375 * aa + b;
    const double beta = 0.125 * aa * a - 0.5 * a * b + c;
    const double gamma = -3. / 256. * aa * aa + b * aa * (1. / 16.) - 0.25 * a * c + d;
    const double P = -alpha * alpha * (1. / 12.) - gamma;
    const double Q = -alpha * alpha * alpha * (1. / 108.) + alpha * gamma * (1. / 3.) - 0.125 * beta * beta;
    const Cx R = Add(0.5 * Q, CSqrtReal(0.25 * Q * Q + P * P * P * (1. / 27.)));
    const Cx U = CCbrt(R);
    const Cx y = Add(Sub(-5. / 6. * alpha, U), Div(P, Mul(U, 3.)));
    const Cx W = PSqrt(Add(Add(alpha, y), y));
    const Cx second = Sub(-3. * alpha, Mul(y, 2.));
    const Cx third = Div(2. * beta, W);
    const Cx s1 = PSqrt(Add(second, third)), s2 = PSqrt(Sub(second, third));
    const double m = -0.25 * a;
    z[0] = Add(m, Mul(Sub(Neg(W), s1), 0.5));
    z[1] = Add(m, Mul(Add(Neg(W), s1), 0.5));
    z[2] = Add(m, Mul(Sub(W, s2), 0.5));
    z[3] = Add(m, Mul(Add(W, s2), 0.5));
  }

  // The smallest real root >= lo of the monic quartic that `accept` takes
  // (kInfinity if none), from Ferrari's roots certified a posteriori: a
  // polynomial of degree 4 has a root within 4 |f(z)| / |f'(z)| of any z
  // (evaluation rounding added to |f| and taken off |f'|).  Where these
  // disks -- widened to real-centred ones for the roots within their radius
  // of the real axis -- are pairwise disjoint, each holds exactly one root,
  // and a real-centred one a real root (a non-real root would bring its
  // conjugate into the same disk), and simple: Newton steps from the centre
  // converge to it.  Returns false (result unset) where the disks overlap
  // (near-multiple roots, ill-conditioned cases) or the steps do not settle
  // inside the disk.
  template <class Accept>
  static VGA_HD bool CertifiedFirstRoot(double a, double b, double c, double d, double lo, Accept accept,
                                        double &root) {
    auto f = [=](double x, double &fx, double &dfx) {
      fx = (((x + a) * x + b) * x + c) * x + d;
      dfx = ((4. * x + 3. * a) * x + 2. * b) * x + c;
    };
    Cx z[4];
    Ferrari(a, b, c, d, z);
    double cen_re[4], cen_im[4], rad[4];
    bool real[4];
    const double aa = Abs(a), ab = Abs(b), ac = Abs(c), ad = Abs(d);
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const Cx x = z[k];
      const Cx f = Add(d, Mul(Add(c, Mul(Add(b, Mul(Add(a, x), x)), x)), x));
      const Cx fp = Add(c, Mul(Add(2. * b, Mul(Add(3. * a, Mul(x, 4.)), x)), x));
      // |f| <= |Re f| + |Im f|, |f'| >= max(|Re f'|, |Im f'|), |x| <= |Re x| + |Im x|.
      const double r = Abs(x.re) + Abs(x.im);
      const double S = (((r + aa) * r + ab) * r + ac) * r + ad;
      const double Sp = ((4. * r + 3. * aa) * r + 2. * ab) * r + ac;
      const double den = Max(Abs(fp.re), Abs(fp.im)) - 8e-15 * Sp;
      const double rho = den > 0. ? 4. * (Abs(f.re) + Abs(f.im) + 8e-15 * S) / den : kInfinity;
      real[k] = Abs(x.im) <= rho;
      cen_re[k] = x.re;
      cen_im[k] = real[k] ? 0. : x.im;
      rad[k] = real[k] ? rho + Abs(x.im) : rho;
    }
    bool ok = true;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
#pragma unroll
      for (int j = i + 1; j < 4; ++j) {
        const double dx = cen_re[i] - cen_re[j], dy = cen_im[i] - cen_im[j];
        // false for NaN or infinite radii as well
        ok = ok && (dx * dx + dy * dy > (rad[i] + rad[j]) * (rad[i] + rad[j]));
      }
    }
    if (!ok) return false;
    // Real roots in ascending order; the first accepted one.
    double prev = -kInfinity;
    for (int pass = 0; pass < 4; ++pass) {
      int k = -1;
      for (int j = 0; j < 4; ++j)
        if (real[j] && cen_re[j] > prev && (k < 0 || cen_re[j] < cen_re[k])) k = j;
      if (k < 0) break;
      prev = cen_re[k];
      if (cen_re[k] + rad[k] < lo) continue;
      // Newton steps from the centre towards the one simple root of the
      // disk; a step that has not settled to rounding, or an iterate that
      // has left the disk, hands the ray to the bracketing solver.
      double x = cen_re[k], dx = 0.;
#pragma unroll
      for (int it = 0; it < 3; ++it) {
        double fx, dfx;
        f(x, fx, dfx);
        dx = fx / dfx;
        x -= dx;
      }
      if (!(Abs(dx) <= 1e-14 * Abs(x) + 1e-20) || !(Abs(x - cen_re[k]) <= rad[k])) return false;
      if (x >= lo && accept(x)) {
        root = x;
        return true;
      }
    }
    root = kInfinity;
    return true;
  }


  // The smallest real root in [lo, hi] of the monic quartic
  // f = x^4 + a x^3 + b x^2 + c x + d that `accept` takes, or kInfinity.
  // The zeros of f'' (closed form) cut [lo, hi] into pieces on which f' is
  // monotonic, so each holds at most one zero of f'; cut again at those,
  // f is monotonic on every sub-piece, so a sub-piece holds a root exactly
  // when f changes sign over it, and the root is bracketed there.  The
  // sub-pieces are visited in ascending order and the walk stops at the
  // first accepted root.  A double root without a sign change in floating
  // point (a ray tangent to the surface to within rounding) is passed over.
  template <class Accept>
  static VGA_HD double FirstQuarticRoot(double a, double b, double c, double d, double lo, double hi,
                                        Accept accept) {
    auto f = [=](double x, double &fx, double &dfx) {
      fx = (((x + a) * x + b) * x + c) * x + d;
      dfx = ((4. * x + 3. * a) * x + 2. * b) * x + c;
    };
    auto g = [=](double x, double &gx, double &dgx) {
      gx = ((4. * x + 3. * a) * x + 2. * b) * x + c;
      dgx = (12. * x + 6. * a) * x + 2. * b;
    };
    // Zeros of f''/12 = x^2 + (a/2) x + b/6, clipped to [lo, hi].
    double brk[3] = {hi, hi, hi};
    const double disc = 0.0625 * a * a - b * (1. / 6.);
    if (disc > 0.) {
      const double sq = Sqrt(disc), m = -0.25 * a;
      brk[0] = Min(Max(m - sq, lo), hi);
      brk[1] = Min(Max(m + sq, lo), hi);
    }
    double x0 = lo, f0, df, g0, dg;
    f(lo, f0, df);
    if (f0 == 0. && accept(lo)) return lo;
    g(lo, g0, dg);
    double u = lo;
    for (int k = 0; k < 3; ++k) {
      const double w = brk[k];
      if (!(w > u)) continue;
      double gw;
      g(w, gw, dg);
      // A zero of f' inside (u, w) splits the piece.
      double ends[2] = {w, w};
      int ne = 1;
      if (SignOf(g0) * SignOf(gw) < 0) {
        ends[0] = Bracketed(g, u, w, g0, gw, 1e-12);
        ne = 2;
      }
      for (int e = 0; e < ne; ++e) {
        const double x1 = ends[e];
        double f1;
        f(x1, f1, df);
        if (SignOf(f0) * SignOf(f1) < 0) {
          const double r = Bracketed(f, x0, x1, f0, f1, 4.5e-16);
          if (accept(r)) return r;
        } else if (f1 == 0. && f0 != 0. && accept(x1)) {
          return x1;
        }
        x0 = x1;
        f0 = f1;
      }
      u = w;
      g0 = gw;
    }
    return kInfinity;
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

  // The reference's ToBoundary (TorusImplementation.h:764-887): closed-form
  // roots, acceptance |Im| < 1e-9 and Re > -1e-9, two Newton steps on the
  // smallest accepted root (only when the last root examined was accepted),
  // then the surface rule for distances below 1e-9.
  static VGA_HD double ClosedFormToBoundary(const P &t, V3 p, V3 dir, double radius, bool forRmin, bool out) {
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
      if (forRmin) n = V3{-n.x, -n.y, -n.z};
      const double dot = Dot(dir, n);
      if (!out && (dot > 0.)) valid = second;
      if (out && (dot < 0.)) valid = second;
    }
    return valid;
  }

  static VGA_HD double ToBoundary(const P &t, V3 p, V3 dir, double radius, bool forRmin, bool out) {
    const double r0sq = p.x * p.x + p.y * p.y + p.z * p.z;
    const double rdotn = p.x * dir.x + p.y * dir.y + p.z * dir.z;
    const double rsumsq = t.rtor2 + radius * radius;
    const double a = 4. * rdotn;
    const double b = 2. * (r0sq + 2. * rdotn * rdotn - rsumsq + 2. * t.rtor2 * dir.z * dir.z);
    const double c = 4. * (r0sq * rdotn - rsumsq * rdotn + 2. * t.rtor2 * p.z * dir.z);
    const double d = r0sq * r0sq - 2. * r0sq * rsumsq + 4. * t.rtor2 * p.z * p.z +
                     (t.rtor2 - radius * radius) * (t.rtor2 - radius * radius);
    const bool needphi = t.dphi < kTwoPi;
    // A real root is a point q on the tube surface, so |q| <= rtor + radius
    // and t <= |p| + rtor + radius (with a margin for a direction off unit
    // length by rounding).
    const double hi = (Sqrt(r0sq) + t.rtor + radius) * (1. + 1e-6) + 1e-6;
    auto accept = [&](double s) {
      return !needphi || WedgeS::ContainsWithBoundary(t.w, V3{p.x + s * dir.x, p.y + s * dir.y, p.z + s * dir.z});
    };
    double valid;
    if (!CertifiedFirstRoot(a, b, c, d, -1E-9, accept, valid)) valid = FirstQuarticRoot(a, b, c, d, -1E-9, hi, accept);
    // A start within ~1e-8 of this surface: the reference's surface rule
    // (threshold 1e-9 on its own closed-form root) decides, so take the
    // reference's computation whole.
    if (valid < 1E-8) return ClosedFormToBoundary(t, p, dir, radius, forRmin, out);
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
    // Without a phi section the ray crosses the outer surface before the
    // inner one when it starts outside the outer one, and the inner one
    // first when it starts inside the inner one: the other surface's first
    // root comes later and cannot be the minimum.
    const double rxy = Sqrt(lp.x * lp.x + lp.y * lp.y);
    const double dist2 = (rxy - t.rtor) * (rxy - t.rtor) + lp.z * lp.z;
    const bool full = t.dphi >= kTwoPi;
    const bool outerOnly = full && dist2 > (t.rmax + 1e-6) * (t.rmax + 1e-6);
    const bool innerOnly = full && t.rmin > 0 && dist2 < (t.rmin - 1e-6) * (t.rmin - 1e-6);
    // One copy of the solver, run for each surface in turn.
    double minDist = kInfinity;
#pragma unroll 1
    for (int k = 0; k < 2; ++k) {
      const bool inner = k == 1;
      if (inner ? (t.rmin <= 0 || outerOnly) : innerOnly) continue;
      minDist = Min(minDist, ToBoundary(t, lp, v, inner ? t.rmin : t.rmax, inner, false));
    }
    if (t.dphi < kTwoPi) {
      double d1, d2;
      WedgeS::DistanceToIn(t.w, lp, v, d1, d2);
      minDist = PhiLimit(t, lp, v, d1, d2, minDist);
    }
    return inBounds ? minDist : minDist + tubeDistance;
  }

  static VGA_HD double DistanceToOut(const P &t, V3 p, V3 v, double) {
    double distance = kInfinity;
#pragma unroll 1
    for (int k = 0; k < 2; ++k) {
      const bool inner = k == 1;
      if (inner && t.rmin <= 0) continue;
      distance = Min(distance, ToBoundary(t, p, v, inner ? t.rmin : t.rmax, inner, true));
    }
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
