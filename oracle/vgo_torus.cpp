// ORACLE -- test infrastructure only (see vgo.h).
//
// Torus restated from the reference:
//   cached values       volumes/UnplacedTorus.h:43-69 (bounding tube rtor-rmax-tol
//                       .. rtor+rmax+tol, half-length rmax)
//   complex helpers     volumes/kernel/TorusImplementation.h:36-300
//   solveQuartic2       TorusImplementation.h:422-505 (Ferrari, complex)
//   Contains / Inside   TorusImplementation.h:522-640
//   ToBoundary          TorusImplementation.h:768-887 (roots, 2 Newton steps)
//   DistanceToIn        TorusImplementation.h:889-1020 -- the bounding tube is
//                       the HollowTube specialisation (rmin treated, phi not)
//   DistanceToOut       TorusImplementation.h:1022-1098
//   SafetyToIn / Out    TorusImplementation.h:1100-1170
// Deviation: the reference's placed DistanceToIn transforms the direction
// with the point transform (TorusImplementation.h:906); here the placement
// rotates the direction only (DirectionUsesPointTransform is false).
#include "vgo.h"
#include "vgo_wedge.h"

namespace vgo {

Shape *MakeTube(const double *p, int n);

namespace {

// Complex<T> of the reference with its operators' operation order.
struct Cx {
  double re, im;
};
inline Cx Add(Cx x, Cx y) { return {x.re + y.re, x.im + y.im}; }
inline Cx Add(double s, Cx x) { return {x.re + s, x.im}; }  // T + Complex and Complex + T
inline Cx Sub(Cx x, Cx y) { return {x.re - y.re, x.im - y.im}; }
inline Cx Sub(double s, Cx x) { return {s - x.re, -x.im}; }
inline Cx Neg(Cx x) { return {-x.re, -x.im}; }
inline Cx Mul(Cx x, Cx y) { return {x.re * y.re - x.im * y.im, x.re * y.im + x.im * y.re}; }
inline Cx Mul(Cx x, double s) { return {x.re * s, x.im * s}; }
inline Cx Conj(Cx x) { return {x.re, -x.im}; }
inline Cx DivReal(Cx x, double y) {
  const double inv = 1. / y;
  return {x.re * inv, x.im * inv};
}
inline Cx Div(double s, Cx y) {  // T / Complex
  const Cx t = Mul(y, Conj(y));
  return DivReal(Mul(Conj(y), s), t.re);
}
inline Cx CSqrt(Cx x) {
  const double l = std::sqrt(x.re * x.re + x.im * x.im);
  return {std::sqrt(0.5 * (l + x.re)), std::sqrt(0.5 * (l - x.re))};
}
inline Cx CSqrtReal(double x) {
  const double imcoef = (x >= 0.) ? 0. : 1.;
  const double realcoef = (x >= 0) ? 1. : 0.;
  const double l = std::sqrt(Abs(x));
  return {realcoef * l, imcoef * l};
}
inline Cx CCbrt(Cx x) {
  double r, s, c;
  if (x.im != 0.0) {
    r = std::sqrt(x.re * x.re + x.im * x.im);
    const double angle = ATan2(x.im, x.re);
    const double na = angle / 3.;
    s = std::sin(na);
    c = std::cos(na);
  } else {
    r = x.re;
    s = 0.;
    c = 1.;
  }
  const double rc = std::cbrt(r);
  return {rc * c, rc * s};
}

constexpr double inv256 = 1. / 256, inv16 = 1. / 16, inv3 = 1. / 3, inv27 = 1. / 27, inv6 = 1. / 6,
                 inv108 = 1. / 108, inv12 = 1. / 12;

// solveQuartic2: monic x^4 + b x^3 + c x^2 + d x + e.
void SolveQuartic(double b, double c, double d, double e, Cx *roots) {
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

struct Torus : Shape {
  double rmin, rmax, rtor, sphi, dphi, rmin2, rmax2, rtor2;
  Wedge wedge;
  Shape *btube = nullptr;   // bounding tube, phi not treated
  Shape *btubeX = nullptr;  // bounding tube with the phi section (Extent only)

  Torus(const double *a) : rmin(a[0]), rmax(a[1]), rtor(a[2]), sphi(a[3]), dphi(a[4]) {
    rmin2 = rmin * rmin;
    rmax2 = rmax * rmax;
    rtor2 = rtor * rtor;
    wedge.Init(dphi, sphi);
    const double t[5] = {rtor - rmax - kTolerance, rtor + rmax + kTolerance, rmax, 0., kTwoPi};
    btube = MakeTube(t, 5);
    const double tx[5] = {rtor - rmax - kTolerance, rtor + rmax + kTolerance, rmax, sphi, dphi};
    btubeX = MakeTube(tx, 5);
  }
  ~Torus() override {
    delete btube;
    delete btubeX;
  }

  // GenericKernelForContainsAndInside<ForInside, notForDisk>
  template <bool ForInside, bool Phi>
  void Kernel(V3 p, bool &cin, bool &cout) const {
    const double rxy = std::sqrt(p.x * p.x + p.y * p.y);
    const double radsq = (rxy - rtor) * (rxy - rtor) + p.z * p.z;
    cout = radsq > MakePlusTolerantSquare<ForInside>(rmax, rmax2);
    if (ForInside) cin = radsq < MakeMinusTolerantSquare<ForInside>(rmax, rmax2);
    if (cout) return;
    cout |= radsq < MakeMinusTolerantSquare<ForInside>(rmin, rmin2);
    if (ForInside) cin &= radsq > MakePlusTolerantSquare<ForInside>(rmin, rmin2);
    if (cout) return;
    if (dphi < kTwoPi && Phi) {
      bool wi, wo;
      wedge.Kernel<ForInside>(p, wi, wo);
      cout |= wo;
      if (ForInside) cin &= wi;
    }
  }
  bool Contains(V3 p) const override {
    bool cin = false, cout;
    Kernel<false, true>(p, cin, cout);
    return !cout;
  }
  bool ContainsDisk(V3 p) const {
    bool cin = false, cout;
    Kernel<false, false>(p, cin, cout);
    return !cout;
  }
  int Inside(V3 p) const override {
    bool cin = false, cout;
    Kernel<true, true>(p, cin, cout);
    int r = kSurface;
    if (cout) r = kOutside;
    if (cin) r = kInside;
    return r;
  }

  static double CheckZero(double b, double c, double d, double e, double x) {
    const double x2 = x * x;
    return x2 * x2 + b * x2 * x + c * x2 + d * x + e;
  }
  static double NewtonIter(double b, double c, double d, double e, double x, double fold) {
    (void)e;
    const double x2 = x * x;
    const double fprime = 4 * x2 * x + 3 * b * x2 + 2 * c * x + d;
    return x - fold / fprime;
  }

  template <bool ForRmin>
  double ToBoundary(V3 p, V3 dir, double radius, bool out) const {
    const double r0sq = p.x * p.x + p.y * p.y + p.z * p.z;
    const double rdotn = p.x * dir.x + p.y * dir.y + p.z * dir.z;
    const double rsumsq = rtor2 + radius * radius;
    const double a = 4. * rdotn;
    const double b = 2. * (r0sq + 2. * rdotn * rdotn - rsumsq + 2. * rtor2 * dir.z * dir.z);
    const double c = 4. * (r0sq * rdotn - rsumsq * rdotn + 2. * rtor2 * p.z * dir.z);
    const double d = r0sq * r0sq - 2. * r0sq * rsumsq + 4. * rtor2 * p.z * p.z +
                     (rtor2 - radius * radius) * (rtor2 - radius * radius);
    const bool needphi = dphi < kTwoPi;
    Cx roots[4];
    SolveQuartic(a, b, c, d, roots);
    double valid = kInfinity, second = kInfinity;
    bool have = false;
    for (int k = 0; k < 4; ++k) {
      have = Abs(roots[k].im) < 1E-9 && roots[k].re > -1E-9;
      if (needphi && have) {
        const double t = roots[k].re;
        have = wedge.ContainsWithBoundary(V3{p.x + t * dir.x, p.y + t * dir.y, p.z + t * dir.z});
      }
      if (k == 0) {
        if (have) valid = roots[0].re;
        if (out && have) second = valid;
      } else {
        if (have) valid = Min(roots[k].re, valid);
        if (out && have) second = Max(roots[k].re, valid);
      }
    }
    // The reference refines only when the last root examined was valid.
    have = have && (valid < kInfinity);
    if (have) {
      valid = NewtonIter(a, b, c, d, valid, CheckZero(a, b, c, d, valid));
      valid = NewtonIter(a, b, c, d, valid, CheckZero(a, b, c, d, valid));
    }
    if (valid < 1E-9) {
      valid = 0.;
      const double rho = std::sqrt(p.x * p.x + p.y * p.y);
      const double red = (rho - rtor) / rho;
      V3 norm{p.x * red, p.y * red, p.z};
      if (ForRmin) norm = V3{-norm.x, -norm.y, -norm.z};
      const double dot = Dot(dir, norm);
      if (!out && (dot > 0.)) valid = second;
      if (out && (dot < 0.)) valid = second;
    }
    return valid;
  }

  // Phi planes shared by DistanceToIn and DistanceToOut.
  double PhiLimit(V3 p, V3 v, double dist1, double dist2, double current) const {
    if (dist1 < current) {
      const V3 ip{p.x + v.x * dist1, p.y + v.y * dist1, p.z + v.z * dist1};
      if (ContainsDisk(ip)) {
        const double diri = ip.x * wedge.a1x + ip.y * wedge.a1y;
        if (diri >= 0) current = dist1;
      }
    }
    if (dist2 < current) {
      const V3 ip{p.x + v.x * dist2, p.y + v.y * dist2, p.z + v.z * dist2};
      if (ContainsDisk(ip)) {
        const double diri2 = ip.x * wedge.a2x + ip.y * wedge.a2y;
        if (diri2 >= 0) current = dist2;
      }
    }
    return current;
  }

  double DistanceToIn(V3 p, V3 v, double stepMax) const override {
    const bool inBounds = btube->Contains(p);
    const double tubeDistance = btube->DistanceToIn(p, v, stepMax);
    if (!inBounds && tubeDistance == kInfinity) return kInfinity;
    V3 lp = p;
    if (!inBounds) lp = V3{p.x + v.x * tubeDistance, p.y + v.y * tubeDistance, p.z + v.z * tubeDistance};
    const double dout = ToBoundary<false>(lp, v, rmax, false);
    double din = kInfinity;
    if (rmin > 0) din = ToBoundary<true>(lp, v, rmin, false);
    double minDist = Min(dout, din);
    if (dphi < kTwoPi) {
      double d1, d2;
      wedge.DistanceToIn(lp, v, d1, d2);
      minDist = PhiLimit(lp, v, d1, d2, minDist);
    }
    return inBounds ? minDist : minDist + tubeDistance;
  }

  double DistanceToOut(V3 p, V3 v, double) const override {
    const double dout = ToBoundary<false>(p, v, rmax, true);
    double din = kInfinity;
    if (rmin > 0) din = ToBoundary<true>(p, v, rmin, true);
    double distance = Min(dout, din);
    if (dphi < kTwoPi) {
      double d1, d2;
      wedge.DistanceToOut(p, v, d1, d2);
      distance = PhiLimit(p, v, d1, d2, distance);
    }
    return distance;
  }

  double SafetyToIn(V3 p) const override {
    const double rxy = std::sqrt(p.x * p.x + p.y * p.y);
    const double rad = std::sqrt((rxy - rtor) * (rxy - rtor) + p.z * p.z);
    double safety = rad - rmax;
    if (rmin != 0) safety = Max(rmin - rad, rad - rmax);
    if (dphi < kTwoPi && (rxy != 0.)) safety = Max(wedge.SafetyToIn(p), safety);
    return safety;
  }
  double SafetyToOut(V3 p) const override {
    const double rxy = std::sqrt(p.x * p.x + p.y * p.y);
    const double rad = std::sqrt((rxy - rtor) * (rxy - rtor) + p.z * p.z);
    double safety = rmax - rad;
    if (rmin != 0) safety = Min(rad - rmin, rmax - rad);
    if (dphi < kTwoPi) safety = Min(wedge.SafetyToOut(p), safety);
    return safety;
  }
  void Extent(V3 &lo, V3 &hi) const override { btubeX->Extent(lo, hi); }
};

}  // namespace

Shape *MakeTorus(const double *p, int n) {
  if (n != 5 || !(p[2] > p[1] + kTolerance) || !(p[1] > p[0]) || !(p[0] >= 0)) return nullptr;
  return new Torus(p);
}

}  // namespace vgo
