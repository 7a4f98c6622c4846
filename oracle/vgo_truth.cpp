// ORACLE -- test infrastructure only.
//
// Exact-arithmetic stand-in for the Torus ray crossings, used to judge
// distance lanes on which two FP64 solutions of the reference's quartic
// (volumes/kernel/TorusImplementation.h:764-887) disagree beyond the FP64
// bar.  Everything is evaluated in IEEE binary128 (__float128, 113-bit
// significand) from the double inputs, so rounding in the quartic's
// coefficients and in its evaluation is ~1e-34 relative instead of ~1e-16.
//
// For a ray p + t v (t >= 0) and the torus solid
//   S = { rmin <= dist(q) <= rmax } (intersected with the phi wedge),
//   dist(q) = sqrt((sqrt(qx^2 + qy^2) - rtor)^2 + qz^2),
// depth(q) is the signed distance-like depth of q in S: the minimum of
// rmax - dist, dist - rmin (if rmin > 0) and the wedge's plane depth
// (min of the two half-plane distances for an opening <= pi, max of them
// beyond pi).  A distance answer within the surface tolerance tol = 1e-9 mm
// of the first boundary crossing lies in the window
//   outside start (depth(p) < -tol): [first t with depth >= -tol, first t with depth >= tol]
//   inside start  (depth(p) >  tol): [first t with depth <= tol,  first t with depth <= -tol]
// ("first t" = +inf if never).  Only the starts the query is defined for get
// a window -- outside for DistanceToIn, inside for DistanceToOut -- the rest
// (and starts within tol of the surface) get NaN; for those,
// vgo_torus_on_surface tells whether a distance lands on a boundary surface.
// The crossing times of a level L are among the real roots of the quartic for the radii rmax - L and rmin + L and the crossings of the
// two wedge planes; between consecutive candidates the sign of depth - L is
// constant and is read at the midpoint, which also settles roots that are
// double to within the arithmetic.
#include <quadmath.h>

#include <algorithm>
#include <cmath>
#include <vector>

namespace {

using Q = __float128;

Q Absq(Q x) { return x < 0 ? -x : x; }

struct Torus {
  Q rmin, rmax, rtor;
  bool phi, convex;
  Q n1x, n1y, n2x, n2y;  // inward normals of the start and end half-planes
};

Q Depth(const Torus &T, Q x, Q y, Q z) {
  const Q rho = sqrtq(x * x + y * y);
  const Q dist = sqrtq((rho - T.rtor) * (rho - T.rtor) + z * z);
  Q d = T.rmax - dist;
  if (T.rmin > 0) d = std::min(d, dist - T.rmin);
  if (T.phi) {
    const Q w1 = T.n1x * x + T.n1y * y, w2 = T.n2x * x + T.n2y * y;
    d = std::min(d, T.convex ? std::min(w1, w2) : std::max(w1, w2));
  }
  return d;
}

// Root of f on [xl, xh] with a sign change, Newton kept inside the bracket.
template <class Fn>
Q Bracketed(Fn fn, Q xl, Q xh, Q fl) {
  if (fl > 0) std::swap(xl, xh);
  Q x = (xl + xh) / 2, dxold = Absq(xh - xl), dx = dxold, f, df;
  fn(x, f, df);
  for (int it = 0; it < 400 && f != 0; ++it) {
    if (((x - xh) * df - f) * ((x - xl) * df - f) > 0 || Absq(2 * f) > Absq(dxold * df)) {
      dxold = dx;
      dx = (xh - xl) / 2;
      x = xl + dx;
    } else {
      dxold = dx;
      dx = f / df;
      x -= dx;
    }
    if (Absq(dx) <= Q(1e-33) * Absq(x) + Q(1e-60)) break;
    fn(x, f, df);
    if (f < 0)
      xl = x;
    else
      xh = x;
  }
  return x;
}

int Sgn(Q v) { return (v > 0) - (v < 0); }

// Real roots in [lo, hi] of c4 x^4 + c3 x^3 + c2 x^2 + c1 x + c0 (c4 > 0),
// located between the zeros of the derivatives.
void QuarticRootsQ(const Q c[5], Q lo, Q hi, std::vector<Q> &out) {
  auto f = [&](Q x, Q &fx, Q &dfx) {
    fx = (((c[4] * x + c[3]) * x + c[2]) * x + c[1]) * x + c[0];
    dfx = ((4 * c[4] * x + 3 * c[3]) * x + 2 * c[2]) * x + c[1];
  };
  auto g = [&](Q x, Q &gx, Q &dgx) {
    gx = ((4 * c[4] * x + 3 * c[3]) * x + 2 * c[2]) * x + c[1];
    dgx = (12 * c[4] * x + 6 * c[3]) * x + 2 * c[2];
  };
  std::vector<Q> brk{lo};
  // f'' = 12 c4 x^2 + 6 c3 x + 2 c2
  const Q A = 12 * c[4], B = 6 * c[3], C = 2 * c[2];
  const Q disc = B * B - 4 * A * C;
  if (disc > 0) {
    const Q s = sqrtq(disc);
    for (Q r : {(-B - s) / (2 * A), (-B + s) / (2 * A)})
      if (r > lo && r < hi) brk.push_back(r);
  }
  brk.push_back(hi);
  std::vector<Q> cut{lo};
  Q gu, dg, gw;
  g(lo, gu, dg);
  for (size_t k = 1; k < brk.size(); ++k) {
    g(brk[k], gw, dg);
    if (Sgn(gu) * Sgn(gw) < 0)
      cut.push_back(Bracketed(g, brk[k - 1], brk[k], gu));
    else if (gw == 0 && k + 1 < brk.size())
      cut.push_back(brk[k]);
    gu = gw;
  }
  cut.push_back(hi);
  Q fu, df, fw;
  f(lo, fu, df);
  if (fu == 0) out.push_back(lo);
  for (size_t k = 1; k < cut.size(); ++k) {
    f(cut[k], fw, df);
    if (Sgn(fu) * Sgn(fw) < 0)
      out.push_back(Bracketed(f, cut[k - 1], cut[k], fu));
    else if (fw == 0 && fu != 0)
      out.push_back(cut[k]);
    fu = fw;
  }
}

// Real t >= 0 where the ray meets the tube surface of radius r.
void TubeCrossings(const Torus &T, const Q P[3], const Q V[3], Q r, std::vector<Q> &out) {
  // (|q|^2 + R^2 - r^2)^2 - 4 R^2 (qx^2 + qy^2) = 0, q = P + t V.
  const Q R = T.rtor;
  const Q vv = V[0] * V[0] + V[1] * V[1] + V[2] * V[2];
  const Q pv = P[0] * V[0] + P[1] * V[1] + P[2] * V[2];
  const Q pp = P[0] * P[0] + P[1] * P[1] + P[2] * P[2];
  const Q k = pp + R * R - r * r;
  const Q a2 = V[0] * V[0] + V[1] * V[1], a1 = 2 * (P[0] * V[0] + P[1] * V[1]), a0 = P[0] * P[0] + P[1] * P[1];
  const Q c[5] = {k * k - 4 * R * R * a0, 4 * pv * k - 4 * R * R * a1, 4 * pv * pv + 2 * vv * k - 4 * R * R * a2,
                  4 * vv * pv, vv * vv};
  // Every root is within Fujiwara's bound.
  Q bound = 0;
  for (int i = 0; i < 4; ++i) {
    Q m = Absq(c[i] / c[4]);
    if (i == 0) m /= 2;
    bound = std::max(bound, powq(m, Q(1) / Q(4 - i)));
  }
  QuarticRootsQ(c, Q(0), 2 * bound * (1 + Q(1e-20)) + 1, out);
}

// First t >= 0 with depth >= L (enter) or depth <= L (!enter); +inf if never.
Q FirstLevel(const Torus &T, const Q P[3], const Q V[3], Q L, bool enter) {
  std::vector<Q> cand{Q(0)};
  TubeCrossings(T, P, V, T.rmax - L, cand);
  if (T.rmin > 0) TubeCrossings(T, P, V, T.rmin + L, cand);
  if (T.phi) {
    for (int i = 0; i < 2; ++i) {
      const Q nx = i ? T.n2x : T.n1x, ny = i ? T.n2y : T.n1y;
      const Q den = nx * V[0] + ny * V[1];
      if (den != 0) {
        const Q t = (L - nx * P[0] - ny * P[1]) / den;
        if (t > 0) cand.push_back(t);
      }
    }
  }
  std::sort(cand.begin(), cand.end());
  auto holds = [&](Q t) {
    const Q d = Depth(T, P[0] + t * V[0], P[1] + t * V[1], P[2] + t * V[2]);
    return enter ? d >= L : d <= L;
  };
  for (size_t k = 0; k < cand.size(); ++k) {
    const Q a = cand[k];
    const Q b = k + 1 < cand.size() ? cand[k + 1] : 2 * a + 1;
    if (b > a && holds((a + b) / 2)) return a;
  }
  return Q(HUGE_VAL);
}

Torus MakeTorus(const double *params) {
  Torus T;
  T.rmin = params[0];
  T.rmax = params[1];
  T.rtor = params[2];
  const double sphi = params[3], dphi = params[4];
  T.phi = dphi < 2 * M_PI;
  T.convex = dphi <= M_PI;
  const Q s = sphi, e = Q(sphi) + Q(dphi);
  T.n1x = -sinq(s);
  T.n1y = cosq(s);
  T.n2x = sinq(e);
  T.n2y = -cosq(e);
  return T;
}

}  // namespace

extern "C" {

// params: rmin, rmax, rtor, sphi, dphi; out: 0 for DistanceToIn, 1 for
// DistanceToOut; tol: the surface tolerance per ray.  For each ray (p, v)
// writes the window [lo, hi] described above, NaN where the start is not one
// the query is defined for.
void vgo_torus_window(const double *params, int out, long n, const double *px, const double *py, const double *pz,
                      const double *vx, const double *vy, const double *vz, const double *tol, double *lo,
                      double *hi) {
  const Torus T = MakeTorus(params);
#pragma omp parallel for schedule(dynamic, 16)
  for (long i = 0; i < n; ++i) {
    const Q P[3] = {px[i], py[i], pz[i]}, V[3] = {vx[i], vy[i], vz[i]};
    const Q tq = tol[i];
    const Q d0 = Depth(T, P[0], P[1], P[2]);
    const bool enter = d0 < 0;
    if (Absq(d0) <= tq || enter == (out != 0)) {
      lo[i] = hi[i] = NAN;
      continue;
    }
    const Q a = FirstLevel(T, P, V, enter ? -tq : tq, enter);
    const Q b = FirstLevel(T, P, V, enter ? tq : -tq, enter);
    lo[i] = (double)a;
    hi[i] = (double)b;
  }
}

// For each ray and finite distance t: 1 if p + t v lies within tol[i] of a
// boundary surface of the torus (a tube surface, or a phi plane within the
// tube radii), else 0.
void vgo_torus_on_surface(const double *params, long n, const double *px, const double *py, const double *pz,
                          const double *vx, const double *vy, const double *vz, const double *t, const double *tol,
                          unsigned char *on) {
  const Torus T = MakeTorus(params);
  for (long i = 0; i < n; ++i) {
    const Q tq = tol[i];
    on[i] = 0;
    if (!std::isfinite(t[i])) continue;
    const Q tt = t[i];
    const Q x = Q(px[i]) + tt * Q(vx[i]), y = Q(py[i]) + tt * Q(vy[i]), z = Q(pz[i]) + tt * Q(vz[i]);
    const Q rho = sqrtq(x * x + y * y);
    const Q dist = sqrtq((rho - T.rtor) * (rho - T.rtor) + z * z);
    bool hit = Absq(dist - T.rmax) <= tq || (T.rmin > 0 && Absq(dist - T.rmin) <= tq);
    if (T.phi && dist <= T.rmax + tq && dist >= T.rmin - tq)
      hit = hit || Absq(T.n1x * x + T.n1y * y) <= tq || Absq(T.n2x * x + T.n2y * y) <= tq;
    on[i] = hit ? 1 : 0;
  }
}

}  // extern "C"
