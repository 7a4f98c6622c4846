// ORACLE -- test infrastructure only (see vgo.h).
// Wedge: the phi section shared by Tube, Cone, Sphere and Torus
// (volumes/Wedge.h:42-376, source/Wedge.cpp:16-43).
#pragma once
#include "vgo.h"

namespace vgo {

struct Wedge {
  double sphi, dphi;
  double a1x, a1y, a2x, a2y;  // along vectors
  double n1x, n1y, n2x, n2y;  // inward normals
  void Init(double angle, double zero) {
    sphi = zero;
    dphi = angle;
    a1x = std::cos(zero);
    a1y = std::sin(zero);
    a2x = std::cos(zero + angle);
    a2y = std::sin(zero + angle);
    n1x = -std::sin(zero);
    n1y = std::cos(zero);
    n2x = std::sin(zero + angle);
    n2y = -std::cos(zero + angle);
  }
  // Wedge::IsOnSurfaceGeneric, Wedge.h:227-241
  static bool OnSurface(double ax, double ay, double nx, double ny, V3 p) {
    bool c1 = ax * p.x + ay * p.y >= 0.;
    if (!c1) return false;
    return Abs(nx * p.x + ny * p.y) < kTolerance;
  }
  // Wedge::GenericKernelForContainsAndInside, Wedge.h:183-224
  template <bool ForInside>
  void Kernel(V3 p, bool &cin, bool &cout) const {
    double x = p.x, y = p.y;
    double startCheck = (-x * a1y + y * a1x);
    double endCheck = (-a2x * y + a2y * x);
    cout = startCheck < 0.;
    if (dphi < kPi)
      cout |= endCheck < 0.;
    else
      cout &= endCheck < 0.;
    if (ForInside) {
      bool onSurface = OnSurface(a1x, a1y, n1x, n1y, p) || OnSurface(a2x, a2y, n2x, n2y, p);
      cout &= !onSurface;
      cin = !onSurface && !cout;
    }
  }
  // Wedge::ContainsWithBoundary, Wedge.h:160-170
  bool ContainsWithBoundary(V3 p) const {
    bool cin, cout;
    Kernel<true>(p, cin, cout);
    return !cout;
  }
  // Wedge::SafetyToOut / SafetyToIn, Wedge.h:243-297 (used by Sphere/Torus)
  double SafetyToOut(V3 p) const {
    double d1 = p.x * n1x + p.y * n1y;
    double d2 = p.x * n2x + p.y * n2y;
    if (dphi < kPi) return Min(d1, d2);
    return Max(d1, d2);
  }
  double SafetyToIn(V3 p) const {
    double d1 = p.x * n1x + p.y * n1y;
    double d2 = p.x * n2x + p.y * n2y;
    if (dphi < kPi) return Max(-1 * d1, -1 * d2);
    return Min(-1 * d1, -1 * d2);
  }
  // Wedge::DistanceToIn / DistanceToOut, Wedge.h:299-372
  void DistanceToIn(V3 p, V3 dir, double &dw1, double &dw2) const {
    dw1 = kInfinity;
    dw2 = kInfinity;
    double comp1 = dir.x * n1x + dir.y * n1y;
    double comp2 = dir.x * n2x + dir.y * n2y;
    if (comp1 > 0.) {
      double tmp = -(p.x * n1x + p.y * n1y) / comp1;
      if (tmp > 0.) dw1 = tmp;
    }
    if (comp2 > 0.) {
      double tmp = -(p.x * n2x + p.y * n2y) / comp2;
      if (tmp > 0.) dw2 = tmp;
    }
  }
  void DistanceToOut(V3 p, V3 dir, double &dw1, double &dw2) const {
    double comp1 = dir.x * n1x + dir.y * n1y;
    double comp2 = dir.x * n2x + dir.y * n2y;
    dw1 = kInfinity;
    dw2 = kInfinity;
    if (comp1 < 0.) {
      double tmp = -(p.x * n1x + p.y * n1y) / comp1;
      if (tmp > 0.) dw1 = tmp;
    }
    if (comp2 < 0.) {
      double tmp = -(p.x * n2x + p.y * n2y) / comp2;
      if (tmp > 0.) dw2 = tmp;
    }
  }
};

}  // namespace vgo
