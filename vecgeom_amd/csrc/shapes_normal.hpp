// Surface normals: VPlacedVolume::Normal(localPoint, normal&) for the shapes
// whose reference implements it (volumes/PlacedVolume.h:266-272 is the
// asserting default) -- Box, Tube, Orb, Sphere, Cone, Trd, Trapezoid,
// Polycone -- evaluated on a local point, with the reference's conventions:
// the normal starts as a zero vector and is set as the reference sets it, the
// return value is its valid flag.  Each function names the reference code it
// follows; where the reference's code does something other than its comments
// say (the Trd, the Polycone), the function does what the code does.
//
// Transcendentals of the point (the Sphere's and the Cone's atan2, the
// Sphere's approximate normal's sin and cos) come from the device math
// library; everything else is separately rounded as the reference's scalar
// code computes it (-ffp-contract=off).
#pragma once

#include "shapes_box_tube.hpp"
#include "shapes_cone.hpp"
#include "shapes_polycone.hpp"
#include "shapes_simple.hpp"
#include "shapes_sphere.hpp"
#include "shapes_trap.hpp"

namespace vga {

// BoxImplementation::NormalKernel, BoxImplementation.h:842-898.
VGA_HD bool ShapeNormal(const BoxP &b, V3 p, V3 &n) {
  const double delta = 100. * kTolerance;
  const double c[3] = {p.x, p.y, p.z};
  double m[3] = {0, 0, 0};
  double nsurf = 0;
#pragma unroll
  for (int dim = 0; dim < 3; ++dim) {
    if (Abs(Abs(c[dim]) - b.d[dim]) < delta) {
      nsurf = nsurf + 1.;
      m[dim] = m[dim] + (c[dim] < 0 ? -1. : 1.);
    }
  }
  const double scale = (nsurf == 3.) ? 0.5773502691896258 : (nsurf == 2.) ? 0.7071067811865475 : 1.;
  n = (nsurf >= 2.) ? V3{m[0] * scale, m[1] * scale, m[2] * scale} : V3{m[0], m[1], m[2]};
  return nsurf > 0;
}

// OrbImplementation::NormalKernel, OrbImplementation.h:208-229.
VGA_HD bool ShapeNormal(const OrbP &o, V3 p, V3 &n) {
  const double rad2 = p.x * p.x + p.y * p.y + p.z * p.z;
  const double rad = Sqrt(rad2);
  n = V3{p.x / rad, p.y / rad, p.z / rad};
  return (rad2 <= o.rTolO * o.rTolO) && (rad2 >= o.rTolI * o.rTolI);
}

// UnplacedTube::Normal, source/UnplacedTube.cpp:141-229: top and bottom lids
// (z within kTolerance, radius in [rmin, rmax] exactly), inner and outer
// walls (radius within kTolerance, z within kTolerance of the range), the
// two phi planes (outward: the wedge's inward normals negated); the first
// surface found sets the vector, later ones add, several are divided by
// sqrt(count).
VGA_HD bool ShapeNormal(const TubeP &t, V3 p, V3 &n) {
  n = V3{0, 0, 0};
  int ns = 0;
  const double x2y2 = Sqrt(p.x * p.x + p.y * p.y);
  const bool inZ = (p.z < t.z + kTolerance) && (p.z > -t.z - kTolerance);
  const bool inR = (x2y2 >= t.rmin) && (x2y2 <= t.rmax);
  if (inR && Abs(p.z - t.z) <= kTolerance) {
    n = V3{0., 0., 1.};
    ns++;
  }
  if (inR && Abs(p.z + t.z) <= kTolerance) {
    if (ns > 0)
      n.z += -1;
    else
      n = V3{0.0, 0.0, -1};
    ns++;
  }
  if (t.rmin > 0. && inZ && Abs(x2y2 - t.rmin) <= kTolerance) {
    const double inv = 1. / x2y2;
    if (ns == 0) {
      n = V3{-p.x * inv, -p.y * inv, 0.0};
    } else {
      n.x += -p.x * inv;
      n.y += -p.y * inv;
    }
    ns++;
  }
  if (inZ && Abs(x2y2 - t.rmax) <= kTolerance) {
    const double inv = 1. / x2y2;
    if (ns > 0) {
      n.x += p.x * inv;
      n.y += p.y * inv;
    } else {
      n = V3{p.x * inv, p.y * inv, 0.0};
    }
    ns++;
  }
  if (t.dphi < kTwoPi && inR) {
    const WedgeP &w = t.w;
    if (WedgeS::OnSurface(w.a1x, w.a1y, w.n1x, w.n1y, p)) {
      const V3 o1{-w.n1x, -w.n1y, -0.0};
      n = (ns == 0) ? o1 : n + o1;
      ns++;
    }
    if (WedgeS::OnSurface(w.a2x, w.a2y, w.n2x, w.n2y, p)) {
      const V3 o2{-w.n2x, -w.n2y, -0.0};
      n = (ns == 0) ? o2 : n + o2;
      ns++;
    }
  }
  if (ns > 1) {
    const double s = Sqrt((double)ns);
    n = V3{n.x / s, n.y / s, n.z / s};
  }
  return ns != 0;
}

// UnplacedTrd::Normal, source/UnplacedTrd.cpp:154-199, as it executes: its
// six "if (onSurf && Fac) norm[0] = ...; norm[1] = ...; norm[2] = ...;
// nosurface++;" lines guard only the x assignment, and each Fac is a double
// converted to bool (nonzero or NaN).  x is the x-component of the last
// face whose Fac holds (+x and -x faces: 4 dy1 dz; the others 0), y is 0,
// z is 4 dx2 dy1, and the result is always valid.
VGA_HD bool ShapeNormal(const TrdP &t, V3 p, V3 &n) {
  const double xx1 = t.dx1, xx2 = t.dx2, yy1 = t.dy1, yy2 = t.dy2, zz = t.dz;
  const double XplusX = 2.0 * yy1 * 2.0 * zz;
  const double XplusZ = 2.0 * yy1 * (xx1 - xx2), XminusZ = 2.0 * yy1 * (-xx1 + xx2);
  const double YplusY = -2.0 * xx1 * 2.0 * zz;
  const double YplusZ = 2.0 * xx1 * (-yy1 + yy2), YminusZ = 2.0 * xx1 * (yy1 - yy2);
  const double ZplusZ = -2.0 * xx2 * 2.0 * yy2, ZminusZ = 2.0 * xx2 * 2.0 * yy1;
  const double zero = 0.0;
  const double ax = p.x - xx2, bx = p.x + xx2, ay = p.y - yy2, by = p.y + yy2, az = p.z - zz;
  const bool fpx = (XplusX * ax + zero * ay + XplusZ * az) != 0;
  const bool fmx = (XplusX * bx + zero * ay + XminusZ * az) != 0;
  const bool fpy = (zero * ax + YplusY * ay + YplusZ * az) != 0;
  const bool fmy = (zero * ax + YplusY * by + YminusZ * az) != 0;
  const bool fpz = (zero * ax + zero * ay + ZplusZ * az) != 0;
  const bool fmz = (zero * ax + zero * by + ZminusZ * az) != 0;
  // The last true Fac decides x; only the +x and -x faces carry a nonzero x.
  const bool lastIsX = !(fpy || fmy || fpz || fmz) && (fpx || fmx);
  n = V3{lastIsX ? XplusX : 0.0, 0.0, ZminusZ};
  return true;
}

// UnplacedTrapezoid::Normal, source/UnplacedTrapezoid.cpp:352-425, and its
// ApproxSurfaceNormal, :583-612 (the PlaneShell form, the reference's
// default build).
VGA_HD bool ShapeNormal(const TrapP &t, V3 p, V3 &n) {
  int ns = 0;
  V3 sum{0., 0., 0.};
  double safe = kInfinity;
  int imin = 0;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const double dist = t.A[i] * p.x + t.B[i] * p.y + t.C[i] * p.z + t.D[i];
    if (Abs(dist) <= kHalfTolerance) {
      ns++;
      sum = sum + V3{t.A[i], t.B[i], t.C[i]};
    }
    if (Abs(dist) < safe) {
      safe = Abs(dist);
      imin = i;
    }
  }
  const double distz = Abs(p.z) - t.dz;
  if (Abs(distz) <= kHalfTolerance) {
    ns++;
    sum = (p.z >= 0.) ? sum + V3{0., 0., 1.} : sum - V3{0., 0., 1.};
  }
  if (ns == 0) {
    const double safez = Abs(Abs(p.z) - t.dz);
    if (safe < safez)
      n = V3{t.A[imin], t.B[imin], t.C[imin]};
    else
      n = (p.z > 0) ? V3{0, 0, 1} : V3{0, 0, -1};
  } else if (ns == 1) {
    n = sum;
  } else {
    const double s = Sqrt(sum.x * sum.x + sum.y * sum.y + sum.z * sum.z + kMinimum);
    n = V3{sum.x / s, sum.y / s, sum.z / s};
  }
  return ns != 0;
}

// The cone's two normals: ConeImplementation::NormalKernel
// (ConeImplementation.h:2002-2128, a placed Cone's Normal; Kernel = true)
// and UnplacedCone::Normal (source/UnplacedCone.cpp:31-141, a Polycone
// section's; Kernel = false).  They differ in the sec(alpha) scaling (the
// kernel multiplies by the cached inverse) and in their atan2 (the kernel
// calls the backend's, which maps x == 0 to +-pi/2 or 0).
template <bool Kernel>
VGA_HD bool ConeNormal(const ConeP &c, V3 p, V3 &n) {
  const double delta = 0.5 * kTolerance;
  int ns = 0;
  V3 sum{0., 0., 0.};
  const double distZ = Abs(Abs(p.z) - c.dz);
  const double rho = Sqrt(p.x * p.x + p.y * p.y);
  const double aMin = Abs((rho - p.z * c.tanRMin) - (c.rmin2 - c.dz * c.tanRMin));
  const double aMax = Abs((rho - p.z * c.tanRMax) - (c.rmax2 - c.dz * c.tanRMax));
  const double distRMin = Kernel ? aMin * c.invSecRMin : aMin / c.secRMin;
  const double distRMax = Kernel ? aMax * c.invSecRMax : aMax / c.secRMax;
  const bool hasRmin = c.rmin1 != 0. || c.rmin2 != 0.;
  const bool full = c.fullPhi != 0.;
  double distSPhi = kInfinity, distEPhi = kInfinity;
  if (!full) {
    if (rho != 0.) {
      double pPhi = Kernel ? ConeS::ATan2(p.y, p.x) : atan2(p.y, p.x);
      if (pPhi < c.sphi - delta)
        pPhi += kTwoPi;
      else if (pPhi > c.sphi + c.dphi + delta)
        pPhi -= kTwoPi;
      distSPhi = Abs(pPhi - c.sphi);
      distEPhi = Abs(pPhi - c.sphi - c.dphi);
    } else if (!(c.rmin1 != 0.) || !(c.rmin2 != 0.)) {
      distSPhi = 0.;
      distEPhi = 0.;
    }
  }
  if (distRMax <= delta) {
    ns++;
    if (rho > delta)
      sum = sum + V3{p.x / rho / c.secRMax, p.y / rho / c.secRMax,
                     Kernel ? -c.tanRMax * c.invSecRMax : -c.tanRMax / c.secRMax};
    else
      sum = sum + V3{0, 0, 0};
  }
  if (hasRmin && distRMin <= delta) {
    ns++;
    if (rho > delta)
      sum = sum + V3{-p.x / rho / c.secRMin, -p.y / rho / c.secRMin,
                     Kernel ? c.tanRMin * c.invSecRMin : c.tanRMin / c.secRMin};
    else
      sum = sum + V3{0., 0., 0.};
  }
  if (!full) {
    if (distSPhi <= delta) {
      ns++;
      sum = sum + V3{c.sinSPhi, -c.cosSPhi, 0};
    }
    if (distEPhi <= delta) {
      ns++;
      sum = sum + V3{-c.sinEPhi, c.cosEPhi, 0};
    }
  }
  if (distZ <= delta) {
    ns++;
    sum = (p.z >= 0.) ? sum + V3{0., 0., 1.} : sum - V3{0., 0., 1.};
  }
  if (ns == 0) {
    n = V3{0, 0, 0};
  } else if (ns == 1) {
    n = sum;
  } else {
    const double s = Sqrt(sum.x * sum.x + sum.y * sum.y + sum.z * sum.z + kMinimum);
    n = V3{sum.x / s, sum.y / s, sum.z / s};
  }
  return ns != 0;
}

VGA_HD bool ShapeNormal(const ConeP &c, V3 p, V3 &n) { return ConeNormal<true>(c, p, n); }

// UnplacedPolycone::Normal, source/UnplacedPolycone.cpp:684-698: -z below
// the sections, +z above (both valid), else the section cone's
// UnplacedCone::Normal of the point as given (the reference does not move it
// into the section's frame).
VGA_HD bool ShapeNormal(const PolyconeP &q, V3 p, V3 &n) {
  const int index = PolyconeS::SectionIndex(q, p.z);
  if (index < 0) {
    n = V3{0., 0., index == -1 ? -1. : 1.};
    return true;
  }
  return ConeNormal<false>(PolyconeS::Sec(q, index), p, n);
}

// SphereImplementation::ApproxSurfaceNormalKernel, SphereImplementation.h:
// 463-610: the nearest surface's normal for a point on none.  The accessors
// the reference calls are swapped (UnplacedSphere.h:262-270):
// IsFullPhiSphere() is the flag for a full sphere in phi and theta.
VGA_HD V3 SphereApproxNormal(const SphereP &s, V3 p, double radius, double rho, double pPhi, double pTheta) {
  const bool phiCut = !(s.fullPhi != 0. && s.fullTheta != 0.) && rho > 0.;
  const bool thetaCut = s.fullTheta == 0. && radius > 0.;
  const double distRMax = Abs(radius - s.fRmax), distRMin = Abs(radius - s.fRmin);
  int side;
  double distMin;
  if (s.fRmin > 0. && distRMin < distRMax) {
    distMin = distRMin;
    side = 0;
  } else {
    distMin = distRMax;
    side = 1;
  }
  if (pPhi < 0.) pPhi = pPhi + (2 * kPi);
  double distSPhi = 0., distEPhi = 0.;
  if (phiCut) {
    distSPhi = (s.fSPhi < 0. ? Abs(pPhi - (s.fSPhi + 2 * kPi)) : Abs(pPhi - s.fSPhi)) * rho;
    distEPhi = Abs(pPhi - s.fSPhi - s.fDPhi) * rho;
    if (distSPhi < distEPhi) {
      if (distSPhi < distMin) {
        distMin = distSPhi;
        side = 2;
      }
    } else if (distEPhi < distMin) {
      distMin = distEPhi;
      side = 3;
    }
  }
  if (thetaCut) {
    const double distSTheta = Abs(pTheta - s.fSTheta) * radius;
    const double distETheta = Abs(pTheta - s.fSTheta - s.fDTheta) * radius;
    if (distSTheta < distETheta) {
      if (distSTheta < distMin) side = 4;
    } else if (distETheta < distMin) {
      side = 5;
    }
  }
  switch (side) {
    case 0: return V3{-p.x / radius, -p.y / radius, -p.z / radius};
    case 1: return V3{p.x / radius, p.y / radius, p.z / radius};
    case 2: return V3{s.sinSPhi, -s.cosSPhi, 0};
    case 3: return V3{-s.sinEPhi, s.cosEPhi, 0};
    case 4: return V3{-s.cosSTheta * cos(pPhi), -s.cosSTheta * sin(pPhi), s.sinSTheta};
    default: return V3{s.cosETheta * cos(pPhi), s.cosETheta * sin(pPhi), s.sinETheta};
  }
}

// SphereImplementation::NormalKernel, SphereImplementation.h:619-760: the sum
// of the normals of every surface within tolerance (r shells 0.5e-9 mm, phi
// planes and theta cones half the angular tolerance), unit length when
// several, the approximate normal (invalid) when none.
VGA_HD bool ShapeNormal(const SphereP &s, V3 p, V3 &n) {
  const double radius = Sqrt(p.x * p.x + p.y * p.y + p.z * p.z);
  const double rho = Sqrt(p.x * p.x + p.y * p.y);
  const double pPhi0 = SphereS::ATan2(p.y, p.x);
  const double pTheta = atan2(rho, p.z);
  const double halfAng = 0.5 * SphereS::kAngTol, halfCar = 0.5 * 1e-9;
  const bool fullPhiOnly = s.fullPhi != 0.;                 // IsFullSphere()
  const bool fullBoth = s.fullPhi != 0. && s.fullTheta != 0.;  // IsFullPhiSphere()
  const bool fullTheta = s.fullTheta != 0.;                  // IsFullThetaSphere()
  double pPhi = pPhi0;
  if (rho > 0. && !fullPhiOnly && pPhi < s.fSPhi - halfAng) pPhi = pPhi + (2 * kPi);
  if (rho > 0. && !fullPhiOnly && pPhi > s.ePhi + halfAng) pPhi = pPhi - (2 * kPi);
  double distSPhi = kInfinity, distEPhi = kInfinity, distSTheta = kInfinity, distETheta = kInfinity;
  if (!fullBoth && rho > 0.) {
    distSPhi = Abs(pPhi - s.fSPhi);
    distEPhi = Abs(pPhi - s.ePhi);
  }
  if (!fullBoth && !(s.fRmin != 0.)) {
    distSPhi = 0.;
    distEPhi = 0.;
  }
  V3 nTs{0, 0, 0}, nTe{0, 0, 0};
  if (!fullTheta && rho > 0.) {
    distSTheta = Abs(pTheta - s.fSTheta);
    distETheta = Abs(pTheta - s.eTheta);
    nTs = V3{-s.cosSTheta * p.x / rho, -s.cosSTheta * p.y / rho, s.sinSTheta};
  }
  if (!fullTheta && rho != 0.) nTe = V3{s.cosETheta * p.x / rho, s.cosETheta * p.y / rho, -s.sinETheta};
  if (!fullTheta && !(s.fRmin != 0.) && s.fSTheta != 0.) {
    distSTheta = 0.;
    nTs = V3{0., 0., -1.};
  }
  if (!fullTheta && !(s.fRmin != 0.) && s.eTheta < kPi) {
    distETheta = 0.;
    nTe = V3{0., 0., 1.};
  }
  V3 nR{0, 0, 0};
  if (radius != 0.) nR = V3{p.x / radius, p.y / radius, p.z / radius};
  const double distRMax = Abs(radius - s.fRmax);
  const double distRMin = (s.fRmin > 0) ? Abs(radius - s.fRmin) : 0.;
  int ns = 0;
  V3 sum{0., 0., 0.};
  if (distRMax <= halfCar) {
    ns++;
    sum = sum + nR;
  }
  if (s.fRmin != 0. && distRMin <= halfCar) {
    ns++;
    sum = sum - nR;
  }
  if (!fullBoth && distSPhi <= halfAng) {
    ns++;
    sum = sum + V3{s.sinSPhi, -s.cosSPhi, 0};
  }
  if (!fullBoth && distEPhi <= halfAng) {
    ns++;
    sum = sum + V3{-s.sinEPhi, s.cosEPhi, 0};
  }
  // With a theta cut fullBoth is false, so the reference's nZ alternative
  // for a point at the centre of a full-phi sphere never applies here.
  if (!fullTheta && distSTheta <= halfAng && s.fSTheta > 0.) {
    ns++;
    sum = sum + nTs;
  }
  if (!fullTheta && distETheta <= halfAng && s.eTheta < kPi) {
    ns++;
    sum = sum + nTe;
    if (sum.z == 0.) sum = sum + V3{0., 0., 1.};
  }
  if (ns == 0) {
    n = SphereApproxNormal(s, p, radius, rho, pPhi0, pTheta);
  } else if (ns == 1) {
    n = sum;
  } else {
    const double mag = Sqrt(sum.x * sum.x + sum.y * sum.y + sum.z * sum.z);
    n = V3{sum.x * 1. / mag, sum.y * 1. / mag, sum.z * 1. / mag};
  }
  return ns > 0;
}

}  // namespace vga
