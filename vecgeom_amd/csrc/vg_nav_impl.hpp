// Navigation functions shared by the device kernels (vg_nav.hip) and a
// host-only build used to check their logic on the CPU.  Behaviour and
// reference citations are described in vg_nav.hip.
//
// With VGA_NAV_HOST_CHECKS defined (host builds only) every table access is
// bounds-checked and a violation aborts with a message.
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#if defined(VGA_NAV_HOST_CHECKS)
#include <cstdio>
#include <cstdlib>
#define VGA_NAV_CHECK(cond, msg)                                  \
  do {                                                            \
    if (!(cond)) {                                                \
      std::fprintf(stderr, "navigation check failed: %s\n", msg); \
      std::abort();                                               \
    }                                                             \
  } while (0)
#else
#define VGA_NAV_CHECK(cond, msg) ((void)0)
#endif
#define VGA_NAV inline __host__ __device__

#include "shapes.hpp"
#include "vg_internal.hpp"

namespace vga {
namespace nav {

constexpr float kInfF = __builtin_huge_valf();

struct NavDev {
  const PlacedRec *placed;
  const double *params;
  const int32_t *daughters;
  const float *aabb;
  const B32 *b32;        // FP32 bound records (mixed-precision navigator)
  const float *b32side;  // their variable-length data
  const float *slot_box; // daughter boxes in slot order, 8 floats each
  int32_t world;
  int32_t n_placed;
  int32_t n_daughters;
  int32_t pad;
  // Slots [lds_lo, lds_hi) -- the world volume's daughters -- read from a
  // copy in LDS made by the trace kernel (none elsewhere).
  const float *lds_box = nullptr;
  int32_t lds_lo = 0;
  int32_t lds_hi = 0;
  // The FP32 bound records of the same slots' daughters, in slot order
  // (mixed trace kernel only; null elsewhere).
  const B32 *lds_bound = nullptr;
  VGA_NAV const PlacedRec &Pv(int id) const {
    VGA_NAV_CHECK(id >= 0 && id < n_placed, "placed id out of range");
    return placed[id];
  }
  VGA_NAV int Daughter(int k) const {
    VGA_NAV_CHECK(k >= 0 && k < n_daughters, "daughter slot out of range");
    return daughters[k];
  }
  VGA_NAV const float *Box(int id) const {
    VGA_NAV_CHECK(id >= 0 && id < n_placed, "box id out of range");
    return aabb + 6 * (size_t)id;
  }
  VGA_NAV const float *SlotBox(int k) const {
    VGA_NAV_CHECK(k >= 0 && k < n_daughters, "box slot out of range");
    if (k >= lds_lo && k < lds_hi) return lds_box + 8 * (k - lds_lo);
    return slot_box + 8 * (size_t)k;
  }
  VGA_NAV const B32 &Bound(int id) const {
    VGA_NAV_CHECK(id >= 0 && id < n_placed, "bound id out of range");
    return b32[id];
  }
  // The bound record of daughter slot k, whose placed id is id.
  VGA_NAV const B32 &SlotBound(int k, int id) const {
    if (lds_bound && k >= lds_lo && k < lds_hi) return lds_bound[k - lds_lo];
    return Bound(id);
  }
};

// Shape sets of specialised navigators (bit t: shape type t).
constexpr uint32_t kAllShapes = (1u << VG_BOX) | (1u << VG_TUBE) | (1u << VG_ORB) | (1u << VG_SPHERE) |
                                (1u << VG_CONE) | (1u << VG_TRD) | (1u << VG_TRAPEZOID) |
                                (1u << VG_PARALLELEPIPED) | (1u << VG_PARABOLOID) | (1u << VG_TORUS) |
                                (1u << VG_POLYCONE) | (1u << VG_POLYHEDRON);
// Box, Tube, Orb and Trd: the fused trace kernel needs 145 VGPRs with these
// alone (three waves per SIMD) against 230 with every shape (two).
constexpr uint32_t kBasicShapes = (1u << VG_BOX) | (1u << VG_TUBE) | (1u << VG_ORB) | (1u << VG_TRD);

// ------------------------------------------------------------ shape switch
// Shape switch over the compile-time shape set M (bit t set: shape type t
// may occur).  A navigator specialised for the shapes of one geometry
// carries only their code, which keeps its register count down.
#define VGA_SHAPE_CASES(TYPE, BODY)                   \
  switch (TYPE) {                                     \
    case VG_BOX:                                      \
      if constexpr ((M >> VG_BOX) & 1u) {             \
        using S = BoxS;                               \
        BODY;                                         \
      }                                               \
      break;                                          \
    case VG_TUBE:                                     \
      if constexpr ((M >> VG_TUBE) & 1u) {            \
        using S = TubeS;                              \
        BODY;                                         \
      }                                               \
      break;                                          \
    case VG_ORB:                                      \
      if constexpr ((M >> VG_ORB) & 1u) {             \
        using S = OrbS;                               \
        BODY;                                         \
      }                                               \
      break;                                          \
    case VG_TRD:                                      \
      if constexpr ((M >> VG_TRD) & 1u) {             \
        using S = TrdS;                               \
        BODY;                                         \
      }                                               \
      break;                                          \
    case VG_PARALLELEPIPED:                           \
      if constexpr ((M >> VG_PARALLELEPIPED) & 1u) {  \
        using S = ParaS;                              \
        BODY;                                         \
      }                                               \
      break;                                          \
    case VG_PARABOLOID:                               \
      if constexpr ((M >> VG_PARABOLOID) & 1u) {      \
        using S = ParabS;                             \
        BODY;                                         \
      }                                               \
      break;                                          \
    case VG_TRAPEZOID:                                \
      if constexpr ((M >> VG_TRAPEZOID) & 1u) {       \
        using S = TrapS;                              \
        BODY;                                         \
      }                                               \
      break;                                          \
    case VG_SPHERE:                                   \
      if constexpr ((M >> VG_SPHERE) & 1u) {          \
        using S = SphereS;                            \
        BODY;                                         \
      }                                               \
      break;                                          \
    case VG_CONE:                                     \
      if constexpr ((M >> VG_CONE) & 1u) {            \
        using S = ConeS;                              \
        BODY;                                         \
      }                                               \
      break;                                          \
    case VG_POLYCONE:                                 \
      if constexpr ((M >> VG_POLYCONE) & 1u) {        \
        using S = PolyconeS;                          \
        BODY;                                         \
      }                                               \
      break;                                          \
    case VG_TORUS:                                    \
      if constexpr ((M >> VG_TORUS) & 1u) {           \
        using S = TorusS;                             \
        BODY;                                         \
      }                                               \
      break;                                          \
    case VG_POLYHEDRON:                               \
      if constexpr ((M >> VG_POLYHEDRON) & 1u) {      \
        using S = PolyhedronS;                        \
        BODY;                                         \
      }                                               \
      break;                                          \
  }

#define VGA_P(blk) (ParamView<S>::Get(blk))

template <uint32_t M = kAllShapes>
VGA_NAV bool DContains(int type, const double *blk, V3 p) {
  VGA_SHAPE_CASES(type, return S::Contains(VGA_P(blk), p));
  return false;
}
template <uint32_t M = kAllShapes>
VGA_NAV int DInside(int type, const double *blk, V3 p) {
  VGA_SHAPE_CASES(type, return S::Inside(VGA_P(blk), p));
  return kOutside;
}
template <uint32_t M = kAllShapes>
VGA_NAV double DDistIn(int type, const double *blk, V3 p, V3 d, double sm) {
  VGA_SHAPE_CASES(type, return S::DistanceToIn(VGA_P(blk), p, d, sm));
  return kInfinity;
}
template <uint32_t M = kAllShapes>
VGA_NAV double DDistOut(int type, const double *blk, V3 p, V3 d, double sm) {
  VGA_SHAPE_CASES(type, return S::DistanceToOut(VGA_P(blk), p, d, sm));
  return kInfinity;
}
template <uint32_t M = kAllShapes>
VGA_NAV double DSafeIn(int type, const double *blk, V3 p) {
  VGA_SHAPE_CASES(type, return S::SafetyToIn(VGA_P(blk), p));
  return kInfinity;
}
template <uint32_t M = kAllShapes>
VGA_NAV double DSafeOut(int type, const double *blk, V3 p) {
  VGA_SHAPE_CASES(type, return S::SafetyToOut(VGA_P(blk), p));
  return kInfinity;
}

#undef VGA_P
#undef VGA_SHAPE_CASES

// ------------------------------------------------------------------ path
template <int D>
struct DPath {
  int v[D];
  int level;
  int boundary;
  VGA_NAV void Clear() {
#pragma unroll
    for (int k = 0; k < D; ++k) v[k] = -1;
    level = 0;
    boundary = 0;
  }
  VGA_NAV int Get(int l) const {
    int r = -1;
#pragma unroll
    for (int k = 0; k < D; ++k) r = (k == l) ? v[k] : r;
    return r;
  }
  VGA_NAV void Set(int l, int x) {
#pragma unroll
    for (int k = 0; k < D; ++k) v[k] = (k == l) ? x : v[k];
  }
  VGA_NAV int Top() const { return level > 0 ? Get(level - 1) : -1; }
  // The host checks that the geometry depth is at most D, so a push never
  // runs past the array (a path holds at most one volume per depth).
  VGA_NAV void Push(int x) {
    VGA_NAV_CHECK(level < D, "navigation path deeper than its register array");
    Set(level, x);
    ++level;
  }
  VGA_NAV void Pop() {
    if (level > 0) {
      --level;
      Set(level, -1);
    }
  }
};

// --------------------------------------------------------------- matrices
// Transformation3D::MultiplyFromRight (base/Transformation3D.h:662-708)
// with the right-hand placement's translation / rotation flags.
VGA_NAV void MultiplyFromRight(Xf &m, const Xf &rhs, int kind) {
  if (kind == kXfIdentity) return;
  if (kind == kXfTranslation || kind == kXfGeneral) {
    double t0 = m.t[0], t1 = m.t[1], t2 = m.t[2];
    t0 += m.r[0] * rhs.t[0];
    t0 += m.r[1] * rhs.t[1];
    t0 += m.r[2] * rhs.t[2];
    t1 += m.r[3] * rhs.t[0];
    t1 += m.r[4] * rhs.t[1];
    t1 += m.r[5] * rhs.t[2];
    t2 += m.r[6] * rhs.t[0];
    t2 += m.r[7] * rhs.t[1];
    t2 += m.r[8] * rhs.t[2];
    m.t[0] = t0;
    m.t[1] = t1;
    m.t[2] = t2;
  }
  if (kind == kXfRotation || kind == kXfGeneral) {
#pragma unroll
    for (int row = 0; row < 3; ++row) {
      const double tx = m.r[3 * row], ty = m.r[3 * row + 1], tz = m.r[3 * row + 2];
      double a = tx * rhs.r[0], b = tx * rhs.r[1], c = tx * rhs.r[2];
      a += ty * rhs.r[3];
      b += ty * rhs.r[4];
      c += ty * rhs.r[5];
      a += tz * rhs.r[6];
      b += tz * rhs.r[7];
      c += tz * rhs.r[8];
      m.r[3 * row] = a;
      m.r[3 * row + 1] = b;
      m.r[3 * row + 2] = c;
    }
  }
}

VGA_NAV void LoadXf(const PlacedRec &r, Xf &x) {
#pragma unroll
  for (int i = 0; i < 3; ++i) x.t[i] = r.xf.t[i];
#pragma unroll
  for (int i = 0; i < 9; ++i) x.r[i] = r.xf.r[i];
}

// The unspecialised Transform (all terms), used on TopMatrix and on the
// pushed point after a boundary.
VGA_NAV V3 ApplyGeneric(const Xf &m, V3 p) { return Rotate(m, V3{p.x - m.t[0], p.y - m.t[1], p.z - m.t[2]}); }

// InverseTransformKernel<false> (base/Transformation3D.h:601-632).
VGA_NAV V3 InversePoint(const Xf &X, V3 l) {
  double x = X.t[0], y = X.t[1], z = X.t[2];
  x += l.x * X.r[0];
  x += l.y * X.r[1];
  x += l.z * X.r[2];
  y += l.x * X.r[3];
  y += l.y * X.r[4];
  y += l.z * X.r[5];
  z += l.x * X.r[6];
  z += l.y * X.r[7];
  z += l.z * X.r[8];
  return {x, y, z};
}

template <int D>
VGA_NAV Xf TopMatrix(const NavDev &g, const DPath<D> &s) {
  Xf m;
#pragma unroll
  for (int i = 0; i < 3; ++i) m.t[i] = 0.;
#pragma unroll
  for (int i = 0; i < 9; ++i) m.r[i] = (i % 4 == 0) ? 1. : 0.;
  for (int l = 1; l < s.level; ++l) {
    const PlacedRec &r = g.Pv(s.Get(l));
    Xf x;
    LoadXf(r, x);
    MultiplyFromRight(m, x, r.xfkind);
  }
  return m;
}

// Contains of a placed volume on a point of its mother frame; the local
// point is returned through `local` (meaningful when the result is true).
// Mixed: the FP32 filter decides first (vg_bound32.hpp: a point farther
// from the surface than the bound's rounding margin is classified exactly
// as the FP64 Contains classifies it), and the FP64 Contains runs only for
// points it leaves undecided.
template <uint32_t M = kAllShapes>
VGA_NAV bool PlacedContains(const NavDev &g, int pv, V3 p, V3 &local, bool mixed = false) {
  int c = -1;
  if (mixed) {
    c = Classify32<M>(g.Bound(pv), g.b32side, F3{(float)p.x, (float)p.y, (float)p.z});
    if (c == kOutside) return false;
  }
  const PlacedRec &r = g.Pv(pv);
  Xf x;
  LoadXf(r, x);
  local = XfPointRt(x, r.xfkind, p);
  return c == kInside || DContains<M>(r.type, g.params + r.param_off, local);
}

template <uint32_t M = kAllShapes>
VGA_NAV double PlacedDistIn(const NavDev &g, int pv, V3 p, V3 d, double sm) {
  const PlacedRec &r = g.Pv(pv);
  Xf x;
  LoadXf(r, x);
  return DDistIn<M>(r.type, g.params + r.param_off, XfPointRt(x, r.xfkind, p), XfDirRt(x, r.xfkind, d), sm);
}

// ----------------------------------------------------------- locate family
template <int D, uint32_t M = kAllShapes>
VGA_NAV int LocatePoint(const NavDev &g, int vol, V3 point, DPath<D> &path, bool top, bool mixed = false) {
  int cand = vol;
  V3 tmp = point;
  if (top) {
    const PlacedRec &r = g.Pv(vol);
    cand = DContains<M>(r.type, g.params + r.param_off, point) ? vol : -1;
  }
  if (cand >= 0) {
    path.Push(cand);
    bool deeper = true;
    while (deeper) {
      const int db = g.Pv(cand).dbegin, de = g.Pv(cand).dend;
      if (de <= db) break;
      deeper = false;
      for (int k = db; k < de; ++k) {
        const int nx = g.Daughter(k);
        V3 tp;
        if (PlacedContains<M>(g, nx, tmp, tp, mixed)) {
          path.Push(nx);
          tmp = tp;
          cand = nx;
          deeper = true;
          break;
        }
      }
    }
  }
  return cand;
}

template <int D, uint32_t M = kAllShapes>
VGA_NAV int LocatePointExcl(const NavDev &g, int vol, int excluded, V3 point, DPath<D> &path, bool mixed = false) {
  int cand = vol;
  V3 tmp = point;
  path.Push(cand);
  bool deeper = true;
  while (deeper) {
    const int db = g.Pv(cand).dbegin, de = g.Pv(cand).dend;
    if (de <= db) break;
    deeper = false;
    for (int k = de - 1; k >= db; --k) {
      const int nx = g.Daughter(k);
      V3 tp;
      if (nx != excluded && PlacedContains<M>(g, nx, tmp, tp, mixed)) {
        path.Push(nx);
        tmp = tp;
        cand = nx;
        deeper = true;
        break;
      }
    }
  }
  return cand;
}

template <int D, uint32_t M = kAllShapes>
VGA_NAV int Relocate(const NavDev &g, V3 localpoint, DPath<D> &path, bool mixed = false) {
  int mother = path.Top();
  if (mother >= 0) {
    V3 tmp = localpoint;
    while (mother >= 0) {
      const PlacedRec &r = g.Pv(mother);
      if (DContains<M>(r.type, g.params + r.param_off, tmp)) break;
      path.Pop();
      Xf x;
      LoadXf(r, x);
      tmp = InversePoint(x, tmp);
      mother = path.Top();
    }
    if (mother >= 0) {
      path.Pop();
      return LocatePointExcl<D, M>(g, mother, mother, tmp, path, mixed);
    }
  }
  return mother;
}

// ------------------------------------------------------------ ABBox search
struct Ray {
  float p[3];
  float inv[3];
  int sign[3];
};

// IntersectCachedKernel2 on a float box, t0 = 0, t1 = +inf.  Max/Min are
// written as the reference's std::max / std::min (second operand when the
// first compares false), not fmaxf / fminf.
VGA_NAV float IntersectBox(const float *b, const Ray &r) {
  const float c0x = r.sign[0] ? b[3] : b[0], c1x = r.sign[0] ? b[0] : b[3];
  const float c0y = r.sign[1] ? b[4] : b[1], c1y = r.sign[1] ? b[1] : b[4];
  const float c0z = r.sign[2] ? b[5] : b[2], c1z = r.sign[2] ? b[2] : b[5];
  float tmin = (c0x - r.p[0]) * r.inv[0];
  const float tymax = (c1y - r.p[1]) * r.inv[1];
  if (tmin > tymax) return kInfF;
  float tmax = (c1x - r.p[0]) * r.inv[0];
  const float tymin = (c0y - r.p[1]) * r.inv[1];
  if (tymin > tmax) return kInfF;
  tmin = (tmin < tymin) ? tymin : tmin;
  tmax = (tymax < tmax) ? tymax : tmax;
  const float tzmin = (c0z - r.p[2]) * r.inv[2];
  const float tzmax = (c1z - r.p[2]) * r.inv[2];
  if ((tmin > tzmax) || (tzmin > tmax)) return kInfF;
  tmin = (tmin < tzmin) ? tzmin : tmin;
  tmax = (tzmax < tmax) ? tzmax : tmax;
  if (!((tmin < kInfF) && (tmax > 0.f))) tmin = kInfF;
  return tmin;
}

VGA_NAV Ray MakeRay(V3 lp, V3 ld) {
  Ray r;
  r.inv[0] = 1.f / (float)ld.x;
  r.inv[1] = 1.f / (float)ld.y;
  r.inv[2] = 1.f / (float)ld.z;
  r.p[0] = (float)lp.x;
  r.p[1] = (float)lp.y;
  r.p[2] = (float)lp.z;
#pragma unroll
  for (int i = 0; i < 3; ++i) r.sign[i] = r.inv[i] < 0;
  return r;
}

// The candidate walk of ABBoxNavigator::FindNextBoundaryAndStep
// (ABBoxNavigator.h:296-333) over daughters [db, de).  On entry `step`
// holds the mother's DistanceToOut (kInfinity if negative).  Returns 1 if
// the exact sorted-order walk was needed.  ExactOnly skips the order-free
// walk (used by the host-side checks of the exact walk).
//
// Mixed (VG_NAV_ABBOX_MIXED): before its FP64 Contains and DistanceToIn, a
// hit daughter is skipped when its FP32 lower bound on the distance from the
// point (vg_bound32.hpp) exceeds the best step found so far by more than
// the bound's rounding margin and 1e-6 mm.  Such a daughter cannot contain
// the point, and its DistanceToIn (the ray's distance to its surface, never
// shorter than the Euclidean distance by more than the tolerance) exceeds
// the final step, so it can neither win nor tie: the walk's result is
// unchanged.
//
// In the mixed form the walk runs in two phases per group of up to 64
// daughters: the box tests alone (a short loop whose loads do not depend on
// each other), recording hits in a bit mask, then the hit daughters in index
// order -- the order of the single loop of the FP64 form, which measured
// faster there.
template <bool ExactOnly = false, uint32_t M = kAllShapes, bool Mixed = false>
VGA_NAV int AbboxCandidates(const NavDev &g, int db, int de, V3 lp, V3 ld, double &step, int &next) {
  const Ray ray = MakeRay(lp, ld);
  const double step0 = step;
  double best = step0;
  int bk = -1, ties = 0;
  bool exact = ExactOnly;
  const F3 lp32{(float)lp.x, (float)lp.y, (float)lp.z};
  // One hit daughter (slot k, box entry b); false ends the walk (exact path).
  auto visit = [&](int k, float b) -> bool {
    const int cand = g.Daughter(k);
    if (Mixed) {
      float so, si, m;
      Bound32<M>(g.SlotBound(k, cand), g.b32side, lp32, so, si, m);
      if ((double)so - (double)m > Max(best, 0.) + 1e-6) return true;
    }
    const PlacedRec &r = g.Pv(cand);
    Xf x;
    LoadXf(r, x);
    const V3 tp = XfPointRt(x, r.xfkind, lp);
    const double *blk = g.params + r.param_off;
    if (b < 0.f && DContains<M>(r.type, blk, tp)) return false;
    const double dd = DDistIn<M>(r.type, blk, tp, XfDirRt(x, r.xfkind, ld), step0);
    if (dd < 0. || dd < (double)b) return false;
    if (dd < best) {
      best = dd;
      bk = k - db;
      ties = 1;
    } else if (dd == best) {
      ++ties;
    }
    return true;
  };
  if constexpr (!ExactOnly && Mixed) {
    for (int base = db; base < de && !exact; base += 64) {
      const int cnt = de - base < 64 ? de - base : 64;
      uint64_t hits = 0;
      for (int j = 0; j < cnt; ++j) hits |= (uint64_t)(IntersectBox(g.SlotBox(base + j), ray) < kInfF) << j;
      while (hits && !exact) {
        const int k = base + __builtin_ctzll(hits);
        hits &= hits - 1;
        exact = !visit(k, IntersectBox(g.SlotBox(k), ray));
      }
    }
  } else if constexpr (!ExactOnly) {
    for (int k = db; k < de && !exact; ++k) {
      const float b = IntersectBox(g.SlotBox(k), ray);
      if (b < kInfF) exact = !visit(k, b);
    }
  }
  if (!exact && ties <= 1) {
    step = best;
    next = bk;
    return 0;
  }
  // Exact walk: hits in ascending (entry distance, daughter index) order,
  // the order of a stable sort of the hit list.
  step = step0;
  next = -1;
  float lastb = -kInfF;
  int lastk = -1;
  while (true) {
    float bb = kInfF;
    int kk = -1;
    for (int k = db; k < de; ++k) {
      const float b = IntersectBox(g.SlotBox(k), ray);
      if (!(b < kInfF)) continue;
      const int kl = k - db;
      const bool after = (b > lastb) || (b == lastb && kl > lastk);
      if (after && (kk < 0 || b < bb)) {
        bb = b;
        kk = kl;
      }
    }
    if (kk < 0) break;
    lastb = bb;
    lastk = kk;
    const double hd = (double)bb;
    if (step < hd) break;
    const int cand = g.Daughter(db + kk);
    if (hd < 0) {
      V3 tp;
      if (PlacedContains<M>(g, cand, lp, tp)) {
        step = 0;
        next = kk;
        break;
      }
    }
    const double dd = PlacedDistIn<M>(g, cand, lp, ld, step);
    if (dd < step) {
      next = kk;
      step = dd;
    }
  }
  return 1;
}

// FindNextBoundaryAndStep of both navigators (kind 0 ABBox, 1 Simple).
// Returns 1 if the ABBox walk took the exact path.
// Mixed: the navigator of VG_NAV_ABBOX_MIXED (kind must then be that
// navigator): FP32 filters in the daughter walk and in the locate scans.
template <int D, bool ExactOnly = false, uint32_t M = kAllShapes, bool Mixed = false>
VGA_NAV int FindNextBoundaryAndStep(const NavDev &g, V3 gp, V3 gd, const DPath<D> &cur, DPath<D> &nst,
                                       double pstep, double &step, int kind) {
  const Xf m = TopMatrix(g, cur);
  const V3 lp = ApplyGeneric(m, gp);
  const V3 ld = Rotate(m, gd);
  const int cv = cur.Top();
  const PlacedRec &C = g.Pv(cv);
  const int db = C.dbegin, de = C.dend;
  int next = -1;
  int exact = 0;
  step = DDistOut<M>(C.type, g.params + C.param_off, lp, ld, pstep);
  nst = cur;
  constexpr bool mixed = Mixed;
  const bool abbox = kind == VG_NAV_ABBOX || kind == VG_NAV_ABBOX_MIXED;
  if (abbox) {
    if (step < 0.) step = kInfinity;
    if (de > db) exact = AbboxCandidates<ExactOnly, M, Mixed>(g, db, de, lp, ld, step, next);
    if (step > pstep) {
      step = pstep;
      nst.boundary = 0;
      return exact;
    }
    nst.boundary = 1;
    if (step > 1E30) {
      step = 0;
      nst.Pop();
      return exact;
    }
    if (step < 0.) step = 0.;
  } else {
    if (step <= 0.) step = kInfinity;
    for (int k = db; k < de; ++k) {
      const double dd = PlacedDistIn<M>(g, g.Daughter(k), lp, ld, step);
      const bool valid = dd < step && !__builtin_isinf(dd);
      next = valid ? k - db : next;
      step = valid ? dd : step;
    }
    if (step == kInfinity && pstep > 0.) {
      step = 0;
      nst.Pop();
      nst.boundary = 1;
      return 0;
    }
    if (step > pstep) {
      step = pstep;
      nst.boundary = 0;
      return 0;
    }
    nst.boundary = 1;
    if (step < 0.) step = 0.;
  }
  const double push = step + 1E-6;
  const V3 np{ld.x * push + lp.x, ld.y * push + lp.y, ld.z * push + lp.z};
  if (next != -1) {
    const int nv = g.Daughter(db + next);
    Xf x;
    LoadXf(g.Pv(nv), x);
    LocatePoint<D, M>(g, nv, ApplyGeneric(x, np), nst, false, mixed);
    return exact;
  }
  Relocate<D, M>(g, np, nst, mixed);
  if (abbox) {
    while (nst.Top() == cv) nst.Pop();
  }
  return exact;
}

// SimpleNavigator::FindNextBoundaryAndStep for baskets
// (navigation/SimpleNavigator.h:638-743), for one track of a basket whose
// current volume is `vol` (the top of the basket's first state, :672): the
// track's own TopMatrix (:660-667); the volume's DistanceToOut with pstep as
// step limit and no repair of values <= 0 (:678-679); DistanceToInMinimize
// over the daughters in order, each with the current distance as step
// limit, taken if strictly smaller (ShapeImplementationHelper.h:637-658);
// distance > pstep -> physics step off boundary (:708-713); otherwise push
// by 1e-9 and LocatePoint into the daughter or RelocatePointFromPath
// (:715-740).
template <int D, uint32_t M = kAllShapes>
VGA_NAV void FindNextBoundaryAndStepBasket(const NavDev &g, V3 gp, V3 gd, const DPath<D> &cur, DPath<D> &nst,
                                           double pstep, double &step, int vol) {
  const Xf m = TopMatrix(g, cur);
  const V3 lp = ApplyGeneric(m, gp);
  const V3 ld = Rotate(m, gd);
  const PlacedRec &C = g.Pv(vol);
  step = DDistOut<M>(C.type, g.params + C.param_off, lp, ld, pstep);
  int next = -1;
  for (int k = C.dbegin; k < C.dend; ++k) {
    const double dd = PlacedDistIn<M>(g, g.Daughter(k), lp, ld, step);
    if (dd < step) {
      step = dd;
      next = k - C.dbegin;
    }
  }
  nst = cur;
  if (step > pstep) {
    step = pstep;
    nst.boundary = 0;
    return;
  }
  nst.boundary = 1;
  const double push = step + 1E-9;
  const V3 np{ld.x * push + lp.x, ld.y * push + lp.y, ld.z * push + lp.z};
  if (next > -1) {
    const int nv = g.Daughter(C.dbegin + next);
    Xf x;
    LoadXf(g.Pv(nv), x);
    LocatePoint<D, M>(g, nv, ApplyGeneric(x, np), nst, false);
    return;
  }
  Relocate<D, M>(g, np, nst);
}

template <int D, uint32_t M = kAllShapes>
VGA_NAV double GetSafety(const NavDev &g, V3 gp, const DPath<D> &cur) {
  const Xf m = TopMatrix(g, cur);
  const V3 lp = ApplyGeneric(m, gp);
  const PlacedRec &C = g.Pv(cur.Top());
  double safety = DSafeOut<M>(C.type, g.params + C.param_off, lp);
  for (int k = C.dbegin; k < C.dend; ++k) {
    const PlacedRec &r = g.Pv(g.Daughter(k));
    Xf x;
    LoadXf(r, x);
    const double t = DSafeIn<M>(r.type, g.params + r.param_off, XfPointRt(x, r.xfkind, lp));
    safety = Min(safety, t);
  }
  return safety;
}

// The TraceTrack loop for one track (TraceTrack.cpp:170-220): locate from
// the world, then FindNextBoundaryAndStep and p += d * (step + 1e-6) until the
// state leaves the world or max_steps steps were taken.  Returns the number
// of steps; p ends at the final point.  The hash folds
// (new top id + 1) * 2 + boundary after every step.
template <int D, bool ExactOnly = false, uint32_t M = kAllShapes, bool Mixed = false>
VGA_NAV int TraceOne(const NavDev &g, V3 &p, V3 d, double pstep, int kind, int max_steps, double &dist, uint64_t &h,
                     int &nexact) {
  DPath<D> cur, nst;
  cur.Clear();
  LocatePoint<D, M>(g, g.world, p, cur, true, Mixed);
  int k = 0;
  nexact = 0;
  dist = 0;
  h = 1469598103934665603ULL;
  while (cur.level > 0 && k < max_steps) {
    double step;
    nexact += FindNextBoundaryAndStep<D, ExactOnly, M, Mixed>(g, p, d, cur, nst, pstep, step, kind);
    dist += step;
    const double s = step + 1E-6;
    p = V3{p.x + d.x * s, p.y + d.y * s, p.z + d.z * s};
    cur = nst;
    const uint64_t v = (uint64_t)(uint32_t)(cur.Top() + 1) * 2ULL + (cur.boundary ? 1ULL : 0ULL);
    h = (h ^ v) * 1099511628211ULL;
    ++k;
  }
  return k;
}

}  // namespace nav
}  // namespace vga

// Launches KERNEL<D, M, MX> with the path depth D (8 or 16) that holds the
// geometry, the shape set M (the basic shapes only when the geometry uses
// nothing else, all shapes otherwise) and MX = the mixed-precision
// navigator (MIXED, a run-time bool).
#define VGA_NAV_VARIANTS_MX(G, MIXED, KERNEL, ...)                                        \
  do {                                                                                    \
    const bool light_ = ((G)->type_mask & ~::vga::nav::kBasicShapes) == 0;                \
    const bool d8_ = (G)->max_depth <= 8;                                                 \
    if (MIXED) {                                                                          \
      if (d8_ && light_)                                                                  \
        hipLaunchKernelGGL((KERNEL<8, ::vga::nav::kBasicShapes, true>), __VA_ARGS__);     \
      else if (d8_)                                                                       \
        hipLaunchKernelGGL((KERNEL<8, ::vga::nav::kAllShapes, true>), __VA_ARGS__);       \
      else if (light_)                                                                    \
        hipLaunchKernelGGL((KERNEL<16, ::vga::nav::kBasicShapes, true>), __VA_ARGS__);    \
      else                                                                                \
        hipLaunchKernelGGL((KERNEL<16, ::vga::nav::kAllShapes, true>), __VA_ARGS__);      \
    } else {                                                                              \
      if (d8_ && light_)                                                                  \
        hipLaunchKernelGGL((KERNEL<8, ::vga::nav::kBasicShapes, false>), __VA_ARGS__);    \
      else if (d8_)                                                                       \
        hipLaunchKernelGGL((KERNEL<8, ::vga::nav::kAllShapes, false>), __VA_ARGS__);      \
      else if (light_)                                                                    \
        hipLaunchKernelGGL((KERNEL<16, ::vga::nav::kBasicShapes, false>), __VA_ARGS__);   \
      else                                                                                \
        hipLaunchKernelGGL((KERNEL<16, ::vga::nav::kAllShapes, false>), __VA_ARGS__);     \
    }                                                                                     \
  } while (0)

// Launches KERNEL<D, M> with the path depth D (8 or 16) that holds the
// geometry and the shape set M: the basic shapes only when the geometry uses
// nothing else, all shapes otherwise.
#define VGA_NAV_VARIANTS(G, KERNEL, ...)                                          \
  do {                                                                            \
    const bool light_ = ((G)->type_mask & ~::vga::nav::kBasicShapes) == 0;        \
    if ((G)->max_depth <= 8) {                                                    \
      if (light_)                                                                 \
        hipLaunchKernelGGL((KERNEL<8, ::vga::nav::kBasicShapes>), __VA_ARGS__);   \
      else                                                                        \
        hipLaunchKernelGGL((KERNEL<8, ::vga::nav::kAllShapes>), __VA_ARGS__);     \
    } else {                                                                      \
      if (light_)                                                                 \
        hipLaunchKernelGGL((KERNEL<16, ::vga::nav::kBasicShapes>), __VA_ARGS__);  \
      else                                                                        \
        hipLaunchKernelGGL((KERNEL<16, ::vga::nav::kAllShapes>), __VA_ARGS__);    \
    }                                                                             \
  } while (0)
