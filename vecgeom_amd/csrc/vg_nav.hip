// Device navigation for gfx950: LocatePoint, FindNextBoundaryAndStep of the
// ABBox and Simple navigators, GetSafety and the TraceTrack loop, one track
// per lane.
//
// Reference behaviour followed (file:line):
//   NavigationState::TopMatrix      navigation/NavigationState.h:398-408
//   SimpleNavigator::LocatePoint    navigation/SimpleNavigator.h:205-241
//   LocatePointExclVolume           navigation/SimpleNavigator.h:246-289
//   RelocatePointFromPath           navigation/SimpleNavigator.h:292-321
//   SimpleNavigator FNBAS           navigation/SimpleNavigator.h:341-482
//   ABBoxNavigator FNBAS            navigation/ABBoxNavigator.h:201-424
//   ABBoxNavigator::GetSafety       navigation/ABBoxNavigator.h:428-452
//   GetHitCandidates (float boxes)  source/ABBoxManager.cpp:80-126
//   IntersectCachedKernel2          volumes/kernel/BoxImplementation.h:334-384
//   track loop                      test/globalbenchmarks/TraceTrack.cpp:170-220
//
// Layout and execution:
//  * the geometry is the uploaded flat tree: PlacedRec (type, parameter
//    offset, daughter range, placement) per placed volume, the daughter
//    list, the shape parameter blocks, and float AABBs (6 per placed
//    volume, in the mother frame) -- a few hundred KB for thousands of
//    placements, resident in L2;
//  * a track's path lives in registers (template depth D, unrolled selects
//    instead of indexed private memory); the global-to-local matrix is
//    recomputed from the path every step, as TopMatrix does;
//  * the shape of a volume is dispatched at run time by a switch over the
//    type tag (the reference makes a virtual call);
//  * ABBox daughter loop.  The reference sorts the boxes hit by the ray by
//    entry distance and walks them in that order until a box starts beyond
//    the current step.  A lane cannot hold a sorted list of up to
//    hundreds of hits, so it first walks the daughters in index order,
//    computing DistanceToIn for every hit box with stepMax = the initial
//    step.  The result equals the sorted walk when (a) no hit box with a
//    negative entry distance contains the point, (b) every DistanceToIn
//    result is >= 0 and >= its box entry distance (boxes bound their
//    shapes, so a box skipped by the sorted walk cannot hold a nearer
//    surface), and (c) the smallest distance is reached by one daughter
//    only.  The lane checks (a)-(c) and otherwise redoes the step in exact
//    sorted order by repeated selection of the next (distance, index) pair.
//    Every device shape's DistanceToIn returns the same value for any
//    stepMax larger than that value (Box and Parallelepiped use stepMax
//    only to return kInfinity early; the others ignore it), which is what
//    makes the larger stepMax of the order-free walk safe.
#include <hip/hip_runtime.h>

#include <cstring>

#include "vg_nav_impl.hpp"

namespace vga {

namespace {

using namespace nav;

constexpr int kNavBlock = 128;

// ------------------------------------------------------------------ kernels
template <int D>
VGA_NAV void LoadState(const vg_nav_states &s, uint64_t i, DPath<D> &p) {
  p.level = s.level[i];
  p.boundary = s.boundary[i];
#pragma unroll
  for (int l = 0; l < D; ++l) p.v[l] = (l < s.depth && l < p.level) ? s.path[(uint64_t)l * s.capacity + i] : -1;
}

// A caller-supplied state is navigated only if its level fits the path
// array and every id on its path is a placed volume of this geometry.
template <int D>
VGA_NAV bool ValidState(const NavDev &g, const DPath<D> &p) {
  if (p.level < 1 || p.level > D) return false;
  bool ok = true;
#pragma unroll
  for (int l = 0; l < D; ++l) ok = ok && (l >= p.level || (p.v[l] >= 0 && p.v[l] < g.n_placed));
  return ok;
}

template <int D>
VGA_NAV void StoreState(const vg_nav_states &s, uint64_t i, const DPath<D> &p) {
  s.level[i] = p.level;
  s.boundary[i] = (uint8_t)p.boundary;
  for (int l = 0; l < s.depth; ++l) s.path[(uint64_t)l * s.capacity + i] = l < D ? p.Get(l) : -1;
}

template <int D, uint32_t M>
__global__ __launch_bounds__(kNavBlock) void LocateKernel(const NavDev g, const double *__restrict__ x,
                                                          const double *__restrict__ y, const double *__restrict__ z,
                                                          uint64_t n, const vg_nav_states s) {
  const uint64_t i = (uint64_t)blockIdx.x * kNavBlock + threadIdx.x;
  if (i >= n) return;
  DPath<D> p;
  p.Clear();
  LocatePoint<D, M>(g, g.world, V3{x[i], y[i], z[i]}, p, true);
  StoreState(s, i, p);
}

struct StepArgs {
  const double *x, *y, *z, *dx, *dy, *dz, *pstep;
  double *step;
  uint64_t n;
  int32_t kind;
  int32_t pad;
};

template <int D, uint32_t M, bool MX>
__global__ __launch_bounds__(kNavBlock) void StepKernel(const NavDev g, const StepArgs a, const vg_nav_states cur,
                                                        const vg_nav_states nxt) {
  const uint64_t i = (uint64_t)blockIdx.x * kNavBlock + threadIdx.x;
  if (i >= a.n) return;
  DPath<D> c, o;
  LoadState(cur, i, c);
  double step = 0;
  const double ps = a.pstep ? a.pstep[i] : kInfinity;
  if (a.kind == VG_NAV_SIMPLE_BASKET) {
    // The basket's current volume: the top of the first track's state.
    DPath<D> c0;
    LoadState(cur, 0, c0);
    if (ValidState(g, c0) && ValidState(g, c)) {
      FindNextBoundaryAndStepBasket<D, M>(g, V3{a.x[i], a.y[i], a.z[i]}, V3{a.dx[i], a.dy[i], a.dz[i]}, c, o, ps,
                                          step, c0.Top());
    } else {
      o = c;
      if (c.level != 0) step = __builtin_nan("");
    }
  } else if (ValidState(g, c)) {
    FindNextBoundaryAndStep<D, false, M, MX>(g, V3{a.x[i], a.y[i], a.z[i]}, V3{a.dx[i], a.dy[i], a.dz[i]}, c, o, ps,
                                         step, a.kind);
  } else {
    // A state outside the world has no volume to step in and stays outside
    // (step 0); a malformed state is left as it is, with step NaN.
    o = c;
    if (c.level != 0) step = __builtin_nan("");
  }
  StoreState(nxt, i, o);
  a.step[i] = step;
}

template <int D, uint32_t M>
__global__ __launch_bounds__(kNavBlock) void SafetyKernel(const NavDev g, const double *__restrict__ x,
                                                          const double *__restrict__ y, const double *__restrict__ z,
                                                          uint64_t n, const vg_nav_states s, double *__restrict__ out) {
  const uint64_t i = (uint64_t)blockIdx.x * kNavBlock + threadIdx.x;
  if (i >= n) return;
  DPath<D> c;
  LoadState(s, i, c);
  out[i] = ValidState(g, c) ? GetSafety<D, M>(g, V3{x[i], y[i], z[i]}, c) : (c.level == 0 ? 0. : __builtin_nan(""));
}

struct TraceArgs {
  const double *x, *y, *z, *dx, *dy, *dz;
  int32_t *nsteps;
  double *fx, *fy, *fz, *travelled;
  uint64_t *hash;
  int32_t *exact;
  uint64_t n;
  double pstep;
  int32_t kind;
  int32_t max_steps;
  int32_t stage_lo, stage_hi;  // daughter slots copied to LDS first (none when equal)
  int32_t stage_bounds;        // 1: their FP32 bound records too (mixed kernel)
  int32_t pad;
};

// The world volume's daughter-slot boxes (8 floats each) are staged in LDS
// by the trace kernel when they take at most this many bytes: 20 resident
// workgroups per CU (five waves per SIMD, two waves a workgroup) then still
// fit in 160 KiB.
constexpr size_t kNavBoxLds = 8192;
// ... and, in the mixed kernel, their FP32 bound records (144 bytes each)
// when those take at most this many bytes more.
constexpr size_t kNavBoundLds = 9216;

// Occupancy floor of the fused trace kernel over the basic shape set: it is
// latency-bound, and 5 waves per SIMD (instead of the 3 its 145 VGPRs allow,
// at the price of spilled registers) measured 0.91 -> 1.01 G steps/s on the
// nested geometry (4 waves: 1.00, 6: 0.89).  The per-step and serving
// kernels measured no better with a floor.
constexpr int TraceWaves(uint32_t m) { return m == kBasicShapes ? 5 : 1; }

template <int D, uint32_t M, bool MX>
__global__ __launch_bounds__(kNavBlock) __attribute__((amdgpu_waves_per_eu(TraceWaves(M)))) void TraceKernel(
    const NavDev g, const TraceArgs a) {
  extern __shared__ __align__(16) float nav_boxes[];
  NavDev gl = g;
  if (a.stage_hi > a.stage_lo) {  // uniform over the grid
    const int nf = 8 * (a.stage_hi - a.stage_lo);
    const float *src = g.slot_box + 8 * (size_t)a.stage_lo;
    for (int j = threadIdx.x; j < nf; j += kNavBlock) nav_boxes[j] = src[j];
    __syncthreads();
    gl.lds_box = nav_boxes;
    gl.lds_lo = a.stage_lo;
    gl.lds_hi = a.stage_hi;
    if constexpr (MX) {
      if (a.stage_bounds) {  // uniform over the grid
        // Records follow the boxes (8 floats a slot keep them 16-byte aligned).
        constexpr int kRec = (int)(sizeof(B32) / sizeof(uint32_t));
        uint32_t *dst = reinterpret_cast<uint32_t *>(nav_boxes + nf);
        const int nr = kRec * (a.stage_hi - a.stage_lo);
        for (int j = threadIdx.x; j < nr; j += kNavBlock) {
          const int slot = a.stage_lo + j / kRec;
          dst[j] = reinterpret_cast<const uint32_t *>(g.b32 + g.daughters[slot])[j % kRec];
        }
        __syncthreads();
        gl.lds_bound = reinterpret_cast<const B32 *>(dst);
      }
    }
  }
  const uint64_t i = (uint64_t)blockIdx.x * kNavBlock + threadIdx.x;
  if (i >= a.n) return;
  V3 p{a.x[i], a.y[i], a.z[i]};
  const V3 d{a.dx[i], a.dy[i], a.dz[i]};
  double dist;
  uint64_t h;
  int nexact;
  const int k = TraceOne<D, false, M, MX>(gl, p, d, a.pstep, a.kind, a.max_steps, dist, h, nexact);
  a.nsteps[i] = k;
  a.fx[i] = p.x;
  a.fy[i] = p.y;
  a.fz[i] = p.z;
  a.travelled[i] = dist;
  a.hash[i] = h;
  if (a.exact) a.exact[i] = nexact;
}

NavDev Dev(const vg_geometry *g) {
  NavDev d;
  d.placed = g->d_placed;
  d.params = g->d_params;
  d.daughters = g->d_daughters;
  d.aabb = g->d_aabb;
  d.b32 = g->d_b32;
  d.b32side = g->d_b32side;
  d.slot_box = g->d_slot_aabb;
  d.world = g->world;
  d.n_placed = (int32_t)g->placed.size();
  d.n_daughters = (int32_t)g->daughters.size();
  d.pad = 0;
  return d;
}

unsigned Blocks(uint64_t n) { return (unsigned)((n + kNavBlock - 1) / kNavBlock); }

}  // namespace

int NavLocate(const vg_geometry *g, const double *x, const double *y, const double *z, uint64_t n,
              const vg_nav_states &s, hipStream_t st) {
  if (n == 0) return VG_OK;
  VGA_NAV_VARIANTS(g, LocateKernel, dim3(Blocks(n)), dim3(kNavBlock), 0, st, Dev(g), x, y, z, n, s);
  return CheckHip(hipGetLastError(), "locate launch");
}

int NavStep(const vg_geometry *g, int kind, const double *x, const double *y, const double *z, const double *dx,
            const double *dy, const double *dz, const double *pstep, uint64_t n, const vg_nav_states &cur,
            const vg_nav_states &nxt, double *step, hipStream_t st) {
  if (n == 0) return VG_OK;
  StepArgs a{x, y, z, dx, dy, dz, pstep, step, n, kind, 0};
  VGA_NAV_VARIANTS_MX(g, kind == VG_NAV_ABBOX_MIXED, StepKernel, dim3(Blocks(n)), dim3(kNavBlock), 0, st, Dev(g), a,
                      cur, nxt);
  return CheckHip(hipGetLastError(), "step launch");
}

int NavSafety(const vg_geometry *g, const double *x, const double *y, const double *z, uint64_t n,
              const vg_nav_states &s, double *out, hipStream_t st) {
  if (n == 0) return VG_OK;
  VGA_NAV_VARIANTS(g, SafetyKernel, dim3(Blocks(n)), dim3(kNavBlock), 0, st, Dev(g), x, y, z, n, s, out);
  return CheckHip(hipGetLastError(), "safety launch");
}

int NavTrace(const vg_geometry *g, int kind, const double *x, const double *y, const double *z, const double *dx,
             const double *dy, const double *dz, uint64_t n, double pstep, int max_steps, int32_t *nsteps, double *fx,
             double *fy, double *fz, double *travelled, uint64_t *hash, int32_t *exact, hipStream_t st) {
  if (n == 0) return VG_OK;
  TraceArgs a{x, y, z, dx, dy, dz, nsteps, fx, fy, fz, travelled, hash, exact, n, pstep, kind, max_steps, 0, 0};
  const PlacedRec *w = g->world >= 0 ? &g->placed[g->world] : nullptr;
  const size_t box_bytes = w ? 8 * sizeof(float) * (size_t)(w->dend - w->dbegin) : 0;
  size_t lds = 0;
  if (box_bytes > 0 && box_bytes <= kNavBoxLds) {
    a.stage_lo = w->dbegin;
    a.stage_hi = w->dend;
    lds = box_bytes;
    const size_t bound_bytes = sizeof(B32) * (size_t)(w->dend - w->dbegin);
    if (kind == VG_NAV_ABBOX_MIXED && g->d_b32 && bound_bytes <= kNavBoundLds) {
      a.stage_bounds = 1;
      lds += bound_bytes;
    }
  }
  VGA_NAV_VARIANTS_MX(g, kind == VG_NAV_ABBOX_MIXED, TraceKernel, dim3(Blocks(n)), dim3(kNavBlock), lds, st, Dev(g),
                      a);
  return CheckHip(hipGetLastError(), "trace launch");
}

}  // namespace vga
