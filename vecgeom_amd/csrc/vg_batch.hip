// Batched multi-shape query launch for gfx950.
//
// A batch is a table of tasks, each one basket query on one placed volume
// (the per-shape work queues: any mix of shapes, queries and basket sizes).
// The tasks are cut into tiles of kBatchBlock * IT items; one launch covers
// every tile of every task, so a step of many small and large baskets costs
// one dispatch instead of one per (shape, query) and no launch leaves the
// chip idle at its tail while the next waits.  Each tile belongs to a single
// task, so every lane of a workgroup evaluates the same shape and query: the
// task record, the shape parameters and the placement are uniform per tile
// (scalar loads), and the switch over shape and query is a uniform branch.
//
// Shapes are split into weight classes by the registers their kernels need
// (light: Box, Tube, Orb, Trd, Trapezoid, Parallelepiped, Paraboloid;
// medium: Cone, Sphere; heavy: Torus; variable-length parameter blocks:
// Polycone, Polyhedron -- staged in LDS per workgroup when the block fits,
// read from global memory otherwise).  Each class has its own kernel, so a
// Box tile is not held to the register budget of the Torus quartic; a batch
// costs one launch per class it uses, and the class kernels run
// concurrently.
//
// Launch shapes: IT = 1, 2 or 4 items per lane (at most 2 for the heavy
// classes); the light class runs one workgroup per tile or a persistent grid
// of `blocks` workgroups striding over the tiles, the others one workgroup
// per tile.  vg_batch_autotune picks IT and the grid by timing the batch.
//
// Inputs may be double (the FP64 form) or float (FP32 storage: each value is
// promoted to double in registers, so results equal the FP64 kernels on the
// same float-valued inputs; distances keep FP64 arithmetic, which the
// 1e-9 mm bar requires -- one float ulp at 50 mm is 4e-6 mm).  Mixed tasks
// (float inputs) run an FP32 filter first (vg_bound32.hpp): Contains and
// Inside are decided in FP32 where the point lies farther from the surface
// than the bound's rigorous rounding margin, and DistanceToIn returns
// kInfinity for rays that provably miss the shape's extent box; every other
// lane runs the FP64 code on the promoted inputs, so results equal the FP32
// storage form's bit for bit.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "shapes.hpp"
#include "vg_stage.hpp"
#include "vg_walk.hpp"
#include "vg_internal.hpp"

namespace vga {

namespace {

constexpr int kBatchBlock = 256;

struct TaskDev {
  int32_t type, query, xfkind, in_type;
  int32_t param_off, vec_ok;
  int32_t param_len, pad;  // parameter block length (doubles)
  uint64_t n, tile_begin;
  const void *x, *y, *z, *dx, *dy, *dz;
  const double *sm;
  void *out;
  uint64_t *count;      // FP64-evaluated queries (nullable)
  const B32 *b32;       // the placed volume's FP32 bound record
  const float *b32side; // its variable-length data
  const float *ext;     // its local extent box, lo xyz, hi xyz (float)
  Xf xf;
};

// Adds each lane's n to *count with one atomic per wave; callable with any
// set of active lanes (the sum is taken bit by bit over ballots).
__device__ __forceinline__ void WaveCount(uint64_t *count, uint32_t n) {
  uint64_t total = 0;
  for (int b = 0; b < 32; ++b) {
    total += (uint64_t)__popcll(__ballot((n >> b) & 1u)) << b;
    if (__ballot((n >> b) > 1u) == 0) break;
  }
  const uint64_t act = __ballot(1);
  if (total && (uint32_t)__lane_id() == (uint32_t)(__ffsll((long long)act) - 1)) atomicAdd((unsigned long long *)count, total);
}

// The shape kind of a shape kernel type (bit of the FP32 bound masks).
template <class S> struct KindOf;
template <> struct KindOf<BoxS> { static constexpr int v = VG_BOX; };
template <> struct KindOf<TubeS> { static constexpr int v = VG_TUBE; };
template <> struct KindOf<OrbS> { static constexpr int v = VG_ORB; };
template <> struct KindOf<TrdS> { static constexpr int v = VG_TRD; };
template <> struct KindOf<TrapS> { static constexpr int v = VG_TRAPEZOID; };
template <> struct KindOf<ParaS> { static constexpr int v = VG_PARALLELEPIPED; };
template <> struct KindOf<ParabS> { static constexpr int v = VG_PARABOLOID; };
template <> struct KindOf<SphereS> { static constexpr int v = VG_SPHERE; };
template <> struct KindOf<ConeS> { static constexpr int v = VG_CONE; };
template <> struct KindOf<TorusS> { static constexpr int v = VG_TORUS; };
template <> struct KindOf<PolyconeS> { static constexpr int v = VG_POLYCONE; };
template <> struct KindOf<PolyhedronS> { static constexpr int v = VG_POLYHEDRON; };

// FP32 filter of a mixed task of shape S for query Q at float point p
// (direction d): true with the result in r when decided.
template <class S, int Q>
__device__ __forceinline__ bool Filter32(const TaskDev &t, F3 p, F3 d, double &r) {
  if constexpr (Q == kQDistIn) {
    if (!RayMissesBox32(*t.b32, t.ext, p, d)) return false;
    if constexpr (KindOf<S>::v == VG_ORB) {  // the Orb's miss value depends on the ray's sense
      const int sg = RadialSign32(*t.b32, p, d);
      if (sg == 0) return false;
      r = sg < 0 ? __builtin_nan("") : kInfinity;
      return true;
    }
    r = kInfinity;
    return true;
  } else {
    const int c = Classify32<1u << KindOf<S>::v>(*t.b32, t.b32side, p);
    if (c < 0) return false;
    r = Q == kQContains ? (c == kInside ? 1. : 0.) : (double)c;
    return true;
  }
}

template <class T, int IT>
struct alignas(sizeof(T) * IT) Vec {
  T v[IT];
};

template <class T, int IT>
__device__ __forceinline__ void Load(const void *base, uint64_t i0, bool full, bool vec, uint64_t n, double (&o)[IT]) {
  const T *p = static_cast<const T *>(base) + i0;
  if (full && vec) {
    const Vec<T, IT> t = *reinterpret_cast<const Vec<T, IT> *>(p);
#pragma unroll
    for (int k = 0; k < IT; ++k) o[k] = (double)t.v[k];
  } else {
#pragma unroll
    for (int k = 0; k < IT; ++k) o[k] = (i0 + k < n) ? (double)p[k] : 0.;
  }
}

template <int IT>
__device__ __forceinline__ void LoadAll(const TaskDev t, uint64_t i0, bool full, const void *a, double (&o)[IT]) {
  const bool vec = t.vec_ok != 0;
  if (t.in_type != VG_F64)
    Load<float, IT>(a, i0, full, vec, t.n, o);
  else
    Load<double, IT>(a, i0, full, vec, t.n, o);
}

__device__ __forceinline__ double LoadOne(const TaskDev &t, const void *a, uint64_t i) {
  return t.in_type != VG_F64 ? (double)static_cast<const float *>(a)[i] : static_cast<const double *>(a)[i];
}

// The FP64 evaluation of query Q for one item at promoted point p (and
// direction d, step max sm); writes out[i].
template <class S, int Q>
__device__ __forceinline__ void Eval64(const TaskDev &t, const typename S::P &P, uint64_t i, V3 p, V3 d, double sm) {
  if constexpr (Q == kQContains) {
    static_cast<uint8_t *>(t.out)[i] = S::Contains(P, XfPointRt(t.xf, t.xfkind, p)) ? 1 : 0;
  } else if constexpr (Q == kQInside) {
    static_cast<int32_t *>(t.out)[i] = S::Inside(P, XfPointRt(t.xf, t.xfkind, p));
  } else if constexpr (Q == kQDistIn) {
    static_cast<double *>(t.out)[i] = S::DistanceToIn(P, XfPointRt(t.xf, t.xfkind, p), XfDirRt(t.xf, t.xfkind, d), sm);
  } else if constexpr (Q == kQDistOut) {
    static_cast<double *>(t.out)[i] = S::DistanceToOut(P, p, d, sm);
  } else if constexpr (Q == kQPlacedDistOut) {
    static_cast<double *>(t.out)[i] = S::DistanceToOut(P, XfPointRt(t.xf, t.xfkind, p), XfDirRt(t.xf, t.xfkind, d), sm);
  } else if constexpr (Q == kQSafeIn) {
    static_cast<double *>(t.out)[i] = S::SafetyToIn(P, XfPointRt(t.xf, t.xfkind, p));
  } else {
    static_cast<double *>(t.out)[i] = S::SafetyToOut(P, p);
  }
}

// Writes a result decided by the FP32 filter.
template <int Q>
__device__ __forceinline__ void WriteDecided(const TaskDev &t, uint64_t i, double r) {
  if constexpr (Q == kQContains)
    static_cast<uint8_t *>(t.out)[i] = (uint8_t)r;
  else if constexpr (Q == kQInside)
    static_cast<int32_t *>(t.out)[i] = (int32_t)r;
  else
    static_cast<double *>(t.out)[i] = r;
}

// blk: the task's parameter block (in global memory or staged in LDS).  MX:
// a mixed-precision task (FP32 filter first, per lane).
template <class S, int Q, int IT, bool MX>
__device__ __forceinline__ void RunTile(const TaskDev t, const double *blk, uint64_t i0) {
  const auto &P = ParamView<S>::Get(blk);
  const bool full = i0 + IT <= t.n;
  constexpr bool dir = (Q == kQDistIn || Q == kQDistOut || Q == kQPlacedDistOut);
  double px[IT], py[IT], pz[IT], dx[IT], dy[IT], dz[IT], sm[IT];
  LoadAll<IT>(t, i0, full, t.x, px);
  LoadAll<IT>(t, i0, full, t.y, py);
  LoadAll<IT>(t, i0, full, t.z, pz);
  if constexpr (dir) {
    LoadAll<IT>(t, i0, full, t.dx, dx);
    LoadAll<IT>(t, i0, full, t.dy, dy);
    LoadAll<IT>(t, i0, full, t.dz, dz);
#pragma unroll
    for (int k = 0; k < IT; ++k) sm[k] = (t.sm && i0 + k < t.n) ? t.sm[i0 + k] : kInfinity;
  }
  constexpr bool filt = MX && (Q == kQContains || Q == kQInside || Q == kQDistIn);
  uint32_t n64 = 0;
#pragma unroll
  for (int k = 0; k < IT; ++k) {
    if (i0 + k >= t.n) break;
    const V3 p{px[k], py[k], pz[k]};
    const uint64_t i = i0 + k;
    if constexpr (filt) {
      double r;
      const F3 d32 = dir ? F3{(float)dx[k], (float)dy[k], (float)dz[k]} : F3{0.f, 0.f, 0.f};
      if (Filter32<S, Q>(t, F3{(float)p.x, (float)p.y, (float)p.z}, d32, r)) {
        WriteDecided<Q>(t, i, r);
        continue;
      }
    }
    ++n64;
    Eval64<S, Q>(t, P, i, p, V3{dx[k], dy[k], dz[k]}, sm[k]);
  }
  if (t.count) WaveCount(t.count, n64);
}

// Mixed-precision tile of the compute-heavy classes with lane compaction:
// every lane runs the FP32 filter on its items and writes the decided ones;
// the undecided items are listed in LDS and then evaluated in FP64 by the
// first lanes of the workgroup, so the FP64 code runs on dense waves
// instead of on the few lanes of each wave that need it.  The inputs of a
// listed item are read again (from cache) in the second pass.  Called by
// every thread of the workgroup.
template <class S, int Q, int IT>
__device__ __forceinline__ void RunTileCompact(const TaskDev t, const double *blk, uint64_t tile_first) {
  constexpr int kItems = kBatchBlock * IT;
  constexpr bool dir = (Q == kQDistIn);
  __shared__ int count;
  __shared__ int item[kItems];
  if (threadIdx.x == 0) count = 0;
  __syncthreads();
  const uint64_t i0 = tile_first + (uint64_t)threadIdx.x * IT;
  const bool full = i0 + IT <= t.n;
  uint32_t n64 = 0;
  if (i0 < t.n) {
    double px[IT], py[IT], pz[IT], dx[IT], dy[IT], dz[IT];
    LoadAll<IT>(t, i0, full, t.x, px);
    LoadAll<IT>(t, i0, full, t.y, py);
    LoadAll<IT>(t, i0, full, t.z, pz);
    if constexpr (dir) {
      LoadAll<IT>(t, i0, full, t.dx, dx);
      LoadAll<IT>(t, i0, full, t.dy, dy);
      LoadAll<IT>(t, i0, full, t.dz, dz);
    }
#pragma unroll
    for (int k = 0; k < IT; ++k) {
      const uint64_t i = i0 + k;
      if (i >= t.n) break;
      double r;
      const F3 d32 = dir ? F3{(float)dx[k], (float)dy[k], (float)dz[k]} : F3{0.f, 0.f, 0.f};
      if (Filter32<S, Q>(t, F3{(float)px[k], (float)py[k], (float)pz[k]}, d32, r)) {
        WriteDecided<Q>(t, i, r);
      } else {
        ++n64;
        item[atomicAdd(&count, 1)] = (int)(i - tile_first);
      }
    }
  }
  if (t.count) WaveCount(t.count, n64);
  __syncthreads();
  const auto &P = ParamView<S>::Get(blk);
  const int c = count;
  for (int j = threadIdx.x; j < c; j += kBatchBlock) {
    const uint64_t i = tile_first + item[j];
    const V3 p{LoadOne(t, t.x, i), LoadOne(t, t.y, i), LoadOne(t, t.z, i)};
    V3 d{0., 0., 0.};
    if constexpr (dir) d = V3{LoadOne(t, t.dx, i), LoadOne(t, t.dy, i), LoadOne(t, t.dz, i)};
    Eval64<S, Q>(t, P, i, p, d, (dir && t.sm) ? t.sm[i] : kInfinity);
  }
  __syncthreads();  // the list is reused by the next tile of a persistent loop
}

template <class S, int IT, bool MX>
__device__ __forceinline__ void RunShape(const TaskDev t, const double *blk, uint64_t i0) {
  switch (t.query) {
    case kQContains: RunTile<S, kQContains, IT, MX>(t, blk, i0); break;
    case kQInside: RunTile<S, kQInside, IT, MX>(t, blk, i0); break;
    case kQDistIn: RunTile<S, kQDistIn, IT, MX>(t, blk, i0); break;
    case kQDistOut: RunTile<S, kQDistOut, IT, MX>(t, blk, i0); break;
    case kQPlacedDistOut: RunTile<S, kQPlacedDistOut, IT, MX>(t, blk, i0); break;
    case kQSafeIn: RunTile<S, kQSafeIn, IT, MX>(t, blk, i0); break;
    case kQSafeOut: RunTile<S, kQSafeOut, IT, MX>(t, blk, i0); break;
  }
}

// The compacted mixed-precision tile of shape S for the filtered queries;
// false for the others (run per lane).  Uniform over the workgroup.
template <class S, int IT>
__device__ __forceinline__ bool RunShapeCompact(const TaskDev t, const double *blk, uint64_t tile_first) {
  switch (t.query) {
    case kQContains: RunTileCompact<S, kQContains, IT>(t, blk, tile_first); return true;
    case kQInside: RunTileCompact<S, kQInside, IT>(t, blk, tile_first); return true;
    case kQDistIn: RunTileCompact<S, kQDistIn, IT>(t, blk, tile_first); return true;
  }
  return false;
}

// Classes 3 and 4 hold the shapes with variable-length parameter blocks:
// class 3 those whose block fits in LDS (staged per workgroup), class 4 the
// others (read from global memory).
//
// Mixed-precision tasks form classes of their own (base class +
// kBaseClasses): their kernels carry the FP32 filter, and the FP64 kernels
// do not pay its registers.
constexpr int kBaseClasses = 5;
constexpr int kClasses = 2 * kBaseClasses;
constexpr int kStagedClass = 3;
constexpr int kMaxStageDoubles = 6144;  // 48 KiB of LDS
constexpr int Base(int c) { return c % kBaseClasses; }
constexpr bool MixedClass(int c) { return c >= kBaseClasses; }

// Occupancy floor (waves per SIMD) of the staged class's kernel: latency
// bound, it runs best at 6 waves (a few spilled registers) rather than at
// the 4 its 121 VGPRs allow; measured 88 -> 100 G queries/s on the batched
// zoo step (7 waves: 99, 8: 93).  The Torus class measured best unbounded.
constexpr int ClassWaves(int c) { return Base(c) == kStagedClass ? 6 : 1; }

// Queries whose code is short for every shape: a Torus task of these runs in
// the medium class, away from the register budget (and low occupancy) of the
// Torus class's quartic.
constexpr bool CheapQuery(int q) { return q == kQContains || q == kQInside || q == kQSafeIn || q == kQSafeOut; }

constexpr int ClassOf(int type) {
  return (type == VG_CONE || type == VG_SPHERE)           ? 1
         : (type == VG_TORUS)                              ? 2
         : (type == VG_POLYCONE || type == VG_POLYHEDRON) ? 3
                                                           : 0;
}

// Items per lane of class c when the batch asks for `it`: the light class
// takes it as given, the compute-heavy classes at most 2.
constexpr int ClassIT(int c, int it) { return Base(c) == 0 ? it : (it < 2 ? it : 2); }

// The Polyhedron's distance tasks in the staged class run with lane refill
// (vg_walk.hpp) over tiles kWalkTileMul times longer, so that each wave
// refills its lanes from 4 * 64 * IT queries.  (The Polycone's measured
// slower as walks, DistanceToIn and DistanceToOut alike.)
constexpr int kWalkTileMul = 4;
constexpr bool WalkTask(int type, int query) {
  return type == VG_POLYHEDRON && (query == kQDistIn || query == kQDistOut || query == kQPlacedDistOut);
}
constexpr uint64_t TileItems(int c, int it, int type, int query) {
  return (uint64_t)kBatchBlock * ClassIT(c, it) * ((Base(c) == kStagedClass && WalkTask(type, query)) ? kWalkTileMul : 1);
}

// The four short queries of a shape (CheapQuery), FP64 or FP32 storage.
template <class S, int IT>
__device__ __forceinline__ void RunShapeCheap(const TaskDev t, const double *blk, uint64_t i0) {
  switch (t.query) {
    case kQContains: RunTile<S, kQContains, IT, false>(t, blk, i0); break;
    case kQInside: RunTile<S, kQInside, IT, false>(t, blk, i0); break;
    case kQSafeIn: RunTile<S, kQSafeIn, IT, false>(t, blk, i0); break;
    case kQSafeOut: RunTile<S, kQSafeOut, IT, false>(t, blk, i0); break;
  }
}

template <int IT, int CLS>
__device__ __forceinline__ void RunTask(const TaskDev t, const double *blk, uint64_t i0) {
  constexpr int B = Base(CLS);
  constexpr bool MX = MixedClass(CLS);
  if constexpr (B == 0) {
    switch (t.type) {
      case VG_BOX: RunShape<BoxS, IT, MX>(t, blk, i0); break;
      case VG_TUBE: RunShape<TubeS, IT, MX>(t, blk, i0); break;
      case VG_ORB: RunShape<OrbS, IT, MX>(t, blk, i0); break;
      case VG_TRD: RunShape<TrdS, IT, MX>(t, blk, i0); break;
      case VG_TRAPEZOID: RunShape<TrapS, IT, MX>(t, blk, i0); break;
      case VG_PARALLELEPIPED: RunShape<ParaS, IT, MX>(t, blk, i0); break;
      case VG_PARABOLOID: RunShape<ParabS, IT, MX>(t, blk, i0); break;
    }
  } else if constexpr (B == 1) {
    switch (t.type) {
      case VG_SPHERE: RunShape<SphereS, IT, MX>(t, blk, i0); break;
      case VG_CONE: RunShape<ConeS, IT, MX>(t, blk, i0); break;
      case VG_TORUS:
        if constexpr (!MX) RunShapeCheap<TorusS, IT>(t, blk, i0);
        break;
    }
  } else if constexpr (B == 2) {
    if (t.type == VG_TORUS) RunShape<TorusS, IT, MX>(t, blk, i0);
  } else {  // classes 3 and 4
    switch (t.type) {
      case VG_POLYCONE: RunShape<PolyconeS, IT, MX>(t, blk, i0); break;
      case VG_POLYHEDRON: RunShape<PolyhedronS, IT, MX>(t, blk, i0); break;
    }
  }
}

// The compacted mixed-precision tile of a compute-heavy class (not the
// light one); false when the task's query runs per lane.
template <int IT, int CLS>
__device__ __forceinline__ bool RunTaskCompact(const TaskDev t, const double *blk, uint64_t tile_first) {
  constexpr int B = Base(CLS);
  if constexpr (B == 0) {
    switch (t.type) {
      case VG_BOX: return RunShapeCompact<BoxS, IT>(t, blk, tile_first);
      case VG_TUBE: return RunShapeCompact<TubeS, IT>(t, blk, tile_first);
      case VG_ORB: return RunShapeCompact<OrbS, IT>(t, blk, tile_first);
      case VG_TRD: return RunShapeCompact<TrdS, IT>(t, blk, tile_first);
      case VG_TRAPEZOID: return RunShapeCompact<TrapS, IT>(t, blk, tile_first);
      case VG_PARALLELEPIPED: return RunShapeCompact<ParaS, IT>(t, blk, tile_first);
      case VG_PARABOLOID: return RunShapeCompact<ParabS, IT>(t, blk, tile_first);
    }
    return false;
  } else if constexpr (B == 1) {
    if (t.type == VG_SPHERE) return RunShapeCompact<SphereS, IT>(t, blk, tile_first);
    return RunShapeCompact<ConeS, IT>(t, blk, tile_first);
  } else if constexpr (B == 2) {
    return RunShapeCompact<TorusS, IT>(t, blk, tile_first);
  } else if constexpr (B >= 3) {
    if (t.type == VG_POLYCONE) return RunShapeCompact<PolyconeS, IT>(t, blk, tile_first);
    return RunShapeCompact<PolyhedronS, IT>(t, blk, tile_first);
  }
  return false;
}

// DistanceToIn of one tile with lane compaction (the Torus, whose exact
// bounding-tube rejection discards most rays): every lane tests its items,
// writes kInfinity for the rejected ones and appends the others to a
// workgroup list in LDS; then the list is processed by the first lanes, so
// the quartic runs on dense waves.  Called by every thread of the workgroup.
template <class S, int IT, bool MX>
__device__ __forceinline__ void RunTileCompactDistIn(const TaskDev t, const double *blk, uint64_t tile_first) {
  constexpr int kItems = kBatchBlock * IT;
  __shared__ int count;
  __shared__ int item[kItems];
  __shared__ double qp[3][kItems], qd[3][kItems], qsm[kItems];
  const auto &P = ParamView<S>::Get(blk);
  if (threadIdx.x == 0) count = 0;
  __syncthreads();
  const uint64_t i0 = tile_first + (uint64_t)threadIdx.x * IT;
  const bool full = i0 + IT <= t.n;
  double px[IT], py[IT], pz[IT], dx[IT], dy[IT], dz[IT];
  uint32_t n64 = 0;
  if (i0 < t.n) {
    LoadAll<IT>(t, i0, full, t.x, px);
    LoadAll<IT>(t, i0, full, t.y, py);
    LoadAll<IT>(t, i0, full, t.z, pz);
    LoadAll<IT>(t, i0, full, t.dx, dx);
    LoadAll<IT>(t, i0, full, t.dy, dy);
    LoadAll<IT>(t, i0, full, t.dz, dz);
#pragma unroll
    for (int k = 0; k < IT; ++k) {
      const uint64_t i = i0 + k;
      if (i >= t.n) break;
      double r;
      if (MX && Filter32<S, kQDistIn>(t, F3{(float)px[k], (float)py[k], (float)pz[k]},
                                   F3{(float)dx[k], (float)dy[k], (float)dz[k]}, r)) {
        static_cast<double *>(t.out)[i] = r;
        continue;
      }
      ++n64;
      const V3 p = XfPointRt(t.xf, t.xfkind, V3{px[k], py[k], pz[k]});
      const V3 d = XfDirRt(t.xf, t.xfkind, V3{dx[k], dy[k], dz[k]});
      const double sm = t.sm ? t.sm[i] : kInfinity;
      if (S::DistanceToInQuick(P, p, d, sm, r)) {
        static_cast<double *>(t.out)[i] = r;
      } else {
        const int j = atomicAdd(&count, 1);
        item[j] = (int)(i - tile_first);
        qp[0][j] = p.x;
        qp[1][j] = p.y;
        qp[2][j] = p.z;
        qd[0][j] = d.x;
        qd[1][j] = d.y;
        qd[2][j] = d.z;
        qsm[j] = sm;
      }
    }
  }
  if (t.count) WaveCount(t.count, n64);
  __syncthreads();
  const int c = count;
  for (int j = threadIdx.x; j < c; j += kBatchBlock) {
    static_cast<double *>(t.out)[tile_first + item[j]] =
        S::DistanceToIn(P, V3{qp[0][j], qp[1][j], qp[2][j]}, V3{qd[0][j], qd[1][j], qd[2][j]}, qsm[j]);
  }
}


// Items [first, first + count) of a walk task with lane refill: each of the
// workgroup's waves takes a quarter of the range.  Local says the inputs are
// in the shape's frame already (DistanceToOut).  Called by every thread of
// the workgroup.
template <class S, class Ops, bool Local, bool MX>
__device__ __forceinline__ void RunTileWalk(const TaskDev &t, const double *blk, uint64_t first, uint64_t count) {
  constexpr int kWaves = kBatchBlock / 64;
  const uint64_t per = (count + kWaves - 1) / kWaves;
  const uint64_t lim = first + count < t.n ? first + count : t.n;
  const uint64_t b = first + (uint64_t)(threadIdx.x / 64) * per;
  const uint64_t begin = b < lim ? b : lim;
  const uint64_t end = begin + per < lim ? begin + per : lim;
  double *out = static_cast<double *>(t.out);
  const bool filt = MX && !Local && t.query == kQDistIn;
  uint32_t n64 = 0;
  WalkRefill<Ops>(
      ParamView<S>::Get(blk), begin, end,
      [&](uint64_t i, V3 &p, V3 &d, double &sm, double &r) {
        p = V3{LoadOne(t, t.x, i), LoadOne(t, t.y, i), LoadOne(t, t.z, i)};
        d = V3{LoadOne(t, t.dx, i), LoadOne(t, t.dy, i), LoadOne(t, t.dz, i)};
        if (filt && Filter32<S, kQDistIn>(t, F3{(float)p.x, (float)p.y, (float)p.z}, F3{(float)d.x, (float)d.y, (float)d.z}, r))
          return true;
        ++n64;
        if (!Local) {
          p = XfPointRt(t.xf, t.xfkind, p);
          d = XfDirRt(t.xf, t.xfkind, d);
        }
        sm = t.sm ? t.sm[i] : kInfinity;
        return false;
      },
      [&](uint64_t i, double r) { out[i] = r; });
  if (t.count) WaveCount(t.count, n64);
}

// Copies a task's parameter block to LDS (whole workgroup).
__device__ __forceinline__ void Stage(double *stage, const double *src, int len) {
  StageBlock(stage, src, len, kBatchBlock);
}

// Tile -> task: the last task whose first tile is <= tile (tasks sorted by
// tile_begin; every task has at least one tile).
__device__ __forceinline__ int FindTask(const TaskDev *tasks, int ntasks, uint64_t tile) {
  int lo = 0, hi = ntasks - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (tasks[mid].tile_begin <= tile)
      lo = mid;
    else
      hi = mid - 1;
  }
  return lo;
}

// The light class runs on a persistent grid if asked (blocks striding over
// the tiles).  The compute-heavy classes take one tile per workgroup: a
// persistent loop around their bodies keeps the task's uniform values live
// across iterations and more than doubles their VGPR count.
template <int IT, int CLS>
__global__ __launch_bounds__(kBatchBlock) __attribute__((amdgpu_waves_per_eu(ClassWaves(CLS)))) void BatchKernel(const TaskDev *__restrict__ tasks, int ntasks,
                                                           uint64_t ntiles, const double *__restrict__ params) {
  constexpr int B = Base(CLS);
  constexpr bool MX = MixedClass(CLS);
  if constexpr (B == 0) {
    for (uint64_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
      const int k = FindTask(tasks, ntasks, tile);
      const TaskDev t = tasks[k];
      if constexpr (MX) {
        if (RunTaskCompact<IT, CLS>(t, params + t.param_off, (tile - t.tile_begin) * kBatchBlock * IT)) continue;
      }
      const uint64_t i0 = ((tile - t.tile_begin) * kBatchBlock + threadIdx.x) * IT;
      if (i0 < t.n) RunTask<IT, CLS>(t, params + t.param_off, i0);
    }
  } else {
    const uint64_t tile = blockIdx.x;
    if (tile >= ntiles) return;
    const int k = FindTask(tasks, ntasks, tile);
    const TaskDev t = tasks[k];
    const uint64_t tile_first = (tile - t.tile_begin) * kBatchBlock * IT;
    const uint64_t i0 = tile_first + threadIdx.x * IT;
    const double *blk = params + t.param_off;
    if constexpr (B == kStagedClass) {
      extern __shared__ __align__(16) double stage[];
      Stage(stage, params + t.param_off, t.param_len);
      __syncthreads();
      blk = stage;
      if (WalkTask(t.type, t.query)) {  // uniform over the workgroup
        const uint64_t per = TileItems(CLS, IT, t.type, t.query), first = (tile - t.tile_begin) * per;
        if (t.query == kQDistIn) {
          RunTileWalk<PolyhedronS, PolyhedronS::DistInOps, false, MX>(t, stage, first, per);
        } else if (t.query == kQDistOut) {
          RunTileWalk<PolyhedronS, PolyhedronS::DistOutOps, true, MX>(t, stage, first, per);
        } else {
          RunTileWalk<PolyhedronS, PolyhedronS::DistOutOps, false, MX>(t, stage, first, per);
        }
        return;
      }
    }
    if constexpr (B == 2) {
      if (t.query == kQDistIn) {  // uniform over the workgroup
        RunTileCompactDistIn<TorusS, IT, MX>(t, blk, tile_first);
        return;
      }
    }
    if constexpr (MX) {
      if (RunTaskCompact<IT, CLS>(t, blk, tile_first)) return;
    }
    if (i0 < t.n) RunTask<IT, CLS>(t, blk, i0);
  }
}

// One task passed by value (kernel argument): the single-query form of the
// same code, for callers that do not keep a batch.
template <int IT, int CLS>
__global__ __launch_bounds__(kBatchBlock) __attribute__((amdgpu_waves_per_eu(ClassWaves(CLS)))) void TaskKernel(const TaskDev t, const double *__restrict__ params) {
  const uint64_t i0 = ((uint64_t)blockIdx.x * kBatchBlock + threadIdx.x) * IT;
  if constexpr (Base(CLS) == kStagedClass) {
    extern __shared__ __align__(16) double stage[];
    if (t.param_len <= kMaxStageDoubles) {
      Stage(stage, params + t.param_off, t.param_len);
      __syncthreads();
      if (i0 < t.n) RunTask<IT, CLS>(t, stage, i0);
      return;
    }
  }
  if (i0 < t.n) RunTask<IT, CLS>(t, params + t.param_off, i0);
}

}  // namespace

}  // namespace vga

using namespace vga;

struct vg_batch {
  const vg_geometry *g = nullptr;
  std::vector<vg_query_task> tasks;
  std::vector<TaskDev> host;                  // sorted by weight class
  int first[vga::kClasses + 1] = {};   // class c owns host[first[c] .. first[c+1])
  uint64_t ntiles[vga::kClasses] = {};  // tiles per class for the current IT
  int stage_len = 0;                               // LDS doubles for the staged class
  // One device task table per items-per-lane value (1, 2, 4), each written
  // once at create: changing the launch shape only selects another table,
  // so a hipGraph captured earlier keeps a table that still matches its
  // own grid and kernel.
  TaskDev *d_table[3] = {nullptr, nullptr, nullptr};
  TaskDev *d_tasks = nullptr;  // the table of the current IT
  int it = 2;            // items per lane
  unsigned blocks = 0;   // 0 = one workgroup per tile, else a persistent grid per class
  uint64_t items = 0;
  // Classes run concurrently: the first non-empty class on the caller's
  // stream, the others on side streams forked from and joined back to it
  // with events (a pattern hipGraph capture records as parallel branches).
  hipStream_t side[vga::kClasses - 1] = {};
  hipEvent_t fork = nullptr;
  hipEvent_t join[vga::kClasses - 1] = {};
  // Per-class timing (vg_batch_set_timing): a ring of event pairs recorded
  // on the stream each class kernel is launched on, around that kernel.
  static constexpr int kRing = 64;
  bool timing = false;
  int ring_used = 0, ring_next = 0;
  hipEvent_t t0[vga::kClasses][kRing] = {};
  hipEvent_t t1[vga::kClasses][kRing] = {};
  bool recorded[vga::kClasses][kRing] = {};
  uint64_t class_items[vga::kClasses] = {};
  uint64_t class_bytes[vga::kClasses] = {};
};

namespace {

constexpr int kITs[3] = {1, 2, 4};

bool Aligned(const void *p, size_t a) { return (reinterpret_cast<uintptr_t>(p) % a) == 0; }

int QueryCode(int32_t q) {
  switch (q) {
    case VG_Q_CONTAINS: return kQContains;
    case VG_Q_INSIDE: return kQInside;
    case VG_Q_DISTANCE_TO_IN: return kQDistIn;
    case VG_Q_DISTANCE_TO_OUT: return kQDistOut;
    case VG_Q_PLACED_DISTANCE_TO_OUT: return kQPlacedDistOut;
    case VG_Q_SAFETY_TO_IN: return kQSafeIn;
    case VG_Q_SAFETY_TO_OUT: return kQSafeOut;
  }
  return -1;
}

// Relative cost of a query kind (distances walk and intersect, safeties
// walk, classifications only test).
int QueryWeight(int q) {
  return q == kQDistIn ? 3 : (q == kQContains || q == kQInside) ? 1 : 2;
}

// Mixed-precision policy (vg_set_mixed_policy).  On gfx950 an FP64 FMA
// issues at the rate of an unpacked FP32 one, so the FP32 filter pays only
// where a decided lane skips long FP64 work; measured per (shape, query) on
// the mixed zoo workload (filter time / FP32-storage time, single-task
// batches of 2^21 queries): Polycone DistanceToIn 0.94; every other pair
// 0.99-3.9 (classifications 1.4-3.9: the FP64 Contains / Inside of these
// shapes is a handful of comparisons, cheaper than the FP32 bound and its
// margin).  Under the measured policy a mixed task without a paying filter
// runs in the FP64 class kernels as FP32 storage (identical results).
int g_mixed_policy = VG_MIXED_POLICY_MEASURED;

bool FilterPays(int type, int query) { return type == VG_POLYCONE && query == kQDistIn; }

bool Filtered(const TaskDev &t) {
  if (t.in_type != VG_MIXED) return false;
  if (t.query != kQContains && t.query != kQInside && t.query != kQDistIn) return false;
  return g_mixed_policy == VG_MIXED_POLICY_ALL || FilterPays(t.type, t.query);
}

int TaskClass(const TaskDev &t) {
  int c = ClassOf(t.type);
  if (t.type == VG_TORUS && CheapQuery(t.query) && !Filtered(t)) c = 1;
  if (c == kStagedClass && t.param_len > kMaxStageDoubles) c = kStagedClass + 1;
  return c + (Filtered(t) ? kBaseClasses : 0);
}

// Fills tile_begin (numbered within each class) for items per lane `it`.
void Schedule(vg_batch *b) {
  for (int c = 0; c < kClasses; ++c) {
    uint64_t tile = 0;
    for (int k = b->first[c]; k < b->first[c + 1]; ++k) {
      TaskDev &x = b->host[k];
      const uint64_t per = TileItems(c, b->it, x.type, x.query);
      x.tile_begin = tile;
      tile += (x.n + per - 1) / per;
    }
    b->ntiles[c] = tile;
  }
}

template <int IT, int CLS>
void LaunchClass(const vg_batch *b, unsigned grid, hipStream_t s) {
  const size_t lds = Base(CLS) == kStagedClass ? (size_t)b->stage_len * sizeof(double) : 0;
  hipLaunchKernelGGL((BatchKernel<IT, CLS>), dim3(grid), dim3(kBatchBlock), lds, s, b->d_tasks + b->first[CLS],
                     b->first[CLS + 1] - b->first[CLS], b->ntiles[CLS], b->g->d_params);
}

template <int CLS>
void LaunchIT(const vg_batch *b, unsigned grid, hipStream_t s) {
  const int it = ClassIT(CLS, b->it);
  if (it == 1)
    LaunchClass<1, CLS>(b, grid, s);
  else if (it == 2)
    LaunchClass<2, CLS>(b, grid, s);
  else if constexpr (ClassIT(CLS, 4) == 4)
    LaunchClass<4, CLS>(b, grid, s);
}

// Algorithmic bytes of one task: every input element read once, every
// output element written once.
uint64_t TaskBytes(const TaskDev &t) {
  const uint64_t w = t.in_type == VG_F64 ? 8 : 4;
  const bool dir = t.query == kQDistIn || t.query == kQDistOut || t.query == kQPlacedDistOut;
  uint64_t per = 3 * w;
  if (dir) per += 3 * w + (t.sm ? 8 : 0);
  per += t.query == kQContains ? 1 : (t.query == kQInside ? 4 : 8);
  return per * t.n;
}

int Launch(vg_batch *b, hipStream_t s) {
  const int slot = b->ring_next;
  if (b->timing) {
    b->ring_next = (b->ring_next + 1) % vg_batch::kRing;
    if (b->ring_used < vg_batch::kRing) ++b->ring_used;
  }
  int classes = 0;
  for (int c = 0; c < kClasses; ++c) classes += b->ntiles[c] ? 1 : 0;
  int rc = VG_OK;
  if (classes > 1) rc = CheckHip(hipEventRecord(b->fork, s), "batch fork");
  int nside = 0;
  bool first = true;
  // (Launching the heaviest classes first measured slower on the zoo step.)
  for (int c = 0; c < kClasses && !rc; ++c) {
    const uint64_t nt = b->ntiles[c];
    if (nt == 0) continue;
    hipStream_t cs = s;
    if (!first) {
      // later classes: a side stream that waits for the fork
      cs = b->side[nside];
      rc = CheckHip(hipStreamWaitEvent(cs, b->fork, 0), "batch fork wait");
      if (rc) break;
    }
    first = false;
    const unsigned grid = (Base(c) == 0 && b->blocks) ? (unsigned)std::min<uint64_t>(b->blocks, nt) : (unsigned)nt;
    if (b->timing) {
      rc = CheckHip(hipEventRecord(b->t0[c][slot], cs), "batch timing");
      if (rc) break;
    }
    switch (c) {
      case 0: LaunchIT<0>(b, grid, cs); break;
      case 1: LaunchIT<1>(b, grid, cs); break;
      case 2: LaunchIT<2>(b, grid, cs); break;
      case 3: LaunchIT<3>(b, grid, cs); break;
      case 4: LaunchIT<4>(b, grid, cs); break;
      case 5: LaunchIT<5>(b, grid, cs); break;
      case 6: LaunchIT<6>(b, grid, cs); break;
      case 7: LaunchIT<7>(b, grid, cs); break;
      case 8: LaunchIT<8>(b, grid, cs); break;
      default: LaunchIT<9>(b, grid, cs); break;
    }
    rc = CheckHip(hipGetLastError(), "batch launch");
    if (!rc && b->timing) {
      rc = CheckHip(hipEventRecord(b->t1[c][slot], cs), "batch timing");
      b->recorded[c][slot] = true;
    }
    if (!rc && cs != s) {
      rc = CheckHip(hipEventRecord(b->join[nside], cs), "batch join");
      if (!rc) rc = CheckHip(hipStreamWaitEvent(s, b->join[nside], 0), "batch join wait");
      ++nside;
    }
  }
  return rc;
}

void DestroyTiming(vg_batch *b) {
  for (int c = 0; c < kClasses; ++c)
    for (int k = 0; k < vg_batch::kRing; ++k) {
      if (b->t0[c][k]) (void)hipEventDestroy(b->t0[c][k]);
      if (b->t1[c][k]) (void)hipEventDestroy(b->t1[c][k]);
      b->t0[c][k] = b->t1[c][k] = nullptr;
    }
}

void DestroySync(vg_batch *b) {
  DestroyTiming(b);
  for (auto &x : b->side)
    if (x) (void)hipStreamDestroy(x);
  for (auto &x : b->join)
    if (x) (void)hipEventDestroy(x);
  if (b->fork) (void)hipEventDestroy(b->fork);
}

int CreateSync(vg_batch *b) {
  int rc = CheckHip(hipEventCreateWithFlags(&b->fork, hipEventDisableTiming), "batch event");
  for (int k = 0; k < kClasses - 1 && !rc; ++k) {
    rc = CheckHip(hipStreamCreateWithFlags(&b->side[k], hipStreamNonBlocking), "batch stream");
    if (!rc) rc = CheckHip(hipEventCreateWithFlags(&b->join[k], hipEventDisableTiming), "batch event");
  }
  return rc;
}

int ITIndex(int it) { return it == 1 ? 0 : (it == 2 ? 1 : 2); }

// Writes the device task table of every IT (at create; synchronous), then
// selects the current one.
int UploadAll(vg_batch *b) {
  const int keep = b->it;
  int rc = VG_OK;
  for (int k = 0; k < 3 && !rc; ++k) {
    b->it = kITs[k];
    Schedule(b);
    rc = CheckHip(hipMalloc(&b->d_table[k], b->host.size() * sizeof(TaskDev)), "hipMalloc(batch table)");
    if (!rc)
      rc = CheckHip(hipMemcpy(b->d_table[k], b->host.data(), b->host.size() * sizeof(TaskDev), hipMemcpyHostToDevice),
                    "batch table upload");
  }
  b->it = keep;
  Schedule(b);
  b->d_tasks = b->d_table[ITIndex(b->it)];
  return rc;
}

// Selects the tables and tile counts of the current IT (host only: no
// device memory is rewritten, so launches in flight and captured graphs are
// unaffected).
void Select(vg_batch *b) {
  Schedule(b);
  b->d_tasks = b->d_table[ITIndex(b->it)];
}

void FreeTables(vg_batch *b) {
  for (auto &t : b->d_table)
    if (t) (void)hipFree(t);
}

// Validates one task and fills its device record; returns a status.
int MakeTask(const vg_geometry *g, const vg_query_task &q, int i, TaskDev &t) {
  if (q.placed < 0 || (size_t)q.placed >= g->placed.size())
    return SetError(VG_ERR_INVALID, "task " + std::to_string(i) + ": bad placed id");
  const PlacedRec &pr = g->placed[q.placed];
  if (!ShapeHasKernels(pr.type))
    return SetError(VG_ERR_UNSUPPORTED, "task " + std::to_string(i) + ": no device kernels for the shape");
  if (pr.type == VG_BOOLEAN)
    return SetError(VG_ERR_UNSUPPORTED, "task " + std::to_string(i) +
                                            ": Boolean volumes run through the basket entry points, not batches");
  const int qc = QueryCode(q.query);
  if (qc < 0 || (q.input_type != VG_F64 && q.input_type != VG_F32 && q.input_type != VG_MIXED))
    return SetError(VG_ERR_INVALID, "task " + std::to_string(i) + ": bad query or input type");
  const bool dir = qc == kQDistIn || qc == kQDistOut || qc == kQPlacedDistOut;
  if (q.n > 0 && (!q.x || !q.y || !q.z || !q.out || (dir && (!q.dx || !q.dy || !q.dz))))
    return SetError(VG_ERR_INVALID, "task " + std::to_string(i) + ": null array");
  std::memset(&t, 0, sizeof(t));
  t.type = pr.type;
  t.query = qc;
  t.xfkind = pr.xfkind;
  t.in_type = q.input_type;
  t.param_off = pr.param_off;
  const double *blk = g->params.data() + pr.param_off;
  t.param_len = pr.type == VG_POLYCONE     ? ParamView<PolyconeS>::Length(blk)
                : pr.type == VG_POLYHEDRON ? ParamView<PolyhedronS>::Length(blk)
                                           : 0;
  t.n = q.n;
  t.x = q.x;
  t.y = q.y;
  t.z = q.z;
  t.dx = q.dx;
  t.dy = q.dy;
  t.dz = q.dz;
  t.sm = q.step_max;
  t.out = q.out;
  t.count = q.fp64_count;
  t.b32 = g->d_b32 + q.placed;
  t.b32side = g->d_b32side;
  t.ext = g->d_extent32 + 6 * (size_t)q.placed;
  t.xf = pr.xf;
  // Vector loads need every input array aligned to a whole vector of the
  // largest IT (4 items).
  const size_t a = (t.in_type == VG_F64 ? 8 : 4) * 4;
  bool ok = Aligned(t.x, a) && Aligned(t.y, a) && Aligned(t.z, a);
  if (dir) ok = ok && Aligned(t.dx, a) && Aligned(t.dy, a) && Aligned(t.dz, a);
  t.vec_ok = ok ? 1 : 0;
  return VG_OK;
}

template <int CLS>
void LaunchTaskIT(int it, dim3 gr, size_t lds, hipStream_t s, const TaskDev &t, const double *params) {
  const size_t l = Base(CLS) == kStagedClass ? lds : 0;
  if (it == 2)
    hipLaunchKernelGGL((TaskKernel<2, CLS>), gr, dim3(kBatchBlock), l, s, t, params);
  else
    hipLaunchKernelGGL((TaskKernel<1, CLS>), gr, dim3(kBatchBlock), l, s, t, params);
}

void LaunchTask(int cls, int it, dim3 gr, size_t lds, hipStream_t s, const TaskDev &t, const double *params) {
  switch (cls) {
    case 0: LaunchTaskIT<0>(it, gr, lds, s, t, params); break;
    case 1: LaunchTaskIT<1>(it, gr, lds, s, t, params); break;
    case 2: LaunchTaskIT<2>(it, gr, lds, s, t, params); break;
    case 3: LaunchTaskIT<3>(it, gr, lds, s, t, params); break;
    case 5: LaunchTaskIT<5>(it, gr, lds, s, t, params); break;
    case 6: LaunchTaskIT<6>(it, gr, lds, s, t, params); break;
    case 7: LaunchTaskIT<7>(it, gr, lds, s, t, params); break;
    default: LaunchTaskIT<8>(it, gr, lds, s, t, params); break;
  }
}

// ------------------------------------------------ the no-allocation batch
// vg_batch_run_noalloc keeps its task table in this static device array
// instead of memory allocated per batch.  The array's name is a host-side
// stand-in; host code uses only the address hipGetSymbolAddress returns for
// it on each device.  Every call on a device shares that device's array.
constexpr int kScratchTasks = 256;
static __device__ __align__(256) TaskDev scratch[kScratchTasks];

constexpr int kMaxDevices = 64;
TaskDev *g_scratch_ptr[kMaxDevices] = {};
std::mutex g_scratch_mutex;

// The device address of `scratch` on the current device, looked up once per
// device.
int ScratchPtr(TaskDev **out) {
  int dev = -1;
  int rc = CheckHip(hipGetDevice(&dev), "hipGetDevice");
  if (rc) return rc;
  if (dev < 0 || dev >= kMaxDevices) return SetError(VG_ERR_STATE, "device index beyond the scratch table");
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (!g_scratch_ptr[dev]) {
    void *p = nullptr;
    rc = CheckHip(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
    if (rc) return rc;
    g_scratch_ptr[dev] = static_cast<TaskDev *>(p);
  }
  *out = g_scratch_ptr[dev];
  return VG_OK;
}

template <int CLS>
void LaunchNoAlloc(const TaskDev *tasks, int ntasks, uint64_t ntiles, int stage_len, const double *params,
                   hipStream_t s) {
  constexpr int IT = ClassIT(CLS, 2);
  const size_t lds = Base(CLS) == kStagedClass ? (size_t)stage_len * sizeof(double) : 0;
  hipLaunchKernelGGL((BatchKernel<IT, CLS>), dim3((unsigned)ntiles), dim3(kBatchBlock), lds, s, tasks, ntasks, ntiles,
                     params);
}

}  // namespace

extern "C" {

int vg_batch_run_noalloc(const vg_geometry *g, const vg_query_task *tasks, int32_t ntasks, void *stream) {
  if (!g || ntasks < 0 || (ntasks > 0 && !tasks)) return SetError(VG_ERR_INVALID, "null argument");
  if (ntasks > kScratchTasks)
    return SetError(VG_ERR_CAPACITY, "vg_batch_run_noalloc takes at most " + std::to_string(kScratchTasks) +
                                         " tasks (" + std::to_string(ntasks) + " given)");
  int rc = EnsureDevice();
  if (rc) return rc;
  TaskDev *dev = nullptr;
  if ((rc = ScratchPtr(&dev))) return rc;
  // Host staging of the table: a fixed per-thread array (no allocation).
  static thread_local TaskDev host[kScratchTasks];
  int n = 0;
  for (int i = 0; i < ntasks; ++i) {
    TaskDev t;
    if ((rc = MakeTask(g, tasks[i], i, t))) return rc;
    if (t.n > 0) host[n++] = t;
  }
  if (n == 0) return VG_OK;
  std::stable_sort(host, host + n, [](const TaskDev &a, const TaskDev &c) { return TaskClass(a) < TaskClass(c); });
  int first[kClasses + 1];
  uint64_t ntiles[kClasses];
  int stage_len = 0;
  for (int c = 0, k = 0; c < kClasses; ++c) {
    first[c] = k;
    uint64_t tile = 0;
    for (; k < n && TaskClass(host[k]) == c; ++k) {
      const uint64_t per = TileItems(c, 2, host[k].type, host[k].query);
      host[k].tile_begin = tile;
      tile += (host[k].n + per - 1) / per;
      if (Base(c) == kStagedClass) stage_len = std::max(stage_len, host[k].param_len);
    }
    ntiles[c] = tile;
  }
  first[kClasses] = n;
  const hipStream_t s = (hipStream_t)stream;
  // A pageable-memory copy returns once the host array has been staged, so
  // the array may be refilled by the next call.
  rc = CheckHip(hipMemcpyAsync(dev, host, n * sizeof(TaskDev), hipMemcpyHostToDevice, s), "scratch task table upload");
  for (int c = 0; c < kClasses && !rc; ++c) {
    if (!ntiles[c]) continue;
    const TaskDev *tt = dev + first[c];
    const int cnt = first[c + 1] - first[c];
    switch (c) {
      case 0: LaunchNoAlloc<0>(tt, cnt, ntiles[c], stage_len, g->d_params, s); break;
      case 1: LaunchNoAlloc<1>(tt, cnt, ntiles[c], stage_len, g->d_params, s); break;
      case 2: LaunchNoAlloc<2>(tt, cnt, ntiles[c], stage_len, g->d_params, s); break;
      case 3: LaunchNoAlloc<3>(tt, cnt, ntiles[c], stage_len, g->d_params, s); break;
      case 4: LaunchNoAlloc<4>(tt, cnt, ntiles[c], stage_len, g->d_params, s); break;
      case 5: LaunchNoAlloc<5>(tt, cnt, ntiles[c], stage_len, g->d_params, s); break;
      case 6: LaunchNoAlloc<6>(tt, cnt, ntiles[c], stage_len, g->d_params, s); break;
      case 7: LaunchNoAlloc<7>(tt, cnt, ntiles[c], stage_len, g->d_params, s); break;
      case 8: LaunchNoAlloc<8>(tt, cnt, ntiles[c], stage_len, g->d_params, s); break;
      default: LaunchNoAlloc<9>(tt, cnt, ntiles[c], stage_len, g->d_params, s); break;
    }
    rc = CheckHip(hipGetLastError(), "no-allocation batch launch");
  }
  return rc;
}

int vg_batch_create(const vg_geometry *g, const vg_query_task *tasks, int32_t ntasks, vg_batch **out) {
  if (!g || !out || ntasks < 0 || (ntasks > 0 && !tasks)) return SetError(VG_ERR_INVALID, "null argument");
  *out = nullptr;
  int rc = EnsureDevice();
  if (rc) return rc;
  auto *b = new vg_batch;
  b->g = g;
  for (int i = 0; i < ntasks; ++i) {
    TaskDev t;
    rc = MakeTask(g, tasks[i], i, t);
    if (rc) {
      delete b;
      return rc;
    }
    b->tasks.push_back(tasks[i]);
    if (t.n == 0) continue;
    b->items += t.n;
    b->host.push_back(t);
  }
  // By class, and within a class the costlier queries first: tiles are
  // dispatched roughly in order, so a launch ends on short tiles instead of
  // on a few long ones (neutral on the zoo step, whose classes are balanced).
  std::stable_sort(b->host.begin(), b->host.end(), [](const TaskDev &a, const TaskDev &c) {
    const int ka = TaskClass(a), kc = TaskClass(c);
    return ka != kc ? ka < kc : QueryWeight(a.query) > QueryWeight(c.query);
  });
  for (int c = 0; c <= kClasses; ++c) {
    int k = 0;
    while (k < (int)b->host.size() && TaskClass(b->host[k]) < c) ++k;
    b->first[c] = k;
  }
  for (int c : {kStagedClass, kStagedClass + kBaseClasses})
    for (int k = b->first[c]; k < b->first[c + 1]; ++k) b->stage_len = std::max(b->stage_len, b->host[k].param_len);
  for (int c = 0; c < kClasses; ++c)
    for (int k = b->first[c]; k < b->first[c + 1]; ++k) {
      b->class_items[c] += b->host[k].n;
      b->class_bytes[c] += TaskBytes(b->host[k]);
    }
  rc = CreateSync(b);
  if (!rc && !b->host.empty()) rc = UploadAll(b);
  if (rc) {
    FreeTables(b);
    DestroySync(b);
    delete b;
    return rc;
  }
  *out = b;
  return VG_OK;
}

int vg_query(const vg_geometry *g, const vg_query_task *task, void *stream) {
  if (!g || !task) return SetError(VG_ERR_INVALID, "null argument");
  int rc = EnsureDevice();
  if (rc) return rc;
  TaskDev t;
  if ((rc = MakeTask(g, *task, 0, t))) return rc;
  if (t.n == 0) return VG_OK;
  const int it = t.vec_ok ? 2 : 1;
  const uint64_t blocks = (t.n + (uint64_t)it * kBatchBlock - 1) / ((uint64_t)it * kBatchBlock);
  const hipStream_t s = (hipStream_t)stream;
  const size_t lds = t.param_len <= kMaxStageDoubles ? (size_t)t.param_len * sizeof(double) : 0;
  // Class 4 (blocks too large for LDS) runs the class-3 code, which reads
  // the block from global memory when it does not fit.
  const int cls = std::min(ClassOf(t.type), kStagedClass) + (Filtered(t) ? kBaseClasses : 0);
  LaunchTask(cls, it, dim3((unsigned)blocks), lds, s, t, g->d_params);
  return CheckHip(hipGetLastError(), "query launch");
}

int vg_set_mixed_policy(int32_t policy) {
  if (policy != VG_MIXED_POLICY_MEASURED && policy != VG_MIXED_POLICY_ALL)
    return SetError(VG_ERR_INVALID, "unknown mixed-precision policy");
  g_mixed_policy = policy;
  return VG_OK;
}

int vg_batch_run(vg_batch *b, void *stream) {
  if (!b) return SetError(VG_ERR_INVALID, "null batch");
  int rc = EnsureDevice();
  if (rc) return rc;
  return Launch(b, (hipStream_t)stream);
}

int vg_batch_set_launch(vg_batch *b, int32_t items_per_lane, uint32_t blocks) {
  if (!b) return SetError(VG_ERR_INVALID, "null batch");
  if (items_per_lane != 1 && items_per_lane != 2 && items_per_lane != 4)
    return SetError(VG_ERR_INVALID, "items per lane must be 1, 2 or 4");
  int rc = EnsureDevice();
  if (rc) return rc;
  b->it = items_per_lane;
  b->blocks = blocks;
  Select(b);
  return VG_OK;
}

int vg_batch_get_launch(const vg_batch *b, int32_t *items_per_lane, uint32_t *blocks, uint64_t *items) {
  if (!b) return SetError(VG_ERR_INVALID, "null batch");
  if (items_per_lane) *items_per_lane = b->it;
  if (blocks) *blocks = b->blocks;
  if (items) *items = b->items;
  return VG_OK;
}

int vg_batch_autotune(vg_batch *b, int32_t repetitions, void *stream, float *best_ms) {
  if (!b || repetitions < 1) return SetError(VG_ERR_INVALID, "null batch or repetitions < 1");
  int rc = EnsureDevice();
  if (rc) return rc;
  hipStream_t s = (hipStream_t)stream;
  int dev = 0, cus = 256;
  if (hipGetDevice(&dev) == hipSuccess) (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  // Small persistent grids leave wave slots to the other classes, which then
  // run beside the light class (it is bound by HBM, they by latency).
  const unsigned grids[12] = {0u,
                              (unsigned)cus,
                              (unsigned)cus * 9 / 8,
                              (unsigned)cus * 5 / 4,
                              (unsigned)cus * 11 / 8,
                              (unsigned)cus * 3 / 2,
                              (unsigned)cus * 7 / 4,
                              (unsigned)cus * 2,
                              (unsigned)cus * 3,
                              (unsigned)cus * 4,
                              (unsigned)cus * 8,
                              (unsigned)cus * 16};
  hipEvent_t e0, e1;
  if ((rc = CheckHip(hipEventCreate(&e0), "hipEventCreate"))) return rc;
  if ((rc = CheckHip(hipEventCreate(&e1), "hipEventCreate"))) {
    (void)hipEventDestroy(e0);
    return rc;
  }
  float best = 1e30f;
  int bestIt = b->it;
  unsigned bestBlocks = b->blocks;
  for (int it : kITs) {
    for (unsigned gr : grids) {
      b->it = it;
      b->blocks = gr;
      Select(b);
      if ((rc = Launch(b, s))) break;  // warm-up
      (void)hipEventRecord(e0, s);
      for (int r = 0; r < repetitions && !rc; ++r) rc = Launch(b, s);
      (void)hipEventRecord(e1, s);
      if (rc || (rc = CheckHip(hipEventSynchronize(e1), "autotune sync"))) break;
      float ms = 0;
      (void)hipEventElapsedTime(&ms, e0, e1);
      ms /= repetitions;
      if (ms < best) {
        best = ms;
        bestIt = it;
        bestBlocks = gr;
      }
    }
    if (rc) break;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  b->it = bestIt;
  b->blocks = bestBlocks;
  Select(b);
  if (best_ms) *best_ms = best;
  return rc;
}

int vg_batch_set_timing(vg_batch *b, int32_t enable) {
  if (!b) return SetError(VG_ERR_INVALID, "null batch");
  int rc = EnsureDevice();
  if (rc) return rc;
  if (enable && !b->t0[0][0]) {
    for (int c = 0; c < kClasses && !rc; ++c)
      for (int k = 0; k < vg_batch::kRing && !rc; ++k) {
        rc = CheckHip(hipEventCreate(&b->t0[c][k]), "hipEventCreate");
        if (!rc) rc = CheckHip(hipEventCreate(&b->t1[c][k]), "hipEventCreate");
      }
    if (rc) {
      DestroyTiming(b);
      return rc;
    }
  }
  b->timing = enable != 0;
  b->ring_used = b->ring_next = 0;
  std::memset(b->recorded, 0, sizeof(b->recorded));
  return VG_OK;
}

int vg_batch_class_stats(vg_batch *b, int32_t cls, uint64_t *items, uint64_t *bytes, int32_t *runs, float *mean_ms) {
  if (!b || cls < 0 || cls >= kClasses) return SetError(VG_ERR_INVALID, "null batch or class out of range");
  if (items) *items = b->class_items[cls];
  if (bytes) *bytes = b->class_bytes[cls];
  int n = 0;
  double sum = 0;
  if (mean_ms || runs) {
    int rc = EnsureDevice();
    if (rc) return rc;
    for (int k = 0; k < b->ring_used; ++k) {
      if (!b->recorded[cls][k]) continue;
      if ((rc = CheckHip(hipEventSynchronize(b->t1[cls][k]), "batch timing sync"))) return rc;
      float ms = 0;
      if ((rc = CheckHip(hipEventElapsedTime(&ms, b->t0[cls][k], b->t1[cls][k]), "batch timing read"))) return rc;
      sum += ms;
      ++n;
    }
  }
  if (runs) *runs = n;
  if (mean_ms) *mean_ms = n ? (float)(sum / n) : 0.f;
  return VG_OK;
}

int vg_batch_free(vg_batch *b) {
  if (!b) return SetError(VG_ERR_INVALID, "null batch");
  FreeTables(b);
  DestroySync(b);
  delete b;
  return VG_OK;
}

}  // extern "C"
