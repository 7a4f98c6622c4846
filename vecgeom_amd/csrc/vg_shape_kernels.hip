// Basket shape-query kernels for gfx950.
//
// One launch serves one placed volume and one query over an SOA3D basket
// (the reference's ShapeImplementationHelper basket loops,
// volumes/ShapeImplementationHelper.h:588-748, and the per-point CUDA
// kernels of source/benchmarking/Benchmarker.cu:13-79).  Design:
//  * the shape parameter block and the placement transform are kernel
//    arguments passed by value, so every lane reads them from SGPRs: no
//    pointer chase per point (the reference does a device virtual call);
//  * the transform kind is a template parameter, as Transform<t,r> is in the
//    reference, so identity placements do no arithmetic;
//  * each lane handles IT consecutive points with 16-byte loads and stores
//    from each SOA array (global_load_dwordx4), keeping enough bytes in
//    flight per CU to run at HBM speed;
//  * 256-thread blocks (4 wave64), grid = ceil(n / (256*IT)).
#include <hip/hip_runtime.h>

#include <cstring>

#include "shapes.hpp"
#include "vg_stage.hpp"
#include "vg_walk.hpp"
#include "vg_internal.hpp"

namespace vga {

namespace {

constexpr int kBlock = 256;

template <int IT>
struct alignas(8 * IT) DVec {
  double v[IT];
};
template <int IT>
struct alignas(4 * IT) IVec {
  int32_t v[IT];
};
template <int IT>
struct alignas(IT) BVec {
  uint8_t v[IT];
};

template <int IT>
__device__ __forceinline__ void LoadD(const double *__restrict__ p, double (&o)[IT]) {
  DVec<IT> t = *reinterpret_cast<const DVec<IT> *>(p);
#pragma unroll
  for (int k = 0; k < IT; ++k) o[k] = t.v[k];
}

template <class S, int XK, int Q>
__device__ __forceinline__ void Eval(const typename S::P &P, const Xf &X, V3 p, V3 d, double sm, double &rd,
                                     int32_t &ri) {
  if constexpr (Q == kQContains) {
    ri = S::Contains(P, XfPoint<XK>(X, p)) ? 1 : 0;
  } else if constexpr (Q == kQInside) {
    ri = S::Inside(P, XfPoint<XK>(X, p));
  } else if constexpr (Q == kQDistIn || Q == kQDistInMin) {
    rd = S::DistanceToIn(P, XfPoint<XK>(X, p), XfDir<XK>(X, d), sm);
  } else if constexpr (Q == kQDistOut) {
    rd = S::DistanceToOut(P, p, d, sm);
  } else if constexpr (Q == kQPlacedDistOut) {
    rd = S::DistanceToOut(P, XfPoint<XK>(X, p), XfDir<XK>(X, d), sm);
  } else if constexpr (Q == kQSafeIn || Q == kQSafeInMin) {
    rd = S::SafetyToIn(P, XfPoint<XK>(X, p));
  } else {
    rd = S::SafetyToOut(P, p);
  }
}

constexpr bool NeedsDir(int q) { return q == kQDistIn || q == kQDistInMin || q == kQDistOut || q == kQPlacedDistOut; }
constexpr bool NeedsStep(int q) { return q == kQDistIn || q == kQDistOut || q == kQPlacedDistOut; }

// Variable-length parameter blocks (Polycone, Polyhedron) are read many
// times per query; with Stage the workgroup first copies the block to LDS
// and every lane reads it from there.
constexpr int kMaxStageDoubles = 6144;  // 48 KiB of LDS

// Occupancy floors (waves per SIMD) for the staged kernels, which are bound
// by latency (LDS reads, dependent FP64 chains, divergence) rather than by
// issue: left alone the compiler spends up to 128 VGPRs on them and holds a
// SIMD to 4 waves.  Asking for 6 (7 for the refill walk) costs a few
// spilled registers in the safeties and measured 4-13 % faster per launch on
// the zoo workload (Polyhedron DistanceToIn 249 -> 216 us, SafetyToIn 214 ->
// 200 us); 8 measured slower.
constexpr int kStagedWaves = 6;
constexpr int kWalkWaves = 7;

// One lane's IT consecutive queries starting at i0.
template <class S, int XK, int Q, int IT>
__device__ __forceinline__ void QueryItems(const typename S::P &P, const Xf &X, const QArgs &a, uint64_t i0) {
  if (i0 >= a.n) return;
  double px[IT], py[IT], pz[IT], dx[IT], dy[IT], dz[IT], sm[IT];
  const bool full = i0 + IT <= a.n;
  if (full) {
    LoadD<IT>(a.x + i0, px);
    LoadD<IT>(a.y + i0, py);
    LoadD<IT>(a.z + i0, pz);
    if constexpr (NeedsDir(Q)) {
      LoadD<IT>(a.dx + i0, dx);
      LoadD<IT>(a.dy + i0, dy);
      LoadD<IT>(a.dz + i0, dz);
    }
    if constexpr (NeedsStep(Q)) {
      if (a.sm) {
        LoadD<IT>(a.sm + i0, sm);
      } else {
#pragma unroll
        for (int k = 0; k < IT; ++k) sm[k] = kInfinity;
      }
    }
    if constexpr (Q == kQDistInMin || Q == kQSafeInMin || Q == kQSafeOutMin) {
      // Minimize variants read the running minimum in place.
      LoadD<IT>(reinterpret_cast<const double *>(a.out) + i0, sm);
    }
  } else {
#pragma unroll
    for (int k = 0; k < IT; ++k) {
      const uint64_t i = i0 + k < a.n ? i0 + k : i0;
      px[k] = a.x[i];
      py[k] = a.y[i];
      pz[k] = a.z[i];
      if constexpr (NeedsDir(Q)) {
        dx[k] = a.dx[i];
        dy[k] = a.dy[i];
        dz[k] = a.dz[i];
      }
      if constexpr (NeedsStep(Q)) sm[k] = a.sm ? a.sm[i] : kInfinity;
      if constexpr (Q == kQDistInMin || Q == kQSafeInMin || Q == kQSafeOutMin)
        sm[k] = reinterpret_cast<const double *>(a.out)[i];
    }
  }

  double rd[IT];
  int32_t ri[IT];
#pragma unroll
  for (int k = 0; k < IT; ++k) {
    V3 p{px[k], py[k], pz[k]};
    V3 d{0, 0, 0};
    if constexpr (NeedsDir(Q)) d = V3{dx[k], dy[k], dz[k]};
    double s = 0;
    if constexpr (NeedsStep(Q) || Q == kQDistInMin) s = sm[k];
    Eval<S, XK, Q>(P, X, p, d, s, rd[k], ri[k]);
  }

  if constexpr (Q == kQContains) {
    uint8_t *o = reinterpret_cast<uint8_t *>(a.out) + i0;
    if (full) {
      BVec<IT> t;
#pragma unroll
      for (int k = 0; k < IT; ++k) t.v[k] = (uint8_t)ri[k];
      *reinterpret_cast<BVec<IT> *>(o) = t;
    } else {
      for (int k = 0; i0 + k < a.n; ++k) o[k] = (uint8_t)ri[k];
    }
  } else if constexpr (Q == kQInside) {
    int32_t *o = reinterpret_cast<int32_t *>(a.out) + i0;
    if (full) {
      IVec<IT> t;
#pragma unroll
      for (int k = 0; k < IT; ++k) t.v[k] = ri[k];
      *reinterpret_cast<IVec<IT> *>(o) = t;
    } else {
      for (int k = 0; i0 + k < a.n; ++k) o[k] = ri[k];
    }
  } else if constexpr (Q == kQDistInMin) {
    // ShapeImplementationHelper.h:637-658
    double *o = reinterpret_cast<double *>(a.out) + i0;
    for (int k = 0; k < IT && i0 + k < a.n; ++k) {
      if (rd[k] < sm[k]) {
        o[k] = rd[k];
        a.next[i0 + k] = a.did;
      }
    }
  } else if constexpr (Q == kQSafeInMin || Q == kQSafeOutMin) {
    // ShapeImplementationHelper.h:707-746
    double *o = reinterpret_cast<double *>(a.out) + i0;
    for (int k = 0; k < IT && i0 + k < a.n; ++k) o[k] = (rd[k] < sm[k]) ? rd[k] : sm[k];
  } else {
    double *o = reinterpret_cast<double *>(a.out) + i0;
    if (full) {
      DVec<IT> t;
#pragma unroll
      for (int k = 0; k < IT; ++k) t.v[k] = rd[k];
      *reinterpret_cast<DVec<IT> *>(o) = t;
    } else {
      for (int k = 0; i0 + k < a.n; ++k) o[k] = rd[k];
    }
    if constexpr (Q == kQDistOut) {
      // ShapeImplementationHelper.h:677-692
      if (a.next) {
        for (int k = 0; k < IT && i0 + k < a.n; ++k) a.next[i0 + k] = (rd[k] < sm[k]) ? -1 : -2;
      }
    }
  }
}

// One workgroup per kBlock * IT queries.  (A grid of only the resident
// workgroups, each staging once and striding over the tiles, measured slower
// for the expensive queries: the per-tile cost varies too much for a static
// split, and the hardware's dispatch of short workgroups balances it.)
template <class S, int XK, int Q, int IT, bool Stage>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(ParamView<S>::kIndirect ? kStagedWaves : 1)))
void ShapeQueryKernel(const typename S::P Pg, const Xf X, const QArgs a,
                                                           int stage_len) {
  if constexpr (Stage) {
    extern __shared__ __align__(16) double stage[];
    StageBlock(stage, Pg.blk, stage_len, kBlock);
    __syncthreads();
    QueryItems<S, XK, Q, IT>(typename S::P{stage}, X, a, ((uint64_t)blockIdx.x * kBlock + threadIdx.x) * IT);
  } else {
    QueryItems<S, XK, Q, IT>(Pg, X, a, ((uint64_t)blockIdx.x * kBlock + threadIdx.x) * IT);
  }
}

template <class S, class = void>
struct HasQuick {
  static constexpr bool value = false;
};
template <class S>
struct HasQuick<S, decltype((void)S::kHasQuick)> {
  static constexpr bool value = S::kHasQuick;
};

// DistanceToIn with lane compaction, for shapes whose DistanceToIn starts
// with a cheap exact rejection (a bounding volume the ray misses).  Phase 1:
// every lane runs the rejection and writes kInfinity for rejected rays; the
// others append their query to a workgroup list in LDS.  Phase 2: the first
// `count` lanes run the full DistanceToIn on the listed queries, so the
// expensive part runs on dense waves instead of waves where most lanes idle.
template <class S, int XK, bool Stage>
__global__ __launch_bounds__(kBlock) void CompactDistInKernel(const typename S::P Pg, const Xf X, const QArgs a,
                                                              int stage_len) {
  typename S::P P = Pg;
  if constexpr (Stage) {
    extern __shared__ __align__(16) double stage[];
    StageBlock(stage, Pg.blk, stage_len, kBlock);
    P = typename S::P{stage};
  }
  __shared__ int count;
  __shared__ int lane[kBlock];
  __shared__ double qp[3][kBlock], qd[3][kBlock], qsm[kBlock];
  if (threadIdx.x == 0) count = 0;
  __syncthreads();
  const uint64_t base = (uint64_t)blockIdx.x * kBlock;
  const uint64_t i = base + threadIdx.x;
  if (i < a.n) {
    const V3 p = XfPoint<XK>(X, V3{a.x[i], a.y[i], a.z[i]});
    const V3 d = XfDir<XK>(X, V3{a.dx[i], a.dy[i], a.dz[i]});
    const double sm = a.sm ? a.sm[i] : kInfinity;
    double r;
    if (S::DistanceToInQuick(P, p, d, sm, r)) {
      reinterpret_cast<double *>(a.out)[i] = r;
    } else {
      const int k = atomicAdd(&count, 1);
      lane[k] = threadIdx.x;
      qp[0][k] = p.x;
      qp[1][k] = p.y;
      qp[2][k] = p.z;
      qd[0][k] = d.x;
      qd[1][k] = d.y;
      qd[2][k] = d.z;
      qsm[k] = sm;
    }
  }
  __syncthreads();
  const int k = threadIdx.x;
  if (k < count) {
    reinterpret_cast<double *>(a.out)[base + lane[k]] =
        S::DistanceToIn(P, V3{qp[0][k], qp[1][k], qp[2][k]}, V3{qd[0][k], qd[1][k], qd[2][k]}, qsm[k]);
  }
}

// Walk flags of a shape: S::kWalkIn / S::kWalkOut when it declares them.
template <class S, class = void>
struct Walks {
  static constexpr bool in = false, out = false;
};
template <class S>
struct Walks<S, decltype((void)S::kWalkOut)> {
  static constexpr bool in = S::kWalkIn, out = S::kWalkOut;
};

// Distances with lane refill (vg_walk.hpp) for shapes whose DistanceToIn /
// DistanceToOut is a walk of variable length: each wave owns per_lane * 64
// consecutive queries.  Ops is S::DistInOps (points and directions
// transformed by XK) or S::DistOutOps (XK the identity for local inputs, the
// placement for PlacedDistanceToOut); NextIndex writes the DistanceToOut
// node index.
template <class S, class Ops, int XK, bool Stage, bool NextIndex>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(kWalkWaves))) void WalkKernel(
    const typename S::P Pg, const Xf X, const QArgs a, int stage_len, int per_lane) {
  typename S::P P = Pg;
  if constexpr (Stage) {
    extern __shared__ __align__(16) double stage[];
    StageBlock(stage, Pg.blk, stage_len, kBlock);
    __syncthreads();
    P = typename S::P{stage};
  }
  const uint64_t chunk = 64 * (uint64_t)per_lane;
  const uint64_t begin = ((uint64_t)blockIdx.x * kBlock + threadIdx.x) / 64 * chunk;
  const uint64_t end = begin + chunk < a.n ? begin + chunk : a.n;
  double *out = reinterpret_cast<double *>(a.out);
  WalkRefill<Ops>(
      P, begin, end,
      [&](uint64_t i, V3 &p, V3 &d, double &sm, double &) {
        p = XfPoint<XK>(X, V3{a.x[i], a.y[i], a.z[i]});
        d = XfDir<XK>(X, V3{a.dx[i], a.dy[i], a.dz[i]});
        sm = a.sm ? a.sm[i] : kInfinity;
        return false;
      },
      [&](uint64_t i, double r) {
        out[i] = r;
        // ShapeImplementationHelper.h:677-692 (DistanceToOut with node index)
        if (NextIndex && a.next) a.next[i] = (r < (a.sm ? a.sm[i] : kInfinity)) ? -1 : -2;
      });
}

// Queries per lane of a walk launch: up to 8, fewer for small baskets so
// that the grid still holds about 4096 waves.
int WalkPerLane(uint64_t n) {
  const uint64_t q = n / (64ull * 4096ull);
  return q < 1 ? 1 : (q > 8 ? 8 : (int)q);
}

template <class S, class Ops, int XK, bool NextIndex = false>
int LaunchWalk(const typename S::P &P, const Xf &X, const QArgs &a, int stage_len, hipStream_t s) {
  const int per_lane = WalkPerLane(a.n);
  const uint64_t waves = (a.n + 64ull * per_lane - 1) / (64ull * per_lane);
  const uint64_t blocks = (waves + kBlock / 64 - 1) / (kBlock / 64);
  if (stage_len > 0 && stage_len <= kMaxStageDoubles)
    hipLaunchKernelGGL((WalkKernel<S, Ops, XK, true, NextIndex>), dim3((unsigned)blocks), dim3(kBlock),
                       stage_len * sizeof(double), s, P, X, a, stage_len, per_lane);
  else
    hipLaunchKernelGGL((WalkKernel<S, Ops, XK, false, NextIndex>), dim3((unsigned)blocks), dim3(kBlock), 0, s, P, X, a, 0,
                       per_lane);
  return CheckHip(hipGetLastError(), "shape query launch");
}

// Plain FP64 form (the benchmark denominator): one thread per query, the
// placement record and shape parameters read from global memory through the
// placed-volume id (as the reference's device virtual call chases its object
// pointers), the transform kind switched at run time, 8-byte loads.
template <class S, int Q>
__global__ __launch_bounds__(kBlock) void PlainQueryKernel(const PlacedRec *__restrict__ placed,
                                                           const double *__restrict__ params, int32_t pid,
                                                           const QArgs a) {
  const uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= a.n) return;
  const PlacedRec &pr = placed[pid];
  const auto &P = ParamView<S>::Get(params + pr.param_off);
  const V3 p{a.x[i], a.y[i], a.z[i]};
  V3 d{0, 0, 0};
  if constexpr (NeedsDir(Q)) d = V3{a.dx[i], a.dy[i], a.dz[i]};
  double sm = kInfinity;
  if constexpr (NeedsStep(Q)) sm = a.sm ? a.sm[i] : kInfinity;
  if constexpr (Q == kQContains) {
    reinterpret_cast<uint8_t *>(a.out)[i] = S::Contains(P, XfPointRt(pr.xf, pr.xfkind, p)) ? 1 : 0;
  } else if constexpr (Q == kQInside) {
    reinterpret_cast<int32_t *>(a.out)[i] = S::Inside(P, XfPointRt(pr.xf, pr.xfkind, p));
  } else if constexpr (Q == kQDistIn) {
    reinterpret_cast<double *>(a.out)[i] =
        S::DistanceToIn(P, XfPointRt(pr.xf, pr.xfkind, p), XfDirRt(pr.xf, pr.xfkind, d), sm);
  } else if constexpr (Q == kQDistOut) {
    reinterpret_cast<double *>(a.out)[i] = S::DistanceToOut(P, p, d, sm);
  } else if constexpr (Q == kQSafeIn) {
    reinterpret_cast<double *>(a.out)[i] = S::SafetyToIn(P, XfPointRt(pr.xf, pr.xfkind, p));
  } else if constexpr (Q == kQSafeOut) {
    reinterpret_cast<double *>(a.out)[i] = S::SafetyToOut(P, p);
  }
}

bool Aligned16(const void *p) { return (reinterpret_cast<uintptr_t>(p) & 15u) == 0; }
bool Aligned32(const void *p) { return (reinterpret_cast<uintptr_t>(p) & 31u) == 0; }

int g_variant = 0;  // 0 = tuned kernels, 1 = plain FP64 form

template <class S, int XK, int Q>
int LaunchQ(const typename S::P &P, const Xf &X, const QArgs &a, int stage_len, hipStream_t s) {
  if (a.n == 0) return VG_OK;
  bool al = Aligned16(a.x) && Aligned16(a.y) && Aligned16(a.z) && Aligned16(a.out);
  if (NeedsDir(Q)) al = al && Aligned16(a.dx) && Aligned16(a.dy) && Aligned16(a.dz);
  if (NeedsStep(Q) && a.sm) al = al && Aligned16(a.sm);
  if constexpr (Q == kQDistIn && Walks<S>::in) return LaunchWalk<S, typename S::DistInOps, XK>(P, X, a, stage_len, s);
  if constexpr (Q == kQDistOut && Walks<S>::out)
    return LaunchWalk<S, typename S::DistOutOps, kXfIdentity, true>(P, X, a, stage_len, s);
  if constexpr (Q == kQPlacedDistOut && Walks<S>::out)
    return LaunchWalk<S, typename S::DistOutOps, XK>(P, X, a, stage_len, s);
  if constexpr (Q == kQDistIn && HasQuick<S>::value) {
    const uint64_t blocks = (a.n + kBlock - 1) / kBlock;
    if constexpr (ParamView<S>::kIndirect) {
      if (stage_len > 0 && stage_len <= kMaxStageDoubles) {
        hipLaunchKernelGGL((CompactDistInKernel<S, XK, true>), dim3((unsigned)blocks), dim3(kBlock),
                           stage_len * sizeof(double), s, P, X, a, stage_len);
        return CheckHip(hipGetLastError(), "shape query launch");
      }
    }
    hipLaunchKernelGGL((CompactDistInKernel<S, XK, false>), dim3((unsigned)blocks), dim3(kBlock), 0, s, P, X, a, 0);
    return CheckHip(hipGetLastError(), "shape query launch");
  }
  if constexpr (ParamView<S>::kIndirect) {
    if (stage_len > 0 && stage_len <= kMaxStageDoubles) {
      const uint64_t blocks = (a.n + kBlock - 1) / kBlock;
      hipLaunchKernelGGL((ShapeQueryKernel<S, XK, Q, 1, true>), dim3((unsigned)blocks), dim3(kBlock),
                         stage_len * sizeof(double), s, P, X, a, stage_len);
      return CheckHip(hipGetLastError(), "shape query launch");
    }
  }
  if constexpr (Q == kQContains) {
    // Four queries per lane: the flags leave as one 4-byte store per lane
    // (two queries per lane would store 2 bytes).
    if (Aligned32(a.x) && Aligned32(a.y) && Aligned32(a.z) && (reinterpret_cast<uintptr_t>(a.out) & 3u) == 0) {
      constexpr int IT = 4;
      const uint64_t blocks = (a.n + (uint64_t)kBlock * IT - 1) / ((uint64_t)kBlock * IT);
      hipLaunchKernelGGL((ShapeQueryKernel<S, XK, Q, IT, false>), dim3((unsigned)blocks), dim3(kBlock), 0, s, P, X, a,
                         0);
      return CheckHip(hipGetLastError(), "shape query launch");
    }
  }
  if (al) {
    constexpr int IT = 2;
    const uint64_t blocks = (a.n + (uint64_t)kBlock * IT - 1) / ((uint64_t)kBlock * IT);
    hipLaunchKernelGGL((ShapeQueryKernel<S, XK, Q, IT, false>), dim3((unsigned)blocks), dim3(kBlock), 0, s, P, X, a,
                       0);
  } else {
    const uint64_t blocks = (a.n + kBlock - 1) / kBlock;
    hipLaunchKernelGGL((ShapeQueryKernel<S, XK, Q, 1, false>), dim3((unsigned)blocks), dim3(kBlock), 0, s, P, X, a,
                       0);
  }
  return CheckHip(hipGetLastError(), "shape query launch");
}

template <class S, int Q>
int LaunchXf(const typename S::P &P, const PlacedRec &pr, const QArgs &a, int stage_len, hipStream_t s) {
  switch (pr.xfkind) {
    case kXfIdentity: return LaunchQ<S, kXfIdentity, Q>(P, pr.xf, a, stage_len, s);
    case kXfTranslation: return LaunchQ<S, kXfTranslation, Q>(P, pr.xf, a, stage_len, s);
    case kXfRotation: return LaunchQ<S, kXfRotation, Q>(P, pr.xf, a, stage_len, s);
    default: return LaunchQ<S, kXfGeneral, Q>(P, pr.xf, a, stage_len, s);
  }
}

template <class S, int Q>
int LaunchPlain(const vg_geometry *g, int32_t pid, const QArgs &a, hipStream_t s) {
  const uint64_t blocks = (a.n + kBlock - 1) / kBlock;
  hipLaunchKernelGGL((PlainQueryKernel<S, Q>), dim3((unsigned)blocks), dim3(kBlock), 0, s, g->d_placed, g->d_params,
                     pid, a);
  return CheckHip(hipGetLastError(), "plain shape query launch");
}

template <class S>
int LaunchShapePlain(const vg_geometry *g, int32_t pid, int q, const QArgs &a, hipStream_t s) {
  switch (q) {
    case kQContains: return LaunchPlain<S, kQContains>(g, pid, a, s);
    case kQInside: return LaunchPlain<S, kQInside>(g, pid, a, s);
    case kQDistIn: return LaunchPlain<S, kQDistIn>(g, pid, a, s);
    case kQDistOut: return LaunchPlain<S, kQDistOut>(g, pid, a, s);
    case kQSafeIn: return LaunchPlain<S, kQSafeIn>(g, pid, a, s);
    case kQSafeOut: return LaunchPlain<S, kQSafeOut>(g, pid, a, s);
  }
  return SetError(VG_ERR_UNSUPPORTED, "the plain form covers the six basic queries only");
}

template <class S>
int LaunchShape(const vg_geometry *g, int32_t pid, const double *block, const PlacedRec &pr, int q, const QArgs &a,
                hipStream_t s) {
  if (g_variant == 1) {
    if (a.n == 0) return VG_OK;
    return LaunchShapePlain<S>(g, pid, q, a, s);
  }
  typename S::P P;
  int len = 0;
  if constexpr (ParamView<S>::kIndirect) {
    P = ParamView<S>::Get(g->d_params + pr.param_off);
    len = ParamView<S>::Length(block);
  } else {
    std::memcpy(&P, block, sizeof(P));
  }
  switch (q) {
    case kQContains: return LaunchXf<S, kQContains>(P, pr, a, len, s);
    case kQInside: return LaunchXf<S, kQInside>(P, pr, a, len, s);
    case kQDistIn: return LaunchXf<S, kQDistIn>(P, pr, a, len, s);
    case kQDistInMin: return LaunchXf<S, kQDistInMin>(P, pr, a, len, s);
    case kQDistOut: return LaunchQ<S, kXfIdentity, kQDistOut>(P, pr.xf, a, len, s);
    case kQPlacedDistOut: return LaunchXf<S, kQPlacedDistOut>(P, pr, a, len, s);
    case kQSafeIn: return LaunchXf<S, kQSafeIn>(P, pr, a, len, s);
    case kQSafeInMin: return LaunchXf<S, kQSafeInMin>(P, pr, a, len, s);
    case kQSafeOut: return LaunchQ<S, kXfIdentity, kQSafeOut>(P, pr.xf, a, len, s);
    case kQSafeOutMin: return LaunchQ<S, kXfIdentity, kQSafeOutMin>(P, pr.xf, a, len, s);
  }
  return SetError(VG_ERR_INVALID, "unknown query");
}

}  // namespace

int SetKernelVariant(int v) {
  if (v != 0 && v != 1) return SetError(VG_ERR_INVALID, "kernel variant must be 0 (tuned) or 1 (plain FP64)");
  g_variant = v;
  return VG_OK;
}

int LaunchQuery(const vg_geometry *g, int32_t placed, int q, const QArgs &a, hipStream_t s) {
  const PlacedRec &pr = g->placed[placed];
  const double *block = g->params.data() + pr.param_off;
  switch (pr.type) {
    case VG_BOOLEAN: return LaunchBooleanQuery(g, placed, q, a, s);
    case VG_BOX: return LaunchShape<BoxS>(g, placed, block, pr, q, a, s);
    case VG_TUBE: return LaunchShape<TubeS>(g, placed, block, pr, q, a, s);
    case VG_ORB: return LaunchShape<OrbS>(g, placed, block, pr, q, a, s);
    case VG_TRD: return LaunchShape<TrdS>(g, placed, block, pr, q, a, s);
    case VG_PARALLELEPIPED: return LaunchShape<ParaS>(g, placed, block, pr, q, a, s);
    case VG_PARABOLOID: return LaunchShape<ParabS>(g, placed, block, pr, q, a, s);
    case VG_TRAPEZOID: return LaunchShape<TrapS>(g, placed, block, pr, q, a, s);
    case VG_SPHERE: return LaunchShape<SphereS>(g, placed, block, pr, q, a, s);
    case VG_CONE: return LaunchShape<ConeS>(g, placed, block, pr, q, a, s);
    case VG_POLYCONE: return LaunchShape<PolyconeS>(g, placed, block, pr, q, a, s);
    case VG_POLYHEDRON: return LaunchShape<PolyhedronS>(g, placed, block, pr, q, a, s);
    case VG_TORUS: return LaunchShape<TorusS>(g, placed, block, pr, q, a, s);
  }
  return SetError(VG_ERR_UNSUPPORTED, "no device kernels for shape type " + std::to_string(pr.type));
}

}  // namespace vga
