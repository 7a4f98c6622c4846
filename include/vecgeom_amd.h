/*
 * vecgeom_amd.h -- C-ABI of the MI355X (gfx950) shape-query and navigation
 * library libvecgeom_amd.so.
 *
 * Every entry point takes plain pointers and sizes.  Array arguments of the
 * query entry points are DEVICE pointers (hipMalloc'd or torch tensors on the
 * GPU); `stream` is a hipStream_t passed as void* (NULL = default stream).
 * Calls are asynchronous on `stream` unless stated otherwise.  Every entry
 * point returns VG_OK (0) or a negative status; vg_last_error() gives the
 * message for the calling thread.
 *
 * Frames follow the reference (volumes/PlacedVolume.h:144-250): Contains,
 * Inside, DistanceToIn and SafetyToIn take points in the MOTHER frame of the
 * placed volume (its placement transform is applied inside); DistanceToOut
 * and SafetyToOut take points in the volume's LOCAL frame;
 * PlacedDistanceToOut takes mother-frame points.
 *
 * Outputs follow the Scalar backend's virtual methods
 * (volumes/ShapeImplementationHelper.h:135-318): a miss is +inf, Box
 * DistanceToIn of an inside point is -1, Inside returns 0 (inside), 1
 * (surface) or 2 (outside).
 */
#ifndef VECGEOM_AMD_H_
#define VECGEOM_AMD_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------ status */
#define VG_OK 0
#define VG_ERR_INVALID (-1)     /* bad argument, handle or id              */
#define VG_ERR_HIP (-2)         /* a HIP runtime call failed               */
#define VG_ERR_STATE (-3)       /* vg_init not called / device mismatch    */
#define VG_ERR_UNSUPPORTED (-4) /* shape or feature not available          */
#define VG_ERR_CAPACITY (-5)    /* request exceeds a configured limit      */

/* Message of the last failed call on this thread ("" if none). */
const char *vg_last_error(void);

/* Library version string. */
const char *vg_version(void);

/*
 * Selects the HIP device used by this process and prepares per-device
 * state.  Replaces the device set-up of CudaManager
 * (source/CudaManager.cpp:21-38).  Idempotent for the same device.
 */
int vg_init(int device);

/* Releases per-device state; geometries must be freed first. */
int vg_shutdown(void);

/* ------------------------------------------------------------ shape types */
/* Constructor parameters, in this order, for each type:
 *   VG_BOX            dx dy dz                         (UnplacedBox.h:30)
 *   VG_TUBE           rmin rmax dz sphi dphi [rad]     (UnplacedTube.h:68)
 *   VG_ORB            r                                (UnplacedOrb.h:45)
 *   VG_SPHERE         rmin rmax sphi dphi stheta dtheta [rad] (UnplacedSphere.h:219)
 *   VG_CONE           rmin1 rmax1 rmin2 rmax2 dz sphi dphi [rad] (UnplacedCone.h:108)
 *   VG_TRD            dx1 dx2 dy1 dy2 dz               (UnplacedTrd.h:73)
 *   VG_TRAPEZOID      dz theta phi dy1 dx1 dx2 tanAlpha1 dy2 dx3 dx4 tanAlpha2
 *                                                      (UnplacedTrapezoid.h:93)
 *   VG_PARALLELEPIPED x y z alpha theta phi [deg]      (UnplacedParallelepiped.h:34)
 *   VG_PARABOLOID     rlo rhi dz                       (UnplacedParaboloid.h:85)
 *   VG_TORUS          rmin rmax rtor sphi dphi [rad]   (UnplacedTorus.h:61)
 *   VG_POLYCONE       sphi dphi nz rmin[nz] rmax[nz] z[nz] (UnplacedPolycone.h:82)
 *   VG_POLYHEDRON     phiStart phiDelta [deg] sides nz z[nz] rmin[nz] rmax[nz]
 *                                                      (UnplacedPolyhedron.h:169)
 *   VG_BOOLEAN        op, then for the left and the right operand: type
 *                     nparams params[nparams] placement[12]
 *                     (UnplacedBooleanVolume(op, left, right),
 *                     UnplacedBooleanVolume.h:39-50; op 0 union,
 *                     1 intersection, 2 subtraction; each operand a
 *                     primitive shape placed in the Boolean's frame,
 *                     placement = translation then the 9 rotation entries).
 *                     Basket and scalar shape queries only: batches and
 *                     navigation refuse Boolean volumes.
 */
enum vg_shape_type {
  VG_BOX = 1,
  VG_TUBE = 2,
  VG_ORB = 3,
  VG_SPHERE = 4,
  VG_CONE = 5,
  VG_TRD = 6,
  VG_TRAPEZOID = 7,
  VG_PARALLELEPIPED = 8,
  VG_PARABOLOID = 9,
  VG_TORUS = 10,
  VG_POLYCONE = 11,
  VG_POLYHEDRON = 12,
  VG_BOOLEAN = 13
};

/* Returns 1 if the library has device kernels for this shape type. */
int vg_shape_supported(int32_t shape_type);

/* Selects the kernels used by the basket queries of this process:
 * 0 = the tuned kernels (default); 1 = the plain FP64 form, one thread per
 * query with the placement and shape parameters read from global memory per
 * thread, as the reference's CUDA benchmark kernels do with a device virtual
 * call (source/benchmarking/Benchmarker.cu:13-79).  The plain form exists as
 * the benchmark denominator; results are identical. */
int vg_set_kernel_variant(int variant);

/* ------------------------------------------------------ geometry upload */
/*
 * Flattened volume hierarchy.  Host pointers; the library copies what it
 * needs during vg_geometry_upload.  This replaces the per-object device
 * construction of CudaManager::LoadGeometry + Synchronize
 * (management/CudaManager.h:97-160, source/CudaManager.cpp:39-322) and the
 * VPlacedVolume/VUnplacedVolume/LogicalVolume/Transformation3D CopyToGpu
 * methods (volumes/PlacedVolume.h:283-316, volumes/UnplacedVolume.h:53-78,
 * source/Transformation3D.cpp:206-221, source/LogicalVolume.cpp:151-168).
 *
 * shapes:    shape_type[n_shapes]; constructor parameters of shape s are
 *            params[shape_param_begin[s] .. shape_param_begin[s+1]).
 * logical:   logical_shape[n_logical]; daughters of logical volume l are the
 *            placed ids daughters[logical_daughter_begin[l] .. [l+1]) in
 *            LogicalVolume::PlaceDaughter order (source/LogicalVolume.cpp:84-102).
 * placed:    placed_logical[n_placed]; placed_transform[12*p .. 12*p+12) =
 *            tx ty tz r0..r8 as in Transformation3D(tx,..,r8)
 *            (base/Transformation3D.h:95-101; local = R^T (master - t)).
 * world:     placed id of the world volume (GeoManager::SetWorld).
 */
typedef struct vg_flat_geometry {
  int32_t n_shapes;
  const int32_t *shape_type;
  const int32_t *shape_param_begin; /* n_shapes + 1 */
  const double *params;
  int32_t n_logical;
  const int32_t *logical_shape;
  const int32_t *logical_daughter_begin; /* n_logical + 1 */
  const int32_t *daughters;
  int32_t n_placed;
  const int32_t *placed_logical;
  const double *placed_transform; /* 12 * n_placed */
  int32_t world;
} vg_flat_geometry;

typedef struct vg_geometry vg_geometry;

/* Validates, precomputes cached shape values and bounding boxes, and copies
 * everything to device memory (synchronous). */
int vg_geometry_upload(const vg_flat_geometry *geometry, vg_geometry **out);

/* Frees device memory of a geometry (CudaManager::CleanGpu,
 * source/CudaManager.cpp:173-192). */
int vg_geometry_free(vg_geometry *geometry);

/* Maximum depth of the placed-volume tree (GeoManager::getMaxDepth,
 * management/GeoManager.h:129-200); the world is depth 1. */
int vg_geometry_max_depth(const vg_geometry *geometry, int32_t *depth);

/* Local-frame extent of the shape of placed volume `placed`
 * (VPlacedVolume::Extent -> UnplacedX::Extent, volumes/PlacedVolume.h:260).
 * lo[3], hi[3] host pointers. */
int vg_geometry_placed_extent(const vg_geometry *geometry, int32_t placed, double *lo, double *hi);

/* Axis-aligned bounding box of placed volume `placed` in its mother frame,
 * padded as ABBoxManager::ComputeABBox does (source/ABBoxManager.cpp:128-161).
 * lo[3], hi[3] host pointers. */
int vg_geometry_placed_aabb(const vg_geometry *geometry, int32_t placed, double *lo, double *hi);

/* --------------------------------------------------- basket shape queries */
/* SOA3D basket forms of the VPlacedVolume virtuals
 * (volumes/PlacedVolume.h:182-250; loops in
 * volumes/ShapeImplementationHelper.h:588-748).  `placed` is a placed-volume
 * id of the uploaded geometry.  x,y,z (and dx,dy,dz) are device arrays of
 * length n; n may be any value >= 0.  step_max may be NULL (= +inf, the
 * reference's default argument). */

/* VPlacedVolume::Contains(point, localPoint&) (volumes/PlacedVolume.h:153-156)
 * in basket form: out[i] as Contains, and the point transformed to the
 * volume's local frame in lx, ly, lz (device arrays; each nullable). */
int vg_contains_local(const vg_geometry *g, int32_t placed, const double *x, const double *y, const double *z,
                      size_t n, uint8_t *out, double *lx, double *ly, double *lz, void *stream);
/* VPlacedVolume::UnplacedContains(localPoint) (volumes/PlacedVolume.h:158-160)
 * in basket form: the unplaced shape's Contains of local points. */
int vg_unplaced_contains(const vg_geometry *g, int32_t placed, const double *x, const double *y, const double *z,
                         size_t n, uint8_t *out, void *stream);

/* VPlacedVolume::Normal(localPoint, normal&) (volumes/PlacedVolume.h:266-272)
 * in basket form, for the shapes whose reference overrides it: Box
 * (PlacedBox.h:90, BoxImplementation.h:842-898), Tube (PlacedTube.h:94,
 * source/UnplacedTube.cpp:141), Orb (PlacedOrb.h:95), Sphere
 * (PlacedSphere.h:223), Cone (PlacedCone.h:96), Trd (PlacedTrd.h:96,
 * source/UnplacedTrd.cpp:154), Trapezoid (PlacedTrapezoid.h:111,
 * source/UnplacedTrapezoid.cpp:352) and Polycone (PlacedPolycone.h:87,
 * source/UnplacedPolycone.cpp:684).  x,y,z are LOCAL points; nx,ny,nz get the
 * normal as the reference computes it (also for points on no surface, where
 * the reference returns a zero or approximate vector) and valid[i] its
 * return value (1/0).  VG_ERR_UNSUPPORTED for the other shapes, whose
 * reference Normal asserts. */
int vg_normal(const vg_geometry *g, int32_t placed, const double *x, const double *y, const double *z, size_t n,
              double *nx, double *ny, double *nz, uint8_t *valid, void *stream);
/* The USolids-compatible DistanceToOut(point, direction, normal&, convex&,
 * stepMax) (volumes/ShapeImplementationHelper.h:277-295 and
 * ScalarShapeImplementationHelper.h:249-264; volumes/USolidsInterfaceHelper.h
 * declares it) in basket form on LOCAL rays: out[i] = DistanceToOut, the
 * Normal of point + out[i] * direction in nx,ny,nz with its valid flag, and
 * convex[i] = 1 (the reference's constant; convex may be NULL).  step_max
 * may be NULL (+inf).  Same shapes as vg_normal. */
int vg_distance_to_out_normal(const vg_geometry *g, int32_t placed, const double *x, const double *y,
                              const double *z, const double *dx, const double *dy, const double *dz,
                              const double *step_max, size_t n, double *out, double *nx, double *ny, double *nz,
                              uint8_t *valid, uint8_t *convex, void *stream);

/* VPlacedVolume::Contains(SOA3D, bool*)  (volumes/PlacedVolume.h:147) */
int vg_contains(const vg_geometry *g, int32_t placed, const double *x, const double *y, const double *z, size_t n,
                uint8_t *out, void *stream);

/* EInside codes of Inside (base/Global.h:198-205). */
#define VG_INSIDE_INSIDE 0
#define VG_INSIDE_SURFACE 1
#define VG_INSIDE_OUTSIDE 2

/* VPlacedVolume::Inside(SOA3D, Inside_t*)  (volumes/PlacedVolume.h:165) */
int vg_inside(const vg_geometry *g, int32_t placed, const double *x, const double *y, const double *z, size_t n,
              int32_t *out, void *stream);

/* VPlacedVolume::DistanceToIn(SOA3D, SOA3D, step_max*, out*)  (volumes/PlacedVolume.h:176) */
int vg_distance_to_in(const vg_geometry *g, int32_t placed, const double *x, const double *y, const double *z,
                      const double *dx, const double *dy, const double *dz, const double *step_max, size_t n,
                      double *out, void *stream);

/* VPlacedVolume::DistanceToInMinimize  (volumes/PlacedVolume.h:181; semantics
 * ShapeImplementationHelper.h:637-658): result = DistanceToIn(p, d,
 * inout_dist[i]); if result < inout_dist[i] then inout_dist[i] = result and
 * inout_next[i] = daughter_index. */
int vg_distance_to_in_minimize(const vg_geometry *g, int32_t placed, int32_t daughter_index, const double *x,
                               const double *y, const double *z, const double *dx, const double *dy, const double *dz,
                               size_t n, double *inout_dist, int32_t *inout_next, void *stream);

/* VPlacedVolume::DistanceToOut(SOA3D, SOA3D, step_max*, out* [, next*])
 * (volumes/PlacedVolume.h:212-221).  Local-frame points.  When next_node is not
 * NULL it receives -1 where out[i] < step_max[i] and -2 otherwise
 * (ShapeImplementationHelper.h:677-692). */
int vg_distance_to_out(const vg_geometry *g, int32_t placed, const double *x, const double *y, const double *z,
                       const double *dx, const double *dy, const double *dz, const double *step_max, size_t n,
                       double *out, int32_t *next_node, void *stream);

/* VPlacedVolume::PlacedDistanceToOut (volumes/PlacedVolume.h:207) in basket form:
 * mother-frame points, transformed inside. */
int vg_placed_distance_to_out(const vg_geometry *g, int32_t placed, const double *x, const double *y, const double *z,
                              const double *dx, const double *dy, const double *dz, const double *step_max, size_t n,
                              double *out, void *stream);

/* VPlacedVolume::SafetyToIn(SOA3D, out*)  (volumes/PlacedVolume.h:231) */
int vg_safety_to_in(const vg_geometry *g, int32_t placed, const double *x, const double *y, const double *z, size_t n,
                    double *out, void *stream);

/* VPlacedVolume::SafetyToInMinimize  (volumes/PlacedVolume.h:237; semantics
 * ShapeImplementationHelper.h:707-720): inout[i] = min(inout[i], SafetyToIn). */
int vg_safety_to_in_minimize(const vg_geometry *g, int32_t placed, const double *x, const double *y, const double *z,
                             size_t n, double *inout, void *stream);

/* VPlacedVolume::SafetyToOut(SOA3D, out*)  (volumes/PlacedVolume.h:243); local frame. */
int vg_safety_to_out(const vg_geometry *g, int32_t placed, const double *x, const double *y, const double *z, size_t n,
                     double *out, void *stream);

/* VPlacedVolume::SafetyToOutMinimize  (volumes/PlacedVolume.h:249; semantics
 * ShapeImplementationHelper.h:733-746). */
int vg_safety_to_out_minimize(const vg_geometry *g, int32_t placed, const double *x, const double *y, const double *z,
                              size_t n, double *inout, void *stream);

/* ------------------------------------------- tasks, batches, FP32 storage */
/*
 * A task is one basket query on one placed volume.  Query codes: */
#define VG_Q_CONTAINS 0               /* out: uint8  (VPlacedVolume::Contains)      */
#define VG_Q_INSIDE 1                 /* out: int32  (Inside)                        */
#define VG_Q_DISTANCE_TO_IN 2         /* out: double (DistanceToIn, mother frame)    */
#define VG_Q_DISTANCE_TO_OUT 3        /* out: double (DistanceToOut, local frame)    */
#define VG_Q_PLACED_DISTANCE_TO_OUT 4 /* out: double (PlacedDistanceToOut)           */
#define VG_Q_SAFETY_TO_IN 5           /* out: double (SafetyToIn, mother frame)      */
#define VG_Q_SAFETY_TO_OUT 6          /* out: double (SafetyToOut, local frame)      */
/* The scalar VPlacedVolume virtuals (volumes/PlacedVolume.h:144-250):
 * one point (and direction) in host memory, evaluated on the device by the
 * basket kernels, synchronously (the call returns with the result).
 * query: a VG_Q_* code above, or VG_Q_CONTAINS_LOCAL (Contains(point,
 * localPoint&): local receives the local point) or VG_Q_UNPLACED_CONTAINS
 * (point is local).  *result: Contains 0/1, Inside the EInside code,
 * otherwise the distance or safety.  dir is read by the distance queries
 * only.  Calls on one geometry are serialised. */
#define VG_Q_CONTAINS_LOCAL 7
#define VG_Q_UNPLACED_CONTAINS 8
int vg_scalar_query(vg_geometry *g, int32_t placed, int32_t query, const double point[3], const double dir[3],
                    double step_max, double *result, double local[3]);

/* Input element types: VG_F64 = double arrays (the SOA3D<double> of the
 * reference); VG_F32 = float arrays (FP32 storage: every value is promoted to
 * double before any arithmetic, so results equal the FP64 form on the same
 * float-valued inputs); VG_MIXED = float arrays with FP32 arithmetic first
 * (subject to vg_set_mixed_policy below):
 * Contains and Inside are decided in FP32 wherever an FP32 lower bound on
 * the point's distance to the surface exceeds its rigorous rounding margin,
 * DistanceToIn returns kInfinity in FP32 for rays that provably miss the
 * shape's extent box, and every other query (and every undecided lane) is
 * evaluated by the FP64 kernel on the promoted inputs -- so every result
 * equals the VG_F32 form's, bit for bit.  Outputs are as for the FP64 entry
 * points. */
#define VG_F64 0
#define VG_F32 1
#define VG_MIXED 2

/* Which VG_MIXED tasks run the FP32 filter.  VG_MIXED_POLICY_MEASURED (the
 * default): only the (shape, query) pairs where the filter was measured to
 * save time on MI355X (an FP64 FMA issues at the rate of an unpacked FP32
 * one on gfx950, so the filter pays only where it skips long FP64 work);
 * the other VG_MIXED tasks run as VG_F32.  VG_MIXED_POLICY_ALL: every
 * Contains, Inside and DistanceToIn task.  Results are the same under
 * either policy; it applies to batches created and queries issued after
 * the call.  No counterpart in the reference. */
#define VG_MIXED_POLICY_MEASURED 0
#define VG_MIXED_POLICY_ALL 1
int vg_set_mixed_policy(int32_t policy);

typedef struct vg_query_task {
  int32_t placed;     /* placed-volume id */
  int32_t query;      /* VG_Q_* */
  int32_t input_type; /* VG_F64 or VG_F32 */
  int32_t reserved;
  uint64_t n;
  const void *x, *y, *z;    /* device arrays of input_type, length n */
  const void *dx, *dy, *dz; /* directions (distance queries only) */
  const double *step_max;   /* distance queries; NULL = kInfinity */
  void *out;                /* device array, see the query codes */
  uint64_t *fp64_count;     /* nullable device counter: each run adds the
                               number of queries it evaluated in FP64 (all of
                               them for VG_F64 / VG_F32; the undecided ones
                               for VG_MIXED).  Costs one atomic per wave. */
} vg_query_task;

/* One task on `stream`: the SOA3D basket virtuals of VPlacedVolume
 * (volumes/PlacedVolume.h:147-246) with a choice of input storage.  No
 * counterpart in the reference for VG_F32 inputs. */
int vg_query(const vg_geometry *g, const vg_query_task *task, void *stream);

/*
 * A batch: a fixed table of tasks (any mix of placed volumes, queries, basket
 * sizes and input types) run by one kernel launch per shape weight class and
 * precision (at most ten, issued on forked streams and joined back to the
 * caller's), each
 * workgroup tile belonging to one task.  This replaces the per-shape, per-query launches of
 * Benchmarker::Run*Cuda (source/benchmarking/Benchmarker.cu:86-367).  The
 * arrays named in the tasks must stay valid while the batch is used; the
 * launch reads whatever they hold when it runs, so a caller refills inputs
 * and runs again (also from a captured hipGraph: vg_batch_run enqueues its
 * class kernels and the fork/join events, nothing else).
 */
typedef struct vg_batch vg_batch;
int vg_batch_create(const vg_geometry *g, const vg_query_task *tasks, int32_t ntasks, vg_batch **out);
int vg_batch_run(vg_batch *b, void *stream);
/* Launch shape: items per lane (1, 2, 4) and the number of persistent
 * workgroups (0 = one workgroup per tile).  Each items-per-lane value has
 * its own device task table, written once by vg_batch_create, so changing
 * the launch shape rewrites no device memory: launches in flight and
 * hipGraphs captured earlier keep their own shape (capture again to replay
 * the new one). */
int vg_batch_set_launch(vg_batch *b, int32_t items_per_lane, uint32_t blocks);
int vg_batch_get_launch(const vg_batch *b, int32_t *items_per_lane, uint32_t *blocks, uint64_t *items);
/* Times every launch shape on the batch itself (`repetitions` runs each) and
 * keeps the fastest, as vg_batch_set_launch would; *best_ms (nullable)
 * receives its time per run.  Runs the batch, so outputs are written; waits
 * for its own timing events. */
int vg_batch_autotune(vg_batch *b, int32_t repetitions, void *stream, float *best_ms);
/* Per-class timing of later vg_batch_run calls: when enabled, HIP events are
 * recorded around each class kernel on the stream it is launched on (the
 * last 64 runs are kept).  Enabling clears earlier records. */
int vg_batch_set_timing(vg_batch *b, int32_t enable);
/* Weight class `cls` (0 light, 1 Cone/Sphere, 2 Torus, 3 Polycone/Polyhedron
 * staged in LDS, 4 the same read from global memory; 5-9 the same classes
 * for VG_MIXED tasks, which run in kernels of their own): its queries per run,
 * its algorithmic bytes per run (each input element read once, each output
 * written once), and the number of timed runs and their mean kernel
 * duration (waits for the recorded events).  Each pointer is nullable. */
int vg_batch_class_stats(vg_batch *b, int32_t cls, uint64_t *items, uint64_t *bytes, int32_t *runs, float *mean_ms);
int vg_batch_free(vg_batch *b);

/* The tasks of a batch run once on `stream` without allocating anything:
 * no hipMalloc / hipFree, no streams or events created; the task table goes
 * to a static device array (one per device, 256 tasks at most -- more is
 * VG_ERR_CAPACITY before anything is launched) and the class kernels run one
 * after another on the stream.  Results equal vg_batch_create + vg_batch_run
 * on the same tasks.  Calls on one device share its table: calls whose work
 * overlaps on one device (on different streams, or from several threads)
 * race, and each call rewrites the whole table, so nothing carries over.
 * No counterpart in the reference: for callers that cannot allocate (code
 * under graph capture, serving loops, harnesses that forbid hipMalloc). */
int vg_batch_run_noalloc(const vg_geometry *g, const vg_query_task *tasks, int32_t ntasks, void *stream);

/* -------------------------------------------------------------- navigation */
/*
 * Navigation states of a basket of tracks, structure-of-arrays in device
 * memory.  Replaces NavStatePool (navigation/NavStatePool.h:34-110), whose
 * states are NavigationState objects (navigation/NavigationState.h) packed
 * array-of-structures; here
 *   level[i]                 number of volumes on the path of track i
 *                            (0: outside the world)
 *   path[l * capacity + i]   placed id at depth l (0 = world), -1 beyond
 *                            level[i]
 *   boundary[i]              1 if the state is on a boundary
 * Any device buffers of these sizes may be used (torch tensors, for one);
 * vg_nav_states_create allocates them.
 */
typedef struct vg_nav_states {
  int32_t *level;
  int32_t *path;
  uint8_t *boundary;
  uint64_t capacity;
  int32_t depth; /* path rows; at least vg_geometry_max_depth() */
  int32_t reserved;
} vg_nav_states;

/* Navigator selection (navigation/SimpleNavigator.h:341-482,
 * navigation/ABBoxNavigator.h:201-424).  VG_NAV_ABBOX_MIXED is the ABBox
 * navigator with an FP32 filter in its daughter walk: a daughter whose FP32
 * lower bound on the distance from the point (with a rigorous rounding
 * margin) exceeds the best step found so far is not evaluated in FP64.  Its
 * steps, next volumes and states are those of VG_NAV_ABBOX, bit for bit; no
 * counterpart in the reference. */
#define VG_NAV_ABBOX 0
#define VG_NAV_SIMPLE 1
#define VG_NAV_ABBOX_MIXED 2
/* SimpleNavigator::FindNextBoundaryAndStep for baskets
 * (navigation/SimpleNavigator.h:638-743): the basket's tracks share one
 * current volume, taken as the top of the first track's state; the
 * volume's DistanceToOut with pstep as step limit, DistanceToInMinimize over
 * its daughters, push by 1e-9, no repair of steps <= 0 or kInfinity.  For
 * vg_find_next_boundary_and_step only (the trace and serving loops step
 * each track in its own volume). */
#define VG_NAV_SIMPLE_BASKET 3

/* Largest geometry depth the device navigator handles. */
#define VG_NAV_MAX_DEPTH 16

/* NavStatePool(size, depth): allocates the three arrays on the current
 * device (synchronous).  depth <= 0 takes the geometry's depth. */
int vg_nav_states_create(const vg_geometry *g, uint64_t capacity, int32_t depth, vg_nav_states *out);
/* Frees arrays made by vg_nav_states_create. */
int vg_nav_states_free(vg_nav_states *s);
/* NavStatePool::CopyFromGpu / CopyToGpu (NavStatePool.h:62-63): copies the
 * first n states to / from host arrays of the same layout, with row stride
 * n on the host side (synchronous). */
int vg_nav_states_to_host(const vg_nav_states *s, uint64_t n, int32_t *level, int32_t *path, uint8_t *boundary);
int vg_nav_states_from_host(vg_nav_states *s, uint64_t n, const int32_t *level, const int32_t *path,
                            const uint8_t *boundary);

/* SimpleNavigator::LocatePoint(world, point, state, top=true)
 * (navigation/SimpleNavigator.h:205-241) for n global points; the states
 * are cleared first (boundary = 0). */
int vg_locate_points(const vg_geometry *g, const double *x, const double *y, const double *z, size_t n,
                     vg_nav_states *states, void *stream);

/* FindNextBoundaryAndStep(globalpoint, globaldir, currentstate, newstate,
 * pstep, step) of the selected navigator (ABBoxNavigator.h:201-424 /
 * SimpleNavigator.h:341-482) for n tracks.  pstep may be NULL (kInfinity
 * for all).  `cur` and `next` must not alias.  A state outside the world
 * (level 0) stays outside with step 0; a state whose level exceeds
 * VG_NAV_MAX_DEPTH or whose path holds an id that is not a placed volume is
 * copied unchanged with step NaN. */
int vg_find_next_boundary_and_step(const vg_geometry *g, int32_t navigator, const double *x, const double *y,
                                   const double *z, const double *dx, const double *dy, const double *dz,
                                   const double *pstep, size_t n, const vg_nav_states *cur, vg_nav_states *next,
                                   double *step, void *stream);

/* ABBoxNavigator::GetSafety (navigation/ABBoxNavigator.h:428-452); 0 for
 * states outside the world, NaN for malformed states. */
int vg_get_safety(const vg_geometry *g, const double *x, const double *y, const double *z, size_t n,
                  const vg_nav_states *cur, double *safety, void *stream);

/*
 * The track loop of the reference's TraceTrack benchmark
 * (test/globalbenchmarks/TraceTrack.cpp:170-220) run on the device, one
 * track per lane: LocatePoint from the world, then FindNextBoundaryAndStep
 * with `pstep` and p += d * (step + 1e-6) until the state leaves the world
 * or max_steps steps were taken.  Per track: steps taken, final point,
 * summed step length, and a 64-bit FNV-1a-style hash over the sequence of
 * (new top placed id + 1) * 2 + boundary flag, starting from
 * 1469598103934665603 with multiplier 1099511628211.  exact_steps (nullable)
 * receives, per track, the number of ABBox steps that took the exact
 * sorted-order path instead of the order-free one.
 */
int vg_trace_tracks(const vg_geometry *g, int32_t navigator, const double *x, const double *y, const double *z,
                    const double *dx, const double *dy, const double *dz, size_t n, double pstep, int32_t max_steps,
                    int32_t *nsteps, double *fx, double *fy, double *fz, double *travelled, uint64_t *hash,
                    int32_t *exact_steps, void *stream);

/*
 * ---------------------------------------------------------------- serving
 * A navigation server with continuous batching.  The reference drives
 * baskets of tracks through its navigators from its own event loop
 * (test/globalbenchmarks/TraceTrack.cpp:170-220 steps one track at a time;
 * the basket navigation entry points are navigation/NavigatorBase and
 * NavStatePool).  Here a client opens a basket (its tracks become resident
 * on the device together with their navigation states), then submits
 * requests; each request is one step of the TraceTrack loop for every track
 * of the basket: FindNextBoundaryAndStep with the server's pstep, then
 * p += d * (step + 1e-6), plus the step count and the running hash of
 * vg_trace_tracks.  A background thread merges all pending requests into one
 * launch whenever one of two device streams is free, so requests that arrive
 * while others run are batched into the next launch.  Requests for one
 * basket run in submission order; N requests on a basket give exactly the
 * state vg_trace_tracks reaches after N steps.
 *
 * Memory: the track arena is allocated once at create and sized by
 * hbm_budget_bytes (8 * 8 + 3 * 4 + 4 * depth bytes per track); a basket is
 * admitted only if a contiguous run of free slots holds it
 * (VG_ERR_CAPACITY otherwise).  The geometry must outlive the server.
 */
typedef struct vg_server vg_server;

int vg_server_create(const vg_geometry *g, uint64_t hbm_budget_bytes, int32_t navigator, double pstep,
                     vg_server **out);
/* Arena slots in total and currently free. */
int vg_server_capacity(const vg_server *s, uint64_t *tracks_total, uint64_t *tracks_free);
/* Uploads n tracks from host arrays and locates them from the world
 * (synchronous); returns the basket id. */
int vg_server_open(vg_server *s, uint64_t n, const double *x, const double *y, const double *z, const double *dx,
                   const double *dy, const double *dz, int64_t *basket);
/* Queues one navigation step of every track of the basket. */
int vg_server_submit(vg_server *s, int64_t basket, int64_t *request);
/* Blocks until the request has completed; latency_s (nullable) receives the
 * time from submission to completion.  A request id is released by the wait
 * (or by the poll) that reports it complete. */
int vg_server_wait(vg_server *s, int64_t request, double *latency_s);
/* Non-blocking: *done = 1 (and latency_s set) once complete. */
int vg_server_poll(vg_server *s, int64_t request, int32_t *done, double *latency_s);
/* Number of tracks of an open basket. */
int vg_server_basket_size(vg_server *s, int64_t basket, uint64_t *n);
/* Copies the basket's current track states to host arrays (each nullable;
 * synchronous).  `capacity` is the number of entries every non-null array
 * holds: VG_ERR_CAPACITY (nothing copied) if the basket has more tracks.
 * Fails with VG_ERR_STATE while a step of the basket runs. */
int vg_server_read(vg_server *s, int64_t basket, uint64_t capacity, double *x, double *y, double *z, int32_t *level,
                   int32_t *boundary, int32_t *nsteps, uint64_t *hash, double *last_step);
/* Frees the basket's slots; VG_ERR_STATE while it has queued or running steps. */
int vg_server_close(vg_server *s, int64_t basket);
/* Launches made, requests they carried, and track steps taken. */
int vg_server_stats(vg_server *s, uint64_t *launches, uint64_t *requests_launched, uint64_t *track_steps);
/* Launch policy: at most max_requests_per_launch requests (1..4096) merged
 * into one launch, on 1 or 2 streams.  (1, 1) is the per-basket form: one
 * launch per request, one at a time.  Default (4096, 2).  Fails with
 * VG_ERR_STATE while requests are queued or running. */
int vg_server_set_batching(vg_server *s, int32_t max_requests_per_launch, int32_t streams);
/* In-process load generator (the transport client of a serving benchmark):
 * basket i arrives arrival_s[i] seconds after the call (sorted), waits for
 * admission by footprint (FIFO; VG_ERR_CAPACITY only if a basket exceeds the
 * whole arena), is opened from its tracks -- the next sizes[i] entries of
 * the concatenated x..dz arrays -- then takes steps_per_basket steps, each
 * one request submitted when the previous completes, and is closed.
 * latency_s[i * steps_per_basket + k] receives step k's latency;
 * admit_wait_s[i] (nullable) the time from arrival to admission; wall_s
 * (nullable) the elapsed time until the last basket closed. */
int vg_server_replay(vg_server *s, uint64_t nbaskets, const double *arrival_s, const uint64_t *sizes,
                     int32_t steps_per_basket, const double *x, const double *y, const double *z, const double *dx,
                     const double *dy, const double *dz, double *latency_s, double *admit_wait_s, double *wall_s);
/* Drains queued requests, stops the thread and frees the arena. */
int vg_server_destroy(vg_server *s);

#ifdef __cplusplus
} /* extern "C" */
#endif

#endif /* VECGEOM_AMD_H_ */
