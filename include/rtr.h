/*
 * librtr -- MI355X (gfx950) ICP registration and voxel fusion, C ABI.
 *
 * Drop-in boundary for the registration path of ICCD/RealTime_Robot.  The
 * reference reaches it from keyPointICP (RealTimeRobot/function.h:111-148):
 *
 *     pcl::IterativeClosestPoint<pcl::PointXYZ, pcl::PointXYZ> icp;  // :112
 *     icp.setInputCloud(mcloud);                                      // :113
 *     icp.setInputTarget(cloud);                                      // :114
 *     icp.align(Final);                                               // :117
 *     icp.getFinalTransformation();                                   // :119,123
 *     pcl::transformPointCloud(*mcloud, *transformed, T);             // :123
 *
 * and merges transformed clouds into the scene map at
 * RealTimeRobot.cpp:104-114.  Each entry point below names the call it
 * replaces.  Conventions:
 *   - every function returns an int status (RTR_OK = 0); rtr_last_error()
 *     gives a thread-local message.  No C++ exception crosses the ABI;
 *   - point arrays are host pointers to float xyz with a byte stride
 *     (12 = packed xyz, 16 = pcl::PointXYZ with its padding);
 *   - 4x4 transforms are double[16], column-major (Eigen's layout);
 *   - a handle is not thread-safe; use one handle per host thread.  Each
 *     handle owns one HIP stream on its device.
 */
#ifndef RTR_H
#define RTR_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define RTR_OK 0
#define RTR_ERR_INVALID_ARG 1
#define RTR_ERR_EMPTY_INPUT 2
#define RTR_ERR_NO_CORRESPONDENCES 3
#define RTR_ERR_CAPACITY 4
#define RTR_ERR_DEVICE 5
#define RTR_ERR_NOT_READY 6
#define RTR_ERR_IO 7
#define RTR_ERR_FORMAT 8

/* correspondence metric */
#define RTR_MODE_POINT_TO_POINT 0
#define RTR_MODE_POINT_TO_PLANE 1

/* convergence states (PCL DefaultConvergenceCriteria::ConvergenceState) */
#define RTR_NOT_CONVERGED 0
#define RTR_CONV_ITERATIONS 1
#define RTR_CONV_TRANSFORM 2
#define RTR_CONV_ABS_MSE 3
#define RTR_CONV_REL_MSE 4
#define RTR_FAIL_NO_CORRESPONDENCES 5

/* Registration parameters.  rtr_params_default() gives PCL 1.8's defaults:
 * 10 iterations, max correspondence distance sqrt(DBL_MAX), transformation
 * epsilon 0, euclidean fitness epsilon -DBL_MAX, 3 minimum correspondences,
 * point-to-point. */
typedef struct {
  int max_iterations;
  int min_correspondences;
  int mode;                           /* RTR_MODE_* */
  int anderson_memory;                /* 0 (default): PCL's plain iteration.  1..8: Anderson
                                         acceleration with that memory (no PCL counterpart): the
                                         iterate is extrapolated from the last m + 1 ICP steps;
                                         convergence is judged on the plain ICP increment (the
                                         transformation criterion) and the iteration limit only */
  double max_correspondence_distance; /* metres */
  double transformation_epsilon;      /* squared translation threshold; rotation 1-eps */
  double rotation_epsilon;            /* cos threshold; <= 0 means 1 - transformation_epsilon */
  double euclidean_fitness_epsilon;   /* relative MSE threshold */
  float cell_size;                    /* grid cell edge; 0 = autotuned */
  float cells_per_point;              /* grid cell budget per target point; 0 = autotuned */
  int coarse_stride;                  /* 0 or 1 (default): every iteration uses every source point.
                                         s > 1: coarse-to-fine (no PCL counterpart): iterations
                                         start on every s-th source point (in the engine's
                                         spatial order) and switch to all of them once an
                                         increment moves the source by under 1 mm and 1 mrad, or
                                         after coarse_iterations or half of max_iterations; only
                                         full-resolution iterations can converge, so the result
                                         is a fixed point of the full iteration */
  int coarse_iterations;              /* most coarse iterations (with coarse_stride > 1) */
} rtr_params;

/* Outcome of one registration. */
typedef struct {
  double T[16];        /* final transform, column-major */
  double mse;          /* mean squared distance of the last iteration's correspondences */
  double fitness_rmse; /* sqrt(mse): the run-time quality measure */
  int64_t n_corr;      /* correspondences of the last iteration */
  int iterations;
  int converged;       /* PCL hasConverged(): criteria fired (incl. max iterations) */
  int state;           /* RTR_CONV_* or RTR_FAIL_* */
  int coarse_iterations;    /* of those, iterations on the coarse source subsample (rtr_params.coarse_stride) */
} rtr_result;

/* Phase timings of the last align, from HIP events on the handle's stream
 * (filled only when profiling is enabled). */
typedef struct {
  float ms_total;      /* whole align on the device */
  float ms_grid;       /* target grid build + source spatial sort */
  float ms_icp;        /* persistent correspondence + reduction + solve kernel */
  float ms_finish;     /* output transforms / fitness */
  int64_t icp_launches;
  int64_t iterations_total; /* sum of iterations over problems */
  int64_t queries_searched; /* nearest-neighbour searches of the launch (every align) */
} rtr_timing;

const char* rtr_last_error(void);
int rtr_version(void);
int rtr_params_default(rtr_params* p);

/* ------------------------------------------------------------------ */
/* Single registration: pcl::IterativeClosestPoint<PointXYZ, PointXYZ> */
/* ------------------------------------------------------------------ */
typedef struct rtr_icp rtr_icp;

/* `IterativeClosestPoint<PointXYZ,PointXYZ> icp;` (function.h:112) */
int rtr_icp_create(rtr_icp** out, int device);
int rtr_icp_destroy(rtr_icp* h);
/* `icp.setInputCloud(mcloud)` / setInputSource (function.h:113).  Copies the points. */
int rtr_icp_set_input_source(rtr_icp* h, const float* xyz, size_t n, size_t stride_bytes);
/* `icp.setInputTarget(cloud)` (function.h:114).  Copies; the grid is rebuilt lazily. */
int rtr_icp_set_input_target(rtr_icp* h, const float* xyz, size_t n, size_t stride_bytes);
/* Target normals for point-to-plane (PCL PointNormal target, normal_x/y/z fields). */
int rtr_icp_set_target_normals(rtr_icp* h, const float* nxyz, size_t n, size_t stride_bytes);
/* pcl::NormalEstimation with setKSearch(k) and viewpoint vp on the current target. */
int rtr_icp_compute_target_normals(rtr_icp* h, int k, const float vp[3]);
/* Copy out the target normals in target order (n_target x 3 floats). */
int rtr_icp_get_target_normals(rtr_icp* h, float* nxyz);
/* setMaxCorrespondenceDistance / setMaximumIterations / setTransformationEpsilon /
 * setEuclideanFitnessEpsilon, and the whole parameter block at once. */
int rtr_icp_set_max_correspondence_distance(rtr_icp* h, double d);
int rtr_icp_set_maximum_iterations(rtr_icp* h, int n);
int rtr_icp_set_transformation_epsilon(rtr_icp* h, double e);
int rtr_icp_set_euclidean_fitness_epsilon(rtr_icp* h, double e);
int rtr_icp_set_params(rtr_icp* h, const rtr_params* p);
int rtr_icp_get_params(rtr_icp* h, rtr_params* p);
/* `icp.align(Final[, guess])` (function.h:117) followed by getFinalTransformation()
 * (:119,123) and hasConverged().  guess may be NULL (identity).  out_xyz (nullable,
 * n_source points at out_stride bytes) receives transformPointCloud(source, T)
 * (function.h:123).  final_T and converged are nullable. */
int rtr_icp_align(rtr_icp* h, const double guess[16], float* out_xyz, size_t out_stride,
                  double final_T[16], int* converged);
/* Full outcome of the last align (iterations, state, mse, correspondences). */
int rtr_icp_get_result(rtr_icp* h, rtr_result* out);
/* Registration::getFitnessScore(max_range): mean NN squared distance of the
 * transformed source over points with d2 <= max_range; DBL_MAX if none. */
int rtr_icp_get_fitness_score(rtr_icp* h, double max_range, double* out);
/* CorrespondenceEstimation::determineCorrespondences for the source under T:
 * idx[i] = target index (or -1 if d2 > max_dist^2), d2[i] = squared distance
 * (+inf when rejected).  Ties go to the lowest target index. */
int rtr_icp_get_correspondences(rtr_icp* h, const double T[16], int32_t* idx, float* d2);
/* Origin used for the accumulated moments (point-to-plane linearisation point). */
int rtr_icp_get_anchor(rtr_icp* h, float anchor[3]);
int rtr_icp_set_profiling(rtr_icp* h, int on);
int rtr_icp_get_timing(rtr_icp* h, rtr_timing* out);

/* ------------------------------------------------------------------ */
/* Batched registration: many (source, target, guess) problems in one  */
/* persistent launch over a work queue, replayed from a hipGraph.       */
/* No reference counterpart (the reference registers one pair).        */
/* ------------------------------------------------------------------ */
typedef struct rtr_batch rtr_batch;

int rtr_batch_create(rtr_batch** out, int device);
int rtr_batch_destroy(rtr_batch* b);
/* Upload clouds once; they stay resident in HBM.  src[i]/tgt[j] are packed xyz. */
int rtr_batch_set_clouds(rtr_batch* b, int n_src, const float* const* src, const size_t* ns,
                         int n_tgt, const float* const* tgt, const size_t* nt);
/* Optional per-target normals (packed nxyz, same counts as the targets). */
int rtr_batch_set_target_normals(rtr_batch* b, const float* const* nrm);
/* Compute k-NN PCA normals for every target on the device. */
int rtr_batch_compute_target_normals(rtr_batch* b, int k, const float vp[3]);
/* Problems: problem p registers source src_id[p] onto target tgt_id[p]. */
int rtr_batch_set_problems(rtr_batch* b, int n, const int* src_id, const int* tgt_id);
int rtr_batch_set_params(rtr_batch* b, const rtr_params* p);
/* Launch/layout tuning (the autotuner's knobs): requested cell edge as a
 * fraction of the max correspondence distance, cell budget per target point,
 * source points per work item (multiple of 256), persistent workgroups per CU
 * (0 = occupancy limit), LDS candidate tiles on (1) or off (0).  Values < 0
 * (and <= 0 for the first three) keep the current setting. */
int rtr_batch_set_tuning(rtr_batch* b, float cell_factor, float cells_per_point, int item_pts,
                         int blocks_per_cu, int use_tile);
/* Early exit of surplus persistent workgroups: those beyond one per CU leave
 * at the start of an iteration once at most below * n_problems problems are
 * unfinished (0 = never, the default; below in [0, 1]).  Many co-resident
 * workgroups raise throughput while most problems iterate; once few remain,
 * idle workgroups polling the work queue stretch their iterations. */
int rtr_batch_set_retire(rtr_batch* b, float below);
/* Work-item granularity of the late iterations: from iteration `from` on
 * (counted from 0 within one align), each problem's iteration is split into
 * 256-point work items instead of item_pts-point ones, so the few problems
 * still iterating at the end of a batch spread over more workgroups
 * (0 = never, the default).  Changes the summation grouping of the moment
 * sums (FP64), not the correspondences. */
int rtr_batch_set_fine_items(rtr_batch* b, int from);
int rtr_batch_get_fine_items(rtr_batch* b, int* from);
int rtr_batch_get_retire(rtr_batch* b, float* below);
/* The current tuning (any pointer may be null). */
int rtr_batch_get_tuning(rtr_batch* b, float* cell_factor, float* cells_per_point, int* item_pts,
                         int* blocks_per_cu, int* use_tile);
/* Grid policy.  always = 1 (default): every align rebuilds the target grids
 * and source orders inside its launch sequence (a registration pays for its
 * own search structure, as PCL's setInputTarget builds its kd-tree).
 * always = 0: grids are built by the first align after the clouds,
 * parameters or tuning change, or by rtr_batch_build, and reused by later
 * aligns, also across rtr_batch_set_problems. */
int rtr_batch_set_rebuild(rtr_batch* b, int always);
/* Build the target grids and source orders now. */
int rtr_batch_build(rtr_batch* b);
/* Enable/disable hipGraph replay of align (default on). */
int rtr_batch_set_graph(rtr_batch* b, int on);
int rtr_batch_set_profiling(rtr_batch* b, int on);
/* Run every problem from guesses (n*16 doubles, nullable = identity);
 * results (nullable, n entries). */
int rtr_batch_align(rtr_batch* b, const double* guesses, rtr_result* results);
int rtr_batch_get_timing(rtr_batch* b, rtr_timing* out);
/* Grid parameters chosen for target j: cell edge, dims, anchor. */
int rtr_batch_get_grid(rtr_batch* b, int tgt, float* cell, int dims[3], float anchor[3]);
/* Transforms each iteration of problem `prob` queried with, as float 3x4 rows
 * (12 floats per entry): entry i for iteration i, entry `iterations` the final
 * one.  *n_out = iterations + 1. */
int rtr_batch_get_history(rtr_batch* b, int prob, float* rows12, int cap, int* n_out);
/* Launch timeline (diagnostic; off by default): with on = 1, each align
 * records, per problem and iteration, the device wall clock at the start and
 * at the end of the iteration's solve (the last item of the iteration done;
 * the next iteration published).  rtr_batch_get_timeline copies problem
 * `prob`'s 2 * iterations stamps (ticks of *tick_ns nanoseconds). */
int rtr_batch_set_timeline(rtr_batch* b, int on);
int rtr_batch_get_timeline(rtr_batch* b, int prob, uint64_t* stamps, int cap, int* n_out, double* tick_ns);
/* Correspondences of the last iteration the latest align ran for problem
 * `prob`, as its search found them (parity hook; the per-iteration
 * correspondences PCL's ICP keeps, function.h:117): idx[i] the nearest target
 * point of source point i and d2[i] its float squared distance, or -1 / +inf
 * when none lies within the max correspondence distance.  Arrays of the
 * source's size.  Recorded by the global-walk search (use_tile 0) only. */
int rtr_batch_get_last_correspondences(rtr_batch* b, int prob, int32_t* idx, float* d2);
/* Fitness score of every problem at its final transform (see rtr_icp_get_fitness_score). */
int rtr_batch_fitness(rtr_batch* b, double max_range, double* out);
/* Registration judge at each problem's final transform.  A strided sample of
 * up to max_points source points is moved by the final pose; a point is an
 * inlier when its nearest target point is within the max correspondence
 * distance.  For each inlier the k nearest target points are fitted with a
 * plane (PCA); where that neighbourhood is planar (surface variation
 * l0/(l0+l1+l2) <= max_curvature) the inlier's distance to the plane counts.
 * Outputs (each nullable): plane_rmse = RMS of those distances (DBL_MAX with
 * none), inlier_frac = inliers / sampled points, planar_frac = planar inliers
 * / inliers.  No reference counterpart: it is the run-time acceptance test
 * that decides whether a learned proposal needs the multi-start fallback. */
int rtr_batch_judge(rtr_batch* b, int k, int max_points, double max_curvature, double* plane_rmse,
                    double* inlier_frac, double* planar_frac);

/* ------------------------------------------------------------------ */
/* Voxel-hash map: merging transformed clouds into the scene map.       */
/* Replaces transformPointCloud + savePCDFileASCII + viewer overlay     */
/* (RealTimeRobot.cpp:104-114).  key = floor(p / voxel) per axis.       */
/* ------------------------------------------------------------------ */
typedef struct rtr_map rtr_map;

int rtr_map_create(rtr_map** out, int device, float voxel_size, size_t capacity);
int rtr_map_destroy(rtr_map* m);
int rtr_map_clear(rtr_map* m);
/* Insert T * xyz (T nullable = identity): per voxel FP32 sum of xyz and count.
 * The merge of one transformed cloud into the map (RealTimeRobot.cpp:105-109:
 * transformPointCloud, then the merged overlay). */
int rtr_map_integrate(rtr_map* m, const float* xyz, size_t n, size_t stride_bytes,
                      const double T[16]);
/* The same for n_frames clouds in one launch, frame f moved by T + 16 f
 * (column-major; T nullable = identity for all). */
int rtr_map_integrate_frames(rtr_map* m, int n_frames, const float* const* xyz, const size_t* n,
                             size_t stride_bytes, const double* T);
/* Device time of the integration kernel of the last integrate call (ms),
 * recorded while profiling is on. */
int rtr_map_set_profiling(rtr_map* m, int on);
int rtr_map_last_ms(rtr_map* m, float* ms);
int rtr_map_size(rtr_map* m, size_t* n_voxels);
/* Copy out occupied voxels sorted by key: keys (packed 21 bits/axis, +2^20
 * bias), sums (3 per voxel) or centroids when as_centroids != 0, counts. */
int rtr_map_extract(rtr_map* m, int64_t* keys, float* xyz, uint32_t* counts, size_t cap,
                    size_t* n_out, int as_centroids);
/* Points dropped so far (non-finite or key out of range). */
int rtr_map_dropped(rtr_map* m, uint64_t* n);

/* PCL VoxelGrid downsample on the device: centroids ordered by linear voxel
 * index over the cloud's bounding box (x fastest). */
int rtr_voxel_downsample(int device, const float* xyz, size_t n, size_t stride_bytes, float leaf,
                         float* out_xyz, size_t cap, size_t* n_out);

/* ------------------------------------------------------------------ */
/* PCD v0.7 I/O (pcl::io::loadPCDFile / savePCDFileASCII,               */
/* RealTimeRobot.cpp:34-35,108-109; function.h:126-127)                 */
/* ------------------------------------------------------------------ */
/* Number of points and whether normal_x/y/z fields exist. */
int rtr_pcd_info(const char* path, size_t* n_points, int* has_normals);
/* Read xyz (packed, cap points) and optionally normals (nullable). */
int rtr_pcd_read(const char* path, float* xyz, float* normals, size_t cap, size_t* n_out);
/* ASCII writer with 8 significant digits, like savePCDFileASCII. */
int rtr_pcd_write_ascii(const char* path, const float* xyz, size_t n, size_t stride_bytes);

/* Streaming copy of `bytes` (multiple of 16) repeated `reps` times; average
 * milliseconds per copy.  Calibrates profiler byte counters and measures the
 * HBM bandwidth the device sustains.  No reference counterpart. */
int rtr_copy_probe(int device, size_t bytes, int reps, float* ms_avg);

/* ------------------------------------------------------------------ */
/* Truncated distance field (TDF) of an occupancy grid: replaces        */
/* ComputeTDFWithCuda (kernel.cu:34-35, declared key_point.h:35-36).     */
/* out[x + y*dims[0] + z*dims[0]*dims[1]] = min(trunc2, squared distance */
/* in voxels to the nearest occupied voxel) -- the reference's values    */
/* (kernel.cu:9-31; trunc2 = 900 there) for any grid size, computed by   */
/* an exact separable distance transform.  Occupied coordinates outside  */
/* the grid are ignored (the reference indexes without a bounds check,   */
/* matching.h:181-187).                                                  */
/* ------------------------------------------------------------------ */
int rtr_tdf(int device, const int32_t* occ_xyz, size_t n_occ, const int dims[3], float trunc2, float* out);
/* Allocation-free variant of rtr_tdf.  No counterpart in the reference: it
 * is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops).  occ_dev: n_occ x 3 int32 voxel coordinates
 * and out_dev: dims[0]*dims[1]*dims[2] floats (x fastest), both device
 * memory the caller owns; the kernels run asynchronously on `stream` (a
 * hipStream_t; NULL: the default stream), and the call makes no allocation:
 * the intermediate grid lives in the library's static per-device array of
 * 2^21 floats (8 MiB), so grids above 2^21 voxels are refused
 * (RTR_ERR_CAPACITY).  Calls that overlap on one device race on that array. */
int rtr_tdf_noalloc(int device, const int32_t* occ_dev, size_t n_occ, const int dims[3], float trunc2,
                    float* out_dev, void* stream);
/* Same from points: voxel = floor((p - origin) / resolution) inside the
 * grid (the occupancy of key_point.h:287-304 and matching.h:169-186). */
int rtr_tdf_points(int device, const float* xyz, size_t n, size_t stride_bytes, const float origin[3],
                   float resolution, const int dims[3], float trunc2, float* out);

/* ------------------------------------------------------------------ */
/* Keypoint + TDF + yaw-search coarse proposer: the reference's own     */
/* initial alignment (RealTimeRobot.cpp:45-105: keypoints, occupancy    */
/* key_point.h:112-161, TDF key_point.h:251-334, get_Distance's 36 yaw  */
/* steps matching.h:122-222, consensus function.h:35-109), kept as the  */
/* baseline of the learned proposer.  Per pair: 32 FPS keypoints per    */
/* cloud; source keypoint occupancy (1 cm voxels within +-10 cm);       */
/* target keypoint TDF (30^3, cap 900); the best of 36 yaws about z for */
/* every (target, source) keypoint pair; the `candidates` best pairs    */
/* become poses T = Trans(k_t) Rz(yaw) Trans(-k_s) (source -> target,   */
/* column-major) and the lowest truncated chamfer (tau, metres) over    */
/* 1024-point subsamples wins.  Yaw only: the reference assumes the     */
/* gravity axis is shared.                                              */
/* ------------------------------------------------------------------ */
typedef struct rtr_coarse rtr_coarse;
int rtr_coarse_create(rtr_coarse** out, int device);
int rtr_coarse_destroy(rtr_coarse* h);
int rtr_coarse_set_clouds(rtr_coarse* h, int n, const float* const* src, const size_t* ns, const float* const* tgt,
                          const size_t* nt);
/* poses: n * 16 doubles (column-major); scores (nullable): n chamfer scores. */
int rtr_coarse_propose(rtr_coarse* h, int candidates, float resolution, float tau, double* poses, double* scores);
/* The keypoint-pair scores behind rtr_coarse_propose: scores/yaw n*32*32
 * (pair, target keypoint, source keypoint; 3.4e38 = rejected by the
 * occupancy-count filter), keypoints (nullable) 2n*32*3 (cloud 2p = source,
 * 2p+1 = target). */
int rtr_coarse_pair_scores(rtr_coarse* h, float resolution, float* scores, int* yaw, float* keypoints);
/* Source keypoint occupancy of the last scoring: voxels n*32*2048 (index
 * x + 20 y + 400 z, ascending), counts n*32. */
int rtr_coarse_occupancy(rtr_coarse* h, uint16_t* voxels, int* counts);
float rtr_coarse_last_ms(rtr_coarse* h);

/* ------------------------------------------------------------------ */
/* Learned pose proposer: farthest-point sampling (1024 points per      */
/* cloud), shared point MLP 3-64-128-256 + max-pool per cloud, head     */
/* 512-256-9 (6D rotation + translation), BF16 MFMA with FP32           */
/* accumulation, trained with Adam on poses the GPU ICP converged to.   */
/* Replaces the reference's coarse alignment stage that produces the    */
/* initial pose for ICP: keypoint/plane-area matching, TDF yaw search   */
/* and consensus (RealTimeRobot.cpp:45-105, matching.h:122-222,         */
/* function.h:35-109).                                                  */
/* ------------------------------------------------------------------ */
typedef struct rtr_prop rtr_prop;

/* Buffers for up to max_pairs (source, target) pairs per call. */
int rtr_prop_create(rtr_prop** out, int device, int max_pairs);
int rtr_prop_destroy(rtr_prop* h);
/* Number of FP32 parameters (W1 16x64 (xyz rows 0-2, rows 3-15 zero-input
 * padding), b1, W2 64x128, b2, W3 128x256, b3, W4 512x256, b4, W5 256x9, b5;
 * matrices row-major (in x out)). */
int rtr_prop_num_params(size_t* n);
int rtr_prop_set_weights(rtr_prop* h, const float* w);
int rtr_prop_get_weights(rtr_prop* h, float* w);
/* Upload n_pairs (source, target) clouds (packed xyz); they stay resident. */
int rtr_prop_set_clouds(rtr_prop* h, int n_pairs, const float* const* src, const size_t* ns,
                        const float* const* tgt, const size_t* nt);
/* Farthest-point samples of the resident clouds (clouds ordered src0, tgt0,
 * src1, ...): 1024 points per cloud centred on the centroid of a strided
 * subsample of at most 8192 points (pts: 2*n*1024*3), the centroids (2*n*3,
 * nullable) and the chosen subsample indices (2*n*1024, nullable).  Selection
 * starts at subsample index 0; ties go to the lowest index. */
int rtr_prop_fps(rtr_prop* h, float* pts, float* centroids, int32_t* sel);
/* FPS + network on the resident clouds: one 4x4 pose per pair (n*16 doubles,
 * column-major) mapping the source onto the target. */
int rtr_prop_propose(rtr_prop* h, double* poses);
/* k hypotheses per pair (1 <= k <= 64; k*n <= max_pairs): the network on the
 * source samples rotated by each rots[j] (k row-major 3x3), candidate j's pose
 * composed with rots[j], each candidate scored by the truncated chamfer of
 * the farthest-point samples: mean over the source samples of
 * min(squared distance to the nearest target sample, tau^2).  Outputs are
 * pair-major (p*k + j): poses (n*k*16, column-major, nullable), scores (n*k,
 * nullable), and best (n*16), the lowest-scoring candidate of each pair
 * (lowest j on ties). */
int rtr_prop_propose_multi(rtr_prop* h, int k, const double* rots, float tau, double* poses,
                           double* scores, double* best);
/* Inference through the fused point-MLP + max-pool kernel (on = 1, default)
 * or through the layer-by-layer products (0); both give identical features. */
int rtr_prop_set_fused(rtr_prop* h, int on);
/* Device time of the last rtr_prop_propose / rtr_prop_propose_multi, ms. */
int rtr_prop_last_ms(rtr_prop* h, float* ms);
/* Network on given centred FPS points (n*2*1024*3, clouds src0, tgt0, ...).
 * Outputs (each nullable): head outputs O (n*9), poses R|t (n*12 doubles, R
 * row-major, in the centred frames).  With labels (n*12 doubles, same
 * layout): the mean loss mean_i ||(R - R*) s_i + t - t*||^2 over the source
 * points and pairs, and its gradient w.r.t. the FP32 parameters. */
int rtr_prop_forward_backward(rtr_prop* h, int n_pairs, const float* xyz, const double* labels,
                              float* O, double* rt, float* loss, float* grad);
/* One Adam step (beta 0.9/0.999, eps 1e-8) on a labelled batch. */
int rtr_prop_train_step(rtr_prop* h, int n_pairs, const float* xyz, const double* labels, float lr,
                        float* loss);
/* The proposer's BF16 MFMA product on its own: C (M x N, FP32) = op(A) op(B)
 * with A, B rounded to BF16 (A is M x K, or K x M when ta; B is K x N, or
 * N x K when tb), K split over k_split workgroups accumulated atomically. */
int rtr_gemm_bf16(int device, const float* A, int ta, const float* B, int tb, float* C, int M, int N,
                  int K, int k_split);
/* The fused point MLP rounds its activations to BF16 with the hardware
 * conversion (v_cvt_pk_bf16_f32); the layer-by-layer path and the host round
 * in software (round to nearest even).  Counts the float bit patterns, NaNs
 * excepted, on which the two differ: identical features need 0. */
int rtr_bf16_conversion_check(int device, unsigned long long* mismatches);

/* ------------------------------------------------------------------ */
/* Learned-init registration of many pairs per call: the network        */
/* proposes an initial pose per pair, one batched ICP launch refines    */
/* every proposal, the registration judge scores the refined poses, and */
/* the pairs it rejects run plain multi-start ICP, whose best start     */
/* replaces them.  Replaces the reference's coarse stage followed by    */
/* ICP: keypoints, TDF yaw search and consensus (RealTimeRobot.cpp:      */
/* 45-105, matching.h:122-222, function.h:35-109), then                 */
/* pcl::IterativeClosestPoint (function.h:111-123).                     */
/* ------------------------------------------------------------------ */
typedef struct rtr_learned rtr_learned;

typedef struct {
  int hypotheses;              /* network orientations per pair (1..64), default 8 */
  int n_starts;                /* fallback starts per rejected pair (1..64), default 8 */
  float tau;                   /* truncation of the hypotheses' chamfer score, default 0.1 m */
  double rmse_threshold;       /* judge: plane RMSE above this rejects, default 2e-3 m */
  double min_inlier_frac;      /* judge: inlier fraction below this rejects, default 0.8 */
  int judge_k;                 /* judge: neighbours of the PCA planes, default 20 */
  int judge_points;            /* judge: sampled source points per pair, default 1024 */
  double max_curvature;        /* judge: planar when surface variation <= this, default 0.02 */
  int refine_anderson;         /* refine: Anderson memory (0..8), default 5 */
  int refine_coarse_stride;    /* refine: coarse-to-fine stride, default 4 */
  int refine_coarse_iterations;/* refine: most coarse iterations, default 25 */
  int fallback_anderson;       /* fallback: Anderson memory, default 0 (plain ICP) */
  int fallback_item_pts;       /* fallback: source points per work item, default 256 */
  int fallback_blocks_per_cu;  /* fallback: persistent workgroups per CU, default 1 */
} rtr_learned_config;

/* Host wall-clock phases of the last call (ms) and its fallback count. */
typedef struct {
  int n_pairs;
  int n_fallback;
  float ms_propose;        /* the network (the refine grids are built beside it) */
  float ms_build_wait;     /* wait for the grid build after the network */
  float ms_refine;         /* batched ICP refine of every proposal */
  float ms_judge;          /* registration judge */
  float ms_fallback;       /* multi-start ICP of the rejected pairs */
  float ms_total;
  float ms_propose_device; /* device time of the network call */
} rtr_learned_stats;

int rtr_learned_config_default(rtr_learned_config* c);
/* icp: the registration parameters (nullable: rtr_params_default); the
 * refine adds cfg's acceleration and coarse-to-fine stage, the fallback runs
 * the plain iteration.  cfg nullable: rtr_learned_config_default. */
int rtr_learned_create(rtr_learned** out, int device, int max_pairs, const rtr_params* icp,
                       const rtr_learned_config* cfg);
int rtr_learned_destroy(rtr_learned* h);
/* Network weights (rtr_prop_num_params FP32 values, the layout of
 * rtr_prop_set_weights), from memory or from a raw little-endian float32 file. */
int rtr_learned_set_weights(rtr_learned* h, const float* w);
int rtr_learned_load_weights(rtr_learned* h, const char* path);
/* The hypotheses' orientations and the fallback's start rotations (row-major
 * 3x3 each; nullable: the library's default set of that size, the identity
 * then greedy farthest-point picks over SO(3)). */
int rtr_learned_set_rotations(rtr_learned* h, const double* hyp, int n_hyp, const double* starts, int n_starts);
int rtr_learned_get_rotations(rtr_learned* h, double* hyp, double* starts);
/* n <= max_pairs (source, target) pairs, packed xyz.  The handle uploads them
 * and also keeps the pointers (the fallback reads the rejected pairs' clouds
 * from them, as PCL's registration shares its input clouds): they must stay
 * valid until the next rtr_learned_set_clouds or rtr_learned_destroy.
 * Centroids (n*3 doubles, nullable: FP64 means) centre the fallback starts. */
int rtr_learned_set_clouds(rtr_learned* h, int n, const float* const* src, const size_t* ns,
                           const float* const* tgt, const size_t* nt, const double* src_centroids,
                           const double* tgt_centroids);
/* Register every pair: results (n), inits (n*16 column-major, nullable): the
 * pose each final ICP ran from (the proposal, or a rejected pair's best
 * start), rejected (n, nullable): 1 where the judge sent the pair to the
 * fallback. */
int rtr_learned_align(rtr_learned* h, rtr_result* results, double* inits, int* rejected);
/* The same from the caller's proposals (n*16 column-major) instead of the
 * network's. */
int rtr_learned_align_from(rtr_learned* h, const double* proposals, rtr_result* results, double* inits,
                           int* rejected);
int rtr_learned_get_stats(rtr_learned* h, rtr_learned_stats* s);
/* The last call's proposals (n*16) and judge outputs (n each, nullable). */
int rtr_learned_get_proposals(rtr_learned* h, double* poses);
int rtr_learned_get_judge(rtr_learned* h, double* plane_rmse, double* inlier_frac, double* planar_frac);
/* plain = 1: the refine engine runs the plain iteration (the caller's icp
 * parameters) until plain = 0 restores the refine's; for parity checks and
 * tuning through the engine handle. */
int rtr_learned_use_plain_params(rtr_learned* h, int plain);
/* The handle's refine and fallback engines and its network (borrowed: owned
 * and destroyed by the rtr_learned handle), for tuning and profiling. */
int rtr_learned_batches(rtr_learned* h, rtr_batch** refine, rtr_batch** fallback, rtr_prop** proposer);

/* ------------------------------------------------------------------ */
/* Self-check hooks: run the library's own geometry code on the host.   */
/* ------------------------------------------------------------------ */
/* The grid NN search on the CPU over the same cell-sorted layout. */
int rtr_selfcheck_nn(const float* tgt, size_t nt, const float* q, size_t nq, float cell,
                     double max_dist, int32_t* idx, float* d2);
/* Same, also returning per query the number of cell ranges walked and of
 * candidates examined (the search work; input of the cell-size cost model). */
int rtr_selfcheck_nn_stats(const float* tgt, size_t nt, const float* q, size_t nq, float cell,
                           double max_dist, int32_t* idx, float* d2, uint32_t* ranges,
                           uint32_t* cands);
/* Reduce one iteration's moments and solve (FP64) on the CPU:
 * a[i] (transformed source), b[i] (target), nrm (nullable) -> Tinc. */
int rtr_selfcheck_solve(int mode, const float* a, const float* b, const float* nrm, size_t n,
                        const float anchor[3], double Tinc[16]);
/* The Anderson mixing of anderson_memory (no reference counterpart) on the
 * affine map g(x) = A x + c in R^6 (A row-major), from x = 0, with memory m:
 * iterates until max|g(x) - x| <= tol or max_iter mixes; x_out is the last
 * g(x), *iters the number of map evaluations. */
int rtr_selfcheck_anderson(int m, const double A[36], const double c[6], double tol, int max_iter,
                           double x_out[6], int* iters);
/* Rotation chart of the acceleration: v = vector part of a unit quaternion
 * with positive scalar part -> row-major R, then back to v_back.  Invalid
 * argument when |v| > 1 or the angle exceeds about 174 degrees. */
int rtr_selfcheck_so3(const double v[3], double R[9], double v_back[3]);

#ifdef __cplusplus
}
#endif
#endif
