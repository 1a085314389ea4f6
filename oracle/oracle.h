/*
 * CPU oracle for the ICP registration + voxel fusion path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP
 * implementation in realtime_robot_amd/ and the CPU timing baseline in
 * bench.py; the product never links, loads or calls it.
 *
 * Two families of functions live here:
 *   orc_*      FP64 restatement of the algorithm (exact nearest neighbour with
 *              lowest-index ties, FP64 Umeyama / point-to-plane solves, the
 *              PCL 1.8 convergence rules, fitness score, voxel accumulation,
 *              k-NN PCA normals, voxel-grid downsampling).
 *   orc_pcl_*  a PCL-semantics CPU ICP (float kd-tree with leaf size 15,
 *              float Umeyama, cumulative float re-transform of the working
 *              cloud, first-found ties) used only as the timed CPU baseline.
 *
 * The reference calls pcl::IterativeClosestPoint<PointXYZ,PointXYZ>
 * (RealTimeRobot/function.h:111-123) and pcl::transformPointCloud
 * (function.h:67,123; RealTimeRobot.cpp:105).  PCL 1.8.0, FLANN and Eigen are
 * not part of the reference tree, so their published behaviour is restated
 * here; each function names the call site it follows.
 *
 * Conventions: clouds are packed float xyz (n x 3); transforms are double[16]
 * column-major (Eigen's layout: element (r,c) at [c*4+r]).
 */
#ifndef RTR_ORACLE_H
#define RTR_ORACLE_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

enum {
  ORC_NOT_CONVERGED = 0,
  ORC_CONV_ITERATIONS = 1,
  ORC_CONV_TRANSFORM = 2,
  ORC_CONV_ABS_MSE = 3,
  ORC_CONV_REL_MSE = 4,
  ORC_FAIL_NO_CORRESPONDENCES = 5
};

/* float transform, PCL transformPointCloud order: ((m00*x + m01*y) + m02*z) + m03 */
void orc_transform_f32(const double T[16], const float* in, int64_t n, float* out);

/* Round a double squared radius down to the largest float f with (double)f <= r2. */
float orc_r2_to_float(double r2);

/* Brute-force exact NN.  d2 = ((dx*dx) + dy*dy) + dz*dz in float, ties to the
 * lowest target index, non-finite targets skipped.  Accepted iff d2 <= r2f;
 * otherwise idx = -1, d2 = +inf. */
void orc_nn_brute(const float* q, int64_t nq, const float* t, int64_t nt, float r2f,
                  int32_t* idx, float* d2);

/* kd-tree over a target cloud. */
typedef struct orc_kdtree orc_kdtree;
orc_kdtree* orc_kd_build(const float* t, int64_t nt, int leaf_size);
void orc_kd_free(orc_kdtree* kd);
/* exact NN.  flags: ORC_KD_TIE_LOWEST: same result as orc_nn_brute; without
 * it: first found in tree order (FLANN result-set semantics).
 * ORC_KD_BOUNDED: the search is bounded by r2f from the start (same result,
 * fewer nodes visited for queries with nothing in range). */
enum { ORC_KD_TIE_LOWEST = 1, ORC_KD_BOUNDED = 2 };
void orc_kd_nn(const orc_kdtree* kd, const float* q, int64_t nq, float r2f, int flags,
               int32_t* idx, float* d2);
/* exact k nearest, ordered by (d2, index) ascending; missing slots get -1/inf. */
void orc_kd_knn(const orc_kdtree* kd, const float* q, int64_t nq, int k, int32_t* idx, float* d2);

/* 3x3 SVD by one-sided Jacobi (Hestenes), A row-major: A = U diag(S) V^T,
 * S descending, U and V orthonormal. */
void orc_svd3(const double A[9], double U[9], double S[3], double V[9]);
/* Umeyama (1991) rigid fit without scaling: maps src onto tgt. */
int orc_umeyama(const double* src, const double* tgt, int64_t n, double T[16]);
/* Symmetric 3x3 eigen-decomposition by cyclic Jacobi: ascending eigenvalues,
 * eigenvectors in the columns of V (row-major). */
void orc_eig3(const double A[9], double w[3], double V[9]);
/* 6x6 SPD solve by Cholesky; AtA packed upper-triangular row-major (21). */
int orc_cholesky6(const double AtA[21], const double Atb[6], double x[6]);
/* ZYX Euler small-angle solution (alpha, beta, gamma, tx, ty, tz) expressed
 * about `anchor` -> world transform. */
void orc_euler_to_T(const double x[6], const double anchor[3], double T[16]);

typedef struct {
  int max_iter;          /* PCL default 10 */
  int min_corr;          /* PCL default 3 */
  int mode;              /* 0 point-to-point, 1 point-to-plane */
  int pad_;
  double eps_T;          /* transformation epsilon (PCL default 0) */
  double eps_rot;        /* rotation epsilon; <= 0 means 1 - eps_T */
  double eps_fit;        /* euclidean fitness epsilon (PCL default -DBL_MAX) */
  double r_max;          /* max correspondence distance (PCL default sqrt(DBL_MAX)) */
  double anchor[3];      /* origin of the point-to-plane linearisation */
} orc_icp_params;

typedef struct {
  double T[16];
  double mse;            /* mean d2 of the last iteration's correspondences */
  int64_t n_corr;        /* correspondences of the last iteration */
  int iterations;
  int converged;
  int state;
  int pad_;
} orc_icp_result;

/* FP64 oracle ICP.  Queries are the float transform of the ORIGINAL source by
 * float(T) every iteration (no cumulative float re-transform), NN is exact
 * with lowest-index ties, the estimate and composition are FP64, and the
 * stopping rules are PCL 1.8 DefaultConvergenceCriteria. tgt_normals may be
 * NULL for point-to-point.  trace (nullable) receives 16 doubles per
 * iteration (T after the iteration), up to max_iter entries. */
int orc_icp(const float* src, int64_t ns, const float* tgt, const float* tgt_normals, int64_t nt,
            const orc_icp_params* p, const double guess[16], orc_icp_result* out, double* trace);

/* PCL Registration::getFitnessScore(max_range): mean NN d2 of float(T)*src
 * over points with d2 <= max_range (squared compared to unsquared); DBL_MAX
 * if none. */
double orc_fitness(const float* src, int64_t ns, const float* tgt, int64_t nt, const double T[16],
                   double max_range);

/* Voxel accumulation: p = float(T)*x (T nullable = identity), key = floorf(p*inv_h)
 * per axis, packed 21 bits per axis with +2^20 bias.  Outputs are sorted by
 * key; sums in FP64 of the float points.  Returns the voxel count, or -1 if
 * cap is exceeded.  dropped (nullable) counts non-finite/out-of-range points. */
int64_t orc_voxel_accumulate(const float* pts, int64_t n, const double T[16], float inv_h,
                             int64_t* keys, double* sums, int64_t* counts, int64_t cap,
                             int64_t* dropped);
uint64_t orc_voxel_key(float x, float y, float z, float inv_h, int* ok);

/* PCA normals from the k nearest (including the point itself) ordered by
 * (d2, idx); smallest-eigenvalue eigenvector flipped towards vp;
 * curvature = l0/(l0+l1+l2).  normals (n x 3), curvature (n, nullable). */
void orc_normals(const float* pts, int64_t n, int k, const float vp[3], float* normals,
                 float* curvature);

/* PCL VoxelGrid-style downsample: ijk = floor(p*inv), linear index over the
 * cloud's ijk bounding box (x fastest), output ordered by that index,
 * centroid = FP64 mean.  Returns the number of voxels (<= cap) or -1. */
int64_t orc_voxel_downsample(const float* pts, int64_t n, float leaf, float* out, int64_t cap,
                             int64_t* counts);

/* Farthest-point sampling as the learned proposer defines it: a strided
 * subsample of m = min(n, max_sub) points (subsample i = point i*n/m), then k
 * greedy picks starting at subsample index 0, each the point with the largest
 * float squared distance ((dx*dx + dy*dy) + dz*dz) to the chosen set, lowest
 * index on ties.  sel (k) receives subsample indices; centroid (3, FP64 mean
 * of the subsample). */
void orc_fps(const float* pts, int64_t n, int64_t max_sub, int k, int32_t* sel, double* centroid);

/* Truncated distance field by brute force, as kernel.cu:9-31 computes it:
 * out[x + y*dx + z*dx*dy] = min(trunc2, min over occupied voxels of the
 * integer squared distance), with occupied coordinates outside the grid
 * ignored (bounds fixed). */
void orc_tdf(const int32_t* occ, int64_t n_occ, const int dims[3], float trunc2, float* out);

/* ---- PCL-semantics CPU baseline (float, cumulative, kd-tree leaf 15) ---- */
int orc_pcl_icp(const float* src, int64_t ns, const float* tgt, int64_t nt, const orc_icp_params* p,
                const double guess[16], orc_icp_result* out);
/* Runs npairs PCL-semantics registrations on nthreads threads (one pair per
 * thread at a time); includes the kd-tree build of each target. */
int orc_pcl_icp_batch(int npairs, const float* const* src, const int64_t* ns,
                      const float* const* tgt, const int64_t* nt, const orc_icp_params* p,
                      const double* guesses, orc_icp_result* out, int nthreads);

#ifdef __cplusplus
}
#endif
#endif
