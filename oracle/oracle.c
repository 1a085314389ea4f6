/*
 * CPU oracle -- TEST INFRASTRUCTURE ONLY (see oracle.h).  Built with
 * -ffp-contract=off so every float expression rounds exactly as written.
 *
 * Reference anchors (PCL behaviour restated, library not in the tree):
 *   ICP call contract ............ RealTimeRobot/function.h:111-123
 *   transformPointCloud .......... function.h:67,123; RealTimeRobot.cpp:105
 *   voxel keys (floor of p/res) .. key_point.h:143-160, matching.h:181-184
 */
#include "oracle.h"

#include <float.h>
#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------ */
/* transforms                                                               */
/* ------------------------------------------------------------------------ */

static void tf_to_f32(const double T[16], float m[12]) {
  /* rows 0..2 of the column-major 4x4, as PCL's Matrix4f */
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 4; ++c) m[r * 4 + c] = (float)T[c * 4 + r];
}

static inline void apply_f32(const float m[12], const float* p, float* o) {
  /* PCL transformPointCloud evaluates each row left to right. */
  float x = p[0], y = p[1], z = p[2];
  for (int r = 0; r < 3; ++r) {
    float a = m[r * 4 + 0] * x;
    a = a + m[r * 4 + 1] * y;
    a = a + m[r * 4 + 2] * z;
    a = a + m[r * 4 + 3];
    o[r] = a;
  }
}

void orc_transform_f32(const double T[16], const float* in, int64_t n, float* out) {
  float m[12];
  tf_to_f32(T, m);
  for (int64_t i = 0; i < n; ++i) {
    float o[3];
    apply_f32(m, in + 3 * i, o);
    out[3 * i + 0] = o[0];
    out[3 * i + 1] = o[1];
    out[3 * i + 2] = o[2];
  }
}

static void mat4_mul(const double A[16], const double B[16], double C[16]) {
  double R[16];
  for (int c = 0; c < 4; ++c)
    for (int r = 0; r < 4; ++r) {
      double s = 0.0;
      for (int k = 0; k < 4; ++k) s += A[k * 4 + r] * B[c * 4 + k];
      R[c * 4 + r] = s;
    }
  memcpy(C, R, sizeof(R));
}

static void mat4_identity(double T[16]) {
  memset(T, 0, 16 * sizeof(double));
  T[0] = T[5] = T[10] = T[15] = 1.0;
}

float orc_r2_to_float(double r2) {
  if (!(r2 >= 0.0)) return -1.0f;
  if (r2 >= (double)FLT_MAX) return FLT_MAX;
  float f = (float)r2;
  if ((double)f > r2) f = nextafterf(f, 0.0f);
  return f;
}

/* ------------------------------------------------------------------------ */
/* exact nearest neighbour                                                  */
/* ------------------------------------------------------------------------ */

static inline float d2f(const float* q, const float* t) {
  /* FLANN L2_Simple<float>: accumulate left to right, no fused multiply-add */
  float dx = q[0] - t[0];
  float dy = q[1] - t[1];
  float dz = q[2] - t[2];
  float s = dx * dx;
  s = s + dy * dy;
  s = s + dz * dz;
  return s;
}

static inline int finite3(const float* p) {
  return isfinite(p[0]) && isfinite(p[1]) && isfinite(p[2]);
}

void orc_nn_brute(const float* q, int64_t nq, const float* t, int64_t nt, float r2f,
                  int32_t* idx, float* d2) {
  for (int64_t i = 0; i < nq; ++i) {
    float best = INFINITY;
    int64_t bi = -1;
    for (int64_t j = 0; j < nt; ++j) {
      if (!finite3(t + 3 * j)) continue;
      float d = d2f(q + 3 * i, t + 3 * j);
      if (d < best) {
        best = d;
        bi = j;
      }
    }
    if (bi >= 0 && best <= r2f) {
      idx[i] = (int32_t)bi;
      d2[i] = best;
    } else {
      idx[i] = -1;
      d2[i] = INFINITY;
    }
  }
}

/* kd-tree: nodes split on the axis of largest spread at the midpoint of the
 * spread (median if that leaves one side empty); leaves hold <= leaf_size
 * points.  Search descends the near child first and visits the far child
 * when the float squared distance to the split plane does not exceed the
 * current best: that plane distance is a float lower bound of every far
 * point's float d2 (rounding is monotone), so the search is exact. */
typedef struct {
  int32_t left, right; /* children, -1 for leaf */
  int32_t begin, end;  /* leaf range in perm */
  int32_t axis;
  float split;
} kdnode;

struct orc_kdtree {
  const float* pts;
  int64_t n;
  int32_t* perm;
  float* sorted; /* points in perm order, packed xyz */
  kdnode* nodes;
  int32_t nnodes, cap;
  int leaf;
};

static int32_t kd_new_node(orc_kdtree* kd) {
  if (kd->nnodes == kd->cap) {
    kd->cap = kd->cap ? kd->cap * 2 : 1024;
    kd->nodes = (kdnode*)realloc(kd->nodes, (size_t)kd->cap * sizeof(kdnode));
  }
  return kd->nnodes++;
}

static int cmp_axis_ctx_axis;
static const float* cmp_axis_ctx_pts;
static int cmp_axis(const void* a, const void* b) {
  float va = cmp_axis_ctx_pts[3 * (*(const int32_t*)a) + cmp_axis_ctx_axis];
  float vb = cmp_axis_ctx_pts[3 * (*(const int32_t*)b) + cmp_axis_ctx_axis];
  if (va < vb) return -1;
  if (va > vb) return 1;
  return (*(const int32_t*)a) - (*(const int32_t*)b);
}
static pthread_mutex_t kd_sort_mu = PTHREAD_MUTEX_INITIALIZER;

static int32_t kd_build_rec(orc_kdtree* kd, int32_t b, int32_t e) {
  int32_t id = kd_new_node(kd);
  kdnode nd;
  nd.left = nd.right = -1;
  nd.begin = b;
  nd.end = e;
  nd.axis = 0;
  nd.split = 0.0f;
  if (e - b <= kd->leaf) {
    kd->nodes[id] = nd;
    return id;
  }
  float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (int32_t i = b; i < e; ++i) {
    const float* p = kd->pts + 3 * (int64_t)kd->perm[i];
    for (int a = 0; a < 3; ++a) {
      if (p[a] < lo[a]) lo[a] = p[a];
      if (p[a] > hi[a]) hi[a] = p[a];
    }
  }
  int axis = 0;
  float span = hi[0] - lo[0];
  for (int a = 1; a < 3; ++a)
    if (hi[a] - lo[a] > span) {
      span = hi[a] - lo[a];
      axis = a;
    }
  if (!(span > 0.0f)) { /* all points identical: keep as a leaf */
    kd->nodes[id] = nd;
    return id;
  }
  float split = lo[axis] + 0.5f * span;
  /* partition: <= split to the left */
  int32_t i = b, j = e - 1;
  while (i <= j) {
    if (kd->pts[3 * (int64_t)kd->perm[i] + axis] <= split)
      ++i;
    else {
      int32_t tmp = kd->perm[i];
      kd->perm[i] = kd->perm[j];
      kd->perm[j] = tmp;
      --j;
    }
  }
  int32_t mid = i;
  if (mid == b || mid == e) {
    /* degenerate: median split on a sorted range */
    pthread_mutex_lock(&kd_sort_mu);
    cmp_axis_ctx_axis = axis;
    cmp_axis_ctx_pts = kd->pts;
    qsort(kd->perm + b, (size_t)(e - b), sizeof(int32_t), cmp_axis);
    pthread_mutex_unlock(&kd_sort_mu);
    mid = b + (e - b) / 2;
    split = kd->pts[3 * (int64_t)kd->perm[mid - 1] + axis];
    /* move the boundary so that left holds exactly the values <= split */
    while (mid < e && kd->pts[3 * (int64_t)kd->perm[mid] + axis] <= split) ++mid;
    if (mid == e) {
      kd->nodes[id] = nd;
      return id;
    }
  }
  nd.axis = axis;
  nd.split = split;
  int32_t l = kd_build_rec(kd, b, mid);
  int32_t r = kd_build_rec(kd, mid, e);
  nd.left = l;
  nd.right = r;
  kd->nodes[id] = nd;
  return id;
}

orc_kdtree* orc_kd_build(const float* t, int64_t nt, int leaf_size) {
  orc_kdtree* kd = (orc_kdtree*)calloc(1, sizeof(orc_kdtree));
  kd->pts = t;
  kd->leaf = leaf_size > 0 ? leaf_size : 15;
  kd->perm = (int32_t*)malloc((size_t)(nt > 0 ? nt : 1) * sizeof(int32_t));
  int64_t m = 0;
  for (int64_t i = 0; i < nt; ++i)
    if (finite3(t + 3 * i)) kd->perm[m++] = (int32_t)i;
  kd->n = m;
  if (m > 0) kd_build_rec(kd, 0, (int32_t)m);
  kd->sorted = (float*)malloc((size_t)(m > 0 ? m : 1) * 3 * sizeof(float));
  for (int64_t i = 0; i < m; ++i) memcpy(kd->sorted + 3 * i, t + 3 * (int64_t)kd->perm[i], 12);
  return kd;
}

void orc_kd_free(orc_kdtree* kd) {
  if (!kd) return;
  free(kd->perm);
  free(kd->sorted);
  free(kd->nodes);
  free(kd);
}

typedef struct {
  float best;
  int32_t bi;
  int tie_lowest;
} nnres;

/* Lower bound of d2f(q, p) for every point p of a cell, from q's per-axis
 * offsets to the split planes that bound the cell on q's side (0 on axes
 * with none), summed in d2f's order.  Each |off| is at most the matching
 * |q - p| after rounding (rounding of q - x is monotone in x), so each float
 * square is at most d2f's, and float addition of non-negative terms is
 * monotone: the bound never exceeds a point's d2f, and the search stays
 * exact (Arya & Mount's incremental distance, evaluated in d2f's order). */
static inline float off_bound(const float off[3]) {
  float s = off[0] * off[0];
  s = s + off[1] * off[1];
  s = s + off[2] * off[2];
  return s;
}

static void kd_nn_rec(const orc_kdtree* kd, int32_t id, const float* q, float off[3], nnres* r) {
  const kdnode* nd = &kd->nodes[id];
  if (nd->left < 0) {
    for (int32_t i = nd->begin; i < nd->end; ++i) {
      float d = d2f(q, kd->sorted + 3 * (int64_t)i);
      int32_t j = kd->perm[i];
      if (d < r->best || (r->tie_lowest && d == r->best && j < r->bi)) {
        r->best = d;
        r->bi = j;
      }
    }
    return;
  }
  float diff = q[nd->axis] - nd->split;
  int32_t near = diff <= 0.0f ? nd->left : nd->right;
  int32_t far = diff <= 0.0f ? nd->right : nd->left;
  kd_nn_rec(kd, near, q, off, r);
  const float saved = off[nd->axis];
  off[nd->axis] = diff;
  const float lb = off_bound(off);
  if (r->tie_lowest ? (lb <= r->best) : (lb < r->best)) kd_nn_rec(kd, far, q, off, r);
  off[nd->axis] = saved;
}

/* flags: ORC_KD_TIE_LOWEST (ties to the lowest index; else the first found in
 * tree order, FLANN's result set) | ORC_KD_BOUNDED (the search starts from
 * the radius r2f instead of +inf: the same answers, since a subtree it prunes
 * only holds points beyond r2f, which the radius test rejects anyway, and the
 * candidates within r2f are visited in the same order; PCL's own search is
 * unbounded, so its restatement leaves this off). */
void orc_kd_nn(const orc_kdtree* kd, const float* q, int64_t nq, float r2f, int flags,
               int32_t* idx, float* d2) {
  const int tie_lowest = (flags & ORC_KD_TIE_LOWEST) != 0;
  const int bounded = (flags & ORC_KD_BOUNDED) != 0;
  for (int64_t i = 0; i < nq; ++i) {
    /* bounded: best starts at r2f (tie rule: a point at exactly r2f still
     * wins over the sentinel index) or just above it (strict rule) */
    nnres r = {INFINITY, -1, tie_lowest};
    if (bounded) {
      r.best = tie_lowest ? r2f : nextafterf(r2f, INFINITY);
      r.bi = INT32_MAX;
    }
    float off[3] = {0.0f, 0.0f, 0.0f};
    if (kd->n > 0 && finite3(q + 3 * i)) kd_nn_rec(kd, 0, q + 3 * i, off, &r);
    if (r.bi >= 0 && r.bi != INT32_MAX && r.best <= r2f) {
      idx[i] = r.bi;
      d2[i] = r.best;
    } else {
      idx[i] = -1;
      d2[i] = INFINITY;
    }
  }
}

typedef struct {
  int k, cnt;
  float* d;
  int32_t* id;
} knnres;

static inline int lex_less(float da, int32_t ia, float db, int32_t ib) {
  return da < db || (da == db && ia < ib);
}

static void knn_insert(knnres* r, float d, int32_t j) {
  if (r->cnt == r->k && !lex_less(d, j, r->d[r->k - 1], r->id[r->k - 1])) return;
  int pos = r->cnt < r->k ? r->cnt++ : r->k - 1;
  while (pos > 0 && lex_less(d, j, r->d[pos - 1], r->id[pos - 1])) {
    r->d[pos] = r->d[pos - 1];
    r->id[pos] = r->id[pos - 1];
    --pos;
  }
  r->d[pos] = d;
  r->id[pos] = j;
}

static void kd_knn_rec(const orc_kdtree* kd, int32_t id, const float* q, knnres* r) {
  const kdnode* nd = &kd->nodes[id];
  if (nd->left < 0) {
    for (int32_t i = nd->begin; i < nd->end; ++i)
      knn_insert(r, d2f(q, kd->sorted + 3 * (int64_t)i), kd->perm[i]);
    return;
  }
  float diff = q[nd->axis] - nd->split;
  int32_t near = diff <= 0.0f ? nd->left : nd->right;
  int32_t far = diff <= 0.0f ? nd->right : nd->left;
  kd_knn_rec(kd, near, q, r);
  float pd = diff * diff;
  if (r->cnt < r->k || pd <= r->d[r->k - 1]) kd_knn_rec(kd, far, q, r);
}

void orc_kd_knn(const orc_kdtree* kd, const float* q, int64_t nq, int k, int32_t* idx, float* d2) {
  for (int64_t i = 0; i < nq; ++i) {
    knnres r = {k, 0, d2 + (int64_t)k * i, idx + (int64_t)k * i};
    if (kd->n > 0 && finite3(q + 3 * i)) kd_knn_rec(kd, 0, q + 3 * i, &r);
    for (int s = r.cnt; s < k; ++s) {
      r.d[s] = INFINITY;
      r.id[s] = -1;
    }
  }
}

/* ------------------------------------------------------------------------ */
/* small dense linear algebra (FP64)                                        */
/* ------------------------------------------------------------------------ */

void orc_svd3(const double A[9], double U[9], double S[3], double V[9]) {
  /* One-sided Jacobi (Hestenes 1958): rotate column pairs of B = A V until
   * they are mutually orthogonal; then S = column norms, U = B / S. */
  double B[9];
  memcpy(B, A, sizeof(B));
  for (int i = 0; i < 9; ++i) V[i] = (i % 4 == 0) ? 1.0 : 0.0;
  for (int sweep = 0; sweep < 60; ++sweep) {
    double off = 0.0;
    for (int p = 0; p < 2; ++p)
      for (int q = p + 1; q < 3; ++q) {
        double a = 0, b = 0, g = 0;
        for (int r = 0; r < 3; ++r) {
          a += B[r * 3 + p] * B[r * 3 + p];
          b += B[r * 3 + q] * B[r * 3 + q];
          g += B[r * 3 + p] * B[r * 3 + q];
        }
        if (fabs(g) <= 1e-300 || fabs(g) <= 1e-17 * sqrt(a * b)) continue;
        off = fmax(off, fabs(g) / sqrt(a * b));
        double zeta = (b - a) / (2.0 * g);
        double t = (zeta >= 0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
        double c = 1.0 / sqrt(1.0 + t * t), s = c * t;
        for (int r = 0; r < 3; ++r) {
          double bp = B[r * 3 + p], bq = B[r * 3 + q];
          B[r * 3 + p] = c * bp - s * bq;
          B[r * 3 + q] = s * bp + c * bq;
          double vp = V[r * 3 + p], vq = V[r * 3 + q];
          V[r * 3 + p] = c * vp - s * vq;
          V[r * 3 + q] = s * vp + c * vq;
        }
      }
    if (off < 1e-16) break;
  }
  double sn[3];
  for (int c = 0; c < 3; ++c)
    sn[c] = sqrt(B[c] * B[c] + B[3 + c] * B[3 + c] + B[6 + c] * B[6 + c]);
  int ord[3] = {0, 1, 2};
  for (int i = 0; i < 3; ++i)
    for (int j = i + 1; j < 3; ++j)
      if (sn[ord[j]] > sn[ord[i]]) {
        int t = ord[i];
        ord[i] = ord[j];
        ord[j] = t;
      }
  double Bs[9], Vs[9];
  for (int c = 0; c < 3; ++c) {
    S[c] = sn[ord[c]];
    for (int r = 0; r < 3; ++r) {
      Bs[r * 3 + c] = B[r * 3 + ord[c]];
      Vs[r * 3 + c] = V[r * 3 + ord[c]];
    }
  }
  memcpy(V, Vs, sizeof(Vs));
  double tol = (S[0] > 0 ? S[0] : 1.0) * 1e-13;
  int rank = 0;
  for (int c = 0; c < 3; ++c) {
    if (S[c] > tol) {
      for (int r = 0; r < 3; ++r) U[r * 3 + c] = Bs[r * 3 + c] / S[c];
      rank++;
    }
  }
  /* complete U to an orthonormal basis for zero singular values */
  if (rank == 0) {
    for (int i = 0; i < 9; ++i) U[i] = (i % 4 == 0) ? 1.0 : 0.0;
    rank = 1;
  }
  if (rank == 1) {
    double u0[3] = {U[0], U[3], U[6]};
    double e[3] = {0, 0, 0};
    int m = fabs(u0[0]) < fabs(u0[1]) ? (fabs(u0[0]) < fabs(u0[2]) ? 0 : 2)
                                      : (fabs(u0[1]) < fabs(u0[2]) ? 1 : 2);
    e[m] = 1.0;
    double d = e[0] * u0[0] + e[1] * u0[1] + e[2] * u0[2];
    double u1[3] = {e[0] - d * u0[0], e[1] - d * u0[1], e[2] - d * u0[2]};
    double nn = sqrt(u1[0] * u1[0] + u1[1] * u1[1] + u1[2] * u1[2]);
    for (int r = 0; r < 3; ++r) U[r * 3 + 1] = u1[r] / nn;
    rank = 2;
  }
  if (rank == 2) {
    U[2] = U[3] * U[7] - U[6] * U[4];
    U[5] = U[6] * U[1] - U[0] * U[7];
    U[8] = U[0] * U[4] - U[3] * U[1];
  }
}

static double det3(const double M[9]) {
  return M[0] * (M[4] * M[8] - M[5] * M[7]) - M[1] * (M[3] * M[8] - M[5] * M[6]) +
         M[2] * (M[3] * M[7] - M[4] * M[6]);
}

/* R = U diag(1,1,s) V^T, s = sign(det U det V) (Umeyama 1991, eq. 39-40). */
static void rot_from_cov(const double Sigma[9], double R[9]) {
  double U[9], S[3], V[9];
  orc_svd3(Sigma, U, S, V);
  double s = (det3(U) * det3(V) < 0.0) ? -1.0 : 1.0;
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c)
      R[r * 3 + c] = U[r * 3 + 0] * V[c * 3 + 0] + U[r * 3 + 1] * V[c * 3 + 1] +
                     s * U[r * 3 + 2] * V[c * 3 + 2];
}

static void rigid_from_moments(double n, const double sx[3], const double sy[3],
                               const double sxy[9], double T[16]) {
  /* sxy[a*3+b] = sum x_a y_b (x = source, y = target) */
  double mx[3], my[3], Sigma[9], R[9];
  for (int a = 0; a < 3; ++a) {
    mx[a] = sx[a] / n;
    my[a] = sy[a] / n;
  }
  /* Sigma = (1/n) sum (y - my)(x - mx)^T, row index = target axis */
  for (int a = 0; a < 3; ++a)
    for (int b = 0; b < 3; ++b) Sigma[a * 3 + b] = sxy[b * 3 + a] / n - my[a] * mx[b];
  rot_from_cov(Sigma, R);
  mat4_identity(T);
  for (int r = 0; r < 3; ++r) {
    for (int c = 0; c < 3; ++c) T[c * 4 + r] = R[r * 3 + c];
    T[12 + r] = my[r] - (R[r * 3 + 0] * mx[0] + R[r * 3 + 1] * mx[1] + R[r * 3 + 2] * mx[2]);
  }
}

int orc_umeyama(const double* src, const double* tgt, int64_t n, double T[16]) {
  if (n < 1) {
    mat4_identity(T);
    return -1;
  }
  /* two-pass centred accumulation for the oracle */
  double mx[3] = {0, 0, 0}, my[3] = {0, 0, 0};
  for (int64_t i = 0; i < n; ++i)
    for (int a = 0; a < 3; ++a) {
      mx[a] += src[3 * i + a];
      my[a] += tgt[3 * i + a];
    }
  for (int a = 0; a < 3; ++a) {
    mx[a] /= (double)n;
    my[a] /= (double)n;
  }
  double Sigma[9] = {0}, R[9];
  for (int64_t i = 0; i < n; ++i)
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b)
        Sigma[a * 3 + b] += (tgt[3 * i + a] - my[a]) * (src[3 * i + b] - mx[b]);
  for (int k = 0; k < 9; ++k) Sigma[k] /= (double)n;
  rot_from_cov(Sigma, R);
  mat4_identity(T);
  for (int r = 0; r < 3; ++r) {
    for (int c = 0; c < 3; ++c) T[c * 4 + r] = R[r * 3 + c];
    T[12 + r] = my[r] - (R[r * 3 + 0] * mx[0] + R[r * 3 + 1] * mx[1] + R[r * 3 + 2] * mx[2]);
  }
  return 0;
}

void orc_eig3(const double A[9], double w[3], double V[9]) {
  double a[9];
  memcpy(a, A, sizeof(a));
  for (int i = 0; i < 9; ++i) V[i] = (i % 4 == 0) ? 1.0 : 0.0;
  for (int sweep = 0; sweep < 60; ++sweep) {
    double off = fabs(a[1]) + fabs(a[2]) + fabs(a[5]);
    double scale = fabs(a[0]) + fabs(a[4]) + fabs(a[8]);
    if (off <= 1e-18 * scale || off == 0.0) break;
    for (int p = 0; p < 2; ++p)
      for (int q = p + 1; q < 3; ++q) {
        double apq = a[p * 3 + q];
        if (apq == 0.0) continue;
        double theta = (a[q * 3 + q] - a[p * 3 + p]) / (2.0 * apq);
        double t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
        double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
        for (int k = 0; k < 3; ++k) { /* A <- A J */
          double akp = a[k * 3 + p], akq = a[k * 3 + q];
          a[k * 3 + p] = c * akp - s * akq;
          a[k * 3 + q] = s * akp + c * akq;
        }
        for (int k = 0; k < 3; ++k) { /* A <- J^T A */
          double apk = a[p * 3 + k], aqk = a[q * 3 + k];
          a[p * 3 + k] = c * apk - s * aqk;
          a[q * 3 + k] = s * apk + c * aqk;
        }
        for (int k = 0; k < 3; ++k) {
          double vkp = V[k * 3 + p], vkq = V[k * 3 + q];
          V[k * 3 + p] = c * vkp - s * vkq;
          V[k * 3 + q] = s * vkp + c * vkq;
        }
      }
  }
  double d[3] = {a[0], a[4], a[8]};
  int ord[3] = {0, 1, 2};
  for (int i = 0; i < 3; ++i)
    for (int j = i + 1; j < 3; ++j)
      if (d[ord[j]] < d[ord[i]]) {
        int t = ord[i];
        ord[i] = ord[j];
        ord[j] = t;
      }
  double Vs[9];
  for (int c = 0; c < 3; ++c) {
    w[c] = d[ord[c]];
    for (int r = 0; r < 3; ++r) Vs[r * 3 + c] = V[r * 3 + ord[c]];
  }
  memcpy(V, Vs, sizeof(Vs));
}

/* packed upper-triangular index of (r, c), r <= c, 6x6 */
static inline int up6(int r, int c) { return r * 6 - r * (r - 1) / 2 + (c - r); }

int orc_cholesky6(const double AtA[21], const double Atb[6], double x[6]) {
  /* A = L L^T.  A pivot below 1e-15 of the largest diagonal marks a direction
   * the correspondences do not constrain; its unknown is set to zero. */
  double L[36] = {0}, y[6];
  int active[6];
  double dmax = 0;
  for (int i = 0; i < 6; ++i) dmax = fmax(dmax, AtA[up6(i, i)]);
  int ok = 1;
  for (int j = 0; j < 6; ++j) {
    double s = AtA[up6(j, j)];
    for (int k = 0; k < j; ++k) s -= L[j * 6 + k] * L[j * 6 + k];
    if (!(s > 1e-15 * dmax)) {
      active[j] = 0;
      ok = 0;
      L[j * 6 + j] = 1.0;
      for (int i = j + 1; i < 6; ++i) L[i * 6 + j] = 0.0;
      continue;
    }
    active[j] = 1;
    double ljj = sqrt(s);
    L[j * 6 + j] = ljj;
    for (int i = j + 1; i < 6; ++i) {
      double t = AtA[up6(j, i)];
      for (int k = 0; k < j; ++k) t -= L[i * 6 + k] * L[j * 6 + k];
      L[i * 6 + j] = t / ljj;
    }
  }
  for (int i = 0; i < 6; ++i) {
    double s = Atb[i];
    for (int k = 0; k < i; ++k) s -= L[i * 6 + k] * y[k];
    y[i] = active[i] ? s / L[i * 6 + i] : 0.0;
  }
  for (int i = 5; i >= 0; --i) {
    double s = y[i];
    for (int k = i + 1; k < 6; ++k) s -= L[k * 6 + i] * x[k];
    x[i] = active[i] ? s / L[i * 6 + i] : 0.0;
  }
  return ok ? 0 : 1;
}

void orc_euler_to_T(const double x[6], const double anchor[3], double T[16]) {
  /* R = Rz(gamma) Ry(beta) Rx(alpha): the small-angle point-to-plane
   * parameterisation of Low (2004), t in the anchor frame. */
  double ca = cos(x[0]), sa = sin(x[0]);
  double cb = cos(x[1]), sb = sin(x[1]);
  double cg = cos(x[2]), sg = sin(x[2]);
  double R[9] = {cg * cb, -sg * ca + cg * sb * sa, sg * sa + cg * sb * ca,
                 sg * cb, cg * ca + sg * sb * sa,  -cg * sa + sg * sb * ca,
                 -sb,     cb * sa,                 cb * ca};
  mat4_identity(T);
  for (int r = 0; r < 3; ++r) {
    for (int c = 0; c < 3; ++c) T[c * 4 + r] = R[r * 3 + c];
    double Ra = R[r * 3 + 0] * anchor[0] + R[r * 3 + 1] * anchor[1] + R[r * 3 + 2] * anchor[2];
    T[12 + r] = x[3 + r] + anchor[r] - Ra;
  }
}

/* ------------------------------------------------------------------------ */
/* ICP (FP64 oracle)                                                        */
/* ------------------------------------------------------------------------ */

typedef struct {
  int iter_sim;
  double prev_mse;
} conv_state;

/* PCL 1.8 DefaultConvergenceCriteria::hasConverged with
 * max_iterations_similar_transforms_ = 0 and mse_threshold_absolute_ = 1e-12;
 * thresholds as IterativeClosestPoint::computeTransformation sets them. */
static int convergence(const orc_icp_params* p, int iterations, const double Tinc[16], double mse,
                       conv_state* cs) {
  if (iterations >= p->max_iter) return ORC_CONV_ITERATIONS;
  double rot_thr = p->eps_rot > 0 ? p->eps_rot : 1.0 - p->eps_T;
  double cos_angle = 0.5 * (Tinc[0] + Tinc[5] + Tinc[10] - 1.0);
  double tsq = Tinc[12] * Tinc[12] + Tinc[13] * Tinc[13] + Tinc[14] * Tinc[14];
  if (cos_angle >= rot_thr && tsq <= p->eps_T) return ORC_CONV_TRANSFORM;
  double dm = fabs(mse - cs->prev_mse);
  if (dm < 1e-12) return ORC_CONV_ABS_MSE;
  if (dm / cs->prev_mse < p->eps_fit) return ORC_CONV_REL_MSE;
  cs->prev_mse = mse;
  return ORC_NOT_CONVERGED;
}

int orc_icp(const float* src, int64_t ns, const float* tgt, const float* tgt_normals, int64_t nt,
            const orc_icp_params* p, const double guess[16], orc_icp_result* out, double* trace) {
  double T[16];
  if (guess)
    memcpy(T, guess, sizeof(T));
  else
    mat4_identity(T);
  memset(out, 0, sizeof(*out));
  float* q = (float*)malloc((size_t)(ns > 0 ? ns : 1) * 12);
  int32_t* idx = (int32_t*)malloc((size_t)(ns > 0 ? ns : 1) * 4);
  float* d2 = (float*)malloc((size_t)(ns > 0 ? ns : 1) * 4);
  orc_kdtree* kd = orc_kd_build(tgt, nt, 15);
  float r2f = orc_r2_to_float(p->r_max * p->r_max);
  conv_state cs = {0, DBL_MAX};
  int state = ORC_NOT_CONVERGED, iters = 0;
  int64_t ncorr = 0;
  double mse = 0;
  for (;;) {
    orc_transform_f32(T, src, ns, q);
    orc_kd_nn(kd, q, ns, r2f, ORC_KD_TIE_LOWEST | ORC_KD_BOUNDED, idx, d2);
    ncorr = 0;
    double sd2 = 0;
    for (int64_t i = 0; i < ns; ++i)
      if (idx[i] >= 0) {
        ++ncorr;
        sd2 += (double)d2[i];
      }
    if (ncorr < (p->min_corr > 0 ? p->min_corr : 3)) {
      state = ORC_FAIL_NO_CORRESPONDENCES;
      break;
    }
    mse = sd2 / (double)ncorr;
    double Tinc[16];
    if (p->mode == 0) {
      double sx[3] = {0}, sy[3] = {0}, sxy[9] = {0};
      /* centre on the first correspondence pair to keep the moments small */
      double cx[3] = {0}, cy[3] = {0};
      int have_c = 0;
      for (int64_t i = 0; i < ns; ++i) {
        if (idx[i] < 0) continue;
        const float* a = q + 3 * i;
        const float* b = tgt + 3 * (int64_t)idx[i];
        if (!have_c) {
          for (int k = 0; k < 3; ++k) {
            cx[k] = a[k];
            cy[k] = b[k];
          }
          have_c = 1;
        }
        double x[3] = {(double)a[0] - cx[0], (double)a[1] - cx[1], (double)a[2] - cx[2]};
        double y[3] = {(double)b[0] - cy[0], (double)b[1] - cy[1], (double)b[2] - cy[2]};
        for (int k = 0; k < 3; ++k) {
          sx[k] += x[k];
          sy[k] += y[k];
        }
        for (int r = 0; r < 3; ++r)
          for (int c = 0; c < 3; ++c) sxy[r * 3 + c] += x[r] * y[c];
      }
      double Tc[16];
      rigid_from_moments((double)ncorr, sx, sy, sxy, Tc);
      /* undo the centring: x = x' + cx, y = y' + cy -> t = t' + cy - R cx */
      memcpy(Tinc, Tc, sizeof(Tc));
      for (int r = 0; r < 3; ++r)
        Tinc[12 + r] = Tc[12 + r] + cy[r] -
                       (Tc[r] * cx[0] + Tc[4 + r] * cx[1] + Tc[8 + r] * cx[2]);
    } else {
      double AtA[21] = {0}, Atb[6] = {0}, x[6];
      for (int64_t i = 0; i < ns; ++i) {
        if (idx[i] < 0) continue;
        const float* a = q + 3 * i;
        const float* b = tgt + 3 * (int64_t)idx[i];
        const float* nrm = tgt_normals + 3 * (int64_t)idx[i];
        double s[3], d[3], n[3];
        for (int k = 0; k < 3; ++k) {
          s[k] = (double)a[k] - p->anchor[k];
          d[k] = (double)b[k] - p->anchor[k];
          n[k] = nrm[k];
        }
        double row[6] = {n[2] * s[1] - n[1] * s[2], n[0] * s[2] - n[2] * s[0],
                         n[1] * s[0] - n[0] * s[1], n[0], n[1], n[2]};
        double rhs = n[0] * (d[0] - s[0]) + n[1] * (d[1] - s[1]) + n[2] * (d[2] - s[2]);
        for (int r = 0; r < 6; ++r) {
          for (int c = r; c < 6; ++c) AtA[up6(r, c)] += row[r] * row[c];
          Atb[r] += row[r] * rhs;
        }
      }
      orc_cholesky6(AtA, Atb, x);
      orc_euler_to_T(x, p->anchor, Tinc);
    }
    mat4_mul(Tinc, T, T);
    ++iters;
    if (trace && iters <= p->max_iter) memcpy(trace + 16 * (iters - 1), T, sizeof(T));
    state = convergence(p, iters, Tinc, mse, &cs);
    if (state != ORC_NOT_CONVERGED) break;
  }
  memcpy(out->T, T, sizeof(T));
  out->iterations = iters;
  out->state = state;
  out->converged = (state != ORC_NOT_CONVERGED && state != ORC_FAIL_NO_CORRESPONDENCES);
  out->n_corr = ncorr;
  out->mse = mse;
  orc_kd_free(kd);
  free(q);
  free(idx);
  free(d2);
  return 0;
}

double orc_fitness(const float* src, int64_t ns, const float* tgt, int64_t nt, const double T[16],
                   double max_range) {
  float* q = (float*)malloc((size_t)(ns > 0 ? ns : 1) * 12);
  int32_t* idx = (int32_t*)malloc((size_t)(ns > 0 ? ns : 1) * 4);
  float* d2 = (float*)malloc((size_t)(ns > 0 ? ns : 1) * 4);
  orc_transform_f32(T, src, ns, q);
  orc_kdtree* kd = orc_kd_build(tgt, nt, 15);
  orc_kd_nn(kd, q, ns, FLT_MAX, ORC_KD_TIE_LOWEST, idx, d2);
  double s = 0;
  int64_t m = 0;
  for (int64_t i = 0; i < ns; ++i)
    if (idx[i] >= 0 && (double)d2[i] <= max_range) {
      s += d2[i];
      ++m;
    }
  orc_kd_free(kd);
  free(q);
  free(idx);
  free(d2);
  return m > 0 ? s / (double)m : DBL_MAX;
}

/* ------------------------------------------------------------------------ */
/* voxels                                                                   */
/* ------------------------------------------------------------------------ */

#define VOX_BIAS (1 << 20)
uint64_t orc_voxel_key(float x, float y, float z, float inv_h, int* ok) {
  float v[3] = {x * inv_h, y * inv_h, z * inv_h};
  uint64_t key = 0;
  *ok = 1;
  for (int a = 0; a < 3; ++a) {
    if (!isfinite(v[a])) {
      *ok = 0;
      return 0;
    }
    float f = floorf(v[a]);
    if (f < -(float)VOX_BIAS || f >= (float)VOX_BIAS) {
      *ok = 0;
      return 0;
    }
    int64_t k = (int64_t)f + VOX_BIAS;
    key |= ((uint64_t)k) << (21 * a);
  }
  return key;
}

typedef struct {
  uint64_t key;
  int64_t i;
} keyidx;
static int cmp_keyidx(const void* a, const void* b) {
  const keyidx* x = (const keyidx*)a;
  const keyidx* y = (const keyidx*)b;
  if (x->key < y->key) return -1;
  if (x->key > y->key) return 1;
  return (x->i > y->i) - (x->i < y->i);
}

int64_t orc_voxel_accumulate(const float* pts, int64_t n, const double T[16], float inv_h,
                             int64_t* keys, double* sums, int64_t* counts, int64_t cap,
                             int64_t* dropped) {
  float m[12];
  double I[16];
  mat4_identity(I);
  tf_to_f32(T ? T : I, m);
  keyidx* ki = (keyidx*)malloc((size_t)(n > 0 ? n : 1) * sizeof(keyidx));
  float* tp = (float*)malloc((size_t)(n > 0 ? n : 1) * 12);
  int64_t m_ok = 0, drop = 0;
  for (int64_t i = 0; i < n; ++i) {
    float o[3];
    if (T)
      apply_f32(m, pts + 3 * i, o);
    else
      memcpy(o, pts + 3 * i, 12);
    memcpy(tp + 3 * i, o, 12);
    int ok;
    uint64_t k = orc_voxel_key(o[0], o[1], o[2], inv_h, &ok);
    if (!ok) {
      ++drop;
      continue;
    }
    ki[m_ok].key = k;
    ki[m_ok].i = i;
    ++m_ok;
  }
  qsort(ki, (size_t)m_ok, sizeof(keyidx), cmp_keyidx);
  int64_t nv = 0;
  for (int64_t s = 0; s < m_ok;) {
    int64_t e = s;
    while (e < m_ok && ki[e].key == ki[s].key) ++e;
    if (nv >= cap) {
      nv = -1;
      break;
    }
    keys[nv] = (int64_t)ki[s].key;
    double acc[3] = {0, 0, 0};
    for (int64_t j = s; j < e; ++j)
      for (int a = 0; a < 3; ++a) acc[a] += tp[3 * ki[j].i + a];
    for (int a = 0; a < 3; ++a) sums[3 * nv + a] = acc[a];
    counts[nv] = e - s;
    ++nv;
    s = e;
  }
  if (dropped) *dropped = drop;
  free(ki);
  free(tp);
  return nv;
}

void orc_normals(const float* pts, int64_t n, int k, const float vp[3], float* normals,
                 float* curvature) {
  if (k > n) k = (int)n;
  if (k < 1) return;
  orc_kdtree* kd = orc_kd_build(pts, n, 15);
  int32_t* idx = (int32_t*)malloc((size_t)k * 4);
  float* d2 = (float*)malloc((size_t)k * 4);
  for (int64_t i = 0; i < n; ++i) {
    orc_kd_knn(kd, pts + 3 * i, 1, k, idx, d2);
    double mean[3] = {0, 0, 0};
    int cnt = 0;
    for (int j = 0; j < k; ++j)
      if (idx[j] >= 0) {
        for (int a = 0; a < 3; ++a) mean[a] += pts[3 * (int64_t)idx[j] + a];
        ++cnt;
      }
    if (cnt < 3) {
      for (int a = 0; a < 3; ++a) normals[3 * i + a] = NAN;
      if (curvature) curvature[i] = NAN;
      continue;
    }
    for (int a = 0; a < 3; ++a) mean[a] /= cnt;
    double C[9] = {0};
    for (int j = 0; j < k; ++j) {
      if (idx[j] < 0) continue;
      double d[3];
      for (int a = 0; a < 3; ++a) d[a] = pts[3 * (int64_t)idx[j] + a] - mean[a];
      for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) C[r * 3 + c] += d[r] * d[c];
    }
    for (int e = 0; e < 9; ++e) C[e] /= cnt;
    double w[3], V[9];
    orc_eig3(C, w, V);
    double nv[3] = {V[0], V[3], V[6]};
    double vx = vp[0] - pts[3 * i], vy = vp[1] - pts[3 * i + 1], vz = vp[2] - pts[3 * i + 2];
    if (vx * nv[0] + vy * nv[1] + vz * nv[2] < 0) {
      nv[0] = -nv[0];
      nv[1] = -nv[1];
      nv[2] = -nv[2];
    }
    for (int a = 0; a < 3; ++a) normals[3 * i + a] = (float)nv[a];
    if (curvature) {
      double sw = w[0] + w[1] + w[2];
      curvature[i] = sw > 0 ? (float)(w[0] / sw) : 0.0f;
    }
  }
  free(idx);
  free(d2);
  orc_kd_free(kd);
}

int64_t orc_voxel_downsample(const float* pts, int64_t n, float leaf, float* out, int64_t cap,
                             int64_t* counts) {
  float inv = (float)(1.0 / (double)leaf);
  int64_t lo[3] = {INT64_MAX, INT64_MAX, INT64_MAX}, hi[3] = {INT64_MIN, INT64_MIN, INT64_MIN};
  int64_t* ijk = (int64_t*)malloc((size_t)(n > 0 ? n : 1) * 3 * sizeof(int64_t));
  char* ok = (char*)malloc((size_t)(n > 0 ? n : 1));
  for (int64_t i = 0; i < n; ++i) {
    ok[i] = finite3(pts + 3 * i);
    if (!ok[i]) continue;
    for (int a = 0; a < 3; ++a) {
      int64_t v = (int64_t)floorf(pts[3 * i + a] * inv);
      ijk[3 * i + a] = v;
      if (v < lo[a]) lo[a] = v;
      if (v > hi[a]) hi[a] = v;
    }
  }
  keyidx* ki = (keyidx*)malloc((size_t)(n > 0 ? n : 1) * sizeof(keyidx));
  int64_t m = 0;
  int64_t dx = hi[0] - lo[0] + 1, dy = hi[1] - lo[1] + 1;
  for (int64_t i = 0; i < n; ++i) {
    if (!ok[i]) continue;
    ki[m].key = (uint64_t)((ijk[3 * i] - lo[0]) + (ijk[3 * i + 1] - lo[1]) * dx +
                           (ijk[3 * i + 2] - lo[2]) * dx * dy);
    ki[m].i = i;
    ++m;
  }
  qsort(ki, (size_t)m, sizeof(keyidx), cmp_keyidx);
  int64_t nv = 0;
  for (int64_t s = 0; s < m;) {
    int64_t e = s;
    while (e < m && ki[e].key == ki[s].key) ++e;
    if (nv >= cap) {
      nv = -1;
      break;
    }
    double acc[3] = {0, 0, 0};
    for (int64_t j = s; j < e; ++j)
      for (int a = 0; a < 3; ++a) acc[a] += pts[3 * ki[j].i + a];
    for (int a = 0; a < 3; ++a) out[3 * nv + a] = (float)(acc[a] / (double)(e - s));
    if (counts) counts[nv] = e - s;
    ++nv;
    s = e;
  }
  free(ijk);
  free(ok);
  free(ki);
  return nv;
}

/* ------------------------------------------------------------------------ */
/* PCL-semantics CPU baseline                                               */
/* ------------------------------------------------------------------------ */

/* float Umeyama as PCL's TransformationEstimationSVD (Scalar = float): means
 * and the cross-covariance are formed in float, the 3x3 SVD result is
 * rounded back to float. */
static void umeyama_f32(const float* a, const float* b, const int32_t* idx, int64_t ns,
                        int64_t ncorr, float Tf[16]) {
  float mx[3] = {0, 0, 0}, my[3] = {0, 0, 0};
  for (int64_t i = 0; i < ns; ++i) {
    if (idx[i] < 0) continue;
    for (int k = 0; k < 3; ++k) {
      mx[k] += a[3 * i + k];
      my[k] += b[3 * (int64_t)idx[i] + k];
    }
  }
  for (int k = 0; k < 3; ++k) {
    mx[k] /= (float)ncorr;
    my[k] /= (float)ncorr;
  }
  float S[9] = {0};
  for (int64_t i = 0; i < ns; ++i) {
    if (idx[i] < 0) continue;
    float x[3], y[3];
    for (int k = 0; k < 3; ++k) {
      x[k] = a[3 * i + k] - mx[k];
      y[k] = b[3 * (int64_t)idx[i] + k] - my[k];
    }
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) S[r * 3 + c] += y[r] * x[c];
  }
  double Sd[9], R[9];
  for (int k = 0; k < 9; ++k) Sd[k] = (double)(S[k] / (float)ncorr);
  rot_from_cov(Sd, R);
  memset(Tf, 0, 16 * sizeof(float));
  Tf[15] = 1.0f;
  for (int r = 0; r < 3; ++r) {
    for (int c = 0; c < 3; ++c) Tf[c * 4 + r] = (float)R[r * 3 + c];
    float Rm = Tf[r] * mx[0];
    Rm = Rm + Tf[4 + r] * mx[1];
    Rm = Rm + Tf[8 + r] * mx[2];
    Tf[12 + r] = my[r] - Rm;
  }
}

int orc_pcl_icp(const float* src, int64_t ns, const float* tgt, int64_t nt, const orc_icp_params* p,
                const double guess[16], orc_icp_result* out) {
  memset(out, 0, sizeof(*out));
  orc_kdtree* kd = orc_kd_build(tgt, nt, 15); /* initCompute: tree over the target */
  float* w = (float*)malloc((size_t)(ns > 0 ? ns : 1) * 12);
  int32_t* idx = (int32_t*)malloc((size_t)(ns > 0 ? ns : 1) * 4);
  float* d2 = (float*)malloc((size_t)(ns > 0 ? ns : 1) * 4);
  double T[16];
  if (guess)
    memcpy(T, guess, sizeof(T));
  else
    mat4_identity(T);
  orc_transform_f32(T, src, ns, w);
  float r2f = orc_r2_to_float(p->r_max * p->r_max);
  conv_state cs = {0, DBL_MAX};
  int state = ORC_NOT_CONVERGED, iters = 0;
  int64_t ncorr = 0;
  double mse = 0;
  for (;;) {
    orc_kd_nn(kd, w, ns, r2f, 0, idx, d2); /* FLANN: unbounded, first found in tree order */
    ncorr = 0;
    double sd2 = 0;
    for (int64_t i = 0; i < ns; ++i)
      if (idx[i] >= 0) {
        ++ncorr;
        sd2 += d2[i];
      }
    if (ncorr < (p->min_corr > 0 ? p->min_corr : 3)) {
      state = ORC_FAIL_NO_CORRESPONDENCES;
      break;
    }
    mse = sd2 / (double)ncorr;
    float Tf[16];
    umeyama_f32(w, tgt, idx, ns, ncorr, Tf);
    double Tinc[16];
    for (int k = 0; k < 16; ++k) Tinc[k] = Tf[k];
    /* the working cloud is re-transformed cumulatively in float */
    orc_transform_f32(Tinc, w, ns, w);
    double Tn[16];
    mat4_mul(Tinc, T, Tn);
    for (int k = 0; k < 16; ++k) T[k] = (double)(float)Tn[k];
    ++iters;
    state = convergence(p, iters, Tinc, mse, &cs);
    if (state != ORC_NOT_CONVERGED) break;
  }
  memcpy(out->T, T, sizeof(T));
  out->iterations = iters;
  out->state = state;
  out->converged = (state != ORC_NOT_CONVERGED && state != ORC_FAIL_NO_CORRESPONDENCES);
  out->n_corr = ncorr;
  out->mse = mse;
  orc_kd_free(kd);
  free(w);
  free(idx);
  free(d2);
  return 0;
}

typedef struct {
  int npairs;
  const float* const* src;
  const int64_t* ns;
  const float* const* tgt;
  const int64_t* nt;
  const orc_icp_params* p;
  const double* guesses;
  orc_icp_result* out;
  int next;
  pthread_mutex_t mu;
} batch_ctx;

static void* batch_worker(void* arg) {
  batch_ctx* c = (batch_ctx*)arg;
  for (;;) {
    pthread_mutex_lock(&c->mu);
    int i = c->next++;
    pthread_mutex_unlock(&c->mu);
    if (i >= c->npairs) break;
    orc_pcl_icp(c->src[i], c->ns[i], c->tgt[i], c->nt[i], c->p,
                c->guesses ? c->guesses + 16 * (int64_t)i : NULL, &c->out[i]);
  }
  return NULL;
}

int orc_pcl_icp_batch(int npairs, const float* const* src, const int64_t* ns,
                      const float* const* tgt, const int64_t* nt, const orc_icp_params* p,
                      const double* guesses, orc_icp_result* out, int nthreads) {
  batch_ctx c;
  c.npairs = npairs;
  c.src = src;
  c.ns = ns;
  c.tgt = tgt;
  c.nt = nt;
  c.p = p;
  c.guesses = guesses;
  c.out = out;
  c.next = 0;
  pthread_mutex_init(&c.mu, NULL);
  if (nthreads < 1) nthreads = 1;
  pthread_t* th = (pthread_t*)malloc((size_t)nthreads * sizeof(pthread_t));
  for (int t = 0; t < nthreads; ++t) pthread_create(&th[t], NULL, batch_worker, &c);
  for (int t = 0; t < nthreads; ++t) pthread_join(th[t], NULL);
  free(th);
  pthread_mutex_destroy(&c.mu);
  return 0;
}

/* ---- farthest-point sampling (learned proposer input) ---- */
void orc_fps(const float* pts, int64_t n, int64_t max_sub, int k, int32_t* sel, double* centroid) {
  const int64_t m = n < max_sub ? n : max_sub;
  float* sp = (float*)malloc(sizeof(float) * 3 * (size_t)(m > 0 ? m : 1));
  float* dist = (float*)malloc(sizeof(float) * (size_t)(m > 0 ? m : 1));
  double c[3] = {0, 0, 0};
  for (int64_t i = 0; i < m; ++i) {
    const int64_t src = (int64_t)(((uint64_t)i * (uint64_t)n) / (uint64_t)m);
    for (int a = 0; a < 3; ++a) {
      sp[3 * i + a] = pts[3 * src + a];
      c[a] += sp[3 * i + a];
    }
    dist[i] = 3.0e38f;
  }
  for (int a = 0; a < 3; ++a) centroid[a] = m > 0 ? c[a] / (double)m : 0.0;
  int64_t cur = 0;
  for (int s = 0; s < k; ++s) {
    sel[s] = (int32_t)cur;
    if (m == 0) continue;
    const float qx = sp[3 * cur], qy = sp[3 * cur + 1], qz = sp[3 * cur + 2];
    float bv = -1.0f;
    int64_t bi = 0;
    for (int64_t i = 0; i < m; ++i) {
      const float dx = sp[3 * i] - qx, dy = sp[3 * i + 1] - qy, dz = sp[3 * i + 2] - qz;
      float d = dx * dx;
      d = d + dy * dy;
      d = d + dz * dz;
      if (d < dist[i]) dist[i] = d;
      if (dist[i] > bv) {
        bv = dist[i];
        bi = i;
      }
    }
    cur = bi;
  }
  free(sp);
  free(dist);
}

/* ---- truncated distance field (restates kernel.cu:9-31, bounds fixed) ---- */
void orc_tdf(const int32_t* occ, int64_t n_occ, const int dims[3], float trunc2, float* out) {
  const int dx = dims[0], dy = dims[1], dz = dims[2];
  for (int z = 0; z < dz; ++z)
    for (int y = 0; y < dy; ++y)
      for (int x = 0; x < dx; ++x) {
        float d = trunc2;
        for (int64_t i = 0; i < n_occ; ++i) {
          const int ox = occ[3 * i], oy = occ[3 * i + 1], oz = occ[3 * i + 2];
          if (ox < 0 || oy < 0 || oz < 0 || ox >= dx || oy >= dy || oz >= dz) continue;
          const int a = x - ox, b = y - oy, c = z - oz;
          const float t = (float)(a * a + b * b + c * c);
          if (t < d) d = t;
        }
        out[x + (int64_t)dx * (y + (int64_t)dy * z)] = d;
      }
}
