// Plain-data types shared by host and device code.
#pragma once
#include <stdint.h>

#include "../../include/rtr.h"

namespace rtr {

// Moments reduced per registration and iteration.  Slots:
//   point-to-point: n, sum d2, sum a (3), sum b (3), sum a b^T (9)       = 17
//   point-to-plane: n, sum d2, A^T A upper (21), A^T b (6)               = 29
// a = transformed source point - anchor, b = target point - anchor.
enum {
  SUM_N = 0,
  SUM_D2 = 1,
  SUM_A = 2,
  SUM_B = 5,
  SUM_AB = 8,
  SUM_P2P_COUNT = 17,
  SUM_ATA = 2,
  SUM_ATB = 23,
  SUM_P2L_COUNT = 29,
  NSUM_MAX = 32
};

// Read-only view of one cell-sorted uniform grid.
struct GridView {
  const float4* pts;           // cell-sorted (x, y, z, original index bits)
  const uint32_t* cell_start;  // dx*dy*dz + 1 global positions into pts
  int dx, dy, dz;
  uint32_t n_pts;
  uint32_t pt_lo;              // the grid's points are pts[pt_lo, pt_lo + n_pts)
  float ox, oy, oz;
  float h, inv_h, margin;
};

constexpr int kAndersonMax = 8;       // largest Anderson acceleration memory m
constexpr int kAndersonStride = 120;  // doubles of acceleration state per problem: 12 + 12 (m + 1)

struct IcpParams {
  int max_iter;
  int min_corr;
  int mode;
  float r2f;         // max correspondence distance^2 rounded down to float
  double eps_T;      // transformation epsilon
  double eps_rot;    // rotation epsilon (<= 0: 1 - eps_T)
  double eps_fit;    // euclidean fitness epsilon
  int anderson;      // Anderson acceleration memory (0: plain iteration)
  int coarse_stride; // > 1: iterations start on every coarse_stride-th source point
  int coarse_iters;  // most coarse iterations
};

// Per target cloud (device).  Grid fields are written by the grid build.
struct TargetDesc {
  uint32_t raw_off;      // first point in the packed raw xyz array (and in the sorted array)
  uint32_t n;            // raw point count
  uint32_t cell_off;     // first cell slot (cell_start has budget + 1 entries)
  uint32_t cell_budget;  // max cells for this target
  int dx, dy, dz;
  uint32_t n_valid;      // finite points placed in the grid
  float ox, oy, oz;
  float h, inv_h, margin;
  float anchor[3];       // bbox centre, origin of the accumulated moments
  float h_req;           // requested minimum cell edge
  int dt_cap;            // distance-field cap: values > dt_cap are stored as dt_cap + 1
  int skip_ring;         // a query whose cell has distance >= skip_ring has no match
};

// Per source cloud (device).
struct SourceDesc {
  uint32_t off;          // first point in the source arrays
  uint32_t n;
  uint32_t bin_off;      // first bin of the spatial sort
  uint32_t bin_budget;
  int dx, dy, dz;
  float ox, oy, oz, inv_h;
};

// One registration = (source, target, initial guess).
struct Problem {
  uint32_t src, tgt;
  uint32_t item_off, n_items;    // its work items (item_pts source points each)
  uint32_t pt_off;               // prefix sum of source sizes over problems (per-point outputs)
  uint32_t fitem_off, n_fitems;  // its fine work items (kFinePts each) for the passes from fine_from on
  uint32_t pad_;
};

constexpr int kFinePts = 256;  // source points per fine work item

struct ProblemState {
  double T[16];      // current estimate, column-major
  double prev_mse;
  double mse;        // MSE of the last iteration's correspondences
  double n_corr;     // correspondences of the last iteration
  int iter;
  int state;         // RTR_CONV_* / RTR_FAIL_*
  int finished;
  int gen;           // number of completed iterations visible to waiters
  int stage;         // 1 while iterating on the coarse source subsample
  int coarse_its;    // iterations run on it
};

}  // namespace rtr
