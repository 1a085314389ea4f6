// Device-side view of a registration engine, passed by value to kernels,
// and the launchers implemented in kernels_icp.hip.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "rtr_types.hpp"

namespace rtr {

constexpr int kBlock = 256;        // threads per workgroup (4 wave64)
constexpr int kWaves = kBlock / 64;
constexpr int kXcds = 8;           // MI355X: 8 XCDs, workgroups dealt round-robin by blockIdx

struct DevEngine {
  // targets: raw packed xyz, descriptors, cell-sorted points and grids
  const float* tgt_raw;
  const float* tgt_nrm_raw;  // packed normals aligned with tgt_raw (nullable)
  TargetDesc* tgts;
  int n_tgt;
  float4* tgt_sorted;        // (x, y, z, original index bits)
  float4* tgt_nrm_sorted;    // normals in sorted order (nullable)
  uint32_t* cell_count;
  uint32_t* cell_start;
  uint32_t* cell_id;         // per raw target point
  uint32_t* cell_rank;       // per raw target point: its slot within its cell
  uint8_t* cell_dist;        // per cell: Chebyshev distance (in cells) to the nearest occupied cell
  uint8_t* cell_dist_tmp;    // ping-pong buffer of the separable distance transform
  float r_max;               // max correspondence distance (inf when unbounded)
  uint32_t n_cells_total;    // sum of (budget + 1)
  // sources: raw packed xyz, spatially sorted SoA and the permutation
  const float* src_raw;
  SourceDesc* srcs;
  int n_src;
  float* sx;
  float* sy;
  float* sz;
  uint32_t* perm;            // sorted position -> original index
  uint32_t* bin_count;
  uint32_t* bin_start;
  uint32_t* bin_id;          // per raw source point
  uint32_t* bin_rank;        // per raw source point: its slot within its bin
  uint32_t n_bins_total;
  // problems and the work queue
  const Problem* probs;
  int n_prob;
  ProblemState* st;
  const double* guesses;     // n_prob * 16 (column-major)
  float* hist;               // n_prob * (max_iter + 1) * 12: float rows of T used by each iteration
  uint64_t* tline;           // n_prob * max_iter * 2 wall-clock stamps: start and end of each iteration's
                             // solve (nullable: timeline off)
  double* aa;                // n_prob * kAndersonStride: acceleration state (null when off)
  uint32_t* nn_seed;         // per problem and sorted source point (at Problem::pt_off): cell-sorted
                             // slot of its last nearest neighbour, the next search's seed (nullable)
  const uint2* items;        // (problem, first sorted source point)
  int n_items;
  int item_pts;              // points per item
  int use_tile;              // stage candidate tiles in LDS (else global gathers only)
  int variant;               // search kernel (point-to-point): 0 4-deep gathers at 3 waves per SIMD, 1 8-deep
                             // at 2, 2 4-deep at 4 (point-to-plane: always 8-deep at 2)
  int retire_from;           // workgroups from this index on leave at the start of an iteration ...
  uint32_t retire_active;    // ... once at most this many problems are unfinished
  double* partials;          // n_items * NSUM_MAX FP64 moment sums
  uint32_t* claim;           // claim counter per (pass, XCD queue), claim_stride words apart
  int claim_stride;          // 32 (one 128-byte line per counter) unless the iteration limit is huge
  uint32_t xcd_off[kXcds + 1];  // item range of each XCD queue
  uint32_t fxcd_off[kXcds + 1]; // fine item range of each XCD queue
  int fine_from;             // passes from this one on run the fine items (INT_MAX: never)
  uint32_t* done;            // per problem: items finished so far
  uint32_t* problems_done;
  int* error;                // nonzero: a bounded wait expired
  unsigned long long* queries_searched;  // nearest-neighbour searches of the launch
  uint32_t* scan_tmp;        // block sums for the scans
  uint32_t* bbox_tmp;        // per target, then per source: 8 words of bounding-box partials
  IcpParams prm;
};

// launchers (stream-ordered, no host synchronisation)
void launch_fill_u32(uint32_t* p, uint32_t v, size_t n, hipStream_t s);
void launch_exclusive_scan(const uint32_t* in, uint32_t* out, uint32_t n, uint32_t* tmp,
                           hipStream_t s);
size_t scan_tmp_words(uint32_t n);
void launch_target_grid(const DevEngine& e, int max_tgt_pts, int max_cells, hipStream_t s);
void launch_source_sort(const DevEngine& e, int max_src_pts, uint32_t src_total, hipStream_t s);
void launch_icp(const DevEngine& e, int grid_blocks, hipStream_t s);
int icp_blocks_per_cu(int mode, int use_tile, int variant);
// correspondences / fitness / transformed output for problem set; T per
// problem from `T` (n_prob*16 column-major, device) or from the state when null.
void launch_correspondences(const DevEngine& e, int prob, const double* T, float r2f, int32_t* idx,
                            float* d2, hipStream_t s);
void launch_transform_out(const DevEngine& e, int prob, float* out, hipStream_t s);
void launch_fitness_d2(const DevEngine& e, float* d2_all, hipStream_t s);
constexpr int kJudgeSplit = 8;  // workgroups per problem in the registration judge
void launch_judge(const DevEngine& e, int k, int max_pts, float r2f, double max_curv, double* out,
                  hipStream_t s);
void launch_target_normals(const DevEngine& e, int k, float vpx, float vpy, float vpz,
                           int max_tgt_pts, hipStream_t s);

}  // namespace rtr
