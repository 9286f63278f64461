// Host-side launch wrappers of the training kernels (kernels.hip).
#pragma once
#include <hip/hip_runtime.h>
#include "bh_common.hpp"

namespace bh {

// weight: per-row weights scaling g and h (null: unweighted).  obj 0 mse, 1 logloss,
// 2 poisson.
// fx (4 words): the FP32 histogram's fixed-point grids, s_g and s_h in fx[2..3], set from
// these gradients' maxima (fx[0..1] are zero between passes).
void launch_grad_f32(hipStream_t s, float const *score, float const *label, float const *weight, u64 n, int obj,
                     float2 *gh, u32 *fx);
// The grids of host-supplied FP32 gradients.
void launch_fx_from(hipStream_t s, float2 const *gh, u64 n, u32 *fx);
// Quantized gradients of iteration ctl->iter with the max-abs in maxbits slot iter & 1
// (two slots of two words); clears the other slot for the score update to fill.
void launch_grad_quant(hipStream_t s, float const *score, float const *label, float const *weight, u64 n, int obj,
                       u32 *maxbits, u64 seed, TreeCtl const *ctl, u32 *ghq, double *scales);
// Fills maxbits slot ctl->iter & 1 from the current scores and labels (when they were
// set outside the iteration loop).
void launch_grad_prime(hipStream_t s, float const *score, float const *label, float const *weight, u64 n, int obj,
                       u32 *maxbits, TreeCtl const *ctl);
void launch_quant_from(hipStream_t s, float2 const *gh, u64 n, u32 *maxbits, u64 seed, u32 iter, u32 *ghq,
                       double *scales_dev);
hipError_t launch_hist(hipStream_t s, bool q, HistChunk const *chunks, TreeCtl const *ctl, u8 const *pl0,
                       u8 const *pl1, u8 const *plp, void const *g0, void const *g1, void const *gp, u32 Fs, u32 F,
                       u32 SB, u32 fsl, u32 n_waves, u32 grid_x, void *partials, void *hist_next, u32 const *fx,
                       u32 *cnt_next = nullptr, PartOp const *routes = nullptr, bool group = false);
hipError_t set_hist_lds_limits();
// F when launch_hist with these arguments runs the rotated row-in-lane kernel, whose unit
// partials are in (bin, feature) order (launch_finish's tF), else 0.
u32 hist_transpose_F(bool q, u32 Fs, u32 F, bool group = false);
// advance: the level ends with this completion; the frontier moves on with it.
void launch_finish(hipStream_t s, bool q, FinishOp const *ops, TreeCtl *ctl, void const *partials,
                   void const *hist_cur, void *hist_next, u32 node_cells, u32 max_ops,
                   u32 const *cnt_cur = nullptr, u32 *cnt_next = nullptr, bool advance = false, u32 tF = 0);
void launch_find(hipStream_t s, bool q, NodeInfo *front, TreeCtl const *ctl, void const *hist, u16 const *n_bins,
                 u32 F, u32 SB, GrowCfg const &cfg, double const *scales, SplitRec *out, u32 max_front,
                 u32 front_base, u32 n_front_override, FindScratch const &fs, u32 const *cnt = nullptr);
u32 find_blocks_per_node(u32 F, u32 tile); // workgroups per node for tile-feature tiles
void launch_plan_level(hipStream_t s, NodeInfo const *front, SplitRec const *splits, TreeCtl *ctl, DevNode *tree,
                       PartOp *ops, PartChunk *pch, u16 const *n_bins, GrowCfg const &cfg, double const *scales,
                       u32 dst, u32 chunk_rows);
// Rows per partition chunk for rows of Fs bytes (kPartChunk, smaller for wide rows).
u32 part_chunk_rows(u32 Fs);
void launch_part_count(hipStream_t s, PartChunk const *pch, PartOp const *ops, TreeCtl const *ctl, u8 const *pl0,
                       u8 const *pl1, u8 const *plp, u32 Fs, u32 *chunk_left, u32 grid);
void launch_part_scan(hipStream_t s, bool q, PartOp *ops, PartChunk *pch, u32 const *chunk_left, TreeCtl *ctl,
                      NodeInfo const *front, NodeInfo *next, SplitRec const *splits, DevNode *tree, HistChunk *hch,
                      FinishOp *fops, GrowCfg const &cfg, double const *scales, u32 is_last, u32 oblivious);
void launch_part_scatter(hipStream_t s, PartChunk const *pch, PartOp const *ops, TreeCtl const *ctl,
                         u8 const *pl0, u8 const *pl1, u8 const *plp, void const *g0, void const *g1,
                         void const *gp, u8 *o0, u8 *o1, void *go0, void *go1, u32 Fs, u32 gh_bytes, u32 grid,
                         u32 const *chunk_left = nullptr);
u32 scatter_tile_rows(u32 Fs, u32 gh_bytes);
void launch_plan_counted(hipStream_t s, NodeInfo const *front, SplitRec const *splits, TreeCtl *ctl, DevNode *tree,
                         PartOp *ops, PartChunk *pch, NodeInfo *next, HistChunk *hch, FinishOp *fops,
                         u16 const *n_bins, GrowCfg const &cfg, double const *scales, u32 dst, u32 is_last,
                         u32 chunk_rows, u32 mode = 0, PartOp4 *ops4 = nullptr, u32 unit_rows = kHistChunk,
                         u32 *cursors = nullptr);
// Four-way partition of grandparent segments (quantized depthwise, rows of at most 64
// bytes): per-chunk class counts, then the LDS-staged scatter; with cursors (zeroed by
// the planner), no counting pass and tiles reserve their class runs atomically.
void launch_part_scatter4(hipStream_t s, PartChunk const *pch, PartOp4 const *ops4, TreeCtl *ctl, u8 const *pl0,
                          u8 const *pl1, u8 const *plp, u32 const *g0, u32 const *g1, u32 const *gp, u8 *o0,
                          u8 *o1, u32 *go0, u32 *go1, u32 Fs, u32 grid, uint4 const *chunk_cls,
                          u32 *cursors = nullptr);
void launch_part_count4(hipStream_t s, PartChunk const *pch, PartOp4 const *ops4, TreeCtl const *ctl, u8 const *pl0,
                        u8 const *pl1, u8 const *plp, u32 Fs, uint4 *chunk_cls, u32 grid);
void launch_finalize(hipStream_t s, NodeInfo const *front, TreeCtl const *ctl, DevNode *tree, GrowCfg const &cfg,
                     double const *scales, u32 max_front);
void launch_next_level(hipStream_t s, TreeCtl *ctl);
// Validation loss of one round (objective 0 MSE, 1 logloss, 2 poisson) and the early-stopping rule
// in the control block; eval_blocks(n) partials of scratch.
u32 eval_blocks(u32 n);
void launch_eval(hipStream_t s, int objective, float const *scores, float const *labels, u32 n, double *partial,
                 TreeCtl *ctl, float *hist, u32 hist_cap, u32 patience);
// unit_rows: rows per root histogram unit (kHistChunk, or a group of sub-units).
void launch_tree_begin(hipStream_t s, TreeCtl *ctl, NodeInfo *front, DevNode *tree, HistChunk *hch,
                       FinishOp *fops, RootSel const &root, u32 unit_rows = kHistChunk);
// Device Bernoulli row sampler: the rows of [0, n) kept in this iteration (threshold
// thr on 32 random bits), ascending, with their gradient words, into out_plane/out_gh;
// their count into ctl->root_n.  chunk_cnt: n / kPartChunk + 1 words of scratch.
void launch_bernoulli(hipStream_t s, TreeCtl *ctl, u32 n, u64 seed, u32 thr, u32 *chunk_cnt, u8 const *plane,
                      u32 Fs, void const *gh, u32 gh_bytes, u8 *out_plane, void *out_gh);
// Per-tree feature selection for iteration ctl->iter: mask[f] = 1 for the k smallest
// (key, f) pairs.
void launch_feature_draw(hipStream_t s, TreeCtl const *ctl, u32 F, u32 k, u64 seed, u8 *mask);
// Monotone / interaction constraints at the start of a tree: node 0's [lo, hi] box =
// (-inf, +inf) and its path empty (either pointer may be null).
void launch_constraints_root(hipStream_t s, double *bounds, u32 *paths, u32 path_words);
// maxbits (training, quantized): also the next iteration's gradient maxima, from the
// updated scores and the labels, into slot (ctl->iter + 1) & 1.
void launch_score_update(hipStream_t s, u8 const *plane, u32 Fs, u64 n, DevNode const *tree, TreeCtl const *ctl,
                         u16 const *n_bins, u32 F, float lr, float *score, u32 *leaf_of_row, u32 max_nodes,
                         float const *label = nullptr, int objective = 0, u32 *maxbits = nullptr,
                         float const *weight = nullptr);
void launch_iter_advance(hipStream_t s, TreeCtl *ctl);
void launch_copy_tree(hipStream_t s, DevNode const *tree, TreeCtl const *ctl, DevNode *out_trees, u32 *out_sizes,
                      u32 max_nodes);
void set_tuning(Tuning const &t);
Tuning const &tuning();
void launch_copy_calibrate(hipStream_t s, void const *src, void *dst, u64 bytes, u32 mode = 0, u32 grid = 0);
void launch_bin_rows(hipStream_t s, float const *X, u64 n, u32 F, u32 Fs, float const *cuts, u32 const *coff,
                     u32 n_cuts_total, u8 *plane);
// Four columns of features f0..f0+G-1 (column-major, leading dimension ld) into the plane.
void launch_bin_cols4(hipStream_t s, float const *cols, u64 ld, u64 n, u32 G, float const *cuts, u32 const *coff,
                      u32 f0, u32 Fs, u8 *plane);
// out[i][:] = X[rows[i]][:] (float rows of F features).
void launch_gather_f32_rows(hipStream_t s, float const *X, u32 const *rows, u64 n_out, u32 F, float *out);
void launch_gather_rows(hipStream_t s, u8 const *plane, u32 Fs, void const *gh, u32 gh_bytes, u32 const *rows,
                        u64 n, u8 *out_plane, void *out_gh, int gh_is_row_id);

} // namespace bh
