/*
 * CPU oracle for the bonsai GBDT hot path — TEST INFRASTRUCTURE ONLY.
 *
 * A plain C++ restatement of the reference's CPU algorithm (bin mapping, logloss/mse
 * gradients, histograms, node and level split finders, stable partition, the
 * depthwise / leafwise / levelwise control loops, bin-space predict, TreeSHAP and the
 * rank-sum AUC).  Each function cites the reference file:line it follows.
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load this
 * library, and only as the checker or the CPU baseline.  The product (bonsai_amd/)
 * never links it.
 *
 * Parity pins (see tests/): golden vectors restated from the reference's own unit
 * tests, closed forms, brute force on tiny inputs, and invariants.  Functions without
 * such a check say "parity unpinned" beside their declaration.
 */
#ifndef BONSAI_ORACLE_H
#define BONSAI_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* TreeConfig subset the hot path reads (include/bonsai/config/tree_config.hpp:9-29). */
typedef struct orc_tree_cfg {
    float    min_child_hess;
    float    min_gain_to_split;
    float    lambda_l2;
    float    lambda_l1;
    int32_t  max_depth;
    int32_t  min_data_in_leaf;
    int32_t  max_leaves;
    int32_t  _pad;
} orc_tree_cfg;

/* SplitOutput (include/bonsai/split.hpp:66-73). */
typedef struct orc_split {
    double   gain;
    uint32_t feature;
    uint32_t bin;
    uint32_t default_left;
    uint32_t valid;
} orc_split;

/* One dense-tree node in bin space (include/bonsai/tree.hpp:32-48 plus covers and
 * split gains).  feature < 0 marks a leaf whose value is `value`. */
typedef struct orc_node {
    int32_t  feature;
    uint32_t bin;
    uint32_t default_left;
    uint32_t left;
    uint32_t right;
    float    value;
    float    cover;
    float    gain;
} orc_node;

/* Histogram precisions the growers can run under. */
enum {
    ORC_HIST_F32_SERIAL = 0, /* float cells, one sequential pass in row order (the
                                reference CPU single-block fill, test_populate.cpp:160-182) */
    ORC_HIST_F32_CHUNKED = 1,/* float cells: sequential per fixed row chunk, chunk partials
                                folded in chunk order (the device's fixed-order reduction) */
    ORC_HIST_INT = 2,        /* int64 sums of int8/int16 quantized gradients (exact) */
    ORC_HIST_F32_OMP = 3,    /* float cells, OpenMP row blocks with per-thread partials
                                reduced in block order (CPU baseline fill) */
    ORC_HIST_F32_FIXED = 4   /* float cells: per fixed row chunk an exact fixed-point sum
                                (grid 2^-s, s = orc_fx_shift of the gradients) rounded once
                                to float, chunk partials folded in chunk order (the
                                device's FP32 histogram) */
};

/* Fixed-point grid of ORC_HIST_F32_FIXED: s = 49 - e with max|v| < 2^e over v[0..n)
 * (frexpf), 0 when every value is 0 or the maximum is not finite.  A value goes on the
 * grid as rintf(ldexpf(v, s)) (|q| < 2^49, so 8,192 of them sum exactly in int64) and a
 * chunk's sum comes back as ldexpf((float)sum, -s). */
int32_t orc_fx_shift(const float *v, uint64_t n);
/* ORC_HIST_F32_FIXED histogram of rows with the given grids (sg for g, sh for h). */
void orc_hist_f32_fixed(const uint8_t *plane, uint64_t row_stride, uint32_t n_features,
                        uint32_t bin_stride, const uint32_t *rows, uint64_t n_rows, const float *g,
                        const float *h, int32_t sg, int32_t sh, uint32_t chunk, float *out);

/* ---- bin mapping: src/bin_mapper.cpp ---- */
/* Cuts for one column (n <= n_samples path: every non-NaN value), returns #cuts. */
int32_t orc_fit_cuts(const float *col, uint64_t n, uint64_t stride, int32_t max_bin,
                     float *cuts_out, int32_t cap);
/* lower_bound with NaN -> last bin (src/bin_mapper.cpp:284-291). */
void orc_transform(const float *cuts, int32_t n_cuts, const float *x, uint64_t n,
                   uint64_t x_stride, uint8_t *out, uint64_t out_stride);

/* ---- objectives: src/objective.cpp ---- */
void  orc_logloss_grad(const float *score, const float *label, uint64_t n, float *g,
                       float *h);
void  orc_mse_grad(const float *score, const float *target, uint64_t n, float *g,
                   float *h);
float orc_init_score(int32_t objective, const float *label, uint64_t n); /* 0 mse, 1 logloss, 2 poisson */
float orc_logloss_eval(const float *score, const float *label, uint64_t n);

/* Stochastic-rounding quantizer (int8 g, int16 h, per-call max-abs scale). */
void orc_poisson_grad(const float *score, const float *label, uint64_t n, float *g, float *h);
void orc_quantize(const float *g, const float *h, uint64_t n, uint64_t seed,
                  uint32_t iter, int8_t *gq, int16_t *hq, float *max_g, float *max_h);
/* Bernoulli row sampler (BernoulliSampler::sample, src/sampler.cpp:38-55) with the
 * product's counter-based draw: rows r of [0, n) whose 32 random bits of (seed, iter, r)
 * are below thr, ascending, into out; returns their count. */
uint64_t orc_bernoulli_rows(uint64_t n, uint32_t thr, uint64_t seed, uint32_t iter, uint32_t *out);
/* Per-tree feature subsample for the finders and growers that follow (mask[f] != 0:
 * selected; NULL or n 0: every feature). */
void orc_set_feature_mask(const uint8_t *mask, uint32_t n);
/* Monotone directions (+1, -1, 0 by sign) for the depthwise and leafwise growers that
 * follow; NULL or n = 0 clears. */
void orc_set_monotone(const int8_t *dirs, uint32_t n);
/* Interaction groups (group g: group_sizes[g] ids, concatenated in members) for the
 * depthwise and leafwise growers that follow; n_groups 0 clears. */
void orc_set_interaction(const uint32_t *members, const uint32_t *group_sizes, uint32_t n_groups);
/* The product's device draw for feature_fraction < 1: k = max(1, ceil(fraction * F))
 * features with the smallest counter-based keys of (seed, iter, f); mask out [F]. */
void orc_feature_draw(uint32_t n_features, float fraction, uint64_t seed, uint32_t iter, uint8_t *mask);

/* ---- histograms over a row list ---- */
void orc_hist_f64(const uint8_t *plane, uint64_t row_stride, uint32_t n_features,
                  uint32_t bin_stride, const uint32_t *rows, uint64_t n_rows,
                  const float *g, const float *h, double *out);
void orc_hist_f32(const uint8_t *plane, uint64_t row_stride, uint32_t n_features,
                  uint32_t bin_stride, const uint32_t *rows, uint64_t n_rows,
                  const float *g, const float *h, int32_t mode, uint32_t chunk,
                  float *out);
void orc_hist_int(const uint8_t *plane, uint64_t row_stride, uint32_t n_features,
                  uint32_t bin_stride, const uint32_t *rows, uint64_t n_rows,
                  const int8_t *gq, const int16_t *hq, int64_t *out);

/* ---- split finding: src/split.cpp ---- */
void orc_find_node_f32(const float *hist, const uint16_t *n_bins, uint32_t n_features,
                       uint32_t bin_stride, float tot_g, float tot_h, uint64_t n_rows,
                       const orc_tree_cfg *cfg, orc_split *out);
void orc_find_node_int(const int64_t *hist, const uint16_t *n_bins,
                       uint32_t n_features, uint32_t bin_stride, int64_t tot_g,
                       int64_t tot_h, double scale_g, double scale_h, uint64_t n_rows,
                       const orc_tree_cfg *cfg, orc_split *out);
/* Level (oblivious) finder over n_nodes node histograms laid out [node][F][SB][2]. */
void orc_find_level_f32(const float *hists, const float *totals, uint32_t n_nodes,
                        const uint16_t *n_bins, uint32_t n_features,
                        uint32_t bin_stride, const orc_tree_cfg *cfg, orc_split *out);
void orc_find_level_int(const int64_t *hists, const int64_t *totals, uint32_t n_nodes,
                        const uint16_t *n_bins, uint32_t n_features,
                        uint32_t bin_stride, double scale_g, double scale_h,
                        const orc_tree_cfg *cfg, orc_split *out);
/* Node totals = sum of the first feature's cells (include/bonsai/histogram.hpp:76-87). */
void orc_totals_f32(const float *hist, uint32_t n_bins0, float *tot_g, float *tot_h);
float orc_leaf_value(double g, double h, const orc_tree_cfg *cfg);

/* ---- stable partition: src/step/primitives.hpp:113-151 ---- */
uint64_t orc_partition(const uint8_t *plane, uint64_t row_stride, const uint32_t *rows,
                       uint64_t n, uint32_t feature, uint32_t split_bin,
                       uint32_t default_left, uint32_t last_bin, uint32_t *out);

/* ---- growers: src/grower_impl.hpp ---- */
/* Returns #nodes written (<= cap) or -1 on error.  hist_mode picks the cell type; for
 * ORC_HIST_INT gq/hq and the scales are read, otherwise g/h (ORC_HIST_F32_FIXED: nonzero
 * scales are the fixed-point grids s_g, s_h, else orc_fx_shift of these rows' g and h).  chunk: rows per
 * ORC_HIST_F32_CHUNKED partial (0: 8192), or ORC_CHUNK_ADAPTIVE | T: each built node of
 * n rows uses ceil(ceil(n / T) / 256) * 256 rows per partial, clamped to [1024, 8192]
 * (the device's leafwise work units with unit target T). */
#define ORC_CHUNK_ADAPTIVE 0x80000000u
int32_t orc_grow_depthwise(const uint8_t *plane, uint64_t row_stride, uint64_t n_rows,
                           uint32_t n_features, const uint16_t *n_bins,
                           const float *g, const float *h, const int8_t *gq,
                           const int16_t *hq, double scale_g, double scale_h,
                           const orc_tree_cfg *cfg, int32_t hist_mode, uint32_t chunk,
                           orc_node *nodes, int32_t cap, uint32_t *leaf_of_row);
int32_t orc_grow_leafwise(const uint8_t *plane, uint64_t row_stride, uint64_t n_rows,
                          uint32_t n_features, const uint16_t *n_bins, const float *g,
                          const float *h, const int8_t *gq, const int16_t *hq,
                          double scale_g, double scale_h, const orc_tree_cfg *cfg,
                          int32_t hist_mode, uint32_t chunk, orc_node *nodes,
                          int32_t cap, uint32_t *leaf_of_row);
/* Oblivious tree: level_out[depth] = (feature, bin, default_left) triples, leaf table
 * of 2^depth values and covers.  Returns depth or -1. */
int32_t orc_grow_levelwise(const uint8_t *plane, uint64_t row_stride, uint64_t n_rows,
                           uint32_t n_features, const uint16_t *n_bins,
                           const float *g, const float *h, const int8_t *gq,
                           const int16_t *hq, double scale_g, double scale_h,
                           const orc_tree_cfg *cfg, int32_t hist_mode, uint32_t chunk,
                           uint32_t *level_out, float *leaf_values,
                           float *leaf_covers, float *level_gains,
                           uint32_t *leaf_of_row);

/* Whole fit on the CPU (the CPU baseline): logloss or mse objective, depthwise
 * grower, OpenMP fill.  Trees written as orc_node arrays of stride max_nodes.
 * Returns #trees or -1. */
int32_t orc_train_depthwise(const uint8_t *plane, uint64_t row_stride, uint64_t n_rows,
                            uint32_t n_features, const uint16_t *n_bins,
                            const float *label, int32_t objective, uint32_t n_iters,
                            float learning_rate, const orc_tree_cfg *cfg,
                            int32_t hist_mode, int32_t n_threads, orc_node *trees,
                            int32_t max_nodes, int32_t *tree_sizes,
                            float *init_score_out, float *scores_out);

/* ---- predict in bin space: include/bonsai/booster.hpp:831-846 ---- */
void orc_predict(const orc_node *trees, const int32_t *tree_offsets, uint32_t n_trees,
                 const uint8_t *plane, uint64_t row_stride, uint64_t n_rows,
                 const uint16_t *last_bin, float learning_rate, float init_score,
                 float *out);
void orc_predict_oblivious(const uint32_t *levels, const int32_t *depths,
                           const int32_t *level_offsets, const float *leaves,
                           const int32_t *leaf_offsets, uint32_t n_trees,
                           const uint8_t *plane, uint64_t row_stride, uint64_t n_rows,
                           const uint16_t *last_bin, float learning_rate,
                           float init_score, float *out);

/* ---- TreeSHAP: src/shap.cpp (Algorithm 2) and src/shap_paths.cpp (packed paths) ---- */
double orc_tree_expected_value(const orc_node *nodes);
/* Algorithm 2 in bin space; phi has n_features + 1 doubles (bias last), accumulated. */
void orc_tree_shap(const orc_node *nodes, int32_t max_depth, const uint8_t *row_bins,
                   const uint16_t *last_bin, uint32_t n_features, double *phi);
/* Conditional expectation E[f | x_S] used by brute-force Shapley (src/shap.cpp:246-276). */
double orc_cond_expectation(const orc_node *nodes, const uint8_t *row_bins,
                            const uint16_t *last_bin, const uint8_t *in_subset);
/* Packed leaf paths for one ensemble.  Element = (feature|missing_ok<<15, lo, hi, z). */
int32_t orc_pack_paths(const orc_node *trees, const int32_t *tree_offsets,
                       uint32_t n_trees, const uint16_t *last_bin, uint8_t *elems,
                       int32_t elem_cap, uint8_t *heads, int32_t head_cap,
                       int32_t *n_heads_out, int32_t *max_len_out);
void orc_shap_path_weights(uint32_t max_len, double *out);
void orc_eval_paths(const uint8_t *elems, const uint8_t *heads, int32_t n_heads,
                    int32_t max_len, const uint16_t *last_bin, const uint8_t *row_bins,
                    double *phi);
/* Full pred_contribs composition (include/bonsai/booster.hpp:902-941). */
void orc_pred_contribs(const orc_node *trees, const int32_t *tree_offsets,
                       uint32_t n_trees, int32_t max_depth, const uint8_t *plane,
                       uint64_t row_stride, uint64_t n_rows, const uint16_t *last_bin,
                       uint32_t n_features, float learning_rate, float init_score,
                       double *out);

/* ---- AUC: src/metric.cpp:95-137 ---- */
float orc_auc(const float *probs, const float *labels, uint64_t n);

#ifdef __cplusplus
}
#endif
#endif
