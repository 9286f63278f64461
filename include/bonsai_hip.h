/*
 * bonsai_hip.h — C ABI of the MI355X-native GBDT hot path (libbonsai_hip.so).
 *
 * The reference has no C ABI: its CUDA engine is a same-compiler C++ class behind the
 * engine concepts (include/bonsai/grower.hpp:78-162) with a pimpl
 * (include/bonsai/cuda/histogram_engine.hpp:281-283).  This header is the boundary a
 * C++ adapter satisfying those concepts (or a ctypes / cgo / JNI binding) calls.  Every
 * declaration cites the reference interface it replaces.
 *
 * Conventions: plain pointers and sizes, host memory unless stated; every function
 * returns 0 on success, > 0 for a configuration/capacity refusal (the reference's
 * ConfigError, src/cuda/detail/device_context.cu:60-87) and < 0 for a device error
 * (the reference's std::runtime_error("cuda: ..."), src/cuda/detail/device_buffer.cuh:96-103).
 * bh_last_error() returns the thread-local message of the last failure.
 * A context is not re-entrant; distinct contexts may be used from distinct threads.
 */
#ifndef BONSAI_HIP_H
#define BONSAI_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct bh_context bh_context;
typedef struct bh_predictor bh_predictor;

/* Growth policies: the reference's cuda_depthwise / cuda_leafwise / cuda_levelwise
 * registry names (include/bonsai/registry/names.hpp:81-94), here hip_*. */
enum { BH_DEPTHWISE = 0, BH_LEAFWISE = 1, BH_LEVELWISE = 2 };
/* Objectives with a device gradient (DeviceObjectiveKind,
 * include/bonsai/objective_traits.hpp:21-27; gh_from_scores_kernel, kernels.cuh:972-1009):
 * squared error, logistic, Poisson with a log link (raw score clamped to +-30 before exp). */
enum { BH_MSE = 0, BH_LOGLOSS = 1, BH_POISSON = 2 };

/* TreeConfig + BoosterConfig subset plus the device knobs
 * (include/bonsai/config/tree_config.hpp:9-29, booster_config.hpp:8-24). */
typedef struct bh_fit_config {
    int32_t  grower;            /* BH_DEPTHWISE / BH_LEAFWISE / BH_LEVELWISE */
    int32_t  objective;         /* BH_MSE / BH_LOGLOSS */
    int32_t  quantized;         /* 0: FP32 fixed-order histograms; 1: int8 g / int16 h */
    int32_t  use_graph;         /* 1: one hipGraph per boosting iteration */
    int32_t  host_sync;         /* 1: host round-trip after every level's find and
                                   partition (the reference's cadence, for comparison) */
    int32_t  max_depth;
    int32_t  min_data_in_leaf;
    int32_t  max_leaves;        /* leafwise only */
    float    min_child_hess;
    float    min_gain_to_split;
    float    lambda_l2;
    float    lambda_l1;
    float    learning_rate;
    uint32_t _pad;
    uint64_t seed;              /* stochastic-rounding key */
} bh_fit_config;

/* Tree node in bin space (DenseTree::Node, include/bonsai/tree.hpp:32-48, with the
 * split bin instead of the float threshold, plus cover and split gain). */
typedef struct bh_node {
    int32_t  feature;           /* < 0: leaf */
    uint32_t bin;               /* rows with bin <= bin go left (missing bin: default_left) */
    uint32_t default_left;
    uint32_t left;
    uint32_t right;
    float    value;             /* leaf value */
    float    cover;
    float    gain;
} bh_node;

/* Split decision (SplitOutput, include/bonsai/split.hpp:66-73) plus the child sums the
 * finder derived (the reference's FeatBest, src/cuda/detail/device_buffer.cuh:84-88). */
typedef struct bh_split {
    double   gain;
    uint32_t feature, bin, default_left, valid;
    double   gL, hL, gR, hR;
    int64_t  igL, ihL, igR, ihR;
    uint32_t n_left, n_right;   /* child row counts (quantized fits) */
} bh_split;

/* ---- devices and errors ---- */
/* cuda_device_count / cuda_available (include/bonsai/cuda/histogram_engine.hpp:36-41). */
int         bh_device_count(int *count);
/* Block until every operation queued on the context's stream has finished. */
int         bh_synchronize(bh_context *ctx);
/* Calibration (no counterpart in the reference): copy `bytes` (a multiple of 16) between
 * two fresh device buffers `reps` times with a plain vector copy kernel and return the
 * mean kernel time (ms, hipEvents).  Its traffic is exactly 2 x bytes, which calibrates
 * HBM counters and gives the achievable copy bandwidth beside the histogram's. */
int         bh_copy_calibrate(int device, uint64_t bytes, uint32_t reps, float *ms_mean);
/* Calibration (no counterpart in the reference): the same timing for mode 0 (the plain
 * copy above), 1 (copy with four independent 16-byte nontemporal loads in flight per
 * lane; traffic 2 x bytes) or 2 (read only, eight loads in flight per lane; traffic
 * bytes), over `grid` workgroups of 256 lanes (modes 1 and 2; 1 .. 65536).  The best
 * of these is the bandwidth a streaming kernel reaches on the device. */
int         bh_stream_calibrate(int device, uint64_t bytes, uint32_t reps, uint32_t mode, uint32_t grid,
                                float *ms_mean);
/* Message of the last failure on this thread. */
const char *bh_last_error(void);

/* ---- context (CudaDeviceContext, src/cuda/detail/device_context.cuh:40-493) ---- */
/* cuda_select_device + engine construction (src/cuda/histogram_engine.cu:54-83). */
int bh_create(int device, bh_context **out);
int bh_destroy(bh_context *ctx);

/* Adopt a binned plane: row-major u8 bins [n_rows][row_stride], n_bins[f] per feature
 * (<= 256).  Replaces the reference's tiled upload stage_tiled / cuda_ingest
 * (src/cuda/detail/device_context.cu:499-571, src/cuda/histogram_engine.cu:381). */
int bh_upload_bins(bh_context *ctx, const uint8_t *bins, uint64_t n_rows, uint32_t n_features,
                   uint64_t row_stride, const uint16_t *n_bins);
/* Device bin transform (the reference's transform_bin, src/cuda/detail/ingest_kernels.cuh:27-50,
 * twin of BinMapper::transform, src/bin_mapper.cpp:284-291): X is row-major float
 * [n_rows][n_features]; feature f's ascending cuts are cuts[cut_offsets[f] .. cut_offsets[f+1]),
 * n_bins[f] = cut count (<= 256).  bin = lower_bound(cuts, x), NaN -> last bin.  Replaces a
 * previous plane like bh_upload_bins. */
int bh_upload_features(bh_context *ctx, const float *X, uint64_t n_rows, uint32_t n_features,
                       const float *cuts, const uint32_t *cut_offsets);
/* Column ingest (cuda_ingest(ColumnBatch), src/cuda/histogram_engine.cu:420-470, with
 * bin_col_kernel, src/cuda/detail/ingest_kernels.cuh:89-102): columns[f] points at feature
 * f's n_rows floats in host memory; cuts as for bh_upload_features.  Columns are staged four
 * at a time and each row's four bins are written as one dword.  Same plane as
 * bh_upload_features on the row-major matrix. */
int bh_upload_feature_columns(bh_context *ctx, const float *const *columns, uint64_t n_rows,
                              uint32_t n_features, const float *cuts, const uint32_t *cut_offsets);
/* Device-resident ingest (cuda_ingest_device, src/cuda/histogram_engine.cu:495-526): X_dev is
 * a row-major float [n_rows][n_features] matrix in device memory of the context's device
 * (other memory is refused), written completely before the call; no host copy is made. */
int bh_upload_features_device(bh_context *ctx, const float *X_dev, uint64_t n_rows, uint32_t n_features,
                              const float *cuts, const uint32_t *cut_offsets);
/* Sample of a device-resident matrix for the host cut fitting (cuda_gather_rows,
 * src/cuda/histogram_engine.cu:473-493, gather_rows_kernel, ingest_kernels.cuh:76-86):
 * out[i][:] = X_dev[rows[i]][:] for i < n_sample (every rows[i] < n_rows, else an error
 * before any copy); n_sample == 0 copies the whole matrix (out holds n_rows * n_features
 * floats).  out is host memory. */
int bh_gather_rows(int device, const float *X_dev, uint64_t n_rows, uint32_t n_features, const uint32_t *rows,
                   uint64_t n_sample, float *out);
/* Copy the resident pristine bin plane back (row-major, out_stride bytes per row). */
int bh_download_bins(bh_context *ctx, uint8_t *out, uint64_t out_stride);
/* Labels of the resident objective (resident_begin, device_context.cu:1541-1579). */
int bh_upload_labels(bh_context *ctx, const float *labels);
/* Per-row sample weights of the resident objective (resident_begin uploads them for a
 * weighted Dataset, include/bonsai/cuda/histogram_engine.hpp:249-250): the device gradient
 * scales g and h by w[r] (gh_from_scores_kernel's Weighted arm, kernels.cuh:972-1009; the
 * host multiply of include/bonsai/booster.hpp:546-553).  n_rows finite non-negative
 * weights; NULL clears (unweighted).  Host gradients (bh_set_gradients) are taken as
 * given. */
int bh_upload_weights(bh_context *ctx, const float *weights);
/* Configure the fit; allocates every device buffer the fit needs (the reference's
 * capacity gates refuse_hist_budget / refuse_leaf_pool, device_context.cu:60-87). */
int bh_configure(bh_context *ctx, const bh_fit_config *cfg);

/* ---- whole-fit path (Booster::update_one_iter + resident_round,
 *      include/bonsai/booster.hpp:500-676) ---- */
/* Set every row's score (init_score on the first call of a fit). */
int bh_set_scores(bh_context *ctx, const float *scores);
int bh_get_scores(bh_context *ctx, float *scores);
/* Run n_iters boosting iterations on the device (gradient, quantize, grow, score
 * update).  With use_graph each iteration is one hipGraph launch. */
int bh_train(bh_context *ctx, uint32_t n_iters);
/* Validation plane (eval_begin / eval_accumulate, src/cuda/detail/device_context.cu:1666-1760,
 * driven by the fit loop's early-stopping rule, src/cli/pipeline.cpp:443-476): binned rows
 * (row-major u8, same features and bins as the training plane), labels, the seed score.
 * Every bh_train round then walks the new tree on these rows and records the mean loss of
 * the configured objective (MSE or logloss).  early_stopping_rounds > 0: once that many
 * rounds follow the best loss without beating it, training stops and keeps the trees up
 * to the best round (later bh_train calls return at once).  bh_set_scores disarms it. */
int bh_set_validation(bh_context *ctx, const uint8_t *bins, uint64_t n_rows, uint64_t row_stride,
                      const float *labels, float init_score, uint32_t early_stopping_rounds);
/* Losses recorded so far (up to cap), their count, the best round and the stop flag. */
int bh_get_eval(bh_context *ctx, float *history, uint32_t cap, uint32_t *n, uint32_t *best_iter,
                int32_t *stopped);
/* Row subset of every following tree (the row_indices the booster's sampler hands to
 * grow, include/bonsai/booster.hpp:666-676; the root over them, SplitInput make_root,
 * src/step/level.hpp:171-187): rows ascending, unique, below n_rows.  The tree is grown
 * on those rows only; every row's score (and bh_grow_tree's leaf_of_row) still comes
 * from the finished tree (route_unsampled, src/grower_impl.hpp:430-482).  rows NULL,
 * n 0 or all n_rows rows: no subset.  Replaces a Bernoulli sampler. */
int bh_set_row_subset(bh_context *ctx, const uint32_t *rows, uint64_t n);
/* Bernoulli row sampler on the device, drawn afresh every iteration (BernoulliSampler,
 * src/sampler.cpp:27-55; allowed on the resident path, include/bonsai/booster.hpp:632-635):
 * row r of iteration t is kept when 32 counter-based random bits (seed, t, r) fall below
 * floor(subsample * 2^32).  subsample must be > 0 (the reference's ConfigError);
 * subsample >= 1 keeps every row.  The draw is this library's own stream, not the
 * reference's std::mt19937 one (bh_set_row_subset takes that one's indices).  Replaces
 * a row subset. */
int bh_set_bernoulli(bh_context *ctx, float subsample, uint64_t seed);
/* Per-tree feature selection handed over (the `selected` list of sample_features,
 * src/grower_impl.hpp:165-180 and :527-528): mask[f] != 0 selects feature f for every
 * following tree; the split finders skip the others (src/split.cpp:106-110).  n_features
 * must equal the uploaded plane's; a mask that selects nothing is refused (the
 * reference's refuse_empty_selection, src/cuda/detail/device_context.cu:81-86).  mask
 * NULL or all selected: every feature.  Replaces a feature fraction. */
int bh_set_feature_mask(bh_context *ctx, const uint8_t *mask, uint32_t n_features);
/* Per-tree feature draw on the device (tree.feature_fraction / feature_seed,
 * include/bonsai/config/tree_config.hpp:15-19): tree t selects
 * k = max(1, ceil(fraction * n_features)) features, those with the k smallest
 * counter-based 64-bit keys (seed, t, f); fraction >= 1 (or one feature) selects all.
 * NaN is refused.  The draw is this library's own stream, not std::sample over the
 * reference's std::mt19937 (bh_set_feature_mask takes that one's selection).  Replaces
 * a feature mask. */
int bh_set_feature_fraction(bh_context *ctx, float fraction, uint64_t seed);
/* Monotone constraints (tree.monotone_constraints, include/bonsai/config/tree_config.hpp:
 * 20-22): dirs[f] > 0 increasing, < 0 decreasing, 0 free; features past n are free
 * (monotone_constraint_of, include/bonsai/split.hpp:160-166).  The finders skip a
 * candidate whose bounded child weights break its feature's direction
 * (src/split.cpp:143-153), children inherit the parent's leaf-weight box fenced at the
 * midpoint of the child weights (propagate_monotone_bounds, src/grower_impl.hpp:47-75),
 * and leaf values are clamped to their box (bounded_leaf_weight,
 * include/bonsai/split.hpp:141-150).  Kept across bh_configure; refused with the
 * levelwise grower (src/grower_impl.hpp:583-588): while directions are set, bh_configure
 * refuses that grower and keeps the previous fit.  dirs NULL / n 0 clears. */
int bh_set_monotone(bh_context *ctx, const int8_t *dirs, uint32_t n);
/* Interaction constraints (tree.interaction_constraints, parsed as
 * parse_interaction_groups does, src/grower_impl.hpp:84-107): n_groups groups, group g
 * holding group_sizes[g] feature ids, concatenated in members (empty groups dropped; at
 * most 256 groups, 4,096 features).  A node may split on the union of the groups that
 * contain every distinct feature on its path, plus that feature when the path has one;
 * the root on any (allowed_features, src/grower_impl.hpp:111-137; the finder's skip,
 * src/split.cpp:111-114).  Kept across bh_configure; refused with the levelwise grower
 * (src/grower_impl.hpp:589-595; while groups are set, bh_configure refuses that grower
 * and keeps the previous fit) and for ids beyond n_features.  n_groups 0 clears. */
int bh_set_interaction(bh_context *ctx, const uint32_t *members, const uint32_t *group_sizes, uint32_t n_groups);
/* Trees grown so far: nodes[t * max_nodes + i], sizes[t]; returns via *n_trees. */
int bh_get_trees(bh_context *ctx, bh_node *nodes, uint32_t max_nodes, uint32_t *sizes,
                 uint32_t cap_trees, uint32_t *n_trees);
/* Per-tree node capacity of the configured fit. */
int bh_max_nodes(bh_context *ctx, uint32_t *max_nodes);

/* ---- engine seam, one tree from host gradients (HistogramEngine::begin_tree +
 *      the grow loop, include/bonsai/grower.hpp:54-162) ---- */
/* Host gradients (begin_tree with host grad/hess, device_context.cu:573-600).  With a
 * quantized fit the device quantizes them (seed, iter) and returns the quantized
 * values and scales when gq/hq/scales are non-null. */
int bh_set_gradients(bh_context *ctx, const float *grad, const float *hess, uint32_t iter,
                     int8_t *gq, int16_t *hq, double *scales);
/* Read back the gradients of the current iteration in row order: after bh_train, those
 * the last iteration computed on the device from the scores before its tree (the
 * resident gh_from_scores_kernel, src/cuda/detail/kernels.cuh:972-1009, plus this
 * build's per-iteration quantizer); after bh_set_gradients, those handed over.  FP32
 * fits fill grad / hess; quantized fits fill gq / hq and scales (g = gq * scales[0],
 * h = hq * scales[1]).  Any pointer may be NULL. */
int bh_get_gradients(bh_context *ctx, float *grad, float *hess, int8_t *gq, int16_t *hq, double *scales);
/* Grow one tree from the current gradients without touching scores.  leaf_of_row
 * (optional, n_rows entries) receives each row's leaf node id (stamp_kernel,
 * kernels.cuh:479-488); for the levelwise grower the leaf-table index. */
int bh_grow_tree(bh_context *ctx, bh_node *nodes, uint32_t cap, uint32_t *n_nodes,
                 uint32_t *leaf_of_row);

/* ---- kernel-level entry points (parity tests and profiling) ---- */
/* Histogram of the rows listed (ascending) under the current gradients
 * (launch_hist, device_context.cu:451-497).  FP32 fits write float pairs
 * [n_features][bin_stride][2] into out_f32; quantized fits write int64 pairs into
 * out_i64.  bin_stride = max n_bins. */
int bh_histogram(bh_context *ctx, const uint32_t *rows, uint64_t n, float *out_f32,
                 int64_t *out_i64);
/* No counterpart in the reference: bh_histogram for callers that cannot allocate (code
 * under HIP graph capture, tight serving or engine loops, harnesses that forbid
 * hipMalloc).  The row list (n row ids below n_rows, device memory) and the output (the
 * histogram in bh_histogram's layout, device memory: int64 pairs for quantized fits,
 * float pairs for FP32 ones) are the caller's; the work list, control block and unit
 * partials live in a fixed 32 MiB per-GPU scratch array that the library holds until
 * the process ends, so the call allocates nothing and returns once its kernels are
 * queued on hip_stream (NULL: the context's stream).  The histogram equals bh_histogram's
 * for the same rows bit for bit.  Row counts whose unit partials (8,192-row units of
 * features x bins cells) exceed the scratch are refused with a capacity error naming the
 * limit.  Every call on one GPU shares that GPU's scratch: calls that overlap on one GPU
 * (other streams, other contexts) race, and like bh_histogram the call reuses the
 * context's gather buffers. */
int bh_histogram_noalloc(bh_context *ctx, const uint32_t *d_rows, uint64_t n, void *d_out, void *hip_stream);
/* Node split find over a host histogram in the layout above (find_splits_many,
 * device_context.cu:997-1048).  Totals are computed from the histogram unless
 * tot_given != 0. */
int bh_find_split(bh_context *ctx, const float *hist_f32, const int64_t *hist_i64,
                  uint64_t n_rows, uint32_t depth, int32_t tot_given, double tot_g,
                  double tot_h, const double *scales, bh_split *out);
/* Stable partition of the listed rows by one split (partition_level,
 * device_context.cu:790-881); writes the reordered row ids, returns n_left. */
int bh_partition(bh_context *ctx, const uint32_t *rows, uint64_t n, uint32_t feature,
                 uint32_t bin, uint32_t default_left, uint32_t *out_rows, uint64_t *n_left);
/* Time the histogram kernel on the resident plane: builds the root histogram `reps`
 * times and returns the mean kernel time (ms) measured with hipEvents on the kernel's
 * own stream. */
int bh_time_root_histogram(bh_context *ctx, uint32_t reps, float *ms_mean);
/* Run n_iters depthwise iterations without a graph, timing every histogram launch with
 * hipEvents on the engine stream.  Totals of kernel time (ms), launches, and bytes two
 * ways: algorithmic (every built node's rows read once, rows x (features + gradient-pair
 * bytes), plus 8 bytes per built (node, feature, bin) cell) with the built rows, and
 * scanned (the rows the launches actually read: a level whose rows were not moved scans
 * each parent's segment) with the scanned rows.  Trees are appended like bh_train. */
int bh_profile_hist(bh_context *ctx, uint32_t n_iters, double *ms_total, double *bytes_total,
                    uint64_t *launches, uint64_t *rows_total, double *scanned_bytes_total,
                    uint64_t *scanned_rows_total);
/* Per-iteration timing of bh_train (the reference's lap profile, src/cuda/detail/profile.cuh:39-40):
 * when enabled, hipEvents bracket every iteration; bh_get_timing writes ms[i] = device
 * time of iteration i of the last bh_train call (0 beyond its count). */
int bh_set_timing(bh_context *ctx, int32_t enable);
int bh_get_timing(bh_context *ctx, float *ms, uint32_t n);

/* ---- engine steps: GPULevelEngine / GPULeafEngine (include/bonsai/grower.hpp:91-162;
 *      CudaHistogramEngine, include/bonsai/cuda/histogram_engine.hpp:90-218) ----
 * For growers that keep the control loop on the host (LevelStep / LeafStep,
 * src/step/level.hpp:159-366, src/step/leaf.hpp:124-267), one host round trip per step.
 * They grow from the current gradients (bh_set_gradients, or the last bh_train
 * iteration's) over every row or the row subset (bh_set_row_subset); a feature mask
 * (bh_set_feature_mask) narrows the finders.  Rows never cross the bus: partitions move
 * each parent's segment into the other work plane (left rows ascending, then right rows,
 * at the parent's offsets) and histograms stay in slot-indexed device buffers.  The
 * device feature draw, the Bernoulli draw and monotone / interaction constraints are
 * refused (the host grower owns them in this mode). */
typedef struct bh_level_op {      /* LevelOp (histogram_engine.hpp:103-108) */
    uint32_t parent_slot, small_slot, large_slot;
} bh_level_op;
typedef struct bh_partition_op {  /* PartitionOp (histogram_engine.hpp:112-120) */
    uint32_t parent_slot, left_slot, right_slot, feature, bin, default_left;
} bh_partition_op;
typedef struct bh_leaf_stamp {    /* LeafStamp (histogram_engine.hpp:122-126) */
    uint32_t slot, node_id;
} bh_leaf_stamp;
typedef struct bh_leaf_round {    /* LeafRound (histogram_engine.hpp:212-218) */
    uint32_t left_slot, right_slot, left_count, right_count;
} bh_leaf_round;
/* A node handed to a finder (the SplitInput fields the device path reads, split.hpp:
 * 15-53): depth in; the node's rows and totals out (Histogram::totals of its cells, the
 * first selected feature's: FP32 double sums rounded to float; quantized exact integers,
 * isum_*, times the scales, sum_*). */
typedef struct bh_step_node {
    uint32_t depth, _pad;
    uint64_t rows;
    double   sum_g, sum_h;
    int64_t  isum_g, isum_h;
} bh_step_node;

/* begin_root (histogram_engine.hpp:144-151): the root's histogram into slot 0, its
 * totals (as bh_step_node's) and rows out.  Level plane: depthwise or levelwise fits. */
int bh_begin_root(bh_context *ctx, double *sums, uint64_t *rows);
/* partition_level (:153-158): child_counts[2i], [2i+1] = (n_left, n_right) of ops[i]. */
int bh_partition_level(bh_context *ctx, const bh_partition_op *ops, uint32_t n_ops, uint32_t *child_counts);
/* advance_level (:159-160): small children built from their rows, large = parent - small;
 * the child level becomes current. */
int bh_advance_level(bh_context *ctx, const bh_level_op *ops, uint32_t n_ops);
/* advance_level with the reference's launch cadence (one histogram build and one
 * subtraction launch per node, src/cuda/detail/device_context.cu:913-988), for measuring
 * against the one-launch-per-level build; same results. */
int bh_advance_level_per_node(bh_context *ctx, const bh_level_op *ops, uint32_t n_ops);
/* advance_layout_only (:161-163): the child level becomes current without histograms. */
int bh_advance_layout_only(bh_context *ctx);
/* find_splits_many (:169-176): node i is slot i of the current level. */
int bh_find_splits_many(bh_context *ctx, bh_step_node *nodes, uint32_t n, bh_split *out);
/* find_level_split (:177-183), levelwise fits: one split for the whole frontier; out[i]
 * carries it with node i's child sums under it (valid = 0: no feasible split). */
int bh_find_level_split(bh_context *ctx, bh_step_node *nodes, uint32_t n, bh_split *out);
/* stamp_leaves / leaf_stamp (:150-152, :239-242): the rows of each named slot (current
 * level, or the leaf pool) end in node node_id. */
int bh_stamp_leaves(bh_context *ctx, const bh_leaf_stamp *stamps, uint32_t n);
/* finalize_tree (:164-168): every row's stamped node id (0xFFFFFFFF: an unstamped leaf)
 * and node_values[id] (0 beyond n_values); either output may be NULL.  Rows outside the
 * row subset are routed through the recorded splits too. */
int bh_finalize_tree(bh_context *ctx, const float *node_values, uint32_t n_values, float *values,
                     uint32_t *leaf_ids);
/* leaf_begin_root (:220-226), leafwise fits: the root into pool slot 0. */
int bh_leaf_begin_root(bh_context *ctx, double *sums, uint64_t *rows);
/* leaf_split (:227-228): both counts nonzero: the smaller child (ties: left) takes the
 * next free slot and the larger inherits the parent's; otherwise no slot is taken. */
int bh_leaf_split(bh_context *ctx, uint32_t parent_slot, uint32_t feature, uint32_t bin, uint32_t default_left,
                  bh_leaf_round *out);
/* leaf_build (:229-233): small built from its rows, large = parent - small in place. */
int bh_leaf_build(bh_context *ctx, uint32_t small_slot, uint32_t large_slot);
/* leaf_find (:234-238): nodes[i]'s histogram is pool slot slots[i]. */
int bh_leaf_find(bh_context *ctx, bh_step_node *nodes, const uint32_t *slots, uint32_t n, bh_split *out);

/* ---- launch-shape tuning (no counterpart in the reference, whose launch shapes are
 *      fixed constants, src/cuda/detail/device_context.cu:451-497) ---- */
typedef struct bh_tuning {
    uint32_t scatter_lds_bytes; /* LDS per partition-scatter workgroup (4,096..162,816 bytes): sets its tile rows */
    uint32_t find_blocks;       /* workgroups sharing one node's split find (1..8) */
    uint32_t hist_threads;      /* quantized histogram workgroup size: 256, 512 or 1024 */
    uint32_t part_grid;         /* partition workgroups */
    uint32_t hist_layout;       /* quantized histogram: 0 row-dword lanes (default), 1 feature-in-lane (rows <= 32
                                   bytes), 2 row-in-lane (rows <= 64 bytes), 3 rotated row-in-lane with
                                   feature-minor LDS cells (rows <= 64 bytes, single-sub-unit levels; every
                                   level, routed ones included), 4 row-dword lanes on every level, routed
                                   ones included, 5 rotated row-dword lanes with feature-minor LDS cells
                                   (as 3 otherwise); with 0-2, levels whose units filter their parent's
                                   rows use 2 whenever rows are <= 64 bytes */
    uint32_t pair_levels;       /* quantized depthwise, rows <= 64 bytes: 1 = move rows every second level
                                   (four-way partitions; the levels between scan parent segments) */
    uint32_t hist_group;        /* quantized depthwise: 8192-row sub-units per histogram work unit on levels
                                   that scan every row (1..4; fewer, larger unit partials) */
} bh_tuning;
/* Set / read the context's launch shapes (the iteration graph is re-captured). */
int bh_set_tuning(bh_context *ctx, const bh_tuning *t);
int bh_get_tuning(bh_context *ctx, bh_tuning *t);
/* Autotuner: coordinate descent over the knobs above on the resident data and the
 * configured fit, `iters` timed boosting iterations per trial (hipEvents).  Scores,
 * trees and the iteration counter are restored afterwards; the best shapes stay set. */
int bh_autotune(bh_context *ctx, uint32_t iters, bh_tuning *chosen, float *ms_per_iter);

/* ---- predict and TreeSHAP (cuda_predict_plan / cuda_predict,
 *      src/cuda/predict.cu:147-266; cuda_shap_plan / cuda_pred_contribs,
 *      src/cuda/shap.cu:212-380) ---- */
/* An immutable ensemble resident on one device.  Trees in bin space; last_bin[f] =
 * n_bins[f] - 1.  Safe to share between serving streams. */
int bh_predictor_create(int device, const bh_node *nodes, const uint32_t *tree_offsets,
                        uint32_t n_trees, const uint16_t *last_bin, uint32_t n_features,
                        float learning_rate, float init_score, bh_predictor **out);
int bh_predictor_destroy(bh_predictor *p);
/* init + lr * sum_t leaf_t(row), bit-equal to the host bin-space walk
 * (include/bonsai/booster.hpp:831-846).  bins host, row-major [n][row_stride]. */
int bh_predict(bh_predictor *p, const uint8_t *bins, uint64_t n, uint64_t row_stride,
               float *out);
/* TreeSHAP contributions: out[n][n_features + 1] (bias last), as pred_contribs
 * (include/bonsai/booster.hpp:902-941). */
int bh_shap(bh_predictor *p, const uint8_t *bins, uint64_t n, uint64_t row_stride,
            double *out);
/* Same two on device-resident inputs/outputs (already in HBM) on a caller stream.
 * bh_shap_device may use d_out as its 64-bit fixed-point accumulator before writing
 * the doubles (batches under 65,536 rows), so d_out must not overlap d_bins.  It
 * allocates nothing and issues kernels only (no memset), so a caller may capture it
 * in a graph. */
int bh_predict_device(bh_predictor *p, const uint8_t *d_bins, uint64_t n, uint64_t row_stride,
                      float *d_out, void *hip_stream);
int bh_shap_device(bh_predictor *p, const uint8_t *d_bins, uint64_t n, uint64_t row_stride,
                   double *d_out, void *hip_stream);

/* ---- serving: continuous batching over a resident ensemble (no counterpart in the
 *      reference, whose Python binding serves one call at a time through a plan cache,
 *      src/python/module.cpp:603-640) ---- */
typedef struct bh_serve_config {
    uint32_t n_slots;        /* batches in flight at once (one HIP stream each) */
    uint32_t do_shap;        /* 1: TreeSHAP contributions as well as predictions */
    uint64_t max_batch_rows; /* row capacity of one batch */
    uint64_t hbm_budget;     /* bytes of request data (rows + outputs) in flight */
} bh_serve_config;

typedef struct bh_serve_stats {
    double   wall_s;
    double   rows_per_s;
    double   p50_ms, p99_ms, mean_ms;
    uint64_t rows;
    uint32_t requests;
    uint32_t max_inflight;           /* requests admitted and not yet completed, at most */
    uint32_t max_device_concurrency; /* batches whose kernels ran on the device at once, at most */
    uint32_t _pad;
    double   device_overlap_ms;      /* device time with two or more batches' kernels running */
} bh_serve_stats;

/* Serve n_req requests whose rows are rows[req_offsets[i] .. req_offsets[i+1]) and
 * which arrive at arrival_s[i] seconds after the call starts (ascending).  Requests
 * that have arrived are packed into a batch whenever a slot frees, within the row
 * capacity and the HBM budget.  Outputs land at each request's own rows;
 * latency_ms[i] = completion - arrival. */
int bh_serve(bh_predictor *p, const uint8_t *rows, uint64_t row_stride,
             const uint64_t *req_offsets, const double *arrival_s, uint32_t n_req,
             const bh_serve_config *cfg, float *pred_out, double *shap_out,
             double *latency_ms, bh_serve_stats *stats);

#ifdef __cplusplus
}
#endif
#endif
