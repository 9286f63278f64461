// Shared host/device layouts of the MI355X GBDT engine.
//
// Data layout in HBM (one context = one resident training set):
//   plane0   u8 [N][Fs]   pristine bin plane, original row order (Fs = F rounded up to 4)
//   work0/1  u8 [N][Fs]   ping-pong planes: rows physically partitioned per node, each
//                         node's rows one contiguous segment, ascending original order
//   gh0      gradient pairs in original order (f32x2, or one packed u32 when quantized:
//                         high 16 bits int16 g in [-127,127], low 16 bits uint16 h)
//   ghw0/1   gradient pairs travelling with the partitioned rows
//   hist[2]  node histograms of the current / next level, [slot][F][SB][2] cells
//            (f32 pairs, or i64 pairs when quantized)
//   cnt[2]   quantized fits: row count per (slot, feature, bin), u32 [slot][F][SB]
#pragma once
#include <stdint.h>

namespace bh {

typedef uint8_t  u8;
typedef uint16_t u16;
typedef uint32_t u32;
typedef uint64_t u64;
typedef int64_t  i64;

constexpr u32 kNone       = 0xFFFFFFFFu;
constexpr u32 kSrcPlane0  = 2; // segment lives in the pristine plane
// Rows per histogram work unit.  Fixed: it is part of the FP32 summation order (the
// oracle's chunked restatement uses the same size), and it bounds the packed quantized
// cells, which hold one unit's sums in 64 bits: h in bits [0, 28), the row count in
// [28, 42), signed g in [42, 64) (h <= 32767 * 8192 < 2^28, count <= 2^13,
// |g| <= 127 * 8192 < 2^21).
constexpr u32 kHistChunk  = 8192;
constexpr u32 kPackCntShift = 28;
constexpr u32 kPackGShift = 42;
constexpr u64 kPackHMask = (1ull << 28) - 1ull;
constexpr u64 kPackCntMask = (1ull << 14) - 1ull;
// Quantized depthwise fits may group up to this many kHistChunk sub-units in one
// histogram work unit (the kernels add each sub-unit's packed cells into the unit's wide
// cells); the group size is a tuning knob (1..kHistGroupMax).
constexpr u32 kHistGroupMax = 4;
constexpr u32 kPartChunk  = 4096;  // rows per partition work unit
constexpr u32 kCursorStride = 16;  // u32 between the class cursors of four-way partitions (64 B)
constexpr u32 kCursorMinOps = 8;   // four-way ops of a level from which tiles reserve their runs
// Rows wider than 64 bytes: partition chunks shrink (powers of two, at least 64 rows) to
// about this many bytes, so a level of wide rows still has many chunks.
constexpr u64 kPartChunkBytes = 256u * 1024u;
// Leafwise rounds partition one leaf and build one child's histogram; the work units are
// sized to the leaf so that a small leaf is still spread over many workgroups: partition
// chunks of kLwMinChunk..kPartChunk rows, about kLwChunkTarget of them, and histogram
// units of kLwMinUnit..kHistChunk rows, at most the context's unit target (quantized sums
// are exact in any grouping; FP32 sums follow the units, a rule the oracle restates).
constexpr u32 kLwChunkTarget = 1024;
constexpr u32 kLwMinChunk = 256;
constexpr u32 kLwUnitTarget = 256;
constexpr u32 kLwMinUnit = 1024;
constexpr u32 kMaxTreeNodes = 8192;
// Tail padding of planes (bytes) and gradient buffers (elements): kernels that read
// aligned 16-byte words around a row range may touch up to 15 bytes past its end.
constexpr size_t kTailPadBytes = 64;
// Split find: up to this many workgroups share one node's feature tiles.
constexpr u32 kFindMaxBlocks = 8;
constexpr u32 kTailPad = 16;


// Tree node in bin space, the layout the host reads back (feature < 0: leaf).
struct DevNode {
    int32_t feature;
    u32     bin;
    u32     default_left;
    u32     left;
    u32     right;
    float   value;
    float   cover;
    float   gain;
};
static_assert(sizeof(DevNode) == 32, "DevNode");

// One frontier node of the tree being grown.
struct NodeInfo {
    u32    id;       // node id in the tree
    u32    off;      // first row of its segment
    u32    cnt;      // rows in the segment
    u32    src;      // 0/1 working plane, 2 = pristine plane
    u32    slot;     // histogram slot (frontier position for level growth)
    u32    depth;
    float  tg, th;   // FP32 totals (Histogram::totals semantics)
    i64    itg, ith; // quantized totals
    u32    totals_ready; // 1 when totals were set from a parent split (last level)
    u32    pad;
};
static_assert(sizeof(NodeInfo) == 56, "NodeInfo");

// Split decision of one node (SplitOutput plus the child sums the finder saw).
struct SplitRec {
    double gain;
    u32    feature, bin, default_left, valid;
    double gL, hL, gR, hR;     // FP32 path child sums (double)
    i64    igL, ihL, igR, ihR; // quantized path child sums (exact)
    u32    n_left, n_right;    // child row counts (quantized finder, from the count cells)
};
static_assert(sizeof(SplitRec) == 96, "SplitRec");

// A histogram work unit: rows [begin, end) of one node's segment, or of its parent's
// segment when the node's rows have not been moved yet (route != kNone): then only the
// rows that the parent's split (ops[route]) sends to `side` (0 left, 1 right) count.
struct HistChunk {
    u32 slot;  // destination node histogram slot (next-level buffer)
    u32 src;   // which plane/gh buffer holds the rows
    u32 begin, end;
    u32 part;  // partial index, or kNone when the chunk is the node's only one
    u32 route = kNone; // PartOp index of the parent's split, kNone: every row counts
    u32 side = 0;
    u32 wide = 0;      // quantized partial format: 1 wide (g, h, count planes), 0 packed 64-bit cells
};

// Completion of one built node: small = sum of partials (fixed order), large = parent -
// small.  large_slot == kNone for the root.
struct FinishOp {
    u32 small_slot, large_slot, parent_slot;
    u32 part_first, part_count; // part_count == 0: written directly (or empty)
    u32 empty;                  // small child has no rows: zero its histogram
    u32 unit_first, unit_count; // histogram units of the small child
    u32 small_off, small_cnt, src; // rows the units scan (the parent's segment when routed)
    u32 route = kNone, side = 0; // HistChunk::route / side of the units
    u32 wide = 0;                // HistChunk::wide of the units
    u32 pad[2];
};
static_assert(sizeof(FinishOp) == 64, "FinishOp");

// A split's partition: parent frontier index and the routing rule.
struct PartOp {
    u32 parent;       // frontier index
    u32 feature, bin, default_left, last_bin;
    u32 off, cnt, src, dst;
    u32 chunk_first, chunk_count;
    u32 n_left;       // filled by the scan
    u32 left_id, right_id;
    u32 alive;        // 0 after demotion
    u32 base;         // scan scratch: global prefix at the op's first unit
};

// Two levels of splits applied in one pass (quantized depthwise): the rows of a
// grandparent segment [off, off + cnt) go to its four grandchildren, each a contiguous
// run of the same range: AL at off, AR at off + n_al, BL at off + n_a, BR at
// off + n_a + n_bl (A = left child, B = right child).  Rows of a child that did not
// split are not moved (no later level reads them).
struct PartOp4 {
    u32 off, cnt, src, dst;
    u32 feature, bin, default_left, last_bin;  // the grandparent's split
    u32 fa, ba, da, la;                        // A's split (feature, bin, default_left, last_bin)
    u32 fb, bb, db, lb;                        // B's split
    u32 n_a, n_al, n_bl;
    u32 alive;                                 // bit 0: A splits, bit 1: B splits
    u32 chunk_first, chunk_count;
    u32 pad[2];
};
static_assert(sizeof(PartOp4) == 96, "PartOp4");

struct PartChunk {
    u32 op;
    u32 begin, end; // rows relative to the op's segment
    u32 left_before; // exclusive prefix of left rows over the op's earlier chunks
};

// Device-side control block of one tree under construction.
struct TreeCtl {
    u32 n_front;        // frontier size of the current level
    u32 n_nodes;        // nodes allocated in the tree
    u32 n_ops;          // split ops of the current level
    u32 n_part_chunks;
    u32 n_hist_chunks;
    u32 n_finish;
    u32 n_parts;        // partial histograms in use
    u32 next_front;     // frontier size after the partition
    u32 iter;           // boosting iteration (RNG key and tree slot)
    u32 depth;          // level being processed
    u32 done;           // leafwise: growth finished
    u32 live;           // leafwise: live leaves
    u32 n_cand;         // leafwise: candidates in the pool
    u32 n_pending;
    u32 sel;            // leafwise: selected candidate index
    u32 max_depth_seen;
    u32 oblivious_depth; // levelwise: levels split so far
    u32 part_error;     // a partition chunk's prefix disagreed with its op's row counts
    // validation / early stopping (the fit loop's rule, src/cli/pipeline.cpp:443-476)
    u32 es_n;           // validation losses recorded since the validation set was armed
    u32 es_best_iter;   // index of the best (lowest) loss
    u32 es_stopped;     // 1 once es_n - 1 - es_best_iter reached the patience
    float es_best;      // best loss
    u32 root_n;         // rows of the root segment when a device sampler chose them
    u32 pad[9];
};
static_assert(sizeof(TreeCtl) == 128, "TreeCtl");

// Where a tree's root rows are: every row in the pristine plane, or a row subset (a
// sampler's kept rows, ascending) gathered into work plane 1, with its count known on
// the host or left in TreeCtl::root_n by the device sampler.
struct RootSel {
    u32 n;          // rows (when known on the host)
    u32 src;        // kSrcPlane0 or 1
    u32 n_from_ctl; // 1: the count is TreeCtl::root_n
};

// Launch-shape knobs of one context, picked by the autotuner (bh_autotune).
struct Tuning {
    u32 scatter_lds = 20u * 1024u; // LDS bytes per partition-scatter workgroup (sets its tile rows)
    u32 find_blocks = kFindMaxBlocks; // workgroups sharing one node's split find (1..kFindMaxBlocks)
    u32 hist_threads = 512;      // quantized histogram workgroup size (256, 512 or 1024)
    u32 part_grid = 2048;        // partition count / scatter workgroups
    u32 hist_layout = 0;         // quantized histogram: 0 row-dword lanes, 1 feature-in-lane (rows <= 32 B),
                                 // 2 row-in-lane (rows <= 64 B; routed levels always use it)
    u32 pair_levels = 1;         // quantized depthwise, rows <= 64 bytes: move rows every second level
    u32 hist_group = 1;          // quantized depthwise: sub-units per histogram unit of full-scan levels
};

// Per-node partial split records and completion counters of a multi-block split find
// (capacity: front indices x kFindMaxBlocks records; counters start and end at 0).
struct FindScratch {
    SplitRec *part;
    u32 *done;
};

// Tree-growth settings shared by every kernel of a fit.
// Interaction constraints: at most kMaxGroups groups over at most 32 * kMaxPathWords
// features.
constexpr u32 kMaxGroups = 256, kMaxPathWords = 128;

struct GrowCfg {
    double min_child_hess, min_gain, l2, l1;
    int32_t max_depth, min_data_in_leaf, max_leaves, quantized;
    // Per-tree feature selection (sample_features, src/grower_impl.hpp:165-180): the
    // finders skip a feature whose byte is 0 (src/split.cpp:106-110); nullptr selects all.
    u8 const *fmask;
    // Monotone constraints (tree.monotone_constraints, include/bonsai/config/tree_config.hpp:
    // 20-22): per-feature direction (+1, -1, 0), nullptr for none; then bounds holds the
    // [lo, hi] leaf-weight box of every tree node id (2 doubles per node).
    int8_t const *mono;
    double *bounds;
    // Interaction constraints (tree.interaction_constraints, tree_config.hpp:23-26):
    // n_groups feature sets as path_words-word bitmasks, and every tree node id's path
    // (the distinct features split on above it) in the same form; n_groups 0: none.
    u32 const *groups;
    u32 *paths;
    u32 n_groups, path_words;
};

// Compact node for the bin-space walks (score update, predict).
struct WalkNode {
    u16 feature; // 0xFFFF: leaf
    u8  bin;
    u8  default_left;
    u16 left, right;
    float value;
};
static_assert(sizeof(WalkNode) == 12, "WalkNode");

} // namespace bh
