// Host runtime of the MI355X GBDT engine: device buffers, the three grow loops as
// device-resident kernel sequences, hipGraph capture of a whole boosting iteration, and
// the extern "C" boundary declared in include/bonsai_hip.h.
//
// The grow loops keep the reference's control semantics (src/grower_impl.hpp) but run
// their decisions on the device: the host only enqueues fixed-shape launches, so an
// iteration has no host round-trip unless host_sync asks for the reference's cadence.
#include <hip/hip_runtime.h>
#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/bonsai_hip.h"
#include "bh_common.hpp"
#include "kernels.hpp"
#include "kernels_ext.hpp"

using namespace bh;

namespace {

thread_local std::string g_err;

int fail(int code, std::string const &msg)
{
    g_err = msg;
    return code;
}

#define HIP_TRY(expr)                                                                        \
    do {                                                                                     \
        hipError_t const e_ = (expr);                                                        \
        if (e_ != hipSuccess)                                                                \
            return fail(-1, std::string("hip: ") + #expr + ": " + hipGetErrorString(e_));    \
    } while (0)

template <typename T> hipError_t dalloc(T **p, size_t count)
{
    *p = nullptr;
    if (count == 0) count = 1;
    return hipMalloc(reinterpret_cast<void **>(p), count * sizeof(T));
}

// Device buffers of one ingest call, freed on every return path.
template <typename T> struct DevScratch {
    T *p = nullptr;
    DevScratch() = default;
    DevScratch(DevScratch const &) = delete;
    DevScratch &operator=(DevScratch const &) = delete;
    ~DevScratch()
    {
        if (p) (void)hipFree(p);
    }
};

u32 round_up4(u32 x) { return (x + 3u) & ~3u; }

} // namespace

// ---------------------------------------------------------------------------------
// Scratch of the no-allocation histogram entry point (bh_histogram_noalloc).  One copy
// per GPU, held until the process ends: the histogram work list, its completion op and
// control block, then the unit partials.  Host code never passes the array's name to a
// kernel (on the host it is a stand-in, not a device address): bh_create resolves each
// device's address once with hipGetSymbolAddress and every use goes through that.
// ---------------------------------------------------------------------------------
namespace {

constexpr size_t kNoAllocScratchBytes = 32ull << 20;
constexpr size_t kNoAllocCtlOff = 0;                 // TreeCtl (128 bytes)
constexpr size_t kNoAllocFinishOff = 256;            // FinishOp (64 bytes)
constexpr size_t kNoAllocChunkOff = 512;             // HistChunk[kNoAllocMaxUnits]
constexpr u32 kNoAllocMaxUnits = 2048;
constexpr size_t kNoAllocPartOff =
    (kNoAllocChunkOff + kNoAllocMaxUnits * sizeof(HistChunk) + 255) / 256 * 256; // unit partials
static_assert(sizeof(TreeCtl) <= kNoAllocFinishOff - kNoAllocCtlOff, "TreeCtl fits its scratch slot");
static_assert(sizeof(FinishOp) <= kNoAllocChunkOff - kNoAllocFinishOff, "FinishOp fits its scratch slot");

__device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
unsigned char *g_scratch_ptr[kMaxDevices] = {};

// Resolves this device's copy of the scratch array (the device must be current).
int resolve_scratch(int device)
{
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (device < 0 || device >= kMaxDevices) return fail(1, "device id beyond the scratch table");
    if (g_scratch_ptr[device]) return 0;
    void *ptr = nullptr;
    hipError_t const e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || ptr == nullptr)
        return fail(-1, std::string("hip: scratch address: ") + hipGetErrorString(e));
    g_scratch_ptr[device] = static_cast<unsigned char *>(ptr);
    return 0;
}

// One node over rows [0, n) of work plane 0: `units` histogram units of kHistChunk rows
// (the same decomposition bh_histogram uploads), their completion op, and a control
// block cleared except for the unit and op counts.  One workgroup.
__global__ __launch_bounds__(256) void k_noalloc_plan(TreeCtl *__restrict__ ctl, FinishOp *__restrict__ fop,
                                                      HistChunk *__restrict__ chunks, u32 units, u32 n)
{
    u32 *cw = reinterpret_cast<u32 *>(ctl);
    for (u32 i = threadIdx.x; i < sizeof(TreeCtl) / 4u; i += blockDim.x) cw[i] = 0u;
    for (u32 k = threadIdx.x; k < units; k += blockDim.x) {
        HistChunk h{};
        h.slot = 0;
        h.src = 0;
        h.begin = k * kHistChunk;
        h.end = min(n, (k + 1u) * kHistChunk);
        h.part = units > 1u ? k : kNone;
        chunks[k] = h;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        FinishOp f{};
        f.small_slot = 0;
        f.large_slot = kNone;
        f.part_first = 0;
        f.part_count = units > 1u ? units : 0u;
        f.empty = n == 0u ? 1u : 0u;
        *fop = f;
        ctl->n_hist_chunks = units;
        ctl->n_finish = 1;
    }
}

} // namespace

void bh_set_error(std::string const &m) { g_err = m; }

// One slot's rows for the engine steps: a segment of a plane (0/1 work, 2 pristine)
// and the routing-tree node the rows sit in.
struct StepSeg {
    u32 off = 0, cnt = 0, src = 2, route = 0;
    bool live = false;
};

struct bh_context {
    int device = 0;
    hipStream_t stream = nullptr;
    // data plane
    u64 N = 0;
    u32 F = 0, Fs = 0, SB = 0;
    std::vector<u16> n_bins;
    u8 *plane0 = nullptr, *work[2] = {nullptr, nullptr};
    u16 *d_nbins = nullptr;
    float *label = nullptr, *score = nullptr;
    float *weight = nullptr; // per-row weights (bh_upload_weights), null: unweighted
    bool have_labels = false;
    // quantized fits: the gradient maxima of the current scores are not in maxbits yet
    // (scores, labels or the fit changed outside the iteration loop)
    bool max_stale = true;
    // validation plane (eval_begin / eval_accumulate, device_context.cu:1666-1760)
    u8 *vplane = nullptr;
    float *vlabel = nullptr, *vscore = nullptr, *vhist = nullptr;
    double *vpartial = nullptr;
    u64 vN = 0;
    u32 vhist_cap = 0, es_rounds = 0, es_base = 0;
    bool have_valid = false, es_done = false;
    // gradients
    float2 *gh0f = nullptr, *ghwf[2] = {nullptr, nullptr};
    u32 *gh0q = nullptr, *ghwq[2] = {nullptr, nullptr};
    u32 *maxbits = nullptr;
    double *scales = nullptr;
    u32 *fx = nullptr; // FP32 histogram grids: maxima of |g|, h (zero between passes), s_g, s_h
    // histograms and work lists
    void *hist[2] = {nullptr, nullptr};
    void *partials = nullptr;
    TreeCtl *ctl = nullptr;
    NodeInfo *front[2] = {nullptr, nullptr};
    SplitRec *splits = nullptr;
    FindScratch fscr{nullptr, nullptr};
    u32 *cnt[2] = {nullptr, nullptr}; // quantized depthwise: row count per (slot, feature, bin)
    PartOp4 *ops4 = nullptr;          // quantized depthwise: four-way partition ops
    uint4 *chunk_cls = nullptr;       // and their per-chunk class counts,
    u32 *cursors4 = nullptr;          // or their class cursors (kCursorStride apart)
    Tuning tune;
    DevNode *tree = nullptr;
    PartOp *ops = nullptr;
    PartChunk *pch = nullptr;
    u32 *chunk_left = nullptr;
    HistChunk *hch = nullptr;
    FinishOp *fops = nullptr;
    u32 *leaf_rows = nullptr;
    ExtState ext{};
    // outputs
    DevNode *trees_out = nullptr;
    u32 *tree_sizes = nullptr;
    u32 trees_cap = 0, trees_done = 0, max_nodes = 0;
    // capacities
    u32 max_front = 0, max_ops = 0, max_pch = 0, max_hch = 0, max_parts = 0, hist_slots = 0;
    u32 lw_units = 0; // quantized leafwise: histogram units per child at most
    // row sampling of the tree's root (bh_set_row_subset / bh_set_bernoulli)
    u32 samp_kind = 0;           // 0 every row, 1 a fixed subset, 2 a Bernoulli draw per iteration
    u32 *samp_rows = nullptr;    // kind 1: the subset, ascending
    u32 samp_n = 0;
    u32 samp_thr = 0;            // kind 2: keep when 32 random bits < thr
    u64 samp_seed = 0;
    u32 *samp_cnt = nullptr;     // kind 2: per-chunk kept counts, then their prefix
    // feature sampling (bh_set_feature_mask / bh_set_feature_fraction)
    u32 fsamp_kind = 0;          // 0 every feature, 1 a fixed mask, 2 a draw per iteration
    u8 *fmask = nullptr;         // [F] 1 = selected
    u32 fsamp_k = 0;             // kind 2: features selected per tree
    u64 fsamp_seed = 0;
    // monotone constraints (bh_set_monotone): directions by sign, per-node boxes
    std::vector<int8_t> mono_host;
    int8_t *mono_dev = nullptr;  // [F]
    double *bounds = nullptr;    // [max_nodes][2]
    // interaction constraints (bh_set_interaction): groups as feature bitmasks, node paths
    std::vector<std::vector<u32>> groups_host;
    u32 *groups_dev = nullptr;   // [n_groups][path_words]
    u32 *paths = nullptr;        // [max_nodes][path_words]
    // configuration
    bh_fit_config fc{};
    GrowCfg gc{};
    bool configured = false;
    bool quant = false;
    // graph
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    // histogram profiling (per-launch hipEvents, non-graph iterations only)
    bool prof = false;
    double prof_ms = 0.0, prof_bytes = 0.0, prof_scanned_bytes = 0.0;
    u64 prof_launches = 0, prof_rows = 0, prof_scanned = 0;
    // engine steps (bh_begin_root ... bh_finalize_tree, bh_leaf_*): the segment of every
    // slot of the current and next level (level plane) or of the slot pool (leaf plane),
    // the histogram buffer of the current level, and the routing tree of the partitions
    int step_mode = 0; // 0 none, 1 level plane, 2 leaf plane
    std::vector<StepSeg> seg_cur, seg_next, pool;
    u32 step_hist = 0, step_next_slot = 0;
    std::vector<DevNode> route;
    std::vector<u32> route_stamp;
    // timing
    bool timing = false;
    std::vector<float> iter_ms;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;

    // Features per histogram slice: all of them up to 64, else up to 128, as long as the
    // slice's LDS cells fit 128 KiB (8 B per quantized cell, 16 B per FP32 g / h pair).
    u32 hist_fsl() const
    {
        u32 const cell = quant ? 8u : 16u;
        u32 f = F <= 64 ? round_up4(F) : 128u;
        while (f > 4 && static_cast<size_t>(f) * SB * cell > (128u << 10)) f -= 4;
        return f;
    }
    u32 hist_waves() const
    {
        u32 const groups = (std::min(F, hist_fsl()) + 3u) / 4u;
        // Quantized cells are order-free: two waves share each feature group.
        u32 const w = quant ? 2u * groups : groups;
        return std::max(1u, std::min(w, 16u));
    }
    size_t cell_bytes() const { return quant ? 16 : 8; }
    // unit partials: float2 cells (FP32), three u32 planes g / h / count (quantized)
    size_t part_bytes() const { return quant ? 12 : 8; }
    // Rows per histogram work unit of the counted (quantized depthwise) path on levels
    // that scan every live row (the root and routed levels): groups of hist_group
    // sub-units, unless the feature-in-lane layout (single sub-units) is set.  Levels that
    // build the small children only keep single sub-units: their fewer units would leave
    // workgroup slots idle.
    u32 unit_rows() const { return quant && tune.hist_layout != 1u ? kHistChunk * tune.hist_group : kHistChunk; }
    u32 node_cells() const { return F * SB; }
};

namespace {

void free_fit(bh_context *c)
{
    if (c->graph_exec) hipGraphExecDestroy(c->graph_exec);
    if (c->graph) hipGraphDestroy(c->graph);
    c->graph_exec = nullptr;
    c->graph = nullptr;
    void *ptrs[] = {c->gh0f, c->ghwf[0], c->ghwf[1], c->gh0q, c->ghwq[0], c->ghwq[1], c->maxbits, c->scales, c->fx,
                    c->hist[0], c->hist[1], c->partials, c->ctl, c->front[0], c->front[1], c->splits, c->tree,
                    c->ops, c->pch, c->chunk_left, c->hch, c->fops, c->trees_out, c->tree_sizes, c->leaf_rows,
                    c->fscr.part, c->fscr.done, c->cnt[0], c->cnt[1], c->ops4, c->chunk_cls, c->cursors4,
                    c->mono_dev,
                    c->bounds, c->groups_dev, c->paths};
    for (void *p : ptrs)
        if (p) hipFree(p);
    c->mono_dev = nullptr;
    c->bounds = nullptr;
    c->gc.mono = nullptr;
    c->gc.bounds = nullptr;
    c->groups_dev = c->paths = nullptr;
    c->gc.groups = c->gc.paths = nullptr;
    c->gc.n_groups = c->gc.path_words = 0;
    c->gh0f = c->ghwf[0] = c->ghwf[1] = nullptr;
    c->gh0q = c->ghwq[0] = c->ghwq[1] = nullptr;
    c->maxbits = nullptr;
    c->scales = nullptr;
    c->fx = nullptr;
    c->hist[0] = c->hist[1] = c->partials = nullptr;
    c->ctl = nullptr;
    c->front[0] = c->front[1] = nullptr;
    c->splits = nullptr;
    c->tree = nullptr;
    c->ops = nullptr;
    c->pch = nullptr;
    c->chunk_left = nullptr;
    c->hch = nullptr;
    c->fops = nullptr;
    c->trees_out = nullptr;
    c->tree_sizes = nullptr;
    c->leaf_rows = nullptr;
    c->fscr = FindScratch{nullptr, nullptr};
    c->cnt[0] = c->cnt[1] = nullptr;
    c->ops4 = nullptr;
    c->chunk_cls = nullptr;
    c->cursors4 = nullptr;
    ext_free(c->ext);
    c->configured = false;
    c->trees_cap = c->trees_done = 0;
}

// Node capacity of one tree under the configured policy.
u32 tree_capacity(bh_fit_config const &f)
{
    if (f.grower == BH_LEAFWISE) return std::min<u32>(kMaxTreeNodes, 4u * static_cast<u32>(f.max_leaves) + 8u);
    return std::min<u32>(kMaxTreeNodes, (2u << static_cast<u32>(f.max_depth)) + 2u);
}

// Device copies of the monotone directions and the per-node boxes for the configured
// fit; none when every direction is 0.  The levelwise grower refuses constraints
// (src/grower_impl.hpp:583-588).
int arm_monotone(bh_context *c)
{
    if (c->mono_dev) (void)hipFree(c->mono_dev);
    if (c->bounds) (void)hipFree(c->bounds);
    c->mono_dev = nullptr;
    c->bounds = nullptr;
    c->gc.mono = nullptr;
    c->gc.bounds = nullptr;
    bool any = false;
    for (int8_t v : c->mono_host) any = any || v != 0;
    if (!any || !c->configured) return 0;
    if (c->fc.grower == BH_LEVELWISE) return fail(1, "monotone_constraints are not supported by the levelwise grower");
    std::vector<int8_t> d(c->F, 0);
    for (size_t f = 0; f < c->mono_host.size() && f < c->F; ++f) d[f] = c->mono_host[f];
    HIP_TRY(dalloc(&c->mono_dev, c->F));
    HIP_TRY(dalloc(&c->bounds, 2 * static_cast<size_t>(c->max_nodes)));
    HIP_TRY(hipMemcpy(c->mono_dev, d.data(), c->F, hipMemcpyHostToDevice));
    c->gc.mono = c->mono_dev;
    c->gc.bounds = c->bounds;
    return 0;
}

// Device copies of the interaction groups and the per-node paths for the configured
// fit; none without groups.  The levelwise grower refuses them
// (src/grower_impl.hpp:589-595).
int arm_interaction(bh_context *c)
{
    if (c->groups_dev) (void)hipFree(c->groups_dev);
    if (c->paths) (void)hipFree(c->paths);
    c->groups_dev = c->paths = nullptr;
    c->gc.groups = c->gc.paths = nullptr;
    c->gc.n_groups = c->gc.path_words = 0;
    if (c->groups_host.empty() || !c->configured) return 0;
    if (c->fc.grower == BH_LEVELWISE)
        return fail(1, "interaction_constraints are not supported by the levelwise grower");
    u32 const W = (c->F + 31u) / 32u;
    if (W > kMaxPathWords) return fail(1, "interaction constraints support at most 4096 features");
    std::vector<u32> bits(c->groups_host.size() * W, 0u);
    for (size_t g = 0; g < c->groups_host.size(); ++g)
        for (u32 f : c->groups_host[g]) {
            if (f >= c->F) return fail(1, "interaction group names a feature beyond n_features");
            bits[g * W + f / 32u] |= 1u << (f % 32u);
        }
    HIP_TRY(dalloc(&c->groups_dev, bits.size()));
    HIP_TRY(dalloc(&c->paths, static_cast<size_t>(c->max_nodes) * W));
    HIP_TRY(hipMemcpy(c->groups_dev, bits.data(), bits.size() * sizeof(u32), hipMemcpyHostToDevice));
    c->gc.groups = c->groups_dev;
    c->gc.paths = c->paths;
    c->gc.n_groups = static_cast<u32>(c->groups_host.size());
    c->gc.path_words = W;
    return 0;
}

int alloc_fit(bh_context *c)
{
    bh_fit_config const &f = c->fc;
    u32 const N = static_cast<u32>(c->N);
    c->quant = f.quantized != 0;
    c->max_stale = true;
    // Frontier capacity: 2^max_depth for level growth, 2 for leafwise rounds.
    u32 const depth = static_cast<u32>(f.max_depth);
    c->max_front = f.grower == BH_LEAFWISE ? 2u : (1u << std::min(depth, 14u));
    c->max_ops = std::max(1u, c->max_front);
    c->max_pch = N / part_chunk_rows(c->Fs) + c->max_ops + 2;
    c->max_hch = N / kHistChunk + c->max_ops + 2;
    c->max_parts = N / kHistChunk + c->max_ops + 2;
    c->hist_slots = f.grower == BH_LEAFWISE ? static_cast<u32>(f.max_leaves) + 4u : std::max(1u, c->max_front / 2u);
    c->max_nodes = tree_capacity(f);
    size_t const nc = c->node_cells();
    c->lw_units = 0;
    if (f.grower == BH_LEAFWISE) {
        // Leaf-sized partition chunks (at most kLwChunkTarget beyond the full-size ones)
        // and up to lw_units histogram units per built node, their partials within about
        // 256 MiB.
        c->max_pch += kLwChunkTarget;
        size_t const per = std::max<size_t>(1, nc * c->part_bytes());
        c->lw_units = static_cast<u32>(std::min<size_t>(kLwUnitTarget, std::max<size_t>(16, (256ull << 20) / per)));
        c->max_hch = std::max(c->max_hch, c->lw_units + 2u);
        c->max_parts = std::max(c->max_parts, c->lw_units + 2u);
    }
    size_t const hist_bytes = static_cast<size_t>(c->hist_slots) * nc * c->cell_bytes();
    if (c->quant) {
        HIP_TRY(dalloc(&c->gh0q, N + kTailPad));
        HIP_TRY(dalloc(&c->ghwq[0], N + kTailPad));
        HIP_TRY(dalloc(&c->ghwq[1], N + kTailPad));
    } else {
        HIP_TRY(dalloc(&c->gh0f, N + kTailPad));
        HIP_TRY(dalloc(&c->ghwf[0], N + kTailPad));
        HIP_TRY(dalloc(&c->ghwf[1], N + kTailPad));
    }
    HIP_TRY(dalloc(&c->maxbits, 4));
    HIP_TRY(dalloc(&c->scales, 4));
    HIP_TRY(hipMemset(c->scales, 0, 4 * sizeof(double)));
    HIP_TRY(dalloc(&c->fx, 4));
    HIP_TRY(hipMemset(c->fx, 0, 4 * sizeof(u32)));
    HIP_TRY(hipMalloc(&c->hist[0], hist_bytes));
    if (f.grower != BH_LEAFWISE) HIP_TRY(hipMalloc(&c->hist[1], hist_bytes));
    HIP_TRY(hipMalloc(&c->partials, static_cast<size_t>(c->max_parts) * nc * c->part_bytes()));
    if (c->quant && f.grower == BH_DEPTHWISE) {
        size_t const cells = static_cast<size_t>(c->hist_slots) * nc;
        HIP_TRY(dalloc(&c->cnt[0], cells));
        HIP_TRY(dalloc(&c->cnt[1], cells));
        HIP_TRY(dalloc(&c->ops4, c->max_ops / 2 + 2));
        HIP_TRY(dalloc(&c->chunk_cls, c->max_pch));
        HIP_TRY(dalloc(&c->cursors4, (static_cast<size_t>(c->max_ops) / 2 + 2) * 4u * kCursorStride));
    }
    if (c->quant && f.grower == BH_LEAFWISE) HIP_TRY(dalloc(&c->cnt[0], static_cast<size_t>(c->hist_slots) * nc));
    HIP_TRY(dalloc(&c->ctl, 1));
    HIP_TRY(hipMemset(c->ctl, 0, sizeof(TreeCtl)));
    u32 const fcap = std::max(c->max_front * 2u, 4u);
    HIP_TRY(dalloc(&c->front[0], fcap));
    HIP_TRY(dalloc(&c->front[1], fcap));
    HIP_TRY(dalloc(&c->splits, fcap));
    HIP_TRY(dalloc(&c->fscr.part, static_cast<size_t>(fcap) * kFindMaxBlocks));
    HIP_TRY(dalloc(&c->fscr.done, fcap));
    HIP_TRY(hipMemset(c->fscr.done, 0, fcap * sizeof(u32)));
    HIP_TRY(dalloc(&c->tree, c->max_nodes));
    HIP_TRY(dalloc(&c->ops, c->max_ops + 2));
    HIP_TRY(dalloc(&c->pch, c->max_pch));
    HIP_TRY(dalloc(&c->chunk_left, c->max_pch));
    HIP_TRY(dalloc(&c->hch, c->max_hch));
    HIP_TRY(dalloc(&c->fops, c->max_ops + 2));
    HIP_TRY(dalloc(&c->leaf_rows, N));
    if (ext_alloc(c->ext, f, c->F, c->SB, c->max_front, c->hist_slots) != hipSuccess)
        return fail(-1, "hip: extension buffers");
    c->gc.min_child_hess = f.min_child_hess;
    c->gc.min_gain = f.min_gain_to_split;
    c->gc.l2 = f.lambda_l2;
    c->gc.l1 = f.lambda_l1;
    c->gc.max_depth = f.max_depth;
    c->gc.min_data_in_leaf = f.min_data_in_leaf;
    c->gc.max_leaves = f.max_leaves;
    c->gc.quantized = f.quantized;
    c->configured = true;
    int r = arm_monotone(c);
    if (r == 0) r = arm_interaction(c);
    if (r != 0) free_fit(c); // a fit whose constraints could not be armed is no fit
    return r;
}

int ensure_tree_out(bh_context *c, u32 n_more)
{
    u32 const need = c->trees_done + n_more;
    if (need <= c->trees_cap) return 0;
    u32 const cap = std::max(need, c->trees_cap * 2u + 16u);
    DevNode *nt = nullptr;
    u32 *ns = nullptr;
    HIP_TRY(dalloc(&nt, static_cast<size_t>(cap) * c->max_nodes));
    HIP_TRY(dalloc(&ns, cap));
    if (c->trees_out) {
        HIP_TRY(hipMemcpy(nt, c->trees_out, static_cast<size_t>(c->trees_cap) * c->max_nodes * sizeof(DevNode),
                          hipMemcpyDeviceToDevice));
        HIP_TRY(hipMemcpy(ns, c->tree_sizes, c->trees_cap * sizeof(u32), hipMemcpyDeviceToDevice));
        hipFree(c->trees_out);
        hipFree(c->tree_sizes);
    }
    c->trees_out = nt;
    c->tree_sizes = ns;
    c->trees_cap = cap;
    // the captured graph holds the old pointers
    if (c->graph_exec) hipGraphExecDestroy(c->graph_exec);
    if (c->graph) hipGraphDestroy(c->graph);
    c->graph_exec = nullptr;
    c->graph = nullptr;
    return 0;
}

void const *gh_src(bh_context *c, int which)
{
    if (c->quant) return which == 2 ? static_cast<void const *>(c->gh0q) : c->ghwq[which];
    return which == 2 ? static_cast<void const *>(c->gh0f) : c->ghwf[which];
}

hipError_t host_sync_point(bh_context *c)
{
    // The reference's blocking fetch after find / partition (device_context.cu:862,1038):
    // a small D2H copy and a stream synchronisation.
    TreeCtl tmp;
    hipError_t e = hipMemcpyAsync(&tmp, c->ctl, sizeof(TreeCtl), hipMemcpyDeviceToHost, c->stream);
    if (e != hipSuccess) return e;
    return hipStreamSynchronize(c->stream);
}

// Profiling: time one histogram launch with hipEvents on its stream; after its
// completion op, count its bytes two ways.  Algorithmic (the scope table's B_hist):
// every built node k reads its n_k rows once, n_k x (F bin bytes + gradient pair bytes),
// and writes one 8-byte cell per (feature, bin); rows are physically compacted per node,
// so no row-index bytes are read.  n_k comes from the built node's count cells (feature
// 0's bins sum to its rows) or, without count cells, from its units (which then cover
// exactly its rows).  Scanned: the rows the units actually read (a routed unit scans
// its parent's whole segment and keeps the rows routed to the built child).
float profiled_hist(bh_context *c, void *hist_next, u32 grid_units, u32 *cnt_next, PartOp const *routes, bool group)
{
    (void)hipEventRecord(c->ev0, c->stream);
    launch_hist(c->stream, c->quant, c->hch, c->ctl, c->work[0], c->work[1], c->plane0, gh_src(c, 0),
                gh_src(c, 1), gh_src(c, 2), c->Fs, c->F, c->SB, c->hist_fsl(), c->hist_waves(), grid_units,
                c->partials, hist_next, c->fx, cnt_next, routes, group);
    (void)hipEventRecord(c->ev1, c->stream);
    (void)hipEventSynchronize(c->ev1);
    float ms = 0.0F;
    (void)hipEventElapsedTime(&ms, c->ev0, c->ev1);
    return ms;
}

void profiled_count(bh_context *c, float ms, u32 const *cnt_next)
{
    (void)hipStreamSynchronize(c->stream);
    TreeCtl ctl;
    (void)hipMemcpy(&ctl, c->ctl, sizeof(ctl), hipMemcpyDeviceToHost);
    if (ctl.n_hist_chunks == 0 || ctl.n_finish == 0) return;
    std::vector<HistChunk> hc(ctl.n_hist_chunks);
    (void)hipMemcpy(hc.data(), c->hch, hc.size() * sizeof(HistChunk), hipMemcpyDeviceToHost);
    std::vector<FinishOp> fo(ctl.n_finish);
    (void)hipMemcpy(fo.data(), c->fops, fo.size() * sizeof(FinishOp), hipMemcpyDeviceToHost);
    u64 scanned = 0;
    for (auto const &u : hc) scanned += u.end - u.begin;
    u64 built = 0, nodes = 0;
    u32 const nb0 = c->n_bins[0];
    std::vector<u32> cn(nb0);
    for (auto const &op : fo) {
        if (op.empty) continue;
        ++nodes;
        if (cnt_next) {
            (void)hipMemcpy(cn.data(), cnt_next + static_cast<size_t>(op.small_slot) * c->node_cells(),
                            nb0 * sizeof(u32), hipMemcpyDeviceToHost);
            for (u32 v : cn) built += v;
        } else {
            for (auto const &u : hc)
                if (u.slot == op.small_slot) built += u.end - u.begin;
        }
    }
    double sum_bins = 0.0;
    for (u16 b : c->n_bins) sum_bins += b;
    double const cells = static_cast<double>(nodes) * sum_bins * 8.0;
    double const row_bytes = c->F + (c->quant ? 4.0 : 8.0);
    c->prof_ms += ms;
    c->prof_rows += built;
    c->prof_scanned += scanned;
    c->prof_bytes += static_cast<double>(built) * row_bytes + cells;
    c->prof_scanned_bytes += static_cast<double>(scanned) * row_bytes + cells;
    c->prof_launches += 1;
}

// Histogram units -> partials / slots, then the fixed-order completion.
// routes: the split rules of the level's routed units (levels whose rows were not
// moved), else null.
void enqueue_build(bh_context *c, void *hist_cur, void *hist_next, u32 max_ops, u32 grid_units,
                   u32 const *cnt_cur = nullptr, u32 *cnt_next = nullptr, bool advance = false,
                   PartOp const *routes = nullptr, bool group = false)
{
    float ms = 0.0F;
    if (c->prof)
        ms = profiled_hist(c, hist_next, grid_units, cnt_next, routes, group);
    else
        launch_hist(c->stream, c->quant, c->hch, c->ctl, c->work[0], c->work[1], c->plane0, gh_src(c, 0),
                    gh_src(c, 1), gh_src(c, 2), c->Fs, c->F, c->SB, c->hist_fsl(), c->hist_waves(), grid_units,
                    c->partials, hist_next, c->fx, cnt_next, routes, group);
    launch_finish(c->stream, c->quant, c->fops, c->ctl, c->partials, hist_cur, hist_next, c->node_cells(), max_ops,
                  cnt_cur, cnt_next, advance, hist_transpose_F(c->quant, c->Fs, c->F, group));
    if (c->prof) profiled_count(c, ms, cnt_next);
}

u32 hist_grid(bh_context *c) { return std::max(1u, std::min(c->max_hch, 4096u)); }

// The tree's root rows: the pristine plane, or the sampled rows in work plane 1
// (SplitInput make_root over row_indices, src/step/level.hpp:171-187; the rows the
// sampler dropped still get their scores from the tree walk over every row, the
// reference's route_unsampled, src/grower_impl.hpp:430-482).
RootSel root_sel(bh_context *c)
{
    if (c->samp_kind == 1) return RootSel{c->samp_n, 1u, 0u};
    if (c->samp_kind == 2) return RootSel{0u, 1u, 1u};
    return RootSel{static_cast<u32>(c->N), kSrcPlane0, 0u};
}

// Gathers this iteration's sampled rows (and their gradient words) into work plane 1.
void enqueue_sample(bh_context *c)
{
    if (c->fsamp_kind == 2)
        launch_feature_draw(c->stream, c->ctl, c->F, c->fsamp_k, c->fsamp_seed, c->fmask);
    u32 const gh_bytes = c->quant ? 4 : 8;
    if (c->samp_kind == 1)
        launch_gather_rows(c->stream, c->plane0, c->Fs, gh_src(c, 2), gh_bytes, c->samp_rows, c->samp_n, c->work[1],
                           const_cast<void *>(gh_src(c, 1)), 0);
    else if (c->samp_kind == 2)
        launch_bernoulli(c->stream, c->ctl, static_cast<u32>(c->N), c->samp_seed, c->samp_thr, c->samp_cnt, c->plane0,
                         c->Fs, gh_src(c, 2), gh_bytes, c->work[1], const_cast<void *>(gh_src(c, 1)));
}
u32 part_grid(bh_context *c) { return std::max(1u, std::min(c->max_pch, c->tune.part_grid)); }

// Quantized depthwise: the histograms carry row counts, so each level is find (child
// counts from the count cells) -> plan (one workgroup: tree, demotion, next frontier,
// work lists) -> per-chunk counts -> scatter (chunk prefixes summed in-kernel) ->
// histograms of the smaller children.  The last level moves no rows and reads none.
// With paired levels (rows of at most 64 bytes) the rows move after every second level
// only: a scan of the parent segment is cheaper than a partition pass.
void enqueue_depthwise_counted(bh_context *c)
{
    hipStream_t s = c->stream;
    u32 const D = static_cast<u32>(c->fc.max_depth);
    u32 const gh_bytes = 4;
    // Paired levels: even levels leave the rows where they are (their children's units
    // scan the parent segments, routed), odd levels move them two splits at once.
    bool const paired = c->tune.pair_levels != 0 && c->Fs <= 64;
    u32 moves = 0; // partitions so far: they alternate between the two work planes
    u32 const unit_rows = c->unit_rows();
    bool const group = unit_rows > kHistChunk;
    launch_tree_begin(s, c->ctl, c->front[0], c->tree, c->hch, c->fops, root_sel(c), unit_rows);
    enqueue_build(c, c->hist[1], c->hist[0], 1, hist_grid(c), c->cnt[1], c->cnt[0], false, nullptr, group);
    for (u32 d = 0; d < D; ++d) {
        NodeInfo *cur = c->front[d & 1u], *nxt = c->front[(d + 1) & 1u];
        void *hcur = c->hist[d & 1u], *hnext = c->hist[(d + 1) & 1u];
        u32 *ccur = c->cnt[d & 1u], *cnext = c->cnt[(d + 1) & 1u];
        u32 const width = 1u << std::min(d, 14u);
        bool const last = d + 1 == D;
        u32 const mode = !paired || last ? 0u : ((d & 1u) ? 2u : 1u);
        u32 const dst = paired ? (moves & 1u) : (d & 1u);
        launch_find(s, true, cur, c->ctl, hcur, c->d_nbins, c->F, c->SB, c->gc, c->scales, c->splits, width, 0, 0,
                    c->fscr, ccur);
        launch_plan_counted(s, cur, c->splits, c->ctl, c->tree, c->ops, c->pch, nxt, c->hch, c->fops, c->d_nbins,
                            c->gc, c->scales, dst, last ? 1u : 0u, part_chunk_rows(c->Fs), mode, c->ops4,
                            mode == 1 ? unit_rows : kHistChunk, c->cursors4);
        if (c->fc.host_sync) host_sync_point(c);
        if (!last) {
            if (mode == 0) {
                launch_part_count(s, c->pch, c->ops, c->ctl, c->work[0], c->work[1], c->plane0, c->Fs,
                                  c->chunk_left, part_grid(c));
                launch_part_scatter(s, c->pch, c->ops, c->ctl, c->work[0], c->work[1], c->plane0, gh_src(c, 0),
                                    gh_src(c, 1), gh_src(c, 2), c->work[0], c->work[1],
                                    const_cast<void *>(gh_src(c, 0)), const_cast<void *>(gh_src(c, 1)), c->Fs,
                                    gh_bytes, part_grid(c), c->chunk_left);
                ++moves;
            } else if (mode == 2) {
                // A level of 2^(d-1) parents: with few of them every tile would meet on
                // the same cursors, so those levels count first (stable order); wider
                // levels reserve their runs and skip the counting pass.
                bool const reserve = width / 2u >= kCursorMinOps;
                if (!reserve)
                    launch_part_count4(s, c->pch, c->ops4, c->ctl, c->work[0], c->work[1], c->plane0, c->Fs,
                                       c->chunk_cls, part_grid(c));
                launch_part_scatter4(s, c->pch, c->ops4, c->ctl, c->work[0], c->work[1], c->plane0,
                                     static_cast<u32 const *>(gh_src(c, 0)), static_cast<u32 const *>(gh_src(c, 1)),
                                     static_cast<u32 const *>(gh_src(c, 2)), c->work[0], c->work[1],
                                     static_cast<u32 *>(const_cast<void *>(gh_src(c, 0))),
                                     static_cast<u32 *>(const_cast<void *>(gh_src(c, 1))), c->Fs, part_grid(c),
                                     c->chunk_cls, reserve ? c->cursors4 : nullptr);
                ++moves;
            }
            enqueue_build(c, hcur, hnext, std::min(width, c->max_ops), hist_grid(c), ccur, cnext, true,
                          mode == 1 ? c->ops : nullptr, group && mode == 1);
        } else {
            launch_next_level(s, c->ctl);
        }
        if (c->fc.host_sync) host_sync_point(c);
    }
    launch_finalize(s, c->front[D & 1u], c->ctl, c->tree, c->gc, c->scales, 1u << std::min(D, 14u));
}

// DepthwiseGrower::grow (src/grower_impl.hpp:501-577) as a device launch sequence.
void enqueue_depthwise(bh_context *c)
{
    if (c->quant) {
        enqueue_depthwise_counted(c);
        return;
    }
    hipStream_t s = c->stream;
    u32 const D = static_cast<u32>(c->fc.max_depth);
    u32 const gh_bytes = c->quant ? 4 : 8;
    launch_tree_begin(s, c->ctl, c->front[0], c->tree, c->hch, c->fops, root_sel(c));
    enqueue_build(c, c->hist[1], c->hist[0], 1, hist_grid(c));
    for (u32 d = 0; d < D; ++d) {
        NodeInfo *cur = c->front[d & 1u], *nxt = c->front[(d + 1) & 1u];
        void *hcur = c->hist[d & 1u], *hnext = c->hist[(d + 1) & 1u];
        u32 const width = 1u << std::min(d, 14u);
        bool const last = d + 1 == D;
        launch_find(s, c->quant, cur, c->ctl, hcur, c->d_nbins, c->F, c->SB, c->gc, c->scales, c->splits, width, 0, 0,
                    c->fscr);
        launch_plan_level(s, cur, c->splits, c->ctl, c->tree, c->ops, c->pch, c->d_nbins, c->gc, c->scales, d & 1u,
                          part_chunk_rows(c->Fs));
        if (c->fc.host_sync) host_sync_point(c);
        launch_part_count(s, c->pch, c->ops, c->ctl, c->work[0], c->work[1], c->plane0, c->Fs, c->chunk_left,
                          part_grid(c));
        launch_part_scan(s, c->quant, c->ops, c->pch, c->chunk_left, c->ctl, cur, nxt, c->splits, c->tree, c->hch,
                         c->fops, c->gc, c->scales, last ? 1u : 0u, 0u);
        if (c->fc.host_sync) host_sync_point(c);
        if (!last) {
            launch_part_scatter(s, c->pch, c->ops, c->ctl, c->work[0], c->work[1], c->plane0, gh_src(c, 0),
                                gh_src(c, 1), gh_src(c, 2), c->work[0], c->work[1],
                                const_cast<void *>(gh_src(c, 0)), const_cast<void *>(gh_src(c, 1)), c->Fs,
                                gh_bytes, part_grid(c));
            enqueue_build(c, hcur, hnext, std::min(width, c->max_ops), hist_grid(c));
        }
        launch_next_level(s, c->ctl);
        if (c->fc.host_sync) host_sync_point(c);
    }
    launch_finalize(s, c->front[D & 1u], c->ctl, c->tree, c->gc, c->scales, 1u << std::min(D, 14u));
}

void enqueue_grow(bh_context *c)
{
    enqueue_sample(c);
    if (c->gc.mono || c->gc.n_groups) launch_constraints_root(c->stream, c->bounds, c->paths, c->gc.path_words);
    if (c->fc.grower == BH_DEPTHWISE) enqueue_depthwise(c);
    else if (c->fc.grower == BH_LEVELWISE) enqueue_levelwise(c->ext, c->stream, c->quant, grow_view(c));
    else enqueue_leafwise(c->ext, c->stream, c->quant, grow_view(c));
}

// Quantized fits: the first iteration after scores, labels or the fit were set takes
// its gradient maxima from a pass of its own (later ones from the score update).
void prime_gradients(bh_context *c)
{
    if (!c->quant || !c->max_stale) return;
    launch_grad_prime(c->stream, c->score, c->label, c->weight, c->N, c->fc.objective, c->maxbits, c->ctl);
    c->max_stale = false;
}

void enqueue_gradients(bh_context *c)
{
    if (c->quant)
        launch_grad_quant(c->stream, c->score, c->label, c->weight, c->N, c->fc.objective, c->maxbits, c->fc.seed,
                          c->ctl, c->gh0q, c->scales);
    else
        launch_grad_f32(c->stream, c->score, c->label, c->weight, c->N, c->fc.objective, c->gh0f, c->fx);
}

// One boosting iteration: gradients, tree, score epilogue, tree copy-out.
void enqueue_iteration(bh_context *c)
{
    enqueue_gradients(c);
    enqueue_grow(c);
    launch_score_update(c->stream, c->plane0, c->Fs, c->N, c->tree, c->ctl, c->d_nbins, c->F, c->fc.learning_rate,
                        c->score, nullptr, c->max_nodes, c->label, c->fc.objective, c->quant ? c->maxbits : nullptr,
                        c->weight);
    if (c->have_valid) {
        // the new tree on the validation plane, then the round's loss and the
        // early-stopping rule (eval_accumulate + the fit loop's check)
        launch_score_update(c->stream, c->vplane, c->Fs, c->vN, c->tree, c->ctl, c->d_nbins, c->F,
                            c->fc.learning_rate, c->vscore, nullptr, c->max_nodes);
        launch_eval(c->stream, c->fc.objective, c->vscore, c->vlabel, static_cast<u32>(c->vN), c->vpartial, c->ctl,
                    c->vhist, c->vhist_cap, c->es_rounds);
    }
    launch_copy_tree(c->stream, c->tree, c->ctl, c->trees_out, c->tree_sizes, c->max_nodes);
    launch_iter_advance(c->stream, c->ctl);
}

// The counted scatter reports a chunk whose prefix disagrees with its op's row counts
// through the control block instead of writing outside the segment; surface it as a
// device error.
int check_partition(bh_context *c)
{
    if (!c->cnt[0]) return 0;
    TreeCtl t;
    HIP_TRY(hipMemcpy(&t, c->ctl, sizeof(TreeCtl), hipMemcpyDeviceToHost));
    if (t.part_error == 0) return 0;
    u32 const zero = 0;
    HIP_TRY(hipMemcpy(reinterpret_cast<unsigned char *>(c->ctl) + offsetof(TreeCtl, part_error), &zero, sizeof(u32),
                      hipMemcpyHostToDevice));
    return fail(-2, "partition: chunk prefix inconsistent with the split's row counts");
}

int check_ready(bh_context *c)
{
    if (!c) return fail(1, "null context");
    if (!c->plane0) return fail(1, "no bins uploaded");
    if (!c->configured) return fail(1, "not configured");
    set_tuning(c->tune);
    return 0;
}

} // namespace

// View of the context the leafwise / levelwise loops (kernels_ext.hip) need.
GrowView grow_view(bh_context *c)
{
    GrowView v;
    v.N = c->N;
    v.F = c->F;
    v.Fs = c->Fs;
    v.SB = c->SB;
    v.plane0 = c->plane0;
    v.work[0] = c->work[0];
    v.work[1] = c->work[1];
    v.gh[0] = gh_src(c, 0);
    v.gh[1] = gh_src(c, 1);
    v.gh[2] = gh_src(c, 2);
    v.d_nbins = c->d_nbins;
    v.hist[0] = c->hist[0];
    v.hist[1] = c->hist[1];
    v.partials = c->partials;
    v.ctl = c->ctl;
    v.front[0] = c->front[0];
    v.front[1] = c->front[1];
    v.splits = c->splits;
    v.fscr = c->fscr;
    v.tree = c->tree;
    v.ops = c->ops;
    v.pch = c->pch;
    v.chunk_left = c->chunk_left;
    v.hch = c->hch;
    v.fops = c->fops;
    v.scales = c->scales;
    v.gc = c->gc;
    v.fc = c->fc;
    v.max_pch = c->max_pch;
    v.max_hch = c->max_hch;
    v.max_ops = c->max_ops;
    v.cnt = c->cnt[0];
    v.lw_units = c->lw_units;
    v.root = root_sel(c);
    v.fsl = c->hist_fsl();
    v.fx = c->fx;
    v.waves = c->hist_waves();
    v.host_sync = c->fc.host_sync != 0;
    return v;
}

extern "C" {

const char *bh_last_error(void) { return g_err.c_str(); }

static int stream_calibrate(int device, uint64_t bytes, uint32_t reps, uint32_t mode, uint32_t grid, float *ms_mean)
{
    if (bytes < 16 || bytes % 16 != 0) return fail(1, "bytes must be a positive multiple of 16");
    if (reps == 0) return fail(1, "reps must be positive");
    if (mode > 2) return fail(1, "mode must be 0 (plain copy), 1 (wide copy) or 2 (wide read)");
    if (mode != 0 && (grid == 0 || grid > 65536)) return fail(1, "grid must be in [1, 65536]");
    HIP_TRY(hipSetDevice(device));
    void *src = nullptr, *dst = nullptr;
    hipStream_t st = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipError_t e = hipMalloc(&src, bytes);
    if (e == hipSuccess) e = hipMalloc(&dst, bytes);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    if (e == hipSuccess) e = hipMemsetAsync(src, 1, bytes, st);
    if (e == hipSuccess) {
        launch_copy_calibrate(st, src, dst, bytes, mode, grid);   // warm
        (void)hipEventRecord(e0, st);
        for (u32 r = 0; r < reps; ++r) launch_copy_calibrate(st, src, dst, bytes, mode, grid);
        (void)hipEventRecord(e1, st);
        e = hipEventSynchronize(e1);
    }
    float ms = 0.0F;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (st) (void)hipStreamDestroy(st);
    if (src) (void)hipFree(src);
    if (dst) (void)hipFree(dst);
    if (e != hipSuccess) return fail(-1, std::string("hip: ") + hipGetErrorString(e));
    *ms_mean = ms / static_cast<float>(reps);
    return 0;
}

int bh_copy_calibrate(int device, uint64_t bytes, uint32_t reps, float *ms_mean)
{
    return stream_calibrate(device, bytes, reps, 0u, 0u, ms_mean);
}

int bh_stream_calibrate(int device, uint64_t bytes, uint32_t reps, uint32_t mode, uint32_t grid, float *ms_mean)
{
    return stream_calibrate(device, bytes, reps, mode, grid, ms_mean);
}

int bh_synchronize(bh_context *c)
{
    if (!c) return fail(1, "null context");
    HIP_TRY(hipSetDevice(c->device));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return 0;
}

int bh_device_count(int *count)
{
    hipError_t const e = hipGetDeviceCount(count);
    if (e != hipSuccess) {
        *count = 0;
        return fail(-1, std::string("hip: hipGetDeviceCount: ") + hipGetErrorString(e));
    }
    return 0;
}

int bh_create(int device, bh_context **out)
{
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return fail(-1, "hip: no device");
    if (device < 0 || device >= n) return fail(1, "device id out of range");
    HIP_TRY(hipSetDevice(device));
    auto *c = new bh_context();
    c->device = device;
    hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete c;
        return fail(-1, "hip: stream");
    }
    hipEventCreate(&c->ev0);
    hipEventCreate(&c->ev1);
    if (int r = resolve_scratch(device)) {
        hipStreamDestroy(c->stream);
        delete c;
        return r;
    }
    e = set_hist_lds_limits();
    if (e == hipSuccess) e = ext_set_limits();
    if (e != hipSuccess) {
        hipStreamDestroy(c->stream);
        delete c;
        return fail(-1, std::string("hip: LDS limit: ") + hipGetErrorString(e));
    }
    *out = c;
    return 0;
}

int bh_destroy(bh_context *c)
{
    if (!c) return 0;
    hipSetDevice(c->device);
    hipStreamSynchronize(c->stream);
    free_fit(c);
    void *ptrs[] = {c->plane0, c->work[0], c->work[1], c->d_nbins, c->label, c->score,
                    c->vplane, c->vlabel, c->vscore, c->vhist, c->vpartial, c->samp_rows, c->samp_cnt, c->fmask,
                    c->weight};
    for (void *p : ptrs)
        if (p) hipFree(p);
    if (c->ev0) hipEventDestroy(c->ev0);
    if (c->ev1) hipEventDestroy(c->ev1);
    hipStreamDestroy(c->stream);
    delete c;
    return 0;
}

// (Re)allocates the pristine plane, the two work planes, labels and scores for an
// n_rows x n_features dataset; the plane is zeroed (padding columns stay 0).
static int alloc_plane(bh_context *c, uint64_t n_rows, uint32_t n_features, const uint16_t *n_bins)
{
    if (n_rows == 0 || n_features == 0) return fail(1, "empty plane");
    if (n_rows >= 0xFFFFFFF0ull) return fail(1, "n_rows exceeds the 32-bit row index");
    u32 SB = 0;
    for (u32 f = 0; f < n_features; ++f) {
        if (n_bins[f] < 1 || n_bins[f] > 256) return fail(1, "n_bins must be in [1, 256] (u8 bins)");
        SB = std::max<u32>(SB, n_bins[f]);
    }
    HIP_TRY(hipSetDevice(c->device));
    hipStreamSynchronize(c->stream);
    free_fit(c);
    void *ptrs[] = {c->plane0, c->work[0], c->work[1], c->d_nbins, c->label, c->score, c->samp_rows, c->samp_cnt,
                    c->fmask, c->weight};
    for (void *p : ptrs)
        if (p) hipFree(p);
    c->weight = nullptr;
    c->plane0 = c->work[0] = c->work[1] = nullptr;
    c->samp_rows = c->samp_cnt = nullptr;
    c->samp_kind = c->samp_n = 0;
    c->fmask = nullptr;
    c->fsamp_kind = 0;
    c->gc.fmask = nullptr;
    c->d_nbins = nullptr;
    c->label = c->score = nullptr;
    c->have_labels = false;
    c->max_stale = true;
    c->N = n_rows;
    c->F = n_features;
    c->Fs = round_up4(n_features);
    c->SB = SB;
    c->n_bins.assign(n_bins, n_bins + n_features);
    size_t const bytes = static_cast<size_t>(n_rows) * c->Fs;
    HIP_TRY(hipMalloc(&c->plane0, bytes + kTailPadBytes));
    HIP_TRY(hipMalloc(&c->work[0], bytes + kTailPadBytes));
    HIP_TRY(hipMalloc(&c->work[1], bytes + kTailPadBytes));
    HIP_TRY(hipMemset(c->plane0, 0, bytes + kTailPadBytes));
    HIP_TRY(dalloc(&c->d_nbins, n_features));
    HIP_TRY(hipMemcpy(c->d_nbins, n_bins, n_features * sizeof(u16), hipMemcpyHostToDevice));
    HIP_TRY(dalloc(&c->label, n_rows));
    HIP_TRY(dalloc(&c->score, n_rows));
    HIP_TRY(hipMemset(c->label, 0, n_rows * sizeof(float)));
    HIP_TRY(hipMemset(c->score, 0, n_rows * sizeof(float)));
    return 0;
}

int bh_upload_bins(bh_context *c, const uint8_t *bins, uint64_t n_rows, uint32_t n_features, uint64_t row_stride,
                   const uint16_t *n_bins)
{
    if (!c) return fail(1, "null context");
    if (row_stride < n_features) return fail(1, "row_stride < n_features");
    if (int r = alloc_plane(c, n_rows, n_features, n_bins)) return r;
    size_t const bytes = static_cast<size_t>(n_rows) * c->Fs;
    if (row_stride == c->Fs) {
        HIP_TRY(hipMemcpy(c->plane0, bins, bytes, hipMemcpyHostToDevice));
    } else {
        HIP_TRY(hipMemcpy2D(c->plane0, c->Fs, bins, row_stride, n_features, n_rows, hipMemcpyHostToDevice));
    }
    return 0;
}

// What every ingest arm shares: the cut tables are validated (ascending per feature is
// the caller's contract, as for BinMapper cuts), the plane is (re)allocated and the tables
// are uploaded.  Rows are chunked so that staging stays near kIngestChunkBytes.
constexpr u64 kIngestChunkBytes = u64{64} << 20;
static int begin_ingest(bh_context *c, uint64_t n_rows, uint32_t n_features, const float *cuts,
                        const uint32_t *cut_offsets, DevScratch<float> &d_cuts, DevScratch<u32> &d_off)
{
    if (!c) return fail(1, "null context");
    if (n_features == 0) return fail(1, "empty plane");
    if (!cuts || !cut_offsets) return fail(1, "null cut tables");
    if (cut_offsets[0] != 0) return fail(1, "cut_offsets[0] must be 0");
    std::vector<u16> nb(n_features);
    for (u32 f = 0; f < n_features; ++f) {
        if (cut_offsets[f + 1] <= cut_offsets[f]) return fail(1, "every feature needs at least one cut");
        u32 const k = cut_offsets[f + 1] - cut_offsets[f];
        if (k > 256) return fail(1, "more than 256 cuts for a feature (u8 bins)");
        nb[f] = static_cast<u16>(k);
    }
    if (int r = alloc_plane(c, n_rows, n_features, nb.data())) return r;
    u32 const n_cuts = cut_offsets[n_features];
    HIP_TRY(dalloc(&d_cuts.p, n_cuts));
    HIP_TRY(dalloc(&d_off.p, n_features + 1));
    HIP_TRY(hipMemcpy(d_cuts.p, cuts, n_cuts * sizeof(float), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(d_off.p, cut_offsets, (n_features + 1) * sizeof(u32), hipMemcpyHostToDevice));
    return 0;
}

// X_dev must be device memory of `device` (the reference's DeviceMatrix is a cudaMalloc'd
// buffer); anything else is refused before a kernel could read it.
static int check_device_matrix(const float *X_dev, int device)
{
    hipPointerAttribute_t a{};
    hipError_t const e = hipPointerGetAttributes(&a, X_dev);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return fail(1, "X_dev is not device memory");
    }
    if (a.type != hipMemoryTypeDevice || a.device != device) return fail(1, "X_dev is not device memory of the device");
    return 0;
}

int bh_upload_features(bh_context *c, const float *X, uint64_t n_rows, uint32_t n_features, const float *cuts,
                       const uint32_t *cut_offsets)
{
    if (!X) return fail(1, "null matrix");
    DevScratch<float> d_cuts, d_x;
    DevScratch<u32> d_off;
    if (int r = begin_ingest(c, n_rows, n_features, cuts, cut_offsets, d_cuts, d_off)) return r;
    u64 const chunk = std::min<u64>(n_rows, std::max<u64>(1, (u64{256} << 20) / (u64{4} * n_features)));
    HIP_TRY(dalloc(&d_x.p, chunk * n_features));
    for (u64 r0 = 0; r0 < n_rows; r0 += chunk) {
        u64 const n = std::min(chunk, n_rows - r0);
        HIP_TRY(hipMemcpyAsync(d_x.p, X + r0 * n_features, n * n_features * sizeof(float), hipMemcpyHostToDevice,
                               c->stream));
        launch_bin_rows(c->stream, d_x.p, n, n_features, c->Fs, d_cuts.p, d_off.p, cut_offsets[n_features],
                        c->plane0 + r0 * c->Fs);
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipStreamSynchronize(c->stream));
    }
    return 0;
}

int bh_upload_feature_columns(bh_context *c, const float *const *columns, uint64_t n_rows, uint32_t n_features,
                              const float *cuts, const uint32_t *cut_offsets)
{
    if (!columns) return fail(1, "null columns");
    for (u32 f = 0; f < n_features; ++f)
        if (!columns[f]) return fail(1, "null column");
    DevScratch<float> d_cuts, d_cols;
    DevScratch<u32> d_off;
    if (int r = begin_ingest(c, n_rows, n_features, cuts, cut_offsets, d_cuts, d_off)) return r;
    // four columns (one dword of every row) per staged group
    u64 const chunk = std::min<u64>(n_rows, kIngestChunkBytes / (4u * sizeof(float)));
    HIP_TRY(dalloc(&d_cols.p, 4u * chunk));
    for (u32 f0 = 0; f0 < n_features; f0 += 4u) {
        u32 const G = std::min(4u, n_features - f0);
        for (u64 r0 = 0; r0 < n_rows; r0 += chunk) {
            u64 const n = std::min(chunk, n_rows - r0);
            for (u32 g = 0; g < G; ++g)
                HIP_TRY(hipMemcpyAsync(d_cols.p + g * chunk, columns[f0 + g] + r0, n * sizeof(float),
                                       hipMemcpyHostToDevice, c->stream));
            launch_bin_cols4(c->stream, d_cols.p, chunk, n, G, d_cuts.p, d_off.p, f0, c->Fs, c->plane0 + r0 * c->Fs);
            HIP_TRY(hipGetLastError());
            HIP_TRY(hipStreamSynchronize(c->stream));
        }
    }
    return 0;
}

int bh_upload_features_device(bh_context *c, const float *X_dev, uint64_t n_rows, uint32_t n_features,
                              const float *cuts, const uint32_t *cut_offsets)
{
    if (!c) return fail(1, "null context");
    if (!X_dev) return fail(1, "null matrix");
    HIP_TRY(hipSetDevice(c->device));
    if (int r = check_device_matrix(X_dev, c->device)) return r;
    DevScratch<float> d_cuts;
    DevScratch<u32> d_off;
    if (int r = begin_ingest(c, n_rows, n_features, cuts, cut_offsets, d_cuts, d_off)) return r;
    launch_bin_rows(c->stream, X_dev, n_rows, n_features, c->Fs, d_cuts.p, d_off.p, cut_offsets[n_features],
                    c->plane0);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(c->stream));
    return 0;
}

int bh_gather_rows(int device, const float *X_dev, uint64_t n_rows, uint32_t n_features, const uint32_t *rows,
                   uint64_t n_sample, float *out)
{
    if (!X_dev || !out) return fail(1, "null matrix or output");
    if (n_rows == 0 || n_features == 0) return fail(1, "empty matrix");
    HIP_TRY(hipSetDevice(device));
    if (int r = check_device_matrix(X_dev, device)) return r;
    if (n_sample == 0) { // the whole matrix
        HIP_TRY(hipMemcpy(out, X_dev, n_rows * n_features * sizeof(float), hipMemcpyDeviceToHost));
        return 0;
    }
    if (!rows) return fail(1, "null sample rows");
    for (u64 i = 0; i < n_sample; ++i)
        if (rows[i] >= n_rows) return fail(1, "sample row out of range");
    u64 const chunk = std::min<u64>(n_sample, std::max<u64>(1, kIngestChunkBytes / (u64{4} * n_features)));
    DevScratch<u32> d_rows;
    DevScratch<float> d_out;
    HIP_TRY(dalloc(&d_rows.p, chunk));
    HIP_TRY(dalloc(&d_out.p, chunk * n_features));
    hipStream_t s = nullptr;
    HIP_TRY(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    hipError_t e = hipSuccess;
    for (u64 i0 = 0; e == hipSuccess && i0 < n_sample; i0 += chunk) {
        u64 const n = std::min(chunk, n_sample - i0);
        e = hipMemcpyAsync(d_rows.p, rows + i0, n * sizeof(u32), hipMemcpyHostToDevice, s);
        if (e != hipSuccess) break;
        launch_gather_f32_rows(s, X_dev, d_rows.p, n, n_features, d_out.p);
        e = hipGetLastError();
        if (e == hipSuccess)
            e = hipMemcpyAsync(out + i0 * n_features, d_out.p, n * n_features * sizeof(float), hipMemcpyDeviceToHost,
                               s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
    }
    (void)hipStreamDestroy(s);
    if (e != hipSuccess) return fail(-1, std::string("hip: ") + hipGetErrorString(e));
    return 0;
}

int bh_download_bins(bh_context *c, uint8_t *out, uint64_t out_stride)
{
    if (!c || !c->plane0) return fail(1, "no bins uploaded");
    if (out_stride < c->F) return fail(1, "out_stride < n_features");
    HIP_TRY(hipSetDevice(c->device));
    HIP_TRY(hipStreamSynchronize(c->stream));
    HIP_TRY(hipMemcpy2D(out, out_stride, c->plane0, c->Fs, c->F, c->N, hipMemcpyDeviceToHost));
    return 0;
}

int bh_upload_labels(bh_context *c, const float *labels)
{
    if (!c || !c->plane0) return fail(1, "no bins uploaded");
    HIP_TRY(hipSetDevice(c->device));
    HIP_TRY(hipMemcpy(c->label, labels, c->N * sizeof(float), hipMemcpyHostToDevice));
    c->have_labels = true;
    c->max_stale = true;
    return 0;
}

static void drop_graph(bh_context *c);

int bh_upload_weights(bh_context *c, const float *weights)
{
    if (!c || !c->plane0) return fail(1, "no bins uploaded");
    HIP_TRY(hipSetDevice(c->device));
    HIP_TRY(hipStreamSynchronize(c->stream));
    if (weights) {
        for (u64 r = 0; r < c->N; ++r)
            if (!(weights[r] >= 0.0F) || std::isinf(weights[r]))
                return fail(1, "weights must be finite and non-negative");
        if (!c->weight) HIP_TRY(dalloc(&c->weight, c->N));
        HIP_TRY(hipMemcpy(c->weight, weights, c->N * sizeof(float), hipMemcpyHostToDevice));
    } else if (c->weight) {
        HIP_TRY(hipFree(c->weight));
        c->weight = nullptr;
    }
    drop_graph(c); // the gradient kernels' arguments change
    c->max_stale = true;
    return 0;
}

int bh_configure(bh_context *c, const bh_fit_config *cfg)
{
    if (!c || !c->plane0) return fail(1, "no bins uploaded");
    if (cfg->grower < 0 || cfg->grower > 2) return fail(1, "unknown grower");
    if (cfg->objective != BH_MSE && cfg->objective != BH_LOGLOSS && cfg->objective != BH_POISSON)
        return fail(1, "unknown objective");
    if (cfg->max_depth < 1 || cfg->max_depth > 12) return fail(1, "max_depth must be in [1, 12]");
    if (cfg->min_data_in_leaf < 0) return fail(1, "min_data_in_leaf < 0");
    if (cfg->grower == BH_LEAFWISE && (cfg->max_leaves < 2 || cfg->max_leaves > 1024))
        return fail(1, "max_leaves must be in [2, 1024] for the leafwise grower");
    if (cfg->grower == BH_LEVELWISE && cfg->max_depth > 11) return fail(1, "levelwise max_depth must be <= 11");
    if (cfg->grower == BH_LEVELWISE) {
        // refused before anything is freed, so the previous fit stays as it was
        // (src/grower_impl.hpp:583-595)
        bool mono = false;
        for (int8_t v : c->mono_host) mono = mono || v != 0;
        if (mono) return fail(1, "monotone_constraints are not supported by the levelwise grower");
        if (!c->groups_host.empty())
            return fail(1, "interaction_constraints are not supported by the levelwise grower");
    }
    HIP_TRY(hipSetDevice(c->device));
    hipStreamSynchronize(c->stream);
    free_fit(c);
    c->fc = *cfg;
    return alloc_fit(c);
}


int bh_set_scores(bh_context *c, const float *scores)
{
    if (int r = check_ready(c)) return r;
    HIP_TRY(hipSetDevice(c->device));
    HIP_TRY(hipMemcpy(c->score, scores, c->N * sizeof(float), hipMemcpyHostToDevice));
    c->max_stale = true;
    // A new fit starts: iteration counter and tree store reset; a validation set armed
    // for the previous fit is disarmed (its scores track that fit).
    TreeCtl z{};
    HIP_TRY(hipMemcpy(c->ctl, &z, sizeof(TreeCtl), hipMemcpyHostToDevice));
    c->trees_done = 0;
    if (c->have_valid) {
        c->have_valid = false;
        c->es_done = false;
        drop_graph(c);
    }
    return 0;
}

int bh_set_validation(bh_context *c, const uint8_t *bins, uint64_t n_rows, uint64_t row_stride, const float *labels,
                      float init_score, uint32_t early_stopping_rounds)
{
    if (int r = check_ready(c)) return r;
    if (n_rows == 0) return fail(1, "empty validation set");
    if (n_rows > 0xFFFFFFFFull) return fail(1, "validation set larger than 2^32 rows");
    if (row_stride < c->F) return fail(1, "row_stride < n_features");
    HIP_TRY(hipSetDevice(c->device));
    HIP_TRY(hipStreamSynchronize(c->stream));
    void *old[] = {c->vplane, c->vlabel, c->vscore, c->vhist, c->vpartial};
    for (void *p : old)
        if (p) (void)hipFree(p);
    c->vplane = nullptr;
    c->vlabel = c->vscore = c->vhist = nullptr;
    c->vpartial = nullptr;
    c->have_valid = false;
    size_t const bytes = static_cast<size_t>(n_rows) * c->Fs;
    HIP_TRY(dalloc(&c->vplane, bytes + kTailPadBytes));
    HIP_TRY(hipMemset(c->vplane, 0, bytes + kTailPadBytes));
    HIP_TRY(hipMemcpy2D(c->vplane, c->Fs, bins, row_stride, c->F, n_rows, hipMemcpyHostToDevice));
    HIP_TRY(dalloc(&c->vlabel, n_rows));
    HIP_TRY(hipMemcpy(c->vlabel, labels, n_rows * sizeof(float), hipMemcpyHostToDevice));
    std::vector<float> seed(n_rows, init_score);
    HIP_TRY(dalloc(&c->vscore, n_rows));
    HIP_TRY(hipMemcpy(c->vscore, seed.data(), n_rows * sizeof(float), hipMemcpyHostToDevice));
    c->vhist_cap = 1u << 16;
    HIP_TRY(dalloc(&c->vhist, c->vhist_cap));
    HIP_TRY(dalloc(&c->vpartial, 1024));
    // early-stopping state starts empty; the rounds count from here (warm start)
    TreeCtl t;
    HIP_TRY(hipMemcpy(&t, c->ctl, sizeof(TreeCtl), hipMemcpyDeviceToHost));
    t.es_n = t.es_best_iter = t.es_stopped = 0;
    t.es_best = 0.0F;
    HIP_TRY(hipMemcpy(c->ctl, &t, sizeof(TreeCtl), hipMemcpyHostToDevice));
    c->vN = n_rows;
    c->es_rounds = early_stopping_rounds;
    c->es_base = c->trees_done;
    c->es_done = false;
    c->have_valid = true;
    drop_graph(c); // the iteration graph gains the validation launches
    return 0;
}

int bh_get_eval(bh_context *c, float *history, uint32_t cap, uint32_t *n, uint32_t *best_iter, int32_t *stopped)
{
    if (!c) return fail(1, "null context");
    if (!c->have_valid) return fail(1, "no validation set armed");
    HIP_TRY(hipSetDevice(c->device));
    HIP_TRY(hipStreamSynchronize(c->stream));
    TreeCtl t;
    HIP_TRY(hipMemcpy(&t, c->ctl, sizeof(TreeCtl), hipMemcpyDeviceToHost));
    u32 const k = std::min(t.es_n, c->vhist_cap);
    if (history && cap) HIP_TRY(hipMemcpy(history, c->vhist, std::min(k, cap) * sizeof(float), hipMemcpyDeviceToHost));
    if (n) *n = k;
    if (best_iter) *best_iter = t.es_best_iter;
    if (stopped) *stopped = static_cast<int32_t>(t.es_stopped);
    return 0;
}

int bh_set_row_subset(bh_context *c, const uint32_t *rows, uint64_t n)
{
    if (!c) return fail(1, "null context");
    if (!c->plane0) return fail(1, "no bins uploaded");
    HIP_TRY(hipSetDevice(c->device));
    HIP_TRY(hipStreamSynchronize(c->stream));
    if (c->samp_rows) (void)hipFree(c->samp_rows);
    c->samp_rows = nullptr;
    c->samp_kind = c->samp_n = 0;
    drop_graph(c); // the iteration graph gains or loses the gather
    if (!rows || n == 0 || n == c->N) {
        if (rows && n == c->N)
            for (u64 i = 0; i < n; ++i)
                if (rows[i] != i) return fail(1, "row subset must be ascending, unique and below n_rows");
        return 0; // every row: the plain root
    }
    if (n > c->N) return fail(1, "row subset larger than the plane");
    for (u64 i = 0; i < n; ++i)
        if (rows[i] >= c->N || (i > 0 && rows[i] <= rows[i - 1]))
            return fail(1, "row subset must be ascending, unique and below n_rows");
    HIP_TRY(dalloc(&c->samp_rows, n));
    HIP_TRY(hipMemcpy(c->samp_rows, rows, n * sizeof(u32), hipMemcpyHostToDevice));
    c->samp_n = static_cast<u32>(n);
    c->samp_kind = 1;
    return 0;
}

int bh_set_bernoulli(bh_context *c, float subsample, uint64_t seed)
{
    if (!c) return fail(1, "null context");
    if (!c->plane0) return fail(1, "no bins uploaded");
    if (!(subsample > 0.0F)) return fail(1, "subsample must be > 0");
    HIP_TRY(hipSetDevice(c->device));
    HIP_TRY(hipStreamSynchronize(c->stream));
    if (c->samp_rows) (void)hipFree(c->samp_rows);
    c->samp_rows = nullptr;
    c->samp_kind = c->samp_n = 0;
    drop_graph(c);
    if (subsample >= 1.0F) return 0; // the reference keeps every row for p >= 1
    if (!c->samp_cnt) HIP_TRY(dalloc(&c->samp_cnt, c->N / kPartChunk + 2));
    double const t = std::floor(static_cast<double>(subsample) * 4294967296.0);
    c->samp_thr = static_cast<u32>(std::min(t, 4294967295.0));
    c->samp_seed = seed;
    c->samp_kind = 2;
    return 0;
}

int bh_set_monotone(bh_context *c, const int8_t *dirs, uint32_t n)
{
    if (!c) return fail(1, "null context");
    if (dirs == nullptr && n != 0) return fail(1, "null monotone directions");
    HIP_TRY(hipSetDevice(c->device));
    HIP_TRY(hipStreamSynchronize(c->stream));
    std::vector<int8_t> prev = std::move(c->mono_host);
    c->mono_host.clear();
    for (u32 f = 0; f < n; ++f) c->mono_host.push_back(dirs[f] > 0 ? 1 : (dirs[f] < 0 ? -1 : 0));
    drop_graph(c); // the finders' configuration changes
    if (int r = arm_monotone(c)) {
        c->mono_host = std::move(prev);
        (void)arm_monotone(c);
        return r;
    }
    return 0;
}

int bh_set_interaction(bh_context *c, const uint32_t *members, const uint32_t *group_sizes, uint32_t n_groups)
{
    if (!c) return fail(1, "null context");
    if (n_groups > kMaxGroups) return fail(1, "at most 256 interaction groups");
    if (n_groups && (!members || !group_sizes)) return fail(1, "null interaction groups");
    HIP_TRY(hipSetDevice(c->device));
    HIP_TRY(hipStreamSynchronize(c->stream));
    std::vector<std::vector<u32>> prev = std::move(c->groups_host);
    c->groups_host.clear();
    size_t k = 0;
    for (u32 g = 0; g < n_groups; ++g) {
        std::vector<u32> grp(members + k, members + k + group_sizes[g]);
        k += group_sizes[g];
        if (!grp.empty()) c->groups_host.push_back(std::move(grp)); // empty entries are dropped
    }
    drop_graph(c);
    if (int r = arm_interaction(c)) {
        c->groups_host = std::move(prev);
        (void)arm_interaction(c);
        return r;
    }
    return 0;
}

// Selection state shared by both feature setters: clears it, dropping the graph (the
// finders' configuration and the draw launch change).
int reset_feature_sampling(bh_context *c)
{
    HIP_TRY(hipSetDevice(c->device));
    HIP_TRY(hipStreamSynchronize(c->stream));
    c->fsamp_kind = 0;
    c->gc.fmask = nullptr;
    drop_graph(c);
    if (!c->fmask) HIP_TRY(dalloc(&c->fmask, c->F));
    return 0;
}

int bh_set_feature_mask(bh_context *c, const uint8_t *mask, uint32_t n_features)
{
    if (!c) return fail(1, "null context");
    if (!c->plane0) return fail(1, "no bins uploaded");
    if (mask && n_features != c->F) return fail(1, "feature mask length differs from n_features");
    u32 n_sel = 0;
    if (mask)
        for (u32 f = 0; f < n_features; ++f) n_sel += mask[f] ? 1u : 0u;
    if (mask && n_sel == 0) return fail(1, "the feature mask selects no features");
    if (int r = reset_feature_sampling(c)) return r;
    if (!mask || n_sel == c->F) return 0; // every feature
    std::vector<u8> m(mask, mask + n_features);
    for (u8 &v : m) v = v ? 1 : 0;
    HIP_TRY(hipMemcpy(c->fmask, m.data(), c->F, hipMemcpyHostToDevice));
    c->fsamp_kind = 1;
    c->gc.fmask = c->fmask;
    return 0;
}

int bh_set_feature_fraction(bh_context *c, float fraction, uint64_t seed)
{
    if (!c) return fail(1, "null context");
    if (!c->plane0) return fail(1, "no bins uploaded");
    if (std::isnan(fraction)) return fail(1, "feature fraction is NaN");
    if (int r = reset_feature_sampling(c)) return r;
    if (fraction >= 1.0F || c->F <= 1) return 0; // sample_features returns every feature
    double const k = std::ceil(static_cast<double>(fraction * static_cast<float>(c->F)));
    c->fsamp_k = static_cast<u32>(std::min(static_cast<double>(c->F), std::max(1.0, k)));
    c->fsamp_seed = seed;
    c->fsamp_kind = 2;
    c->gc.fmask = c->fmask;
    return 0;
}

int bh_get_scores(bh_context *c, float *scores)
{
    if (int r = check_ready(c)) return r;
    HIP_TRY(hipSetDevice(c->device));
    HIP_TRY(hipStreamSynchronize(c->stream));
    HIP_TRY(hipMemcpy(scores, c->score, c->N * sizeof(float), hipMemcpyDeviceToHost));
    return 0;
}

int bh_set_timing(bh_context *c, int32_t enable)
{
    if (!c) return fail(1, "null context");
    c->timing = enable != 0;
    return 0;
}

int bh_get_timing(bh_context *c, float *ms, uint32_t n)
{
    if (!c) return fail(1, "null context");
    for (u32 i = 0; i < n; ++i) ms[i] = i < c->iter_ms.size() ? c->iter_ms[i] : 0.0F;
    return 0;
}

int bh_train(bh_context *c, uint32_t n_iters)
{
    if (int r = check_ready(c)) return r;
    if (!c->have_labels) return fail(1, "no labels uploaded");
    if (c->es_done) return 0; // the fit stopped early: its model is final
    HIP_TRY(hipSetDevice(c->device));
    if (int r = ensure_tree_out(c, n_iters)) return r;
    c->iter_ms.clear();
    bool const graph = c->fc.use_graph && !c->fc.host_sync;
    prime_gradients(c);
    if (graph && !c->graph_exec) {
        HIP_TRY(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        enqueue_iteration(c);
        hipError_t const e = hipStreamEndCapture(c->stream, &c->graph);
        if (e != hipSuccess) return fail(-1, std::string("hip: capture: ") + hipGetErrorString(e));
        HIP_TRY(hipGraphInstantiate(&c->graph_exec, c->graph, nullptr, nullptr, 0));
    }
    // With early stopping the device records the rule's verdict every round; the host
    // looks at it every kEsCheck rounds, so at most kEsCheck - 1 rounds past the stop
    // are grown and dropped.
    constexpr u32 kEsCheck = 8;
    bool const es = c->have_valid && c->es_rounds > 0;
    u32 const step = es ? kEsCheck : std::max(1u, n_iters);
    for (u32 done = 0; done < n_iters; done += step) {
        u32 const k = std::min(step, n_iters - done);
        for (u32 it = 0; it < k; ++it) {
            if (c->timing) HIP_TRY(hipEventRecord(c->ev0, c->stream));
            if (graph) HIP_TRY(hipGraphLaunch(c->graph_exec, c->stream));
            else enqueue_iteration(c);
            if (c->timing) {
                HIP_TRY(hipEventRecord(c->ev1, c->stream));
                HIP_TRY(hipEventSynchronize(c->ev1));
                float ms = 0.0F;
                hipEventElapsedTime(&ms, c->ev0, c->ev1);
                c->iter_ms.push_back(ms);
            }
        }
        if (es) {
            HIP_TRY(hipStreamSynchronize(c->stream));
            TreeCtl t;
            HIP_TRY(hipMemcpy(&t, c->ctl, sizeof(TreeCtl), hipMemcpyDeviceToHost));
            if (t.es_stopped) {
                if (int r = check_partition(c)) return r;
                c->trees_done = c->es_base + t.es_best_iter + 1; // Booster::truncate(best + 1)
                c->es_done = true;
                return 0;
            }
        }
        c->trees_done += k;
    }
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(c->stream));
    if (int r = check_partition(c)) return r;
    return 0;
}

static void drop_graph(bh_context *c)
{
    if (c->graph_exec) (void)hipGraphExecDestroy(c->graph_exec);
    if (c->graph) (void)hipGraphDestroy(c->graph);
    c->graph_exec = nullptr;
    c->graph = nullptr;
}

static int validate_tuning(bh_tuning const *t)
{
    if (t->scatter_lds_bytes < 4096u || t->scatter_lds_bytes > 162816u)
        return fail(1, "scatter_lds_bytes must be in [4096, 162816]");
    if (t->find_blocks < 1u || t->find_blocks > kFindMaxBlocks) return fail(1, "find_blocks must be in [1, 8]");
    if (t->hist_threads != 256u && t->hist_threads != 512u && t->hist_threads != 1024u)
        return fail(1, "hist_threads must be 256, 512 or 1024");
    if (t->part_grid < 64u || t->part_grid > 65536u) return fail(1, "part_grid must be in [64, 65536]");
    if (t->hist_layout > 5u) return fail(1, "hist_layout must be 0 .. 5");
    if (t->pair_levels > 1u) return fail(1, "pair_levels must be 0 or 1");
    if (t->hist_group < 1u || t->hist_group > kHistGroupMax) return fail(1, "hist_group must be in [1, 4]");
    return 0;
}

int bh_set_tuning(bh_context *c, const bh_tuning *t)
{
    if (!c) return fail(1, "null context");
    if (int r = validate_tuning(t)) return r;
    HIP_TRY(hipSetDevice(c->device));
    HIP_TRY(hipStreamSynchronize(c->stream));
    c->tune = Tuning{t->scatter_lds_bytes, t->find_blocks, t->hist_threads, t->part_grid, t->hist_layout,
                     t->pair_levels, t->hist_group};
    drop_graph(c); // the iteration graph is re-captured with the new shapes
    return 0;
}

int bh_get_tuning(bh_context *c, bh_tuning *t)
{
    if (!c) return fail(1, "null context");
    *t = bh_tuning{c->tune.scatter_lds, c->tune.find_blocks, c->tune.hist_threads, c->tune.part_grid,
                   c->tune.hist_layout, c->tune.pair_levels, c->tune.hist_group};
    return 0;
}

// Coordinate descent over the launch-shape knobs on the resident data and the
// configured fit: each trial restores the scores and the tree state, runs one warm
// iteration (graph capture) and times `iters` iterations with hipEvents.  The model
// state after the call is the one before it.
int bh_autotune(bh_context *c, uint32_t iters, bh_tuning *chosen, float *ms_per_iter)
{
    if (int r = check_ready(c)) return r;
    if (!c->have_labels) return fail(1, "no labels uploaded");
    if (iters == 0) return fail(1, "iters must be positive");
    if (c->have_valid) return fail(1, "autotune before arming a validation set (its scores would move)");
    HIP_TRY(hipSetDevice(c->device));
    HIP_TRY(hipStreamSynchronize(c->stream));
    float *saved = nullptr;
    HIP_TRY(dalloc(&saved, c->N));
    TreeCtl saved_ctl;
    hipError_t e = hipMemcpy(saved, c->score, c->N * sizeof(float), hipMemcpyDeviceToDevice);
    if (e == hipSuccess) e = hipMemcpy(&saved_ctl, c->ctl, sizeof(TreeCtl), hipMemcpyDeviceToHost);
    if (e != hipSuccess) {
        (void)hipFree(saved);
        return fail(-1, std::string("hip: ") + hipGetErrorString(e));
    }
    u32 const saved_trees = c->trees_done;
    bool const saved_timing = c->timing;
    c->timing = false;
    auto restore = [&]() -> int {
        HIP_TRY(hipMemcpy(c->score, saved, c->N * sizeof(float), hipMemcpyDeviceToDevice));
        HIP_TRY(hipMemcpy(c->ctl, &saved_ctl, sizeof(TreeCtl), hipMemcpyHostToDevice));
        c->trees_done = saved_trees;
        c->max_stale = true; // the restored scores' gradient maxima are not in maxbits
        return 0;
    };
    auto trial = [&](Tuning const &t, float *ms) -> int {
        c->tune = t;
        drop_graph(c);
        if (int r = restore()) return r;
        if (int r = bh_train(c, 1)) return r;
        if (int r = restore()) return r;
        HIP_TRY(hipEventRecord(c->ev0, c->stream));
        if (int r = bh_train(c, iters)) return r;
        HIP_TRY(hipEventRecord(c->ev1, c->stream));
        HIP_TRY(hipEventSynchronize(c->ev1));
        float total = 0.0F;
        HIP_TRY(hipEventElapsedTime(&total, c->ev0, c->ev1));
        *ms = total / static_cast<float>(iters);
        return 0;
    };
    Tuning best = c->tune;
    float best_ms = 0.0F;
    int rc = trial(best, &best_ms);
    static const u32 lds_c[] = {8192u, 12288u, 20480u, 32768u, 40992u, 49152u, 98304u, 147456u};
    static const u32 thr_c[] = {256u, 512u, 1024u};
    static const u32 fb_c[] = {1u, 2u, 4u, 8u};
    static const u32 pg_c[] = {512u, 1024u, 2048u, 4096u};
    static const u32 hl_c[] = {0u, 1u, 2u, 3u, 4u, 5u};
    static const u32 pl_c[] = {0u, 1u};
    static const u32 hg_c[] = {1u, 2u, 4u};
    struct Knob {
        u32 Tuning::*field;
        u32 const *cand;
        int n;
    };
    static const Knob knobs[] = {{&Tuning::scatter_lds, lds_c, 8}, {&Tuning::hist_threads, thr_c, 3},
                                 {&Tuning::find_blocks, fb_c, 4},  {&Tuning::part_grid, pg_c, 4},
                                 {&Tuning::hist_layout, hl_c, 6},  {&Tuning::pair_levels, pl_c, 2},
                                 {&Tuning::hist_group, hg_c, 3}};
    for (Knob const &kn : knobs) {
        for (int k = 0; rc == 0 && k < kn.n; ++k) {
            Tuning t = best;
            u32 &slot = t.*kn.field;
            if (slot == kn.cand[k]) continue;
            slot = kn.cand[k];
            float ms = 0.0F;
            rc = trial(t, &ms);
            if (rc == 0 && ms < best_ms) {
                best_ms = ms;
                best = t;
            }
        }
    }
    c->tune = best;
    drop_graph(c);
    int const rr = restore();
    (void)hipFree(saved);
    c->timing = saved_timing;
    if (rc) return rc;
    if (rr) return rr;
    set_tuning(c->tune);
    if (chosen)
        *chosen = bh_tuning{best.scatter_lds, best.find_blocks, best.hist_threads, best.part_grid, best.hist_layout,
                            best.pair_levels, best.hist_group};
    if (ms_per_iter) *ms_per_iter = best_ms;
    return 0;
}

int bh_max_nodes(bh_context *c, uint32_t *max_nodes)
{
    if (int r = check_ready(c)) return r;
    *max_nodes = c->max_nodes;
    return 0;
}

int bh_get_trees(bh_context *c, bh_node *nodes, uint32_t max_nodes, uint32_t *sizes, uint32_t cap_trees,
                 uint32_t *n_trees)
{
    if (int r = check_ready(c)) return r;
    if (max_nodes < c->max_nodes) return fail(1, "max_nodes smaller than the fit's tree capacity");
    HIP_TRY(hipSetDevice(c->device));
    HIP_TRY(hipStreamSynchronize(c->stream));
    u32 const n = std::min(cap_trees, c->trees_done);
    std::vector<u32> sz(n);
    if (n) HIP_TRY(hipMemcpy(sz.data(), c->tree_sizes, n * sizeof(u32), hipMemcpyDeviceToHost));
    std::vector<DevNode> buf(c->max_nodes);
    for (u32 t = 0; t < n; ++t) {
        HIP_TRY(hipMemcpy(buf.data(), c->trees_out + static_cast<size_t>(t) * c->max_nodes,
                          c->max_nodes * sizeof(DevNode), hipMemcpyDeviceToHost));
        std::memcpy(nodes + static_cast<size_t>(t) * max_nodes, buf.data(), sz[t] * sizeof(DevNode));
        sizes[t] = sz[t];
    }
    *n_trees = n;
    return 0;
}

int bh_set_gradients(bh_context *c, const float *grad, const float *hess, uint32_t iter, int8_t *gq, int16_t *hq,
                     double *scales)
{
    if (int r = check_ready(c)) return r;
    HIP_TRY(hipSetDevice(c->device));
    std::vector<float2> gh(c->N);
    for (u64 i = 0; i < c->N; ++i) gh[i] = make_float2(grad[i], hess[i]);
    c->max_stale = true; // the host gradients' maxima overwrite slot 0
    if (!c->quant) {
        HIP_TRY(hipMemcpyAsync(c->gh0f, gh.data(), c->N * sizeof(float2), hipMemcpyHostToDevice, c->stream));
        launch_fx_from(c->stream, c->gh0f, c->N, c->fx);
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipStreamSynchronize(c->stream));
        return 0;
    }
    float2 *tmp = nullptr;
    HIP_TRY(dalloc(&tmp, c->N));
    HIP_TRY(hipMemcpy(tmp, gh.data(), c->N * sizeof(float2), hipMemcpyHostToDevice));
    launch_quant_from(c->stream, tmp, c->N, c->maxbits, c->fc.seed, iter, c->gh0q, c->scales);
    HIP_TRY(hipStreamSynchronize(c->stream));
    hipFree(tmp);
    u32 mb[2];
    HIP_TRY(hipMemcpy(mb, c->maxbits, sizeof(mb), hipMemcpyDeviceToHost));
    float mg, mh;
    std::memcpy(&mg, &mb[0], 4);
    std::memcpy(&mh, &mb[1], 4);
    double const sc[2] = {static_cast<double>(mg) / 127.0, static_cast<double>(mh) / 32767.0};
    HIP_TRY(hipMemcpy(c->scales, sc, sizeof(sc), hipMemcpyHostToDevice));
    if (scales) {
        scales[0] = sc[0];
        scales[1] = sc[1];
    }
    if (gq || hq) {
        std::vector<u32> q(c->N);
        HIP_TRY(hipMemcpy(q.data(), c->gh0q, c->N * sizeof(u32), hipMemcpyDeviceToHost));
        for (u64 i = 0; i < c->N; ++i) {
            if (gq) gq[i] = static_cast<int8_t>(static_cast<int16_t>(q[i] >> 16));
            if (hq) hq[i] = static_cast<int16_t>(q[i] & 0xFFFFu);
        }
    }
    return 0;
}

int bh_get_gradients(bh_context *c, float *grad, float *hess, int8_t *gq, int16_t *hq, double *scales)
{
    if (int r = check_ready(c)) return r;
    HIP_TRY(hipSetDevice(c->device));
    HIP_TRY(hipStreamSynchronize(c->stream));
    if (!c->quant) {
        if (gq || hq || scales) return fail(1, "an FP32 fit keeps float gradients: pass grad / hess");
        if (grad || hess) {
            std::vector<float2> gh(c->N);
            HIP_TRY(hipMemcpy(gh.data(), c->gh0f, c->N * sizeof(float2), hipMemcpyDeviceToHost));
            for (u64 i = 0; i < c->N; ++i) {
                if (grad) grad[i] = gh[i].x;
                if (hess) hess[i] = gh[i].y;
            }
        }
        return 0;
    }
    if (grad || hess) return fail(1, "a quantized fit keeps integer gradients: pass gq / hq / scales");
    if (scales) HIP_TRY(hipMemcpy(scales, c->scales, 2 * sizeof(double), hipMemcpyDeviceToHost));
    if (gq || hq) {
        std::vector<u32> q(c->N);
        HIP_TRY(hipMemcpy(q.data(), c->gh0q, c->N * sizeof(u32), hipMemcpyDeviceToHost));
        for (u64 i = 0; i < c->N; ++i) {
            if (gq) gq[i] = static_cast<int8_t>(static_cast<int16_t>(q[i] >> 16));
            if (hq) hq[i] = static_cast<int16_t>(q[i] & 0xFFFFu);
        }
    }
    return 0;
}

int bh_grow_tree(bh_context *c, bh_node *nodes, uint32_t cap, uint32_t *n_nodes, uint32_t *leaf_of_row)
{
    if (int r = check_ready(c)) return r;
    HIP_TRY(hipSetDevice(c->device));
    enqueue_grow(c);
    launch_score_update(c->stream, c->plane0, c->Fs, c->N, c->tree, c->ctl, c->d_nbins, c->F, 0.0F, nullptr,
                        c->leaf_rows, c->max_nodes);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(c->stream));
    if (int r = check_partition(c)) return r;
    TreeCtl ctl;
    HIP_TRY(hipMemcpy(&ctl, c->ctl, sizeof(TreeCtl), hipMemcpyDeviceToHost));
    if (ctl.n_nodes > cap) return fail(1, "node capacity too small");
    HIP_TRY(hipMemcpy(nodes, c->tree, ctl.n_nodes * sizeof(DevNode), hipMemcpyDeviceToHost));
    *n_nodes = ctl.n_nodes;
    if (leaf_of_row) HIP_TRY(hipMemcpy(leaf_of_row, c->leaf_rows, c->N * sizeof(u32), hipMemcpyDeviceToHost));
    return 0;
}

int bh_histogram(bh_context *c, const uint32_t *rows, uint64_t n, float *out_f32, int64_t *out_i64)
{
    if (int r = check_ready(c)) return r;
    if (n > c->N) return fail(1, "more rows than the plane holds");
    HIP_TRY(hipSetDevice(c->device));
    u32 *d_rows = nullptr;
    HIP_TRY(dalloc(&d_rows, n));
    HIP_TRY(hipMemcpy(d_rows, rows, n * sizeof(u32), hipMemcpyHostToDevice));
    u32 const gh_bytes = c->quant ? 4 : 8;
    launch_gather_rows(c->stream, c->plane0, c->Fs, gh_src(c, 2), gh_bytes, d_rows, n, c->work[0],
                       const_cast<void *>(gh_src(c, 0)), 0);
    // One node over [0, n) of work[0]: units + finish op, as the level loop emits them.
    u32 const units = static_cast<u32>((n + kHistChunk - 1) / kHistChunk);
    std::vector<HistChunk> hc(units);
    for (u32 k = 0; k < units; ++k) {
        hc[k] = HistChunk{};
        hc[k].slot = 0;
        hc[k].src = 0;
        hc[k].begin = k * kHistChunk;
        hc[k].end = static_cast<u32>(std::min<u64>(n, static_cast<u64>(k + 1) * kHistChunk));
        hc[k].part = units > 1 ? k : kNone;
    }
    FinishOp fo{};
    fo.small_slot = 0;
    fo.large_slot = kNone;
    fo.part_count = units > 1 ? units : 0;
    fo.empty = n == 0;
    TreeCtl ctl{};
    ctl.n_hist_chunks = units;
    ctl.n_finish = 1;
    if (units) HIP_TRY(hipMemcpy(c->hch, hc.data(), units * sizeof(HistChunk), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(c->fops, &fo, sizeof(fo), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(c->ctl, &ctl, sizeof(ctl), hipMemcpyHostToDevice));
    enqueue_build(c, c->hist[0], c->hist[0], 1, std::max(1u, std::min(units, 4096u)));
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(c->stream));
    hipFree(d_rows);
    size_t const nc = c->node_cells();
    if (c->quant) {
        if (!out_i64) return fail(1, "quantized fit: pass out_i64");
        HIP_TRY(hipMemcpy(out_i64, c->hist[0], nc * 2 * sizeof(int64_t), hipMemcpyDeviceToHost));
    } else {
        if (!out_f32) return fail(1, "FP32 fit: pass out_f32");
        HIP_TRY(hipMemcpy(out_f32, c->hist[0], nc * 2 * sizeof(float), hipMemcpyDeviceToHost));
    }
    return 0;
}

int bh_histogram_noalloc(bh_context *c, const uint32_t *d_rows, uint64_t n, void *d_out, void *hip_stream)
{
    if (int r = check_ready(c)) return r;
    if (n > c->N) return fail(1, "more rows than the plane holds");
    if (n > 0 && !d_rows) return fail(1, "null row list");
    if (!d_out) return fail(1, "null output");
    u32 const units = static_cast<u32>((n + kHistChunk - 1) / kHistChunk);
    size_t const part_bytes = static_cast<size_t>(units > 1u ? units : 0u) * c->node_cells() * c->part_bytes();
    if (units > kNoAllocMaxUnits || kNoAllocPartOff + part_bytes > kNoAllocScratchBytes)
        return fail(1, "bh_histogram_noalloc: " + std::to_string(n) + " rows need " +
                           std::to_string(kNoAllocPartOff + part_bytes) + " bytes of scratch, more than its " +
                           std::to_string(kNoAllocScratchBytes) + " (at most " +
                           std::to_string(std::min<size_t>(kNoAllocMaxUnits,
                                                           (kNoAllocScratchBytes - kNoAllocPartOff) /
                                                               (c->node_cells() * c->part_bytes()))) +
                           " units of 8192 rows at this plane's features and bins)");
    HIP_TRY(hipSetDevice(c->device));
    unsigned char *const sp = g_scratch_ptr[c->device];
    if (!sp) return fail(-1, "scratch not resolved for this device");
    hipStream_t const st = hip_stream ? static_cast<hipStream_t>(hip_stream) : c->stream;
    TreeCtl *const ctl = reinterpret_cast<TreeCtl *>(sp + kNoAllocCtlOff);
    FinishOp *const fop = reinterpret_cast<FinishOp *>(sp + kNoAllocFinishOff);
    HistChunk *const chunks = reinterpret_cast<HistChunk *>(sp + kNoAllocChunkOff);
    void *const partials = sp + kNoAllocPartOff;
    u32 const gh_bytes = c->quant ? 4 : 8;
    if (n > 0)
        launch_gather_rows(st, c->plane0, c->Fs, gh_src(c, 2), gh_bytes, d_rows, n, c->work[0],
                           const_cast<void *>(gh_src(c, 0)), 0);
    hipLaunchKernelGGL(k_noalloc_plan, dim3(1), dim3(256), 0, st, ctl, fop, chunks, units, static_cast<u32>(n));
    HIP_TRY(hipGetLastError());
    HIP_TRY(launch_hist(st, c->quant, chunks, ctl, c->work[0], c->work[1], c->plane0, gh_src(c, 0), gh_src(c, 1),
                        gh_src(c, 2), c->Fs, c->F, c->SB, c->hist_fsl(), c->hist_waves(),
                        std::max(1u, std::min(units, 4096u)), partials, d_out, c->fx));
    launch_finish(st, c->quant, fop, ctl, partials, d_out, d_out, c->node_cells(), 1, nullptr, nullptr, false,
                  hist_transpose_F(c->quant, c->Fs, c->F));
    HIP_TRY(hipGetLastError());
    return 0;
}

int bh_find_split(bh_context *c, const float *hist_f32, const int64_t *hist_i64, uint64_t n_rows, uint32_t depth,
                  int32_t tot_given, double tot_g, double tot_h, const double *scales, bh_split *out)
{
    if (int r = check_ready(c)) return r;
    HIP_TRY(hipSetDevice(c->device));
    size_t const nc = c->node_cells();
    if (c->quant) {
        if (!hist_i64) return fail(1, "quantized fit: pass hist_i64");
        HIP_TRY(hipMemcpy(c->hist[0], hist_i64, nc * 2 * sizeof(int64_t), hipMemcpyHostToDevice));
        if (scales) HIP_TRY(hipMemcpy(c->scales, scales, 2 * sizeof(double), hipMemcpyHostToDevice));
    } else {
        if (!hist_f32) return fail(1, "FP32 fit: pass hist_f32");
        HIP_TRY(hipMemcpy(c->hist[0], hist_f32, nc * 2 * sizeof(float), hipMemcpyHostToDevice));
    }
    NodeInfo nd{};
    nd.cnt = static_cast<u32>(n_rows);
    nd.depth = depth;
    nd.slot = 0;
    if (tot_given) {
        nd.totals_ready = 1;
        nd.tg = static_cast<float>(tot_g);
        nd.th = static_cast<float>(tot_h);
        nd.itg = static_cast<i64>(tot_g);
        nd.ith = static_cast<i64>(tot_h);
    }
    HIP_TRY(hipMemcpy(c->front[0], &nd, sizeof(nd), hipMemcpyHostToDevice));
    launch_find(c->stream, c->quant, c->front[0], c->ctl, c->hist[0], c->d_nbins, c->F, c->SB, c->gc, c->scales,
                c->splits, 1, 0, 1, c->fscr);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(c->stream));
    SplitRec rec;
    HIP_TRY(hipMemcpy(&rec, c->splits, sizeof(rec), hipMemcpyDeviceToHost));
    static_assert(sizeof(SplitRec) == sizeof(bh_split), "bh_split layout");
    std::memcpy(out, &rec, sizeof(rec));
    return 0;
}

int bh_partition(bh_context *c, const uint32_t *rows, uint64_t n, uint32_t feature, uint32_t bin,
                 uint32_t default_left, uint32_t *out_rows, uint64_t *n_left)
{
    if (int r = check_ready(c)) return r;
    if (n == 0 || n > c->N) return fail(1, "row count out of range");
    if (feature >= c->F) return fail(1, "feature out of range");
    HIP_TRY(hipSetDevice(c->device));
    u32 *d_rows = nullptr;
    HIP_TRY(dalloc(&d_rows, n));
    HIP_TRY(hipMemcpy(d_rows, rows, n * sizeof(u32), hipMemcpyHostToDevice));
    u32 const gh_bytes = c->quant ? 4 : 8;
    // Rows into work[0] with their ids as the travelling gradient pair.
    launch_gather_rows(c->stream, c->plane0, c->Fs, gh_src(c, 2), gh_bytes, d_rows, n, c->work[0],
                       const_cast<void *>(gh_src(c, 0)), 1);
    NodeInfo nd{};
    nd.id = 0;
    nd.cnt = static_cast<u32>(n);
    nd.src = 0;
    nd.totals_ready = 1;
    SplitRec s{};
    s.valid = 1;
    s.feature = feature;
    s.bin = bin;
    s.default_left = default_left;
    TreeCtl ctl{};
    ctl.n_front = 1;
    ctl.n_nodes = 1;
    DevNode root{};
    root.feature = -1;
    HIP_TRY(hipMemcpy(c->front[0], &nd, sizeof(nd), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(c->splits, &s, sizeof(s), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(c->ctl, &ctl, sizeof(ctl), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(c->tree, &root, sizeof(root), hipMemcpyHostToDevice));
    GrowCfg gc = c->gc;
    launch_plan_level(c->stream, c->front[0], c->splits, c->ctl, c->tree, c->ops, c->pch, c->d_nbins, gc, c->scales,
                      1, part_chunk_rows(c->Fs));
    launch_part_count(c->stream, c->pch, c->ops, c->ctl, c->work[0], c->work[1], c->plane0, c->Fs, c->chunk_left,
                      part_grid(c));
    // oblivious = 1: no demotion, so an all-one-side split still scatters.
    launch_part_scan(c->stream, c->quant, c->ops, c->pch, c->chunk_left, c->ctl, c->front[0], c->front[1],
                     c->splits, c->tree, c->hch, c->fops, gc, c->scales, 1, 1);
    launch_part_scatter(c->stream, c->pch, c->ops, c->ctl, c->work[0], c->work[1], c->plane0, gh_src(c, 0),
                        gh_src(c, 1), gh_src(c, 2), c->work[0], c->work[1], const_cast<void *>(gh_src(c, 0)),
                        const_cast<void *>(gh_src(c, 1)), c->Fs, gh_bytes, part_grid(c));
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(c->stream));
    hipFree(d_rows);
    PartOp op;
    HIP_TRY(hipMemcpy(&op, c->ops, sizeof(op), hipMemcpyDeviceToHost));
    *n_left = op.n_left;
    if (gh_bytes == 8) {
        std::vector<uint2> v(n);
        HIP_TRY(hipMemcpy(v.data(), gh_src(c, 1), n * sizeof(uint2), hipMemcpyDeviceToHost));
        for (u64 i = 0; i < n; ++i) out_rows[i] = v[i].x;
    } else {
        HIP_TRY(hipMemcpy(out_rows, gh_src(c, 1), n * sizeof(u32), hipMemcpyDeviceToHost));
    }
    return 0;
}

int bh_profile_hist(bh_context *c, uint32_t n_iters, double *ms_total, double *bytes_total, uint64_t *launches,
                    uint64_t *rows_total, double *scanned_bytes_total, uint64_t *scanned_rows_total)
{
    if (int r = check_ready(c)) return r;
    if (c->fc.grower != BH_DEPTHWISE) return fail(1, "histogram profiling covers the depthwise grower");
    HIP_TRY(hipSetDevice(c->device));
    if (int r = ensure_tree_out(c, n_iters)) return r;
    c->prof = true;
    c->prof_ms = c->prof_bytes = c->prof_scanned_bytes = 0.0;
    c->prof_launches = c->prof_rows = c->prof_scanned = 0;
    prime_gradients(c);
    for (u32 it = 0; it < n_iters; ++it) enqueue_iteration(c);
    c->prof = false;
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(c->stream));
    if (int r = check_partition(c)) return r;
    c->trees_done += n_iters;
    *ms_total = c->prof_ms;
    *bytes_total = c->prof_bytes;
    *launches = c->prof_launches;
    *rows_total = c->prof_rows;
    if (scanned_bytes_total) *scanned_bytes_total = c->prof_scanned_bytes;
    if (scanned_rows_total) *scanned_rows_total = c->prof_scanned;
    return 0;
}

int bh_time_root_histogram(bh_context *c, uint32_t reps, float *ms_mean)
{
    if (int r = check_ready(c)) return r;
    HIP_TRY(hipSetDevice(c->device));
    // Root units over the pristine plane and the current gradients.
    launch_tree_begin(c->stream, c->ctl, c->front[0], c->tree, c->hch, c->fops,
                      RootSel{static_cast<u32>(c->N), kSrcPlane0, 0u});
    HIP_TRY(hipStreamSynchronize(c->stream));
    float total = 0.0F;
    for (u32 r = 0; r < reps; ++r) {
        HIP_TRY(hipEventRecord(c->ev0, c->stream));
        launch_hist(c->stream, c->quant, c->hch, c->ctl, c->work[0], c->work[1], c->plane0, gh_src(c, 0),
                    gh_src(c, 1), gh_src(c, 2), c->Fs, c->F, c->SB, c->hist_fsl(), c->hist_waves(), hist_grid(c),
                    c->partials, c->hist[0], c->fx);
        HIP_TRY(hipEventRecord(c->ev1, c->stream));
        HIP_TRY(hipEventSynchronize(c->ev1));
        float ms = 0.0F;
        HIP_TRY(hipEventElapsedTime(&ms, c->ev0, c->ev1));
        total += ms;
    }
    *ms_mean = reps ? total / static_cast<float>(reps) : 0.0F;
    return 0;
}


// =================================================================================
// Engine steps: the reference's GPULevelEngine / GPULeafEngine methods
// (include/bonsai/grower.hpp:91-162; CudaHistogramEngine,
// include/bonsai/cuda/histogram_engine.hpp:90-218) for growers that keep the control loop
// on the host (LevelStep / LeafStep, src/step/level.hpp:159-366, src/step/leaf.hpp:124-267).
// Rows never cross the bus: a partition moves each parent's segment into the other work
// plane (left rows ascending, then right rows, at the parent's offsets), histograms stay
// in slot-indexed device buffers, and only ops, counts and split records move.  A host
// routing tree records every partition, so finalize walks each row of the pristine plane
// to its stamped leaf.
// =================================================================================
namespace {

// Host <-> device copies of the step path, ordered on the engine stream and complete on
// return.  The engine stream is non-blocking, so a plain hipMemcpy (null stream) would
// neither wait for the step's kernels before reading their output nor keep an upload
// from landing while a kernel still reads the old contents.
hipError_t step_copy(bh_context *c, void *dst, void const *src, size_t bytes, hipMemcpyKind kind)
{
    hipError_t const e = hipMemcpyAsync(dst, src, bytes, kind, c->stream);
    return e == hipSuccess ? hipStreamSynchronize(c->stream) : e;
}

// The control block fields the step launches read.
int step_ctl(bh_context *c, u32 n_front, u32 n_ops, u32 n_part_chunks, u32 n_hist_chunks, u32 n_finish)
{
    TreeCtl t{};
    t.n_front = n_front;
    t.n_ops = n_ops;
    t.n_part_chunks = n_part_chunks;
    t.n_hist_chunks = n_hist_chunks;
    t.n_finish = n_finish;
    HIP_TRY(step_copy(c, c->ctl, &t, sizeof(t), hipMemcpyHostToDevice));
    return 0;
}

int step_ready(bh_context *c)
{
    if (int r = check_ready(c)) return r;
    if (c->gc.mono || c->gc.n_groups)
        return fail(1, "engine steps take no monotone or interaction constraints (the host grower owns them)");
    if (c->fsamp_kind == 2)
        return fail(1, "engine steps take a handed-over feature mask, not the device feature draw");
    return 0;
}

// The first selected feature (its cells give node totals, include/bonsai/split.hpp:36-53).
u32 step_totals_feature(bh_context *c)
{
    if (c->fsamp_kind != 1 || !c->fmask) return 0;
    std::vector<u8> m(c->F);
    if (step_copy(c, m.data(), c->fmask, c->F, hipMemcpyDeviceToHost) != hipSuccess) return 0;
    for (u32 f = 0; f < c->F; ++f)
        if (m[f]) return f;
    return 0;
}

// Build the histograms of `builds` (small slot, its segment, large slot, parent slot):
// 8,192-row units, partials folded in unit order, large = parent - small.
struct StepBuild {
    u32 small, large, parent;
    StepSeg seg;
};
int step_build(bh_context *c, std::vector<StepBuild> const &builds, void *hist_cur, void *hist_next)
{
    std::vector<HistChunk> hc;
    std::vector<FinishOp> fo;
    u32 parts = 0;
    for (auto const &b : builds) {
        u32 const units = (b.seg.cnt + kHistChunk - 1) / kHistChunk;
        FinishOp f{};
        f.small_slot = b.small;
        f.large_slot = b.large;
        f.parent_slot = b.parent;
        f.part_first = parts;
        f.part_count = units > 1 ? units : 0;
        f.empty = b.seg.cnt == 0 ? 1u : 0u;
        for (u32 k = 0; k < units; ++k) {
            HistChunk h{};
            h.slot = b.small;
            h.src = b.seg.src;
            h.begin = b.seg.off + k * kHistChunk;
            h.end = b.seg.off + std::min(b.seg.cnt, (k + 1) * kHistChunk);
            h.part = units > 1 ? parts + k : kNone;
            hc.push_back(h);
        }
        if (units > 1) parts += units;
        fo.push_back(f);
    }
    if (hc.size() > c->max_hch || parts > c->max_parts || fo.size() > c->max_ops + 2)
        return fail(1, "engine step: more histogram units than the configured fit holds");
    if (!hc.empty()) HIP_TRY(step_copy(c, c->hch, hc.data(), hc.size() * sizeof(HistChunk), hipMemcpyHostToDevice));
    if (!fo.empty()) HIP_TRY(step_copy(c, c->fops, fo.data(), fo.size() * sizeof(FinishOp), hipMemcpyHostToDevice));
    if (int r = step_ctl(c, 0, 0, 0, static_cast<u32>(hc.size()), static_cast<u32>(fo.size()))) return r;
    HIP_TRY(launch_hist(c->stream, c->quant, c->hch, c->ctl, c->work[0], c->work[1], c->plane0, gh_src(c, 0),
                        gh_src(c, 1), gh_src(c, 2), c->Fs, c->F, c->SB, c->hist_fsl(), c->hist_waves(),
                        std::max(1u, std::min(static_cast<u32>(hc.size()), 4096u)), c->partials, hist_next, c->fx));
    launch_finish(c->stream, c->quant, c->fops, c->ctl, c->partials, hist_cur, hist_next, c->node_cells(),
                  std::max<u32>(1u, static_cast<u32>(fo.size())), nullptr, nullptr, false,
                  hist_transpose_F(c->quant, c->Fs, c->F));
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(c->stream));
    return 0;
}

// Stable partition of each segment by its rule into the other work plane at the same
// offsets; n_left per op out.
struct StepPart {
    StepSeg seg;
    u32 feature, bin, default_left;
};
int step_partition(bh_context *c, std::vector<StepPart> const &parts, std::vector<u32> &n_left)
{
    u32 const chunk_rows = part_chunk_rows(c->Fs);
    std::vector<PartOp> ops(parts.size());
    std::vector<PartChunk> pch;
    for (size_t i = 0; i < parts.size(); ++i) {
        StepPart const &p = parts[i];
        PartOp &o = ops[i];
        o = PartOp{};
        o.parent = static_cast<u32>(i);
        o.feature = p.feature;
        o.bin = p.bin;
        o.default_left = p.default_left;
        o.last_bin = static_cast<u32>(c->n_bins[p.feature]) - 1u;
        o.off = p.seg.off;
        o.cnt = p.seg.cnt;
        o.src = p.seg.src;
        o.dst = p.seg.src == 0 ? 1u : 0u;
        o.chunk_first = static_cast<u32>(pch.size());
        o.alive = 1;
        for (u32 b = 0; b < p.seg.cnt; b += chunk_rows)
            pch.push_back(PartChunk{static_cast<u32>(i), b, std::min(p.seg.cnt, b + chunk_rows), 0u});
        o.chunk_count = static_cast<u32>(pch.size()) - o.chunk_first;
    }
    if (ops.size() > c->max_ops + 2 || pch.size() > c->max_pch)
        return fail(1, "engine step: more partition work than the configured fit holds");
    n_left.assign(parts.size(), 0u);
    if (pch.empty()) return 0;
    HIP_TRY(step_copy(c, c->ops, ops.data(), ops.size() * sizeof(PartOp), hipMemcpyHostToDevice));
    HIP_TRY(step_copy(c, c->pch, pch.data(), pch.size() * sizeof(PartChunk), hipMemcpyHostToDevice));
    if (int r = step_ctl(c, 0, static_cast<u32>(ops.size()), static_cast<u32>(pch.size()), 0, 0)) return r;
    launch_part_count(c->stream, c->pch, c->ops, c->ctl, c->work[0], c->work[1], c->plane0, c->Fs, c->chunk_left,
                      part_grid(c));
    std::vector<u32> cl(pch.size());
    HIP_TRY(step_copy(c, cl.data(), c->chunk_left, cl.size() * sizeof(u32), hipMemcpyDeviceToHost));
    for (size_t k = 0; k < pch.size(); ++k) {
        if (cl[k] > pch[k].end - pch[k].begin) return fail(-2, "engine step: partition count beyond its chunk");
        PartOp &o = ops[pch[k].op];
        pch[k].left_before = o.n_left;
        o.n_left += cl[k];
    }
    for (size_t i = 0; i < ops.size(); ++i) n_left[i] = ops[i].n_left;
    HIP_TRY(step_copy(c, c->ops, ops.data(), ops.size() * sizeof(PartOp), hipMemcpyHostToDevice));
    HIP_TRY(step_copy(c, c->pch, pch.data(), pch.size() * sizeof(PartChunk), hipMemcpyHostToDevice));
    u32 const gh_bytes = c->quant ? 4 : 8;
    launch_part_scatter(c->stream, c->pch, c->ops, c->ctl, c->work[0], c->work[1], c->plane0, gh_src(c, 0),
                        gh_src(c, 1), gh_src(c, 2), c->work[0], c->work[1], const_cast<void *>(gh_src(c, 0)),
                        const_cast<void *>(gh_src(c, 1)), c->Fs, gh_bytes, part_grid(c));
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(c->stream));
    return 0;
}

// Node totals as the finders take them: the first selected feature's cells of the node's
// histogram, summed in bin order (FP32: double sums stored as float, Histogram::totals,
// include/bonsai/histogram.hpp:76-87; quantized: exact integers, times the scales).
int step_totals(bh_context *c, std::vector<NodeInfo> const &front, void const *hist, bh_step_node *nodes)
{
    u32 const f0 = step_totals_feature(c);
    u32 const nb = c->n_bins[f0];
    double sc[2] = {0.0, 0.0};
    if (c->quant) HIP_TRY(step_copy(c, sc, c->scales, sizeof(sc), hipMemcpyDeviceToHost));
    for (size_t i = 0; i < front.size(); ++i) {
        size_t const off = static_cast<size_t>(front[i].slot) * c->node_cells() + static_cast<size_t>(f0) * c->SB;
        bh_step_node &o = nodes[i];
        o.rows = front[i].cnt;
        if (c->quant) {
            std::vector<i64> cells(2 * nb);
            HIP_TRY(step_copy(c, cells.data(), static_cast<i64 const *>(hist) + 2 * off, cells.size() * sizeof(i64),
                              hipMemcpyDeviceToHost));
            i64 g = 0, h = 0;
            for (u32 b = 0; b < nb; ++b) {
                g += cells[2 * b];
                h += cells[2 * b + 1];
            }
            o.isum_g = g;
            o.isum_h = h;
            o.sum_g = static_cast<double>(g) * sc[0];
            o.sum_h = static_cast<double>(h) * sc[1];
        } else {
            std::vector<float> cells(2 * nb);
            HIP_TRY(step_copy(c, cells.data(), static_cast<float const *>(hist) + 2 * off,
                              cells.size() * sizeof(float), hipMemcpyDeviceToHost));
            double g = 0.0, h = 0.0;
            for (u32 b = 0; b < nb; ++b) {
                g += cells[2 * b];
                h += cells[2 * b + 1];
            }
            o.sum_g = static_cast<float>(g);
            o.sum_h = static_cast<float>(h);
            o.isum_g = o.isum_h = 0;
        }
    }
    return 0;
}

// The split records of nodes laid out as NodeInfo (slot = histogram slot), with their
// totals and rows into nodes.
int step_find(bh_context *c, std::vector<NodeInfo> const &front, void const *hist, bh_split *out,
              bh_step_node *nodes)
{
    u32 const n = static_cast<u32>(front.size());
    if (n == 0) return 0;
    if (int r = step_totals(c, front, hist, nodes)) return r;
    if (n > std::max(c->max_front * 2u, 4u)) return fail(1, "engine step: more nodes than the fit's frontier");
    HIP_TRY(step_copy(c, c->front[0], front.data(), n * sizeof(NodeInfo), hipMemcpyHostToDevice));
    if (int r = step_ctl(c, n, 0, 0, 0, 0)) return r;
    launch_find(c->stream, c->quant, c->front[0], c->ctl, hist, c->d_nbins, c->F, c->SB, c->gc, c->scales, c->splits,
                n, 0, n, c->fscr);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(c->stream));
    static_assert(sizeof(SplitRec) == sizeof(bh_split), "bh_split layout");
    HIP_TRY(step_copy(c, out, c->splits, n * sizeof(SplitRec), hipMemcpyDeviceToHost));
    return 0;
}

// Root of a step-driven tree: every row (pristine plane) or the row subset (gathered
// into work plane 1), its histogram into slot 0 of hist[0], the totals from it.
int step_root(bh_context *c, double *sums, uint64_t *rows)
{
    StepSeg root;
    root.live = true;
    if (c->samp_kind == 2) return fail(1, "engine steps take a handed-over row subset, not the device Bernoulli draw");
    if (c->samp_kind == 1) {
        launch_gather_rows(c->stream, c->plane0, c->Fs, gh_src(c, 2), c->quant ? 4 : 8, c->samp_rows, c->samp_n,
                           c->work[1], const_cast<void *>(gh_src(c, 1)), 0);
        root.src = 1;
        root.cnt = c->samp_n;
    } else {
        root.src = kSrcPlane0;
        root.cnt = static_cast<u32>(c->N);
    }
    if (int r = step_build(c, {StepBuild{0u, kNone, kNone, root}}, c->hist[1] ? c->hist[1] : c->hist[0], c->hist[0]))
        return r;
    u32 const f0 = step_totals_feature(c);
    u32 const nb = c->n_bins[f0];
    size_t const off = static_cast<size_t>(f0) * c->SB;
    double g = 0.0, h = 0.0;
    if (c->quant) {
        std::vector<i64> cells(2 * nb);
        double sc[2];
        HIP_TRY(step_copy(c, cells.data(), static_cast<i64 const *>(c->hist[0]) + 2 * off, cells.size() * sizeof(i64),
                          hipMemcpyDeviceToHost));
        HIP_TRY(step_copy(c, sc, c->scales, sizeof(sc), hipMemcpyDeviceToHost));
        i64 ig = 0, ih = 0;
        for (u32 b = 0; b < nb; ++b) {
            ig += cells[2 * b];
            ih += cells[2 * b + 1];
        }
        g = static_cast<double>(ig) * sc[0];
        h = static_cast<double>(ih) * sc[1];
    } else {
        std::vector<float> cells(2 * nb);
        HIP_TRY(step_copy(c, cells.data(), static_cast<float const *>(c->hist[0]) + 2 * off,
                          cells.size() * sizeof(float), hipMemcpyDeviceToHost));
        for (u32 b = 0; b < nb; ++b) {
            g += cells[2 * b];
            h += cells[2 * b + 1];
        }
        g = static_cast<float>(g); // Histogram::totals stores the double sums as float
        h = static_cast<float>(h);
    }
    sums[0] = g;
    sums[1] = h;
    *rows = root.cnt;
    c->route.assign(1, DevNode{-1, 0, 0, 0, 0, 0.0F, 0.0F, 0.0F});
    c->route_stamp.assign(1, kNone);
    return 0;
}

// Record a partition in the routing tree: the parent's node gets the split, its two
// children are new leaves.
void route_split(bh_context *c, u32 node, u32 feature, u32 bin, u32 dl, u32 &left, u32 &right)
{
    left = static_cast<u32>(c->route.size());
    right = left + 1u;
    DevNode &p = c->route[node];
    p.feature = static_cast<int32_t>(feature);
    p.bin = bin;
    p.default_left = dl;
    p.left = left;
    p.right = right;
    c->route.push_back(DevNode{-1, 0, 0, 0, 0, 0.0F, 0.0F, 0.0F});
    c->route.push_back(DevNode{-1, 0, 0, 0, 0, 0.0F, 0.0F, 0.0F});
    c->route_stamp.push_back(kNone);
    c->route_stamp.push_back(kNone);
}

NodeInfo step_node(StepSeg const &g, u32 slot, u32 depth)
{
    NodeInfo nd{};
    nd.id = g.route;
    nd.off = g.off;
    nd.cnt = g.cnt;
    nd.src = g.src;
    nd.slot = slot;
    nd.depth = depth;
    return nd;
}

} // namespace

int bh_begin_root(bh_context *c, double *sums, uint64_t *rows)
{
    if (int r = step_ready(c)) return r;
    if (c->fc.grower == BH_LEAFWISE) return fail(1, "the level plane needs a depthwise or levelwise fit");
    if (!sums || !rows) return fail(1, "null output");
    HIP_TRY(hipSetDevice(c->device));
    c->step_mode = 0;
    if (int r = step_root(c, sums, rows)) return r;
    StepSeg root;
    root.live = true;
    root.src = c->samp_kind == 1 ? 1u : kSrcPlane0;
    root.cnt = static_cast<u32>(*rows);
    c->seg_cur.assign(1, root);
    c->seg_next.clear();
    c->step_hist = 0;
    c->step_mode = 1;
    return 0;
}

int bh_partition_level(bh_context *c, const bh_partition_op *ops, uint32_t n_ops, uint32_t *child_counts)
{
    if (!c || c->step_mode != 1) return fail(1, "no level-plane tree begun (bh_begin_root)");
    if (n_ops > 0 && (!ops || !child_counts)) return fail(1, "null ops or counts");
    HIP_TRY(hipSetDevice(c->device));
    std::vector<StepPart> parts(n_ops);
    u32 max_slot = 0;
    for (u32 i = 0; i < n_ops; ++i) {
        bh_partition_op const &o = ops[i];
        if (o.parent_slot >= c->seg_cur.size() || !c->seg_cur[o.parent_slot].live)
            return fail(1, "partition op names a slot with no rows at this level");
        if (o.feature >= c->F) return fail(1, "partition op feature out of range");
        if (o.left_slot == o.right_slot) return fail(1, "partition op children share a slot");
        max_slot = std::max(max_slot, std::max(o.left_slot, o.right_slot));
        parts[i] = StepPart{c->seg_cur[o.parent_slot], o.feature, o.bin, o.default_left ? 1u : 0u};
    }
    if (n_ops && max_slot >= std::max(c->max_front, 2u)) return fail(1, "child slot beyond the fit's frontier");
    std::vector<u32> nl;
    if (int r = step_partition(c, parts, nl)) return r;
    c->seg_next.assign(n_ops ? max_slot + 1u : 0u, StepSeg{});
    for (u32 i = 0; i < n_ops; ++i) {
        bh_partition_op const &o = ops[i];
        StepSeg const &p = c->seg_cur[o.parent_slot];
        u32 const dst = p.src == 0 ? 1u : 0u;
        u32 lr = 0, rr = 0;
        route_split(c, p.route, o.feature, o.bin, o.default_left ? 1u : 0u, lr, rr);
        StepSeg &L = c->seg_next[o.left_slot];
        StepSeg &R = c->seg_next[o.right_slot];
        if (L.live || R.live) return fail(1, "two partition ops name the same child slot");
        L = StepSeg{p.off, nl[i], dst, lr, true};
        R = StepSeg{p.off + nl[i], p.cnt - nl[i], dst, rr, true};
        child_counts[2 * i] = nl[i];
        child_counts[2 * i + 1] = p.cnt - nl[i];
    }
    return 0;
}

int bh_advance_level(bh_context *c, const bh_level_op *ops, uint32_t n_ops)
{
    if (!c || c->step_mode != 1) return fail(1, "no level-plane tree begun (bh_begin_root)");
    if (n_ops > 0 && !ops) return fail(1, "null ops");
    HIP_TRY(hipSetDevice(c->device));
    std::vector<StepBuild> b(n_ops);
    for (u32 i = 0; i < n_ops; ++i) {
        bh_level_op const &o = ops[i];
        if (o.parent_slot >= c->seg_cur.size() || o.small_slot >= c->seg_next.size() ||
            o.large_slot >= c->seg_next.size() || !c->seg_next[o.small_slot].live)
            return fail(1, "level op names a slot the partition did not make");
        if (std::max(o.small_slot, o.large_slot) >= c->hist_slots)
            return fail(1, "level op slot beyond the fit's histogram slots");
        b[i] = StepBuild{o.small_slot, o.large_slot, o.parent_slot, c->seg_next[o.small_slot]};
    }
    void *cur = c->hist[c->step_hist], *nxt = c->hist[1u - c->step_hist];
    if (!nxt) return fail(1, "the level plane needs two histogram buffers");
    if (int r = step_build(c, b, cur, nxt)) return r;
    c->seg_cur = std::move(c->seg_next);
    c->seg_next.clear();
    c->step_hist = 1u - c->step_hist;
    return 0;
}

int bh_advance_level_per_node(bh_context *c, const bh_level_op *ops, uint32_t n_ops)
{
    if (!c || c->step_mode != 1) return fail(1, "no level-plane tree begun (bh_begin_root)");
    if (n_ops > 0 && !ops) return fail(1, "null ops");
    HIP_TRY(hipSetDevice(c->device));
    void *cur = c->hist[c->step_hist], *nxt = c->hist[1u - c->step_hist];
    if (!nxt) return fail(1, "the level plane needs two histogram buffers");
    // Every node's units and completion op uploaded once; then one histogram launch and
    // one completion launch per node, each with its own unit count in the control block.
    std::vector<HistChunk> hc;
    std::vector<FinishOp> fo;
    std::vector<u32> first(n_ops + 1, 0);
    u32 parts = 0;
    for (u32 i = 0; i < n_ops; ++i) {
        bh_level_op const &o = ops[i];
        if (o.parent_slot >= c->seg_cur.size() || o.small_slot >= c->seg_next.size() ||
            o.large_slot >= c->seg_next.size() || !c->seg_next[o.small_slot].live)
            return fail(1, "level op names a slot the partition did not make");
        if (std::max(o.small_slot, o.large_slot) >= c->hist_slots)
            return fail(1, "level op slot beyond the fit's histogram slots");
        StepSeg const &g = c->seg_next[o.small_slot];
        u32 const units = (g.cnt + kHistChunk - 1) / kHistChunk;
        FinishOp f{};
        f.small_slot = o.small_slot;
        f.large_slot = o.large_slot;
        f.parent_slot = o.parent_slot;
        f.part_first = parts;
        f.part_count = units > 1 ? units : 0;
        f.empty = g.cnt == 0 ? 1u : 0u;
        first[i] = static_cast<u32>(hc.size());
        for (u32 k = 0; k < units; ++k) {
            HistChunk h{};
            h.slot = o.small_slot;
            h.src = g.src;
            h.begin = g.off + k * kHistChunk;
            h.end = g.off + std::min(g.cnt, (k + 1) * kHistChunk);
            h.part = units > 1 ? parts + k : kNone;
            hc.push_back(h);
        }
        if (units > 1) parts += units;
        fo.push_back(f);
    }
    first[n_ops] = static_cast<u32>(hc.size());
    if (hc.size() > c->max_hch || parts > c->max_parts || fo.size() > c->max_ops + 2)
        return fail(1, "engine step: more histogram units than the configured fit holds");
    if (!hc.empty()) HIP_TRY(step_copy(c, c->hch, hc.data(), hc.size() * sizeof(HistChunk), hipMemcpyHostToDevice));
    if (!fo.empty()) HIP_TRY(step_copy(c, c->fops, fo.data(), fo.size() * sizeof(FinishOp), hipMemcpyHostToDevice));
    TreeCtl *hctl = nullptr;
    HIP_TRY(hipHostMalloc(reinterpret_cast<void **>(&hctl), std::max<u32>(1u, n_ops) * sizeof(TreeCtl), 0));
    u32 const tF = hist_transpose_F(c->quant, c->Fs, c->F);
    hipError_t e = hipSuccess;
    for (u32 i = 0; i < n_ops && e == hipSuccess; ++i) {
        hctl[i] = TreeCtl{};
        hctl[i].n_hist_chunks = first[i + 1] - first[i];
        hctl[i].n_finish = 1;
        e = hipMemcpyAsync(c->ctl, &hctl[i], sizeof(TreeCtl), hipMemcpyHostToDevice, c->stream);
        if (e == hipSuccess)
            e = launch_hist(c->stream, c->quant, c->hch + first[i], c->ctl, c->work[0], c->work[1], c->plane0,
                            gh_src(c, 0), gh_src(c, 1), gh_src(c, 2), c->Fs, c->F, c->SB, c->hist_fsl(),
                            c->hist_waves(), std::max(1u, std::min(first[i + 1] - first[i], 4096u)), c->partials,
                            nxt, c->fx);
        if (e == hipSuccess)
            launch_finish(c->stream, c->quant, c->fops + i, c->ctl, c->partials, cur, nxt, c->node_cells(), 1,
                          nullptr, nullptr, false, tF);
    }
    hipError_t const s2 = hipStreamSynchronize(c->stream);
    (void)hipHostFree(hctl);
    if (e != hipSuccess) return fail(-1, std::string("hip: per-node build: ") + hipGetErrorString(e));
    if (s2 != hipSuccess) return fail(-1, std::string("hip: per-node build: ") + hipGetErrorString(s2));
    HIP_TRY(hipGetLastError());
    c->seg_cur = std::move(c->seg_next);
    c->seg_next.clear();
    c->step_hist = 1u - c->step_hist;
    return 0;
}

int bh_advance_layout_only(bh_context *c)
{
    if (!c || c->step_mode != 1) return fail(1, "no level-plane tree begun (bh_begin_root)");
    c->seg_cur = std::move(c->seg_next);
    c->seg_next.clear();
    return 0;
}

int bh_find_splits_many(bh_context *c, bh_step_node *nodes, uint32_t n, bh_split *out)
{
    if (!c || c->step_mode != 1) return fail(1, "no level-plane tree begun (bh_begin_root)");
    if (n > c->seg_cur.size()) return fail(1, "more nodes than slots at this level");
    if (n > 0 && (!nodes || !out)) return fail(1, "null nodes or output");
    HIP_TRY(hipSetDevice(c->device));
    std::vector<NodeInfo> f(n);
    for (u32 i = 0; i < n; ++i) {
        if (!c->seg_cur[i].live) return fail(1, "a node's slot holds no rows at this level");
        f[i] = step_node(c->seg_cur[i], i, nodes[i].depth);
    }
    return step_find(c, f, c->hist[c->step_hist], out, nodes);
}

int bh_find_level_split(bh_context *c, bh_step_node *nodes, uint32_t n, bh_split *out)
{
    if (!c || c->step_mode != 1) return fail(1, "no level-plane tree begun (bh_begin_root)");
    if (c->fc.grower != BH_LEVELWISE) return fail(1, "the level split needs a levelwise fit");
    if (n == 0 || n > c->seg_cur.size()) return fail(1, "node count out of range");
    if (!nodes || !out) return fail(1, "null nodes or output");
    HIP_TRY(hipSetDevice(c->device));
    std::vector<NodeInfo> f(n);
    for (u32 i = 0; i < n; ++i) {
        if (!c->seg_cur[i].live) return fail(1, "a node's slot holds no rows at this level");
        f[i] = step_node(c->seg_cur[i], i, nodes[i].depth);
    }
    if (int r = step_totals(c, f, c->hist[c->step_hist], nodes)) return r;
    HIP_TRY(step_copy(c, c->front[0], f.data(), n * sizeof(NodeInfo), hipMemcpyHostToDevice));
    if (int r = step_ctl(c, n, 0, 0, 0, 0)) return r;
    SplitRec *pick = c->splits + std::max(c->max_front * 2u, 4u) - 1u; // last record: the pick
    if (n >= std::max(c->max_front * 2u, 4u)) return fail(1, "more nodes than the fit's frontier");
    launch_level_find_step(c->ext, c->stream, c->quant, grow_view(c), c->front[0], n, c->hist[c->step_hist], pick);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(c->stream));
    SplitRec p;
    HIP_TRY(step_copy(c, &p, pick, sizeof(p), hipMemcpyDeviceToHost));
    std::vector<SplitRec> per(n);
    if (p.valid) HIP_TRY(step_copy(c, per.data(), c->splits, n * sizeof(SplitRec), hipMemcpyDeviceToHost));
    for (u32 i = 0; i < n; ++i) {
        SplitRec r = p.valid ? per[i] : p;
        std::memcpy(&out[i], &r, sizeof(r));
    }
    return 0;
}

int bh_stamp_leaves(bh_context *c, const bh_leaf_stamp *stamps, uint32_t n)
{
    if (!c || c->step_mode == 0) return fail(1, "no tree begun");
    if (n > 0 && !stamps) return fail(1, "null stamps");
    std::vector<StepSeg> const &segs = c->step_mode == 1 ? c->seg_cur : c->pool;
    for (u32 i = 0; i < n; ++i) {
        if (stamps[i].slot >= segs.size() || !segs[stamps[i].slot].live)
            return fail(1, "stamp names a slot with no rows");
        c->route_stamp[segs[stamps[i].slot].route] = stamps[i].node_id;
    }
    return 0;
}

int bh_finalize_tree(bh_context *c, const float *node_values, uint32_t n_values, float *values,
                     uint32_t *leaf_ids)
{
    if (!c || c->step_mode == 0) return fail(1, "no tree begun");
    if (c->route.size() > c->max_nodes) return fail(1, "the routing tree exceeds the fit's node capacity");
    HIP_TRY(hipSetDevice(c->device));
    HIP_TRY(step_copy(c, c->tree, c->route.data(), c->route.size() * sizeof(DevNode), hipMemcpyHostToDevice));
    TreeCtl t{};
    t.n_nodes = static_cast<u32>(c->route.size());
    HIP_TRY(step_copy(c, c->ctl, &t, sizeof(t), hipMemcpyHostToDevice));
    launch_score_update(c->stream, c->plane0, c->Fs, c->N, c->tree, c->ctl, c->d_nbins, c->F, 0.0F, nullptr,
                        c->leaf_rows, c->max_nodes);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(c->stream));
    std::vector<u32> lr(c->N);
    HIP_TRY(step_copy(c, lr.data(), c->leaf_rows, c->N * sizeof(u32), hipMemcpyDeviceToHost));
    for (u64 r = 0; r < c->N; ++r) {
        u32 const id = c->route_stamp[lr[r]];
        if (leaf_ids) leaf_ids[r] = id;
        if (values) values[r] = node_values && id < n_values ? node_values[id] : 0.0F;
    }
    return 0;
}

int bh_leaf_begin_root(bh_context *c, double *sums, uint64_t *rows)
{
    if (int r = step_ready(c)) return r;
    if (c->fc.grower != BH_LEAFWISE) return fail(1, "the leaf plane needs a leafwise fit");
    if (!sums || !rows) return fail(1, "null output");
    HIP_TRY(hipSetDevice(c->device));
    c->step_mode = 0;
    if (int r = step_root(c, sums, rows)) return r;
    StepSeg root;
    root.live = true;
    root.src = c->samp_kind == 1 ? 1u : kSrcPlane0;
    root.cnt = static_cast<u32>(*rows);
    c->pool.assign(1, root);
    c->step_next_slot = 1;
    c->step_mode = 2;
    return 0;
}

int bh_leaf_split(bh_context *c, uint32_t parent_slot, uint32_t feature, uint32_t bin, uint32_t default_left,
                  bh_leaf_round *out)
{
    if (!c || c->step_mode != 2) return fail(1, "no leaf-plane tree begun (bh_leaf_begin_root)");
    if (parent_slot >= c->pool.size() || !c->pool[parent_slot].live) return fail(1, "parent slot holds no rows");
    if (feature >= c->F) return fail(1, "feature out of range");
    if (!out) return fail(1, "null output");
    HIP_TRY(hipSetDevice(c->device));
    StepSeg const p = c->pool[parent_slot];
    std::vector<u32> nl;
    if (int r = step_partition(c, {StepPart{p, feature, bin, default_left ? 1u : 0u}}, nl)) return r;
    u32 const dst = p.src == 0 ? 1u : 0u;
    u32 const nr = p.cnt - nl[0];
    *out = bh_leaf_round{0u, 0u, nl[0], nr};
    if (nl[0] == 0 || nr == 0) {
        // one side emptied: no slot taken, the parent keeps its (moved, same order) rows
        c->pool[parent_slot].src = dst;
        return 0;
    }
    // the smaller child (ties: the left) takes the next free slot, the larger inherits
    // the parent's
    bool const left_small = nl[0] <= nr;
    u32 const small_slot = c->step_next_slot;
    if (small_slot >= c->hist_slots) return fail(1, "leaf slot pool exhausted (max_leaves)");
    c->step_next_slot += 1;
    if (c->pool.size() <= small_slot) c->pool.resize(small_slot + 1u);
    u32 lr = 0, rr = 0;
    route_split(c, p.route, feature, bin, default_left ? 1u : 0u, lr, rr);
    StepSeg const L{p.off, nl[0], dst, lr, true}, R{p.off + nl[0], nr, dst, rr, true};
    u32 const ls = left_small ? small_slot : parent_slot, rs = left_small ? parent_slot : small_slot;
    c->pool[ls] = L;
    c->pool[rs] = R;
    *out = bh_leaf_round{ls, rs, nl[0], nr};
    return 0;
}

int bh_leaf_build(bh_context *c, uint32_t small_slot, uint32_t large_slot)
{
    if (!c || c->step_mode != 2) return fail(1, "no leaf-plane tree begun (bh_leaf_begin_root)");
    if (small_slot >= c->pool.size() || large_slot >= c->pool.size() || !c->pool[small_slot].live ||
        !c->pool[large_slot].live || small_slot == large_slot)
        return fail(1, "leaf build names slots the split did not make");
    HIP_TRY(hipSetDevice(c->device));
    // large = parent - small in place: the large child inherited the parent's slot
    return step_build(c, {StepBuild{small_slot, large_slot, large_slot, c->pool[small_slot]}}, c->hist[0],
                      c->hist[0]);
}

int bh_leaf_find(bh_context *c, bh_step_node *nodes, const uint32_t *slots, uint32_t n, bh_split *out)
{
    if (!c || c->step_mode != 2) return fail(1, "no leaf-plane tree begun (bh_leaf_begin_root)");
    if (n > 0 && (!nodes || !slots || !out)) return fail(1, "null nodes, slots or output");
    HIP_TRY(hipSetDevice(c->device));
    std::vector<NodeInfo> f(n);
    for (u32 i = 0; i < n; ++i) {
        if (slots[i] >= c->pool.size() || !c->pool[slots[i]].live) return fail(1, "slot holds no rows");
        f[i] = step_node(c->pool[slots[i]], slots[i], nodes[i].depth);
    }
    return step_find(c, f, c->hist[0], out, nodes);
}

} // extern "C"
