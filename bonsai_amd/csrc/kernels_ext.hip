// Leafwise and levelwise grow loops on the device.
//
// Both reuse the depthwise data plane (partition count/scan/scatter, histogram units,
// fixed-order completion, node finder) and add their own control kernels:
//   levelwise (ObliviousGrower::grow, src/grower_impl.hpp:598-734): one split per level
//     chosen over the whole frontier (k_find_level + k_level_pick), broadcast to every
//     node, empty children kept;
//   leafwise (LeafwiseGrower::grow, src/grower_impl.hpp:743-862): a device-resident
//     candidate pool replaces the host max-heap; every round pops the best leaf (gain,
//     then lowest node id), partitions it, builds the smaller child, subtracts in place
//     and scores both children.
#include <hip/hip_runtime.h>
#include <algorithm>
#include "bh_common.hpp"
#include "kernels.hpp"
#include "kernels_ext.hpp"
#include "bh_device.hpp"
#include "bh_find.hpp"

namespace bh {

namespace {

__device__ __forceinline__ double l1_thr(double g, double l1)
{
    if (g > l1) return g - l1;
    if (g < -l1) return g + l1;
    return 0.0;
}
__device__ __forceinline__ double score_of(double g, double h, double l1, double l2)
{
    double const t = l1_thr(g, l1);
    double const d = h + l2;
    return d > 0.0 ? (t * t) / d : 0.0;
}
__device__ __forceinline__ float leaf_weight(double g, double h, double l1, double l2)
{
    double const d = h + l2;
    double const w = d > 0.0 ? -l1_thr(g, l1) / d : 0.0;
    return static_cast<float>(w);
}
__device__ __forceinline__ float node_value(NodeInfo const &n, GrowCfg const &c, double const *scales)
{
    if (c.mono) // write_leaf's bounded weight (src/step/primitives.hpp:42-50)
        return c.quantized ? fd_leaf_value(n.id, static_cast<double>(n.itg) * scales[0],
                                           static_cast<double>(n.ith) * scales[1], c)
                           : fd_leaf_value(n.id, n.tg, n.th, c);
    if (c.quantized)
        return leaf_weight(static_cast<double>(n.itg) * scales[0], static_cast<double>(n.ith) * scales[1], c.l1,
                           c.l2);
    return leaf_weight(n.tg, n.th, c.l1, c.l2);
}
__device__ __forceinline__ void put_leaf(DevNode *tree, u32 id, float v)
{
    DevNode n = tree[id];
    n.feature = -1;
    n.bin = n.default_left = n.left = n.right = 0;
    n.value = v;
    tree[id] = n;
}

// Node totals of a frontier (Histogram::totals, include/bonsai/histogram.hpp:76-87)
// from the first selected feature's cells (include/bonsai/split.hpp:36-53) when no
// parent split supplied them.  One workgroup per node.
template <bool Q>
__global__ __launch_bounds__(64) void k_totals(NodeInfo *__restrict__ front, TreeCtl const *__restrict__ ctl,
                                               void const *__restrict__ hist, u16 const *__restrict__ n_bins, u32 F,
                                               u32 SB, u8 const *__restrict__ fmask)
{
    u32 const i = blockIdx.x;
    if (i >= ctl->n_front || ctl->done) return;
    NodeInfo nd = front[i];
    if (nd.totals_ready) return;
    size_t const nc = static_cast<size_t>(F) * SB;
    u32 f0 = 0;
    if (fmask)
        while (f0 + 1u < F && !fmask[f0]) ++f0;
    u32 const nb0 = n_bins[f0];
    size_t const off = static_cast<size_t>(f0) * SB;
    if (Q) {
        i64 const *c = reinterpret_cast<i64 const *>(hist) + (static_cast<size_t>(nd.slot) * nc + off) * 2;
        i64 g = 0, h = 0;
        for (u32 b = threadIdx.x; b < nb0; b += 64) {
            g += c[2 * b];
            h += c[2 * b + 1];
        }
        for (int o = 32; o > 0; o >>= 1) {
            g += __shfl_xor(g, o);
            h += __shfl_xor(h, o);
        }
        if (threadIdx.x == 0) {
            nd.itg = g;
            nd.ith = h;
            nd.totals_ready = 1;
            front[i] = nd;
        }
    } else if (threadIdx.x == 0) {
        float2 const *c = reinterpret_cast<float2 const *>(hist) + static_cast<size_t>(nd.slot) * nc + off;
        double g = 0.0, h = 0.0;
        u32 b = 0;
        for (; b + 8 <= nb0; b += 8) {
            float2 v[8];
#pragma unroll
            for (int k = 0; k < 8; ++k) v[k] = c[b + k];
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                g += v[k].x;
                h += v[k].y;
            }
        }
        for (; b < nb0; ++b) {
            g += c[b].x;
            h += c[b].y;
        }
        nd.tg = static_cast<float>(g);
        nd.th = static_cast<float>(h);
        nd.totals_ready = 1;
        front[i] = nd;
    }
}

// ---------------------------------------------------------------------------------
// Level (oblivious) finder: HistogramLevelSplitFinder::find (src/split.cpp:164-235,
// :350-372).  One workgroup per feature.  Frontier nodes are taken in tiles: the tile's
// prefix sums (fill_prefix: double running sums stored as float, or exact integers) go
// to LDS, then each (bin, direction) candidate adds the tile's child scores to its
// running sum in node order; an infeasible node contributes its parent score.
// ---------------------------------------------------------------------------------
constexpr u32 kLevelTileF32 = 16, kLevelTileQ = 8;

template <bool Q>
__global__ __launch_bounds__(256) void k_find_level(NodeInfo const *__restrict__ front,
                                                    TreeCtl const *__restrict__ ctl, void const *__restrict__ hist,
                                                    u16 const *__restrict__ n_bins, u32 F, u32 SB, GrowCfg cfg,
                                                    double const *__restrict__ scales, SplitRec *__restrict__ fbest)
{
    extern __shared__ __align__(16) unsigned char smem[];
    u32 const f = blockIdx.x;
    if (f >= F) return;
    if (ctl->done) return;
    u32 const TN = Q ? kLevelTileQ : kLevelTileF32;
    u32 const P = ctl->n_front;
    u32 const nb = (cfg.fmask && !cfg.fmask[f]) ? 0u : n_bins[f]; // unselected: no cut
    u32 const ncut = nb >= 2 ? nb - 2 : 0;
    SplitRec best{};
    best.default_left = 1;
    if (ncut == 0 || P == 0) {
        if (threadIdx.x == 0) fbest[f] = best;
        return;
    }
    // LDS: per tile node prefix (g, h) over the cut bins + node scalars.
    float2 *pf = reinterpret_cast<float2 *>(smem);  // F32: [TN][ncut]
    i64 *pq = reinterpret_cast<i64 *>(smem);        // Q:   [TN][ncut][2]
    __shared__ double s_ps[kLevelTileF32], s_rg[kLevelTileF32], s_rh[kLevelTileF32], s_mg[kLevelTileF32],
        s_mh[kLevelTileF32];
    __shared__ i64 s_itg[kLevelTileF32], s_ith[kLevelTileF32], s_img[kLevelTileF32], s_imh[kLevelTileF32];
    __shared__ double s_sum_parent;
    __shared__ u32 s_nonempty;
    size_t const nc = static_cast<size_t>(F) * SB;
    double const sg = Q ? scales[0] : 0.0, sh = Q ? scales[1] : 0.0;
    // Two candidates per thread at most (ncut <= 254): k = threadIdx.x and + 256.
    double acc[2] = {0.0, 0.0};
    if (threadIdx.x == 0) {
        s_sum_parent = 0.0;
        s_nonempty = 0;
    }
    __syncthreads();
    for (u32 p0 = 0; p0 < P; p0 += TN) {
        u32 const tn = min(TN, P - p0);
        u32 const t = threadIdx.x;
        if (t < tn) {
            NodeInfo const nd = front[p0 + t];
            if (Q) {
                i64 const *c = reinterpret_cast<i64 const *>(hist) +
                               (static_cast<size_t>(nd.slot) * nc + static_cast<size_t>(f) * SB) * 2;
                i64 g = 0, h = 0;
                for (u32 b = 0; b < ncut; ++b) {
                    g += c[2 * b];
                    h += c[2 * b + 1];
                    pq[(t * ncut + b) * 2] = g;
                    pq[(t * ncut + b) * 2 + 1] = h;
                }
                i64 const mg = c[2 * (nb - 1)], mh = c[2 * (nb - 1) + 1];
                s_img[t] = mg;
                s_imh[t] = mh;
                s_itg[t] = nd.itg;
                s_ith[t] = nd.ith;
                s_ps[t] = score_of(static_cast<double>(nd.itg) * sg, static_cast<double>(nd.ith) * sh, cfg.l1,
                                   cfg.l2);
                if (mg != 0 || mh != 0) atomicOr(&s_nonempty, 1u);
            } else {
                float2 const *c = reinterpret_cast<float2 const *>(hist) + static_cast<size_t>(nd.slot) * nc +
                                  static_cast<size_t>(f) * SB;
                double g = 0.0, h = 0.0;
                for (u32 b = 0; b < ncut; ++b) {
                    g = g + c[b].x;
                    h = h + c[b].y;
                    pf[t * ncut + b] = make_float2(static_cast<float>(g), static_cast<float>(h));
                }
                float2 const m = c[nb - 1];
                s_mg[t] = m.x;
                s_mh[t] = m.y;
                s_ps[t] = score_of(nd.tg, nd.th, cfg.l1, cfg.l2);
                s_rg[t] = static_cast<double>(nd.tg) - m.x;
                s_rh[t] = static_cast<double>(nd.th) - m.y;
                if (m.x != 0.0F || m.y != 0.0F) atomicOr(&s_nonempty, 1u);
            }
        }
        __syncthreads();
        if (threadIdx.x == 0)
            for (u32 k = 0; k < tn; ++k) s_sum_parent += s_ps[k];
        for (u32 s = 0; s < 2; ++s) {
            u32 const k = threadIdx.x + s * blockDim.x;
            if (k >= 2 * ncut) continue;
            u32 const b = k >> 1;
            bool const dl = (k & 1u) == 0;
            double sum = acc[s];
            for (u32 p = 0; p < tn; ++p) {
                double cs;
                bool ok;
                if (Q) {
                    i64 const lg = pq[(p * ncut + b) * 2], lh = pq[(p * ncut + b) * 2 + 1];
                    i64 const mg = s_img[p], mh = s_imh[p];
                    i64 const hL = lh + (dl ? mh : 0);
                    i64 const hR = ((s_ith[p] - mh) - lh) + (!dl ? mh : 0);
                    i64 const gL = lg + (dl ? mg : 0);
                    i64 const gR = ((s_itg[p] - mg) - lg) + (!dl ? mg : 0);
                    double const dhL = static_cast<double>(hL) * sh, dhR = static_cast<double>(hR) * sh;
                    ok = !(dhL < cfg.min_child_hess || dhR < cfg.min_child_hess);
                    cs = ok ? score_of(static_cast<double>(gL) * sg, dhL, cfg.l1, cfg.l2) +
                                  score_of(static_cast<double>(gR) * sg, dhR, cfg.l1, cfg.l2)
                            : 0.0;
                } else {
                    float2 const lp = pf[p * ncut + b];
                    double const lg = lp.x, lh = lp.y;
                    double const mg = s_mg[p], mh = s_mh[p];
                    double const right_g = s_rg[p] - lg, right_h = s_rh[p] - lh;
                    double const gL = lg + (dl ? mg : 0.0), hL = lh + (dl ? mh : 0.0);
                    double const gR = right_g + (!dl ? mg : 0.0), hR = right_h + (!dl ? mh : 0.0);
                    ok = !(hL < cfg.min_child_hess || hR < cfg.min_child_hess);
                    cs = ok ? score_of(gL, hL, cfg.l1, cfg.l2) + score_of(gR, hR, cfg.l1, cfg.l2) : 0.0;
                }
                sum += ok ? cs : s_ps[p];
            }
            acc[s] = sum;
        }
        __syncthreads();
    }
    // Per-feature winner: max gain, earliest (bin, dl=true first); an all-empty missing
    // cell set scans one direction (identical sums, so the first direction wins anyway).
    __shared__ double s_g[256];
    __shared__ u32 s_k[256];
    double bg = 0.0;
    u32 bk = kNone;
    double const sp = s_sum_parent;
    bool const one_dir = s_nonempty == 0;
    for (u32 s = 0; s < 2; ++s) {
        u32 const k = threadIdx.x + s * blockDim.x;
        if (k >= 2 * ncut) continue;
        if (one_dir && (k & 1u)) continue;
        double const gain = acc[s] - sp;
        if (!(gain > 0.0 && gain >= cfg.min_gain)) continue;
        if (gain > bg || (gain == bg && k < bk)) {
            bg = gain;
            bk = k;
        }
    }
    s_g[threadIdx.x] = bg;
    s_k[threadIdx.x] = bk;
    __syncthreads();
    for (u32 st = blockDim.x / 2; st > 0; st >>= 1) {
        if (threadIdx.x < st) {
            double const og = s_g[threadIdx.x + st];
            u32 const ok = s_k[threadIdx.x + st];
            if (ok != kNone &&
                (s_k[threadIdx.x] == kNone || og > s_g[threadIdx.x] || (og == s_g[threadIdx.x] && ok < s_k[threadIdx.x]))) {
                s_g[threadIdx.x] = og;
                s_k[threadIdx.x] = ok;
            }
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        if (s_k[0] != kNone) {
            best.gain = s_g[0];
            best.feature = f;
            best.bin = s_k[0] >> 1;
            best.default_left = (s_k[0] & 1u) == 0 ? 1u : 0u;
            best.valid = 1;
        }
        fbest[f] = best;
    }
}

// reduce_in_feature_order (src/split.cpp:240-253) + the level's bookkeeping
// (src/grower_impl.hpp:633-645): record the split or stop growth.
__global__ __launch_bounds__(256) void k_level_pick(SplitRec const *__restrict__ fbest, u32 F,
                                                    TreeCtl *__restrict__ ctl, u32 *__restrict__ level_splits,
                                                    float *__restrict__ level_gains, SplitRec *__restrict__ pick)
{
    if (ctl->done) return;
    __shared__ double s_g[256];
    __shared__ u32 s_f[256];
    double bg = 0.0;
    u32 bf = kNone;
    for (u32 f = threadIdx.x; f < F; f += blockDim.x) {
        SplitRec const c = fbest[f];
        if (c.valid && (c.gain > bg || bf == kNone || (c.gain == bg && f < bf))) {
            if (bf == kNone || c.gain > bg || (c.gain == bg && f < bf)) {
                bg = c.gain;
                bf = f;
            }
        }
    }
    s_g[threadIdx.x] = bg;
    s_f[threadIdx.x] = bf;
    __syncthreads();
    for (u32 st = blockDim.x / 2; st > 0; st >>= 1) {
        if (threadIdx.x < st) {
            double const og = s_g[threadIdx.x + st];
            u32 const of = s_f[threadIdx.x + st];
            if (of != kNone &&
                (s_f[threadIdx.x] == kNone || og > s_g[threadIdx.x] || (og == s_g[threadIdx.x] && of < s_f[threadIdx.x]))) {
                s_g[threadIdx.x] = og;
                s_f[threadIdx.x] = of;
            }
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        u32 const d = ctl->oblivious_depth;
        if (s_f[0] == kNone) {
            ctl->done = 1;
            SplitRec z{};
            pick[0] = z;
            return;
        }
        SplitRec const s = fbest[s_f[0]];
        pick[0] = s;
        level_splits[3 * d] = s.feature;
        level_splits[3 * d + 1] = s.bin;
        level_splits[3 * d + 2] = s.default_left;
        level_gains[d] = static_cast<float>(s.gain);
        ctl->oblivious_depth = d + 1;
    }
}

// Broadcast the level split to every frontier node: one partition op per node (empty
// nodes included), the node's child sums under the split (for leaves made at the last
// level), and the partition units.
template <bool Q>
__global__ __launch_bounds__(1024) void k_plan_oblivious(NodeInfo const *__restrict__ front,
                                                         TreeCtl *__restrict__ ctl, SplitRec const *__restrict__ pick,
                                                         void const *__restrict__ hist, u16 const *__restrict__ n_bins,
                                                         u32 F, u32 SB, SplitRec *__restrict__ splits,
                                                         PartOp *__restrict__ ops, PartChunk *__restrict__ pch, u32 dst,
                                                         u32 is_last, u32 chunk_rows)
{
    __shared__ u32 scratch[1024];
    __shared__ u32 s_chunks;
    if (ctl->done) {
        if (threadIdx.x == 0) {
            ctl->n_ops = 0;
            ctl->n_part_chunks = 0;
        }
        return;
    }
    SplitRec const s = pick[0];
    u32 const P = ctl->n_front;
    u32 const nb = n_bins[s.feature];
    bool const dl = s.default_left != 0;
    size_t const nc = static_cast<size_t>(F) * SB;
    if (threadIdx.x == 0) s_chunks = 0;
    __syncthreads();
    for (u32 t0 = 0; t0 < P; t0 += blockDim.x) {
        u32 const i = t0 + threadIdx.x;
        bool const in = i < P;
        u32 nch = 0;
        NodeInfo nd{};
        if (in) {
            nd = front[i];
            nch = (nd.cnt + chunk_rows - 1) / chunk_rows;
        }
        u32 ctot;
        u32 const cfirst = block_exscan(nch, scratch, &ctot) + s_chunks;
        if (in) {
            SplitRec r = s;
            if (!is_last) {
                // child sums are only read for leaves made at the last level
            } else if (Q) {
                i64 const *c = reinterpret_cast<i64 const *>(hist) +
                               (static_cast<size_t>(nd.slot) * nc + static_cast<size_t>(s.feature) * SB) * 2;
                i64 lg = 0, lh = 0;
                for (u32 b = 0; b <= s.bin; ++b) {
                    lg += c[2 * b];
                    lh += c[2 * b + 1];
                }
                i64 const mg = c[2 * (nb - 1)], mh = c[2 * (nb - 1) + 1];
                r.igL = lg + (dl ? mg : 0);
                r.ihL = lh + (dl ? mh : 0);
                r.igR = ((nd.itg - mg) - lg) + (!dl ? mg : 0);
                r.ihR = ((nd.ith - mh) - lh) + (!dl ? mh : 0);
            } else {
                float2 const *c = reinterpret_cast<float2 const *>(hist) + static_cast<size_t>(nd.slot) * nc +
                                  static_cast<size_t>(s.feature) * SB;
                double g = 0.0, h = 0.0;
                for (u32 b = 0; b <= s.bin; ++b) {
                    g = g + c[b].x;
                    h = h + c[b].y;
                }
                double const lg = static_cast<float>(g), lh = static_cast<float>(h);
                double const mg = c[nb - 1].x, mh = c[nb - 1].y;
                double const rg = static_cast<double>(nd.tg) - mg, rh = static_cast<double>(nd.th) - mh;
                r.gL = lg + (dl ? mg : 0.0);
                r.hL = lh + (dl ? mh : 0.0);
                r.gR = (rg - lg) + (!dl ? mg : 0.0);
                r.hR = (rh - lh) + (!dl ? mh : 0.0);
            }
            splits[i] = r;
            PartOp op{};
            op.parent = i;
            op.feature = s.feature;
            op.bin = s.bin;
            op.default_left = s.default_left;
            op.last_bin = nb - 1u;
            op.off = nd.off;
            op.cnt = nd.cnt;
            op.src = nd.src;
            op.dst = dst;
            op.chunk_first = cfirst;
            op.chunk_count = nch;
            op.alive = 1;
            ops[i] = op;
        }
        __syncthreads();
        if (threadIdx.x == 0) s_chunks += ctot;
        __syncthreads();
    }
    u32 const n_ch = s_chunks;
    for (u32 c = threadIdx.x; c < n_ch; c += blockDim.x) {
        u32 lo = 0, hi = P;
        while (hi - lo > 1) {
            u32 const mid = (lo + hi) >> 1;
            if (ops[mid].chunk_first <= c) lo = mid;
            else hi = mid;
        }
        PartOp const op = ops[lo];
        u32 const k = c - op.chunk_first;
        PartChunk pc;
        pc.op = lo;
        pc.begin = k * chunk_rows;
        pc.end = min(op.cnt, (k + 1) * chunk_rows);
        pc.left_before = 0;
        pch[c] = pc;
    }
    if (threadIdx.x == 0) {
        ctl->n_ops = P;
        ctl->n_part_chunks = n_ch;
    }
}

// Level swap for the oblivious loop: a stopped tree keeps its frontier (copied into
// the buffer the next level reads).
__global__ __launch_bounds__(256) void k_next_level_obl(TreeCtl *__restrict__ ctl, NodeInfo const *__restrict__ cur,
                                                        NodeInfo *__restrict__ nxt)
{
    bool const done = ctl->done != 0;
    u32 const P = ctl->n_front;
    if (done)
        for (u32 i = threadIdx.x; i < P; i += blockDim.x) nxt[i] = cur[i];
    __syncthreads();
    if (threadIdx.x == 0) {
        if (!done) ctl->n_front = ctl->next_front;
        ctl->depth += 1;
    }
}

// The finished oblivious tree as a perfect dense tree (ObliviousTree ->
// dense_equivalent, src/tree.cpp:21-50): internal node i at level lvl(i) carries that
// level's split, children 2i+1 / 2i+2, leaves 2^D-1 .. in leaf-table order, covers
// aggregated bottom-up.
template <bool Q>
__global__ __launch_bounds__(1024) void k_oblivious_tree(NodeInfo const *__restrict__ front,
                                                         TreeCtl *__restrict__ ctl, u32 const *__restrict__ level_splits,
                                                         float const *__restrict__ level_gains, DevNode *__restrict__ tree,
                                                         GrowCfg cfg, double const *__restrict__ scales)
{
    u32 const D = ctl->oblivious_depth;
    u32 const n_int = (1u << D) - 1u, n_leaf = 1u << D;
    for (u32 j = threadIdx.x; j < n_leaf; j += blockDim.x) {
        NodeInfo const nd = front[j];
        DevNode n{};
        n.feature = -1;
        n.value = node_value(nd, cfg, scales);
        n.cover = static_cast<float>(nd.cnt);
        tree[n_int + j] = n;
    }
    __syncthreads();
    // Internal nodes level by level from the bottom.
    for (int lvl = static_cast<int>(D) - 1; lvl >= 0; --lvl) {
        u32 const first = (1u << lvl) - 1u, count = 1u << lvl;
        for (u32 k = threadIdx.x; k < count; k += blockDim.x) {
            u32 const i = first + k;
            DevNode n{};
            n.feature = static_cast<int32_t>(level_splits[3 * lvl]);
            n.bin = level_splits[3 * lvl + 1];
            n.default_left = level_splits[3 * lvl + 2];
            n.left = 2 * i + 1;
            n.right = 2 * i + 2;
            n.gain = level_gains[lvl];
            n.cover = tree[2 * i + 1].cover + tree[2 * i + 2].cover;
            tree[i] = n;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) ctl->n_nodes = n_int + n_leaf;
}

// ---------------------------------------------------------------------------------
// Leafwise rounds.
// ---------------------------------------------------------------------------------
struct LwCtl {
    u32 n_cand;
    u32 n_free;
    u32 parent_slot;
    u32 demoted;
    u32 small_left;
    u32 n_push;      // children of the last round that still have to join the pool
    u32 grid_done;   // workgroups of the running split find that have finished
    u32 pad[9];
};

// Rows per partition chunk of a leaf of `rows` rows: about kLwChunkTarget chunks,
// multiples of 256 rows, kLwMinChunk..kPartChunk, and at most cap (the byte-sized chunk
// of wide rows).
__device__ __forceinline__ u32 lw_chunk_rows(u32 rows, u32 cap)
{
    u32 const per = ((rows + kLwChunkTarget - 1u) / kLwChunkTarget + 255u) & ~255u;
    return min(cap, min(kPartChunk, max(kLwMinChunk, per)));
}

// Rows per histogram unit of a built node of `rows` rows: at most `target` units of
// kLwMinUnit..kHistChunk rows (multiples of 256); target 0 keeps kHistChunk.  FP32 sums
// run in ascending rows within a unit and over units in order, so this rule is part of
// the FP32 leafwise summation order (the oracle restates it).
__device__ __forceinline__ u32 lw_unit_rows(u32 rows, u32 target)
{
    if (target == 0) return kHistChunk;
    u32 const per = ((rows + target - 1u) / target + 255u) & ~255u;
    return min(kHistChunk, max(kLwMinUnit, per));
}

// queue_children (src/grower_impl.hpp:404-428): a child with a valid split joins the
// pool; one without becomes a leaf now and frees its slot.  One thread.
__device__ __forceinline__ void lw_push(LwCtl *lw, NodeInfo const *kids, SplitRec const *splits, u32 n, bool is_root, Cand *cands,
                        u32 *free_slots, DevNode *tree, GrowCfg const &cfg, double const *scales)
{
    for (u32 k = 0; k < n; ++k) {
        NodeInfo const nd = kids[k];
        SplitRec const s = splits[k];
        bool const ok = s.valid && (!is_root || cfg.max_depth > 0);
        if (ok) {
            Cand c;
            c.node = nd;
            c.split = s;
            cands[lw->n_cand++] = c;
        } else {
            put_leaf(tree, nd.id, node_value(nd, cfg, scales));
            if (nd.slot != kNone) free_slots[lw->n_free++] = nd.slot;
        }
    }
}

// The partitioned leaf's children (commit_pop, src/grower_impl.hpp:370-400), one
// thread: demotion when every row went one way; else two children, the larger keeping
// the parent's histogram slot (completed in place) and the smaller getting a fresh slot
// and the histogram units in *fo, unless the children sit at max_depth (then they only
// need totals, from the parent split's sums).  split_totals: the children's totals are
// the split's exact integer sums (quantized), so the finder need not derive them.
// Returns the smaller child's unit count (0: no histogram this round).
template <bool Q>
__device__ __forceinline__ u32 lw_children(PartOp &op, u32 nL, TreeCtl *ctl, LwCtl *lw, NodeInfo const &par, SplitRec const &s,
                           u32 *free_slots, NodeInfo *kids, DevNode *tree, GrowCfg const &cfg,
                           double const *scales, bool split_totals, u32 unit_target, FinishOp &fo,
                           u32 &unit_rows)
{
    u32 const nR = op.cnt - nL;
    op.n_left = nL;
    bool const demote = (nL == 0) != (nR == 0);
    op.alive = demote ? 0u : 1u;
    lw->n_push = 0;
    if (demote) {
        put_leaf(tree, par.id, node_value(par, cfg, scales));
        tree[par.id].gain = 0.0F;
        free_slots[lw->n_free++] = par.slot;
        ctl->n_front = 0;
        ctl->n_finish = 0;
        ctl->n_hist_chunks = 0;
        lw->demoted = 1;
        return 0;
    }
    lw->demoted = 0;
    tree[op.left_id].cover = static_cast<float>(nL);
    tree[op.right_id].cover = static_cast<float>(nR);
    NodeInfo L{}, R{};
    L.id = op.left_id;
    R.id = op.right_id;
    L.off = op.off;
    L.cnt = nL;
    R.off = op.off + nL;
    R.cnt = nR;
    L.src = R.src = op.dst;
    L.depth = R.depth = par.depth + 1;
    bool const left_small = nL <= nR;
    bool const leaves = static_cast<int>(L.depth) >= cfg.max_depth;
    ctl->live += 1;
    ctl->max_depth_seen = max(ctl->max_depth_seen, L.depth);
    if (leaves || split_totals) {
        L.totals_ready = R.totals_ready = 1;
        if (Q) {
            L.itg = s.igL;
            L.ith = s.ihL;
            R.itg = s.igR;
            R.ith = s.ihR;
        } else {
            L.tg = static_cast<float>(s.gL);
            L.th = static_cast<float>(s.hL);
            R.tg = static_cast<float>(s.gR);
            R.th = static_cast<float>(s.hR);
        }
    }
    u32 units = 0;
    if (leaves) {
        L.slot = R.slot = kNone;
        free_slots[lw->n_free++] = par.slot;
        ctl->n_finish = 0;
        ctl->n_hist_chunks = 0;
    } else {
        u32 const fresh = free_slots[--lw->n_free];
        L.slot = left_small ? fresh : par.slot;
        R.slot = left_small ? par.slot : fresh;
        u32 const small_cnt = left_small ? nL : nR;
        unit_rows = lw_unit_rows(small_cnt, unit_target);
        units = (small_cnt + unit_rows - 1) / unit_rows;
        fo = FinishOp{};
        fo.small_slot = fresh;
        fo.large_slot = par.slot;
        fo.parent_slot = par.slot;
        fo.part_first = 0;
        fo.part_count = units > 1 ? units : 0;
        fo.empty = small_cnt == 0;
        fo.unit_first = 0;
        fo.unit_count = units;
        fo.small_off = left_small ? L.off : R.off;
        fo.small_cnt = small_cnt;
        fo.src = op.dst;
        ctl->n_finish = 1;
        ctl->n_hist_chunks = units;
    }
    lw->small_left = left_small ? 1u : 0u;
    lw->n_push = 2;
    kids[0] = L;
    kids[1] = R;
    ctl->n_front = 2;
    return units;
}

// Histogram units of the smaller child: rows [off, off + cnt) of plane `src` in
// unit_rows pieces, each with its own partial when there is more than one.  u[] holds
// (slot, src, off, cnt) in LDS.
__device__ __forceinline__ void lw_expand_units(u32 const *u, u32 units, u32 unit_rows, HistChunk *hch)
{
    for (u32 k = threadIdx.x; k < units; k += blockDim.x) {
        HistChunk hc{};
        hc.slot = u[0];
        hc.src = u[1];
        hc.begin = u[2] + k * unit_rows;
        hc.end = u[2] + min(u[3], (k + 1) * unit_rows);
        hc.part = units > 1 ? k : kNone;
        hch[k] = hc;
    }
}

__device__ __forceinline__ void lw_stash_units(u32 *u, FinishOp const &fo)
{
    u[0] = fo.small_slot;
    u[1] = fo.src;
    u[2] = fo.small_off;
    u[3] = fo.small_cnt;
}

// Round start: the last round's children join the pool, then the best candidate is
// popped (max gain, ties to the lower node id; the reference's gain_less,
// src/grower_impl.hpp:763-770) unless the pool is empty or the leaf budget is spent,
// its split committed (commit_split_node, :188-210) and the one partition op emitted,
// in chunks sized to the leaf.  COUNTED (quantized): the split carries its child row
// counts, so the children, their slots and the histogram units are set here and the
// scatter takes each chunk's prefix from the chunk counts.  Runs as its own one-block
// kernel for the first round and in the last workgroup of every round's split find for
// the next; any workgroup size that is a multiple of 64.
template <bool COUNTED>
__device__ __forceinline__ void lw_select_round(TreeCtl *__restrict__ ctl, LwCtl *__restrict__ lw, Cand *__restrict__ cands,
                                NodeInfo *__restrict__ kids, SplitRec const *__restrict__ kid_splits,
                                u32 *__restrict__ free_slots, NodeInfo *__restrict__ parent_out,
                                SplitRec *__restrict__ split_out, DevNode *__restrict__ tree, PartOp *__restrict__ ops,
                                PartChunk *__restrict__ pch, HistChunk *__restrict__ hch,
                                u16 const *__restrict__ n_bins, FinishOp *__restrict__ fops, GrowCfg const &cfg,
                                double const *__restrict__ scales, u32 unit_target, u32 chunk_cap)
{
    __shared__ double s_wg[16];
    __shared__ u32 s_wid[16], s_wix[16];
    __shared__ u32 s_sel, s_go, s_n, s_units, s_urows, s_fo[4], s_ccnt, s_crow, s_nch;
    if (ctl->done) {
        if (threadIdx.x == 0) {
            ctl->n_ops = 0;
            ctl->n_part_chunks = 0;
            ctl->n_front = 0;
            ctl->n_finish = 0;
            ctl->n_hist_chunks = 0;
        }
        return;
    }
    if (threadIdx.x == 0) {
        u32 const np = lw->n_push;
        if (np) lw_push(lw, kids, kid_splits, np, false, cands, free_slots, tree, cfg, scales);
        lw->n_push = 0;
        s_n = lw->n_cand;
    }
    __syncthreads();
    u32 const n = s_n;
    double bg = 0.0;
    u32 bid = kNone, bix = kNone;
    for (u32 k = threadIdx.x; k < n; k += blockDim.x) {
        Cand const &c = cands[k];
        double const g = c.split.gain;
        u32 const id = c.node.id;
        if (bix == kNone || g > bg || (g == bg && id < bid)) {
            bg = g;
            bid = id;
            bix = k;
        }
    }
    // argmax (gain, then lower id: a total order, so the butterfly is exact) in each wave,
    // then over the wave winners in wave 0
    for (int o = 32; o > 0; o >>= 1) {
        double const og = __shfl_xor(bg, o);
        u32 const oid = __shfl_xor(bid, o), oix = __shfl_xor(bix, o);
        if (oix != kNone && (bix == kNone || og > bg || (og == bg && oid < bid))) {
            bg = og;
            bid = oid;
            bix = oix;
        }
    }
    if ((threadIdx.x & 63u) == 0) {
        s_wg[threadIdx.x >> 6] = bg;
        s_wid[threadIdx.x >> 6] = bid;
        s_wix[threadIdx.x >> 6] = bix;
    }
    __syncthreads();
    if (threadIdx.x < 64u) {
        u32 const nw = blockDim.x >> 6;
        bg = threadIdx.x < nw ? s_wg[threadIdx.x] : 0.0;
        bid = threadIdx.x < nw ? s_wid[threadIdx.x] : kNone;
        bix = threadIdx.x < nw ? s_wix[threadIdx.x] : kNone;
        for (int o = 32; o > 0; o >>= 1) {
            double const og = __shfl_xor(bg, o);
            u32 const oid = __shfl_xor(bid, o), oix = __shfl_xor(bix, o);
            if (oix != kNone && (bix == kNone || og > bg || (og == bg && oid < bid))) {
                bg = og;
                bid = oid;
                bix = oix;
            }
        }
    }
    if (threadIdx.x == 0) {
        bool const budget = cfg.max_leaves == 0 || ctl->live < static_cast<u32>(cfg.max_leaves);
        s_go = (n > 0 && budget) ? 1u : 0u;
        s_sel = bix;
        if (!s_go) {
            ctl->done = 1;
            ctl->n_ops = 0;
            ctl->n_part_chunks = 0;
            ctl->n_front = 0;
            ctl->n_finish = 0;
            ctl->n_hist_chunks = 0;
        }
    }
    __syncthreads();
    if (!s_go) return;
    if (threadIdx.x == 0) {
        Cand const c = cands[s_sel];
        u32 const crow = lw_chunk_rows(c.node.cnt, chunk_cap);
        u32 const nch = (c.node.cnt + crow - 1) / crow;
        s_ccnt = c.node.cnt;
        s_crow = crow;
        s_nch = nch;
        // remove from the pool (order inside the pool is irrelevant)
        cands[s_sel] = cands[n - 1];
        lw->n_cand = n - 1;
        u32 const lid = ctl->n_nodes, rid = lid + 1;
        ctl->n_nodes = lid + 2;
        DevNode pn = tree[c.node.id];
        pn.feature = static_cast<int32_t>(c.split.feature);
        pn.bin = c.split.bin;
        pn.default_left = c.split.default_left;
        pn.left = lid;
        pn.right = rid;
        pn.value = 0.0F;
        pn.gain = static_cast<float>(c.split.gain);
        tree[c.node.id] = pn;
        DevNode leaf{};
        leaf.feature = -1;
        tree[lid] = leaf;
        tree[rid] = leaf;
        parent_out[0] = c.node;
        split_out[0] = c.split;
        fd_propagate_bounds(cfg, scales, c.node.id, c.split, lid, rid);
        fd_propagate_path(cfg, c.node.id, c.split.feature, lid, rid);
        PartOp op{};
        op.parent = 0;
        op.feature = c.split.feature;
        op.bin = c.split.bin;
        op.default_left = c.split.default_left;
        op.last_bin = static_cast<u32>(n_bins[c.split.feature]) - 1u;
        op.off = c.node.off;
        op.cnt = c.node.cnt;
        op.src = c.node.src;
        op.dst = c.node.src == 0 ? 1u : 0u;
        op.chunk_first = 0;
        op.chunk_count = nch;
        op.left_id = lid;
        op.right_id = rid;
        op.alive = 1;
        lw->parent_slot = c.node.slot;
        if (COUNTED) {
            FinishOp fo{};
            u32 urows = kHistChunk;
            s_units = lw_children<true>(op, c.split.n_left, ctl, lw, c.node, c.split, free_slots, kids, tree, cfg,
                                        scales, true, unit_target, fo, urows);
            lw_stash_units(s_fo, fo);
            s_urows = urows;
            if (s_units) fops[0] = fo;
        }
        ops[0] = op;
        ctl->n_ops = 1;
        ctl->n_part_chunks = nch;
    }
    __syncthreads();
    u32 const crow = s_crow, nch = s_nch, ccnt = s_ccnt;
    for (u32 k = threadIdx.x; k < nch; k += blockDim.x) {
        PartChunk pc;
        pc.op = 0;
        pc.begin = k * crow;
        pc.end = min(ccnt, (k + 1) * crow);
        pc.left_before = 0;
        pch[k] = pc;
    }
    if (COUNTED) lw_expand_units(s_fo, s_units, s_urows, hch);
}

template <bool COUNTED>
__global__ __launch_bounds__(256) void k_lw_select(TreeCtl *__restrict__ ctl, LwCtl *__restrict__ lw,
                                                   Cand *__restrict__ cands, NodeInfo *__restrict__ kids,
                                                   SplitRec const *__restrict__ kid_splits,
                                                   u32 *__restrict__ free_slots, NodeInfo *__restrict__ parent_out,
                                                   SplitRec *__restrict__ split_out, DevNode *__restrict__ tree,
                                                   PartOp *__restrict__ ops, PartChunk *__restrict__ pch,
                                                   HistChunk *__restrict__ hch, u16 const *__restrict__ n_bins,
                                                   FinishOp *__restrict__ fops, GrowCfg cfg,
                                                   double const *__restrict__ scales, u32 unit_target, u32 chunk_cap)
{
    lw_select_round<COUNTED>(ctl, lw, cands, kids, kid_splits, free_slots, parent_out, split_out, tree, ops, pch, hch,
                             n_bins, fops, cfg, scales, unit_target, chunk_cap);
}

// Split find of a round's children (k_find's body) whose last workgroup to finish -- all
// others have stopped reading anything the select writes -- starts the next round
// (next != 0).
struct LwNext {
    Cand *cands;
    u32 *free_slots;
    NodeInfo *parent_out;
    SplitRec *split_out;
    DevNode *tree;
    PartOp *ops;
    PartChunk *pch;
    HistChunk *hch;
    FinishOp *fops;
    u32 unit_target;
    u32 chunk_cap;
    u32 next;
};

template <bool Q, bool COUNTED, bool CON>
__global__ __launch_bounds__(kFindThreads) void k_lw_find(NodeInfo *__restrict__ kids, TreeCtl *__restrict__ ctl,
                                                        LwCtl *__restrict__ lw, void const *__restrict__ hist,
                                                        u16 const *__restrict__ n_bins, u32 F, u32 SB, GrowCfg cfg,
                                                        double const *__restrict__ scales,
                                                        SplitRec *__restrict__ splits, SplitRec *__restrict__ part,
                                                        u32 *__restrict__ done, u32 const *__restrict__ cnt, LwNext nx)
{
    __shared__ u32 s_last_wg;
    find_node<Q, CON, kLwFindTile>(kids, ctl, hist, n_bins, F, SB, cfg, scales, splits, 0u, 0u, part, done, cnt);
    if (!nx.next) return;
    __syncthreads();
    if (threadIdx.x == 0) {
        __threadfence();
        s_last_wg = atomicAdd(&lw->grid_done, 1u) == gridDim.x * gridDim.y - 1u ? 1u : 0u;
    }
    __syncthreads();
    if (!s_last_wg) return;
    if (threadIdx.x == 0) {
        lw->grid_done = 0u;
        __threadfence();
    }
    __syncthreads();
    lw_select_round<COUNTED>(ctl, lw, nx.cands, kids, splits, nx.free_slots, nx.parent_out, nx.split_out, nx.tree,
                             nx.ops, nx.pch, nx.hch, n_bins, nx.fops, cfg, scales, nx.unit_target, nx.chunk_cap);
}

// Round middle of the FP32 grower: left count from the chunk prefix, then the children
// (lw_children) and the smaller child's histogram units.
template <bool Q>
__global__ __launch_bounds__(1024) void k_lw_scan(PartOp *__restrict__ ops, PartChunk *__restrict__ pch,
                                                  u32 const *__restrict__ chunk_left, TreeCtl *__restrict__ ctl,
                                                  LwCtl *__restrict__ lw, NodeInfo const *__restrict__ parent_in,
                                                  SplitRec const *__restrict__ split_in, u32 *__restrict__ free_slots,
                                                  NodeInfo *__restrict__ kids, DevNode *__restrict__ tree,
                                                  HistChunk *__restrict__ hch, FinishOp *__restrict__ fops,
                                                  GrowCfg cfg, double const *__restrict__ scales, u32 unit_target)
{
    __shared__ u32 scratch[1024];
    __shared__ u32 s_run, s_units, s_urows, s_fo[4];
    if (ctl->n_ops == 0) return;
    u32 const n_ch = ctl->n_part_chunks;
    if (threadIdx.x == 0) s_run = 0;
    __syncthreads();
    for (u32 t0 = 0; t0 < n_ch; t0 += blockDim.x) {
        u32 const c = t0 + threadIdx.x;
        u32 const v = c < n_ch ? chunk_left[c] : 0u;
        u32 tot;
        u32 const pre = block_exscan(v, scratch, &tot) + s_run;
        if (c < n_ch) pch[c].left_before = pre;
        __syncthreads();
        if (threadIdx.x == 0) s_run += tot;
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        PartOp op = ops[0];
        FinishOp fo{};
        u32 urows = kHistChunk;
        s_units = lw_children<Q>(op, s_run, ctl, lw, parent_in[0], split_in[0], free_slots, kids, tree, cfg, scales,
                                 false, unit_target, fo, urows);
        ops[0] = op;
        lw_stash_units(s_fo, fo);
        s_urows = urows;
        if (s_units) fops[0] = fo;
    }
    __syncthreads();
    lw_expand_units(s_fo, s_units, s_urows, hch);
}

// The root's split joins the pool (or the root becomes a leaf).
__global__ void k_lw_push_root(LwCtl *__restrict__ lw, NodeInfo const *__restrict__ kids,
                               SplitRec const *__restrict__ splits, Cand *__restrict__ cands,
                               u32 *__restrict__ free_slots, DevNode *__restrict__ tree, GrowCfg cfg,
                               double const *__restrict__ scales)
{
    if (threadIdx.x != 0) return;
    lw_push(lw, kids, splits, 1, true, cands, free_slots, tree, cfg, scales);
}

// Tree end: the last round's children join the pool, then what is left in the pool
// becomes leaves (src/grower_impl.hpp:838-846).
__global__ void k_lw_finalize(TreeCtl *__restrict__ ctl, LwCtl *__restrict__ lw, Cand *__restrict__ cands,
                              NodeInfo const *__restrict__ kids, SplitRec const *__restrict__ kid_splits,
                              u32 *__restrict__ free_slots, DevNode *__restrict__ tree, GrowCfg cfg,
                              double const *__restrict__ scales)
{
    __shared__ u32 s_n;
    if (threadIdx.x == 0) {
        u32 const np = lw->n_push;
        if (np) lw_push(lw, kids, kid_splits, np, false, cands, free_slots, tree, cfg, scales);
        lw->n_push = 0;
        s_n = lw->n_cand;
    }
    __syncthreads();
    u32 const n = s_n;
    for (u32 k = threadIdx.x; k < n; k += blockDim.x) put_leaf(tree, cands[k].node.id, node_value(cands[k].node, cfg, scales));
    (void)ctl;
}

// Leafwise tree start: root frontier (slot 0), free list 1.., pool empty.
__global__ __launch_bounds__(256) void k_lw_begin(TreeCtl *__restrict__ ctl, LwCtl *__restrict__ lw,
                                                  NodeInfo *__restrict__ kids, DevNode *__restrict__ tree,
                                                  HistChunk *__restrict__ hch, FinishOp *__restrict__ fops,
                                                  u32 *__restrict__ free_slots, u32 n_slots, RootSel root,
                                                  u32 unit_target)
{
    u32 const n_rows = root.n_from_ctl ? ctl->root_n : root.n;
    u32 const urows = lw_unit_rows(n_rows, unit_target);
    u32 const n_units = (n_rows + urows - 1) / urows;
    u32 const multi = n_units > 1;
    for (u32 k = threadIdx.x; k < n_units; k += blockDim.x) {
        HistChunk hc{};
        hc.slot = 0;
        hc.src = root.src;
        hc.begin = k * urows;
        hc.end = min(n_rows, (k + 1) * urows);
        hc.part = multi ? k : kNone;
        hch[k] = hc;
    }
    for (u32 k = threadIdx.x; k + 1 < n_slots; k += blockDim.x) free_slots[k] = n_slots - 1 - k;
    if (threadIdx.x == 0) {
        NodeInfo rn{};
        rn.cnt = n_rows;
        rn.src = root.src;
        rn.slot = 0;
        kids[0] = rn;
        DevNode r{};
        r.feature = -1;
        r.cover = static_cast<float>(n_rows);
        tree[0] = r;
        FinishOp fo{};
        fo.small_slot = 0;
        fo.large_slot = kNone;
        fo.part_count = multi ? n_units : 0;
        fo.empty = n_rows == 0;
        fops[0] = fo;
        ctl->n_front = 1;
        ctl->n_nodes = 1;
        ctl->n_ops = 0;
        ctl->n_part_chunks = 0;
        ctl->n_hist_chunks = n_units;
        ctl->n_finish = 1;
        ctl->done = 0;
        ctl->live = 1;
        ctl->max_depth_seen = 0;
        lw->n_cand = 0;
        lw->n_free = n_slots - 1;
        lw->n_push = 0;
        lw->grid_done = 0;
    }
}

__global__ void k_obl_begin(TreeCtl *ctl)
{
    if (threadIdx.x == 0) {
        ctl->done = 0;
        ctl->oblivious_depth = 0;
    }
}

} // namespace

hipError_t ext_alloc(ExtState &s, bh_fit_config const &f, u32 F, u32 SB, u32 max_front, u32 hist_slots)
{
    (void)SB;
    (void)max_front;
    s.n_features = F;
    hipError_t e;
    if (f.grower == BH_LEAFWISE) {
        s.cap = static_cast<u32>(f.max_leaves) + 8u;
        if ((e = hipMalloc(&s.cands, sizeof(Cand) * s.cap)) != hipSuccess) return e;
        if ((e = hipMalloc(&s.free_slots, sizeof(u32) * (hist_slots + 4))) != hipSuccess) return e;
        if ((e = hipMalloc(&s.pending, sizeof(LwCtl))) != hipSuccess) return e;
        if ((e = hipMalloc(&s.feat_best, sizeof(SplitRec) * 4)) != hipSuccess) return e; // parent split
    }
    if (f.grower == BH_LEVELWISE) {
        if ((e = hipMalloc(&s.feat_best, sizeof(SplitRec) * (F + 2))) != hipSuccess) return e;
        if ((e = hipMalloc(&s.level_splits, sizeof(u32) * 3 * 16)) != hipSuccess) return e;
        if ((e = hipMalloc(&s.level_gains, sizeof(float) * 16)) != hipSuccess) return e;
    }
    s.hist_slots = hist_slots;
    return hipSuccess;
}

void ext_free(ExtState &s)
{
    void *p[] = {s.cands, s.pending, s.free_slots, s.feat_best, s.level_splits, s.level_gains};
    for (void *q : p)
        if (q) (void)hipFree(q);
    s = ExtState{};
}

// The leafwise finders use the device's full per-workgroup LDS like k_find.
hipError_t ext_set_limits()
{
    int dev = 0, max_shared = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&max_shared, hipDeviceAttributeMaxSharedMemoryPerBlock, dev);
    if (e != hipSuccess) return e;
    void const *fns[] = {reinterpret_cast<void const *>(&k_lw_find<true, true, false>),
                         reinterpret_cast<void const *>(&k_lw_find<false, false, false>),
                         reinterpret_cast<void const *>(&k_lw_find<true, true, true>),
                         reinterpret_cast<void const *>(&k_lw_find<false, false, true>)};
    for (void const *fn : fns) {
        hipFuncAttributes a{};
        if (hipFuncGetAttributes(&a, fn) != hipSuccess) continue;
        int const dyn = max_shared - static_cast<int>(a.sharedSizeBytes);
        if (dyn > 0) (void)hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, dyn);
    }
    (void)hipGetLastError();
    return hipSuccess;
}

// ObliviousGrower::grow as a launch sequence (fixed shape: every level is enqueued, a
// stopped tree turns the remaining levels into no-ops).
void enqueue_levelwise(ExtState &s, hipStream_t st, bool q, GrowView const &v)
{
    u32 const D = static_cast<u32>(v.fc.max_depth);
    u32 const gh_bytes = q ? 4 : 8;
    SplitRec *pick = s.feat_best + v.F; // one spare record after the per-feature bests
    launch_tree_begin(st, v.ctl, v.front[0], v.tree, v.hch, v.fops, v.root);
    hipLaunchKernelGGL(k_obl_begin, dim3(1), dim3(64), 0, st, v.ctl);
    launch_hist(st, q, v.hch, v.ctl, v.work[0], v.work[1], v.plane0, v.gh[0], v.gh[1], v.gh[2], v.Fs, v.F, v.SB,
                v.fsl, v.waves, std::max(1u, std::min(v.max_hch, 4096u)), v.partials, v.hist[0], v.fx);
    launch_finish(st, q, v.fops, v.ctl, v.partials, v.hist[1], v.hist[0], v.F * v.SB, 1, nullptr, nullptr, false,
                  hist_transpose_F(q, v.Fs, v.F));
    u32 const TN = q ? kLevelTileQ : kLevelTileF32;
    size_t const lds_level = static_cast<size_t>(TN) * v.SB * (q ? 16 : 8);
    for (u32 d = 0; d < D; ++d) {
        NodeInfo *cur = v.front[d & 1u], *nxt = v.front[(d + 1) & 1u];
        void *hcur = v.hist[d & 1u], *hnext = v.hist[(d + 1) & 1u];
        u32 const width = 1u << d;
        bool const last = d + 1 == D;
        if (q)
            hipLaunchKernelGGL(k_totals<true>, dim3(width), dim3(64), 0, st, cur, v.ctl, hcur, v.d_nbins, v.F, v.SB,
                               v.gc.fmask);
        else
            hipLaunchKernelGGL(k_totals<false>, dim3(width), dim3(64), 0, st, cur, v.ctl, hcur, v.d_nbins, v.F, v.SB,
                               v.gc.fmask);
        if (q)
            hipLaunchKernelGGL(k_find_level<true>, dim3(v.F), dim3(256), lds_level, st, cur, v.ctl, hcur, v.d_nbins, v.F,
                               v.SB, v.gc, v.scales, s.feat_best);
        else
            hipLaunchKernelGGL(k_find_level<false>, dim3(v.F), dim3(256), lds_level, st, cur, v.ctl, hcur, v.d_nbins,
                               v.F, v.SB, v.gc, v.scales, s.feat_best);
        hipLaunchKernelGGL(k_level_pick, dim3(1), dim3(256), 0, st, s.feat_best, v.F, v.ctl, s.level_splits,
                           s.level_gains, pick);
        if (q)
            hipLaunchKernelGGL(k_plan_oblivious<true>, dim3(1), dim3(1024), 0, st, cur, v.ctl, pick, hcur, v.d_nbins,
                               v.F, v.SB, v.splits, v.ops, v.pch, d & 1u, last ? 1u : 0u, part_chunk_rows(v.Fs));
        else
            hipLaunchKernelGGL(k_plan_oblivious<false>, dim3(1), dim3(1024), 0, st, cur, v.ctl, pick, hcur,
                               v.d_nbins, v.F, v.SB, v.splits, v.ops, v.pch, d & 1u, last ? 1u : 0u,
                               part_chunk_rows(v.Fs));
        launch_part_count(st, v.pch, v.ops, v.ctl, v.work[0], v.work[1], v.plane0, v.Fs, v.chunk_left,
                          std::max(1u, std::min(v.max_pch, tuning().part_grid)));
        launch_part_scan(st, q, v.ops, v.pch, v.chunk_left, v.ctl, cur, nxt, v.splits, v.tree, v.hch, v.fops, v.gc,
                         v.scales, last ? 1u : 0u, 1u);
        if (!last) {
            launch_part_scatter(st, v.pch, v.ops, v.ctl, v.work[0], v.work[1], v.plane0, v.gh[0], v.gh[1], v.gh[2],
                                v.work[0], v.work[1], const_cast<void *>(v.gh[0]), const_cast<void *>(v.gh[1]), v.Fs,
                                gh_bytes, std::max(1u, std::min(v.max_pch, tuning().part_grid)));
            launch_hist(st, q, v.hch, v.ctl, v.work[0], v.work[1], v.plane0, v.gh[0], v.gh[1], v.gh[2], v.Fs, v.F,
                        v.SB, v.fsl, v.waves, std::max(1u, std::min(v.max_hch, 4096u)), v.partials, hnext, v.fx);
            launch_finish(st, q, v.fops, v.ctl, v.partials, hcur, hnext, v.F * v.SB, std::min(width, v.max_ops),
                          nullptr, nullptr, false, hist_transpose_F(q, v.Fs, v.F));
        }
        hipLaunchKernelGGL(k_next_level_obl, dim3(1), dim3(256), 0, st, v.ctl, cur, nxt);
        if (v.host_sync) (void)hipStreamSynchronize(st);
    }
    if (q)
        hipLaunchKernelGGL(k_oblivious_tree<true>, dim3(1), dim3(1024), 0, st, v.front[D & 1u], v.ctl, s.level_splits,
                           s.level_gains, v.tree, v.gc, v.scales);
    else
        hipLaunchKernelGGL(k_oblivious_tree<false>, dim3(1), dim3(1024), 0, st, v.front[D & 1u], v.ctl,
                           s.level_splits, s.level_gains, v.tree, v.gc, v.scales);
}

// find_level_split as one step over a frontier the host laid out (front[0 .. n) with
// ctl->n_front = n): node totals, the per-feature level scan, the pick (pick[0], valid
// = 0 when no split is feasible), and every node's child sums under the pick in
// v.splits[i] (k_plan_oblivious's last-level arm; its partition ops are not used).
void launch_level_find_step(ExtState &s, hipStream_t st, bool q, GrowView const &v, NodeInfo *front, u32 n,
                            void const *hist, SplitRec *pick)
{
    u32 const TN = q ? kLevelTileQ : kLevelTileF32;
    size_t const lds_level = static_cast<size_t>(TN) * v.SB * (q ? 16 : 8);
    hipLaunchKernelGGL(k_obl_begin, dim3(1), dim3(64), 0, st, v.ctl);
    if (q) {
        hipLaunchKernelGGL(k_totals<true>, dim3(n), dim3(64), 0, st, front, v.ctl, hist, v.d_nbins, v.F, v.SB,
                           v.gc.fmask);
        hipLaunchKernelGGL(k_find_level<true>, dim3(v.F), dim3(256), lds_level, st, front, v.ctl, hist, v.d_nbins,
                           v.F, v.SB, v.gc, v.scales, s.feat_best);
    } else {
        hipLaunchKernelGGL(k_totals<false>, dim3(n), dim3(64), 0, st, front, v.ctl, hist, v.d_nbins, v.F, v.SB,
                           v.gc.fmask);
        hipLaunchKernelGGL(k_find_level<false>, dim3(v.F), dim3(256), lds_level, st, front, v.ctl, hist, v.d_nbins,
                           v.F, v.SB, v.gc, v.scales, s.feat_best);
    }
    hipLaunchKernelGGL(k_level_pick, dim3(1), dim3(256), 0, st, s.feat_best, v.F, v.ctl, s.level_splits, s.level_gains,
                       pick);
    if (q)
        hipLaunchKernelGGL(k_plan_oblivious<true>, dim3(1), dim3(1024), 0, st, front, v.ctl, pick, hist, v.d_nbins, v.F,
                           v.SB, v.splits, v.ops, v.pch, 0u, 1u, part_chunk_rows(v.Fs));
    else
        hipLaunchKernelGGL(k_plan_oblivious<false>, dim3(1), dim3(1024), 0, st, front, v.ctl, pick, hist, v.d_nbins,
                           v.F, v.SB, v.splits, v.ops, v.pch, 0u, 1u, part_chunk_rows(v.Fs));
}

// LeafwiseGrower::grow as a launch sequence: the root, then max_leaves - 1 rounds
// (plus a margin for demoted splits, which spend a round without adding a leaf); rounds
// after the heap empties or the budget is spent are no-ops.
template <bool Q, bool COUNTED>
void launch_lw_find(hipStream_t st, GrowView const &v, LwCtl *lw, u32 *cnt, LwNext const &nx)
{
    size_t const lds = static_cast<size_t>(kLwFindTile) * v.SB * (Q ? (cnt ? 36 : 32) : 24);
    dim3 const grid(find_blocks_per_node(v.F, kLwFindTile), 2);
    bool const con = v.gc.fmask || v.gc.mono || v.gc.n_groups;
    auto const fn = con ? &k_lw_find<Q, COUNTED, true> : &k_lw_find<Q, COUNTED, false>;
    hipLaunchKernelGGL(fn, grid, dim3(kFindThreads), lds, st, v.front[0], v.ctl, lw, v.hist[0], v.d_nbins, v.F, v.SB,
                       v.gc, v.scales, v.splits, v.fscr.part, v.fscr.done, cnt, nx);
}

void enqueue_leafwise(ExtState &s, hipStream_t st, bool q, GrowView const &v)
{
    u32 const gh_bytes = q ? 4 : 8;
    LwCtl *lw = reinterpret_cast<LwCtl *>(s.pending);
    NodeInfo *kids = v.front[0];
    NodeInfo *parent = v.front[1];
    SplitRec *psplit = s.feat_best;
    void *pool = v.hist[0];
    // Quantized: count cells beside the histogram pool give every split its child row
    // counts (the counted round); FP32: the chunk counts are scanned on the device.
    u32 *cnt = q ? v.cnt : nullptr;
    u32 const pgrid = std::max(1u, std::min(v.max_pch, tuning().part_grid));
    u32 const hgrid = std::max(1u, std::min(v.max_hch, 4096u));
    hipLaunchKernelGGL(k_lw_begin, dim3(1), dim3(256), 0, st, v.ctl, lw, kids, v.tree, v.hch, v.fops, s.free_slots,
                       s.hist_slots, v.root, v.lw_units);
    launch_hist(st, q, v.hch, v.ctl, v.work[0], v.work[1], v.plane0, v.gh[0], v.gh[1], v.gh[2], v.Fs, v.F, v.SB,
                v.fsl, v.waves, hgrid, v.partials, pool, v.fx, cnt);
    launch_finish(st, q, v.fops, v.ctl, v.partials, pool, pool, v.F * v.SB, 1, cnt, cnt, false,
                      hist_transpose_F(q, v.Fs, v.F));
    launch_find(st, q, kids, v.ctl, pool, v.d_nbins, v.F, v.SB, v.gc, v.scales, v.splits, 2, 0, 0, v.fscr, cnt);
    hipLaunchKernelGGL(k_lw_push_root, dim3(1), dim3(64), 0, st, lw, kids, v.splits, s.cands, s.free_slots, v.tree,
                       v.gc, v.scales);
    // The first round's select is a kernel of its own; every later one runs in the last
    // workgroup of the previous round's split find.
    if (q)
        hipLaunchKernelGGL(k_lw_select<true>, dim3(1), dim3(256), 0, st, v.ctl, lw, s.cands, kids, v.splits,
                           s.free_slots, parent, psplit, v.tree, v.ops, v.pch, v.hch, v.d_nbins, v.fops, v.gc, v.scales,
                           v.lw_units, part_chunk_rows(v.Fs));
    else
        hipLaunchKernelGGL(k_lw_select<false>, dim3(1), dim3(256), 0, st, v.ctl, lw, s.cands, kids, v.splits,
                           s.free_slots, parent, psplit, v.tree, v.ops, v.pch, v.hch, v.d_nbins, v.fops, v.gc, v.scales,
                           0u, part_chunk_rows(v.Fs));
    u32 const rounds = static_cast<u32>(v.fc.max_leaves) - 1u + 8u;
    for (u32 r = 0; r < rounds; ++r) {
        launch_part_count(st, v.pch, v.ops, v.ctl, v.work[0], v.work[1], v.plane0, v.Fs, v.chunk_left, pgrid);
        if (q) {
            launch_part_scatter(st, v.pch, v.ops, v.ctl, v.work[0], v.work[1], v.plane0, v.gh[0], v.gh[1],
                                v.gh[2], v.work[0], v.work[1], const_cast<void *>(v.gh[0]),
                                const_cast<void *>(v.gh[1]), v.Fs, gh_bytes, pgrid, v.chunk_left);
        } else {
            hipLaunchKernelGGL(k_lw_scan<false>, dim3(1), dim3(1024), 0, st, v.ops, v.pch, v.chunk_left, v.ctl, lw,
                               parent, psplit, s.free_slots, kids, v.tree, v.hch, v.fops, v.gc, v.scales, v.lw_units);
            launch_part_scatter(st, v.pch, v.ops, v.ctl, v.work[0], v.work[1], v.plane0, v.gh[0], v.gh[1],
                                v.gh[2], v.work[0], v.work[1], const_cast<void *>(v.gh[0]),
                                const_cast<void *>(v.gh[1]), v.Fs, gh_bytes, pgrid);
        }
        launch_hist(st, q, v.hch, v.ctl, v.work[0], v.work[1], v.plane0, v.gh[0], v.gh[1], v.gh[2], v.Fs, v.F, v.SB,
                    v.fsl, v.waves, hgrid, v.partials, pool, v.fx, cnt);
        launch_finish(st, q, v.fops, v.ctl, v.partials, pool, pool, v.F * v.SB, 1, cnt, cnt, false,
                      hist_transpose_F(q, v.Fs, v.F));
        LwNext const nx{s.cands, s.free_slots, parent, psplit, v.tree, v.ops, v.pch, v.hch, v.fops, v.lw_units,
                        part_chunk_rows(v.Fs), r + 1 < rounds ? 1u : 0u};
        if (q) launch_lw_find<true, true>(st, v, lw, cnt, nx);
        else launch_lw_find<false, false>(st, v, lw, cnt, nx);
        if (v.host_sync) (void)hipStreamSynchronize(st);
    }
    hipLaunchKernelGGL(k_lw_finalize, dim3(1), dim3(256), 0, st, v.ctl, lw, s.cands, kids, v.splits, s.free_slots,
                       v.tree, v.gc, v.scales);
}

} // namespace bh
