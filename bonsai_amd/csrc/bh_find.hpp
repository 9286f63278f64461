// Node split find (the body of k_find, shared with the leafwise finder that also starts
// the next round).  The algorithm and its reference are described at k_find.
#pragma once
#include <hip/hip_runtime.h>
#include "bh_common.hpp"

namespace bh {

constexpr u32 kFindTile = 7;
// The leafwise finder's tile: its two nodes per launch leave most of the chip idle, so
// smaller tiles (7 workgroups per node at 28 features) shorten the critical path.
constexpr u32 kLwFindTile = 4;
// Threads per split-find workgroup: 16 waves, so each SIMD interleaves four waves of the
// FP64 candidate scoring.
constexpr u32 kFindThreads = 1024;

template <bool Q> struct CellT;
template <> struct CellT<false> { typedef float2 T; typedef double A; };
template <> struct CellT<true> { struct T { i64 x, y; }; typedef i64 A; };

__device__ __forceinline__ double fd_l1_thr(double g, double l1)
{
    if (g > l1) return g - l1;
    if (g < -l1) return g + l1;
    return 0.0;
}
// include/bonsai/split.hpp:135-143
__device__ __forceinline__ double fd_score_of(double g, double h, double l1, double l2)
{
    double const t = fd_l1_thr(g, l1);
    double const d = h + l2;
    return d > 0.0 ? (t * t) / d : 0.0;
}

// bounded_leaf_weight (include/bonsai/split.hpp:141-150): the Newton weight clamped to
// [lo, hi] in std::clamp's comparison order.
__device__ __forceinline__ double fd_bounded_weight(double g, double h, double l1, double l2, double lo, double hi)
{
    double const d = h + l2;
    double const w = d > 0.0 ? -fd_l1_thr(g, l1) / d : 0.0;
    return w < lo ? lo : (hi < w ? hi : w);
}

// propagate_monotone_bounds (src/grower_impl.hpp:47-75): both children inherit the
// parent's box; a split on a monotone feature fences them at the midpoint of the two
// child weights, each bounded by the parent's box.  The child sums are the ones the
// finder scored (exact integers times the scales when quantized).
__device__ __forceinline__ void fd_propagate_bounds(GrowCfg const &c, double const *__restrict__ scales, u32 pid,
                                                    SplitRec const &s, u32 lid, u32 rid)
{
    if (!c.mono) return;
    double const plo = c.bounds[2 * pid], phi = c.bounds[2 * pid + 1];
    double llo = plo, lhi = phi, rlo = plo, rhi = phi;
    int const mc = c.mono[s.feature];
    if (mc != 0) {
        double gL = s.gL, hL = s.hL, gR = s.gR, hR = s.hR;
        if (c.quantized) {
            gL = static_cast<double>(s.igL) * scales[0];
            hL = static_cast<double>(s.ihL) * scales[1];
            gR = static_cast<double>(s.igR) * scales[0];
            hR = static_cast<double>(s.ihR) * scales[1];
        }
        double const wL = fd_bounded_weight(gL, hL, c.l1, c.l2, plo, phi);
        double const wR = fd_bounded_weight(gR, hR, c.l1, c.l2, plo, phi);
        double const mid = 0.5 * (wL + wR);
        if (mc > 0) {
            lhi = mid < lhi ? mid : lhi;
            rlo = rlo < mid ? mid : rlo;
        } else {
            llo = llo < mid ? mid : llo;
            rhi = mid < rhi ? mid : rhi;
        }
    }
    c.bounds[2 * lid] = llo;
    c.bounds[2 * lid + 1] = lhi;
    c.bounds[2 * rid] = rlo;
    c.bounds[2 * rid + 1] = rhi;
}

// propagate_interaction_state (src/grower_impl.hpp:143-161): both children's path is
// the parent's plus the split feature.
__device__ __forceinline__ void fd_propagate_path(GrowCfg const &c, u32 pid, u32 feature, u32 lid, u32 rid)
{
    if (!c.n_groups) return;
    u32 const W = c.path_words;
    for (u32 w = 0; w < W; ++w) {
        u32 const v = c.paths[pid * W + w] | (w == (feature >> 5) ? 1u << (feature & 31u) : 0u);
        c.paths[lid * W + w] = v;
        c.paths[rid * W + w] = v;
    }
}

// A node's leaf weight as float, clamped to its box when constraints are on.
__device__ __forceinline__ float fd_leaf_value(u32 id, double g, double h, GrowCfg const &c)
{
    double lo = -__builtin_huge_val(), hi = __builtin_huge_val();
    if (c.mono) {
        lo = c.bounds[2 * id];
        hi = c.bounds[2 * id + 1];
    }
    return static_cast<float>(fd_bounded_weight(g, h, c.l1, c.l2, lo, hi));
}

// CON: the feature mask and the monotone / interaction constraints are compiled in
// (launches without any use the leaner CON = false body).
template <bool Q, bool CON, u32 TILE = kFindTile>
__device__ void find_node(NodeInfo *__restrict__ front, TreeCtl const *__restrict__ ctl, void const *__restrict__ hist,
                          u16 const *__restrict__ n_bins, u32 F, u32 SB, GrowCfg cfg, double const *__restrict__ scales,
                          SplitRec *__restrict__ out, u32 front_base, u32 n_front_override,
                          SplitRec *__restrict__ part, u32 *__restrict__ done, u32 const *__restrict__ cnt)
{
    typedef typename CellT<Q>::T Cell;
    typedef typename CellT<Q>::A Acc;
    extern __shared__ __align__(16) unsigned char smem[];
    Cell *s_cells = reinterpret_cast<Cell *>(smem);              // [TILE][SB]
    Acc *pg = reinterpret_cast<Acc *>(s_cells + TILE * SB); // [TILE][SB]
    Acc *ph = pg + TILE * SB;                                // [TILE][SB]
    u32 *s_cn = reinterpret_cast<u32 *>(ph + TILE * SB);    // [TILE][SB] counts, then their prefix
    __shared__ double s_best_gain[kFindThreads / 64];
    __shared__ u32 s_best_key[kFindThreads / 64];
    __shared__ double s_tot[2];
    __shared__ i64 s_itot[2];
    __shared__ Acc s_mg[TILE], s_mh[TILE];
    __shared__ u32 s_ncut[TILE], s_dirs[TILE], s_nb[TILE], s_mc[TILE];
    __shared__ u32 s_allow[kMaxPathWords], s_cover[kMaxGroups], s_single;

    // Every launch's frontier buffer holds gridDim.y entries past front_base, so the
    // node loads beside the frontier size.
    u32 const i = front_base + blockIdx.y;
    NodeInfo node = front[i];
    u32 const n_front = n_front_override ? n_front_override : ctl->n_front;
    if (blockIdx.y >= n_front) return;
    u32 const G = gridDim.x, bx = blockIdx.x;
    if (node.slot == kNone) {
        // No histogram (a leafwise child at max_depth): totals came from the parent
        // split and the node cannot split.
        if (bx == 0 && threadIdx.x == 0) {
            SplitRec rec{};
            rec.default_left = 1;
            out[i] = rec;
        }
        return;
    }
    size_t const node_cells = static_cast<size_t>(F) * SB;
    Cell const *hcells = reinterpret_cast<Cell const *>(hist) + static_cast<size_t>(node.slot) * node_cells;
    u32 const *hcnt = cnt ? cnt + static_cast<size_t>(node.slot) * node_cells : nullptr;
    double const sg = Q ? scales[0] : 0.0, sh = Q ? scales[1] : 0.0;
    bool const can = node.cnt >= 2u * static_cast<u32>(cfg.min_data_in_leaf) &&
                     static_cast<int>(node.depth) < cfg.max_depth;

    double best_gain = 0.0;
    u32 best_key = kNone;
    double bgL = 0.0, bhL = 0.0, bgR = 0.0, bhR = 0.0; // winner's child sums (FP32 path)
    i64 bigL = 0, bihL = 0, bigR = 0, bihR = 0;        // winner's child sums (quantized)
    u32 bnl = 0;                                       // winner's left rows (count cells)
    __shared__ u32 s_last;
    double node_score = 0.0, tg = 0.0, th = 0.0;
    double node_lo = -__builtin_huge_val(), node_hi = __builtin_huge_val(); // monotone box
    if (CON && cfg.mono) {
        node_lo = cfg.bounds[2 * node.id];
        node_hi = cfg.bounds[2 * node.id + 1];
    }
    bool const had_totals = node.totals_ready != 0;
    // Quantized totals are exact integers, equal over every feature's cells: a block that
    // scans takes them from the first feature of its first tile (no extra round trip).
    bool const tot_from_tile = Q && can && !had_totals;
    // Phase 0: node totals from the first selected feature's cells (Histogram::totals,
    // include/bonsai/histogram.hpp:76-87; SplitInput::totals skips unselected features,
    // include/bonsai/split.hpp:36-53) unless a parent split already set them.
    if (!had_totals && !tot_from_tile) {
        u32 f0 = 0;
        if (CON && cfg.fmask)
            while (f0 + 1u < F && !cfg.fmask[f0]) ++f0;
        u32 const nb0 = n_bins[f0];
        for (u32 k = threadIdx.x; k < nb0; k += blockDim.x) s_cells[k] = hcells[static_cast<size_t>(f0) * SB + k];
        __syncthreads();
        if (Q) {
            // Exact integer totals: any order (one wave reduces feature 0's cells).
            if (threadIdx.x < 64) {
                i64 g = 0, h = 0;
                for (u32 b = threadIdx.x; b < nb0; b += 64) {
                    g += reinterpret_cast<i64 const *>(&s_cells[b])[0];
                    h += reinterpret_cast<i64 const *>(&s_cells[b])[1];
                }
                for (int o = 32; o > 0; o >>= 1) {
                    g += __shfl_xor(g, o);
                    h += __shfl_xor(h, o);
                }
                if (threadIdx.x == 0) {
                    node.itg = g;
                    node.ith = h;
                }
            }
        } else if (threadIdx.x == 0) {
            // Histogram::totals: the first feature's cells, double in bin order, stored
            // as float.
            double g = 0.0, h = 0.0;
            float2 const *cf = reinterpret_cast<float2 const *>(s_cells);
            u32 b = 0;
            for (; b + 8 <= nb0; b += 8) {
                float2 v[8];
#pragma unroll
                for (int k = 0; k < 8; ++k) v[k] = cf[b + k];
#pragma unroll
                for (int k = 0; k < 8; ++k) {
                    g += v[k].x;
                    h += v[k].y;
                }
            }
            for (; b < nb0; ++b) {
                g += cf[b].x;
                h += cf[b].y;
            }
            node.tg = static_cast<float>(g);
            node.th = static_cast<float>(h);
        }
        if (threadIdx.x == 0) node.totals_ready = 1;
    }
    if (threadIdx.x == 0) {
        s_tot[0] = node.tg;
        s_tot[1] = node.th;
        s_itot[0] = node.itg;
        s_itot[1] = node.ith;
    }
    __syncthreads();
    if (!can) {
        if (bx == 0 && threadIdx.x == 0) {
            if (!had_totals) front[i] = node;
            SplitRec rec{};
            rec.default_left = 1;
            out[i] = rec;
        }
        return;
    }
    // Interaction constraints: the features this node may split on (allowed_features,
    // src/grower_impl.hpp:111-137): the union of every group that contains the node's
    // whole path, plus the path's feature when the path has one; an empty path allows all.
    u32 const W = cfg.path_words;
    if (CON && cfg.n_groups) {
        u32 const *path = cfg.paths + static_cast<size_t>(node.id) * W;
        if (threadIdx.x == 0) {
            u32 pc = 0;
            for (u32 w = 0; w < W; ++w) pc += __popc(path[w]);
            s_single = pc;
        }
        for (u32 g = threadIdx.x; g < cfg.n_groups; g += blockDim.x) {
            u32 const *grp = cfg.groups + static_cast<size_t>(g) * W;
            u32 miss = 0;
            for (u32 w = 0; w < W; ++w) miss |= path[w] & ~grp[w];
            s_cover[g] = miss == 0u;
        }
        __syncthreads();
        for (u32 w = threadIdx.x; w < W; w += blockDim.x) {
            u32 a = 0;
            if (s_single == 0) {
                a = ~0u;
            } else {
                for (u32 g = 0; g < cfg.n_groups; ++g)
                    if (s_cover[g]) a |= cfg.groups[static_cast<size_t>(g) * W + w];
                if (s_single == 1) a |= path[w];
            }
            s_allow[w] = a;
        }
        // (the tile loop's first barrier publishes s_allow)
    }
    // Block bx scans feature tiles bx, bx + G, ...
    for (u32 f0 = bx * TILE; f0 < F; f0 += G * TILE) {
        u32 const nt = min(TILE, F - f0);
        __syncthreads(); // the previous tile (or phase 0) is done with the LDS
        // Stage the tile's cells, bin counts and count cells (coalesced, one round trip).
        // An unselected or disallowed feature offers no cut (zero bins).
        if (threadIdx.x < nt) {
            u32 const f = f0 + threadIdx.x;
            bool const off = CON && ((cfg.fmask && !cfg.fmask[f]) ||
                             (cfg.n_groups && !((s_allow[f >> 5] >> (f & 31u)) & 1u)));
            s_nb[threadIdx.x] = off ? 0u : n_bins[f];
        }
        for (u32 k = threadIdx.x; k < nt * SB; k += blockDim.x) s_cells[k] = hcells[static_cast<size_t>(f0) * SB + k];
        if (hcnt)
            for (u32 k = threadIdx.x; k < nt * SB; k += blockDim.x) s_cn[k] = hcnt[static_cast<size_t>(f0) * SB + k];
        __syncthreads();
        if (tot_from_tile && f0 == bx * TILE) {
            if (threadIdx.x < 64) {
                i64 g = 0, h = 0;
                for (u32 b = threadIdx.x; b < SB; b += 64) {
                    g += reinterpret_cast<i64 const *>(&s_cells[b])[0];
                    h += reinterpret_cast<i64 const *>(&s_cells[b])[1];
                }
                for (int o = 32; o > 0; o >>= 1) {
                    g += __shfl_xor(g, o);
                    h += __shfl_xor(h, o);
                }
                if (threadIdx.x == 0) {
                    node.itg = g;
                    node.ith = h;
                    node.totals_ready = 1;
                    s_itot[0] = g;
                    s_itot[1] = h;
                }
            }
            __syncthreads();
        }
        if (f0 == bx * TILE) {
            tg = Q ? static_cast<double>(s_itot[0]) * sg : s_tot[0];
            th = Q ? static_cast<double>(s_itot[1]) * sh : s_tot[1];
            node_score = fd_score_of(tg, th, cfg.l1, cfg.l2);
        }
        // Phase A: per-feature running sums in bin order (src/split.cpp:131-158).
        // Quantized: exact integers, so a wave scans a feature in parallel (the count
        // cells alongside, in place).
        if (Q) {
            u32 const wv = threadIdx.x >> 6, ln = threadIdx.x & 63u;
            for (u32 t = wv; t < TILE; t += blockDim.x >> 6) {
                u32 ncut = 0, dirs = 1;
                if (t < nt) {
                    u32 const nb = s_nb[t];
                    ncut = nb >= 2 ? nb - 2 : 0;
                    i64 const *ci = reinterpret_cast<i64 const *>(s_cells + t * SB);
                    u32 *cn = s_cn + t * SB;
                    i64 const mg = nb ? ci[2 * (nb - 1)] : 0, mh = nb ? ci[2 * (nb - 1) + 1] : 0;
                    dirs = (mg == 0 && mh == 0) ? 1 : 2;
                    if (ln == 0) {
                        s_mg[t] = mg;
                        s_mh[t] = mh;
                        s_mc[t] = hcnt && nb ? cn[nb - 1] : 0u;
                    }
                    u32 const per = (ncut + 63u) / 64u;
                    u32 const b0 = ln * per;
                    i64 lg[8], lh[8];
                    u32 lc[8];
                    i64 g = 0, h = 0;
                    u32 cc = 0;
#pragma unroll
                    for (u32 k = 0; k < 8; ++k) {
                        u32 const b = b0 + k;
                        if (k < per && b < ncut) {
                            g += ci[2 * b];
                            h += ci[2 * b + 1];
                            if (hcnt) cc += cn[b];
                        }
                        lg[k] = g;
                        lh[k] = h;
                        lc[k] = cc;
                    }
                    // inclusive scan of the lane totals
                    i64 sg2 = g, sh2 = h;
                    u32 sc2 = cc;
                    for (int o = 1; o < 64; o <<= 1) {
                        i64 const ug = __shfl_up(sg2, o), uh = __shfl_up(sh2, o);
                        u32 const uc = __shfl_up(sc2, o);
                        if (static_cast<int>(ln) >= o) {
                            sg2 += ug;
                            sh2 += uh;
                            sc2 += uc;
                        }
                    }
                    i64 const eg = sg2 - g, eh = sh2 - h;
                    u32 const ec = sc2 - cc;
                    Acc *qg = pg + t * SB, *qh = ph + t * SB;
#pragma unroll
                    for (u32 k = 0; k < 8; ++k) {
                        u32 const b = b0 + k;
                        if (k < per && b < ncut) {
                            qg[b] = eg + lg[k];
                            qh[b] = eh + lh[k];
                            if (hcnt) cn[b] = ec + lc[k];
                        }
                    }
                }
                if (ln == 0) {
                    s_ncut[t] = ncut;
                    s_dirs[t] = dirs;
                }
            }
        }
        u32 const t = threadIdx.x;
        if (!Q && t < TILE) {
            u32 ncut = 0, dirs = 1;
            if (t < nt) {
                u32 const nb = s_nb[t];
                ncut = nb >= 2 ? nb - 2 : 0;
                Acc *qg = pg + t * SB, *qh = ph + t * SB;
                float2 const *cf = reinterpret_cast<float2 const *>(s_cells + t * SB);
                float2 const m = nb ? cf[nb - 1] : make_float2(0.0F, 0.0F);
                s_mg[t] = m.x;
                s_mh[t] = m.y;
                dirs = (m.x == 0.0F && m.y == 0.0F) ? 1 : 2;
                double g = 0.0, h = 0.0;
                u32 b = 0;
                for (; b + 4 <= ncut; b += 4) {
                    float2 const v0 = cf[b], v1 = cf[b + 1], v2 = cf[b + 2], v3 = cf[b + 3];
                    g += v0.x; h += v0.y; qg[b] = g; qh[b] = h;
                    g += v1.x; h += v1.y; qg[b + 1] = g; qh[b + 1] = h;
                    g += v2.x; h += v2.y; qg[b + 2] = g; qh[b + 2] = h;
                    g += v3.x; h += v3.y; qg[b + 3] = g; qh[b + 3] = h;
                }
                for (; b < ncut; ++b) {
                    float2 const v = cf[b];
                    g += v.x;
                    h += v.y;
                    qg[b] = g;
                    qh[b] = h;
                }
            }
            s_ncut[t] = ncut;
            s_dirs[t] = dirs;
        }
        __syncthreads();
        // Phase B: every (feature, cut bin, direction) candidate; a thread takes one
        // (feature, bin) pair and scores its directions from one prefix load.
        u32 const n_pairs = nt * SB;
        for (u32 pi = threadIdx.x; pi < n_pairs; pi += blockDim.x) {
            u32 const tt = pi / SB;
            u32 const b = pi - tt * SB;
            if (b >= s_ncut[tt]) continue;
            u32 const dirs = s_dirs[tt];
            Acc const plg = pg[tt * SB + b], plh = ph[tt * SB + b];
            u32 const pnl = hcnt ? s_cn[tt * SB + b] : 0u;
            for (u32 d = 0; d < dirs; ++d) {
                bool const dl = d == 0;
                double gL, hL, gR, hR;
                i64 igL = 0, ihL = 0, igR = 0, ihR = 0;
                if (Q) {
                    i64 const lg = plg, lh = plh;
                    i64 const mg = s_mg[tt], mh = s_mh[tt];
                    i64 const rg = s_itot[0] - mg, rh = s_itot[1] - mh;
                    igL = lg + (dl ? mg : 0);
                    ihL = lh + (dl ? mh : 0);
                    igR = (rg - lg) + (!dl ? mg : 0);
                    ihR = (rh - lh) + (!dl ? mh : 0);
                    gL = static_cast<double>(igL) * sg;
                    hL = static_cast<double>(ihL) * sh;
                    gR = static_cast<double>(igR) * sg;
                    hR = static_cast<double>(ihR) * sh;
                } else {
                    double const lg = plg, lh = plh;
                    double const mg = s_mg[tt], mh = s_mh[tt];
                    double const rg = tg - mg, rh = th - mh;
                    double const right_g = rg - lg, right_h = rh - lh;
                    gL = lg + (dl ? mg : 0.0);
                    hL = lh + (dl ? mh : 0.0);
                    gR = right_g + (!dl ? mg : 0.0);
                    hR = right_h + (!dl ? mh : 0.0);
                }
                if (hL < cfg.min_child_hess || hR < cfg.min_child_hess) continue;
                if (CON && cfg.mono) {
                    // src/split.cpp:143-153: a candidate whose bounded child weights
                    // break the feature's direction is skipped.
                    int const mc = cfg.mono[f0 + tt];
                    if (mc != 0) {
                        double const wl = fd_bounded_weight(gL, hL, cfg.l1, cfg.l2, node_lo, node_hi);
                        double const wr = fd_bounded_weight(gR, hR, cfg.l1, cfg.l2, node_lo, node_hi);
                        if (static_cast<double>(mc) * (wr - wl) < 0.0) continue;
                    }
                }
                double const cs = fd_score_of(gL, hL, cfg.l1, cfg.l2) + fd_score_of(gR, hR, cfg.l1, cfg.l2);
                double const gain = cs - node_score;
                if (!(gain > 0.0 && gain >= cfg.min_gain)) continue;
                u32 const key = (f0 + tt) * (2u * SB) + b * 2u + d;
                if (gain > best_gain || (gain == best_gain && key < best_key)) {
                    best_gain = gain;
                    best_key = key;
                    bgL = gL;
                    bhL = hL;
                    bgR = gR;
                    bhR = hR;
                    bigL = igL;
                    bihL = ihL;
                    bigR = igR;
                    bihR = ihR;
                    // Left rows: bins <= the cut bin, plus the missing bin when it routes left.
                    if (hcnt) bnl = pnl + (dl ? s_mc[tt] : 0u);
                }
            }
        }
        __syncthreads();
    }
    // Block winner (gain, then lowest key; a total order, so the butterfly is exact):
    // within each wave by shuffles, then over the wave winners in wave 0.
    {
        double g = best_gain;
        u32 k = best_key;
        for (int o = 32; o > 0; o >>= 1) {
            double const og = __shfl_xor(g, o);
            u32 const ok = __shfl_xor(k, o);
            if (ok != kNone && (k == kNone || og > g || (og == g && ok < k))) {
                g = og;
                k = ok;
            }
        }
        if ((threadIdx.x & 63u) == 0) {
            s_best_gain[threadIdx.x >> 6] = g;
            s_best_key[threadIdx.x >> 6] = k;
        }
        __syncthreads();
        if (threadIdx.x < 64u) {
            u32 const nw = blockDim.x >> 6;
            g = threadIdx.x < nw ? s_best_gain[threadIdx.x] : 0.0;
            k = threadIdx.x < nw ? s_best_key[threadIdx.x] : kNone;
            for (int o = 32; o > 0; o >>= 1) {
                double const og = __shfl_xor(g, o);
                u32 const ok = __shfl_xor(k, o);
                if (ok != kNone && (k == kNone || og > g || (og == g && ok < k))) {
                    g = og;
                    k = ok;
                }
            }
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            s_best_gain[0] = g;
            s_best_key[0] = k;
        }
        __syncthreads();
    }
    // The thread that scored the winner publishes its child sums and left rows (keys
    // are unique) to the block's partial record; the node's last block to finish merges
    // the partials (gain, then lowest key).
    u32 const key = s_best_key[0];
    SplitRec *dst = part + static_cast<size_t>(i) * G + bx;
    if (key != kNone && best_key == key) {
        SplitRec rec{};
        u32 const f = key / (2u * SB), b = (key >> 1) % SB, d = key & 1u;
        rec.gain = s_best_gain[0];
        rec.feature = f;
        rec.bin = b;
        rec.default_left = d == 0 ? 1u : 0u;
        rec.valid = 1;
        rec.gL = bgL;
        rec.hL = bhL;
        rec.gR = bgR;
        rec.hR = bhR;
        rec.igL = bigL;
        rec.ihL = bihL;
        rec.igR = bigR;
        rec.ihR = bihR;
        if (hcnt) {
            rec.n_left = bnl;
            rec.n_right = node.cnt - bnl;
        }
        *dst = rec;
        __threadfence();
    } else if (key == kNone && threadIdx.x == 0) {
        SplitRec rec{};
        rec.default_left = 1;
        *dst = rec;
        __threadfence();
    }
    __syncthreads();
    if (threadIdx.x == 0) s_last = atomicAdd(done + i, 1u) == G - 1u ? 1u : 0u;
    __syncthreads();
    if (s_last == 0u) return;
    __shared__ SplitRec s_parts[kFindMaxBlocks];
    // The G partial records are loaded in parallel by lanes of wave 0 (G <= 64), after
    // one fence in that wave, then compared.
    if (threadIdx.x < 64u) {
        __threadfence();
        if (threadIdx.x < G) s_parts[threadIdx.x] = part[static_cast<size_t>(i) * G + threadIdx.x];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        u32 bi = kNone, bkey = kNone;
        double bg = 0.0;
        for (u32 k = 0; k < G; ++k) {
            SplitRec const &r = s_parts[k];
            if (!r.valid) continue;
            u32 const rk = r.feature * (2u * SB) + r.bin * 2u + (r.default_left ? 0u : 1u);
            if (bkey == kNone || r.gain > bg || (r.gain == bg && rk < bkey)) {
                bi = k;
                bkey = rk;
                bg = r.gain;
            }
        }
        done[i] = 0u;
        if (!had_totals) front[i] = node;
        if (bi == kNone) {
            SplitRec rec{};
            rec.default_left = 1;
            out[i] = rec;
        } else {
            out[i] = s_parts[bi];
        }
    }
}

} // namespace bh
