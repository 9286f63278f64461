// Leafwise and levelwise grow loops (kernels_ext.hip): their extra device state and the
// view of the engine context they run on.
#pragma once
#include <hip/hip_runtime.h>
#include "../../include/bonsai_hip.h"
#include "bh_common.hpp"

namespace bh {

// Engine buffers a grow loop reads and writes (owned by bh_context).
struct GrowView {
    u64 N;
    u32 F, Fs, SB;
    u8 *plane0, *work[2];
    void const *gh[3];
    u16 const *d_nbins;
    void *hist[2];
    void *partials;
    TreeCtl *ctl;
    NodeInfo *front[2];
    SplitRec *splits;
    FindScratch fscr;
    DevNode *tree;
    PartOp *ops;
    PartChunk *pch;
    u32 *chunk_left;
    HistChunk *hch;
    FinishOp *fops;
    double *scales;
    u32 const *fx;   // FP32 histogram grids (launch_hist)
    GrowCfg gc;
    bh_fit_config fc;
    u32 max_pch, max_hch, max_ops, fsl, waves;
    RootSel root;  // the tree's root rows
    u32 *cnt;      // quantized leafwise: row count per (slot, feature, bin) beside the pool
    u32 lw_units;  // leafwise: histogram units per built node at most
    bool host_sync;
};

// Leafwise candidate pool entry: a leaf with its best split, waiting in the gain heap
// (grower_detail::Candidate, src/step/primitives.hpp:404-415).
struct Cand {
    NodeInfo node;
    SplitRec split;
};

struct ExtState {
    // leafwise
    Cand *cands = nullptr;      // candidate pool (heap entries), capacity max_leaves + 4
    NodeInfo *pending = nullptr;// leaves set aside (no valid split)
    u32 *free_slots = nullptr;  // histogram slot free list
    u32 cap = 0;
    // levelwise
    SplitRec *feat_best = nullptr; // per-feature best of the level scan
    u32 *level_splits = nullptr;   // (feature, bin, default_left) per level
    float *level_gains = nullptr;
    u32 n_features = 0;
    u32 hist_slots = 0;
};

hipError_t ext_alloc(ExtState &s, bh_fit_config const &f, u32 F, u32 SB, u32 max_front, u32 hist_slots);
void ext_free(ExtState &s);
hipError_t ext_set_limits();
void enqueue_levelwise(ExtState &s, hipStream_t st, bool q, GrowView const &v);
void enqueue_leafwise(ExtState &s, hipStream_t st, bool q, GrowView const &v);
// The level finder as one step over a host-laid frontier (front[0 .. n), ctl->n_front =
// n): pick[0] = the level's split, v.splits[i] = node i's child sums under it.
void launch_level_find_step(ExtState &s, hipStream_t st, bool q, GrowView const &v, NodeInfo *front, u32 n,
                            void const *hist, SplitRec *pick);

} // namespace bh

struct bh_context;
bh::GrowView grow_view(bh_context *c);
