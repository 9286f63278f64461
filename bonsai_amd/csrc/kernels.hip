// Training kernels of the MI355X GBDT engine (gfx950, wave64).
//
// Every kernel is launched with a fixed grid and reads its work list from device
// memory (TreeCtl + op/chunk arrays), so a whole boosting iteration can be replayed
// from one hipGraph with no host round-trip.  Work lists are produced on device by the
// single-workgroup control kernels (k_plan_level, k_part_scan, ...).
//
// Reference behaviour each kernel reproduces is cited per kernel.
#include <hip/hip_runtime.h>
#include <math.h>
#include <algorithm>
#include "bh_common.hpp"
#include "bh_device.hpp"
#include "bh_find.hpp"
#include "kernels.hpp"

namespace bh {

// ---------------------------------------------------------------------------------
// Shared device helpers
// ---------------------------------------------------------------------------------
__device__ __forceinline__ bool routes_left(u32 bin, u32 last, u32 split, u32 dl)
{
    // include/bonsai/dataset.hpp:234-238
    return bin == last ? (dl != 0) : (bin <= split);
}

__device__ __forceinline__ double l1_thr(double g, double l1)
{
    if (g > l1) return g - l1;
    if (g < -l1) return g + l1;
    return 0.0;
}
// include/bonsai/split.hpp:135-143
__device__ __forceinline__ double score_of(double g, double h, double l1, double l2)
{
    double const t = l1_thr(g, l1);
    double const d = h + l2;
    return d > 0.0 ? (t * t) / d : 0.0;
}
// include/bonsai/split.hpp:146-152 (bounds are +-inf: no monotone constraints)
__device__ __forceinline__ float leaf_weight(double g, double h, double l1, double l2)
{
    double const d = h + l2;
    double const w = d > 0.0 ? -l1_thr(g, l1) / d : 0.0;
    return static_cast<float>(w);
}

__device__ __forceinline__ u64 splitmix64(u64 z)
{
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}
__device__ __forceinline__ float sr_uniform(u64 seed, u32 iter, u64 row, u32 channel)
{
    u64 const key = seed ^ (static_cast<u64>(iter) * 0xD1B54A32D192ED03ULL) ^
                    (row * 0x9E3779B97F4A7C15ULL) ^
                    (static_cast<u64>(channel + 1) * 0xA0761D6478BD642FULL);
    u64 const z = splitmix64(key);
    return static_cast<float>(static_cast<u32>(z >> 40)) * (1.0F / 16777216.0F);
}

// Stages bytes [a, a + len) of src in LDS with aligned 16-byte loads from the enclosing
// aligned range; returns the byte offset of `a` inside dst.  Planes and gradient buffers
// carry 64 bytes of tail padding, so the over-read past the range stays allocated.
__device__ __forceinline__ u32 lds_load_range(u8 const *__restrict__ src, size_t a, u32 len, uint4 *dst)
{
    size_t const a16 = a & ~static_cast<size_t>(15);
    u32 const sh = static_cast<u32>(a - a16);
    u32 const nw = (sh + len + 15u) >> 4;
    uint4 const *s = reinterpret_cast<uint4 const *>(src + a16);
    for (u32 w = threadIdx.x; w < nw; w += blockDim.x) dst[w] = s[w];
    return sh;
}

// lds_load_range in two halves, so a tile's loads can be in flight while the block
// works on the previous tile: pf_load reads the aligned 16-byte words into registers
// (at most K per thread, w[] with constant indices only), pf_store writes them to LDS.
template <u32 K>
__device__ __forceinline__ u32 pf_load(uint4 (&w)[K], u32 &nw, u8 const *__restrict__ src, size_t a, u32 len)
{
    size_t const a16 = a & ~static_cast<size_t>(15);
    u32 const sh = static_cast<u32>(a - a16);
    nw = (sh + len + 15u) >> 4;
    uint4 const *s = reinterpret_cast<uint4 const *>(src + a16);
#pragma unroll
    for (u32 k = 0; k < K; ++k) {
        u32 const i = threadIdx.x + k * blockDim.x;
        w[k] = i < nw ? s[i] : make_uint4(0, 0, 0, 0);
    }
    return sh;
}

template <u32 K> __device__ __forceinline__ void pf_store(uint4 const (&w)[K], u32 nw, uint4 *dst)
{
#pragma unroll
    for (u32 k = 0; k < K; ++k) {
        u32 const i = threadIdx.x + k * blockDim.x;
        if (i < nw) dst[i] = w[k];
    }
}

__device__ __forceinline__ void write_leaf(DevNode *tree, u32 id, float v)
{
    DevNode n = tree[id];
    n.feature = -1;
    n.bin = 0;
    n.default_left = 0;
    n.left = 0;
    n.right = 0;
    n.value = v;
    tree[id] = n;
}

__device__ __forceinline__ float node_leaf_value(NodeInfo const &n, GrowCfg const &c,
                                                 double const *scales)
{
    if (c.mono) // write_leaf's bounded weight (src/step/primitives.hpp:42-50)
        return c.quantized ? fd_leaf_value(n.id, static_cast<double>(n.itg) * scales[0],
                                           static_cast<double>(n.ith) * scales[1], c)
                           : fd_leaf_value(n.id, n.tg, n.th, c);
    if (c.quantized)
        return leaf_weight(static_cast<double>(n.itg) * scales[0],
                           static_cast<double>(n.ith) * scales[1], c.l1, c.l2);
    return leaf_weight(n.tg, n.th, c.l1, c.l2);
}

// ---------------------------------------------------------------------------------
// Objective: src/objective.cpp:44-101 (mse, logloss) and :253-278 (poisson, log link,
// raw score clamped to +-30 before exp, include/bonsai/objective.hpp:47); device twin of
// gh_from_scores_kernel (src/cuda/detail/kernels.cuh:972-1009) fused with the
// per-iteration quantizer.  Weighted rows scale g and h by w[r] (the kernel's Weighted
// arm; the host multiply of include/bonsai/booster.hpp:546-553).
// ---------------------------------------------------------------------------------
constexpr float kPoissonMaxLog = 30.0F;

__device__ __forceinline__ void grad_of(int objective, float s, float y, float &g, float &h)
{
    if (objective == 1) {
        float const p = 1.0F / (1.0F + expf(-s));
        g = p - y;
        h = p * (1.0F - p);
    } else if (objective == 2) {
        float const mu = expf(fminf(fmaxf(s, -kPoissonMaxLog), kPoissonMaxLog));
        g = mu - y;
        h = mu;
    } else {
        g = s - y;
        h = 1.0F;
    }
}

// Row r's gradient pair, scaled by its weight when the fit is weighted.
__device__ __forceinline__ void grad_of_row(int objective, float s, float y, float const *__restrict__ weight,
                                            u64 r, float &g, float &h)
{
    grad_of(objective, s, y, g, h);
    if (weight) {
        float const w = weight[r];
        g *= w;
        h *= w;
    }
}


__global__ void k_zero_u32(u32 *p, u32 n)
{
    for (u32 i = threadIdx.x; i < n; i += blockDim.x) p[i] = 0u;
}

// max |g| and max h as float bits (non-negative floats order like their bit patterns,
// so the max is independent of the reduction order).
__device__ __forceinline__ void block_max2_atomic(float mg, float mh, u32 *maxbits)
{
    __shared__ float s_m[2][16];
    for (int o = 32; o > 0; o >>= 1) {
        mg = fmaxf(mg, __shfl_xor(mg, o));
        mh = fmaxf(mh, __shfl_xor(mh, o));
    }
    if ((threadIdx.x & 63) == 0) {
        s_m[0][threadIdx.x >> 6] = mg;
        s_m[1][threadIdx.x >> 6] = mh;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (u32 w = 1; w < blockDim.x / 64; ++w) {
            mg = fmaxf(mg, s_m[0][w]);
            mh = fmaxf(mh, s_m[1][w]);
        }
        atomicMax(&maxbits[0], __float_as_uint(mg));
        atomicMax(&maxbits[1], __float_as_uint(mh));
    }
}

// FP32 gradients; the maxima of |g| and h go to fx[0..1] for the histogram's
// fixed-point grid (k_fx_shift).
__global__ __launch_bounds__(256) void k_grad_f32(float const *__restrict__ score, float const *__restrict__ label,
                                                  float const *__restrict__ weight, u64 n, int objective,
                                                  float2 *__restrict__ gh, u32 *__restrict__ fx)
{
    float mg = 0.0F, mh = 0.0F;
    for (u64 r = blockIdx.x * static_cast<u64>(blockDim.x) + threadIdx.x; r < n;
         r += static_cast<u64>(gridDim.x) * blockDim.x) {
        float g, h;
        grad_of_row(objective, score[r], label[r], weight, r, g, h);
        gh[r] = make_float2(g, h);
        mg = fmaxf(mg, fabsf(g));
        mh = fmaxf(mh, fabsf(h));
    }
    block_max2_atomic(mg, mh, fx);
}

// Max-abs of g and h over every row, for the quantizer of the current iteration.  The
// maxima live in two slots of maxbits by iteration parity (2 words each): the training
// score update of iteration t fills slot (t + 1) & 1 for the next iteration, so this
// pass only primes slot t & 1 when scores or labels changed outside the loop.
__global__ void k_grad_max(float const *__restrict__ score, float const *__restrict__ label,
                           float const *__restrict__ weight, u64 n, int objective, u32 *__restrict__ maxbits,
                           TreeCtl const *__restrict__ ctl)
{
    maxbits += 2u * (ctl->iter & 1u);
    float mg = 0.0F, mh = 0.0F;
    for (u64 r = blockIdx.x * static_cast<u64>(blockDim.x) + threadIdx.x; r < n;
         r += static_cast<u64>(gridDim.x) * blockDim.x) {
        float g, h;
        grad_of_row(objective, score[r], label[r], weight, r, g, h);
        mg = fmaxf(mg, fabsf(g));
        mh = fmaxf(mh, fabsf(h));
    }
    block_max2_atomic(mg, mh, maxbits);
}

// Quantizer shared with the oracle (oracle/oracle.cpp orc_quantize): int8 g / int16 h,
// stochastic rounding q = floor(x * (qmax / max) + u), u from a counter RNG keyed on
// (seed, iteration, original row, channel).
__global__ void k_grad_quant(float const *__restrict__ score, float const *__restrict__ label,
                             float const *__restrict__ weight, u64 n, int objective, u32 *__restrict__ maxbits,
                             u64 seed,
                             TreeCtl const *__restrict__ ctl, u32 *__restrict__ ghq,
                             double *__restrict__ scales)
{
    u32 const iter = ctl->iter;
    u32 const par = iter & 1u;
    float const mg = __uint_as_float(maxbits[2u * par]);
    float const mh = __uint_as_float(maxbits[2u * par + 1u]);
    float const ig = mg > 0.0F ? 127.0F / mg : 0.0F;
    float const ih = mh > 0.0F ? 32767.0F / mh : 0.0F;
    // the other slot collects this iteration's score update maxima (nothing reads it now)
    if (blockIdx.x == 0 && threadIdx.x < 2u) maxbits[2u * (par ^ 1u) + threadIdx.x] = 0u;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        scales[0] = static_cast<double>(mg) / 127.0;
        scales[1] = static_cast<double>(mh) / 32767.0;
    }
    for (u64 r = blockIdx.x * static_cast<u64>(blockDim.x) + threadIdx.x; r < n;
         r += static_cast<u64>(gridDim.x) * blockDim.x) {
        float g, h;
        grad_of_row(objective, score[r], label[r], weight, r, g, h);
        float const tg = g * ig + sr_uniform(seed, iter, r, 0);
        float const th = h * ih + sr_uniform(seed, iter, r, 1);
        float qg = fminf(127.0F, fmaxf(-127.0F, floorf(tg)));
        float qh = fminf(32767.0F, fmaxf(0.0F, floorf(th)));
        int const ig8 = static_cast<int>(qg);
        int const ih16 = static_cast<int>(qh);
        ghq[r] = (static_cast<u32>(static_cast<u16>(static_cast<int16_t>(ig8))) << 16) |
                 static_cast<u32>(static_cast<u16>(ih16));
    }
}

// Host-supplied float gradients quantized with the same rule (parity tests).
__global__ void k_quant_from(float2 const *__restrict__ gh, u64 n, u32 const *__restrict__ maxbits,
                             u64 seed, u32 iter, u32 *__restrict__ ghq)
{
    float const mg = __uint_as_float(maxbits[0]);
    float const mh = __uint_as_float(maxbits[1]);
    float const ig = mg > 0.0F ? 127.0F / mg : 0.0F;
    float const ih = mh > 0.0F ? 32767.0F / mh : 0.0F;
    for (u64 r = blockIdx.x * static_cast<u64>(blockDim.x) + threadIdx.x; r < n;
         r += static_cast<u64>(gridDim.x) * blockDim.x) {
        float2 const v = gh[r];
        float const tg = v.x * ig + sr_uniform(seed, iter, r, 0);
        float const th = v.y * ih + sr_uniform(seed, iter, r, 1);
        float qg = fminf(127.0F, fmaxf(-127.0F, floorf(tg)));
        float qh = fminf(32767.0F, fmaxf(0.0F, floorf(th)));
        ghq[r] = (static_cast<u32>(static_cast<u16>(static_cast<int16_t>(static_cast<int>(qg))))
                  << 16) |
                 static_cast<u32>(static_cast<u16>(static_cast<int>(qh)));
    }
}

__global__ void k_gh_max(float2 const *__restrict__ gh, u64 n, u32 *__restrict__ maxbits)
{
    float mg = 0.0F, mh = 0.0F;
    for (u64 r = blockIdx.x * static_cast<u64>(blockDim.x) + threadIdx.x; r < n;
         r += static_cast<u64>(gridDim.x) * blockDim.x) {
        float2 const v = gh[r];
        mg = fmaxf(mg, fabsf(v.x));
        mh = fmaxf(mh, fabsf(v.y));
    }
    block_max2_atomic(mg, mh, maxbits);
}

// ---------------------------------------------------------------------------------
// Histogram build (H6).  Device counterpart of hist_tile_kernel
// (src/cuda/detail/kernels.cuh:186-277); CPU semantics of populate
// (src/grower.cpp:93-139).
//
// One workgroup per work unit (at most kHistChunk rows of one node's contiguous segment)
// and feature slice; the slice's sub-histogram lives in LDS and is written once to a
// partial (multi-unit nodes) or straight into the node's slot (single-unit nodes).
//
// FP32 gradients (k_hist_fx): LDS float atomics cost about 30 times an integer atomic on
// this device (192 against 6.5 clocks per wave instruction and CU, measured), so the
// cells sum on a fixed-point grid instead.  Every g (h) goes on the grid 2^-s with
// s = 49 - e, max|g| < 2^e over the iteration's rows, rounded to nearest even: a unit's
// cells then sum exactly in int64 (8,192 x 2^49 < 2^63), in any order, and each cell is
// rounded once to float when the unit is written.  Units fold in unit order and the
// larger sibling is parent - small in float (k_finish<false>), as before.  Lanes are row
// x dword as in k_hist_q: two ds_add_u64 per (row, feature).
// ---------------------------------------------------------------------------------
__device__ __forceinline__ int fx_shift_of(float m)
{
    if (!(m > 0.0F) || !(m <= 3.402823466e38F)) return 0;
    int e = 0;
    (void)frexpf(m, &e);
    return 49 - e;
}
__device__ __forceinline__ long long fx_quant(float v, int s) { return static_cast<long long>(rintf(ldexpf(v, s))); }
__device__ __forceinline__ float fx_dequant(long long q, int s) { return ldexpf(static_cast<float>(q), -s); }

// The grids of the current gradients: fx[0..1] hold max|g|, max h as float bits (filled
// by the gradient pass), fx[2..3] receive s_g, s_h, and the maxima are cleared for the
// next pass.  One thread.
__global__ void k_fx_shift(u32 *fx)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        fx[2] = static_cast<u32>(fx_shift_of(__uint_as_float(fx[0])));
        fx[3] = static_cast<u32>(fx_shift_of(__uint_as_float(fx[1])));
        fx[0] = 0u;
        fx[1] = 0u;
    }
}

constexpr int kFxUnroll = 4;

__global__ __launch_bounds__(1024) void k_hist_fx(
    HistChunk const *__restrict__ chunks, TreeCtl const *__restrict__ ctl,
    u8 const *__restrict__ pl0, u8 const *__restrict__ pl1, u8 const *__restrict__ plp,
    void const *__restrict__ gw0, void const *__restrict__ gw1, void const *__restrict__ gwp,
    u32 Fs, u32 F, u32 SB, u32 fsl, void *__restrict__ partials, void *__restrict__ hist_next,
    u32 const *__restrict__ fx)
{
    extern __shared__ __align__(16) unsigned char smem[];
    u32 const f_begin = blockIdx.y * fsl;
    if (f_begin >= F) return;
    u32 const nf = min(F - f_begin, fsl);
    u32 const cells = nf * SB;
    unsigned long long *lg = reinterpret_cast<unsigned long long *>(smem); // nf x SB g cells
    unsigned long long *lh = lg + cells;                                    // nf x SB h cells
    int const sg = static_cast<int>(fx[2]), sh = static_cast<int>(fx[3]);
    u32 const Dw = (nf + 3u) / 4u;             // dwords of the slice per row
    u32 const R = 64u / Dw;                    // rows per wave instruction
    u32 const lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    u32 const NW = blockDim.x >> 6;
    u32 const lr = lane / Dw, d = lane - lr * Dw;
    bool const lane_ok = lr < R;
    u32 const n_units = ctl->n_hist_chunks;
    size_t const node_cells = static_cast<size_t>(F) * SB;
    for (u32 c = blockIdx.x; c < n_units; c += gridDim.x) {
        HistChunk const ch = chunks[c];
        for (u32 i = threadIdx.x; i < 2u * cells; i += blockDim.x) lg[i] = 0ull;
        __syncthreads();
        u8 const *plane = ch.src == 0 ? pl0 : (ch.src == 1 ? pl1 : plp);
        float2 const *ghv = static_cast<float2 const *>(ch.src == 0 ? gw0 : (ch.src == 1 ? gw1 : gwp));
        u8 const *cb = plane + f_begin + 4u * d;
        u32 const stride = NW * R;
        auto add_row = [&](u32 dwv, float2 v) {
            unsigned long long const qg = static_cast<unsigned long long>(fx_quant(v.x, sg));
            unsigned long long const qh = static_cast<unsigned long long>(fx_quant(v.y, sh));
#pragma unroll
            for (u32 k = 0; k < 4; ++k) {
                u32 const fl = 4u * d + k;
                if (fl < nf) {
                    u32 const cell = fl * SB + ((dwv >> (8u * k)) & 0xFFu);
                    atomicAdd(lg + cell, qg);
                    atomicAdd(lh + cell, qh);
                }
            }
        };
        if (lane_ok) {
            u32 r = ch.begin + wave * R + lr;
            for (; r + (kFxUnroll - 1) * stride < ch.end; r += kFxUnroll * stride) {
                u32 dw[kFxUnroll];
                float2 v[kFxUnroll];
#pragma unroll
                for (int u = 0; u < kFxUnroll; ++u) {
                    u32 const rr = r + u * stride;
                    dw[u] = *reinterpret_cast<u32 const *>(cb + static_cast<size_t>(rr) * Fs);
                    v[u] = ghv[rr];
                }
#pragma unroll
                for (int u = 0; u < kFxUnroll; ++u) add_row(dw[u], v[u]);
            }
            for (; r < ch.end; r += stride)
                add_row(*reinterpret_cast<u32 const *>(cb + static_cast<size_t>(r) * Fs), ghv[r]);
        }
        __syncthreads();
        float2 *dst = ch.part != kNone
                          ? static_cast<float2 *>(partials) + static_cast<size_t>(ch.part) * node_cells +
                                static_cast<size_t>(f_begin) * SB
                          : static_cast<float2 *>(hist_next) + static_cast<size_t>(ch.slot) * node_cells +
                                static_cast<size_t>(f_begin) * SB;
        for (u32 i = threadIdx.x; i < cells; i += blockDim.x)
            dst[i] = make_float2(fx_dequant(static_cast<long long>(lg[i]), sg),
                                 fx_dequant(static_cast<long long>(lh[i]), sh));
        __syncthreads();
    }
}

// Quantized histogram build, row-and-dword lanes (H6, same unit contract as k_hist_fx).
// Lane l of a wave takes dword d = l % Dw of row l / Dw (Dw = the slice's dwords per
// row), so one wave load reads 64 / Dw whole rows as one contiguous span and a lane's
// four bins feed four ds_add_u64 on the packed (g << 32) + h cell.  Each atomic
// instruction then spreads over 64 / Dw rows x Dw features (7 features x 9 rows at 28
// features) instead of 64 rows of one feature, which removes most same-address
// collisions.  Integer sums are exact and order-free.  The LDS histogram is the slice's
// [feature][SB] cells.
constexpr u32 kHqThreads = 1024;
constexpr int kHqUnroll = 8;

// One row's packed contribution (see kPackGShift): the gradient word carries int16 g in
// its high half and uint16 h in its low half.
__device__ __forceinline__ u64 pack_q(u32 v)
{
    return static_cast<u64>(static_cast<i64>(static_cast<int16_t>(v >> 16)) * (1ll << kPackGShift)) +
           (1ull << kPackCntShift) + (v & 0xFFFFu);
}
__device__ __forceinline__ i64 unpack_g(u64 v) { return static_cast<i64>(v) >> kPackGShift; }
__device__ __forceinline__ i64 unpack_h(u64 v) { return static_cast<i64>(v & kPackHMask); }
__device__ __forceinline__ u32 unpack_cnt(u64 v) { return static_cast<u32>((v >> kPackCntShift) & kPackCntMask); }

// Rows of a unit whose node has not been moved out of its parent's segment count only
// when the parent's split sends them to the node's side.  The routing byte of a row
// sits in one lane of the row's lane group; a wave ballot hands its verdict to the
// group (a feature outside the slice is read directly).
struct HqRoute {
    u32 feature, bin, dl, last, want_left, in_slice, lane_d, shift;
};

template <bool ROUTED>
__device__ __forceinline__ void hist_q_rows(unsigned long long *__restrict__ lh, u8 const *__restrict__ plane,
                                            u32 const *__restrict__ ghv, u8 const *__restrict__ cb, u32 Fs,
                                            u32 SB, u32 nf, u32 d, u32 lr, u32 Dw, u32 r, u32 end, u32 stride,
                                            HqRoute const &rt)
{
    auto keep_row = [&](u32 dwv, u32 row) -> bool {
        if (!ROUTED) return true;
        if (rt.in_slice) {
            bool const mine = d == rt.lane_d &&
                              routes_left((dwv >> rt.shift) & 0xFFu, rt.last, rt.bin, rt.dl) == (rt.want_left != 0);
            u64 const m = __ballot(mine);
            return ((m >> (lr * Dw + rt.lane_d)) & 1ull) != 0;
        }
        u32 const b = plane[static_cast<size_t>(row) * Fs + rt.feature];
        return routes_left(b, rt.last, rt.bin, rt.dl) == (rt.want_left != 0);
    };
    for (; r + (kHqUnroll - 1) * stride < end; r += kHqUnroll * stride) {
        u32 dw[kHqUnroll], v[kHqUnroll];
#pragma unroll
        for (int u = 0; u < kHqUnroll; ++u) {
            u32 const rr = r + u * stride;
            dw[u] = *reinterpret_cast<u32 const *>(cb + static_cast<size_t>(rr) * Fs);
            v[u] = ghv[rr];
        }
#pragma unroll
        for (int u = 0; u < kHqUnroll; ++u) {
            if (!keep_row(dw[u], r + u * stride)) continue;
            u64 const pk = pack_q(v[u]);
#pragma unroll
            for (u32 k = 0; k < 4; ++k) {
                u32 const fl = 4u * d + k;
                if (fl < nf)
                    atomicAdd(lh + fl * SB + ((dw[u] >> (8u * k)) & 0xFFu), static_cast<unsigned long long>(pk));
            }
        }
    }
    for (; r < end; r += stride) {
        u32 const dw1 = *reinterpret_cast<u32 const *>(cb + static_cast<size_t>(r) * Fs);
        u32 const v1 = ghv[r];
        if (!keep_row(dw1, r)) continue;
        u64 const pk = pack_q(v1);
#pragma unroll
        for (u32 k = 0; k < 4; ++k) {
            u32 const fl = 4u * d + k;
            if (fl < nf)
                atomicAdd(lh + fl * SB + ((dw1 >> (8u * k)) & 0xFFu), static_cast<unsigned long long>(pk));
        }
    }
}

// Quantized unit partials: packed 64-bit cells (one 8192-row sub-unit per unit), or wide
// -- three u32 planes of node_cells each (g as int32, h, row count) -- on levels whose
// units group several sub-units (HistChunk::wide).  A grouped work unit (up to
// kHistGroupMax sub-units of one node) adds every sub-unit's packed LDS cells into its
// own partial (or its node's slot) in global memory and clears the LDS for the next
// sub-unit: only this workgroup touches those cells during the launch, and the group
// writes one partial instead of one per sub-unit (sums of 4 sub-units:
// |g| <= 127 * 32768, h <= 32767 * 32768 < 2^32, count <= 2^15).
//
// Flush of the unit's cells [cell0, cell0 + cells) after one sub-unit: `first` writes,
// later sub-units add.  `clear` zeroes the LDS cells for the next sub-unit.
__device__ __forceinline__ void hq_flush(unsigned long long *__restrict__ lh, u32 cells, HistChunk const &ch,
                                         size_t node_cells, size_t cell0, bool first, bool clear,
                                         void *__restrict__ partials, void *__restrict__ hist_next,
                                         u32 *__restrict__ cnt_next)
{
    if (ch.part != kNone && !ch.wide) {
        u64 *P = static_cast<u64 *>(partials) + static_cast<size_t>(ch.part) * node_cells + cell0;
        for (u32 i = threadIdx.x; i < cells; i += blockDim.x) P[i] = lh[i];
    } else if (ch.part != kNone) {
        u32 *P = static_cast<u32 *>(partials) + static_cast<size_t>(ch.part) * node_cells * 3u + cell0;
        for (u32 i = threadIdx.x; i < cells; i += blockDim.x) {
            u64 const v = lh[i];
            if (clear) lh[i] = 0ull;
            u32 g = static_cast<u32>(static_cast<int32_t>(unpack_g(v)));
            u32 h = static_cast<u32>(unpack_h(v));
            u32 n = unpack_cnt(v);
            if (!first) {
                g += P[i];
                h += P[node_cells + i];
                n += P[2 * node_cells + i];
            }
            P[i] = g;
            P[node_cells + i] = h;
            P[2 * node_cells + i] = n;
        }
    } else {
        size_t const base = static_cast<size_t>(ch.slot) * node_cells + cell0;
        i64 *dst = static_cast<i64 *>(hist_next) + base * 2;
        for (u32 i = threadIdx.x; i < cells; i += blockDim.x) {
            u64 const v = lh[i];
            if (clear) lh[i] = 0ull;
            i64 g = unpack_g(v), h = unpack_h(v);
            u32 n = unpack_cnt(v);
            if (!first) {
                g += dst[2 * i];
                h += dst[2 * i + 1];
                if (cnt_next) n += cnt_next[base + i];
            }
            dst[2 * i] = g;
            dst[2 * i + 1] = h;
            if (cnt_next) cnt_next[base + i] = n;
        }
    }
}

__global__ __launch_bounds__(kHqThreads) void k_hist_q(
    HistChunk const *__restrict__ chunks, TreeCtl const *__restrict__ ctl,
    u8 const *__restrict__ pl0, u8 const *__restrict__ pl1, u8 const *__restrict__ plp,
    void const *__restrict__ gw0, void const *__restrict__ gw1, void const *__restrict__ gwp,
    u32 Fs, u32 F, u32 SB, u32 fsl, void *__restrict__ partials, void *__restrict__ hist_next,
    u32 *__restrict__ cnt_next, PartOp const *__restrict__ routes)
{
    extern __shared__ __align__(16) unsigned char smem[];
    unsigned long long *lh = reinterpret_cast<unsigned long long *>(smem);   // fsl x SB
    u32 const f_begin = blockIdx.y * fsl;
    if (f_begin >= F) return;
    u32 const nf = min(F - f_begin, fsl);
    u32 const Dw = (nf + 3u) / 4u;              // dwords of the slice per row (<= 32)
    u32 const R = 64u / Dw;                     // rows per wave instruction
    u32 const lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    u32 const NW = blockDim.x >> 6;
    u32 const lr = lane / Dw, d = lane - lr * Dw;
    bool const lane_ok = lr < R;
    u32 const cells = nf * SB;
    u32 const n_units = ctl->n_hist_chunks;
    size_t const node_cells = static_cast<size_t>(F) * SB;

    for (u32 c = blockIdx.x; c < n_units; c += gridDim.x) {
        HistChunk const ch = chunks[c];
        for (u32 i = threadIdx.x; i < cells; i += blockDim.x) lh[i] = 0ull;
        __syncthreads();
        u8 const *plane = ch.src == 0 ? pl0 : (ch.src == 1 ? pl1 : plp);
        u32 const *ghv = static_cast<u32 const *>(ch.src == 0 ? gw0 : (ch.src == 1 ? gw1 : gwp));
        u8 const *cb = plane + f_begin + 4u * d;
        u32 const stride = NW * R;
        HqRoute rt{};
        if (ch.route != kNone) {
            PartOp const &op = routes[ch.route];
            rt.feature = op.feature;
            rt.bin = op.bin;
            rt.dl = op.default_left;
            rt.last = op.last_bin;
            rt.want_left = ch.side == 0 ? 1u : 0u;
            rt.in_slice = op.feature >= f_begin && op.feature < f_begin + nf ? 1u : 0u;
            rt.lane_d = (op.feature - f_begin) >> 2;
            rt.shift = 8u * ((op.feature - f_begin) & 3u);
        }
        for (u32 s0 = ch.begin; s0 < ch.end; s0 += kHistChunk) {
            u32 const s1 = min(ch.end, s0 + kHistChunk);
            u32 const r = s0 + wave * R + lr;
            if (lane_ok) {
                if (ch.route != kNone)
                    hist_q_rows<true>(lh, plane, ghv, cb, Fs, SB, nf, d, lr, Dw, r, s1, stride, rt);
                else
                    hist_q_rows<false>(lh, plane, ghv, cb, Fs, SB, nf, d, lr, Dw, r, s1, stride, rt);
            }
            __syncthreads();
            hq_flush(lh, cells, ch, node_cells, static_cast<size_t>(f_begin) * SB, s0 == ch.begin, s1 < ch.end,
                     partials, hist_next, cnt_next);
            __syncthreads();
        }
    }
}

// Quantized histogram, rotated row-dword lanes (rows of at most 64 bytes, all features in
// one slice, single 8192-row sub-units; same contract as k_hist_q otherwise).  Loads are
// k_hist_q's: lane (r, d) reads dword d of row r, so one wave load is 64 / Dw whole rows
// of contiguous bytes.  The LDS cells are feature-minor, cell (f, b) at b * FS + f, so a
// cell's bank pair is a function of its feature alone, and atomic instruction k makes
// lane (r, d) add byte (k + r) mod 4 of its dword: the lanes of one instruction then
// cover every feature with about 64 / F lanes each instead of 64 / Dw rows of 4
// features, so random bins no longer collide on banks.  Unit partials are written in
// (bin, feature) order like k_hist_q4's (the completion kernel reads them transposed);
// single-unit nodes write their slot directly.
template <bool ROUTED>
__device__ __forceinline__ void hist_q5_rows(unsigned long long *__restrict__ lh, u8 const *__restrict__ plane,
                                             u32 const *__restrict__ ghv, u8 const *__restrict__ cb, u32 Fs,
                                             u32 F, u32 FS, u32 d, u32 lr, u32 Dw, u32 r, u32 end, u32 stride,
                                             HqRoute const &rt)
{
    u32 const rot = lr & 3u;
    auto keep_row = [&](u32 dwv) -> bool {
        if (!ROUTED) return true;
        bool const mine = d == rt.lane_d &&
                          routes_left((dwv >> rt.shift) & 0xFFu, rt.last, rt.bin, rt.dl) == (rt.want_left != 0);
        u64 const m = __ballot(mine);
        return ((m >> (lr * Dw + rt.lane_d)) & 1ull) != 0;
    };
    auto add = [&](u32 dwv, u32 v) {
        u64 const pk = pack_q(v);
#pragma unroll
        for (u32 k = 0; k < 4; ++k) {
            u32 const j = (k + rot) & 3u;
            u32 const f = 4u * d + j;
            if (f < F)
                atomicAdd(lh + ((dwv >> (8u * j)) & 0xFFu) * FS + f, static_cast<unsigned long long>(pk));
        }
    };
    for (; r + (kHqUnroll - 1) * stride < end; r += kHqUnroll * stride) {
        u32 dw[kHqUnroll], v[kHqUnroll];
#pragma unroll
        for (int u = 0; u < kHqUnroll; ++u) {
            u32 const rr = r + u * stride;
            dw[u] = *reinterpret_cast<u32 const *>(cb + static_cast<size_t>(rr) * Fs);
            v[u] = ghv[rr];
        }
#pragma unroll
        for (int u = 0; u < kHqUnroll; ++u)
            if (keep_row(dw[u])) add(dw[u], v[u]);
    }
    for (; r < end; r += stride) {
        u32 const dw1 = *reinterpret_cast<u32 const *>(cb + static_cast<size_t>(r) * Fs);
        u32 const v1 = ghv[r];
        if (keep_row(dw1)) add(dw1, v1);
    }
}

__global__ __launch_bounds__(kHqThreads) void k_hist_q5(
    HistChunk const *__restrict__ chunks, TreeCtl const *__restrict__ ctl,
    u8 const *__restrict__ pl0, u8 const *__restrict__ pl1, u8 const *__restrict__ plp,
    void const *__restrict__ gw0, void const *__restrict__ gw1, void const *__restrict__ gwp,
    u32 Fs, u32 F, u32 SB, u32 FS, void *__restrict__ partials, void *__restrict__ hist_next,
    u32 *__restrict__ cnt_next, PartOp const *__restrict__ routes)
{
    extern __shared__ __align__(16) unsigned char smem[];
    unsigned long long *lh = reinterpret_cast<unsigned long long *>(smem); // SB x FS, feature-minor
    u32 const Dw = Fs / 4u;
    u32 const R = 64u / Dw;
    u32 const lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    u32 const NW = blockDim.x >> 6;
    u32 const lr = lane / Dw, d = lane - lr * Dw;
    bool const lane_ok = lr < R;
    u32 const lcells = SB * FS;
    u32 const n_units = ctl->n_hist_chunks;
    size_t const node_cells = static_cast<size_t>(F) * SB;
    for (u32 c = blockIdx.x; c < n_units; c += gridDim.x) {
        HistChunk const ch = chunks[c];
        for (u32 i = threadIdx.x; i < lcells; i += blockDim.x) lh[i] = 0ull;
        __syncthreads();
        u8 const *plane = ch.src == 0 ? pl0 : (ch.src == 1 ? pl1 : plp);
        u32 const *ghv = static_cast<u32 const *>(ch.src == 0 ? gw0 : (ch.src == 1 ? gw1 : gwp));
        u8 const *cb = plane + 4u * d;
        u32 const stride = NW * R;
        HqRoute rt{};
        if (ch.route != kNone) {
            PartOp const &op = routes[ch.route];
            rt.feature = op.feature;
            rt.bin = op.bin;
            rt.dl = op.default_left;
            rt.last = op.last_bin;
            rt.want_left = ch.side == 0 ? 1u : 0u;
            rt.in_slice = 1u;
            rt.lane_d = op.feature >> 2;
            rt.shift = 8u * (op.feature & 3u);
        }
        if (lane_ok) {
            u32 const r = ch.begin + wave * R + lr;
            if (ch.route != kNone)
                hist_q5_rows<true>(lh, plane, ghv, cb, Fs, F, FS, d, lr, Dw, r, ch.end, stride, rt);
            else
                hist_q5_rows<false>(lh, plane, ghv, cb, Fs, F, FS, d, lr, Dw, r, ch.end, stride, rt);
        }
        __syncthreads();
        u32 const out_cells = F * SB;
        if (ch.part != kNone) {
            u64 *P = static_cast<u64 *>(partials) + static_cast<size_t>(ch.part) * node_cells;
            for (u32 i = threadIdx.x; i < out_cells; i += blockDim.x) {
                u32 const b = i / F, f = i - b * F;
                P[i] = lh[b * FS + f];
            }
        } else {
            size_t const base = static_cast<size_t>(ch.slot) * node_cells;
            i64 *dst = static_cast<i64 *>(hist_next) + base * 2;
            for (u32 i = threadIdx.x; i < out_cells; i += blockDim.x) {
                u32 const b = i / F, f = i - b * F;
                u64 const v = lh[b * FS + f];
                size_t const o = static_cast<size_t>(f) * SB + b;
                dst[2 * o] = unpack_g(v);
                dst[2 * o + 1] = unpack_h(v);
                if (cnt_next) cnt_next[base + o] = unpack_cnt(v);
            }
        }
        __syncthreads();
    }
}

// Quantized histogram, row-in-lane layout (rows of at most 64 bytes, one feature slice;
// same contract as k_hist_q).  Lane l of a wave owns whole rows: dword k of 64
// consecutive rows is one load instruction spanning 64 rows (the lines of a 64-row
// block are fetched once and re-read from cache for the other dwords), the gradient word
// is loaded once per row and packed once, and U rows per lane are in flight.  Atomic
// instruction (row, k, j) puts 64 rows' bins of one feature into the LDS cells.  A
// routed unit's filtered rows skip their atomics.
template <u32 DW, u32 U>
__global__ __launch_bounds__(kHqThreads) void k_hist_q3(
    HistChunk const *__restrict__ chunks, TreeCtl const *__restrict__ ctl,
    u8 const *__restrict__ pl0, u8 const *__restrict__ pl1, u8 const *__restrict__ plp,
    void const *__restrict__ gw0, void const *__restrict__ gw1, void const *__restrict__ gwp,
    u32 Fs, u32 F, u32 SB, void *__restrict__ partials, void *__restrict__ hist_next,
    u32 *__restrict__ cnt_next, PartOp const *__restrict__ routes)
{
    extern __shared__ __align__(16) unsigned char smem[];
    unsigned long long *lh = reinterpret_cast<unsigned long long *>(smem); // F x SB
    u32 const lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    u32 const NW = blockDim.x >> 6;
    u32 const cells = F * SB;
    u32 const n_units = ctl->n_hist_chunks;
    size_t const node_cells = static_cast<size_t>(F) * SB;
    for (u32 c = blockIdx.x; c < n_units; c += gridDim.x) {
        HistChunk const ch = chunks[c];
        for (u32 i = threadIdx.x; i < cells; i += blockDim.x) lh[i] = 0ull;
        __syncthreads();
        u8 const *plane = ch.src == 0 ? pl0 : (ch.src == 1 ? pl1 : plp);
        u32 const *ghv = static_cast<u32 const *>(ch.src == 0 ? gw0 : (ch.src == 1 ? gw1 : gwp));
        bool const routed = ch.route != kNone;
        PartOp rop{};
        if (routed) rop = routes[ch.route];
        bool const want_left = ch.side == 0;
        u32 const rf_dw = rop.feature >> 2, rf_sh = 8u * (rop.feature & 3u);
        for (u32 s0 = ch.begin; s0 < ch.end; s0 += kHistChunk) {
        u32 const s1 = min(ch.end, s0 + kHistChunk);
        for (u32 r0 = s0 + wave * (U * 64u); r0 < s1; r0 += NW * U * 64u) {
            u32 dw[U][DW], v[U];
            bool ok[U];
#pragma unroll
            for (u32 u = 0; u < U; ++u) {
                u32 const row = r0 + u * 64u + lane;
                ok[u] = row < s1;
                u32 const *rp = reinterpret_cast<u32 const *>(plane + static_cast<size_t>(ok[u] ? row : ch.begin) * Fs);
#pragma unroll
                for (u32 k = 0; k < DW; ++k) dw[u][k] = rp[k];
                v[u] = ghv[ok[u] ? row : ch.begin];
            }
#pragma unroll
            for (u32 u = 0; u < U; ++u) {
                bool keep = ok[u];
                if (routed) {
                    u32 w = 0;
#pragma unroll
                    for (u32 k = 0; k < DW; ++k)
                        if (k == rf_dw) w = dw[u][k];
                    keep = keep && routes_left((w >> rf_sh) & 0xFFu, rop.last_bin, rop.bin, rop.default_left) ==
                                       want_left;
                }
                if (!keep) continue;
                u64 const pk = pack_q(v[u]);
#pragma unroll
                for (u32 k = 0; k < DW; ++k)
#pragma unroll
                    for (u32 j = 0; j < 4; ++j) {
                        u32 const f = 4u * k + j;
                        if (f < F)
                            atomicAdd(lh + f * SB + ((dw[u][k] >> (8u * j)) & 0xFFu),
                                      static_cast<unsigned long long>(pk));
                    }
            }
        }
        __syncthreads();
        hq_flush(lh, cells, ch, node_cells, 0, s0 == ch.begin, s1 < ch.end, partials, hist_next, cnt_next);
        __syncthreads();
        } // sub-units
    }
}

// Quantized histogram for narrow rows (F <= 32), feature-in-lane layout (H6, same
// contract as k_hist_q).  A tile of the unit's rows is staged in LDS with 16-byte loads;
// one ds_add_u64 then covers two rows x the row's features (lane = row parity x
// feature column) on an LDS histogram laid out [bin][32]: column f always maps to bank
// pair f, so the only possible bank conflict is between the two rows of one feature.
// The random bins therefore no longer serialise the LDS atomics.
constexpr u32 kHq2Cols = 32;
constexpr int kHq2Unroll = 4;
// LDS per workgroup kept under half of the CU's 160 KiB so two workgroups share a CU
// and overlap one another's staging and accumulation.
constexpr u32 kHq2LdsBudget = 81408;

__global__ __launch_bounds__(512) void k_hist_q2(
    HistChunk const *__restrict__ chunks, TreeCtl const *__restrict__ ctl,
    u8 const *__restrict__ pl0, u8 const *__restrict__ pl1, u8 const *__restrict__ plp,
    void const *__restrict__ gw0, void const *__restrict__ gw1, void const *__restrict__ gwp,
    u32 Fs, u32 F, u32 SB, u32 tile, void *__restrict__ partials, void *__restrict__ hist_next,
    u32 *__restrict__ cnt_next, PartOp const *__restrict__ routes)
{
    extern __shared__ __align__(16) unsigned char smem[];
    unsigned long long *lh = reinterpret_cast<unsigned long long *>(smem);            // SB x 32
    uint4 *s_rows = reinterpret_cast<uint4 *>(smem + static_cast<size_t>(SB) * kHq2Cols * 8u); // tile*Fs+16
    uint4 *s_gh = reinterpret_cast<uint4 *>(reinterpret_cast<unsigned char *>(s_rows) + tile * Fs + 16u);
    u32 const lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    u32 const NW = blockDim.x >> 6;
    u32 const col = lane & 31u, half = lane >> 5;
    bool const col_ok = col < F;
    u32 const lds_cells = SB * kHq2Cols;
    u32 const n_units = ctl->n_hist_chunks;
    size_t const node_cells = static_cast<size_t>(F) * SB;

    for (u32 c = blockIdx.x; c < n_units; c += gridDim.x) {
        HistChunk const ch = chunks[c];
        for (u32 i = threadIdx.x; i < lds_cells; i += blockDim.x) lh[i] = 0ull;
        u8 const *plane = ch.src == 0 ? pl0 : (ch.src == 1 ? pl1 : plp);
        u8 const *ghb = static_cast<u8 const *>(ch.src == 0 ? gw0 : (ch.src == 1 ? gw1 : gwp));
        // Routed unit: a row counts only when the parent's split sends it to ch.side.
        bool const routed = ch.route != kNone;
        PartOp rop{};
        if (routed) rop = routes[ch.route];
        bool const want_left = ch.side == 0;
        for (u32 t0 = ch.begin; t0 < ch.end; t0 += tile) {
            u32 const rows = min(tile, ch.end - t0);
            __syncthreads(); // zeroing done / previous tile consumed
            u32 const sh = lds_load_range(plane, static_cast<size_t>(t0) * Fs, rows * Fs, s_rows);
            u32 const shg = lds_load_range(ghb, static_cast<size_t>(t0) * 4u, rows * 4u, s_gh);
            __syncthreads();
            u8 const *rb = reinterpret_cast<u8 const *>(s_rows) + sh + col;
            u8 const *rtb = reinterpret_cast<u8 const *>(s_rows) + sh + rop.feature;
            u32 const *gb = reinterpret_cast<u32 const *>(reinterpret_cast<u8 const *>(s_gh) + shg);
            auto keep = [&](u32 rr) -> bool {
                return !routed ||
                       routes_left(rtb[rr * Fs], rop.last_bin, rop.bin, rop.default_left) == want_left;
            };
            if (col_ok) {
                u32 const stride = 2u * NW;
                u32 r = 2u * wave + half;
                for (; r + (kHq2Unroll - 1) * stride < rows; r += kHq2Unroll * stride) {
                    u32 bin[kHq2Unroll], v[kHq2Unroll];
#pragma unroll
                    for (int u = 0; u < kHq2Unroll; ++u) {
                        u32 const rr = r + u * stride;
                        bin[u] = rb[rr * Fs];
                        v[u] = gb[rr];
                    }
#pragma unroll
                    for (int u = 0; u < kHq2Unroll; ++u)
                        if (keep(r + u * stride))
                            atomicAdd(lh + bin[u] * kHq2Cols + col,
                                      static_cast<unsigned long long>(pack_q(v[u])));
                }
                for (; r < rows; r += stride)
                    if (keep(r))
                        atomicAdd(lh + static_cast<u32>(rb[r * Fs]) * kHq2Cols + col,
                                  static_cast<unsigned long long>(pack_q(gb[r])));
            }
        }
        __syncthreads();
        // Flush in (bin, feature) order: LDS reads stay on distinct bank pairs.
        u32 const out_cells = F * SB;
        if (ch.part != kNone) { // packed partial (units of one 8192-row sub-unit)
            u64 *dst = static_cast<u64 *>(partials) + static_cast<size_t>(ch.part) * node_cells;
            for (u32 i = threadIdx.x; i < out_cells; i += blockDim.x) {
                u32 const b = i / F, fl = i - b * F;
                dst[static_cast<size_t>(fl) * SB + b] = lh[b * kHq2Cols + fl];
            }
        } else {
            size_t const base = static_cast<size_t>(ch.slot) * node_cells;
            i64 *dst = reinterpret_cast<i64 *>(hist_next) + base * 2;
            for (u32 i = threadIdx.x; i < out_cells; i += blockDim.x) {
                u32 const b = i / F, fl = i - b * F;
                u64 const v = lh[b * kHq2Cols + fl];
                size_t const o = static_cast<size_t>(fl) * SB + b;
                dst[2 * o] = unpack_g(v);
                dst[2 * o + 1] = unpack_h(v);
                if (cnt_next) cnt_next[base + o] = unpack_cnt(v);
            }
        }
        __syncthreads();
    }
}

// Quantized histogram, rotated row-in-lane layout (rows of at most 64 bytes, one feature
// slice, single 8192-row sub-units; same contract as k_hist_q otherwise).  Lane l owns
// whole rows like k_hist_q3, but rotates which feature it adds at each atomic
// instruction: its k-th dword load takes dword (k + l) mod DW of the row and byte
// (j + l / DW) mod 4 of it, so the 64 lanes of one instruction spread over all 4 * DW
// features instead of sharing one.  The LDS cells are feature-minor, cell (f, b) at
// b * FS + f, so a cell's bank is a function of its feature alone: the lanes of an
// instruction land on distinct bank pairs up to the ~2 lanes per feature, and same-cell
// collisions (few-valued features) fall to those same ~2 lanes.  Multi-unit nodes write
// packed partials in that (bin, feature) order (F cells per bin, coalesced); the
// completion kernel reads them with the transposed index.  Single-unit nodes write
// their slot directly.
constexpr u32 hq4_fs(u32 F) { return F <= 16u ? 16u : (F <= 32u ? 32u : 64u); }

template <u32 DW, u32 U>
__global__ __launch_bounds__(kHqThreads) void k_hist_q4(
    HistChunk const *__restrict__ chunks, TreeCtl const *__restrict__ ctl,
    u8 const *__restrict__ pl0, u8 const *__restrict__ pl1, u8 const *__restrict__ plp,
    void const *__restrict__ gw0, void const *__restrict__ gw1, void const *__restrict__ gwp,
    u32 Fs, u32 F, u32 SB, void *__restrict__ partials, void *__restrict__ hist_next,
    u32 *__restrict__ cnt_next, PartOp const *__restrict__ routes)
{
    extern __shared__ __align__(16) unsigned char smem[];
    unsigned long long *lh = reinterpret_cast<unsigned long long *>(smem); // SB x FS, feature-minor
    constexpr u32 FS = hq4_fs(4u * DW);
    u32 const lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    u32 const NW = blockDim.x >> 6;
    u32 const lcells = SB * FS;
    u32 const n_units = ctl->n_hist_chunks;
    size_t const node_cells = static_cast<size_t>(F) * SB;
    // this lane's rotation: dword offset and byte offset
    u32 const rot_d = lane % DW, rot_b = (lane / DW) & 3u;
    for (u32 c = blockIdx.x; c < n_units; c += gridDim.x) {
        HistChunk const ch = chunks[c];
        for (u32 i = threadIdx.x; i < lcells; i += blockDim.x) lh[i] = 0ull;
        __syncthreads();
        u8 const *plane = ch.src == 0 ? pl0 : (ch.src == 1 ? pl1 : plp);
        u32 const *ghv = static_cast<u32 const *>(ch.src == 0 ? gw0 : (ch.src == 1 ? gw1 : gwp));
        bool const routed = ch.route != kNone;
        PartOp rop{};
        if (routed) rop = routes[ch.route];
        bool const want_left = ch.side == 0;
        u32 const rf_dw = rop.feature >> 2, rf_sh = 8u * (rop.feature & 3u);
        for (u32 r0 = ch.begin + wave * (U * 64u); r0 < ch.end; r0 += NW * U * 64u) {
            u32 dw[U][DW], v[U], rdw[U];
            bool ok[U];
#pragma unroll
            for (u32 u = 0; u < U; ++u) {
                u32 const row = r0 + u * 64u + lane;
                ok[u] = row < ch.end;
                u32 const *rp = reinterpret_cast<u32 const *>(plane + static_cast<size_t>(ok[u] ? row : ch.begin) * Fs);
#pragma unroll
                for (u32 k = 0; k < DW; ++k) {
                    u32 m = rot_d + k;
                    m = m >= DW ? m - DW : m;
                    dw[u][k] = rp[m];
                }
                v[u] = ghv[ok[u] ? row : ch.begin];
                rdw[u] = routed ? rp[rf_dw] : 0u;
            }
#pragma unroll
            for (u32 u = 0; u < U; ++u) {
                bool keep = ok[u];
                if (routed)
                    keep = keep && routes_left((rdw[u] >> rf_sh) & 0xFFu, rop.last_bin, rop.bin, rop.default_left) ==
                                       want_left;
                if (!keep) continue;
                u64 const pk = pack_q(v[u]);
#pragma unroll
                for (u32 k = 0; k < DW; ++k) {
                    u32 m = rot_d + k;
                    m = m >= DW ? m - DW : m;
#pragma unroll
                    for (u32 j = 0; j < 4; ++j) {
                        u32 const jj = (j + rot_b) & 3u;
                        u32 const f = 4u * m + jj;
                        if (f < F)
                            atomicAdd(lh + ((dw[u][k] >> (8u * jj)) & 0xFFu) * FS + f,
                                      static_cast<unsigned long long>(pk));
                    }
                }
            }
        }
        __syncthreads();
        u32 const out_cells = F * SB;
        if (ch.part != kNone) {
            // packed partial, (bin, feature) order: coalesced stores, conflict-free LDS reads
            u64 *P = static_cast<u64 *>(partials) + static_cast<size_t>(ch.part) * node_cells;
            for (u32 i = threadIdx.x; i < out_cells; i += blockDim.x) {
                u32 const b = i / F, f = i - b * F;
                P[i] = lh[b * FS + f];
            }
        } else {
            size_t const base = static_cast<size_t>(ch.slot) * node_cells;
            i64 *dst = static_cast<i64 *>(hist_next) + base * 2;
            for (u32 i = threadIdx.x; i < out_cells; i += blockDim.x) {
                u32 const b = i / F, f = i - b * F;
                u64 const v = lh[b * FS + f];
                size_t const o = static_cast<size_t>(f) * SB + b;
                dst[2 * o] = unpack_g(v);
                dst[2 * o + 1] = unpack_h(v);
                if (cnt_next) cnt_next[base + o] = unpack_cnt(v);
            }
        }
        __syncthreads();
    }
}

// Fixed-order completion of built nodes (H6 reduce + H7 sibling subtraction,
// src/step/primitives.hpp:156-173; device twin subtract_kernel, kernels.cuh:507-520):
// small = p_0 + p_1 + ... in unit order, large = parent - small.
template <bool Q>
__global__ void k_finish(FinishOp const *__restrict__ ops, TreeCtl const *__restrict__ ctl,
                         void const *__restrict__ partials, void const *__restrict__ hist_cur,
                         void *__restrict__ hist_next, u32 node_cells, u32 const *__restrict__ cnt_cur,
                         u32 *__restrict__ cnt_next)
{
    u32 const n_ops = ctl->n_finish;
    for (u32 o = blockIdx.y; o < n_ops; o += gridDim.y) {
        FinishOp const op = ops[o];
        for (u32 i = blockIdx.x * blockDim.x + threadIdx.x; i < node_cells; i += gridDim.x * blockDim.x) {
            if (Q) {
                i64 *nx = reinterpret_cast<i64 *>(hist_next);
                size_t const si = static_cast<size_t>(op.small_slot) * node_cells + i;
                i64 g = 0, h = 0;
                u32 n = 0;
                if (op.empty) {
                    nx[si * 2] = 0;
                    nx[si * 2 + 1] = 0;
                    if (cnt_next) cnt_next[si] = 0;
                } else if (op.part_count == 0) {
                    g = nx[si * 2];
                    h = nx[si * 2 + 1];
                    if (cnt_next) n = cnt_next[si];
                } else {
                    // Partials hold packed unit sums: unpack each before adding.
                    u64 const *pp = reinterpret_cast<u64 const *>(partials) + i;
                    u32 p = 0;
                    for (; p + 8 <= op.part_count; p += 8) {
                        u64 v[8];
#pragma unroll
                        for (int k = 0; k < 8; ++k) v[k] = pp[static_cast<size_t>(op.part_first + p + k) * node_cells];
#pragma unroll
                        for (int k = 0; k < 8; ++k) {
                            g += unpack_g(v[k]);
                            h += unpack_h(v[k]);
                            n += unpack_cnt(v[k]);
                        }
                    }
                    for (; p < op.part_count; ++p) {
                        u64 const v = pp[static_cast<size_t>(op.part_first + p) * node_cells];
                        g += unpack_g(v);
                        h += unpack_h(v);
                        n += unpack_cnt(v);
                    }
                    nx[si * 2] = g;
                    nx[si * 2 + 1] = h;
                    if (cnt_next) cnt_next[si] = n;
                }
                if (op.large_slot != kNone) {
                    i64 const *pc = reinterpret_cast<i64 const *>(hist_cur);
                    size_t const pi = static_cast<size_t>(op.parent_slot) * node_cells + i;
                    size_t const li = static_cast<size_t>(op.large_slot) * node_cells + i;
                    nx[li * 2] = pc[pi * 2] - g;
                    nx[li * 2 + 1] = pc[pi * 2 + 1] - h;
                    if (cnt_next) cnt_next[li] = cnt_cur[pi] - n;
                }
            } else {
                float2 *nx = reinterpret_cast<float2 *>(hist_next);
                float2 s = make_float2(0.0F, 0.0F);
                if (op.empty) {
                    nx[static_cast<size_t>(op.small_slot) * node_cells + i] = s;
                } else if (op.part_count == 0) {
                    s = nx[static_cast<size_t>(op.small_slot) * node_cells + i];
                } else {
                    float2 const *pp = reinterpret_cast<float2 const *>(partials) + i;
                    u32 p = 0;
                    for (; p + 8 <= op.part_count; p += 8) {
                        float2 v[8];
#pragma unroll
                        for (int k = 0; k < 8; ++k) v[k] = pp[static_cast<size_t>(op.part_first + p + k) * node_cells];
#pragma unroll
                        for (int k = 0; k < 8; ++k) {
                            s.x += v[k].x;
                            s.y += v[k].y;
                        }
                    }
                    for (; p < op.part_count; ++p) {
                        float2 const v = pp[static_cast<size_t>(op.part_first + p) * node_cells];
                        s.x += v.x;
                        s.y += v.y;
                    }
                    nx[static_cast<size_t>(op.small_slot) * node_cells + i] = s;
                }
                if (op.large_slot != kNone) {
                    float2 const pv =
                        reinterpret_cast<float2 const *>(hist_cur)[static_cast<size_t>(op.parent_slot) * node_cells + i];
                    nx[static_cast<size_t>(op.large_slot) * node_cells + i] =
                        make_float2(pv.x - s.x, pv.y - s.y);
                }
            }
        }
    }
}

// Quantized completion with the partial range spread over threads (integer sums are
// exact in any order).  A 256-thread workgroup owns kFqCells consecutive cells of one
// op: thread (g, j) sums partials g, g + kFqGroups, ... of cell j (coalesced across the
// cells), the kFqGroups group sums meet in LDS, and one thread per cell writes the small
// child and large = parent - small (with counts when cnt_next is given).
constexpr u32 kFqCells = 16;
constexpr u32 kFqGroups = 16;

// advance != 0: the level ends with this launch, and workgroup (0, 0) also moves the
// control block to the next frontier (no workgroup of the launch reads it).
// tF != 0: the packed partials are in (bin, feature) order with tF features (k_hist_q4):
// a workgroup's cells are consecutive in that order, and each cell's node-layout index
// (feature * SB + bin) is used for the writes and the parent.
__global__ __launch_bounds__(256) void k_finish_q(FinishOp const *__restrict__ ops, TreeCtl *__restrict__ ctl,
                                                  void const *__restrict__ partials, void const *__restrict__ hist_cur,
                                                  void *__restrict__ hist_next, u32 node_cells,
                                                  u32 const *__restrict__ cnt_cur, u32 *__restrict__ cnt_next,
                                                  u32 advance, u32 tF)
{
    __shared__ i64 s_g[kFqGroups][kFqCells], s_h[kFqGroups][kFqCells];
    __shared__ u32 s_n[kFqGroups][kFqCells];
    u32 const n_ops = ctl->n_finish;
    if (advance && blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x == 0) {
        ctl->n_front = ctl->next_front;
        ctl->depth += 1;
    }
    u32 const j = threadIdx.x % kFqCells, grp = threadIdx.x / kFqCells;
    u32 const n_tiles = (node_cells + kFqCells - 1) / kFqCells;
    for (u32 o = blockIdx.y; o < n_ops; o += gridDim.y) {
        FinishOp const op = ops[o];
        i64 *nx = reinterpret_cast<i64 *>(hist_next);
        for (u32 tile = blockIdx.x; tile < n_tiles; tile += gridDim.x) {
            u32 const pi_ = tile * kFqCells + j;          // cell index in the partials' order
            bool const in = pi_ < node_cells;
            u32 i = pi_;                                  // cell index in the node layout
            if (tF) {
                u32 const b = pi_ / tF, f = pi_ - b * tF;
                i = f * (node_cells / tF) + b;
            }
            i64 g = 0, h = 0;
            u32 n = 0;
            if (in && op.part_count > 0 && !op.wide) {
                u64 const *pp = static_cast<u64 const *>(partials) + pi_;
                u32 p = grp;
                for (; p + 3 * kFqGroups < op.part_count; p += 4 * kFqGroups) {
                    u64 v[4];
#pragma unroll
                    for (int k = 0; k < 4; ++k)
                        v[k] = pp[static_cast<size_t>(op.part_first + p + k * kFqGroups) * node_cells];
#pragma unroll
                    for (int k = 0; k < 4; ++k) {
                        g += unpack_g(v[k]);
                        h += unpack_h(v[k]);
                        n += unpack_cnt(v[k]);
                    }
                }
                for (; p < op.part_count; p += kFqGroups) {
                    u64 const v = pp[static_cast<size_t>(op.part_first + p) * node_cells];
                    g += unpack_g(v);
                    h += unpack_h(v);
                    n += unpack_cnt(v);
                }
            } else if (in && op.part_count > 0) {
                // wide partials: three u32 planes (g as int32, h, count) per partial
                u32 const *pp = static_cast<u32 const *>(partials) + i;
                size_t const pst = static_cast<size_t>(node_cells) * 3u;
                u32 p = grp;
                for (; p + 3 * kFqGroups < op.part_count; p += 4 * kFqGroups) {
                    u32 vg[4], vh[4], vn[4];
#pragma unroll
                    for (int k = 0; k < 4; ++k) {
                        u32 const *q = pp + static_cast<size_t>(op.part_first + p + k * kFqGroups) * pst;
                        vg[k] = q[0];
                        vh[k] = q[node_cells];
                        vn[k] = q[2 * node_cells];
                    }
#pragma unroll
                    for (int k = 0; k < 4; ++k) {
                        g += static_cast<int32_t>(vg[k]);
                        h += vh[k];
                        n += vn[k];
                    }
                }
                for (; p < op.part_count; p += kFqGroups) {
                    u32 const *q = pp + static_cast<size_t>(op.part_first + p) * pst;
                    g += static_cast<int32_t>(q[0]);
                    h += q[node_cells];
                    n += q[2 * node_cells];
                }
            }
            s_g[grp][j] = g;
            s_h[grp][j] = h;
            s_n[grp][j] = n;
            __syncthreads();
            if (grp == 0 && in) {
                size_t const si = static_cast<size_t>(op.small_slot) * node_cells + i;
                if (op.empty) {
                    g = h = 0;
                    n = 0;
                    nx[si * 2] = 0;
                    nx[si * 2 + 1] = 0;
                    if (cnt_next) cnt_next[si] = 0;
                } else if (op.part_count == 0) {
                    g = nx[si * 2];
                    h = nx[si * 2 + 1];
                    n = cnt_next ? cnt_next[si] : 0u;
                } else {
                    for (u32 k = 1; k < kFqGroups; ++k) {
                        g += s_g[k][j];
                        h += s_h[k][j];
                        n += s_n[k][j];
                    }
                    nx[si * 2] = g;
                    nx[si * 2 + 1] = h;
                    if (cnt_next) cnt_next[si] = n;
                }
                if (op.large_slot != kNone) {
                    i64 const *pc = reinterpret_cast<i64 const *>(hist_cur);
                    size_t const pi = static_cast<size_t>(op.parent_slot) * node_cells + i;
                    size_t const li = static_cast<size_t>(op.large_slot) * node_cells + i;
                    nx[li * 2] = pc[pi * 2] - g;
                    nx[li * 2 + 1] = pc[pi * 2 + 1] - h;
                    if (cnt_next) cnt_next[li] = cnt_cur[pi] - n;
                }
            }
            __syncthreads();
        }
    }
}

// ---------------------------------------------------------------------------------
// Node split finder (H8): device twin of find_kernel + reduce_kernel
// (src/cuda/detail/kernels.cuh:586-754) with the CPU finder's exact semantics
// (src/split.cpp:33-162, :255-297): double running sums in bin order, dl = true before
// false, strict > (earliest bin, then lowest feature, wins ties), an empty missing cell
// scans one direction, rows < 2 * min_data_in_leaf cannot split.
//
// Up to kFindMaxBlocks workgroups per frontier node, each scanning kFindTile-feature
// tiles.  Phase 0 derives the node totals from its first feature's cells
// (Histogram::totals, include/bonsai/histogram.hpp:76-87) unless a parent split already
// set them (quantized: every feature's cells sum to the same exact totals, so a block
// takes them from its first staged feature).  Phase A: the per-feature prefix in bin
// order into LDS (FP32: one thread per feature, sequential double sums; quantized: one
// wave per feature, and the count cells' prefix beside it).  Phase B: every (feature,
// bin, direction) candidate is scored in parallel; phase C reduces to the block's winner
// with the serial scan's tie order, and the node's last block merges the blocks'
// winners the same way.
// ---------------------------------------------------------------------------------
u32 find_blocks_per_node(u32 F, u32 tile)
{
    return std::max(1u, std::min<u32>(std::min(tuning().find_blocks, kFindMaxBlocks), (F + tile - 1) / tile));
}

template <bool Q, bool CON>
__global__ __launch_bounds__(kFindThreads) void k_find(NodeInfo *__restrict__ front, TreeCtl const *__restrict__ ctl,
                                                     void const *__restrict__ hist, u16 const *__restrict__ n_bins,
                                                     u32 F, u32 SB, GrowCfg cfg, double const *__restrict__ scales,
                                                     SplitRec *__restrict__ out, u32 front_base, u32 n_front_override,
                                                     SplitRec *__restrict__ part, u32 *__restrict__ done,
                                                     u32 const *__restrict__ cnt)
{
    find_node<Q, CON>(front, ctl, hist, n_bins, F, SB, cfg, scales, out, front_base, n_front_override, part, done,
                      cnt);
}

// ---------------------------------------------------------------------------------
// Level control (H13 depthwise): plan_level (src/grower_impl.hpp:212-255) and
// commit_split_node (:188-210) on device.  One workgroup.  Splits get child ids in
// frontier order; non-splits become leaves; the split ops and their partition units
// are emitted for the level's partition.
// ---------------------------------------------------------------------------------

// Index of the last item whose `first` is <= c (items sorted by first).
template <typename T, typename Get>
__device__ __forceinline__ u32 find_owner(T const *items, u32 n, u32 c, Get get)
{
    u32 lo = 0, hi = n; // answer in [lo, hi)
    while (hi - lo > 1) {
        u32 const mid = (lo + hi) >> 1;
        if (get(items[mid]) <= c) lo = mid;
        else hi = mid;
    }
    return lo;
}

__global__ __launch_bounds__(1024) void k_plan_level(NodeInfo const *__restrict__ front,
                                                     SplitRec const *__restrict__ splits,
                                                     TreeCtl *__restrict__ ctl, DevNode *__restrict__ tree,
                                                     PartOp *__restrict__ ops, PartChunk *__restrict__ pch,
                                                     u16 const *__restrict__ n_bins, GrowCfg cfg,
                                                     double const *__restrict__ scales, u32 dst,
                                                     u32 chunk_rows)
{
    __shared__ u32 scratch[1024];
    __shared__ u32 s_base_nodes, s_ops, s_chunks;
    u32 const P = ctl->n_front;
    if (threadIdx.x == 0) {
        s_base_nodes = ctl->n_nodes;
        s_ops = 0;
        s_chunks = 0;
    }
    __syncthreads();
    for (u32 t0 = 0; t0 < P; t0 += blockDim.x) {
        u32 const i = t0 + threadIdx.x;
        bool const in = i < P;
        u32 const valid = in ? splits[i].valid : 0u;
        u32 tot;
        u32 const rank = block_exscan(valid, scratch, &tot) + s_ops;
        u32 nch = 0;
        if (in && valid) nch = (front[i].cnt + chunk_rows - 1) / chunk_rows;
        u32 ctot;
        u32 const cfirst = block_exscan(nch, scratch, &ctot) + s_chunks;
        if (in) {
            NodeInfo const nd = front[i];
            if (valid) {
                SplitRec const s = splits[i];
                u32 const lid = s_base_nodes + 2 * rank, rid = lid + 1;
                DevNode pn = tree[nd.id];
                pn.feature = static_cast<int32_t>(s.feature);
                pn.bin = s.bin;
                pn.default_left = s.default_left;
                pn.left = lid;
                pn.right = rid;
                pn.value = 0.0F;
                pn.gain = static_cast<float>(s.gain);
                tree[nd.id] = pn;
                DevNode leaf{};
                leaf.feature = -1;
                tree[lid] = leaf;
                tree[rid] = leaf;
                PartOp op{};
                op.parent = i;
                op.feature = s.feature;
                op.bin = s.bin;
                op.default_left = s.default_left;
                op.last_bin = static_cast<u32>(n_bins[s.feature]) - 1u;
                op.off = nd.off;
                op.cnt = nd.cnt;
                op.src = nd.src;
                op.dst = dst;
                op.chunk_first = cfirst;
                op.chunk_count = nch;
                op.left_id = lid;
                op.right_id = rid;
                op.alive = 1;
                ops[rank] = op;
                fd_propagate_bounds(cfg, scales, nd.id, s, lid, rid);
                fd_propagate_path(cfg, nd.id, s.feature, lid, rid);
            } else {
                write_leaf(tree, nd.id, node_leaf_value(nd, cfg, scales));
            }
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            s_ops += tot;
            s_chunks += ctot;
        }
        __syncthreads();
    }
    // Expand the ops' unit ranges into the flat unit list.
    u32 const n_ops = s_ops, n_ch = s_chunks;
    for (u32 c = threadIdx.x; c < n_ch; c += blockDim.x) {
        u32 const j = find_owner(ops, n_ops, c, [](PartOp const &o) { return o.chunk_first; });
        PartOp const op = ops[j];
        u32 const k = c - op.chunk_first;
        PartChunk pc;
        pc.op = j;
        pc.begin = k * chunk_rows;
        pc.end = min(op.cnt, (k + 1) * chunk_rows);
        pc.left_before = 0;
        pch[c] = pc;
    }
    if (threadIdx.x == 0) {
        ctl->n_nodes = s_base_nodes + 2 * n_ops;
        ctl->n_ops = n_ops;
        ctl->n_part_chunks = n_ch;
    }
}

// Partition count (H10, route_count_kernel, kernels.cuh:332-370): left rows per unit.
__global__ __launch_bounds__(256) void k_part_count(PartChunk const *__restrict__ pch, PartOp const *__restrict__ ops,
                                                    TreeCtl const *__restrict__ ctl, u8 const *__restrict__ pl0,
                                                    u8 const *__restrict__ pl1, u8 const *__restrict__ plp, u32 Fs,
                                                    u32 *__restrict__ chunk_left)
{
    __shared__ u32 red[4];
    u32 const n = ctl->n_part_chunks;
    for (u32 c = blockIdx.x; c < n; c += gridDim.x) {
        PartChunk const ch = pch[c];
        PartOp const op = ops[ch.op];
        u8 const *plane = op.src == 0 ? pl0 : (op.src == 1 ? pl1 : plp);
        u32 cnt = 0;
        u8 const *col = plane + static_cast<size_t>(op.off) * Fs + op.feature;
        u32 k = ch.begin + threadIdx.x;
        for (; k + 7u * blockDim.x < ch.end; k += 8u * blockDim.x) {
            u32 b[8];
#pragma unroll
            for (u32 u = 0; u < 8; ++u) b[u] = col[static_cast<size_t>(k + u * blockDim.x) * Fs];
#pragma unroll
            for (u32 u = 0; u < 8; ++u) cnt += routes_left(b[u], op.last_bin, op.bin, op.default_left) ? 1u : 0u;
        }
        for (; k < ch.end; k += blockDim.x)
            cnt += routes_left(col[static_cast<size_t>(k) * Fs], op.last_bin, op.bin, op.default_left) ? 1u : 0u;
        for (int o = 32; o > 0; o >>= 1) cnt += __shfl_xor(cnt, o);
        if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = cnt;
        __syncthreads();
        if (threadIdx.x == 0) chunk_left[c] = red[0] + red[1] + red[2] + red[3];
        __syncthreads();
    }
}

// Partition scan + demotion + next frontier (H10 seg_scan_kernel, kernels.cuh:376-413;
// demote_empty_splits, src/grower_impl.hpp:263-297; commit_children :302-324;
// smaller_child tie -> left, src/step/primitives.hpp:105-108).  One workgroup.
// oblivious = 1 (levelwise): no demotion, no tree nodes (the level split is broadcast).
template <bool Q>
__global__ __launch_bounds__(1024) void k_part_scan(
    PartOp *__restrict__ ops, PartChunk *__restrict__ pch, u32 const *__restrict__ chunk_left,
    TreeCtl *__restrict__ ctl, NodeInfo const *__restrict__ front, NodeInfo *__restrict__ next,
    SplitRec const *__restrict__ splits, DevNode *__restrict__ tree, HistChunk *__restrict__ hch,
    FinishOp *__restrict__ fops, GrowCfg cfg, double const *__restrict__ scales, u32 is_last,
    u32 oblivious)
{
    __shared__ u32 scratch[1024];
    __shared__ u32 s_run, s_alive, s_hch, s_parts;
    u32 const n_ops = ctl->n_ops;
    u32 const n_ch = ctl->n_part_chunks;
    if (threadIdx.x == 0) {
        s_run = 0;
        s_alive = 0;
        s_hch = 0;
        s_parts = 0;
    }
    __syncthreads();
    // 1) global exclusive prefix of left counts over all units (units of an op are
    //    contiguous and ops are in order).
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
    u32 const total_left = s_run;
    // 2) per op: n_left from the prefix, demotion, covers.
    for (u32 j = threadIdx.x; j < n_ops; j += blockDim.x) {
        PartOp op = ops[j];
        u32 const b0 = op.chunk_count ? pch[op.chunk_first].left_before : 0u;
        u32 const e = op.chunk_first + op.chunk_count;
        u32 const b1 = op.chunk_count ? (e < n_ch ? pch[e].left_before : total_left) : 0u;
        u32 const nL = b1 - b0;
        u32 const nR = op.cnt - nL;
        op.base = b0;
        op.n_left = nL;
        bool const demote = !oblivious && ((nL == 0) != (nR == 0));
        op.alive = demote ? 0u : 1u;
        ops[j] = op;
        NodeInfo const par = front[op.parent];
        if (demote) {
            write_leaf(tree, par.id, node_leaf_value(par, cfg, scales));
            tree[par.id].gain = 0.0F;
        } else if (!oblivious) {
            tree[op.left_id].cover = static_cast<float>(nL);
            tree[op.right_id].cover = static_cast<float>(nR);
        }
    }
    __syncthreads();
    // 3) per unit: make the prefix op-relative.
    for (u32 c = threadIdx.x; c < n_ch; c += blockDim.x) pch[c].left_before -= ops[pch[c].op].base;
    // 4) next frontier, finish ops, histogram units of the smaller children.
    for (u32 t0 = 0; t0 < n_ops; t0 += blockDim.x) {
        u32 const j = t0 + threadIdx.x;
        bool const in = j < n_ops;
        PartOp op{};
        if (in) op = ops[j];
        u32 const alive = in ? op.alive : 0u;
        u32 atot;
        u32 const rank = block_exscan(alive, scratch, &atot) + s_alive;
        u32 const nL = op.n_left, nR = op.cnt - op.n_left;
        bool const left_small = nL <= nR;
        u32 const small_cnt = left_small ? nL : nR;
        u32 n_units = 0, multi = 0;
        if (in && alive && !is_last) {
            n_units = (small_cnt + kHistChunk - 1) / kHistChunk;
            multi = n_units > 1 ? n_units : 0;
        }
        u32 htot, ptot;
        u32 const hfirst = block_exscan(n_units, scratch, &htot) + s_hch;
        u32 const pfirst = block_exscan(multi, scratch, &ptot) + s_parts;
        if (in && alive) {
            NodeInfo const par = front[op.parent];
            SplitRec const s = splits[op.parent];
            NodeInfo L{}, R{};
            L.id = op.left_id;
            R.id = op.right_id;
            L.off = op.off;
            L.cnt = nL;
            R.off = op.off + nL;
            R.cnt = nR;
            L.src = R.src = op.dst;
            L.slot = 2 * rank;
            R.slot = 2 * rank + 1;
            L.depth = R.depth = par.depth + 1;
            if (is_last) {
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
            next[2 * rank] = L;
            next[2 * rank + 1] = R;
            if (!is_last) {
                FinishOp fo{};
                fo.small_slot = left_small ? L.slot : R.slot;
                fo.large_slot = left_small ? R.slot : L.slot;
                fo.parent_slot = par.slot;
                fo.part_first = multi ? pfirst : 0;
                fo.part_count = multi;
                fo.empty = small_cnt == 0 ? 1u : 0u;
                fo.unit_first = hfirst;
                fo.unit_count = n_units;
                fo.small_off = left_small ? L.off : R.off;
                fo.small_cnt = small_cnt;
                fo.src = op.dst;
                fops[rank] = fo;
            }
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            s_alive += atot;
            s_hch += htot;
            s_parts += ptot;
        }
        __syncthreads();
    }
    u32 const n_alive = s_alive, n_units_total = s_hch;
    if (!is_last) {
        for (u32 c = threadIdx.x; c < n_units_total; c += blockDim.x) {
            u32 const r = find_owner(fops, n_alive, c, [](FinishOp const &o) { return o.unit_first; });
            FinishOp const fo = fops[r];
            u32 const k = c - fo.unit_first;
            HistChunk hc{};
            hc.slot = fo.small_slot;
            hc.src = fo.src;
            hc.begin = fo.small_off + k * kHistChunk;
            hc.end = fo.small_off + min(fo.small_cnt, (k + 1) * kHistChunk);
            hc.part = fo.part_count ? fo.part_first + k : kNone;
            hch[c] = hc;
        }
    }
    if (threadIdx.x == 0) {
        ctl->next_front = 2 * n_alive;
        ctl->n_finish = is_last ? 0 : n_alive;
        ctl->n_hist_chunks = is_last ? 0 : n_units_total;
        ctl->n_parts = s_parts;
    }
}

// Stable scatter (H10 scatter_kernel, kernels.cuh:418-476; host_partition,
// src/step/primitives.hpp:262-374): left rows keep their order at off + ..., right rows
// follow at off + n_left + ....  Rows are contiguous per thread (16 each) so the
// block-wide exclusive scan of per-thread left counts gives every row its stable rank;
// the row bytes and the gradient pair are then copied with coalesced dword moves.
// Left rows of an op's chunks first..c-1, summed by the whole block from their counts
// (uses wsum[0 .. waves)).  Every thread receives the total.
__device__ u32 chunk_prefix(u32 const *__restrict__ chunk_left, u32 first, u32 c, u32 *wsum)
{
    u32 part = 0;
    for (u32 j = first + threadIdx.x; j < c; j += blockDim.x) part += chunk_left[j];
    for (int o = 32; o > 0; o >>= 1) part += __shfl_xor(part, o);
    if ((threadIdx.x & 63u) == 0) wsum[threadIdx.x >> 6] = part;
    __syncthreads();
    u32 total = 0;
    for (u32 k = 0; k < blockDim.x / 64u; ++k) total += wsum[k];
    __syncthreads();
    return total;
}

// A chunk's left and right runs both inside its op's segment.
__device__ __forceinline__ bool chunk_fits(PartOp const &op, PartChunk const &ch, u32 left_before, u32 n_left_chunk)
{
    u32 const rows = ch.end - ch.begin;
    return left_before <= ch.begin && left_before + n_left_chunk <= op.n_left && op.n_left <= op.cnt &&
           (ch.begin - left_before) + (rows - n_left_chunk) <= op.cnt - op.n_left;
}

constexpr u32 kCopyRows = 8;

__global__ __launch_bounds__(256) void k_part_scatter(
    PartChunk const *__restrict__ pch, PartOp const *__restrict__ ops, TreeCtl const *__restrict__ ctl,
    u8 const *__restrict__ pl0, u8 const *__restrict__ pl1, u8 const *__restrict__ plp,
    void const *__restrict__ gw0, void const *__restrict__ gw1, void const *__restrict__ gwp,
    u8 *__restrict__ out0, u8 *__restrict__ out1, void *__restrict__ go0, void *__restrict__ go1, u32 Fs,
    u32 gh_bytes, u32 tpr_log2, u32 const *__restrict__ chunk_left, TreeCtl *ctl_err)
{
    __shared__ u32 dest[kPartChunk];
    __shared__ u32 scratch[256];
    constexpr u32 RPT = kPartChunk / 256; // rows per thread
    u32 const n = ctl->n_part_chunks;
    u32 const D = Fs / 4u;
    u32 const tpr = 1u << tpr_log2;
    for (u32 c = blockIdx.x; c < n; c += gridDim.x) {
        PartChunk const ch = pch[c];
        PartOp const op = ops[ch.op];
        if (!op.alive) continue; // uniform across the block
        u8 const *plane = op.src == 0 ? pl0 : (op.src == 1 ? pl1 : plp);
        u8 const *gin = reinterpret_cast<u8 const *>(op.src == 0 ? gw0 : (op.src == 1 ? gw1 : gwp));
        u8 *oplane = op.dst == 0 ? out0 : out1;
        u8 *gout = reinterpret_cast<u8 *>(op.dst == 0 ? go0 : go1);
        u32 const rows = ch.end - ch.begin;
        u32 const r0 = threadIdx.x * RPT;
        u32 bits = 0;
        for (u32 k = 0; k < RPT; ++k) {
            u32 const r = r0 + k;
            if (r < rows) {
                u32 const b = plane[static_cast<size_t>(op.off + ch.begin + r) * Fs + op.feature];
                if (routes_left(b, op.last_bin, op.bin, op.default_left)) bits |= 1u << k;
            }
        }
        u32 tot;
        u32 const pre = block_exscan(__popc(bits), scratch, &tot);
        u32 left_before = ch.left_before;
        if (chunk_left) {
            left_before = chunk_prefix(chunk_left, op.chunk_first, c, scratch + 64);
            if (!chunk_fits(op, ch, left_before, tot)) {
                if (threadIdx.x == 0) atomicExch(&ctl_err->part_error, 1u);
                continue; // uniform across the block
            }
        }
        u32 lpos = op.off + left_before + pre;
        u32 rpos = op.off + op.n_left + (ch.begin + r0 - left_before - pre);
        for (u32 k = 0; k < RPT; ++k) {
            u32 const r = r0 + k;
            if (r < rows) dest[r] = ((bits >> k) & 1u) ? lpos++ : rpos++;
        }
        __syncthreads();
        // Row bytes: tpr threads per row, dword granularity, kCopyRows rows' loads in
        // flight per thread (a wide row is a long latency chain otherwise).
        u32 const rows_per_pass = 256u >> tpr_log2;
        u32 const sub = threadIdx.x & (tpr - 1u), rsub = threadIdx.x >> tpr_log2;
        size_t const sbase = static_cast<size_t>(op.off) + ch.begin;
        for (u32 r0 = 0; r0 < rows; r0 += rows_per_pass * kCopyRows) {
            for (u32 dw = sub; dw < D; dw += tpr) {
                u32 v[kCopyRows];
#pragma unroll
                for (u32 u = 0; u < kCopyRows; ++u) {
                    u32 const r = r0 + rsub + u * rows_per_pass;
                    v[u] = r < rows ? reinterpret_cast<u32 const *>(plane + (sbase + r) * Fs)[dw] : 0u;
                }
#pragma unroll
                for (u32 u = 0; u < kCopyRows; ++u) {
                    u32 const r = r0 + rsub + u * rows_per_pass;
                    if (r < rows) reinterpret_cast<u32 *>(oplane + static_cast<size_t>(dest[r]) * Fs)[dw] = v[u];
                }
            }
        }
        // Gradient pairs.
        for (u32 r = threadIdx.x; r < rows; r += 256u) {
            size_t const s = static_cast<size_t>(op.off + ch.begin + r);
            if (gh_bytes == 8)
                reinterpret_cast<uint2 *>(gout)[dest[r]] = reinterpret_cast<uint2 const *>(gin)[s];
            else
                reinterpret_cast<u32 *>(gout)[dest[r]] = reinterpret_cast<u32 const *>(gin)[s];
        }
        __syncthreads();
    }
}

// Narrow rows (Fs <= 64 bytes), LDS-staged with 16-byte accesses.  A tile of the unit's
// rows is one contiguous byte range: it is read with aligned 16-byte loads into LDS
// (keeping its sub-16-byte offset), ranked with wave ballots, and each child's run is a
// contiguous destination range written with aligned 16-byte stores (dword head and tail
// at the run's edges).  Planes and gradient buffers carry 64 bytes of tail padding, so
// the aligned over-read past a tile's last row stays inside the allocation.

// k / D as (k * ceil(2^32 / D)) >> 32 in 64-bit arithmetic: exact for k < 2^16, D <= 16.
__device__ __forceinline__ u32 div_magic(u32 k, u64 magic)
{
    return static_cast<u32>((static_cast<u64>(k) * magic) >> 32);
}

// NR runs in one pass: run m writes n[m] dwords to dst[first[m] ...], dword k reading
// src[base + idx_m[k / D] * D + k % D] with idx_m = idx + m * idx_stride.  The aligned
// 16-byte stores of all runs share one index space dealt round-robin over the threads
// (short runs leave no threads idle); the unaligned head and tail dwords of every run
// are one more step.  Run indices stay wave-uniform: no indexed register arrays.
template <int NR> struct Runs {
    size_t first[NR]; // first destination dword of each run
    u32 n[NR];        // dwords of each run
};

constexpr u32 kScatterThreads = 256; // workgroup size of the narrow scatters
template <int NR>
__device__ __forceinline__ void store_runs(u32 *__restrict__ dst, Runs<NR> const rs, u32 D, u64 magic,
                                           u32 const *src, u32 base, u16 const *idx, u32 idx_stride)
{
    constexpr u32 T = kScatterThreads;
    u32 const t = threadIdx.x;
    u32 start = 0; // first vector of run m in the shared index space
#pragma unroll
    for (int m = 0; m < NR; ++m) {
        u32 const n = rs.n[m];
        size_t const first = rs.first[m];
        u32 const head = min(n, static_cast<u32>((4u - (first & 3u)) & 3u));
        u32 const nvec = (n - head) >> 2;
        u16 const *im = idx + m * idx_stride;
        if ((t >> 3) == static_cast<u32>(m)) {
            // e < 4: head dword e; e >= 4: tail dword e - 4
            u32 const e = t & 7u;
            u32 const k = e < 4u ? e : head + 4u * nvec + (e - 4u);
            if (e < 4u ? e < head : k < n) {
                u32 const q = div_magic(k, magic);
                dst[first + k] = src[base + static_cast<u32>(im[q]) * D + (k - q * D)];
            }
        }
        u32 const lag = (t + T - start % T) % T; // first vector of this thread in the run
        if (D >= 4u) {
            // four dwords span at most two rows: both row indices read once
            for (u32 v = lag; v < nvec; v += T) {
                u32 const k0 = head + 4u * v;
                u32 const q = div_magic(k0, magic);
                u32 const rem = k0 - q * D;
                u32 const a0 = base + static_cast<u32>(im[q]) * D;
                u32 const a1 = rem + 3u >= D ? base + static_cast<u32>(im[q + 1u]) * D - D : a0;
                u32 w[4];
#pragma unroll
                for (u32 j = 0; j < 4; ++j) {
                    u32 const c = rem + j;
                    w[j] = src[(c < D ? a0 : a1) + c];
                }
                *reinterpret_cast<uint4 *>(dst + first + k0) = make_uint4(w[0], w[1], w[2], w[3]);
            }
        } else if (D == 1u) {
            // one dword per row (the gradient words)
            for (u32 v = lag; v < nvec; v += T) {
                u32 const k0 = head + 4u * v;
                u32 w[4];
#pragma unroll
                for (u32 j = 0; j < 4; ++j) w[j] = src[base + static_cast<u32>(im[k0 + j])];
                *reinterpret_cast<uint4 *>(dst + first + k0) = make_uint4(w[0], w[1], w[2], w[3]);
            }
        } else {
            for (u32 v = lag; v < nvec; v += T) {
                u32 const k0 = head + 4u * v;
                u32 q = div_magic(k0, magic);
                u32 rem = k0 - q * D;
                u32 w[4];
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    w[j] = src[base + static_cast<u32>(im[q]) * D + rem];
                    if (++rem == D) {
                        rem = 0;
                        ++q;
                    }
                }
                *reinterpret_cast<uint4 *>(dst + first + k0) = make_uint4(w[0], w[1], w[2], w[3]);
            }
        }
        start += nvec;
    }
}

// Register prefetch words per thread (256 threads) of the narrow scatters.
constexpr u32 kPfRows = 8, kPfGh = 2;

__global__ __launch_bounds__(256) void k_part_scatter_v(
    PartChunk const *__restrict__ pch, PartOp const *__restrict__ ops, TreeCtl const *__restrict__ ctl,
    u8 const *__restrict__ pl0, u8 const *__restrict__ pl1, u8 const *__restrict__ plp,
    void const *__restrict__ gw0, void const *__restrict__ gw1, void const *__restrict__ gwp,
    u8 *__restrict__ out0, u8 *__restrict__ out1, void *__restrict__ go0, void *__restrict__ go1, u32 Fs,
    u32 gh_bytes, u32 tile, u64 magic_d, u64 magic_g, u32 const *__restrict__ chunk_left, TreeCtl *ctl_err)
{
    extern __shared__ __align__(16) unsigned char smem[];
    uint4 *s_rows = reinterpret_cast<uint4 *>(smem);                        // tile * Fs + 16 bytes
    uint4 *s_gh = reinterpret_cast<uint4 *>(smem + tile * Fs + 16u);        // tile * gh_bytes + 16
    u16 *s_lidx = reinterpret_cast<u16 *>(smem + tile * (Fs + gh_bytes) + 32u);
    u16 *s_ridx = s_lidx + tile;
    __shared__ u32 s_wsum[4];
    u32 const n = ctl->n_part_chunks;
    u32 const D = Fs / 4u, G = gh_bytes / 4u;
    u32 const lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    for (u32 c = blockIdx.x; c < n; c += gridDim.x) {
        PartChunk const ch = pch[c];
        PartOp const op = ops[ch.op];
        if (!op.alive) continue; // uniform across the block
        u8 const *plane = op.src == 0 ? pl0 : (op.src == 1 ? pl1 : plp);
        u8 const *gin = static_cast<u8 const *>(op.src == 0 ? gw0 : (op.src == 1 ? gw1 : gwp));
        u32 *oplane = reinterpret_cast<u32 *>(op.dst == 0 ? out0 : out1);
        u32 *gout = static_cast<u32 *>(op.dst == 0 ? go0 : go1);
        u32 left_run = ch.left_before; // left rows of the op before the current tile
        if (chunk_left) {
            // Prefix over the op's earlier chunks straight from their counts; every run
            // of this chunk must land inside the op's segment.
            left_run = chunk_prefix(chunk_left, op.chunk_first, c, s_wsum);
            if (!chunk_fits(op, ch, left_run, chunk_left[c])) {
                if (threadIdx.x == 0) atomicExch(&ctl_err->part_error, 1u);
                continue; // uniform across the block
            }
        }
        // Register prefetch (the launcher sizes tiles to fit it): the next tile's loads
        // fly while this one is ranked and stored.
        uint4 pr[kPfRows], pg[kPfGh];
        u32 pr_n, pg_n, pr_sh, pg_sh;
        {
            size_t const srow0 = static_cast<size_t>(op.off) + ch.begin;
            u32 const rows0 = min(tile, ch.end - ch.begin);
            pr_sh = pf_load(pr, pr_n, plane, srow0 * Fs, rows0 * Fs);
            pg_sh = pf_load(pg, pg_n, gin, srow0 * gh_bytes, rows0 * gh_bytes);
        }
        for (u32 t0 = ch.begin; t0 < ch.end; t0 += tile) {
            u32 const rows = min(tile, ch.end - t0);
            pf_store(pr, pr_n, s_rows);
            pf_store(pg, pg_n, s_gh);
            u32 const sh = pr_sh, shg = pg_sh;
            __syncthreads();
            if (t0 + tile < ch.end) {
                size_t const srow1 = static_cast<size_t>(op.off) + t0 + tile;
                u32 const rows1 = min(tile, ch.end - t0 - tile);
                pr_sh = pf_load(pr, pr_n, plane, srow1 * Fs, rows1 * Fs);
                pg_sh = pf_load(pg, pg_n, gin, srow1 * gh_bytes, rows1 * gh_bytes);
            }
            u8 const *srb = reinterpret_cast<u8 const *>(s_rows) + sh;
            u32 run = 0;
            for (u32 p0 = 0; p0 < rows; p0 += blockDim.x) {
                u32 const r = p0 + threadIdx.x;
                bool const ok = r < rows;
                bool const left = ok && routes_left(srb[r * Fs + op.feature], op.last_bin, op.bin, op.default_left);
                u64 const bal = __ballot(left);
                u32 const lrank = static_cast<u32>(__popcll(bal & ((1ull << lane) - 1ull)));
                if (lane == 0) s_wsum[wave] = static_cast<u32>(__popcll(bal));
                __syncthreads();
                u32 pre = run;
                for (u32 k = 0; k < wave; ++k) pre += s_wsum[k];
                u32 tot = 0;
                for (u32 k = 0; k < blockDim.x / 64u; ++k) tot += s_wsum[k];
                __syncthreads();
                if (ok) {
                    u32 const lr = pre + lrank;
                    if (left) s_lidx[lr] = static_cast<u16>(r);
                    else s_ridx[r - lr] = static_cast<u16>(r);
                }
                run += tot;
            }
            __syncthreads();
            u32 const nl = run, nr = rows - run;
            // Left run: destination rows [off + left_run, + nl); right run: after n_left.
            size_t const l0 = static_cast<size_t>(op.off) + left_run;
            size_t const r0 = static_cast<size_t>(op.off) + op.n_left + (t0 - left_run);
            u32 const *sr = reinterpret_cast<u32 const *>(s_rows);
            u32 const *sg = reinterpret_cast<u32 const *>(s_gh);
            store_runs<2>(oplane, Runs<2>{{l0 * D, r0 * D}, {nl * D, nr * D}}, D, magic_d, sr, sh >> 2, s_lidx, tile);
            store_runs<2>(gout, Runs<2>{{l0 * G, r0 * G}, {nl * G, nr * G}}, G, magic_g, sg, shg >> 2, s_lidx, tile);
            left_run += nl;
            __syncthreads();
        }
    }
}

// ---------------------------------------------------------------------------------
// Four-way partition (two levels of splits in one pass, PartOp4).  Row classes: 0 AL,
// 1 AR, 2 BL, 3 BR, 4 = the row's child did not split (the row is not moved).
// ---------------------------------------------------------------------------------
__device__ __forceinline__ u32 route4_bins(u32 b0, u32 ba, u32 bb, PartOp4 const &q)
{
    if (routes_left(b0, q.last_bin, q.bin, q.default_left)) {
        if (!(q.alive & 1u)) return 4u;
        return routes_left(ba, q.la, q.ba, q.da) ? 0u : 1u;
    }
    if (!(q.alive & 2u)) return 4u;
    return routes_left(bb, q.lb, q.bb, q.db) ? 2u : 3u;
}
__device__ __forceinline__ u32 route4(u8 const *row, PartOp4 const &q)
{
    return route4_bins(row[q.feature], row[q.fa], row[q.fb], q);
}

// Rows per class of every partition chunk.  Only three bytes of a row route it, but
// its cache lines are read whole either way, so the chunk is streamed like the scatter
// does it: tiles of whole rows read with aligned 16-byte loads into registers (the next
// tile's loads in flight while this one is counted), staged in LDS, and every lane
// routes its rows from there.  tile * Fs + 16 <= kPfRows * 16 * 256 bytes (launcher).
__global__ __launch_bounds__(256) void k_part_count4(PartChunk const *__restrict__ pch,
                                                     PartOp4 const *__restrict__ ops4,
                                                     TreeCtl const *__restrict__ ctl, u8 const *__restrict__ pl0,
                                                     u8 const *__restrict__ pl1, u8 const *__restrict__ plp, u32 Fs,
                                                     u32 tile, uint4 *__restrict__ chunk_cls)
{
    extern __shared__ __align__(16) unsigned char smem[];
    uint4 *s_rows = reinterpret_cast<uint4 *>(smem); // tile * Fs + 16 bytes
    __shared__ u32 red[4][4];
    u32 const n = ctl->n_part_chunks;
    u32 const lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    for (u32 c = blockIdx.x; c < n; c += gridDim.x) {
        PartChunk const ch = pch[c];
        PartOp4 const q = ops4[ch.op];
        u8 const *plane = q.src == 0 ? pl0 : (q.src == 1 ? pl1 : plp);
        u32 k0 = 0, k1 = 0, k2 = 0, k3 = 0;
        uint4 pr[kPfRows];
        u32 pr_n, pr_sh;
        pr_sh = pf_load(pr, pr_n, plane, (static_cast<size_t>(q.off) + ch.begin) * Fs,
                        min(tile, ch.end - ch.begin) * Fs);
        for (u32 t0 = ch.begin; t0 < ch.end; t0 += tile) {
            u32 const rows = min(tile, ch.end - t0);
            __syncthreads(); // the previous tile is counted
            pf_store(pr, pr_n, s_rows);
            u32 const sh = pr_sh;
            __syncthreads();
            if (t0 + tile < ch.end)
                pr_sh = pf_load(pr, pr_n, plane, (static_cast<size_t>(q.off) + t0 + tile) * Fs,
                                min(tile, ch.end - t0 - tile) * Fs);
            u8 const *srb = reinterpret_cast<u8 const *>(s_rows) + sh;
            for (u32 r = threadIdx.x; r < rows; r += blockDim.x) {
                u32 const cls = route4(srb + r * Fs, q);
                k0 += cls == 0u;
                k1 += cls == 1u;
                k2 += cls == 2u;
                k3 += cls == 3u;
            }
        }
        for (int o = 32; o > 0; o >>= 1) {
            k0 += __shfl_xor(k0, o);
            k1 += __shfl_xor(k1, o);
            k2 += __shfl_xor(k2, o);
            k3 += __shfl_xor(k3, o);
        }
        if (lane == 0) {
            red[0][wave] = k0;
            red[1][wave] = k1;
            red[2][wave] = k2;
            red[3][wave] = k3;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            uint4 t;
            t.x = red[0][0] + red[0][1] + red[0][2] + red[0][3];
            t.y = red[1][0] + red[1][1] + red[1][2] + red[1][3];
            t.z = red[2][0] + red[2][1] + red[2][2] + red[2][3];
            t.w = red[3][0] + red[3][1] + red[3][2] + red[3][3];
            chunk_cls[c] = t;
        }
        __syncthreads();
    }
}

// Per-class rows of an op's chunks first..c-1, summed by the whole block (uses
// wsum[0 .. 4 * waves)).  Every thread receives the totals.
__device__ uint4 chunk_prefix4(uint4 const *__restrict__ chunk_cls, u32 first, u32 c, u32 *wsum)
{
    uint4 p = make_uint4(0, 0, 0, 0);
    for (u32 j = first + threadIdx.x; j < c; j += blockDim.x) {
        uint4 const v = chunk_cls[j];
        p.x += v.x;
        p.y += v.y;
        p.z += v.z;
        p.w += v.w;
    }
    for (int o = 32; o > 0; o >>= 1) {
        p.x += __shfl_xor(p.x, o);
        p.y += __shfl_xor(p.y, o);
        p.z += __shfl_xor(p.z, o);
        p.w += __shfl_xor(p.w, o);
    }
    u32 const waves = blockDim.x / 64u, wave = threadIdx.x >> 6;
    if ((threadIdx.x & 63u) == 0) {
        wsum[0 * waves + wave] = p.x;
        wsum[1 * waves + wave] = p.y;
        wsum[2 * waves + wave] = p.z;
        wsum[3 * waves + wave] = p.w;
    }
    __syncthreads();
    uint4 t = make_uint4(0, 0, 0, 0);
    for (u32 k = 0; k < waves; ++k) {
        t.x += wsum[0 * waves + k];
        t.y += wsum[1 * waves + k];
        t.z += wsum[2 * waves + k];
        t.w += wsum[3 * waves + k];
    }
    __syncthreads();
    return t;
}

// Narrow rows (Fs <= 64), LDS-staged like k_part_scatter_v: a tile is read with
// aligned 16-byte loads, its rows are ranked per class with wave ballots (stable within
// the class), and each class run is written as one contiguous destination range.
//
// CUR = false: a tile's runs start where k_part_count4's counts of the op's earlier
// chunks and the chunk's earlier tiles end, so the partition is stable.
// CUR = true (no counting pass; levels with many ops): the class sizes are the split
// finder's child counts (PartOp4), and each tile reserves its four runs with one atomic
// add per class on the op's cursors (kCursorStride u32 apart, zeroed by the planner).
// Rows keep their order inside a tile, but the tiles of a class follow the order of
// their reservations, which can change from run to run.  Only the quantized depthwise
// grower moves rows this way, and nothing downstream observes the order inside a node's
// segment: its histogram cells are exact integer sums, and scores are routed from the
// pristine plane.  With one op every tile of the level meets on the same four cursors,
// so the planner keeps CUR for levels whose ops spread the reservations.
template <bool CUR>
__global__ __launch_bounds__(256, 4) void k_part_scatter4(
    PartChunk const *__restrict__ pch, PartOp4 const *__restrict__ ops4, TreeCtl *__restrict__ ctl,
    u8 const *__restrict__ pl0, u8 const *__restrict__ pl1, u8 const *__restrict__ plp,
    u32 const *__restrict__ gw0, u32 const *__restrict__ gw1, u32 const *__restrict__ gwp,
    u8 *__restrict__ out0, u8 *__restrict__ out1, u32 *__restrict__ go0, u32 *__restrict__ go1, u32 Fs,
    u32 tile, u64 magic_d, uint4 const *__restrict__ chunk_cls, u32 *__restrict__ cursors)
{
    // Tiles fit the register prefetch: tile * Fs + 16 <= kPfRows * 16 * 256 bytes and
    // tile * 4 + 16 <= kPfGh * 16 * 256 (checked by the launcher).
    extern __shared__ __align__(16) unsigned char smem[];
    uint4 *s_rows = reinterpret_cast<uint4 *>(smem);                  // tile * Fs + 16 bytes
    uint4 *s_gh = reinterpret_cast<uint4 *>(smem + tile * Fs + 16u);  // tile * 4 + 16
    u16 *s_idx = reinterpret_cast<u16 *>(smem + tile * (Fs + 4u) + 32u); // 4 x tile
    __shared__ __align__(16) u32 s_wsum[16]; // [class][wave] rows of one 256-row step
    __shared__ u32 s_res[4];
    u32 const n = ctl->n_part_chunks;
    u32 const D = Fs / 4u;
    u32 const lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    for (u32 c = blockIdx.x; c < n; c += gridDim.x) {
        PartChunk const ch = pch[c];
        PartOp4 const q = ops4[ch.op];
        u8 const *plane = q.src == 0 ? pl0 : (q.src == 1 ? pl1 : plp);
        u32 const *gin = q.src == 0 ? gw0 : (q.src == 1 ? gw1 : gwp);
        u32 *oplane = reinterpret_cast<u32 *>(q.dst == 0 ? out0 : out1);
        u32 *gout = q.dst == 0 ? go0 : go1;
        // Class run starts inside the segment and the room each class has.
        u32 const n_ar = q.n_a - q.n_al, n_b = q.cnt - q.n_a, n_br = n_b - q.n_bl;
        uint4 run = make_uint4(0, 0, 0, 0), mine = make_uint4(0, 0, 0, 0);
        if (!CUR) {
            run = chunk_prefix4(chunk_cls, q.chunk_first, c, s_wsum);
            mine = chunk_cls[c];
        }
        u32 *cur = cursors + static_cast<size_t>(ch.op) * 4u * kCursorStride;
        bool const fits = q.n_a <= q.cnt && q.n_al <= q.n_a && q.n_bl <= n_b && run.x + mine.x <= q.n_al &&
                          run.y + mine.y <= n_ar && run.z + mine.z <= q.n_bl && run.w + mine.w <= n_br;
        if (!fits) {
            if (threadIdx.x == 0) atomicExch(&ctl->part_error, 1u);
            continue; // uniform across the block
        }
        size_t const base[4] = {static_cast<size_t>(q.off), static_cast<size_t>(q.off) + q.n_al,
                                static_cast<size_t>(q.off) + q.n_a, static_cast<size_t>(q.off) + q.n_a + q.n_bl};
        uint4 pr[kPfRows], pg[kPfGh];
        u32 pr_n, pg_n, pr_sh, pg_sh;
        {
            size_t const srow0 = static_cast<size_t>(q.off) + ch.begin;
            u32 const rows0 = min(tile, ch.end - ch.begin);
            pr_sh = pf_load(pr, pr_n, plane, srow0 * Fs, rows0 * Fs);
            pg_sh = pf_load(pg, pg_n, reinterpret_cast<u8 const *>(gin), srow0 * 4u, rows0 * 4u);
        }
        for (u32 t0 = ch.begin; t0 < ch.end; t0 += tile) {
            u32 const rows = min(tile, ch.end - t0);
            pf_store(pr, pr_n, s_rows);
            pf_store(pg, pg_n, s_gh);
            u32 const sh = pr_sh, shg = pg_sh;
            __syncthreads();
            // Next tile's loads fly while this one is ranked and stored.
            if (t0 + tile < ch.end) {
                size_t const srow1 = static_cast<size_t>(q.off) + t0 + tile;
                u32 const rows1 = min(tile, ch.end - t0 - tile);
                pr_sh = pf_load(pr, pr_n, plane, srow1 * Fs, rows1 * Fs);
                pg_sh = pf_load(pg, pg_n, reinterpret_cast<u8 const *>(gin), srow1 * 4u, rows1 * 4u);
            }
            u8 const *srb = reinterpret_cast<u8 const *>(s_rows) + sh;
            u32 tr[4] = {0, 0, 0, 0}; // rows of each class ranked so far in this tile
            for (u32 p0 = 0; p0 < rows; p0 += blockDim.x) {
                u32 const r = p0 + threadIdx.x;
                u32 const cls = r < rows ? route4(srb + r * Fs, q) : 4u;
                u32 rank = 0;
#pragma unroll
                for (u32 m = 0; m < 4; ++m) {
                    u64 const bal = __ballot(cls == m);
                    if (cls == m) rank = static_cast<u32>(__popcll(bal & ((1ull << lane) - 1ull)));
                    if (lane == 0) s_wsum[m * 4u + wave] = static_cast<u32>(__popcll(bal));
                }
                __syncthreads();
                // the four waves' counts of each class in one 16-byte read per class
                uint4 const c0 = reinterpret_cast<uint4 const *>(s_wsum)[0];
                uint4 const c1 = reinterpret_cast<uint4 const *>(s_wsum)[1];
                uint4 const c2 = reinterpret_cast<uint4 const *>(s_wsum)[2];
                uint4 const c3 = reinterpret_cast<uint4 const *>(s_wsum)[3];
                if (cls < 4u) {
                    uint4 const cw = cls == 0u ? c0 : (cls == 1u ? c1 : (cls == 2u ? c2 : c3));
                    u32 const before = (wave > 0u ? cw.x : 0u) + (wave > 1u ? cw.y : 0u) + (wave > 2u ? cw.z : 0u);
                    u32 const pre = (cls == 0u ? tr[0] : (cls == 1u ? tr[1] : (cls == 2u ? tr[2] : tr[3]))) + before;
                    s_idx[cls * tile + pre + rank] = static_cast<u16>(r);
                }
                tr[0] += c0.x + c0.y + c0.z + c0.w;
                tr[1] += c1.x + c1.y + c1.z + c1.w;
                tr[2] += c2.x + c2.y + c2.z + c2.w;
                tr[3] += c3.x + c3.y + c3.z + c3.w;
                __syncthreads();
            }
            u32 const *sr = reinterpret_cast<u32 const *>(s_rows);
            u32 const *sg = reinterpret_cast<u32 const *>(s_gh);
            if (CUR) {
                if (threadIdx.x < 4u) {
                    u32 const m = threadIdx.x;
                    u32 const tm = m == 0u ? tr[0] : (m == 1u ? tr[1] : (m == 2u ? tr[2] : tr[3]));
                    s_res[m] = tm ? atomicAdd(cur + m * kCursorStride, tm) : 0u;
                }
                __syncthreads();
                run = make_uint4(s_res[0], s_res[1], s_res[2], s_res[3]);
            }
            u32 const done[4] = {run.x, run.y, run.z, run.w};
            // The tile's runs stay inside their class ranges (they do whenever the
            // chunk's counts, or the split's child counts, matched its rows; checked
            // before any store).
            bool const tile_fits = static_cast<u64>(run.x) + tr[0] <= q.n_al && static_cast<u64>(run.y) + tr[1] <= n_ar &&
                                   static_cast<u64>(run.z) + tr[2] <= q.n_bl && static_cast<u64>(run.w) + tr[3] <= n_br;
            if (!tile_fits && threadIdx.x == 0) atomicExch(&ctl->part_error, 1u);
            if (tile_fits) { // uniform across the block
                size_t const a0 = base[0] + done[0], a1 = base[1] + done[1], a2 = base[2] + done[2],
                             a3 = base[3] + done[3];
                store_runs<4>(oplane, Runs<4>{{a0 * D, a1 * D, a2 * D, a3 * D}, {tr[0] * D, tr[1] * D, tr[2] * D, tr[3] * D}},
                              D, magic_d, sr, sh >> 2, s_idx, tile);
                store_runs<4>(gout, Runs<4>{{a0, a1, a2, a3}, {tr[0], tr[1], tr[2], tr[3]}}, 1u, 1ull << 32, sg, shg >> 2,
                              s_idx, tile);
            }
            run.x += tr[0];
            run.y += tr[1];
            run.z += tr[2];
            run.w += tr[3];
            __syncthreads();
        }
    }
}

// ---------------------------------------------------------------------------------
// Counted quantized level (depthwise): the split finder already derived each split's
// child row counts from the histogram count cells, so one workgroup plans the whole
// level -- tree nodes (commit_split_node, src/grower_impl.hpp:188-210), empty-split
// demotion (demote_empty_splits, :263-297: ids are committed first, as the reference
// does), covers, partition ops with their chunks, the next frontier, the finish ops and
// the histogram units of the smaller children.  Only the partition's per-chunk left
// counts still need a pass over the rows (none at the last level, where nothing moves).
// chunk_rows = rows per partition chunk.
//
// mode 0: this level's splits move the rows (two-way partition ops, ops[]).
// mode 1: rows stay put; the children's histogram units scan this level's segments and
//         keep the rows routed to the small child (ops[] holds the routing splits).
// mode 2: this level's nodes still share their parent's segment; one four-way
//         partition per parent applies the parent's split (ops[] from mode 1) and this
//         level's splits together (ops4[]).  Node offsets set in mode 1 are the ones
//         the nodes get from this partition, so the next frontier is the same as mode 0's.
// ---------------------------------------------------------------------------------

// A counted split that moves rows: valid and not demoted (both children non-empty, or
// both empty).
__device__ __forceinline__ bool counted_alive(SplitRec const &sp, NodeInfo const &nd)
{
    u32 const nL = sp.n_left, nR = nd.cnt - sp.n_left;
    return sp.valid && ((nL == 0) == (nR == 0));
}
// ---------------------------------------------------------------------------------
constexpr u32 kPlanLdsOps = 4096;

__global__ __launch_bounds__(1024) void k_plan_counted(
    NodeInfo const *__restrict__ front, SplitRec const *__restrict__ splits, TreeCtl *__restrict__ ctl,
    DevNode *__restrict__ tree, PartOp *__restrict__ ops, PartChunk *__restrict__ pch, NodeInfo *__restrict__ next,
    HistChunk *__restrict__ hch, FinishOp *__restrict__ fops, u16 const *__restrict__ n_bins, GrowCfg cfg,
    double const *__restrict__ scales, u32 dst, u32 is_last, u32 chunk_rows, u32 mode, PartOp4 *__restrict__ ops4,
    u32 unit_rows, u32 *__restrict__ cursors)
{
    __shared__ u32 scratch[1024];
    __shared__ u32 s_base_nodes, s_valid, s_alive, s_chunks, s_units, s_parts;
    // First chunk / unit of each op for the owner searches below (levels of at most
    // kPlanLdsOps ops; larger ones search the op arrays in global memory).
    __shared__ u32 s_cfirst[kPlanLdsOps], s_ufirst[kPlanLdsOps];
    u32 const P = ctl->n_front;
    if (threadIdx.x == 0) {
        s_base_nodes = ctl->n_nodes;
        s_valid = s_alive = s_chunks = s_units = s_parts = 0;
    }
    __syncthreads();
    for (u32 t0 = 0; t0 < P; t0 += blockDim.x) {
        u32 const i = t0 + threadIdx.x;
        bool const in = i < P;
        NodeInfo nd{};
        SplitRec sp{};
        if (in) {
            nd = front[i];
            sp = splits[i];
        }
        u32 const valid = in ? sp.valid : 0u;
        u32 const nL = sp.n_left, nR = nd.cnt - sp.n_left;
        bool const demote = valid && ((nL == 0) != (nR == 0));
        u32 const alive = valid && !demote ? 1u : 0u;
        u32 nch = 0, n_units = 0, multi = 0;
        bool const left_small = nL <= nR;
        u32 const small_cnt = left_small ? nL : nR;
        if (alive && !is_last) {
            nch = mode == 0 ? (nd.cnt + chunk_rows - 1) / chunk_rows : 0u;
            u32 const scan = mode == 1 ? (small_cnt ? nd.cnt : 0u) : small_cnt;
            n_units = (scan + unit_rows - 1) / unit_rows;
            multi = n_units > 1 ? n_units : 0;
        }
        u32 vtot, atot, ctot, htot, ptot;
        u32 const vrank = block_exscan(valid, scratch, &vtot) + s_valid;
        u32 const arank = block_exscan(alive, scratch, &atot) + s_alive;
        u32 const cfirst = block_exscan(nch, scratch, &ctot) + s_chunks;
        u32 const hfirst = block_exscan(n_units, scratch, &htot) + s_units;
        u32 const pfirst = block_exscan(multi, scratch, &ptot) + s_parts;
        if (in) {
            if (valid) {
                u32 const lid = s_base_nodes + 2 * vrank, rid = lid + 1;
                DevNode leaf{};
                leaf.feature = -1;
                tree[lid] = leaf;
                tree[rid] = leaf;
                if (demote) {
                    write_leaf(tree, nd.id, node_leaf_value(nd, cfg, scales));
                    tree[nd.id].gain = 0.0F;
                } else {
                    DevNode pn{}; // cover: the node's rows, as its parent set it
                    pn.cover = static_cast<float>(nd.cnt);
                    pn.feature = static_cast<int32_t>(sp.feature);
                    pn.bin = sp.bin;
                    pn.default_left = sp.default_left;
                    pn.left = lid;
                    pn.right = rid;
                    pn.value = 0.0F;
                    pn.gain = static_cast<float>(sp.gain);
                    tree[nd.id] = pn;
                    tree[lid].cover = static_cast<float>(nL);
                    tree[rid].cover = static_cast<float>(nR);
                    fd_propagate_bounds(cfg, scales, nd.id, sp, lid, rid);
                    fd_propagate_path(cfg, nd.id, sp.feature, lid, rid);
                    PartOp op{};
                    op.parent = i;
                    op.feature = sp.feature;
                    op.bin = sp.bin;
                    op.default_left = sp.default_left;
                    op.last_bin = static_cast<u32>(n_bins[sp.feature]) - 1u;
                    op.off = nd.off;
                    op.cnt = nd.cnt;
                    op.src = nd.src;
                    op.dst = dst;
                    op.chunk_first = cfirst;
                    op.chunk_count = nch;
                    op.n_left = nL;
                    op.left_id = lid;
                    op.right_id = rid;
                    op.alive = 1;
                    if (mode != 2) ops[arank] = op;
                    if (arank < kPlanLdsOps) {
                        if (mode == 0) s_cfirst[arank] = cfirst;
                        s_ufirst[arank] = hfirst;
                    }
                    NodeInfo L{}, R{};
                    L.id = lid;
                    R.id = rid;
                    L.off = nd.off;
                    L.cnt = nL;
                    R.off = nd.off + nL;
                    R.cnt = nR;
                    L.src = R.src = mode == 1 ? nd.src : dst;
                    L.slot = 2 * arank;
                    R.slot = 2 * arank + 1;
                    L.depth = R.depth = nd.depth + 1;
                    if (is_last) {
                        L.totals_ready = R.totals_ready = 1;
                        L.itg = sp.igL;
                        L.ith = sp.ihL;
                        R.itg = sp.igR;
                        R.ith = sp.ihR;
                    }
                    next[2 * arank] = L;
                    next[2 * arank + 1] = R;
                    if (!is_last) {
                        FinishOp fo{};
                        fo.small_slot = left_small ? L.slot : R.slot;
                        fo.large_slot = left_small ? R.slot : L.slot;
                        fo.parent_slot = nd.slot;
                        fo.part_first = multi ? pfirst : 0;
                        fo.part_count = multi;
                        fo.empty = small_cnt == 0 ? 1u : 0u;
                        fo.unit_first = hfirst;
                        fo.unit_count = n_units;
                        fo.small_off = mode == 1 ? nd.off : (left_small ? L.off : R.off);
                        fo.small_cnt = mode == 1 ? nd.cnt : small_cnt;
                        fo.src = L.src;
                        fo.route = mode == 1 ? arank : kNone;
                        fo.side = left_small ? 0u : 1u;
                        fo.wide = unit_rows > kHistChunk ? 1u : 0u;
                        fops[arank] = fo;
                    }
                }
            } else {
                write_leaf(tree, nd.id, node_leaf_value(nd, cfg, scales));
            }
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            s_valid += vtot;
            s_alive += atot;
            s_chunks += ctot;
            s_units += htot;
            s_parts += ptot;
        }
        __syncthreads();
    }
    u32 const pairs = mode == 2 && !is_last ? P / 2u : 0u;
    for (u32 t0 = 0; t0 < pairs; t0 += blockDim.x) {
        // Four-way ops, one per parent segment (parent j of frontier nodes 2j, 2j+1).
        u32 const j = t0 + threadIdx.x;
        bool const in = j < pairs;
        PartOp4 q{};
        u32 nch = 0;
        if (in) {
            PartOp const g = ops[j];
            NodeInfo const a = front[2 * j], b = front[2 * j + 1];
            SplitRec const sa = splits[2 * j], sb = splits[2 * j + 1];
            bool const aa = counted_alive(sa, a), ab = counted_alive(sb, b);
            q.off = g.off;
            q.cnt = g.cnt;
            q.src = g.src;
            q.dst = dst;
            q.feature = g.feature;
            q.bin = g.bin;
            q.default_left = g.default_left;
            q.last_bin = g.last_bin;
            if (aa) {
                q.fa = sa.feature;
                q.ba = sa.bin;
                q.da = sa.default_left;
                q.la = static_cast<u32>(n_bins[sa.feature]) - 1u;
            }
            if (ab) {
                q.fb = sb.feature;
                q.bb = sb.bin;
                q.db = sb.default_left;
                q.lb = static_cast<u32>(n_bins[sb.feature]) - 1u;
            }
            q.n_a = a.cnt;
            q.n_al = aa ? sa.n_left : 0u;
            q.n_bl = ab ? sb.n_left : 0u;
            q.alive = (aa ? 1u : 0u) | (ab ? 2u : 0u);
            nch = q.alive ? (g.cnt + chunk_rows - 1) / chunk_rows : 0u;
        }
        u32 ctot;
        u32 const cfirst = block_exscan(nch, scratch, &ctot) + s_chunks;
        if (in) {
            q.chunk_first = cfirst;
            q.chunk_count = nch;
            ops4[j] = q;
            // the four class cursors of k_part_scatter4<true>
#pragma unroll
            for (u32 m = 0; m < 4; ++m) cursors[(static_cast<size_t>(j) * 4u + m) * kCursorStride] = 0u;
            if (j < kPlanLdsOps) s_cfirst[j] = cfirst;
        }
        __syncthreads();
        if (threadIdx.x == 0) s_chunks += ctot;
        __syncthreads();
    }
    u32 const n_alive = s_alive, n_ch = s_chunks, n_units_total = s_units;
    for (u32 c = threadIdx.x; c < n_ch && pairs; c += blockDim.x) {
        u32 const j = pairs <= kPlanLdsOps
                          ? find_owner(s_cfirst, pairs, c, [](u32 v) { return v; })
                          : find_owner(ops4, pairs, c, [](PartOp4 const &o) { return o.chunk_first; });
        PartOp4 const q = ops4[j];
        u32 const k = c - q.chunk_first;
        PartChunk pc;
        pc.op = j;
        pc.begin = k * chunk_rows;
        pc.end = min(q.cnt, (k + 1) * chunk_rows);
        pc.left_before = 0;
        pch[c] = pc;
    }
    for (u32 c = threadIdx.x; c < n_ch && !pairs; c += blockDim.x) {
        u32 const j = n_alive <= kPlanLdsOps
                          ? find_owner(s_cfirst, n_alive, c, [](u32 v) { return v; })
                          : find_owner(ops, n_alive, c, [](PartOp const &o) { return o.chunk_first; });
        PartOp const op = ops[j];
        u32 const k = c - op.chunk_first;
        PartChunk pc;
        pc.op = j;
        pc.begin = k * chunk_rows;
        pc.end = min(op.cnt, (k + 1) * chunk_rows);
        pc.left_before = 0;
        pch[c] = pc;
    }
    for (u32 c = threadIdx.x; c < n_units_total; c += blockDim.x) {
        u32 const r = n_alive <= kPlanLdsOps
                          ? find_owner(s_ufirst, n_alive, c, [](u32 v) { return v; })
                          : find_owner(fops, n_alive, c, [](FinishOp const &o) { return o.unit_first; });
        FinishOp const fo = fops[r];
        u32 const k = c - fo.unit_first;
        HistChunk hc{};
        hc.slot = fo.small_slot;
        hc.src = fo.src;
        hc.begin = fo.small_off + k * unit_rows;
        hc.end = fo.small_off + min(fo.small_cnt, (k + 1) * unit_rows);
        hc.part = fo.part_count ? fo.part_first + k : kNone;
        hc.route = fo.route;
        hc.side = fo.side;
        hc.wide = fo.wide;
        hch[c] = hc;
    }
    if (threadIdx.x == 0) {
        ctl->n_nodes = s_base_nodes + 2 * s_valid;
        ctl->n_ops = n_alive;
        ctl->n_part_chunks = n_ch;
        ctl->next_front = 2 * n_alive;
        ctl->n_finish = is_last ? 0 : n_alive;
        ctl->n_hist_chunks = is_last ? 0 : n_units_total;
        ctl->n_parts = s_parts;
    }
}

// Frontier left at the end of the tree becomes leaves (LevelStep::end_tree,
// src/step/level.hpp:119-128): values from the totals the parent split carried.
__global__ void k_finalize(NodeInfo const *__restrict__ front, TreeCtl const *__restrict__ ctl,
                           DevNode *__restrict__ tree, GrowCfg cfg, double const *__restrict__ scales)
{
    u32 const P = ctl->n_front;
    for (u32 i = blockIdx.x * blockDim.x + threadIdx.x; i < P; i += gridDim.x * blockDim.x) {
        NodeInfo const nd = front[i];
        write_leaf(tree, nd.id, node_leaf_value(nd, cfg, scales));
    }
}

// Frontier swap between levels.
__global__ void k_next_level(TreeCtl *ctl)
{
    if (threadIdx.x == 0) {
        ctl->n_front = ctl->next_front;
        ctl->depth += 1;
    }
}

// Tree start: root frontier, root histogram units over the pristine plane.
__global__ __launch_bounds__(256) void k_tree_begin(TreeCtl *__restrict__ ctl, NodeInfo *__restrict__ front,
                                                    DevNode *__restrict__ tree, HistChunk *__restrict__ hch,
                                                    FinishOp *__restrict__ fops, u32 n_rows_host, u32 unit_rows,
                                                    u32 src, u32 n_from_ctl)
{
    u32 const n_rows = n_from_ctl ? ctl->root_n : n_rows_host;
    u32 const n_units = (n_rows + unit_rows - 1) / unit_rows;
    u32 const multi = n_units > 1;
    for (u32 k = threadIdx.x; k < n_units; k += blockDim.x) {
        HistChunk hc{};
        hc.slot = 0;
        hc.src = src;
        hc.begin = k * unit_rows;
        hc.end = min(n_rows, (k + 1) * unit_rows);
        hc.part = multi ? k : kNone;
        hc.wide = unit_rows > kHistChunk ? 1u : 0u;
        hch[k] = hc;
    }
    if (threadIdx.x == 0) {
        NodeInfo root{};
        root.id = 0;
        root.off = 0;
        root.cnt = n_rows;
        root.src = src;
        root.slot = 0;
        front[0] = root;
        DevNode r{};
        r.feature = -1;
        r.cover = static_cast<float>(n_rows);
        tree[0] = r;
        FinishOp fo{};
        fo.small_slot = 0;
        fo.large_slot = kNone;
        fo.part_first = 0;
        fo.part_count = multi ? n_units : 0;
        fo.empty = n_rows == 0;
        fo.wide = unit_rows > kHistChunk ? 1u : 0u;
        fops[0] = fo;
        ctl->n_front = 1;
        ctl->n_nodes = 1;
        ctl->n_ops = 0;
        ctl->n_part_chunks = 0;
        ctl->n_hist_chunks = n_units;
        ctl->n_finish = 1;
        ctl->n_parts = multi ? n_units : 0;
        ctl->next_front = 0;
        ctl->depth = 0;
        ctl->done = 0;
        ctl->oblivious_depth = 0;
    }
}

// ---------------------------------------------------------------------------------
// Resident score epilogue (H12): route_add_kernel (src/cuda/detail/kernels.cuh:1140-1163)
// and the booster's score loop (include/bonsai/booster.hpp:577-580):
// score += lr * leaf, multiply then add, each rounded.
// ---------------------------------------------------------------------------------
constexpr u32 kWalkLds = 2048;

__device__ __forceinline__ WalkNode to_walk(DevNode const &d)
{
    WalkNode w;
    w.feature = d.feature < 0 ? 0xFFFFu : static_cast<u16>(d.feature);
    w.bin = static_cast<u8>(d.bin);
    w.default_left = static_cast<u8>(d.default_left);
    w.left = static_cast<u16>(d.left);
    w.right = static_cast<u16>(d.right);
    w.value = d.value;
    return w;
}

// Score epilogue (H12, route_add_kernel, kernels.cuh:1140-1163): every row walks the
// tree from the pristine plane, score += lr * leaf (one rounding each), optionally the
// leaf id.  STAGED (rows of <= 64 bytes): a block's 256 rows are read as one contiguous
// byte range with 16-byte loads into LDS and each lane walks its row from there;
// otherwise (wide rows) the lane reads only the bins its path touches.
template <bool STAGED>
__global__ __launch_bounds__(256) void k_score_update(u8 const *__restrict__ plane, u32 Fs, u64 n,
                                                      DevNode const *__restrict__ tree,
                                                      TreeCtl const *__restrict__ ctl,
                                                      u16 const *__restrict__ n_bins, u32 F, float lr,
                                                      float *__restrict__ score, u32 *__restrict__ leaf_of_row,
                                                      u32 node_cap, u32 last_cap, float const *__restrict__ label,
                                                      float const *__restrict__ weight, int objective,
                                                      u32 *__restrict__ maxbits)
{
    // LDS: [node_cap WalkNode][last_cap u16][rows: 256 * Fs + 16, 16-byte aligned]
    extern __shared__ __align__(16) unsigned char smem[];
    WalkNode *nodes = reinterpret_cast<WalkNode *>(smem);
    u16 *last = reinterpret_cast<u16 *>(smem + node_cap * sizeof(WalkNode));
    uint4 *s_rows = reinterpret_cast<uint4 *>(
        smem + ((node_cap * sizeof(WalkNode) + last_cap * sizeof(u16) + 15u) & ~static_cast<size_t>(15)));
    u32 const nn = ctl->n_nodes;
    u32 const nl = min(nn, node_cap);
    float mg = 0.0F, mh = 0.0F;
    for (u32 i = threadIdx.x; i < nl; i += blockDim.x) nodes[i] = to_walk(tree[i]);
    for (u32 f = threadIdx.x; f < F && f < last_cap; f += blockDim.x) last[f] = static_cast<u16>(n_bins[f] - 1u);
    __syncthreads();
    for (u64 r0 = blockIdx.x * static_cast<u64>(blockDim.x); r0 < n; r0 += static_cast<u64>(gridDim.x) * blockDim.x) {
        u64 const r = r0 + threadIdx.x;
        u8 const *row;
        if (STAGED) {
            u32 const rows = static_cast<u32>(min<u64>(blockDim.x, n - r0));
            __syncthreads(); // previous tile walked
            u32 const sh = lds_load_range(plane, r0 * Fs, rows * Fs, s_rows);
            __syncthreads();
            row = reinterpret_cast<u8 const *>(s_rows) + sh + threadIdx.x * Fs;
        } else {
            row = plane + r * Fs;
        }
        if (r >= n) continue;
        // the row's score and label loads fly during the walk
        float const s0 = score ? score[r] : 0.0F;
        float const y = maxbits ? label[r] : 0.0F;
        u32 idx = 0;
        WalkNode w = nl ? nodes[0] : to_walk(tree[0]);
        while (w.feature != 0xFFFFu) {
            u32 const f = w.feature;
            u32 const b = row[f];
            u32 const lb = f < last_cap ? last[f] : static_cast<u32>(n_bins[f]) - 1u;
            idx = routes_left(b, lb, w.bin, w.default_left) ? w.left : w.right;
            w = idx < nl ? nodes[idx] : to_walk(tree[idx]);
        }
        if (score) {
            float const s = __fadd_rn(s0, __fmul_rn(lr, w.value));
            score[r] = s;
            if (maxbits) {
                float g, h;
                grad_of_row(objective, s, y, weight, r, g, h);
                mg = fmaxf(mg, fabsf(g));
                mh = fmaxf(mh, fabsf(h));
            }
        }
        if (leaf_of_row) leaf_of_row[r] = idx;
    }
    // the next iteration's gradient maxima (k_grad_max's pass, fused)
    if (maxbits) block_max2_atomic(mg, mh, maxbits + 2u * ((ctl->iter + 1u) & 1u));
}

// Validation loss pass (eval_loss_pass1_kernel, src/cuda/detail/kernels.cuh:1061-1103):
// per-row loss terms in float, summed in double by grid-stride loops, then a tree
// reduction per block; one partial per block.  Objective terms as the reference's
// objectives evaluate them (src/objective.cpp:60-120): squared error, and binary
// cross-entropy from the raw score, max(0, s) + log1p(exp(-|s|)) - y * s.
template <int OBJ>
__global__ __launch_bounds__(256) void k_eval_loss(float const *__restrict__ scores, float const *__restrict__ labels,
                                                   u32 n, double *__restrict__ partial)
{
    __shared__ double sl[256];
    double acc = 0.0;
    for (u32 r = blockIdx.x * blockDim.x + threadIdx.x; r < n; r += gridDim.x * blockDim.x) {
        float const s = scores[r], y = labels[r];
        if (OBJ == 0) {
            double const d = static_cast<double>(s) - static_cast<double>(y);
            acc += d * d;
        } else if (OBJ == 2) {
            float const f = fminf(fmaxf(s, -kPoissonMaxLog), kPoissonMaxLog);
            acc += static_cast<double>(expf(f)) - static_cast<double>(y) * static_cast<double>(f);
        } else {
            float const ax = fabsf(s);
            acc += fmaxf(0.0F, s) + log1pf(expf(-ax)) - y * s;
        }
    }
    sl[threadIdx.x] = acc;
    __syncthreads();
    for (u32 st = blockDim.x / 2; st > 0; st >>= 1) {
        if (threadIdx.x < st) sl[threadIdx.x] += sl[threadIdx.x + st];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = sl[0];
}

// Block partials summed in block order (fixed, so the loss is run-to-run stable), the
// mean loss recorded, and the early-stopping rule applied: a new best (or the first
// loss) resets the patience; once `patience` losses follow the best without beating it
// the fit stops and keeps the trees up to the best one.  After the stop nothing more is
// recorded.  patience == 0: record only.
__global__ void k_eval_finish(double const *__restrict__ partial, u32 blocks, u32 n, TreeCtl *__restrict__ ctl,
                              float *__restrict__ hist, u32 hist_cap, u32 patience)
{
    __shared__ double s_p[1024];
    for (u32 b = threadIdx.x; b < blocks; b += blockDim.x) s_p[b] = partial[b];
    __syncthreads();
    if (threadIdx.x != 0 || ctl->es_stopped) return;
    double total = 0.0;
    for (u32 b = 0; b < blocks; ++b) total += s_p[b];
    float const loss = static_cast<float>(total / static_cast<double>(n));
    u32 const i = ctl->es_n;
    if (i < hist_cap) hist[i] = loss;
    ctl->es_n = i + 1;
    if (i == 0 || loss < ctl->es_best) {
        ctl->es_best = loss;
        ctl->es_best_iter = i;
    } else if (patience && i - ctl->es_best_iter >= patience) {
        ctl->es_stopped = 1;
    }
}

__global__ void k_iter_advance(TreeCtl *ctl)
{
    if (threadIdx.x == 0) ctl->iter += 1;
}

__global__ void k_copy_tree(DevNode const *__restrict__ tree, TreeCtl const *__restrict__ ctl,
                            DevNode *__restrict__ out_trees, u32 *__restrict__ out_sizes, u32 max_nodes)
{
    u32 const it = ctl->iter;
    u32 const nn = min(ctl->n_nodes, max_nodes);
    for (u32 i = threadIdx.x; i < nn; i += blockDim.x)
        out_trees[static_cast<size_t>(it) * max_nodes + i] = tree[i];
    if (threadIdx.x == 0) out_sizes[it] = nn;
}

// Gathers listed rows of the pristine plane (and their gradient pairs, or the row id
// itself as the pair) into a contiguous segment: the entry to the kernel-level tests.
// ---------------------------------------------------------------------------------
// Device Bernoulli row sampler on the resident path (BernoulliSampler::sample,
// src/sampler.cpp:38-55; eligible there, include/bonsai/booster.hpp:632-635): row r of
// iteration t is kept when the top 32 bits of a counter-based draw (splitmix64 of the
// seed, t, r and channel 2; channels 0 and 1 are the gradient roundings) fall below
// thr = floor(p * 2^32).  The kept rows, ascending, are gathered into work plane 1 as
// the tree's root segment; the score update still walks every row.  (The reference
// draws with std::mt19937; its exact stream is reproduced by handing the indices over
// instead, bh_set_row_subset.)
// ---------------------------------------------------------------------------------
__device__ __forceinline__ bool bern_keep(u64 seed, u32 iter, u32 row, u32 thr)
{
    u64 const key = seed ^ (static_cast<u64>(iter) * 0xD1B54A32D192ED03ULL) ^
                    (static_cast<u64>(row) * 0x9E3779B97F4A7C15ULL) ^ (3ULL * 0xA0761D6478BD642FULL);
    return static_cast<u32>(splitmix64(key) >> 32) < thr;
}

// Kept rows per kPartChunk-row chunk.
__global__ __launch_bounds__(256) void k_bern_count(TreeCtl const *__restrict__ ctl, u32 n, u64 seed, u32 thr,
                                                    u32 *__restrict__ chunk_cnt)
{
    __shared__ u32 red[4];
    u32 const iter = ctl->iter;
    u32 const n_ch = (n + kPartChunk - 1) / kPartChunk;
    for (u32 c = blockIdx.x; c < n_ch; c += gridDim.x) {
        u32 const r0 = c * kPartChunk, r1 = min(n, r0 + kPartChunk);
        u32 k = 0;
        for (u32 r = r0 + threadIdx.x; r < r1; r += blockDim.x) k += bern_keep(seed, iter, r, thr) ? 1u : 0u;
        for (int o = 32; o > 0; o >>= 1) k += __shfl_xor(k, o);
        if ((threadIdx.x & 63u) == 0) red[threadIdx.x >> 6] = k;
        __syncthreads();
        if (threadIdx.x == 0) chunk_cnt[c] = red[0] + red[1] + red[2] + red[3];
        __syncthreads();
    }
}

// Exclusive prefix of the chunk counts in place; the total is the root's row count.
__global__ __launch_bounds__(1024) void k_bern_scan(u32 *__restrict__ chunk_cnt, u32 n_ch, TreeCtl *__restrict__ ctl)
{
    __shared__ u32 scratch[1024];
    __shared__ u32 s_run;
    if (threadIdx.x == 0) s_run = 0;
    __syncthreads();
    for (u32 t0 = 0; t0 < n_ch; t0 += blockDim.x) {
        u32 const c = t0 + threadIdx.x;
        u32 const v = c < n_ch ? chunk_cnt[c] : 0u;
        u32 tot;
        u32 const pre = block_exscan(v, scratch, &tot) + s_run;
        if (c < n_ch) chunk_cnt[c] = pre;
        __syncthreads();
        if (threadIdx.x == 0) s_run += tot;
        __syncthreads();
    }
    if (threadIdx.x == 0) ctl->root_n = s_run;
}

// Stable compaction of the kept rows (and their gradient words) into out_plane/out_gh.
__global__ __launch_bounds__(256) void k_bern_gather(TreeCtl const *__restrict__ ctl, u32 n, u64 seed, u32 thr,
                                                     u32 const *__restrict__ chunk_off, u8 const *__restrict__ plane,
                                                     u32 Fs, void const *__restrict__ gh, u32 gh_bytes,
                                                     u8 *__restrict__ out_plane, void *__restrict__ out_gh)
{
    __shared__ u32 s_w[4];
    u32 const iter = ctl->iter;
    u32 const lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    u32 const D = Fs / 4u;
    u32 const n_ch = (n + kPartChunk - 1) / kPartChunk;
    for (u32 c = blockIdx.x; c < n_ch; c += gridDim.x) {
        u32 const r0 = c * kPartChunk, r1 = min(n, r0 + kPartChunk);
        u32 run = chunk_off[c];
        for (u32 p0 = r0; p0 < r1; p0 += blockDim.x) {
            u32 const r = p0 + threadIdx.x;
            bool const keep = r < r1 && bern_keep(seed, iter, r, thr);
            u64 const bal = __ballot(keep);
            if (lane == 0) s_w[wave] = static_cast<u32>(__popcll(bal));
            __syncthreads();
            u32 pre = run;
            for (u32 k = 0; k < wave; ++k) pre += s_w[k];
            u32 const tot = s_w[0] + s_w[1] + s_w[2] + s_w[3];
            __syncthreads();
            if (keep) {
                size_t const dst = pre + static_cast<u32>(__popcll(bal & ((1ull << lane) - 1ull)));
                u32 const *src = reinterpret_cast<u32 const *>(plane + static_cast<size_t>(r) * Fs);
                u32 *o = reinterpret_cast<u32 *>(out_plane + dst * Fs);
                for (u32 d = 0; d < D; ++d) o[d] = src[d];
                if (gh_bytes == 8) reinterpret_cast<uint2 *>(out_gh)[dst] = reinterpret_cast<uint2 const *>(gh)[r];
                else reinterpret_cast<u32 *>(out_gh)[dst] = reinterpret_cast<u32 const *>(gh)[r];
            }
            run += tot;
        }
    }
}

void launch_bernoulli(hipStream_t s, TreeCtl *ctl, u32 n, u64 seed, u32 thr, u32 *chunk_cnt, u8 const *plane,
                      u32 Fs, void const *gh, u32 gh_bytes, u8 *out_plane, void *out_gh)
{
    u32 const n_ch = (n + kPartChunk - 1) / kPartChunk;
    u32 const grid = std::max(1u, std::min(n_ch, 2048u));
    hipLaunchKernelGGL(k_bern_count, dim3(grid), dim3(256), 0, s, ctl, n, seed, thr, chunk_cnt);
    hipLaunchKernelGGL(k_bern_scan, dim3(1), dim3(1024), 0, s, chunk_cnt, n_ch, ctl);
    hipLaunchKernelGGL(k_bern_gather, dim3(grid), dim3(256), 0, s, ctl, n, seed, thr, chunk_cnt, plane, Fs, gh,
                       gh_bytes, out_plane, out_gh);
}

// ---------------------------------------------------------------------------------
// Per-tree feature draw (sample_features, src/grower_impl.hpp:165-180): the
// k = max(1, ceil(fraction * F)) features with the smallest keys are selected, a key
// being the 64-bit splitmix64 draw of (seed, t, f, channel 5) with ties broken by the
// feature id.  Every thread ranks one feature against all F keys (F is at most a few
// thousand: one pass over them through LDS per workgroup).  (The reference draws with
// std::sample over std::mt19937; its exact selection is reproduced by handing the
// mask over instead, bh_set_feature_mask.)
// ---------------------------------------------------------------------------------
__device__ __forceinline__ u64 feature_key(u64 seed, u32 iter, u32 f)
{
    return splitmix64(seed ^ (static_cast<u64>(iter) * 0xD1B54A32D192ED03ULL) ^
                      (static_cast<u64>(f) * 0x9E3779B97F4A7C15ULL) ^ (5ULL * 0xA0761D6478BD642FULL));
}

__global__ __launch_bounds__(256) void k_feature_draw(TreeCtl const *__restrict__ ctl, u32 F, u32 k, u64 seed,
                                                      u8 *__restrict__ mask)
{
    __shared__ u64 s_key[256];
    u32 const iter = ctl->iter;
    u32 const f = blockIdx.x * 256u + threadIdx.x;
    u64 const mine = f < F ? feature_key(seed, iter, f) : 0ULL;
    u32 rank = 0;
    for (u32 g0 = 0; g0 < F; g0 += 256u) {
        __syncthreads();
        s_key[threadIdx.x] = g0 + threadIdx.x < F ? feature_key(seed, iter, g0 + threadIdx.x) : 0ULL;
        __syncthreads();
        u32 const n = min(256u, F - g0);
        for (u32 j = 0; j < n; ++j) {
            u64 const kj = s_key[j];
            rank += (kj < mine || (kj == mine && g0 + j < f)) ? 1u : 0u;
        }
    }
    if (f < F) mask[f] = rank < k ? 1 : 0;
}

// The root's monotone box is unbounded and its interaction path empty (SplitInput
// defaults, include/bonsai/split.hpp:19-34).
__global__ void k_constraints_root(double *__restrict__ bounds, u32 *__restrict__ paths, u32 path_words)
{
    if (bounds && threadIdx.x == 0) {
        bounds[0] = -__builtin_huge_val();
        bounds[1] = __builtin_huge_val();
    }
    if (paths)
        for (u32 w = threadIdx.x; w < path_words; w += blockDim.x) paths[w] = 0u;
}

void launch_constraints_root(hipStream_t s, double *bounds, u32 *paths, u32 path_words)
{
    hipLaunchKernelGGL(k_constraints_root, dim3(1), dim3(64), 0, s, bounds, paths, path_words);
}

void launch_feature_draw(hipStream_t s, TreeCtl const *ctl, u32 F, u32 k, u64 seed, u8 *mask)
{
    hipLaunchKernelGGL(k_feature_draw, dim3((F + 255u) / 256u), dim3(256), 0, s, ctl, F, k, seed, mask);
}

__global__ void k_gather_rows(u8 const *__restrict__ plane, u32 Fs, void const *__restrict__ gh, u32 gh_bytes,
                              u32 const *__restrict__ rows, u64 n, u8 *__restrict__ out_plane,
                              void *__restrict__ out_gh, int gh_is_row_id)
{
    u32 const D = Fs / 4u;
    for (u64 i = blockIdx.x * static_cast<u64>(blockDim.x) + threadIdx.x; i < n;
         i += static_cast<u64>(gridDim.x) * blockDim.x) {
        u32 const r = rows[i];
        u32 const *src = reinterpret_cast<u32 const *>(plane + static_cast<size_t>(r) * Fs);
        u32 *dst = reinterpret_cast<u32 *>(out_plane + i * Fs);
        for (u32 d = 0; d < D; ++d) dst[d] = src[d];
        if (gh_is_row_id) {
            if (gh_bytes == 8) reinterpret_cast<uint2 *>(out_gh)[i] = make_uint2(r, 0u);
            else reinterpret_cast<u32 *>(out_gh)[i] = r;
        } else if (gh_bytes == 8) {
            reinterpret_cast<uint2 *>(out_gh)[i] = reinterpret_cast<uint2 const *>(gh)[r];
        } else {
            reinterpret_cast<u32 *>(out_gh)[i] = reinterpret_cast<u32 const *>(gh)[r];
        }
    }
}

// Device bin transform (src/cuda/detail/ingest_kernels.cuh:27-50 is the reference's
// twin of BinMapper::transform, src/bin_mapper.cpp:284-291): bin = lower_bound(cuts, x),
// NaN -> the last bin.  One thread per (row, padded feature) cell of a chunk of rows; the
// chunk's float rows are read coalesced, the u8 plane rows are written whole, padding
// columns get bin 0.  Cut tables live in LDS when they fit (lds_cuts != 0).
__global__ __launch_bounds__(256) void k_bin_rows(float const *__restrict__ X, u64 n, u32 F, u32 Fs,
                                                  float const *__restrict__ cuts, u32 const *__restrict__ coff,
                                                  u8 *__restrict__ plane, u32 lds_cuts)
{
    extern __shared__ float s_cuts[];
    if (lds_cuts) {
        for (u32 i = threadIdx.x; i < lds_cuts; i += blockDim.x) s_cuts[i] = cuts[i];
        __syncthreads();
    }
    float const *C = lds_cuts ? s_cuts : cuts;
    u64 const total = n * Fs;
    for (u64 e = blockIdx.x * static_cast<u64>(blockDim.x) + threadIdx.x; e < total;
         e += static_cast<u64>(gridDim.x) * blockDim.x) {
        u64 const r = e / Fs;
        u32 const f = static_cast<u32>(e - r * Fs);
        u32 b = 0;
        if (f < F) {
            float const x = X[r * F + f];
            u32 const lo = coff[f], hi = coff[f + 1];
            if (isnan(x)) {
                b = hi - lo - 1u;
            } else {
                u32 first = lo, len = hi - lo;
                while (len > 0) {
                    u32 const half = len >> 1;
                    if (C[first + half] < x) {
                        first += half + 1u;
                        len -= half + 1u;
                    } else {
                        len = half;
                    }
                }
                b = first - lo;
            }
        }
        plane[e] = static_cast<u8>(b);
    }
}

// lower_bound of x over one feature's ascending cuts C[0..len), NaN -> the last bin.
__device__ __forceinline__ u32 bin_of(float x, float const *C, u32 len)
{
    if (isnan(x)) return len - 1u;
    u32 first = 0;
    while (len > 0) {
        u32 const half = len >> 1;
        if (C[first + half] < x) {
            first += half + 1u;
            len -= half + 1u;
        } else {
            len = half;
        }
    }
    return first;
}

// Column ingest (the reference's bin_col_kernel, src/cuda/detail/ingest_kernels.cuh:89-102,
// one column per launch there): up to four columns of the features f0..f0+3 (f0 a
// multiple of 4) arrive column-major in `cols` ([g][ld] floats, rows r0..r0+n of the
// plane); each lane bins one row's four cells against the features' cuts (staged in
// LDS) and writes them as one aligned dword, so the plane's bytes of a row group are
// written whole instead of one strided byte per feature.  Bytes past n_features are 0.
__global__ __launch_bounds__(256) void k_bin_cols4(float const *__restrict__ cols, u64 ld, u64 n, u32 G,
                                                   float const *__restrict__ cuts, u32 const *__restrict__ coff,
                                                   u32 f0, u32 Fs, u8 *__restrict__ plane)
{
    __shared__ float s_cuts[4 * 256];
    __shared__ u32 s_off[5];
    if (threadIdx.x <= G) s_off[threadIdx.x] = coff[f0 + threadIdx.x] - coff[f0];
    __syncthreads();
    for (u32 i = threadIdx.x; i < s_off[G]; i += blockDim.x) s_cuts[i] = cuts[coff[f0] + i];
    __syncthreads();
    for (u64 r = blockIdx.x * static_cast<u64>(blockDim.x) + threadIdx.x; r < n;
         r += static_cast<u64>(gridDim.x) * blockDim.x) {
        u32 word = 0;
        for (u32 g = 0; g < G; ++g) {
            u32 const b = bin_of(cols[g * ld + r], s_cuts + s_off[g], s_off[g + 1] - s_off[g]);
            word |= b << (8u * g);
        }
        *reinterpret_cast<u32 *>(plane + r * Fs + f0) = word;
    }
}

// Sample gather for the host cut fitting (gather_rows_kernel, ingest_kernels.cuh:76-86):
// out[i][:] = X[rows[i]][:], one lane per output float.
__global__ __launch_bounds__(256) void k_gather_f32_rows(float const *__restrict__ X, u32 const *__restrict__ rows,
                                                         u64 n_out, u32 F, float *__restrict__ out)
{
    u64 const cells = n_out * F;
    for (u64 i = blockIdx.x * static_cast<u64>(blockDim.x) + threadIdx.x; i < cells;
         i += static_cast<u64>(gridDim.x) * blockDim.x) {
        u64 const r = i / F;
        out[i] = X[static_cast<u64>(rows[r]) * F + (i - r * F)];
    }
}

// Plain 16-byte-per-lane copy: the calibration kernel of known traffic (reads and
// writes n * 16 bytes) against which HBM counters and bandwidth are read.
__global__ __launch_bounds__(256) void k_copy_calibrate(uint4 const *__restrict__ src, uint4 *__restrict__ dst, u64 n)
{
    for (u64 i = blockIdx.x * static_cast<u64>(blockDim.x) + threadIdx.x; i < n;
         i += static_cast<u64>(gridDim.x) * blockDim.x)
        dst[i] = src[i];
}

// Bandwidth probes of known traffic with U independent 16-byte nontemporal loads in
// flight per lane: a copy (reads and writes n * 16 bytes) and a read (n * 16 bytes
// read; one dword per workgroup written so the loads stay live).
typedef u32 u32x4v __attribute__((ext_vector_type(4)));
template <int U>
__global__ __launch_bounds__(256) void k_copy_wide(u32x4v const *__restrict__ src, u32x4v *__restrict__ dst, u64 n)
{
    u64 const stride = static_cast<u64>(gridDim.x) * 256u * U;
    for (u64 base = blockIdx.x * 256ull * U + threadIdx.x; base < n; base += stride) {
        u32x4v v[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            u64 const i = base + static_cast<u64>(u) * 256u;
            v[u] = i < n ? __builtin_nontemporal_load(src + i) : u32x4v{0u, 0u, 0u, 0u};
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            u64 const i = base + static_cast<u64>(u) * 256u;
            if (i < n) __builtin_nontemporal_store(v[u], dst + i);
        }
    }
}
template <int U>
__global__ __launch_bounds__(256) void k_read_wide(u32x4v const *__restrict__ src, u32 *__restrict__ sink, u64 n)
{
    __shared__ u32 acc;
    if (threadIdx.x == 0) acc = 0u;
    __syncthreads();
    u32 x = 0u;
    u64 const stride = static_cast<u64>(gridDim.x) * 256u * U;
    for (u64 base = blockIdx.x * 256ull * U + threadIdx.x; base < n; base += stride) {
        u32x4v v[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            u64 const i = base + static_cast<u64>(u) * 256u;
            v[u] = i < n ? __builtin_nontemporal_load(src + i) : u32x4v{0u, 0u, 0u, 0u};
        }
#pragma unroll
        for (int u = 0; u < U; ++u) x ^= v[u].x ^ v[u].y ^ v[u].z ^ v[u].w;
    }
    atomicXor(&acc, x);
    __syncthreads();
    if (threadIdx.x == 0) sink[blockIdx.x] = acc;
}

// ---------------------------------------------------------------------------------
// Launch wrappers (host side).
// ---------------------------------------------------------------------------------
// Scatter tile sizing: rows per LDS tile = budget / (row + gradient + 4 index bytes),
// a multiple of 64 in [64, kScatTileMax].
constexpr u32 kScatTileMax = 4096;

// Launch-shape knobs in force for the work the calling thread enqueues (each context
// activates its own before enqueueing; graphs keep the shapes they were captured with).
static thread_local Tuning t_tune;
void set_tuning(Tuning const &t) { t_tune = t; }
Tuning const &tuning() { return t_tune; }

void launch_copy_calibrate(hipStream_t s, void const *src, void *dst, u64 bytes, u32 mode, u32 grid)
{
    u64 const n = bytes / 16;
    if (mode == 1u) {
        hipLaunchKernelGGL(k_copy_wide<4>, dim3(grid), dim3(256), 0, s, static_cast<u32x4v const *>(src),
                           static_cast<u32x4v *>(dst), n);
        return;
    }
    if (mode == 2u) {
        hipLaunchKernelGGL(k_read_wide<8>, dim3(grid), dim3(256), 0, s, static_cast<u32x4v const *>(src),
                           static_cast<u32 *>(dst), n);
        return;
    }
    hipLaunchKernelGGL(k_copy_calibrate, dim3(static_cast<u32>(std::min<u64>((n + 255) / 256, 16384))), dim3(256), 0,
                       s, static_cast<uint4 const *>(src), static_cast<uint4 *>(dst), n);
}
void launch_bin_rows(hipStream_t s, float const *X, u64 n, u32 F, u32 Fs, float const *cuts, u32 const *coff,
                     u32 n_cuts_total, u8 *plane)
{
    u32 const lds_cuts = n_cuts_total * 4u <= 48u * 1024u ? n_cuts_total : 0u;
    u64 const total = n * Fs;
    u32 const grid = static_cast<u32>(std::min<u64>((total + 255) / 256, 8192));
    hipLaunchKernelGGL(k_bin_rows, dim3(grid), dim3(256), lds_cuts * 4u, s, X, n, F, Fs, cuts, coff, plane,
                       lds_cuts);
}
void launch_bin_cols4(hipStream_t s, float const *cols, u64 ld, u64 n, u32 G, float const *cuts, u32 const *coff,
                      u32 f0, u32 Fs, u8 *plane)
{
    u32 const grid = static_cast<u32>(std::max<u64>(1, std::min<u64>((n + 255) / 256, 8192)));
    hipLaunchKernelGGL(k_bin_cols4, dim3(grid), dim3(256), 0, s, cols, ld, n, G, cuts, coff, f0, Fs, plane);
}
void launch_gather_f32_rows(hipStream_t s, float const *X, u32 const *rows, u64 n_out, u32 F, float *out)
{
    u64 const cells = n_out * F;
    u32 const grid = static_cast<u32>(std::max<u64>(1, std::min<u64>((cells + 255) / 256, 8192)));
    hipLaunchKernelGGL(k_gather_f32_rows, dim3(grid), dim3(256), 0, s, X, rows, n_out, F, out);
}
void launch_grad_f32(hipStream_t s, float const *score, float const *label, float const *weight, u64 n, int obj,
                     float2 *gh, u32 *fx)
{
    hipLaunchKernelGGL(k_grad_f32, dim3(2048), dim3(256), 0, s, score, label, weight, n, obj, gh, fx);
    hipLaunchKernelGGL(k_fx_shift, dim3(1), dim3(64), 0, s, fx);
}
void launch_fx_from(hipStream_t s, float2 const *gh, u64 n, u32 *fx)
{
    hipLaunchKernelGGL(k_gh_max, dim3(1024), dim3(256), 0, s, gh, n, fx);
    hipLaunchKernelGGL(k_fx_shift, dim3(1), dim3(64), 0, s, fx);
}
void launch_grad_quant(hipStream_t s, float const *score, float const *label, float const *weight, u64 n, int obj,
                       u32 *maxbits, u64 seed, TreeCtl const *ctl, u32 *ghq, double *scales)
{
    hipLaunchKernelGGL(k_grad_quant, dim3(2048), dim3(256), 0, s, score, label, weight, n, obj, maxbits, seed, ctl,
                       ghq, scales);
}
void launch_grad_prime(hipStream_t s, float const *score, float const *label, float const *weight, u64 n, int obj,
                       u32 *maxbits, TreeCtl const *ctl)
{
    hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(64), 0, s, maxbits, 4u);
    hipLaunchKernelGGL(k_grad_max, dim3(1024), dim3(256), 0, s, score, label, weight, n, obj, maxbits, ctl);
}
void launch_quant_from(hipStream_t s, float2 const *gh, u64 n, u32 *maxbits, u64 seed, u32 iter, u32 *ghq,
                       double *scales_dev)
{
    hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(64), 0, s, maxbits, 2u);
    hipLaunchKernelGGL(k_gh_max, dim3(1024), dim3(256), 0, s, gh, n, maxbits);
    hipLaunchKernelGGL(k_quant_from, dim3(2048), dim3(256), 0, s, gh, n, maxbits, seed, iter, ghq);
    (void)scales_dev;
}

u32 hist_transpose_F(bool q, u32 Fs, u32 F, bool group)
{
    return q && !group && Fs <= 64u && (t_tune.hist_layout == 3u || t_tune.hist_layout == 5u) ? F : 0u;
}

hipError_t launch_hist(hipStream_t s, bool q, HistChunk const *chunks, TreeCtl const *ctl, u8 const *pl0,
                       u8 const *pl1, u8 const *plp, void const *g0, void const *g1, void const *gp, u32 Fs, u32 F,
                       u32 SB, u32 fsl, u32 n_waves, u32 grid_x, void *partials, void *hist_next, u32 const *fx,
                       u32 *cnt_next, PartOp const *routes, bool group)
{
    // group: units may span several 8192-row sub-units (the row-dword and row-in-lane
    // quantized kernels add each sub-unit into the unit's cells; the feature-in-lane one
    // takes single sub-units only).
    u32 const slices = (F + fsl - 1) / fsl;
    size_t const lds = static_cast<size_t>(fsl) * SB * 8;
    dim3 const grid(grid_x, slices), block(64 * n_waves);
    if (hist_transpose_F(q, Fs, F, group) && t_tune.hist_layout == 5u) {
        // rotated row-dword (single sub-units), every level of rows up to 64 bytes
        u32 const FS = hq4_fs(Fs);
        size_t const lds5 = static_cast<size_t>(FS) * SB * 8u;
        hipLaunchKernelGGL(k_hist_q5, dim3(grid_x), dim3(t_tune.hist_threads), lds5, s, chunks, ctl, pl0, pl1, plp,
                           g0, g1, gp, Fs, F, SB, FS, partials, hist_next, cnt_next, routes);
    } else if (hist_transpose_F(q, Fs, F, group)) {
        // rotated row-in-lane (single sub-units), every level of rows up to 64 bytes
        u32 const Dw = Fs / 4u;
        size_t const lds4 = static_cast<size_t>(hq4_fs(Fs)) * SB * 8u;
#define BH_HQ4(DWV, UV)                                                                                          \
    case DWV:                                                                                                     \
        hipLaunchKernelGGL((k_hist_q4<DWV, UV>), dim3(grid_x), dim3(t_tune.hist_threads), lds4, s, chunks, ctl,  \
                           pl0, pl1, plp, g0, g1, gp, Fs, F, SB, partials, hist_next, cnt_next, routes);          \
        break;
        switch (Dw) {
            BH_HQ4(1, 4) BH_HQ4(2, 4) BH_HQ4(3, 4) BH_HQ4(4, 4) BH_HQ4(5, 4) BH_HQ4(6, 4) BH_HQ4(7, 4)
            BH_HQ4(8, 3) BH_HQ4(9, 3) BH_HQ4(10, 2) BH_HQ4(11, 2) BH_HQ4(12, 2) BH_HQ4(13, 2) BH_HQ4(14, 2)
            BH_HQ4(15, 2) BH_HQ4(16, 2)
        default: break;
        }
#undef BH_HQ4
    } else if (q && Fs <= 64u && ((routes && t_tune.hist_layout != 4u) || t_tune.hist_layout == 2u)) {
        // row-in-lane: whole rows per lane, instantiated per row width in dwords.  Routed
        // levels always use it (a filtered row skips all its atomics at once).
        u32 const Dw = Fs / 4u;
        dim3 const gq3(grid_x);
        size_t const lds1 = static_cast<size_t>(F) * SB * 8u;
#define BH_HQ3(DWV, UV)                                                                                          \
    case DWV:                                                                                                     \
        hipLaunchKernelGGL((k_hist_q3<DWV, UV>), gq3, dim3(t_tune.hist_threads), lds1, s, chunks, ctl, pl0, pl1, \
                           plp, g0, g1, gp, Fs, F, SB, partials, hist_next, cnt_next, routes);                  \
        break;
        switch (Dw) {
            BH_HQ3(1, 4) BH_HQ3(2, 4) BH_HQ3(3, 4) BH_HQ3(4, 4) BH_HQ3(5, 4) BH_HQ3(6, 4) BH_HQ3(7, 4)
            BH_HQ3(8, 3) BH_HQ3(9, 3) BH_HQ3(10, 2) BH_HQ3(11, 2) BH_HQ3(12, 2) BH_HQ3(13, 2) BH_HQ3(14, 2)
            BH_HQ3(15, 2) BH_HQ3(16, 2)
        default: break;
        }
#undef BH_HQ3
    } else if (q && !group && Fs <= kHq2Cols && t_tune.hist_layout == 1u) {
        u32 const lh_bytes = SB * kHq2Cols * 8u;
        u32 tile = lh_bytes + 32u + 64u * (Fs + 4u) <= kHq2LdsBudget
                       ? ((kHq2LdsBudget - lh_bytes - 32u) / (Fs + 4u)) & ~63u
                       : 64u;
        tile = std::max(64u, std::min(tile, 1024u));
        size_t const lds2 = static_cast<size_t>(lh_bytes) + tile * (Fs + 4u) + 32u;
        hipLaunchKernelGGL(k_hist_q2, dim3(grid_x), dim3(512), lds2, s, chunks, ctl, pl0, pl1, plp, g0, g1, gp, Fs, F,
                           SB, tile, partials, hist_next, cnt_next, routes);
    } else if (q) {
        hipLaunchKernelGGL(k_hist_q, grid, dim3(t_tune.hist_threads), lds, s, chunks, ctl, pl0, pl1, plp, g0, g1, gp,
                           Fs, F, SB, fsl, partials, hist_next, cnt_next, routes);
    } else {
        if (!fx) return hipErrorInvalidValue;
        hipLaunchKernelGGL(k_hist_fx, grid, dim3(t_tune.hist_threads), lds * 2u, s, chunks, ctl, pl0, pl1, plp, g0,
                           g1, gp, Fs, F, SB, fsl, partials, hist_next, fx);
    }
    (void)n_waves;
    return hipGetLastError();
}

// Opt the LDS-heavy kernels into the device's full per-workgroup LDS (160 KiB on
// gfx950).  Best effort: a runtime that does not gate dynamic LDS behind the attribute
// rejects the call, and launches are then checked against the device limit.
static void raise_lds(void const *fn, int max_shared)
{
    hipFuncAttributes a{};
    if (hipFuncGetAttributes(&a, fn) != hipSuccess) return;
    int const dyn = max_shared - static_cast<int>(a.sharedSizeBytes);
    if (dyn > 0) (void)hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, dyn);
    (void)hipGetLastError();
}

hipError_t set_hist_lds_limits()
{
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    int max_shared = 0;
    e = hipDeviceGetAttribute(&max_shared, hipDeviceAttributeMaxSharedMemoryPerBlock, dev);
    if (e != hipSuccess) return e;
    raise_lds(reinterpret_cast<void const *>(&k_hist_q), max_shared);
    void const *q4[] = {reinterpret_cast<void const *>(&k_hist_q4<1, 4>),  reinterpret_cast<void const *>(&k_hist_q4<2, 4>),
                        reinterpret_cast<void const *>(&k_hist_q4<3, 4>),  reinterpret_cast<void const *>(&k_hist_q4<4, 4>),
                        reinterpret_cast<void const *>(&k_hist_q4<5, 4>),  reinterpret_cast<void const *>(&k_hist_q4<6, 4>),
                        reinterpret_cast<void const *>(&k_hist_q4<7, 4>),  reinterpret_cast<void const *>(&k_hist_q4<8, 3>),
                        reinterpret_cast<void const *>(&k_hist_q4<9, 3>),  reinterpret_cast<void const *>(&k_hist_q4<10, 2>),
                        reinterpret_cast<void const *>(&k_hist_q4<11, 2>), reinterpret_cast<void const *>(&k_hist_q4<12, 2>),
                        reinterpret_cast<void const *>(&k_hist_q4<13, 2>), reinterpret_cast<void const *>(&k_hist_q4<14, 2>),
                        reinterpret_cast<void const *>(&k_hist_q4<15, 2>), reinterpret_cast<void const *>(&k_hist_q4<16, 2>)};
    for (void const *fn : q4) raise_lds(fn, max_shared);
    raise_lds(reinterpret_cast<void const *>(&k_hist_q2), max_shared);
    raise_lds(reinterpret_cast<void const *>(&k_hist_q5), max_shared);
    raise_lds(reinterpret_cast<void const *>(&k_hist_fx), max_shared);
    raise_lds(reinterpret_cast<void const *>(&k_find<true, false>), max_shared);
    raise_lds(reinterpret_cast<void const *>(&k_find<false, false>), max_shared);
    raise_lds(reinterpret_cast<void const *>(&k_find<true, true>), max_shared);
    raise_lds(reinterpret_cast<void const *>(&k_find<false, true>), max_shared);
    raise_lds(reinterpret_cast<void const *>(&k_part_scatter4<false>), max_shared);
    raise_lds(reinterpret_cast<void const *>(&k_part_scatter4<true>), max_shared);
    raise_lds(reinterpret_cast<void const *>(&k_part_scatter_v), max_shared);
    return hipSuccess;
}

void launch_finish(hipStream_t s, bool q, FinishOp const *ops, TreeCtl *ctl, void const *partials,
                   void const *hist_cur, void *hist_next, u32 node_cells, u32 max_ops, u32 const *cnt_cur,
                   u32 *cnt_next, bool advance, u32 tF)
{
    u32 const tiles = min(64u, (node_cells + 255) / 256);
    dim3 const grid(tiles, max(1u, min(max_ops, 1024u)));
    if (q) {
        // Enough workgroups to cover the cells of one op several times over (up to 16,384
        // in all, so a level of many ops does not walk a long chain of cell tiles per
        // workgroup); ops beyond the grid's y extent loop.
        u32 const ops_y = max(1u, min(max_ops, 1024u));
        u32 const cell_tiles = (node_cells + kFqCells - 1) / kFqCells;
        u32 const tx = max(1u, min(cell_tiles, max(1u, 16384u / ops_y)));
        hipLaunchKernelGGL(k_finish_q, dim3(tx, ops_y), dim3(256), 0, s, ops, ctl, partials, hist_cur, hist_next,
                           node_cells, cnt_cur, cnt_next, advance ? 1u : 0u, tF);
    } else {
        hipLaunchKernelGGL(k_finish<false>, grid, dim3(256), 0, s, ops, ctl, partials, hist_cur, hist_next,
                           node_cells, nullptr, nullptr);
        if (advance) launch_next_level(s, ctl);
    }
}

void launch_find(hipStream_t s, bool q, NodeInfo *front, TreeCtl const *ctl, void const *hist, u16 const *n_bins,
                 u32 F, u32 SB, GrowCfg const &cfg, double const *scales, SplitRec *out, u32 max_front,
                 u32 front_base, u32 n_front_override, FindScratch const &fs, u32 const *cnt)
{
    // cells + two prefix planes, and (quantized with count cells) the counts' prefix
    size_t const lds = static_cast<size_t>(kFindTile) * SB * (q ? (cnt ? 36 : 32) : 24);
    u32 const G = find_blocks_per_node(F, kFindTile);
    dim3 const grid(G, max_front);
    // A 1,024-thread workgroup occupies a whole CU (its register use leaves room for 16
    // waves): when the level has more workgroups than CUs, halve them so two fit per CU.
    int dev = 0, cus = 256;
    if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) !=
                                                hipSuccess)
        cus = 256;
    u32 const threads = static_cast<u64>(G) * max_front > static_cast<u64>(cus) ? kFindThreads / 2u : kFindThreads;
    bool const con = cfg.fmask || cfg.mono || cfg.n_groups;
    auto const fn = q ? (con ? &k_find<true, true> : &k_find<true, false>)
                      : (con ? &k_find<false, true> : &k_find<false, false>);
    hipLaunchKernelGGL(fn, grid, dim3(threads), lds, s, front, ctl, hist, n_bins, F, SB, cfg, scales, out, front_base,
                       n_front_override, fs.part, fs.done, q ? cnt : nullptr);
}

void launch_plan_level(hipStream_t s, NodeInfo const *front, SplitRec const *splits, TreeCtl *ctl, DevNode *tree,
                       PartOp *ops, PartChunk *pch, u16 const *n_bins, GrowCfg const &cfg, double const *scales,
                       u32 dst, u32 chunk_rows)
{
    hipLaunchKernelGGL(k_plan_level, dim3(1), dim3(1024), 0, s, front, splits, ctl, tree, ops, pch, n_bins, cfg,
                       scales, dst, chunk_rows);
}

u32 part_chunk_rows(u32 Fs)
{
    if (Fs <= 64) return kPartChunk;
    u32 r = kPartChunk;
    while (r > 64u && static_cast<u64>(r) * Fs > kPartChunkBytes) r >>= 1;
    return r;
}

void launch_part_count(hipStream_t s, PartChunk const *pch, PartOp const *ops, TreeCtl const *ctl, u8 const *pl0,
                       u8 const *pl1, u8 const *plp, u32 Fs, u32 *chunk_left, u32 grid)
{
    hipLaunchKernelGGL(k_part_count, dim3(grid), dim3(256), 0, s, pch, ops, ctl, pl0, pl1, plp, Fs, chunk_left);
}

void launch_part_scan(hipStream_t s, bool q, PartOp *ops, PartChunk *pch, u32 const *chunk_left, TreeCtl *ctl,
                      NodeInfo const *front, NodeInfo *next, SplitRec const *splits, DevNode *tree, HistChunk *hch,
                      FinishOp *fops, GrowCfg const &cfg, double const *scales, u32 is_last, u32 oblivious)
{
    if (q)
        hipLaunchKernelGGL(k_part_scan<true>, dim3(1), dim3(1024), 0, s, ops, pch, chunk_left, ctl, front, next,
                           splits, tree, hch, fops, cfg, scales, is_last, oblivious);
    else
        hipLaunchKernelGGL(k_part_scan<false>, dim3(1), dim3(1024), 0, s, ops, pch, chunk_left, ctl, front, next,
                           splits, tree, hch, fops, cfg, scales, is_last, oblivious);
}

void launch_part_scatter(hipStream_t s, PartChunk const *pch, PartOp const *ops, TreeCtl const *ctl,
                         u8 const *pl0, u8 const *pl1, u8 const *plp, void const *g0, void const *g1,
                         void const *gp, u8 *o0, u8 *o1, void *go0, void *go1, u32 Fs, u32 gh_bytes, u32 grid,
                         u32 const *chunk_left)
{
    u32 const D = Fs / 4;
    if (D <= 16) {
        u32 tile = scatter_tile_rows(Fs, gh_bytes);
        size_t const lds = static_cast<size_t>(tile) * (Fs + gh_bytes + 4u) + 32u;
        u32 const G = gh_bytes / 4u;
        u64 const magic_d = ((1ull << 32) + D - 1u) / D;
        u64 const magic_g = ((1ull << 32) + G - 1u) / G;
        hipLaunchKernelGGL(k_part_scatter_v, dim3(grid), dim3(kScatterThreads), lds, s, pch, ops, ctl, pl0, pl1, plp, g0, g1,
                           gp, o0, o1, go0, go1, Fs, gh_bytes, tile, magic_d, magic_g, chunk_left,
                           const_cast<TreeCtl *>(ctl));
        return;
    }
    u32 tl = 0;
    while ((1u << tl) < D && tl < 8) ++tl;
    hipLaunchKernelGGL(k_part_scatter, dim3(grid), dim3(256), 0, s, pch, ops, ctl, pl0, pl1, plp, g0, g1, gp, o0,
                       o1, go0, go1, Fs, gh_bytes, tl, chunk_left, const_cast<TreeCtl *>(ctl));
}

u32 scatter_tile_rows(u32 Fs, u32 gh_bytes)
{
    u32 const per_row = Fs + gh_bytes + 4u;
    u32 tile = (t_tune.scatter_lds - 32u) / per_row;
    // the register prefetch of k_part_scatter_v holds a whole tile
    u32 const pf_rows = (kPfRows * 16u * 256u - 32u) / Fs, pf_gh = (kPfGh * 16u * 256u - 32u) / gh_bytes;
    tile = std::min(tile, std::min(pf_rows, pf_gh));
    return std::max(64u, std::min<u32>(kScatTileMax, tile & ~63u));
}

void launch_plan_counted(hipStream_t s, NodeInfo const *front, SplitRec const *splits, TreeCtl *ctl, DevNode *tree,
                         PartOp *ops, PartChunk *pch, NodeInfo *next, HistChunk *hch, FinishOp *fops,
                         u16 const *n_bins, GrowCfg const &cfg, double const *scales, u32 dst, u32 is_last,
                         u32 chunk_rows, u32 mode, PartOp4 *ops4, u32 unit_rows, u32 *cursors)
{
    hipLaunchKernelGGL(k_plan_counted, dim3(1), dim3(1024), 0, s, front, splits, ctl, tree, ops, pch, next, hch, fops,
                       n_bins, cfg, scales, dst, is_last, chunk_rows, mode, ops4, unit_rows, cursors);
}

void launch_part_count4(hipStream_t s, PartChunk const *pch, PartOp4 const *ops4, TreeCtl const *ctl, u8 const *pl0,
                        u8 const *pl1, u8 const *plp, u32 Fs, uint4 *chunk_cls, u32 grid)
{
    // whole tiles in the register prefetch (kPfRows 16-byte words per thread)
    u32 const tile = std::max(64u, std::min(kPartChunk, ((kPfRows * 16u * 256u - 16u) / Fs) & ~63u));
    size_t const lds = static_cast<size_t>(tile) * Fs + 16u;
    hipLaunchKernelGGL(k_part_count4, dim3(grid), dim3(256), lds, s, pch, ops4, ctl, pl0, pl1, plp, Fs, tile,
                       chunk_cls);
}

u32 scatter4_tile_rows(u32 Fs)
{
    u32 const per_row = Fs + 4u + 8u;
    u32 tile = (t_tune.scatter_lds - 32u) / per_row;
    // the register prefetch holds a whole tile (rows and gradient words)
    u32 const pf_rows = (kPfRows * 16u * 256u - 32u) / Fs, pf_gh = (kPfGh * 16u * 256u - 32u) / 4u;
    tile = std::min(tile, std::min(pf_rows, pf_gh));
    return std::max(64u, std::min<u32>(kScatTileMax, tile & ~63u));
}

void launch_part_scatter4(hipStream_t s, PartChunk const *pch, PartOp4 const *ops4, TreeCtl *ctl, u8 const *pl0,
                          u8 const *pl1, u8 const *plp, u32 const *g0, u32 const *g1, u32 const *gp, u8 *o0,
                          u8 *o1, u32 *go0, u32 *go1, u32 Fs, u32 grid, uint4 const *chunk_cls, u32 *cursors)
{
    u32 const D = Fs / 4u;
    u32 const tile = scatter4_tile_rows(Fs);
    size_t const lds = static_cast<size_t>(tile) * (Fs + 12u) + 32u;
    u64 const magic_d = ((1ull << 32) + D - 1u) / D;
    if (cursors)
        hipLaunchKernelGGL(k_part_scatter4<true>, dim3(grid), dim3(kScatterThreads), lds, s, pch, ops4, ctl, pl0, pl1, plp, g0, g1,
                           gp, o0, o1, go0, go1, Fs, tile, magic_d, chunk_cls, cursors);
    else
        hipLaunchKernelGGL(k_part_scatter4<false>, dim3(grid), dim3(kScatterThreads), lds, s, pch, ops4, ctl, pl0, pl1, plp, g0,
                           g1, gp, o0, o1, go0, go1, Fs, tile, magic_d, chunk_cls, cursors);
}


void launch_finalize(hipStream_t s, NodeInfo const *front, TreeCtl const *ctl, DevNode *tree, GrowCfg const &cfg,
                     double const *scales, u32 max_front)
{
    hipLaunchKernelGGL(k_finalize, dim3(max(1u, (max_front + 255) / 256)), dim3(256), 0, s, front, ctl, tree, cfg,
                       scales);
}

void launch_next_level(hipStream_t s, TreeCtl *ctl)
{
    hipLaunchKernelGGL(k_next_level, dim3(1), dim3(64), 0, s, ctl);
}

void launch_tree_begin(hipStream_t s, TreeCtl *ctl, NodeInfo *front, DevNode *tree, HistChunk *hch,
                       FinishOp *fops, RootSel const &root, u32 unit_rows)
{
    hipLaunchKernelGGL(k_tree_begin, dim3(1), dim3(256), 0, s, ctl, front, tree, hch, fops, root.n, unit_rows,
                       root.src, root.n_from_ctl);
}

void launch_score_update(hipStream_t s, u8 const *plane, u32 Fs, u64 n, DevNode const *tree, TreeCtl const *ctl,
                         u16 const *n_bins, u32 F, float lr, float *score, u32 *leaf_of_row, u32 max_nodes,
                         float const *label, int objective, u32 *maxbits, float const *weight)
{
    u32 const node_cap = std::min(max_nodes, kWalkLds);
    u32 const last_cap = std::min(F, kWalkLds);
    size_t const head = (node_cap * sizeof(WalkNode) + last_cap * sizeof(u16) + 15u) & ~static_cast<size_t>(15);
    if (Fs <= 64)
        hipLaunchKernelGGL(k_score_update<true>, dim3(2048), dim3(256), head + 256u * Fs + 16u, s, plane, Fs, n, tree,
                           ctl, n_bins, F, lr, score, leaf_of_row, node_cap, last_cap, label, weight, objective,
                           maxbits);
    else
        hipLaunchKernelGGL(k_score_update<false>, dim3(2048), dim3(256), head, s, plane, Fs, n, tree, ctl, n_bins, F,
                           lr, score, leaf_of_row, node_cap, last_cap, label, weight, objective, maxbits);
}

u32 eval_blocks(u32 n) { return std::max(1u, std::min(n / 256u, 1024u)); }

void launch_eval(hipStream_t s, int objective, float const *scores, float const *labels, u32 n, double *partial,
                 TreeCtl *ctl, float *hist, u32 hist_cap, u32 patience)
{
    u32 const blocks = eval_blocks(n);
    if (objective == 1)
        hipLaunchKernelGGL(k_eval_loss<1>, dim3(blocks), dim3(256), 0, s, scores, labels, n, partial);
    else if (objective == 2)
        hipLaunchKernelGGL(k_eval_loss<2>, dim3(blocks), dim3(256), 0, s, scores, labels, n, partial);
    else
        hipLaunchKernelGGL(k_eval_loss<0>, dim3(blocks), dim3(256), 0, s, scores, labels, n, partial);
    hipLaunchKernelGGL(k_eval_finish, dim3(1), dim3(256), 0, s, partial, blocks, n, ctl, hist, hist_cap, patience);
}

void launch_iter_advance(hipStream_t s, TreeCtl *ctl)
{
    hipLaunchKernelGGL(k_iter_advance, dim3(1), dim3(64), 0, s, ctl);
}

void launch_gather_rows(hipStream_t s, u8 const *plane, u32 Fs, void const *gh, u32 gh_bytes, u32 const *rows,
                        u64 n, u8 *out_plane, void *out_gh, int gh_is_row_id)
{
    hipLaunchKernelGGL(k_gather_rows, dim3(1024), dim3(256), 0, s, plane, Fs, gh, gh_bytes, rows, n, out_plane,
                       out_gh, gh_is_row_id);
}

void launch_copy_tree(hipStream_t s, DevNode const *tree, TreeCtl const *ctl, DevNode *out_trees, u32 *out_sizes,
                      u32 max_nodes)
{
    hipLaunchKernelGGL(k_copy_tree, dim3(1), dim3(256), 0, s, tree, ctl, out_trees, out_sizes, max_nodes);
}

} // namespace bh
