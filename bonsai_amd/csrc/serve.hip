// Predict / TreeSHAP plans and the continuous-batching serving loop (H15, H16, cfg4).
//
// A bh_predictor holds one ensemble resident in HBM in two packed forms:
//   walk nodes (8 B each) for the bin-space predict walk, and
//   leaf paths (8 B elements, 12 B heads) for TreeSHAP's closed form,
// and is immutable after creation, so serving streams share it.
#include <hip/hip_runtime.h>
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "../../include/bonsai_hip.h"
#include "bh_common.hpp"

using namespace bh;

namespace {

} // namespace

// The engine's thread-local error slot (bh_last_error).
void bh_set_error(std::string const &m);

namespace {

int sfail(int code, std::string const &m)
{
    bh_set_error(m);
    return code;
}
#define SHIP(expr)                                                                          \
    do {                                                                                    \
        hipError_t const e_ = (expr);                                                       \
        if (e_ != hipSuccess) return sfail(-1, std::string("hip: ") + #expr + ": " + hipGetErrorString(e_)); \
    } while (0)

// Predict walk node: internal {feature, bin, default_left, left, right}; leaf
// {0xFFFF, value}.  Child ids are local to the tree.
struct PNode {
    u16 feature;
    u8 bin;
    u8 dl;
    union {
        struct {
            u16 left, right;
        } c;
        float value;
    } u;
};
static_assert(sizeof(PNode) == 8, "PNode");

// TreeSHAP packed path element / head (src/shap_paths.cpp:33-110 layout).
struct SElem {
    u16 feature; // bit 15: the missing bin satisfies
    u8 lo, hi;
    float z;
};
struct SHead {
    u32 first;
    u16 n;
    u16 klass;
    float value;
};
static_assert(sizeof(SElem) == 8 && sizeof(SHead) == 12, "shap layout");

constexpr int kMaxPath = 16;

// ---------------------------------------------------------------------------------
// Predict: init + lr * sum_t leaf_t(row), the tree sum in tree order, multiply then
// add each rounded (predict_walk_kernel, src/cuda/predict.cu:50-85; host
// predict_at_binned, include/bonsai/booster.hpp:831-846).  A workgroup takes kPR rows;
// its threads walk (row, tree) pairs of one tree tile into LDS, then one thread per row
// adds the tile in tree order.
// ---------------------------------------------------------------------------------
constexpr u32 kPR = 16, kPT = 256;

__global__ __launch_bounds__(256) void k_predict(u8 const *__restrict__ bins, u64 n, u64 stride, u32 F,
                                                 PNode const *__restrict__ nodes, u32 const *__restrict__ roots,
                                                 u32 n_trees, u16 const *__restrict__ last, float lr, float init,
                                                 float *__restrict__ out)
{
    __shared__ float s_val[kPR][kPT];
    __shared__ float s_acc[kPR];
    extern __shared__ __align__(16) unsigned char s_rows[]; // kPR * F bytes
    u64 const r0 = static_cast<u64>(blockIdx.x) * kPR;
    if (r0 >= n) return;
    u32 const nr = static_cast<u32>(min<u64>(kPR, n - r0));
    for (u32 k = threadIdx.x; k < nr * F; k += blockDim.x) {
        u32 const rr = k / F, f = k - rr * F;
        s_rows[k] = bins[(r0 + rr) * stride + f];
    }
    if (threadIdx.x < kPR) s_acc[threadIdx.x] = 0.0F;
    __syncthreads();
    for (u32 t0 = 0; t0 < n_trees; t0 += kPT) {
        u32 const tt = min(kPT, n_trees - t0);
        for (u32 k = threadIdx.x; k < nr * tt; k += blockDim.x) {
            u32 const rr = k / tt, t = k - rr * tt;
            PNode const *tn = nodes + roots[t0 + t];
            u8 const *row = s_rows + rr * F;
            PNode nd = tn[0];
            while (nd.feature != 0xFFFFu) {
                u32 const b = row[nd.feature];
                bool const left = b == last[nd.feature] ? (nd.dl != 0) : (b <= nd.bin);
                nd = tn[left ? nd.u.c.left : nd.u.c.right];
            }
            s_val[rr][t] = nd.u.value;
        }
        __syncthreads();
        if (threadIdx.x < nr) {
            float acc = s_acc[threadIdx.x];
            for (u32 t = 0; t < tt; ++t) acc += s_val[threadIdx.x][t];
            s_acc[threadIdx.x] = acc;
        }
        __syncthreads();
    }
    if (threadIdx.x < nr) out[r0 + threadIdx.x] = __fadd_rn(init, __fmul_rn(lr, s_acc[threadIdx.x]));
}

// ---------------------------------------------------------------------------------
// TreeSHAP, packed closed form (eval_shap_paths, src/shap_paths.cpp:196-278):
//   P(t) = prod_j (z_j + o_j t);  unsatisfied element: -value * sum_i w_i [t^i] P;
//   satisfied element: value * (1 - z) * sum_i w_i [t^i] (P / (t + z)).
// Paths are packed in groups of 64, element j of the group's paths stored contiguously
// ([group][j][lane]), so one coalesced load gives every lane its path's element.  A
// workgroup serves kShapRows rows: each lane loads a path once and applies it to all of
// them.  Contributions are accumulated as 64-bit fixed point (value x scale, scale from
// the ensemble's leaf magnitudes) with LDS integer atomics: exact and order-free, so a
// row's result does not depend on the order of the adds or on what shares its batch.
// ---------------------------------------------------------------------------------
constexpr u32 kShapRows = 32;
// A launch with fewer row blocks than this splits each block's path groups over more
// workgroups, so small batches still spread over the chip.
constexpr u32 kShapFillBlocks = 2048;

// φ of one cell from its 64-bit fixed-point sum.
__device__ __forceinline__ double shap_from_fixed(unsigned long long v, double scale, float lr)
{
    return static_cast<double>(static_cast<long long>(v)) / scale * static_cast<double>(lr);
}

// KM: compile-time bound on the path length (8 covers depth-8 trees), so every per-path
// array lives in registers.
template <int KM>
__global__ __launch_bounds__(256) void k_shap(u8 const *__restrict__ bins, u64 n, u64 stride, u32 F,
                                              SElem const *__restrict__ elems, SHead const *__restrict__ heads,
                                              u32 n_groups, u32 K, float const *__restrict__ weights,
                                              u16 const *__restrict__ last, float lr, double bias, double scale,
                                              double *__restrict__ out)
{
    extern __shared__ __align__(16) unsigned char smem[];
    unsigned long long *s_phi = reinterpret_cast<unsigned long long *>(smem); // [kShapRows][F] fixed point
    u8 *s_row = reinterpret_cast<u8 *>(s_phi + kShapRows * F);                 // [kShapRows][F]
    u64 const r0 = static_cast<u64>(blockIdx.x) * kShapRows;
    if (r0 >= n) return;
    u32 const nr = static_cast<u32>(min<u64>(kShapRows, n - r0));
    for (u32 k = threadIdx.x; k < kShapRows * F; k += blockDim.x) s_phi[k] = 0ull;
    for (u32 k = threadIdx.x; k < nr * F; k += blockDim.x) {
        u32 const rr = k / F, f = k - rr * F;
        s_row[k] = bins[(r0 + rr) * stride + f];
    }
    __syncthreads();
    // gridDim.y workgroups share a row block, each taking every gridDim.y-th set of
    // NW path groups.
    u32 const wave = threadIdx.x >> 6, lane = threadIdx.x & 63u, NW = blockDim.x >> 6;
    for (u32 g = blockIdx.y * NW + wave; g < n_groups; g += NW * gridDim.y) {
        SHead const hd = heads[static_cast<size_t>(g) * 64u + lane];
        u32 const nel = hd.n;
        // The group's longest path (wave-uniform): the KM = 16 body skips element slots
        // past it; the KM = 8 body runs all eight.
        u32 nmax = nel;
        for (int o = 32; o > 0; o >>= 1) nmax = max(nmax, static_cast<u32>(__shfl_xor(static_cast<int>(nmax), o)));
        nmax = __builtin_amdgcn_readfirstlane(nmax);
        if (nmax == 0) continue;
        // Slots past a lane's path hold a neutral element: z = 1 and never satisfied
        // (lo > hi, a missing bin no byte equals), so it multiplies P(t) by exactly 1 and
        // its deflation terms are exact zeros (P's coefficients past the path are 0, as
        // are its weights).  No per-lane guard is needed until the contribution's add.
        u32 feat[KM], lo[KM], hi[KM], lst[KM];
        bool miss[KM];
        float z[KM], wl[KM], vz[KM];
        float const *w = weights + (nel - 1) * nel / 2;
        // The leaf value in fixed-point units, rounded once to float: each contribution
        // is then a float product of the path's float polynomial terms, rounded to the
        // 64-bit accumulator's integer grid (|value| * scale <= 2^59).
        float const vs = static_cast<float>(static_cast<double>(hd.value) * scale);
#pragma unroll
        for (int j = 0; j < KM; ++j) {
            feat[j] = 0;
            lo[j] = 1;
            hi[j] = 0;
            lst[j] = 0xFFFFFFFFu;
            miss[j] = false;
            z[j] = 1.0F;
            wl[j] = 0.0F;
            vz[j] = 0.0F;
            if (static_cast<u32>(j) < nel) {
                SElem const e = elems[(static_cast<size_t>(g) * K + j) * 64u + lane];
                feat[j] = e.feature & 0x7FFFu;
                miss[j] = (e.feature & 0x8000u) != 0;
                lo[j] = e.lo;
                hi[j] = e.hi;
                lst[j] = last[feat[j]]; // the element's missing bin, read once per path
                z[j] = e.z;
                wl[j] = w[j];
                vz[j] = vs * (1.0F - e.z); // a satisfied element's factor before its sum
            }
        }
        // Lanes start on different rows, so one atomic instruction spreads over the
        // rows' accumulators instead of piling onto one row's few features.
        u32 rr = lane % nr;
        for (u32 r = 0; r < nr; ++r, rr = rr + 1 == nr ? 0u : rr + 1) {
            u8 const *row = s_row + rr * F;
            unsigned long long *phi = s_phi + rr * F;
            float poly[KM + 1];
            bool sat[KM];
#pragma unroll
            for (int j = 0; j <= KM; ++j) poly[j] = j == 0 ? 1.0F : 0.0F;
#pragma unroll
            for (int j = 0; j < KM; ++j) {
                sat[j] = false;
                if (KM <= 8 || static_cast<u32>(j) < nmax) {
                    u32 const b = row[feat[j]];
                    bool const one = b == lst[j] ? miss[j] : (b >= lo[j] && b <= hi[j]);
                    sat[j] = one;
                    // multiply by (z + o t)
                    float const zj = z[j];
#pragma unroll
                    for (int i = KM; i >= 1; --i)
                        if (i <= j + 1) poly[i] = __builtin_fmaf(zj, poly[i], one ? poly[i - 1] : 0.0F);
                    poly[0] *= zj;
                }
            }
            float unsat = 0.0F;
#pragma unroll
            for (int i = 0; i < KM; ++i) unsat = __builtin_fmaf(wl[i], poly[i], unsat); // wl[i] = 0 beyond the path
            float const nu = -vs * unsat;
#pragma unroll
            for (int k = 0; k < KM; ++k) {
                if (KM <= 8 || static_cast<u32>(k) < nmax) {
                    // synthetic division by (t + z_k), from the top: deflated[i] =
                    // P[i+1] - z_k deflated[i+1], zero above the path's degree; used only
                    // where element k is satisfied (P then holds the factor t + z_k)
                    float const nzk = -z[k];
                    float sdef = 0.0F, prev = 0.0F;
#pragma unroll
                    for (int i = KM - 1; i >= 0; --i) {
                        if (KM <= 8 || static_cast<u32>(i) < nmax) {
                            prev = __builtin_fmaf(nzk, prev, poly[i + 1]);
                            sdef = __builtin_fmaf(wl[i], prev, sdef);
                        }
                    }
                    float const c = sat[k] ? vz[k] * sdef : nu;
                    if (static_cast<u32>(k) < nel)
                        atomicAdd(phi + feat[k], static_cast<unsigned long long>(__float2ll_rn(c)));
                }
            }
        }
    }
    __syncthreads();
    if (gridDim.y > 1) {
        // Path groups split over workgroups: the block's fixed-point sums go into the
        // zeroed output cells with 64-bit integer atomics (exact in any order), and
        // k_shap_to_double converts them once every workgroup has added its share.
        unsigned long long *acc = reinterpret_cast<unsigned long long *>(out);
        for (u32 k = threadIdx.x; k < nr * F; k += blockDim.x) {
            u32 const rr = k / F, f = k - rr * F;
            if (s_phi[k] != 0ull) atomicAdd(acc + (r0 + rr) * (F + 1) + f, s_phi[k]);
        }
        return;
    }
    for (u32 k = threadIdx.x; k < nr * F; k += blockDim.x) {
        u32 const rr = k / F, f = k - rr * F;
        out[(r0 + rr) * (F + 1) + f] = shap_from_fixed(s_phi[k], scale, lr);
    }
    for (u32 rr = threadIdx.x; rr < nr; rr += blockDim.x) out[(r0 + rr) * (F + 1) + F] = bias;
}

// Split launches: the output cells start at zero (a fill kernel, so the sequence stays
// capturable in a graph) ...
__global__ __launch_bounds__(256) void k_shap_zero(unsigned long long *__restrict__ acc, u64 cells)
{
    for (u64 i = blockIdx.x * 256ull + threadIdx.x; i < cells; i += static_cast<u64>(gridDim.x) * 256ull) acc[i] = 0ull;
}

// ... and end as doubles, each cell converted in place by the thread that reads it.
__global__ __launch_bounds__(256) void k_shap_to_double(double *__restrict__ out, u64 n, u32 F, float lr, double bias,
                                                        double scale)
{
    u64 const cells = n * (F + 1);
    for (u64 i = blockIdx.x * 256ull + threadIdx.x; i < cells; i += static_cast<u64>(gridDim.x) * 256ull) {
        u32 const f = static_cast<u32>(i % (F + 1));
        if (f == F) {
            out[i] = bias;
        } else {
            unsigned long long const v = reinterpret_cast<unsigned long long const *>(out)[i];
            out[i] = shap_from_fixed(v, scale, lr);
        }
    }
}

struct ActiveElem {
    u32 feature;
    int lo, hi;
    bool missing_ok;
    double zero;
};

// pack_shap_paths (src/shap_paths.cpp:57-163): every root-to-leaf path, repeated
// features merged into one interval, zero fractions multiplied along the path.
void pack_paths(bh_node const *t, u16 const *last, u32 id, std::vector<ActiveElem> &act, std::vector<SElem> &el,
                std::vector<SHead> &hd, size_t &max_len)
{
    bh_node const &nd = t[id];
    if (nd.feature < 0) {
        SHead h{};
        h.first = static_cast<u32>(el.size());
        h.n = static_cast<u16>(act.size());
        h.value = nd.value;
        hd.push_back(h);
        for (auto const &a : act) {
            int lo = a.lo, hi = a.hi;
            if (hi < lo) {
                lo = 1;
                hi = 0;
            }
            SElem e{};
            e.feature = static_cast<u16>(a.feature | (a.missing_ok ? 0x8000u : 0u));
            e.lo = static_cast<u8>(lo);
            e.hi = static_cast<u8>(hi);
            e.z = static_cast<float>(a.zero);
            el.push_back(e);
        }
        max_len = std::max(max_len, act.size());
        return;
    }
    u32 const f = static_cast<u32>(nd.feature);
    int const split = static_cast<int>(nd.bin), lb = static_cast<int>(last[f]);
    double const cover = nd.cover;
    auto frac = [&](u32 c) { return cover > 0.0 ? static_cast<double>(t[c].cover) / cover : 0.0; };
    size_t idx = act.size();
    for (size_t k = 0; k < act.size(); ++k)
        if (act[k].feature == f) {
            idx = k;
            break;
        }
    bool const fresh = idx == act.size();
    if (fresh) act.push_back({f, 0, lb - 1, true, 1.0});
    ActiveElem const saved = act[idx];
    act[idx].hi = std::min(saved.hi, split);
    act[idx].missing_ok = saved.missing_ok && nd.default_left != 0;
    act[idx].zero = saved.zero * frac(nd.left);
    pack_paths(t, last, nd.left, act, el, hd, max_len);
    act[idx].lo = std::max(saved.lo, split + 1);
    act[idx].hi = saved.hi;
    act[idx].missing_ok = saved.missing_ok && nd.default_left == 0;
    act[idx].zero = saved.zero * frac(nd.right);
    pack_paths(t, last, nd.right, act, el, hd, max_len);
    if (fresh) act.pop_back();
    else act[idx] = saved;
}

// tree_expected_value (src/shap.cpp:246-292): cover-weighted mean of the leaves.
double expected_value(bh_node const *t, u32 id)
{
    bh_node const &nd = t[id];
    if (nd.feature < 0) return nd.value;
    double const cover = nd.cover;
    if (cover <= 0.0) return 0.0;
    return (static_cast<double>(t[nd.left].cover) * expected_value(t, nd.left) +
            static_cast<double>(t[nd.right].cover) * expected_value(t, nd.right)) /
           cover;
}

} // namespace

struct bh_predictor {
    int device = 0;
    u32 n_trees = 0, F = 0;
    float lr = 0.0F, init = 0.0F;
    double shap_bias = 0.0;
    u32 n_heads = 0, max_len = 0;
    u32 n_groups = 0;          // TreeSHAP path groups of 64
    double shap_scale = 1.0;   // fixed-point scale of the TreeSHAP accumulators
    PNode *d_nodes = nullptr;
    u32 *d_roots = nullptr;
    u16 *d_last = nullptr;
    SElem *d_elems = nullptr;
    SHead *d_heads = nullptr;
    float *d_w = nullptr;
    hipStream_t stream = nullptr;
};

namespace {

int launch_predict(bh_predictor *p, u8 const *d_bins, u64 n, u64 stride, float *d_out, hipStream_t s)
{
    if (n == 0) return 0;
    u32 const blocks = static_cast<u32>((n + kPR - 1) / kPR);
    size_t const lds = static_cast<size_t>(kPR) * p->F;
    hipLaunchKernelGGL(k_predict, dim3(blocks), dim3(256), lds, s, d_bins, n, stride, p->F, p->d_nodes, p->d_roots,
                       p->n_trees, p->d_last, p->lr, p->init, d_out);
    SHIP(hipGetLastError());
    return 0;
}

int launch_shap(bh_predictor *p, u8 const *d_bins, u64 n, u64 stride, double *d_out, hipStream_t s)
{
    if (n == 0) return 0;
    size_t const lds = static_cast<size_t>(kShapRows) * p->F * (sizeof(unsigned long long) + 1u);
    if (lds > 160u * 1024u) return sfail(1, "TreeSHAP: too many features for the LDS accumulators");
    u32 const blocks = static_cast<u32>((n + kShapRows - 1) / kShapRows);
    // Split the path groups of each row block over `split` workgroups while the row
    // blocks alone leave the chip idle; every workgroup keeps at least one group per wave.
    u32 split = 1;
    if (blocks < kShapFillBlocks)
        split = std::min((kShapFillBlocks + blocks - 1) / blocks, std::max(1u, p->n_groups / 4u));
    u64 const cells = n * (p->F + 1);
    u32 const cgrid = static_cast<u32>(std::min<u64>((cells + 255) / 256, 4096));
    if (split > 1) {
        hipLaunchKernelGGL(k_shap_zero, dim3(cgrid), dim3(256), 0, s, reinterpret_cast<unsigned long long *>(d_out),
                           cells);
        SHIP(hipGetLastError());
    }
    dim3 const grid(blocks, split);
    if (p->max_len <= 8)
        hipLaunchKernelGGL(k_shap<8>, grid, dim3(256), lds, s, d_bins, n, stride, p->F, p->d_elems, p->d_heads,
                           p->n_groups, p->max_len, p->d_w, p->d_last, p->lr, p->shap_bias, p->shap_scale, d_out);
    else
        hipLaunchKernelGGL(k_shap<kMaxPath>, grid, dim3(256), lds, s, d_bins, n, stride, p->F, p->d_elems,
                           p->d_heads, p->n_groups, p->max_len, p->d_w, p->d_last, p->lr, p->shap_bias,
                           p->shap_scale, d_out);
    SHIP(hipGetLastError());
    if (split > 1) {
        hipLaunchKernelGGL(k_shap_to_double, dim3(cgrid), dim3(256), 0, s, d_out, n, p->F, p->lr, p->shap_bias,
                           p->shap_scale);
        SHIP(hipGetLastError());
    }
    return 0;
}

} // namespace

extern "C" {

int bh_predictor_create(int device, const bh_node *nodes, const uint32_t *tree_offsets, uint32_t n_trees,
                        const uint16_t *last_bin, uint32_t n_features, float learning_rate, float init_score,
                        bh_predictor **out)
{
    *out = nullptr;
    if (n_features == 0) return sfail(1, "no features");
    SHIP(hipSetDevice(device));
    auto *p = new bh_predictor();
    p->device = device;
    p->n_trees = n_trees;
    p->F = n_features;
    p->lr = learning_rate;
    p->init = init_score;
    std::vector<PNode> pn;
    std::vector<u32> roots(n_trees);
    std::vector<SElem> el;
    std::vector<SHead> hd;
    size_t max_len = 0;
    double ev_sum = 0.0;
    std::vector<ActiveElem> act;
    for (u32 t = 0; t < n_trees; ++t) {
        bh_node const *tn = nodes + tree_offsets[t];
        u32 const end = t + 1 < n_trees ? tree_offsets[t + 1] : 0;
        // node count: walk reachable ids (covers the whole array for grown trees)
        u32 count = 0;
        std::vector<u32> stack{0};
        while (!stack.empty()) {
            u32 const i = stack.back();
            stack.pop_back();
            count = std::max(count, i + 1);
            if (tn[i].feature >= 0) {
                if (static_cast<u32>(tn[i].feature) >= n_features || tn[i].bin > 255) {
                    delete p;
                    return sfail(1, "tree node out of range (feature id or bin)");
                }
                stack.push_back(tn[i].left);
                stack.push_back(tn[i].right);
            }
        }
        (void)end;
        if (count > 65535) {
            delete p;
            return sfail(1, "tree larger than 65535 nodes");
        }
        roots[t] = static_cast<u32>(pn.size());
        for (u32 i = 0; i < count; ++i) {
            PNode q{};
            if (tn[i].feature < 0) {
                q.feature = 0xFFFFu;
                q.u.value = tn[i].value;
            } else {
                q.feature = static_cast<u16>(tn[i].feature);
                q.bin = static_cast<u8>(tn[i].bin);
                q.dl = static_cast<u8>(tn[i].default_left);
                q.u.c.left = static_cast<u16>(tn[i].left);
                q.u.c.right = static_cast<u16>(tn[i].right);
            }
            pn.push_back(q);
        }
        act.clear();
        pack_paths(tn, last_bin, 0, act, el, hd, max_len);
        ev_sum += expected_value(tn, 0);
    }
    if (max_len > static_cast<size_t>(kMaxPath)) {
        delete p;
        return sfail(1, "TreeSHAP paths longer than 16 features are not supported");
    }
    // bias column: sum of per-tree expected values, times lr, plus init
    // (include/bonsai/booster.hpp:902-925).
    p->shap_bias = ev_sum * static_cast<double>(learning_rate) + static_cast<double>(init_score);
    p->n_heads = static_cast<u32>(hd.size());
    p->max_len = static_cast<u32>(std::max<size_t>(max_len, 1));
    // Interleave the paths in groups of 64: element j of the group's paths contiguous.
    u32 const K = p->max_len;
    p->n_groups = (p->n_heads + 63u) / 64u;
    std::vector<SElem> el2(static_cast<size_t>(p->n_groups) * K * 64u);
    std::vector<SHead> hd2(static_cast<size_t>(p->n_groups) * 64u);
    double leaf_abs = 0.0;
    for (size_t q = 0; q < hd.size(); ++q) {
        size_t const g = q / 64u, l = q % 64u;
        hd2[g * 64u + l] = hd[q];
        hd2[g * 64u + l].first = 0;
        for (u32 j = 0; j < hd[q].n; ++j) el2[(g * K + j) * 64u + l] = el[hd[q].first + j];
        leaf_abs += std::fabs(static_cast<double>(hd[q].value));
    }
    // |one path's contribution| <= |value| * sum_i w_i C(n, i) < 4 |value| for n <= 16,
    // so every accumulator stays below 2^61 in magnitude.
    p->shap_scale = std::ldexp(1.0, 61) / std::max(4.0 * leaf_abs, 1e-300);
    el.swap(el2);
    hd.swap(hd2);
    // Permutation weights w_i = 1 / (n C(n-1, i)) (shap_path_weights, src/shap_paths.cpp:165-194).
    std::vector<float> w(kMaxPath * (kMaxPath + 1) / 2, 0.0F);
    {
        std::vector<double> binom(kMaxPath, 0.0);
        binom[0] = 1.0;
        for (size_t nn = 1; nn <= static_cast<size_t>(kMaxPath); ++nn) {
            size_t const rr = nn - 1;
            if (rr > 0) {
                binom[rr] = 0.0;
                for (size_t i = rr + 1; i-- > 1;) binom[i] += binom[i - 1];
            }
            for (size_t i = 0; i < nn; ++i)
                w[rr * nn / 2 + i] = static_cast<float>(1.0 / (static_cast<double>(nn) * binom[i]));
        }
    }
    hipError_t e = hipSuccess;
    auto al = [&](void **ptr, size_t bytes) {
        if (e == hipSuccess) e = hipMalloc(ptr, std::max<size_t>(bytes, 16));
    };
    al(reinterpret_cast<void **>(&p->d_nodes), pn.size() * sizeof(PNode));
    al(reinterpret_cast<void **>(&p->d_roots), roots.size() * sizeof(u32));
    al(reinterpret_cast<void **>(&p->d_last), n_features * sizeof(u16));
    al(reinterpret_cast<void **>(&p->d_elems), el.size() * sizeof(SElem));
    al(reinterpret_cast<void **>(&p->d_heads), hd.size() * sizeof(SHead));
    al(reinterpret_cast<void **>(&p->d_w), w.size() * sizeof(float));
    if (e == hipSuccess && !pn.empty()) e = hipMemcpy(p->d_nodes, pn.data(), pn.size() * sizeof(PNode), hipMemcpyHostToDevice);
    if (e == hipSuccess && n_trees) e = hipMemcpy(p->d_roots, roots.data(), roots.size() * sizeof(u32), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(p->d_last, last_bin, n_features * sizeof(u16), hipMemcpyHostToDevice);
    if (e == hipSuccess && !el.empty()) e = hipMemcpy(p->d_elems, el.data(), el.size() * sizeof(SElem), hipMemcpyHostToDevice);
    if (e == hipSuccess && !hd.empty()) e = hipMemcpy(p->d_heads, hd.data(), hd.size() * sizeof(SHead), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(p->d_w, w.data(), w.size() * sizeof(float), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        bh_predictor_destroy(p);
        return sfail(-1, std::string("hip: predictor upload: ") + hipGetErrorString(e));
    }
    *out = p;
    return 0;
}

int bh_predictor_destroy(bh_predictor *p)
{
    if (!p) return 0;
    (void)hipSetDevice(p->device);
    void *ptrs[] = {p->d_nodes, p->d_roots, p->d_last, p->d_elems, p->d_heads, p->d_w};
    for (void *q : ptrs)
        if (q) (void)hipFree(q);
    if (p->stream) (void)hipStreamDestroy(p->stream);
    delete p;
    return 0;
}

int bh_predict_device(bh_predictor *p, const uint8_t *d_bins, uint64_t n, uint64_t row_stride, float *d_out,
                      void *hip_stream)
{
    if (!p) return sfail(1, "null predictor");
    if (row_stride < p->F) return sfail(1, "row_stride < n_features");
    return launch_predict(p, d_bins, n, row_stride, d_out, static_cast<hipStream_t>(hip_stream));
}

int bh_shap_device(bh_predictor *p, const uint8_t *d_bins, uint64_t n, uint64_t row_stride, double *d_out,
                   void *hip_stream)
{
    if (!p) return sfail(1, "null predictor");
    if (row_stride < p->F) return sfail(1, "row_stride < n_features");
    return launch_shap(p, d_bins, n, row_stride, d_out, static_cast<hipStream_t>(hip_stream));
}

int bh_predict(bh_predictor *p, const uint8_t *bins, uint64_t n, uint64_t row_stride, float *out)
{
    if (!p) return sfail(1, "null predictor");
    if (row_stride < p->F) return sfail(1, "row_stride < n_features");
    if (n == 0) return 0;
    SHIP(hipSetDevice(p->device));
    u8 *d_bins = nullptr;
    float *d_out = nullptr;
    SHIP(hipMalloc(&d_bins, n * row_stride));
    SHIP(hipMalloc(&d_out, n * sizeof(float)));
    SHIP(hipMemcpy(d_bins, bins, n * row_stride, hipMemcpyHostToDevice));
    int rc = launch_predict(p, d_bins, n, row_stride, d_out, p->stream);
    if (rc == 0) {
        hipError_t e = hipStreamSynchronize(p->stream);
        if (e == hipSuccess) e = hipMemcpy(out, d_out, n * sizeof(float), hipMemcpyDeviceToHost);
        if (e != hipSuccess) rc = sfail(-1, std::string("hip: predict: ") + hipGetErrorString(e));
    }
    (void)hipFree(d_bins);
    (void)hipFree(d_out);
    return rc;
}

int bh_shap(bh_predictor *p, const uint8_t *bins, uint64_t n, uint64_t row_stride, double *out)
{
    if (!p) return sfail(1, "null predictor");
    if (row_stride < p->F) return sfail(1, "row_stride < n_features");
    if (n == 0) return 0;
    SHIP(hipSetDevice(p->device));
    u8 *d_bins = nullptr;
    double *d_out = nullptr;
    SHIP(hipMalloc(&d_bins, n * row_stride));
    SHIP(hipMalloc(&d_out, n * (p->F + 1) * sizeof(double)));
    SHIP(hipMemcpy(d_bins, bins, n * row_stride, hipMemcpyHostToDevice));
    int rc = launch_shap(p, d_bins, n, row_stride, d_out, p->stream);
    if (rc == 0) {
        hipError_t e = hipStreamSynchronize(p->stream);
        if (e == hipSuccess) e = hipMemcpy(out, d_out, n * (p->F + 1) * sizeof(double), hipMemcpyDeviceToHost);
        if (e != hipSuccess) rc = sfail(-1, std::string("hip: shap: ") + hipGetErrorString(e));
    }
    (void)hipFree(d_bins);
    (void)hipFree(d_out);
    return rc;
}

} // extern "C"


// ---------------------------------------------------------------------------------
// Continuous-batching serving loop (cfg4).  Requests arrive on a clock; whenever an
// execution slot (one HIP stream with its pinned staging and device buffers) is free,
// every request that has arrived and still fits the slot's row capacity and the HBM
// budget is packed into one batch: rows H2D, predict (+ TreeSHAP), results D2H, all
// asynchronous on the slot's stream, so batches on different slots overlap on the
// device.  Each row's result depends only on that row (per-row kernels, fixed-order
// reductions), so a request's output is identical alone and in any batch.
// ---------------------------------------------------------------------------------
namespace {

struct Slot {
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;
    hipEvent_t k0 = nullptr, k1 = nullptr; // the batch's kernels: first start, last end
    u8 *h_rows = nullptr, *d_rows = nullptr;
    float *h_pred = nullptr, *d_pred = nullptr;
    double *h_shap = nullptr, *d_shap = nullptr;
    std::vector<u32> reqs; // requests in the batch
    u64 rows = 0;
    u64 bytes = 0;
    bool busy = false;
};

} // namespace

extern "C" int bh_serve(bh_predictor *p, const uint8_t *rows, uint64_t row_stride, const uint64_t *req_offsets,
                        const double *arrival_s, uint32_t n_req, const bh_serve_config *cfg, float *pred_out,
                        double *shap_out, double *latency_ms, bh_serve_stats *stats)
{
    if (!p) return sfail(1, "null predictor");
    if (row_stride < p->F) return sfail(1, "row_stride < n_features");
    if (cfg->n_slots == 0 || cfg->max_batch_rows == 0) return sfail(1, "n_slots and max_batch_rows must be > 0");
    SHIP(hipSetDevice(p->device));
    u32 const F = p->F;
    bool const shap = cfg->do_shap != 0;
    u64 const per_row = row_stride + sizeof(float) + (shap ? (F + 1) * sizeof(double) : 0);
    for (u32 i = 0; i < n_req; ++i)
        if (req_offsets[i + 1] - req_offsets[i] > cfg->max_batch_rows)
            return sfail(1, "a request is larger than max_batch_rows");
    std::vector<Slot> slots(cfg->n_slots);
    u64 const cap = cfg->max_batch_rows;
    hipError_t e = hipSuccess;
    for (auto &s : slots) {
        if (e == hipSuccess) e = hipStreamCreateWithFlags(&s.stream, hipStreamNonBlocking);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&s.done, hipEventDisableTiming);
        if (e == hipSuccess) e = hipEventCreate(&s.k0);
        if (e == hipSuccess) e = hipEventCreate(&s.k1);
        if (e == hipSuccess) e = hipHostMalloc(reinterpret_cast<void **>(&s.h_rows), cap * row_stride, 0);
        if (e == hipSuccess) e = hipHostMalloc(reinterpret_cast<void **>(&s.h_pred), cap * sizeof(float), 0);
        if (e == hipSuccess) e = hipMalloc(&s.d_rows, cap * row_stride);
        if (e == hipSuccess) e = hipMalloc(&s.d_pred, cap * sizeof(float));
        if (shap) {
            if (e == hipSuccess) e = hipHostMalloc(reinterpret_cast<void **>(&s.h_shap), cap * (F + 1) * sizeof(double), 0);
            if (e == hipSuccess) e = hipMalloc(&s.d_shap, cap * (F + 1) * sizeof(double));
        }
    }
    auto cleanup = [&]() {
        for (auto &s : slots) {
            if (s.stream) (void)hipStreamSynchronize(s.stream);
            if (s.h_rows) (void)hipHostFree(s.h_rows);
            if (s.h_pred) (void)hipHostFree(s.h_pred);
            if (s.h_shap) (void)hipHostFree(s.h_shap);
            if (s.d_rows) (void)hipFree(s.d_rows);
            if (s.d_pred) (void)hipFree(s.d_pred);
            if (s.d_shap) (void)hipFree(s.d_shap);
            if (s.done) (void)hipEventDestroy(s.done);
            if (s.k0) (void)hipEventDestroy(s.k0);
            if (s.k1) (void)hipEventDestroy(s.k1);
            if (s.stream) (void)hipStreamDestroy(s.stream);
        }
    };
    if (e != hipSuccess) {
        cleanup();
        return sfail(-1, std::string("hip: serve buffers: ") + hipGetErrorString(e));
    }
    // Device-time intervals of every batch's kernels, from a base event on the first
    // slot's stream: they show whether batches of different slots ran on the device at
    // the same time.
    hipEvent_t base = nullptr;
    e = hipEventCreate(&base);
    if (e == hipSuccess) e = hipEventRecord(base, slots[0].stream);
    if (e == hipSuccess) e = hipEventSynchronize(base);
    if (e != hipSuccess) {
        if (base) (void)hipEventDestroy(base);
        cleanup();
        return sfail(-1, std::string("hip: serve events: ") + hipGetErrorString(e));
    }
    std::vector<std::pair<double, double>> kspans;
    using clk = std::chrono::steady_clock;
    auto const t0 = clk::now();
    auto now_s = [&]() { return std::chrono::duration<double>(clk::now() - t0).count(); };
    u32 next = 0, completed = 0, inflight = 0, max_inflight = 0;
    u64 inflight_bytes = 0, total_rows = 0;
    int rc = 0;
    while (completed < n_req && rc == 0) {
        bool progressed = false;
        // completions
        for (auto &s : slots) {
            if (!s.busy) continue;
            hipError_t const q = hipEventQuery(s.done);
            if (q == hipErrorNotReady) continue;
            if (q != hipSuccess) {
                rc = sfail(-1, std::string("hip: serve batch: ") + hipGetErrorString(q));
                break;
            }
            double const t_done = now_s();
            float a_ms = 0.0F, b_ms = 0.0F;
            if (hipEventElapsedTime(&a_ms, base, s.k0) == hipSuccess &&
                hipEventElapsedTime(&b_ms, base, s.k1) == hipSuccess)
                kspans.emplace_back(a_ms, b_ms);
            u64 pos = 0;
            for (u32 r : s.reqs) {
                u64 const a = req_offsets[r], nr = req_offsets[r + 1] - a;
                std::memcpy(pred_out + a, s.h_pred + pos, nr * sizeof(float));
                if (shap && shap_out)
                    std::memcpy(shap_out + a * (F + 1), s.h_shap + pos * (F + 1), nr * (F + 1) * sizeof(double));
                latency_ms[r] = (t_done - arrival_s[r]) * 1000.0;
                pos += nr;
                ++completed;
            }
            inflight -= static_cast<u32>(s.reqs.size());
            inflight_bytes -= s.bytes;
            s.busy = false;
            s.reqs.clear();
            progressed = true;
        }
        if (rc) break;
        // admissions: fill free slots with arrived requests
        double const t = now_s();
        for (auto &s : slots) {
            if (s.busy || next >= n_req || arrival_s[next] > t) continue;
            u64 rows_in = 0, bytes_in = 0;
            while (next < n_req && arrival_s[next] <= t) {
                u64 const nr = req_offsets[next + 1] - req_offsets[next];
                u64 const nb = nr * per_row;
                bool const first = s.reqs.empty();
                if (!first && (rows_in + nr > cap || inflight_bytes + bytes_in + nb > cfg->hbm_budget)) break;
                if (first && inflight > 0 && inflight_bytes + nb > cfg->hbm_budget) break; // wait for memory
                std::memcpy(s.h_rows + rows_in * row_stride, rows + req_offsets[next] * row_stride, nr * row_stride);
                s.reqs.push_back(next);
                rows_in += nr;
                bytes_in += nb;
                ++next;
            }
            if (s.reqs.empty()) continue;
            s.rows = rows_in;
            s.bytes = bytes_in;
            inflight += static_cast<u32>(s.reqs.size());
            inflight_bytes += bytes_in;
            max_inflight = std::max(max_inflight, inflight);
            total_rows += rows_in;
            e = hipMemcpyAsync(s.d_rows, s.h_rows, rows_in * row_stride, hipMemcpyHostToDevice, s.stream);
            if (e == hipSuccess) e = hipEventRecord(s.k0, s.stream);
            if (e == hipSuccess) {
                rc = launch_predict(p, s.d_rows, rows_in, row_stride, s.d_pred, s.stream);
                if (rc == 0 && shap) rc = launch_shap(p, s.d_rows, rows_in, row_stride, s.d_shap, s.stream);
            }
            if (rc) break;
            if (e == hipSuccess) e = hipEventRecord(s.k1, s.stream);
            if (e == hipSuccess) e = hipMemcpyAsync(s.h_pred, s.d_pred, rows_in * sizeof(float), hipMemcpyDeviceToHost, s.stream);
            if (e == hipSuccess && shap)
                e = hipMemcpyAsync(s.h_shap, s.d_shap, rows_in * (F + 1) * sizeof(double), hipMemcpyDeviceToHost, s.stream);
            if (e == hipSuccess) e = hipEventRecord(s.done, s.stream);
            if (e != hipSuccess) {
                rc = sfail(-1, std::string("hip: serve enqueue: ") + hipGetErrorString(e));
                break;
            }
            s.busy = true;
            progressed = true;
        }
        if (!progressed) std::this_thread::yield();
    }
    double const wall = now_s();
    cleanup();
    (void)hipEventDestroy(base);
    if (rc) return rc;
    // Sweep the kernel spans: time with two or more batches on the device, and the most
    // at once.
    std::vector<std::pair<double, int>> ev;
    for (auto const &k : kspans) {
        ev.emplace_back(k.first, +1);
        ev.emplace_back(k.second, -1);
    }
    std::sort(ev.begin(), ev.end(), [](auto const &x, auto const &y) {
        return x.first < y.first || (x.first == y.first && x.second < y.second);
    });
    int live = 0, peak = 0;
    double overlap = 0.0, prev = 0.0;
    for (auto const &v : ev) {
        if (live >= 2) overlap += v.first - prev;
        live += v.second;
        peak = std::max(peak, live);
        prev = v.first;
    }
    std::vector<double> lat(latency_ms, latency_ms + n_req);
    std::sort(lat.begin(), lat.end());
    auto pct = [&](double q) {
        if (lat.empty()) return 0.0;
        size_t const k = std::min(lat.size() - 1, static_cast<size_t>(std::ceil(q * lat.size())) - 1);
        return lat[k];
    };
    double mean = 0.0;
    for (double v : lat) mean += v;
    stats->wall_s = wall;
    stats->rows = total_rows;
    stats->requests = n_req;
    stats->rows_per_s = wall > 0 ? static_cast<double>(total_rows) / wall : 0.0;
    stats->p50_ms = pct(0.50);
    stats->p99_ms = pct(0.99);
    stats->mean_ms = n_req ? mean / n_req : 0.0;
    stats->max_inflight = max_inflight;
    stats->device_overlap_ms = overlap;
    stats->max_device_concurrency = static_cast<uint32_t>(peak);
    return 0;
}
