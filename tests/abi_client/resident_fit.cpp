// A C++ caller of the C ABI (include/bonsai_hip.h), compiled with g++ and linked to
// libbonsai_hip.so, shaped like the reference's grower seam: the growers' resident_begin /
// grow / resident_end / resident (include/bonsai/grower.hpp:254-306) as the booster drives
// them (try_resident_round / resident_round, include/bonsai/booster.hpp:630-676: on a
// resident round grow() gets empty grad / hess views and the sampler's rows), and grow()
// with host gradients for the non-resident path.
//
// TEST INFRASTRUCTURE: it checks every tree against the CPU oracle (liboracle.so) and
// exits 0 when all agree, 1 on a mismatch, 77 without a HIP device.
//
//   resident_fit [depthwise|leafwise]
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "bonsai_hip.h"
#include "oracle.h"

namespace {

// The binned training set the grower sees (the reference's Dataset, reduced to what
// the seam needs here: a row-major u8 plane, bins per feature, labels).
struct Plane {
    uint64_t n_rows = 0;
    uint32_t n_features = 0;
    std::vector<uint8_t> bins;
    std::vector<uint16_t> n_bins;
    std::vector<float> labels;
};

struct Tree {
    std::vector<bh_node> nodes;
};

void check(int rc, char const *what)
{
    if (rc != 0) throw std::runtime_error(std::string(what) + ": " + bh_last_error());
}

// The adapter a maintainer would put behind the reference's TreeGrower seam for the
// hip_depthwise / hip_leafwise registry names: one device context owned by value.
class HipGrower {
  public:
    HipGrower(int grower, int max_depth, int max_leaves, bool quantized)
    {
        cfg_.grower = grower;
        cfg_.objective = BH_LOGLOSS;
        cfg_.quantized = quantized ? 1 : 0;
        cfg_.use_graph = 1;
        cfg_.max_depth = max_depth;
        cfg_.min_data_in_leaf = 20;
        cfg_.max_leaves = max_leaves;
        cfg_.min_child_hess = 1.0F;
        cfg_.lambda_l2 = 1.0F;
        cfg_.learning_rate = 0.1F;
        cfg_.seed = 4242;
    }
    ~HipGrower()
    {
        if (ctx_) bh_destroy(ctx_);
    }
    HipGrower(HipGrower const &) = delete;
    HipGrower &operator=(HipGrower const &) = delete;

    // resident_begin(Dataset const&, DeviceObjectiveKind, std::span<float const> scores,
    // float lr) -> bool: labels and scores go to the device once per fit.
    bool resident_begin(Plane const &ds, std::vector<float> const &scores, float learning_rate)
    {
        arm(ds);
        cfg_.learning_rate = learning_rate;
        check(bh_configure(ctx_, &cfg_), "bh_configure");
        configured_ = true;
        subset_all_ = true;
        check(bh_upload_labels(ctx_, ds.labels.data()), "bh_upload_labels");
        check(bh_set_scores(ctx_, scores.data()), "bh_set_scores");
        resident_ = true;
        trees_ = 0;
        return true;
    }
    // resident_end(std::span<float> scores): the host copy becomes authoritative again.
    void resident_end(std::vector<float> &scores)
    {
        check(bh_get_scores(ctx_, scores.data()), "bh_get_scores");
        resident_ = false;
    }
    bool resident() const { return resident_; }

    // grow(ds, grad, hess, rows).  Resident rounds pass empty gradient views: the device
    // derives them, grows the tree and adds it to the scores (one hipGraph launch).
    Tree grow(Plane const &ds, std::vector<float> const &grad, std::vector<float> const &hess,
              std::vector<uint32_t> const &rows)
    {
        if (!ctx_) arm(ds);
        if (!resident_ && !configured_) {
            check(bh_configure(ctx_, &cfg_), "bh_configure");
            configured_ = true;
        }
        // The sampler's rows: every row, or a subset the device grows the tree on (its
        // scores still come from the finished tree).  Set only when they change.
        bool const all_rows = rows.size() == ds.n_rows;
        if (!all_rows || !subset_all_) {
            check(bh_set_row_subset(ctx_, all_rows ? nullptr : rows.data(), all_rows ? 0 : rows.size()),
                  "bh_set_row_subset");
            subset_all_ = all_rows;
        }
        uint32_t max_nodes = 0;
        check(bh_max_nodes(ctx_, &max_nodes), "bh_max_nodes");
        Tree t;
        if (grad.empty()) {
            if (!resident_) throw std::runtime_error("empty gradients outside a resident fit");
            check(bh_train(ctx_, 1), "bh_train");
            ++trees_;
            std::vector<bh_node> all(static_cast<size_t>(trees_) * max_nodes);
            std::vector<uint32_t> sizes(trees_);
            uint32_t n = 0;
            check(bh_get_trees(ctx_, all.data(), max_nodes, sizes.data(), trees_, &n), "bh_get_trees");
            size_t const base = static_cast<size_t>(n - 1) * max_nodes;
            t.nodes.assign(all.begin() + base, all.begin() + base + sizes[n - 1]);
            return t;
        }
        check(bh_set_gradients(ctx_, grad.data(), hess.data(), 0, nullptr, nullptr, nullptr), "bh_set_gradients");
        t.nodes.resize(max_nodes);
        uint32_t n = 0;
        check(bh_grow_tree(ctx_, t.nodes.data(), max_nodes, &n, nullptr), "bh_grow_tree");
        t.nodes.resize(n);
        return t;
    }

    // Debug seam of this build: the gradients the device used for the last tree.
    void gradients(std::vector<int8_t> &gq, std::vector<int16_t> &hq, double scales[2])
    {
        check(bh_get_gradients(ctx_, nullptr, nullptr, gq.data(), hq.data(), scales), "bh_get_gradients");
    }
    bh_fit_config const &config() const { return cfg_; }

  private:
    void arm(Plane const &ds)
    {
        if (ctx_) return;
        check(bh_create(0, &ctx_), "bh_create");
        check(bh_upload_bins(ctx_, ds.bins.data(), ds.n_rows, ds.n_features, ds.n_features, ds.n_bins.data()),
              "bh_upload_bins");
    }

    bh_fit_config cfg_{};
    bh_context *ctx_ = nullptr;
    bool resident_ = false, configured_ = false, subset_all_ = true;
    uint32_t trees_ = 0;
};

// Deterministic synthetic plane: 40,000 rows x 12 features, 31 value bins plus the
// missing bin (5% missing), labels from a rule on three features plus noise.
Plane make_plane()
{
    Plane p;
    p.n_rows = 40000;
    p.n_features = 12;
    p.n_bins.assign(p.n_features, 32);
    p.bins.resize(p.n_rows * p.n_features);
    p.labels.resize(p.n_rows);
    uint64_t x = 88172645463325252ULL;
    auto next = [&]() {
        x ^= x << 13;
        x ^= x >> 7;
        x ^= x << 17;
        return x;
    };
    for (uint64_t r = 0; r < p.n_rows; ++r) {
        for (uint32_t f = 0; f < p.n_features; ++f) {
            uint64_t const v = next();
            p.bins[r * p.n_features + f] = (v % 100) < 5 ? 31 : static_cast<uint8_t>((v >> 8) % 31);
        }
        uint8_t const *b = &p.bins[r * p.n_features];
        double const z = 0.15 * (b[0] % 31) - 2.0 + (b[3] < 10 ? 1.0 : -0.5) + 0.05 * (b[7] % 31) +
                         ((next() % 1000) / 1000.0 - 0.5);
        p.labels[r] = z > 0.0 ? 1.0F : 0.0F;
    }
    return p;
}

// Leaf value of each row under a bin-space tree (routes_left: the missing bin follows
// default_left, otherwise bin <= split bin goes left).
float leaf_value(std::vector<bh_node> const &t, Plane const &ds, uint64_t r)
{
    uint32_t i = 0;
    while (t[i].feature >= 0) {
        uint32_t const f = static_cast<uint32_t>(t[i].feature);
        uint32_t const b = ds.bins[r * ds.n_features + f];
        bool const left = b == ds.n_bins[f] - 1u ? t[i].default_left != 0 : b <= t[i].bin;
        i = left ? t[i].left : t[i].right;
    }
    return t[i].value;
}

bool same_tree(std::vector<bh_node> const &a, std::vector<orc_node> const &b, std::string &why)
{
    if (a.size() != b.size()) {
        why = "sizes " + std::to_string(a.size()) + " vs " + std::to_string(b.size());
        return false;
    }
    for (size_t i = 0; i < a.size(); ++i) {
        bool const leaf = b[i].feature < 0;
        bool ok = a[i].feature == b[i].feature && std::memcmp(&a[i].value, &b[i].value, 4) == 0 &&
                  std::memcmp(&a[i].cover, &b[i].cover, 4) == 0 && std::memcmp(&a[i].gain, &b[i].gain, 4) == 0;
        if (!leaf)
            ok = ok && a[i].bin == b[i].bin && a[i].default_left == b[i].default_left && a[i].left == b[i].left &&
                 a[i].right == b[i].right;
        if (!ok) {
            why = "node " + std::to_string(i);
            return false;
        }
    }
    return true;
}

std::vector<orc_node> oracle_tree(Plane const &ds, bh_fit_config const &c, std::vector<int8_t> const &gq,
                                  std::vector<int16_t> const &hq, double const sc[2], std::vector<float> const *g,
                                  std::vector<float> const *h)
{
    orc_tree_cfg oc{};
    oc.min_child_hess = c.min_child_hess;
    oc.min_gain_to_split = c.min_gain_to_split;
    oc.lambda_l2 = c.lambda_l2;
    oc.lambda_l1 = c.lambda_l1;
    oc.max_depth = c.max_depth;
    oc.min_data_in_leaf = c.min_data_in_leaf;
    oc.max_leaves = c.max_leaves;
    std::vector<orc_node> out(1u << 14);
    int32_t n;
    bool const q = g == nullptr;
    auto fn = c.grower == BH_LEAFWISE ? orc_grow_leafwise : orc_grow_depthwise;
    // FP32 leafwise histograms sum over leaf-sized units (at most 256 per built node),
    // depthwise ones over 8,192-row units: the oracle restates both orders.
    uint32_t const chunk = c.grower == BH_LEAFWISE ? (0x80000000u | 256u) : 8192u;
    n = fn(ds.bins.data(), ds.n_features, ds.n_rows, ds.n_features, ds.n_bins.data(), q ? nullptr : g->data(),
           q ? nullptr : h->data(), q ? gq.data() : nullptr, q ? hq.data() : nullptr, q ? sc[0] : 0.0,
           q ? sc[1] : 0.0, &oc, q ? ORC_HIST_INT : ORC_HIST_F32_FIXED, chunk, out.data(),
           static_cast<int32_t>(out.size()), nullptr);
    if (n < 0) throw std::runtime_error("oracle grower failed");
    out.resize(static_cast<size_t>(n));
    return out;
}

int run(int grower)
{
    Plane const ds = make_plane();
    int const depth = grower == BH_LEAFWISE ? 10 : 6;
    HipGrower g(grower, depth, 31, true);
    // init_score (LogLossObjective::init_score, src/objective.cpp:122-130) by the oracle
    float const init = orc_init_score(1, ds.labels.data(), ds.n_rows);
    std::vector<float> scores(ds.n_rows, init);
    std::vector<uint32_t> rows(ds.n_rows);
    for (uint64_t r = 0; r < ds.n_rows; ++r) rows[r] = static_cast<uint32_t>(r);

    // Resident fit: five rounds, each tree checked against the oracle grown from the
    // gradients the device used; the scores follow s + lr * leaf on the host.
    if (!g.resident_begin(ds, scores, 0.1F)) return 1;
    std::vector<int8_t> gq(ds.n_rows);
    std::vector<int16_t> hq(ds.n_rows);
    std::vector<float> host = scores;
    for (int round = 0; round < 5; ++round) {
        Tree const t = g.grow(ds, {}, {}, rows);
        double sc[2];
        g.gradients(gq, hq, sc);
        std::vector<orc_node> const ref = oracle_tree(ds, g.config(), gq, hq, sc, nullptr, nullptr);
        std::string why;
        if (!same_tree(t.nodes, ref, why)) {
            std::printf("resident round %d: device tree differs from the oracle's (%s)\n", round, why.c_str());
            return 1;
        }
        for (uint64_t r = 0; r < ds.n_rows; ++r) host[r] = host[r] + 0.1F * leaf_value(t.nodes, ds, r);
    }
    std::vector<float> dev(ds.n_rows);
    g.resident_end(dev);
    if (g.resident() || std::memcmp(dev.data(), host.data(), dev.size() * sizeof(float)) != 0) {
        std::printf("resident scores differ from the host replay\n");
        return 1;
    }

    // Non-resident grow with host gradients over a row subset (every third row): the
    // quantizer's output is read back and the oracle grows from the same integers on the
    // same rows (restricted plane).
    std::vector<float> gr(ds.n_rows), hs(ds.n_rows);
    orc_logloss_grad(dev.data(), ds.labels.data(), ds.n_rows, gr.data(), hs.data());
    std::vector<uint32_t> sub;
    for (uint64_t r = 0; r < ds.n_rows; r += 3) sub.push_back(static_cast<uint32_t>(r));
    Tree const t = g.grow(ds, gr, hs, sub);
    double sc[2];
    g.gradients(gq, hq, sc);
    Plane part;
    part.n_rows = sub.size();
    part.n_features = ds.n_features;
    part.n_bins = ds.n_bins;
    std::vector<int8_t> pg;
    std::vector<int16_t> ph;
    for (uint32_t r : sub) {
        part.bins.insert(part.bins.end(), ds.bins.begin() + r * ds.n_features,
                         ds.bins.begin() + (r + 1) * ds.n_features);
        pg.push_back(gq[r]);
        ph.push_back(hq[r]);
    }
    std::vector<orc_node> const ref = oracle_tree(part, g.config(), pg, ph, sc, nullptr, nullptr);
    std::string why;
    if (!same_tree(t.nodes, ref, why)) {
        std::printf("host-gradient grow on a row subset differs from the oracle's (%s)\n", why.c_str());
        return 1;
    }
    std::printf("resident_fit %s: 5 resident rounds and a row-subset grow match the oracle, scores bit-equal\n",
                grower == BH_LEAFWISE ? "leafwise" : "depthwise");
    return 0;
}

} // namespace

int main(int argc, char **argv)
{
    int n = 0;
    if (bh_device_count(&n) != 0 || n == 0) {
        std::printf("resident_fit: no HIP device\n");
        return 77;
    }
    int const grower = argc > 1 && std::strcmp(argv[1], "leafwise") == 0 ? BH_LEAFWISE : BH_DEPTHWISE;
    try {
        return run(grower);
    } catch (std::exception const &e) {
        std::printf("resident_fit: %s\n", e.what());
        return 1;
    }
}
