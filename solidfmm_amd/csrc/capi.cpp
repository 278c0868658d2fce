// C ABI of the translation engine (include/solidfmm_gpu.h): contexts,
// validated plans, launches, graphs, the autotuner and the host-pool
// convenience path.
//
// Validation follows the reference's checked entry points
// (translations.cpp:634-1001): every item is checked before any work is
// enqueued, so a failing call writes nothing.  Targets that repeat within a
// list are applied in list order: item k of a target goes to "round" k, and
// rounds run one after another on the stream, each round a set of launches
// whose targets are pairwise distinct.
//
// Within a round, items are bucketed by their larger order pb =
// max(p_in, p_out) and the kernel family chosen for (op, pb): the forced
// family of the context, else the autotuned one, else the default.  A
// bucket whose items all have p_in = p_out = pb runs the uniform-order
// kernel; any other bucket runs the per-item-order (RT) variant of the
// kernel of order pb.  Items of a bucket are sorted by (p_in, p_out), so a
// wave mostly holds one order pair and skips the degrees it does not need.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/solidfmm_gpu.h"
#include "mfma_tables.hpp"
#include "translate.hpp"

namespace {
constexpr int kOrders = sfmm::mfma::kMaxOrder + 1;
}

struct sfmm_context {
    int device = 0;
    size_t P = 0;
    int precision = SFMM_FP64;
    int family = SFMM_KERNELS_AUTO;
    int tune[3][kOrders] = {};  // autotuned family per (op, pb); 0 = default
    std::string err;
    std::mutex mu;
    // staging of the blocking host-pool call (sfmm_translate_host)
    std::mutex host_mu;
    sfmm_staging *host_staging = nullptr;
    // streams forked from the caller's stream to run the launches of one
    // round side by side (sfmm_plan_execute), and their fork/join events;
    // created on first use
    std::vector<hipStream_t> branch_streams;
    std::vector<hipEvent_t> branch_events;
};

struct sfmm_plan {
    struct Launch {
        sfmm::KernelChoice k;
        size_t offset, count;
        uint32_t round = 0;  // launches of one round write pairwise distinct targets
    };
    sfmm_context *ctx = nullptr;
    int op = 0;
    size_t n_items = 0;
    size_t src_extent = 0, dst_extent = 0;
    std::vector<Launch> launches;
    // Device memory, one block: items | reduction targets | reduction items
    // | scratch.  The first three are uploaded from `blob`.
    bool scratch = false;  // targets repeat: per-item scratch + reduction
    size_t n_targets = 0;
    size_t off_targets = 0, off_red_items = 0, off_scratch = 0, dev_bytes = 0;
    std::vector<unsigned char> blob;
    unsigned char *d_mem = nullptr;
    bool owns_mem = true;  // false: caller memory (sfmm_plan_bind)
    // captured graph (sfmm_plan_capture)
    hipGraphExec_t exec = nullptr;

    const sfmm_item *d_items() const { return reinterpret_cast<const sfmm_item *>(d_mem); }
};

namespace {

int fail(sfmm_context *ctx, int code, const std::string &msg)
{
    if (ctx) ctx->err = msg;
    return code;
}

int hip_fail(sfmm_context *ctx, hipError_t e, const char *what)
{
    return fail(ctx, SFMM_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

const char *op_name(int op) { return op == SFMM_M2M ? "m2m" : op == SFMM_M2L ? "m2l" : "l2l"; }

int family_for(const sfmm_context *ctx, int op, uint32_t pb)
{
    if (ctx->family != SFMM_KERNELS_AUTO) return ctx->family;
    const int t = ctx->tune[op][pb];
    return t ? t : sfmm::default_family(pb);
}

// Kernel for one item on a context: family and whether it has one.  FP64
// contexts: the family of its bucket max(p_in, p_out).  Reduced-precision
// contexts, square M2L: the precision lane kernel up to kPrecLaneAutoMax
// (FP32; kPrecLaneMaxDW for double-word FP32), above it the FP32 MFMA
// kernel; everything else, and any family the context forces that has no
// kernel for the item, the precision generic kernel.
bool choose(const sfmm_context *ctx, int op, uint32_t p_in, uint32_t p_out, int &family)
{
    const uint32_t pb = std::max(p_in, p_out);
    if (ctx->precision == SFMM_FP64) {
        if (!sfmm::have_kernel(op, p_in, p_out, ctx->family)) return false;
        family = family_for(ctx, op, pb);
        return true;
    }
    if (pb > sfmm::kPrecMaxOrder) return false;
    const bool square_m2l = op == SFMM_M2L && p_in == p_out;
    const uint32_t lane_max = ctx->precision == SFMM_SPLIT32 ? sfmm::kPrecLaneMaxDW : sfmm::kPrecLaneMax;
    const bool lane_ok = square_m2l && pb <= lane_max;
    const bool mfma_ok = square_m2l && ctx->precision == SFMM_FP32;  // pb <= kPrecMaxOrder above
    switch (ctx->family) {
    case SFMM_KERNELS_LANE:
        family = sfmm::kLaneKernels;
        return lane_ok;
    case SFMM_KERNELS_MFMA:
        family = sfmm::kMfmaKernels;
        return mfma_ok;
    case SFMM_KERNELS_GENERIC: family = sfmm::kGenericKernel; return true;
    default:
        if (lane_ok && (ctx->precision != SFMM_FP32 || pb <= sfmm::kPrecLaneAutoMax)) family = sfmm::kLaneKernels;
        else if (mfma_ok) family = sfmm::kMfmaKernels;
        else family = sfmm::kGenericKernel;
        return true;
    }
}

// Targets repeated more often than this are accumulated through scratch
// blocks and a segmented reduction instead of one launch per repeat.
constexpr uint32_t kMaxRounds = 4;

size_t align_up(size_t v, size_t a) { return (v + a - 1) / a * a; }

// Reals a translation writes into its (scratch) target block.
uint32_t written_reals(int op, const sfmm_item &it)
{
    const uint32_t pend = std::min(it.p_in, it.p_out);
    const bool zero = it.x == 0.0 && it.y == 0.0 && it.z == 0.0;
    if (op == SFMM_L2L || (op == SFMM_M2M && zero)) return pend * (pend + 1);
    return it.p_out * (it.p_out + 1);
}

// Open-addressing map from 64-bit keys (never ~0) to 32-bit values that
// grows with its contents; plan building looks up every item's target in
// it, and the distinct targets of a list are usually far fewer than its
// items, so the table stays cache-resident.
class FlatMap {
public:
    // value of k, inserting `v` first if k is new
    uint32_t &get(uint64_t k, uint32_t v, bool &inserted)
    {
        if (2 * (size_ + 1) > keys_.size()) grow();
        size_t i = slot(k);
        while (keys_[i] != kEmpty && keys_[i] != k) i = (i + 1) & mask_;
        inserted = keys_[i] == kEmpty;
        if (inserted) {
            keys_[i] = k;
            vals_[i] = v;
            ++size_;
        }
        return vals_[i];
    }
    size_t size() const { return size_; }

private:
    static constexpr uint64_t kEmpty = ~uint64_t(0);
    size_t slot(uint64_t k) const { return size_t((k * 0x9E3779B97F4A7C15ull) >> shift_) & mask_; }
    void grow()
    {
        const size_t cap = keys_.empty() ? 1024 : keys_.size() * 2;
        std::vector<uint64_t> ok(cap, kEmpty);
        std::vector<uint32_t> ov(cap);
        ok.swap(keys_);
        ov.swap(vals_);
        mask_ = cap - 1;
        shift_ = 64 - __builtin_ctzll(cap);
        size_ = 0;
        for (size_t i = 0; i < ok.size(); ++i)
            if (ok[i] != kEmpty) {
                size_t j = slot(ok[i]);
                while (keys_[j] != kEmpty) j = (j + 1) & mask_;
                keys_[j] = ok[i];
                vals_[j] = ov[i];
                ++size_;
            }
    }
    std::vector<uint64_t> keys_;
    std::vector<uint32_t> vals_;
    size_t mask_ = 0, size_ = 0;
    int shift_ = 64;
};

// Validation and grouping into `plan` (host side only).  Items with an
// order of zero contribute nothing (the reference pads such lanes with
// zeros and writes them to its trash buffer, microkernel_generic.cpp:
// 340-380) and are dropped.  Targets that repeat at most kMaxRounds times
// are applied in list order by rounds (the k-th item of a target runs in
// launch round k); otherwise every item writes its own scratch block and a
// segmented reduction adds a target's blocks in list order (reduce.hip).
// Square FP64 M2L orders 21 and 22 on the MFMA kernels run faster through
// the per-item-order variant, whose output degrees ascend with fewer
// registers live (2e5 translations: the fixed-order kernel is 6 % slower at
// 22, and 1-2 % faster at 18 and 20, where this choice once saved 8-14 %
// before the swaps recomputed their indices).
template <class K>
bool mfma_rt_square(const K &key)
{
    return key.family == sfmm::kMfmaKernels && key.pb >= 21 && key.pb <= 22;
}

int build_plan(sfmm_context *ctx, int op, const sfmm_item *items, size_t n, sfmm_plan &plan)
{
    if (op < SFMM_M2M || op > SFMM_L2L) return fail(ctx, SFMM_ERR_LOGIC, "solidfmm: unknown operator");
    if (n && !items) return fail(ctx, SFMM_ERR_LOGIC, "solidfmm: null item list");
    struct Key {
        uint32_t round;
        int family;
        uint32_t pb, p_in, p_out;
        size_t index;
    };
    std::vector<Key> keys;
    keys.reserve(n);
    FlatMap seen;  // target -> occurrences so far
    uint32_t rounds = 0;
    // kernel family per (p_in, p_out), filled on first use: lists repeat a
    // handful of order pairs millions of times
    constexpr int kNoKernel = -2, kUnknown = -1;
    std::vector<int8_t> fam_of(size_t(ctx->P + 1) * (ctx->P + 1), int8_t(kUnknown));
    for (size_t i = 0; i < n; ++i) {
        const sfmm_item &it = items[i];
        if (it.p_in > ctx->P || it.p_out > ctx->P)
            return fail(ctx, SFMM_ERR_RANGE,
                        std::string("solidfmm::") + op_name(op) + "(): orders exceeding operator data.");
        if ((it.src | it.dst) & 1u)
            return fail(ctx, SFMM_ERR_LOGIC, "solidfmm: expansion offsets must be even (16-byte aligned)");
        if (it.p_in == 0 || it.p_out == 0) continue;
        int8_t &fc = fam_of[size_t(it.p_in) * (ctx->P + 1) + it.p_out];
        if (fc == kUnknown) {
            int f = 0;
            fc = int8_t(choose(ctx, op, it.p_in, it.p_out, f) ? f : kNoKernel);
        }
        if (fc == kNoKernel)
            return fail(ctx, SFMM_ERR_UNSUPPORTED,
                        std::string("solidfmm: no kernel for ") + op_name(op) + " with p_in=" +
                            std::to_string(it.p_in) + ", p_out=" + std::to_string(it.p_out));
        const int fam = fc;
        const uint32_t pb = std::max(it.p_in, it.p_out);
        bool fresh;
        const uint32_t round = seen.get(it.dst, 0, fresh)++;
        rounds = std::max(rounds, round + 1);
        plan.src_extent = std::max<size_t>(plan.src_extent, it.src + size_t(it.p_in) * (it.p_in + 1));
        plan.dst_extent = std::max<size_t>(plan.dst_extent, it.dst + size_t(it.p_out) * (it.p_out + 1));
        // the generic kernel takes every order in one launch
        keys.push_back({round, fam, fam == sfmm::kGenericKernel ? 0u : pb, it.p_in, it.p_out, i});
    }
    plan.scratch = rounds > kMaxRounds;
    if (plan.scratch)
        for (Key &k : keys) k.round = 0;
    // keys are in list order here: remember each item's position in it
    std::vector<uint32_t> list_pos;
    if (plan.scratch) {
        list_pos.resize(n);
        for (size_t k = 0; k < keys.size(); ++k) list_pos[keys[k].index] = uint32_t(k);
    }
    // Stable order by (round, family, pb, p_in, p_out): the packed key of
    // each item, its rank among the (few) distinct keys, then a counting
    // sort by rank.
    {
        auto pack = [](const Key &k) {
            return (uint64_t(k.round) << 40) | (uint64_t(k.family) << 32) | (uint64_t(k.pb) << 16) |
                   (uint64_t(k.p_in) << 8) | uint64_t(k.p_out);
        };
        FlatMap bucket;
        std::vector<uint64_t> distinct;
        std::vector<uint32_t> of(keys.size());
        for (size_t k = 0; k < keys.size(); ++k) {
            bool fresh;
            const uint64_t ck = pack(keys[k]);
            of[k] = bucket.get(ck, uint32_t(distinct.size()), fresh);
            if (fresh) distinct.push_back(ck);
        }
        std::vector<uint32_t> order(distinct.size());
        for (uint32_t b = 0; b < order.size(); ++b) order[b] = b;
        std::sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return distinct[a] < distinct[b]; });
        std::vector<size_t> start(distinct.size() + 1, 0);
        std::vector<uint32_t> rank(distinct.size());
        for (uint32_t r = 0; r < order.size(); ++r) rank[order[r]] = r;
        for (size_t k = 0; k < keys.size(); ++k) start[rank[of[k]] + 1]++;
        for (size_t r = 0; r < distinct.size(); ++r) start[r + 1] += start[r];
        std::vector<Key> sorted(keys.size());
        for (size_t k = 0; k < keys.size(); ++k) sorted[start[rank[of[k]]]++] = keys[k];
        keys.swap(sorted);
    }
    const size_t m = keys.size();
    plan.n_items = m;
    plan.launches.clear();
    std::vector<sfmm_item> ordered;
    ordered.reserve(m);
    std::vector<uint64_t> scr_off;  // scratch offset per list position
    std::vector<uint8_t> is_head;   // per list position: the item's block enters the reduction
    size_t scratch_reals = 0;
    if (plan.scratch) {
        scr_off.resize(m);
        is_head.assign(m, 1);
    }
    // FP64 lane and MFMA kernels in scratch mode share one scratch block
    // among the items of a run: consecutive items of a launch, within one
    // 64-item wave (lane kernels) or 16-item group (MFMA kernels), with the
    // same target and orders and a nonzero shift (the kernels add a run's
    // results across lanes).
    const bool runs = plan.scratch && ctx->precision == SFMM_FP64;
    auto zero = [](const sfmm_item &it) { return it.x == 0.0 && it.y == 0.0 && it.z == 0.0; };
    const sfmm_item *prev = nullptr;
    uint64_t run_off = 0;
    for (size_t k = 0; k < m; ++k) {
        const Key &key = keys[k];
        if (k == 0 || keys[k - 1].round != key.round || keys[k - 1].family != key.family ||
            keys[k - 1].pb != key.pb) {
            plan.launches.push_back(
                {{key.family, key.pb, false, ctx->precision, !plan.scratch}, ordered.size(), 0, key.round});
            plan.launches.back().k.seg =
                runs && (key.family == sfmm::kLaneKernels || key.family == sfmm::kMfmaKernels);
            prev = nullptr;
        }
        sfmm_plan::Launch &L = plan.launches.back();
        if (key.family == sfmm::kGenericKernel) {
            L.k.pk = std::max({L.k.pk, key.p_in, key.p_out});
        } else if (key.p_in != key.pb || key.p_out != key.pb ||
                   (op == SFMM_M2L && ctx->precision == SFMM_FP64 && mfma_rt_square(key))) {
            L.k.rt = true;
        }
        const sfmm_item &src_it = items[key.index];
        sfmm_item it = src_it;
        if (plan.scratch) {
            const size_t wave = key.family == sfmm::kLaneKernels ? 64 : 16;
            const bool join = L.k.seg && prev && L.count % wave != 0 && prev->dst == src_it.dst &&
                              prev->p_in == src_it.p_in && prev->p_out == src_it.p_out && !zero(*prev) &&
                              !zero(src_it);
            if (!join) {  // a block of its own in the scratch pool
                run_off = scratch_reals;
                scratch_reals += size_t(it.p_out) * (it.p_out + 1);
            } else {
                is_head[list_pos[key.index]] = 0;
            }
            scr_off[list_pos[key.index]] = run_off;
            it.dst = run_off;
        }
        prev = &src_it;
        ordered.push_back(it);
        L.count++;
    }
    // Reduction tables: targets in order of first appearance, each with its
    // scratch blocks in the list order of their first items (a run's block
    // holds the sum of its items, added in the kernel).
    std::vector<sfmm::RedTarget> rt;
    std::vector<sfmm::RedItem> ri;
    if (plan.scratch) {
        FlatMap tid;
        std::vector<uint32_t> count;
        std::vector<uint32_t> of_pos(m);  // target id per list position
        size_t pos = 0, heads = 0;
        for (size_t i = 0; i < n; ++i) {
            const sfmm_item &it = items[i];
            if (it.p_in == 0 || it.p_out == 0) continue;
            bool fresh;
            const uint32_t t = tid.get(it.dst, uint32_t(rt.size()), fresh);
            if (fresh) {
                rt.push_back({it.dst, 0, 0, 0, 0});
                count.push_back(0);
            }
            of_pos[pos] = t;
            if (is_head[pos]) {
                count[t]++;
                ++heads;
            }
            ++pos;
            rt[t].len = std::max(rt[t].len, written_reals(op, it));
        }
        uint32_t first = 0;
        for (size_t t = 0; t < rt.size(); ++t) {
            rt[t].first = first;
            first += count[t];
            count[t] = 0;
        }
        ri.resize(heads);
        pos = 0;
        for (size_t i = 0; i < n; ++i) {
            const sfmm_item &it = items[i];
            if (it.p_in == 0 || it.p_out == 0) continue;
            if (is_head[pos]) {
                const uint32_t t = of_pos[pos];
                ri[rt[t].first + count[t]++] = {scr_off[pos], written_reals(op, it), 0};
            }
            ++pos;
        }
        for (size_t t = 0; t < rt.size(); ++t) rt[t].count = count[t];
    }
    plan.n_targets = rt.size();
    // device layout and the host image of its first three parts
    plan.off_targets = align_up(m * sizeof(sfmm_item), 256);
    plan.off_red_items = align_up(plan.off_targets + rt.size() * sizeof(sfmm::RedTarget), 256);
    plan.off_scratch = align_up(plan.off_red_items + ri.size() * sizeof(sfmm::RedItem), 256);
    plan.dev_bytes = plan.scratch ? plan.off_scratch + scratch_reals * sizeof(double) : m * sizeof(sfmm_item);
    const size_t image = plan.scratch ? plan.off_scratch : m * sizeof(sfmm_item);
    plan.blob.assign(image, 0);
    if (m) std::memcpy(plan.blob.data(), ordered.data(), m * sizeof(sfmm_item));
    if (plan.scratch) {
        std::memcpy(plan.blob.data() + plan.off_targets, rt.data(), rt.size() * sizeof(sfmm::RedTarget));
        std::memcpy(plan.blob.data() + plan.off_red_items, ri.data(), ri.size() * sizeof(sfmm::RedItem));
    }
    return SFMM_OK;
}

void drop_graph(sfmm_plan *plan)
{
    if (plan->exec) (void)hipGraphExecDestroy(plan->exec);
    plan->exec = nullptr;
}

hipError_t enqueue(sfmm_plan *plan, const double *src, double *dst, hipStream_t stream)
{
    double *out = plan->scratch ? reinterpret_cast<double *>(plan->d_mem + plan->off_scratch) : dst;
    for (const auto &L : plan->launches) {
        hipError_t e = sfmm::launch_kernel(plan->op, L.k, plan->d_items() + L.offset, L.count, src, out, stream);
        if (e != hipSuccess) return e;
    }
    if (plan->scratch)
        return sfmm::launch_reduce(reinterpret_cast<const sfmm::RedTarget *>(plan->d_mem + plan->off_targets),
                                   plan->n_targets,
                                   reinterpret_cast<const sfmm::RedItem *>(plan->d_mem + plan->off_red_items), out,
                                   dst, stream);
    return hipSuccess;
}

// Relative cost of a launch, to balance the branches of a captured round:
// items times the algorithmic bytes of its order bound.
double launch_cost(const sfmm_plan::Launch &L)
{
    const double p = double(L.k.pk);
    return double(L.count) * (3.0 * p * (p + 1.0) + 3.0);
}

// As enqueue, with the launches of each round spread over `stream` and the
// extra streams `aux` (forked from and joined back into `stream` by
// events): they write pairwise distinct targets, so they may run at once
// and one launch's tail overlaps the next's start.  Rounds stay in order;
// the reduction follows the last round.  Used under graph capture, where
// the branches become parallel graph nodes.
hipError_t enqueue_branched(sfmm_plan *plan, const double *src, double *dst, hipStream_t stream,
                            const std::vector<hipStream_t> &aux, const std::vector<hipEvent_t> &ev)
{
    double *out = plan->scratch ? reinterpret_cast<double *>(plan->d_mem + plan->off_scratch) : dst;
    const size_t nb = aux.size() + 1;
    for (size_t i = 0; i < plan->launches.size();) {
        size_t j = i;
        while (j < plan->launches.size() && plan->launches[j].round == plan->launches[i].round) ++j;
        hipError_t e = hipSuccess;
        if (j - i == 1 || nb == 1) {
            for (size_t k = i; k < j && e == hipSuccess; ++k) {
                const auto &L = plan->launches[k];
                e = sfmm::launch_kernel(plan->op, L.k, plan->d_items() + L.offset, L.count, src, out, stream);
            }
        } else {
            // heaviest launches first, each on the least loaded branch
            std::vector<size_t> order;
            for (size_t k = i; k < j; ++k) order.push_back(k);
            std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) {
                return launch_cost(plan->launches[a]) > launch_cost(plan->launches[b]);
            });
            std::vector<double> load(nb, 0.0);
            std::vector<std::vector<size_t>> branch(nb);
            for (size_t k : order) {
                const size_t b = size_t(std::min_element(load.begin(), load.end()) - load.begin());
                branch[b].push_back(k);
                load[b] += launch_cost(plan->launches[k]);
            }
            e = hipEventRecord(ev[0], stream);
            for (size_t b = 1; b < nb && e == hipSuccess; ++b)
                if (!branch[b].empty()) e = hipStreamWaitEvent(aux[b - 1], ev[0], 0);
            for (size_t b = 0; b < nb && e == hipSuccess; ++b) {
                hipStream_t s = b ? aux[b - 1] : stream;
                for (size_t k : branch[b]) {
                    const auto &L = plan->launches[k];
                    e = sfmm::launch_kernel(plan->op, L.k, plan->d_items() + L.offset, L.count, src, out, s);
                    if (e != hipSuccess) break;
                }
            }
            for (size_t b = 1; b < nb && e == hipSuccess; ++b) {
                if (branch[b].empty()) continue;
                e = hipEventRecord(ev[b], aux[b - 1]);
                if (e == hipSuccess) e = hipStreamWaitEvent(stream, ev[b], 0);
            }
        }
        if (e != hipSuccess) return e;
        i = j;
    }
    if (plan->scratch)
        return sfmm::launch_reduce(reinterpret_cast<const sfmm::RedTarget *>(plan->d_mem + plan->off_targets),
                                   plan->n_targets,
                                   reinterpret_cast<const sfmm::RedItem *>(plan->d_mem + plan->off_red_items), out,
                                   dst, stream);
    return hipSuccess;
}

// Branches used by sfmm_plan_execute and sfmm_plan_capture: the caller's
// stream and three forked ones (default priority).
constexpr size_t kBranches = 4;

bool has_parallel_round(const sfmm_plan *plan)
{
    for (size_t k = 1; k < plan->launches.size(); ++k)
        if (plan->launches[k].round == plan->launches[k - 1].round) return true;
    return false;
}

// The context's branch streams and events, created on first use (with the
// context's device current and its lock held).
hipError_t ensure_branches(sfmm_context *ctx)
{
    hipError_t e = hipSuccess;
    while (e == hipSuccess && ctx->branch_streams.size() + 1 < kBranches) {
        hipStream_t a = nullptr;
        e = hipStreamCreateWithFlags(&a, hipStreamNonBlocking);
        if (e == hipSuccess) ctx->branch_streams.push_back(a);
    }
    while (e == hipSuccess && ctx->branch_events.size() < kBranches) {
        hipEvent_t x = nullptr;
        e = hipEventCreateWithFlags(&x, hipEventDisableTiming);
        if (e == hipSuccess) ctx->branch_events.push_back(x);
    }
    return e;
}

// Device tables the kernels read, created before any capture so that no
// allocation happens inside one.
bool tables_ready() { return sfmm::mfma::device_table() && sfmm::mfma::device_factorials(); }

// Device scratch of the allocation-free entry point (sfmm_translate_noalloc):
// the plan's tables (items, reduction targets and items) and, when targets
// repeat, its per-item scratch blocks.  One copy per GPU, held until the
// process ends.  Host code never uses the array's name as an address: the
// per-device address comes from hipGetSymbolAddress at set-up.
constexpr size_t kNoAllocBytes = SFMM_NOALLOC_SCRATCH_BYTES;
static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

struct ScratchAddresses {
    std::mutex mu;
    std::vector<unsigned char *> per_device;
};

ScratchAddresses &scratch_addresses()
{
    static ScratchAddresses s;
    return s;
}

// Once per device, with that device current.
hipError_t scratch_setup(int device)
{
    ScratchAddresses &st = scratch_addresses();
    std::lock_guard<std::mutex> lock(st.mu);
    if (st.per_device.size() <= size_t(device)) st.per_device.resize(device + 1, nullptr);
    if (st.per_device[device]) return hipSuccess;
    void *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    st.per_device[device] = static_cast<unsigned char *>(ptr);
    return hipSuccess;
}

unsigned char *scratch_address(int device)
{
    ScratchAddresses &st = scratch_addresses();
    std::lock_guard<std::mutex> lock(st.mu);
    return size_t(device) < st.per_device.size() ? st.per_device[device] : nullptr;
}

}  // namespace

extern "C" {

size_t sfmm_max_order(void) { return sfmm::max_kernel_order(); }

int sfmm_create(sfmm_context **out, int device, size_t P, int precision)
{
    if (!out) return SFMM_ERR_LOGIC;
    *out = nullptr;
    if (P == 0) return SFMM_ERR_LOGIC;
    if (precision != SFMM_FP64 && precision != SFMM_FP32 && precision != SFMM_SPLIT32) return SFMM_ERR_LOGIC;
    if (P > (precision == SFMM_FP64 ? sfmm::max_kernel_order() : size_t(sfmm::kPrecMaxOrder)))
        return SFMM_ERR_OVERFLOW;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || device < 0 || device >= count) return SFMM_ERR_HIP;
    // Set-up on the device (its current device restored afterwards): the
    // swap and factorial tables, and the address of the allocation-free
    // entry point's scratch array.
    int prev = 0;
    if (hipGetDevice(&prev) != hipSuccess) return SFMM_ERR_HIP;
    if (hipSetDevice(device) != hipSuccess) return SFMM_ERR_HIP;
    const bool ready = tables_ready() && scratch_setup(device) == hipSuccess;
    (void)hipSetDevice(prev);
    if (!ready) return SFMM_ERR_HIP;
    sfmm_context *ctx = new (std::nothrow) sfmm_context;
    if (!ctx) return SFMM_ERR_ALLOC;
    ctx->device = device;
    ctx->P = P;
    ctx->precision = precision;
    *out = ctx;
    return SFMM_OK;
}

void sfmm_destroy(sfmm_context *ctx)
{
    if (ctx && ctx->host_staging) sfmm_staging_destroy(ctx->host_staging);
    if (ctx) {
        for (hipEvent_t x : ctx->branch_events) (void)hipEventDestroy(x);
        for (hipStream_t a : ctx->branch_streams) (void)hipStreamDestroy(a);
    }
    delete ctx;
}

size_t sfmm_order(const sfmm_context *ctx) { return ctx ? ctx->P : 0; }

int sfmm_device(const sfmm_context *ctx) { return ctx ? ctx->device : -1; }

const char *sfmm_last_error(const sfmm_context *ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int sfmm_set_kernels(sfmm_context *ctx, int family)
{
    if (!ctx) return SFMM_ERR_LOGIC;
    if (family < SFMM_KERNELS_AUTO || family > SFMM_KERNELS_GENERIC)
        return fail(ctx, SFMM_ERR_LOGIC, "solidfmm: unknown kernel family");
    std::lock_guard<std::mutex> lock(ctx->mu);
    ctx->family = family;
    return SFMM_OK;
}

}  // extern "C"

namespace {

// Validation and grouping into a new plan with no device memory yet (its
// device image stays on the host).
int prepare_plan(sfmm_context *ctx, int op, const sfmm_item *items, size_t n, sfmm_plan **out)
{
    sfmm_plan *plan = new (std::nothrow) sfmm_plan;
    if (!plan) return fail(ctx, SFMM_ERR_ALLOC, "solidfmm: out of host memory");
    int rc;
    try {
        rc = build_plan(ctx, op, items, n, *plan);
    } catch (const std::bad_alloc &) {
        rc = fail(ctx, SFMM_ERR_ALLOC, "solidfmm: out of host memory");
    }
    if (rc != SFMM_OK) {
        delete plan;
        return rc;
    }
    plan->ctx = ctx;
    plan->op = op;
    *out = plan;
    return SFMM_OK;
}

}  // namespace

extern "C" {

int sfmm_plan_create(sfmm_context *ctx, int op, const sfmm_item *items, size_t n, sfmm_plan **out)
{
    if (!ctx || !out) return SFMM_ERR_LOGIC;
    std::lock_guard<std::mutex> lock(ctx->mu);
    *out = nullptr;
    sfmm_plan *plan = nullptr;
    int rc = prepare_plan(ctx, op, items, n, &plan);
    if (rc != SFMM_OK) return rc;
    hipError_t e = hipSetDevice(ctx->device);
    if (e == hipSuccess && plan->dev_bytes) {
        e = hipMalloc(&plan->d_mem, plan->dev_bytes);
        if (e != hipSuccess) {
            plan->d_mem = nullptr;
            delete plan;
            return fail(ctx, SFMM_ERR_ALLOC, std::string("hipMalloc: ") + hipGetErrorString(e));
        }
        e = hipMemcpy(plan->d_mem, plan->blob.data(), plan->blob.size(), hipMemcpyHostToDevice);
    }
    if (e != hipSuccess) {
        if (plan->d_mem) (void)hipFree(plan->d_mem);
        delete plan;
        return hip_fail(ctx, e, "plan upload");
    }
    plan->blob.clear();
    plan->blob.shrink_to_fit();
    *out = plan;
    return SFMM_OK;
}

int sfmm_plan_prepare(sfmm_context *ctx, int op, const sfmm_item *items, size_t n, sfmm_plan **out)
{
    if (!ctx || !out) return SFMM_ERR_LOGIC;
    *out = nullptr;
    // build_plan reads only the context's order and kernel choices: copy
    // them, so the preparation itself runs outside the lock
    sfmm_context view;
    {
        std::lock_guard<std::mutex> lock(ctx->mu);
        view.device = ctx->device;
        view.P = ctx->P;
        view.family = ctx->family;
        view.precision = ctx->precision;
        std::memcpy(view.tune, ctx->tune, sizeof(view.tune));
    }
    sfmm_plan *plan = nullptr;
    const int rc = prepare_plan(&view, op, items, n, &plan);
    if (rc != SFMM_OK) {
        std::lock_guard<std::mutex> lock(ctx->mu);
        ctx->err = view.err;
        return rc;
    }
    plan->ctx = ctx;
    plan->owns_mem = false;
    *out = plan;
    return SFMM_OK;
}

size_t sfmm_plan_device_bytes(const sfmm_plan *plan) { return plan ? plan->dev_bytes : 0; }

size_t sfmm_plan_host_bytes(const sfmm_plan *plan) { return plan ? plan->blob.size() : 0; }

int sfmm_plan_bind_staged(sfmm_plan *plan, void *dev_mem, size_t dev_bytes, void *host_mem, size_t host_bytes,
                          void *stream)
{
    if (!plan) return SFMM_ERR_LOGIC;
    sfmm_context *ctx = plan->ctx;
    std::lock_guard<std::mutex> lock(ctx->mu);
    if (plan->owns_mem) return fail(ctx, SFMM_ERR_LOGIC, "solidfmm: plan owns its device memory");
    if (plan->dev_bytes && (!dev_mem || dev_bytes < plan->dev_bytes || (reinterpret_cast<uintptr_t>(dev_mem) & 255u)))
        return fail(ctx, SFMM_ERR_LOGIC, "solidfmm: device buffer missing, misaligned or too small");
    if (!plan->blob.empty() && (!host_mem || host_bytes < plan->blob.size()))
        return fail(ctx, SFMM_ERR_LOGIC, "solidfmm: host staging missing or too small");
    hipError_t e = hipSetDevice(ctx->device);
    if (e == hipSuccess && !plan->blob.empty()) {
        std::memcpy(host_mem, plan->blob.data(), plan->blob.size());
        e = hipMemcpyAsync(dev_mem, host_mem, plan->blob.size(), hipMemcpyHostToDevice,
                           static_cast<hipStream_t>(stream));
    }
    if (e != hipSuccess) return hip_fail(ctx, e, "hipMemcpyAsync");
    plan->d_mem = static_cast<unsigned char *>(dev_mem);
    return SFMM_OK;
}

int sfmm_plan_copy_tables(const sfmm_plan *plan, void *host_mem, size_t host_bytes)
{
    if (!plan) return SFMM_ERR_LOGIC;
    if (!plan->blob.empty() && (!host_mem || host_bytes < plan->blob.size())) {
        std::lock_guard<std::mutex> lock(plan->ctx->mu);
        return fail(plan->ctx, SFMM_ERR_LOGIC, "solidfmm: host staging missing or too small");
    }
    if (!plan->blob.empty()) std::memcpy(host_mem, plan->blob.data(), plan->blob.size());
    return SFMM_OK;
}

int sfmm_plan_bind_pinned(sfmm_plan *plan, void *dev_mem, size_t dev_bytes, const void *host_tables, void *stream)
{
    if (!plan) return SFMM_ERR_LOGIC;
    sfmm_context *ctx = plan->ctx;
    std::lock_guard<std::mutex> lock(ctx->mu);
    if (plan->owns_mem) return fail(ctx, SFMM_ERR_LOGIC, "solidfmm: plan owns its device memory");
    if (plan->dev_bytes && (!dev_mem || dev_bytes < plan->dev_bytes || (reinterpret_cast<uintptr_t>(dev_mem) & 255u)))
        return fail(ctx, SFMM_ERR_LOGIC, "solidfmm: device buffer missing, misaligned or too small");
    if (!plan->blob.empty() && !host_tables) return fail(ctx, SFMM_ERR_LOGIC, "solidfmm: host tables missing");
    hipError_t e = hipSetDevice(ctx->device);
    if (e == hipSuccess && !plan->blob.empty())
        e = hipMemcpyAsync(dev_mem, host_tables, plan->blob.size(), hipMemcpyHostToDevice,
                           static_cast<hipStream_t>(stream));
    if (e != hipSuccess) return hip_fail(ctx, e, "hipMemcpyAsync");
    plan->d_mem = static_cast<unsigned char *>(dev_mem);
    return SFMM_OK;
}

int sfmm_plan_bind(sfmm_plan *plan, void *dev_mem, size_t dev_bytes, void *stream)
{
    if (!plan) return SFMM_ERR_LOGIC;
    sfmm_context *ctx = plan->ctx;
    std::lock_guard<std::mutex> lock(ctx->mu);
    if (plan->owns_mem) return fail(ctx, SFMM_ERR_LOGIC, "solidfmm: plan owns its device memory");
    if (plan->dev_bytes && (!dev_mem || dev_bytes < plan->dev_bytes || (reinterpret_cast<uintptr_t>(dev_mem) & 255u)))
        return fail(ctx, SFMM_ERR_LOGIC, "solidfmm: device buffer missing, misaligned or too small");
    hipError_t e = hipSetDevice(ctx->device);
    if (e == hipSuccess && !plan->blob.empty())
        e = hipMemcpyAsync(dev_mem, plan->blob.data(), plan->blob.size(), hipMemcpyHostToDevice,
                           static_cast<hipStream_t>(stream));
    if (e != hipSuccess) return hip_fail(ctx, e, "hipMemcpyAsync");
    plan->d_mem = static_cast<unsigned char *>(dev_mem);
    return SFMM_OK;
}

void sfmm_plan_destroy(sfmm_plan *plan)
{
    if (!plan) return;
    drop_graph(plan);
    if (plan->d_mem && plan->owns_mem) (void)hipFree(plan->d_mem);
    delete plan;
}

size_t sfmm_plan_launches(const sfmm_plan *plan) { return plan ? plan->launches.size() : 0; }

void sfmm_plan_extents(const sfmm_plan *plan, size_t *src_len, size_t *dst_len)
{
    if (src_len) *src_len = plan ? plan->src_extent : 0;
    if (dst_len) *dst_len = plan ? plan->dst_extent : 0;
}

int sfmm_plan_launch_info(const sfmm_plan *plan, size_t i, int *family, uint32_t *order, int *per_item_orders,
                          size_t *count)
{
    if (!plan || i >= plan->launches.size()) return SFMM_ERR_LOGIC;
    const auto &L = plan->launches[i];
    if (family) *family = L.k.family;
    if (order) *order = L.k.pk;
    if (per_item_orders) *per_item_orders = L.k.rt ? 1 : 0;
    if (count) *count = L.count;
    return SFMM_OK;
}

int sfmm_plan_execute(sfmm_plan *plan, const double *src_pool, size_t src_len, double *dst_pool, size_t dst_len,
                      void *stream)
{
    if (!plan) return SFMM_ERR_LOGIC;
    sfmm_context *ctx = plan->ctx;
    std::lock_guard<std::mutex> lock(ctx->mu);
    if (plan->n_items && (!src_pool || !dst_pool)) return fail(ctx, SFMM_ERR_LOGIC, "solidfmm: null pool");
    if (plan->n_items && !plan->d_mem) return fail(ctx, SFMM_ERR_LOGIC, "solidfmm: plan not bound to device memory");
    if (src_len < plan->src_extent || dst_len < plan->dst_extent)
        return fail(ctx, SFMM_ERR_LOGIC, "solidfmm: pool shorter than the plan's extent");
    hipError_t e = hipSetDevice(ctx->device);
    if (e != hipSuccess) return hip_fail(ctx, e, "hipSetDevice");
    if (has_parallel_round(plan)) {
        e = ensure_branches(ctx);
        if (e != hipSuccess) return hip_fail(ctx, e, "branch streams");
        e = enqueue_branched(plan, src_pool, dst_pool, static_cast<hipStream_t>(stream), ctx->branch_streams,
                             ctx->branch_events);
    } else {
        e = enqueue(plan, src_pool, dst_pool, static_cast<hipStream_t>(stream));
    }
    if (e != hipSuccess) return hip_fail(ctx, e, "kernel launch");
    return SFMM_OK;
}

// One chain of launches on the caller's stream, as sfmm_plan_execute was
// before branching: the request loop (server.cpp) overlaps requests on its
// own slot streams, and branches shared by all slots would couple them.
__attribute__((visibility("hidden"))) int sfmm_plan_execute_chain(sfmm_plan *plan, const double *src_pool,
                                                                 size_t src_len, double *dst_pool, size_t dst_len,
                                                                 void *stream)
{
    if (!plan) return SFMM_ERR_LOGIC;
    sfmm_context *ctx = plan->ctx;
    std::lock_guard<std::mutex> lock(ctx->mu);
    if (plan->n_items && (!src_pool || !dst_pool)) return fail(ctx, SFMM_ERR_LOGIC, "solidfmm: null pool");
    if (plan->n_items && !plan->d_mem) return fail(ctx, SFMM_ERR_LOGIC, "solidfmm: plan not bound to device memory");
    if (src_len < plan->src_extent || dst_len < plan->dst_extent)
        return fail(ctx, SFMM_ERR_LOGIC, "solidfmm: pool shorter than the plan's extent");
    hipError_t e = hipSetDevice(ctx->device);
    if (e == hipSuccess) e = enqueue(plan, src_pool, dst_pool, static_cast<hipStream_t>(stream));
    if (e != hipSuccess) return hip_fail(ctx, e, "kernel launch");
    return SFMM_OK;
}

int sfmm_translate_noalloc(sfmm_context *ctx, int op, const sfmm_item *host_items, size_t n, const double *src_pool,
                           size_t src_len, double *dst_pool, size_t dst_len, void *stream)
{
    if (!ctx) return SFMM_ERR_LOGIC;
    sfmm_plan *plan = nullptr;
    int rc = sfmm_plan_prepare(ctx, op, host_items, n, &plan);  // host memory only
    if (rc != SFMM_OK) return rc;
    std::lock_guard<std::mutex> lock(ctx->mu);
    if (plan->dev_bytes > kNoAllocBytes) {
        rc = fail(ctx, SFMM_ERR_LOGIC,
                  "solidfmm: the list needs " + std::to_string(plan->dev_bytes) +
                      " bytes of device scratch; sfmm_translate_noalloc has " + std::to_string(kNoAllocBytes) +
                      " (SFMM_NOALLOC_SCRATCH_BYTES)");
    } else if (plan->n_items && (!src_pool || !dst_pool)) {
        rc = fail(ctx, SFMM_ERR_LOGIC, "solidfmm: null pool");
    } else if (src_len < plan->src_extent || dst_len < plan->dst_extent) {
        rc = fail(ctx, SFMM_ERR_LOGIC, "solidfmm: pool shorter than the plan's extent");
    } else if (plan->n_items) {
        unsigned char *ptr = scratch_address(ctx->device);
        hipError_t e = ptr ? hipSetDevice(ctx->device) : hipErrorNotInitialized;
        // the tables come from the plan's pageable host memory: the runtime
        // stages such a copy before returning, so the plan may go at once
        if (e == hipSuccess && !plan->blob.empty())
            e = hipMemcpyAsync(ptr, plan->blob.data(), plan->blob.size(), hipMemcpyHostToDevice,
                               static_cast<hipStream_t>(stream));
        if (e == hipSuccess) {
            plan->d_mem = ptr;
            e = enqueue(plan, src_pool, dst_pool, static_cast<hipStream_t>(stream));
        }
        if (e != hipSuccess) rc = hip_fail(ctx, e, "sfmm_translate_noalloc");
    }
    plan->d_mem = nullptr;  // the scratch array is the library's, not the plan's
    delete plan;
    return rc;
}

int sfmm_plan_capture(sfmm_plan *plan, const double *src_pool, size_t src_len, double *dst_pool, size_t dst_len)
{
    if (!plan) return SFMM_ERR_LOGIC;
    sfmm_context *ctx = plan->ctx;
    std::lock_guard<std::mutex> lock(ctx->mu);
    if (plan->n_items && (!src_pool || !dst_pool)) return fail(ctx, SFMM_ERR_LOGIC, "solidfmm: null pool");
    if (plan->n_items && !plan->d_mem) return fail(ctx, SFMM_ERR_LOGIC, "solidfmm: plan not bound to device memory");
    if (src_len < plan->src_extent || dst_len < plan->dst_extent)
        return fail(ctx, SFMM_ERR_LOGIC, "solidfmm: pool shorter than the plan's extent");
    hipError_t e = hipSetDevice(ctx->device);
    if (e != hipSuccess) return hip_fail(ctx, e, "hipSetDevice");
    if (!tables_ready()) return fail(ctx, SFMM_ERR_ALLOC, "solidfmm: device tables");
    drop_graph(plan);
    // the capture stream plus kBranches - 1 forked ones (default priority)
    hipStream_t s = nullptr;
    std::vector<hipStream_t> aux;
    std::vector<hipEvent_t> ev;
    e = hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
    for (size_t b = 1; b < kBranches && e == hipSuccess; ++b) {
        hipStream_t a = nullptr;
        e = hipStreamCreateWithFlags(&a, hipStreamNonBlocking);
        if (e == hipSuccess) aux.push_back(a);
    }
    for (size_t b = 0; b < kBranches && e == hipSuccess; ++b) {
        hipEvent_t x = nullptr;
        e = hipEventCreateWithFlags(&x, hipEventDisableTiming);
        if (e == hipSuccess) ev.push_back(x);
    }
    hipGraph_t graph = nullptr;
    if (e == hipSuccess) {
        e = hipStreamBeginCapture(s, hipStreamCaptureModeRelaxed);
        if (e == hipSuccess) {
            hipError_t le = enqueue_branched(plan, src_pool, dst_pool, s, aux, ev);
            e = hipStreamEndCapture(s, &graph);
            if (le != hipSuccess) e = le;
        }
    }
    if (e == hipSuccess) e = hipGraphInstantiate(&plan->exec, graph, nullptr, nullptr, 0);
    if (graph) (void)hipGraphDestroy(graph);
    for (hipEvent_t x : ev) (void)hipEventDestroy(x);
    for (hipStream_t a : aux) (void)hipStreamDestroy(a);
    if (s) (void)hipStreamDestroy(s);
    if (e != hipSuccess) {
        plan->exec = nullptr;
        return hip_fail(ctx, e, "graph capture");
    }
    return SFMM_OK;
}

int sfmm_plan_replay(sfmm_plan *plan, void *stream)
{
    if (!plan) return SFMM_ERR_LOGIC;
    sfmm_context *ctx = plan->ctx;
    std::lock_guard<std::mutex> lock(ctx->mu);
    if (!plan->exec) return fail(ctx, SFMM_ERR_LOGIC, "solidfmm: plan has no captured graph");
    hipError_t e = hipSetDevice(ctx->device);
    if (e == hipSuccess) e = hipGraphLaunch(plan->exec, static_cast<hipStream_t>(stream));
    if (e != hipSuccess) return hip_fail(ctx, e, "hipGraphLaunch");
    return SFMM_OK;
}

int sfmm_launch(sfmm_context *ctx, int op, uint32_t p_in, uint32_t p_out, const sfmm_item *dev_items, size_t n,
                const double *src_pool, double *dst_pool, void *stream)
{
    if (!ctx) return SFMM_ERR_LOGIC;
    std::lock_guard<std::mutex> lock(ctx->mu);
    if (op < SFMM_M2M || op > SFMM_L2L) return fail(ctx, SFMM_ERR_LOGIC, "solidfmm: unknown operator");
    if (p_in > ctx->P || p_out > ctx->P) return fail(ctx, SFMM_ERR_RANGE, "solidfmm: orders exceeding operator data.");
    if (n == 0 || p_in == 0 || p_out == 0) return SFMM_OK;
    int fam = 0;
    if (!choose(ctx, op, p_in, p_out, fam)) return fail(ctx, SFMM_ERR_UNSUPPORTED, "solidfmm: no kernel for orders");
    hipError_t e = hipSetDevice(ctx->device);
    if (e != hipSuccess) return hip_fail(ctx, e, "hipSetDevice");
    const uint32_t pb = std::max(p_in, p_out);
    const sfmm::KernelChoice k{fam, pb, p_in != p_out, ctx->precision};
    e = sfmm::launch_kernel(op, k, dev_items, n, src_pool, dst_pool, static_cast<hipStream_t>(stream));
    if (e != hipSuccess) return hip_fail(ctx, e, "kernel launch");
    return SFMM_OK;
}

// Autotuner: for each bucket pb of the list, time every family that covers
// pb on a sample of the bucket's own items (up to kSample, sorted as a plan
// sorts them) in scratch pools, and keep the fastest for plans created
// afterwards.
int sfmm_autotune(sfmm_context *ctx, int op, const sfmm_item *items, size_t n, void *stream)
{
    if (!ctx) return SFMM_ERR_LOGIC;
    if (op < SFMM_M2M || op > SFMM_L2L) return fail(ctx, SFMM_ERR_LOGIC, "solidfmm: unknown operator");
    if (n && !items) return fail(ctx, SFMM_ERR_LOGIC, "solidfmm: null item list");
    if (ctx->precision != SFMM_FP64) return SFMM_OK;  // one kernel per case: nothing to choose
    // large enough that every family fills the GPU (256k items: four waves
    // per SIMD for the one-lane-per-translation kernels); a 32k sample left
    // those at half a wave per SIMD and picked MFMA at order 12, 38 % slower
    // on the FMM's 295k-item bucket
    constexpr size_t kSample = size_t(1) << 18;
    constexpr int kReps = 5;
    std::vector<std::vector<sfmm_item>> bucket(kOrders);
    for (size_t i = 0; i < n; ++i) {
        const sfmm_item &it = items[i];
        if (it.p_in == 0 || it.p_out == 0) continue;
        if (it.p_in > ctx->P || it.p_out > ctx->P)
            return fail(ctx, SFMM_ERR_RANGE, "solidfmm: orders exceeding operator data.");
        const uint32_t pb = std::max(it.p_in, it.p_out);
        if (bucket[pb].size() < kSample) bucket[pb].push_back(it);
    }
    std::lock_guard<std::mutex> lock(ctx->mu);
    hipError_t e = hipSetDevice(ctx->device);
    if (e != hipSuccess) return hip_fail(ctx, e, "hipSetDevice");
    if (!tables_ready()) return fail(ctx, SFMM_ERR_ALLOC, "solidfmm: device tables");
    hipStream_t s = static_cast<hipStream_t>(stream);
    // pools for the largest bucket: one block of the bucket's order per
    // sample item
    size_t reals = 0;
    for (uint32_t pb = 1; pb < uint32_t(kOrders); ++pb)
        reals = std::max(reals, bucket[pb].size() * size_t(pb) * (pb + 1));
    if (reals == 0) return SFMM_OK;
    double *src = nullptr, *dst = nullptr;
    hipEvent_t ev[2] = {nullptr, nullptr};
    int rc = SFMM_OK;
    e = hipMalloc(&src, reals * sizeof(double));
    if (e == hipSuccess) e = hipMalloc(&dst, reals * sizeof(double));
    if (e != hipSuccess) {
        rc = fail(ctx, SFMM_ERR_ALLOC, std::string("hipMalloc: ") + hipGetErrorString(e));
    } else {
        // 0x3F bytes: every double is ~4.8e-4 (normal numbers)
        e = hipMemsetAsync(src, 0x3F, reals * sizeof(double), s);
        if (e == hipSuccess) e = hipMemsetAsync(dst, 0x3F, reals * sizeof(double), s);
        if (e == hipSuccess) e = hipEventCreate(&ev[0]);
        if (e == hipSuccess) e = hipEventCreate(&ev[1]);
        if (e != hipSuccess) rc = hip_fail(ctx, e, "autotune setup");
    }
    // Each family is timed on the plan it would get for the bucket's sample:
    // the sample keeps the list's order and which of its items share a
    // target (targets remapped to a compact pool, sources to distinct
    // blocks), so repeated targets are timed in scratch mode with their runs
    // and the reduction, as the real list runs.
    std::unordered_map<uint64_t, uint64_t> tid;
    for (uint32_t pb = 1; rc == SFMM_OK && pb < uint32_t(kOrders); ++pb) {
        std::vector<sfmm_item> &b = bucket[pb];
        if (b.empty()) continue;
        const size_t block = size_t(pb) * (pb + 1);
        tid.clear();
        for (size_t i = 0; i < b.size(); ++i) {
            b[i].src = i * block;
            b[i].dst = tid.emplace(b[i].dst, uint64_t(tid.size())).first->second * block;
        }
        int best = 0;
        float best_ms = 0.0f;
        for (int fam : {int(sfmm::kLaneKernels), int(sfmm::kMfmaKernels), int(sfmm::kGenericKernel)}) {
            if (!sfmm::family_covers(fam, pb)) continue;
            sfmm_context view;
            view.device = ctx->device;
            view.P = ctx->P;
            view.family = SFMM_KERNELS_AUTO;
            view.precision = ctx->precision;
            std::memcpy(view.tune, ctx->tune, sizeof(view.tune));
            view.tune[op][pb] = fam;
            sfmm_plan *plan = nullptr;
            rc = prepare_plan(&view, op, b.data(), b.size(), &plan);
            if (rc != SFMM_OK) {
                ctx->err = view.err;
                break;
            }
            plan->ctx = ctx;
            e = hipSuccess;
            if (plan->dev_bytes) {
                e = hipMalloc(&plan->d_mem, plan->dev_bytes);
                if (e != hipSuccess) plan->d_mem = nullptr;
                if (e == hipSuccess)
                    e = hipMemcpyAsync(plan->d_mem, plan->blob.data(), plan->blob.size(), hipMemcpyHostToDevice, s);
            }
            if (e == hipSuccess) e = enqueue(plan, src, dst, s);  // warm-up
            float ms = 1e30f;
            for (int r = 0; r < kReps && e == hipSuccess; ++r) {
                e = hipEventRecord(ev[0], s);
                if (e == hipSuccess) e = enqueue(plan, src, dst, s);
                if (e == hipSuccess) e = hipEventRecord(ev[1], s);
                if (e == hipSuccess) e = hipEventSynchronize(ev[1]);
                float t = 0.0f;
                if (e == hipSuccess) e = hipEventElapsedTime(&t, ev[0], ev[1]);
                ms = std::min(ms, t);
            }
            if (e == hipSuccess) e = hipStreamSynchronize(s);
            sfmm_plan_destroy(plan);
            if (e != hipSuccess) {
                rc = hip_fail(ctx, e, "autotune launch");
                break;
            }
            if (!best || ms < best_ms) {
                best = fam;
                best_ms = ms;
            }
        }
        if (rc == SFMM_OK) ctx->tune[op][pb] = best;
    }
    if (ev[0]) (void)hipEventDestroy(ev[0]);
    if (ev[1]) (void)hipEventDestroy(ev[1]);
    if (src) (void)hipFree(src);
    if (dst) (void)hipFree(dst);
    return rc;
}

int sfmm_get_tuning(sfmm_context *ctx, int op, int *family_by_order, size_t len)
{
    if (!ctx || !family_by_order) return SFMM_ERR_LOGIC;
    if (op < SFMM_M2M || op > SFMM_L2L) return fail(ctx, SFMM_ERR_LOGIC, "solidfmm: unknown operator");
    std::lock_guard<std::mutex> lock(ctx->mu);
    for (size_t pb = 0; pb < len; ++pb)
        family_by_order[pb] = (pb == 0 || pb >= size_t(kOrders)) ? 0 : family_for(ctx, op, uint32_t(pb));
    return SFMM_OK;
}

int sfmm_set_tuning(sfmm_context *ctx, int op, uint32_t order, int family)
{
    if (!ctx) return SFMM_ERR_LOGIC;
    if (op < SFMM_M2M || op > SFMM_L2L) return fail(ctx, SFMM_ERR_LOGIC, "solidfmm: unknown operator");
    if (order < 1 || order >= uint32_t(kOrders)) return fail(ctx, SFMM_ERR_RANGE, "solidfmm: order out of range");
    if (family != SFMM_KERNELS_AUTO && !sfmm::family_covers(family, order))
        return fail(ctx, SFMM_ERR_UNSUPPORTED, "solidfmm: family has no kernel for this order");
    std::lock_guard<std::mutex> lock(ctx->mu);
    ctx->tune[op][order] = family;
    return SFMM_OK;
}

struct sfmm_staging {
    sfmm_context *ctx = nullptr;
    double *h_src = nullptr, *h_dst = nullptr;  // pinned
    size_t h_src_cap = 0, h_dst_cap = 0;        // reals
    double *d_src = nullptr, *d_dst = nullptr;
    size_t d_src_cap = 0, d_dst_cap = 0;  // reals
    unsigned char *d_plan = nullptr;
    size_t d_plan_cap = 0;  // bytes
    double *h_plan = nullptr;  // pinned: the plan's tables pass through it
    size_t h_plan_cap = 0;     // reals
    hipStream_t stream = nullptr;
};

}  // extern "C"

namespace {

// grow-only pinned host buffer of at least `want` reals
hipError_t grow_host(double *&p, size_t &cap, size_t want)
{
    if (want <= cap) return hipSuccess;
    if (p) (void)hipHostFree(p);
    p = nullptr;
    cap = 0;
    const size_t n = std::max<size_t>(want + want / 4, 1024);
    hipError_t e = hipHostMalloc(reinterpret_cast<void **>(&p), n * sizeof(double), hipHostMallocDefault);
    if (e == hipSuccess) cap = n;
    return e;
}

template <class T>
hipError_t grow_dev(T *&p, size_t &cap, size_t want, size_t unit)
{
    if (want <= cap) return hipSuccess;
    if (p) (void)hipFree(p);
    p = nullptr;
    cap = 0;
    const size_t n = std::max<size_t>(want + want / 4, 1024);
    hipError_t e = hipMalloc(reinterpret_cast<void **>(&p), n * unit);
    if (e == hipSuccess) cap = n;
    return e;
}

// Copy with up to 16 host threads (large copies only).
void parallel_copy(double *dst, const double *src, size_t n)
{
    const size_t bytes = n * sizeof(double);
    const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
    const unsigned nt = unsigned(std::min<size_t>({size_t(16), size_t(hw), bytes / (4u << 20) + 1}));
    if (nt <= 1) {
        std::memcpy(dst, src, bytes);
        return;
    }
    std::vector<std::thread> pool;
    const size_t chunk = (n + nt - 1) / nt;
    for (unsigned t = 0; t < nt; ++t) {
        const size_t b = t * chunk, e = std::min(n, b + chunk);
        if (b >= e) break;
        pool.emplace_back([=] { std::memcpy(dst + b, src + b, (e - b) * sizeof(double)); });
    }
    for (auto &th : pool) th.join();
}

}  // namespace

extern "C" {

int sfmm_staging_create(sfmm_context *ctx, sfmm_staging **out)
{
    if (!ctx || !out) return SFMM_ERR_LOGIC;
    *out = nullptr;
    sfmm_staging *st = new (std::nothrow) sfmm_staging;
    if (!st) return fail(ctx, SFMM_ERR_ALLOC, "solidfmm: out of host memory");
    st->ctx = ctx;
    hipError_t e = hipSetDevice(ctx->device);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&st->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete st;
        std::lock_guard<std::mutex> lock(ctx->mu);
        return hip_fail(ctx, e, "staging stream");
    }
    *out = st;
    return SFMM_OK;
}

void sfmm_staging_destroy(sfmm_staging *st)
{
    if (!st) return;
    (void)hipSetDevice(st->ctx->device);
    if (st->stream) (void)hipStreamSynchronize(st->stream);
    if (st->h_src) (void)hipHostFree(st->h_src);
    if (st->h_dst) (void)hipHostFree(st->h_dst);
    if (st->d_src) (void)hipFree(st->d_src);
    if (st->d_dst) (void)hipFree(st->d_dst);
    if (st->d_plan) (void)hipFree(st->d_plan);
    if (st->h_plan) (void)hipHostFree(st->h_plan);
    if (st->stream) (void)hipStreamDestroy(st->stream);
    delete st;
}

int sfmm_staging_reserve(sfmm_staging *st, size_t src_len, size_t dst_len, double **h_src, double **h_dst)
{
    if (!st || !h_src || !h_dst) return SFMM_ERR_LOGIC;
    hipError_t e = hipSetDevice(st->ctx->device);
    if (e == hipSuccess) e = grow_host(st->h_src, st->h_src_cap, std::max<size_t>(src_len, 2));
    if (e == hipSuccess) e = grow_host(st->h_dst, st->h_dst_cap, std::max<size_t>(dst_len, 2));
    if (e != hipSuccess) {
        std::lock_guard<std::mutex> lock(st->ctx->mu);
        return fail(st->ctx, SFMM_ERR_ALLOC, std::string("hipHostMalloc: ") + hipGetErrorString(e));
    }
    *h_src = st->h_src;
    *h_dst = st->h_dst;
    return SFMM_OK;
}

int sfmm_staging_translate(sfmm_staging *st, int op, const sfmm_item *items, size_t n, size_t src_len,
                           size_t dst_len)
{
    if (!st) return SFMM_ERR_LOGIC;
    sfmm_context *ctx = st->ctx;
    if (src_len > st->h_src_cap || dst_len > st->h_dst_cap) {
        std::lock_guard<std::mutex> lock(ctx->mu);
        return fail(ctx, SFMM_ERR_LOGIC, "solidfmm: staging pools not reserved for these lengths");
    }
    // Bounds of every block against the pools, before any copy.
    for (size_t i = 0; i < n; ++i) {
        const sfmm_item &it = items[i];
        if (it.src + uint64_t(it.p_in) * (it.p_in + 1) > src_len ||
            it.dst + uint64_t(it.p_out) * (it.p_out + 1) > dst_len) {
            std::lock_guard<std::mutex> lock(ctx->mu);
            return fail(ctx, SFMM_ERR_LOGIC, "solidfmm: item outside its pool");
        }
    }
    hipError_t e = hipSetDevice(ctx->device);
    if (e == hipSuccess) e = grow_dev(st->d_src, st->d_src_cap, std::max<size_t>(src_len, 2), sizeof(double));
    if (e == hipSuccess) e = grow_dev(st->d_dst, st->d_dst_cap, std::max<size_t>(dst_len, 2), sizeof(double));
    if (e != hipSuccess) {
        std::lock_guard<std::mutex> lock(ctx->mu);
        return fail(ctx, SFMM_ERR_ALLOC, std::string("hipMalloc: ") + hipGetErrorString(e));
    }
    // copies in first, so the host-side plan preparation overlaps them
    e = hipMemcpyAsync(st->d_src, st->h_src, src_len * sizeof(double), hipMemcpyHostToDevice, st->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(st->d_dst, st->h_dst, dst_len * sizeof(double), hipMemcpyHostToDevice, st->stream);
    if (e != hipSuccess) {
        (void)hipStreamSynchronize(st->stream);
        std::lock_guard<std::mutex> lock(ctx->mu);
        return hip_fail(ctx, e, "hipMemcpyAsync");
    }
    sfmm_plan *plan = nullptr;
    int rc = sfmm_plan_prepare(ctx, op, items, n, &plan);
    if (rc == SFMM_OK && plan->n_items) {
        const size_t bytes = sfmm_plan_device_bytes(plan);
        e = grow_dev(st->d_plan, st->d_plan_cap, bytes, 1);
        if (e != hipSuccess) {
            std::lock_guard<std::mutex> lock(ctx->mu);
            rc = fail(ctx, SFMM_ERR_ALLOC, std::string("hipMalloc: ") + hipGetErrorString(e));
        }
        const size_t hb = sfmm_plan_host_bytes(plan);
        if (rc == SFMM_OK && grow_host(st->h_plan, st->h_plan_cap, (hb + 7) / 8) != hipSuccess) {
            std::lock_guard<std::mutex> lock(ctx->mu);
            rc = fail(ctx, SFMM_ERR_ALLOC, "hipHostMalloc: plan staging");
        }
        if (rc == SFMM_OK)
            rc = sfmm_plan_bind_staged(plan, st->d_plan, st->d_plan_cap, st->h_plan, st->h_plan_cap * sizeof(double),
                                       st->stream);
        if (rc == SFMM_OK) rc = sfmm_plan_execute(plan, st->d_src, src_len, st->d_dst, dst_len, st->stream);
        if (rc == SFMM_OK) {
            e = hipMemcpyAsync(st->h_dst, st->d_dst, dst_len * sizeof(double), hipMemcpyDeviceToHost, st->stream);
            if (e != hipSuccess) {
                std::lock_guard<std::mutex> lock(ctx->mu);
                rc = hip_fail(ctx, e, "hipMemcpyAsync");
            }
        }
    }
    // the plan's tables are read by the bind copy: free them only after it
    e = hipStreamSynchronize(st->stream);
    if (plan) sfmm_plan_destroy(plan);
    if (rc == SFMM_OK && e != hipSuccess) {
        std::lock_guard<std::mutex> lock(ctx->mu);
        rc = hip_fail(ctx, e, "hipStreamSynchronize");
    }
    return rc;
}

int sfmm_translate_host(sfmm_context *ctx, int op, const sfmm_item *items, size_t n, const double *src_pool,
                        size_t src_len, double *dst_pool, size_t dst_len)
{
    if (!ctx) return SFMM_ERR_LOGIC;
    for (size_t i = 0; i < n; ++i) {
        const sfmm_item &it = items[i];
        if (it.src + uint64_t(it.p_in) * (it.p_in + 1) > src_len ||
            it.dst + uint64_t(it.p_out) * (it.p_out + 1) > dst_len) {
            std::lock_guard<std::mutex> lock(ctx->mu);
            return fail(ctx, SFMM_ERR_LOGIC, "solidfmm: item outside its pool");
        }
    }
    if (n == 0) return SFMM_OK;
    // one staging object per context for this blocking convenience; calls
    // on one context are serialised here
    std::lock_guard<std::mutex> guard(ctx->host_mu);
    if (!ctx->host_staging) {
        int rc = sfmm_staging_create(ctx, &ctx->host_staging);
        if (rc != SFMM_OK) return rc;
    }
    sfmm_staging *st = ctx->host_staging;
    double *hs, *hd;
    int rc = sfmm_staging_reserve(st, src_len, dst_len, &hs, &hd);
    if (rc != SFMM_OK) return rc;
    parallel_copy(hs, src_pool, src_len);
    parallel_copy(hd, dst_pool, dst_len);
    rc = sfmm_staging_translate(st, op, items, n, src_len, dst_len);
    if (rc == SFMM_OK) parallel_copy(dst_pool, hd, dst_len);
    return rc;
}

struct sfmm_cellset {
    sfmm_context *ctx = nullptr;
    sfmm_cell *d_cells = nullptr;
    size_t n_cells = 0, n_points = 0, pool_extent = 0;
    int dim = 1;
};

int sfmm_cellset_create(sfmm_context *ctx, const sfmm_cell *cells, size_t n, size_t n_points, sfmm_cellset **out)
{
    if (!ctx || !out) return SFMM_ERR_LOGIC;
    std::lock_guard<std::mutex> lock(ctx->mu);
    *out = nullptr;
    if (n && !cells) return fail(ctx, SFMM_ERR_LOGIC, "solidfmm: null cell list");
    size_t extent = 0;
    const uint32_t dim = n ? std::max(cells[0].dim, 1u) : 1u;
    if (dim > uint32_t(sfmm::kMaxParticleDim))
        return fail(ctx, SFMM_ERR_LOGIC, "solidfmm: particle expansions of dimension 1..4 only");
    for (size_t i = 0; i < n; ++i) {
        const sfmm_cell &c = cells[i];
        if (c.order > ctx->P) return fail(ctx, SFMM_ERR_RANGE, "solidfmm: cell order exceeding operator data.");
        if (c.offset & 1u) return fail(ctx, SFMM_ERR_LOGIC, "solidfmm: expansion offsets must be even");
        if (uint64_t(c.first) + c.count > n_points) return fail(ctx, SFMM_ERR_LOGIC, "solidfmm: cell points out of range");
        if (std::max(c.dim, 1u) != dim) return fail(ctx, SFMM_ERR_LOGIC, "solidfmm: cells of one set must share dim");
        extent = std::max<size_t>(extent, c.offset + size_t(dim) * c.order * (c.order + 1));
    }
    sfmm_cellset *set = new (std::nothrow) sfmm_cellset;
    if (!set) return fail(ctx, SFMM_ERR_ALLOC, "solidfmm: out of host memory");
    set->ctx = ctx;
    set->n_cells = n;
    set->n_points = n_points;
    set->pool_extent = extent;
    set->dim = int(dim);
    hipError_t e = hipSetDevice(ctx->device);
    if (e == hipSuccess && n) {
        e = hipMalloc(&set->d_cells, n * sizeof(sfmm_cell));
        if (e == hipSuccess) e = hipMemcpy(set->d_cells, cells, n * sizeof(sfmm_cell), hipMemcpyHostToDevice);
    }
    if (e != hipSuccess) {
        if (set->d_cells) (void)hipFree(set->d_cells);
        delete set;
        return hip_fail(ctx, e, "cell upload");
    }
    *out = set;
    return SFMM_OK;
}

void sfmm_cellset_destroy(sfmm_cellset *set)
{
    if (!set) return;
    if (set->d_cells) (void)hipFree(set->d_cells);
    delete set;
}

int sfmm_p2m(sfmm_cellset *set, const double *points, size_t n_points, double *pool, size_t pool_len, void *stream)
{
    if (!set) return SFMM_ERR_LOGIC;
    sfmm_context *ctx = set->ctx;
    std::lock_guard<std::mutex> lock(ctx->mu);
    if (n_points < set->n_points || pool_len < set->pool_extent)
        return fail(ctx, SFMM_ERR_LOGIC, "solidfmm: points or pool shorter than the cell set needs");
    if (set->n_cells && (!points || !pool)) return fail(ctx, SFMM_ERR_LOGIC, "solidfmm: null pointer");
    hipError_t e = hipSetDevice(ctx->device);
    if (e == hipSuccess)
        e = sfmm::launch_p2m(points, set->d_cells, set->n_cells, pool, set->dim, static_cast<hipStream_t>(stream));
    if (e != hipSuccess) return hip_fail(ctx, e, "p2m launch");
    return SFMM_OK;
}

int sfmm_l2p(sfmm_cellset *set, const double *points, size_t n_points, const double *pool, size_t pool_len,
             double *phi, double *grad, void *stream)
{
    if (!set) return SFMM_ERR_LOGIC;
    sfmm_context *ctx = set->ctx;
    std::lock_guard<std::mutex> lock(ctx->mu);
    if (n_points < set->n_points || pool_len < set->pool_extent)
        return fail(ctx, SFMM_ERR_LOGIC, "solidfmm: points or pool shorter than the cell set needs");
    if (set->n_cells && (!points || !pool || !phi)) return fail(ctx, SFMM_ERR_LOGIC, "solidfmm: null pointer");
    hipError_t e = hipSetDevice(ctx->device);
    if (e == hipSuccess)
        e = sfmm::launch_l2p(points, set->d_cells, set->n_cells, pool, phi, grad, set->dim,
                             static_cast<hipStream_t>(stream));
    if (e != hipSuccess) return hip_fail(ctx, e, "l2p launch");
    return SFMM_OK;
}

int sfmm_synchronize(sfmm_context *ctx, void *stream)
{
    if (!ctx) return SFMM_ERR_LOGIC;
    hipError_t e = hipStreamSynchronize(static_cast<hipStream_t>(stream));
    if (e != hipSuccess) return hip_fail(ctx, e, "hipStreamSynchronize");
    return SFMM_OK;
}

}  // extern "C"
