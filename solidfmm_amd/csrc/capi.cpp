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
#include <cstring>
#include <mutex>
#include <new>
#include <string>
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
};

struct sfmm_plan {
    struct Launch {
        sfmm::KernelChoice k;
        size_t offset, count;
    };
    sfmm_context *ctx = nullptr;
    int op = 0;
    sfmm_item *d_items = nullptr;
    size_t n_items = 0;
    size_t src_extent = 0, dst_extent = 0;
    std::vector<Launch> launches;
    // captured graph (sfmm_plan_capture)
    hipGraphExec_t exec = nullptr;
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
// contexts: the precision lane kernel for square M2L up to kPrecLaneMax
// (kPrecLaneMaxDW for double-word FP32) unless the generic family is
// forced, else the precision generic kernel.
bool choose(const sfmm_context *ctx, int op, uint32_t p_in, uint32_t p_out, int &family)
{
    const uint32_t pb = std::max(p_in, p_out);
    if (ctx->precision == SFMM_FP64) {
        if (!sfmm::have_kernel(op, p_in, p_out, ctx->family)) return false;
        family = family_for(ctx, op, pb);
        return true;
    }
    if (pb > sfmm::kPrecMaxOrder) return false;
    const uint32_t lane_max = ctx->precision == SFMM_SPLIT32 ? sfmm::kPrecLaneMaxDW : sfmm::kPrecLaneMax;
    const bool lane_ok = op == SFMM_M2L && p_in == p_out && pb <= lane_max;
    switch (ctx->family) {
    case SFMM_KERNELS_LANE:
        family = sfmm::kLaneKernels;
        return lane_ok;
    case SFMM_KERNELS_MFMA: return false;
    case SFMM_KERNELS_GENERIC: family = sfmm::kGenericKernel; return true;
    default: family = lane_ok ? sfmm::kLaneKernels : sfmm::kGenericKernel; return true;
    }
}

// Validation and grouping.  Items with an order of zero contribute nothing
// (the reference pads such lanes with zeros and writes them to its trash
// buffer, microkernel_generic.cpp:340-380) and are dropped.
int build_groups(sfmm_context *ctx, int op, const sfmm_item *items, size_t n, std::vector<sfmm_item> &ordered,
                 std::vector<sfmm_plan::Launch> &launches)
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
    std::unordered_map<uint64_t, uint32_t> seen;
    seen.reserve(n * 2 + 1);
    for (size_t i = 0; i < n; ++i) {
        const sfmm_item &it = items[i];
        if (it.p_in > ctx->P || it.p_out > ctx->P)
            return fail(ctx, SFMM_ERR_RANGE,
                        std::string("solidfmm::") + op_name(op) + "(): orders exceeding operator data.");
        if ((it.src | it.dst) & 1u)
            return fail(ctx, SFMM_ERR_LOGIC, "solidfmm: expansion offsets must be even (16-byte aligned)");
        if (it.p_in == 0 || it.p_out == 0) continue;
        int fam = 0;
        if (!choose(ctx, op, it.p_in, it.p_out, fam))
            return fail(ctx, SFMM_ERR_UNSUPPORTED,
                        std::string("solidfmm: no kernel for ") + op_name(op) + " with p_in=" +
                            std::to_string(it.p_in) + ", p_out=" + std::to_string(it.p_out));
        const uint32_t pb = std::max(it.p_in, it.p_out);
        const uint32_t round = seen[it.dst]++;
        // the generic kernel takes every order in one launch
        keys.push_back({round, fam, fam == sfmm::kGenericKernel ? 0u : pb, it.p_in, it.p_out, i});
    }
    std::stable_sort(keys.begin(), keys.end(), [](const Key &a, const Key &b) {
        if (a.round != b.round) return a.round < b.round;
        if (a.family != b.family) return a.family < b.family;
        if (a.pb != b.pb) return a.pb < b.pb;
        if (a.p_in != b.p_in) return a.p_in < b.p_in;
        return a.p_out < b.p_out;
    });
    ordered.clear();
    ordered.reserve(keys.size());
    launches.clear();
    for (size_t k = 0; k < keys.size(); ++k) {
        const Key &key = keys[k];
        if (k == 0 || keys[k - 1].round != key.round || keys[k - 1].family != key.family ||
            keys[k - 1].pb != key.pb)
            launches.push_back({{key.family, key.pb, false, ctx->precision}, ordered.size(), 0});
        sfmm_plan::Launch &L = launches.back();
        if (key.family == sfmm::kGenericKernel) {
            L.k.pk = std::max({L.k.pk, key.p_in, key.p_out});
        } else if (key.p_in != key.pb || key.p_out != key.pb) {
            L.k.rt = true;
        }
        ordered.push_back(items[key.index]);
        L.count++;
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
    for (const auto &L : plan->launches) {
        hipError_t e = sfmm::launch_kernel(plan->op, L.k, plan->d_items + L.offset, L.count, src, dst, stream);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

// Device tables the kernels read, created before any capture so that no
// allocation happens inside one.
bool tables_ready() { return sfmm::mfma::device_table() && sfmm::mfma::device_factorials(); }

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
    sfmm_context *ctx = new (std::nothrow) sfmm_context;
    if (!ctx) return SFMM_ERR_ALLOC;
    ctx->device = device;
    ctx->P = P;
    ctx->precision = precision;
    *out = ctx;
    return SFMM_OK;
}

void sfmm_destroy(sfmm_context *ctx) { delete ctx; }

size_t sfmm_order(const sfmm_context *ctx) { return ctx ? ctx->P : 0; }

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

int sfmm_plan_create(sfmm_context *ctx, int op, const sfmm_item *items, size_t n, sfmm_plan **out)
{
    if (!ctx || !out) return SFMM_ERR_LOGIC;
    std::lock_guard<std::mutex> lock(ctx->mu);
    *out = nullptr;
    std::vector<sfmm_item> ordered;
    std::vector<sfmm_plan::Launch> launches;
    int rc = build_groups(ctx, op, items, n, ordered, launches);
    if (rc != SFMM_OK) return rc;
    hipError_t e = hipSetDevice(ctx->device);
    if (e != hipSuccess) return hip_fail(ctx, e, "hipSetDevice");
    sfmm_plan *plan = new (std::nothrow) sfmm_plan;
    if (!plan) return fail(ctx, SFMM_ERR_ALLOC, "solidfmm: out of host memory");
    plan->ctx = ctx;
    plan->op = op;
    plan->n_items = ordered.size();
    plan->launches = std::move(launches);
    for (const sfmm_item &it : ordered) {
        plan->src_extent = std::max<size_t>(plan->src_extent, it.src + size_t(it.p_in) * (it.p_in + 1));
        plan->dst_extent = std::max<size_t>(plan->dst_extent, it.dst + size_t(it.p_out) * (it.p_out + 1));
    }
    if (!ordered.empty()) {
        e = hipMalloc(&plan->d_items, ordered.size() * sizeof(sfmm_item));
        if (e != hipSuccess) {
            delete plan;
            return fail(ctx, SFMM_ERR_ALLOC, std::string("hipMalloc: ") + hipGetErrorString(e));
        }
        e = hipMemcpy(plan->d_items, ordered.data(), ordered.size() * sizeof(sfmm_item), hipMemcpyHostToDevice);
        if (e != hipSuccess) {
            (void)hipFree(plan->d_items);
            delete plan;
            return hip_fail(ctx, e, "hipMemcpy");
        }
    }
    *out = plan;
    return SFMM_OK;
}

void sfmm_plan_destroy(sfmm_plan *plan)
{
    if (!plan) return;
    drop_graph(plan);
    if (plan->d_items) (void)hipFree(plan->d_items);
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
    if (src_len < plan->src_extent || dst_len < plan->dst_extent)
        return fail(ctx, SFMM_ERR_LOGIC, "solidfmm: pool shorter than the plan's extent");
    hipError_t e = hipSetDevice(ctx->device);
    if (e != hipSuccess) return hip_fail(ctx, e, "hipSetDevice");
    e = enqueue(plan, src_pool, dst_pool, static_cast<hipStream_t>(stream));
    if (e != hipSuccess) return hip_fail(ctx, e, "kernel launch");
    return SFMM_OK;
}

int sfmm_plan_capture(sfmm_plan *plan, const double *src_pool, size_t src_len, double *dst_pool, size_t dst_len)
{
    if (!plan) return SFMM_ERR_LOGIC;
    sfmm_context *ctx = plan->ctx;
    std::lock_guard<std::mutex> lock(ctx->mu);
    if (plan->n_items && (!src_pool || !dst_pool)) return fail(ctx, SFMM_ERR_LOGIC, "solidfmm: null pool");
    if (src_len < plan->src_extent || dst_len < plan->dst_extent)
        return fail(ctx, SFMM_ERR_LOGIC, "solidfmm: pool shorter than the plan's extent");
    hipError_t e = hipSetDevice(ctx->device);
    if (e != hipSuccess) return hip_fail(ctx, e, "hipSetDevice");
    if (!tables_ready()) return fail(ctx, SFMM_ERR_ALLOC, "solidfmm: device tables");
    drop_graph(plan);
    hipStream_t s = nullptr;
    e = hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
    if (e != hipSuccess) return hip_fail(ctx, e, "hipStreamCreate");
    hipGraph_t graph = nullptr;
    e = hipStreamBeginCapture(s, hipStreamCaptureModeRelaxed);
    if (e == hipSuccess) {
        hipError_t le = enqueue(plan, src_pool, dst_pool, s);
        e = hipStreamEndCapture(s, &graph);
        if (le != hipSuccess) e = le;
    }
    if (e == hipSuccess) e = hipGraphInstantiate(&plan->exec, graph, nullptr, nullptr, 0);
    if (graph) (void)hipGraphDestroy(graph);
    (void)hipStreamDestroy(s);
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
    constexpr size_t kSample = 32768;
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
    size_t most = 0;
    uint32_t pmax = 1;
    for (uint32_t pb = 1; pb < uint32_t(kOrders); ++pb)
        if (!bucket[pb].empty()) {
            most = std::max(most, bucket[pb].size());
            pmax = pb;
        }
    if (most == 0) return SFMM_OK;
    const size_t block = size_t(pmax) * (pmax + 1);
    double *src = nullptr, *dst = nullptr;
    sfmm_item *d_items = nullptr;
    hipEvent_t ev[2] = {nullptr, nullptr};
    int rc = SFMM_OK;
    e = hipMalloc(&src, most * block * sizeof(double));
    if (e == hipSuccess) e = hipMalloc(&dst, most * block * sizeof(double));
    if (e == hipSuccess) e = hipMalloc(&d_items, most * sizeof(sfmm_item));
    if (e != hipSuccess) {
        rc = fail(ctx, SFMM_ERR_ALLOC, std::string("hipMalloc: ") + hipGetErrorString(e));
    } else {
        // 0x3F bytes: every double is ~4.8e-4 (normal numbers)
        e = hipMemsetAsync(src, 0x3F, most * block * sizeof(double), s);
        if (e == hipSuccess) e = hipMemsetAsync(dst, 0x3F, most * block * sizeof(double), s);
        if (e == hipSuccess) e = hipEventCreate(&ev[0]);
        if (e == hipSuccess) e = hipEventCreate(&ev[1]);
        if (e != hipSuccess) rc = hip_fail(ctx, e, "autotune setup");
    }
    for (uint32_t pb = 1; rc == SFMM_OK && pb < uint32_t(kOrders); ++pb) {
        std::vector<sfmm_item> &b = bucket[pb];
        if (b.empty()) continue;
        std::stable_sort(b.begin(), b.end(), [](const sfmm_item &x, const sfmm_item &y) {
            return x.p_in != y.p_in ? x.p_in < y.p_in : x.p_out < y.p_out;
        });
        bool rt = false;
        for (size_t i = 0; i < b.size(); ++i) {
            b[i].src = i * block;
            b[i].dst = i * block;
            rt = rt || b[i].p_in != pb || b[i].p_out != pb;
        }
        e = hipMemcpyAsync(d_items, b.data(), b.size() * sizeof(sfmm_item), hipMemcpyHostToDevice, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        if (e != hipSuccess) {
            rc = hip_fail(ctx, e, "autotune upload");
            break;
        }
        int best = 0;
        float best_ms = 0.0f;
        for (int fam : {int(sfmm::kLaneKernels), int(sfmm::kMfmaKernels), int(sfmm::kGenericKernel)}) {
            if (!sfmm::family_covers(fam, pb)) continue;
            const sfmm::KernelChoice k{fam, pb, rt};
            e = sfmm::launch_kernel(op, k, d_items, b.size(), src, dst, s);  // warm-up
            float ms = 1e30f;
            for (int r = 0; r < kReps && e == hipSuccess; ++r) {
                e = hipEventRecord(ev[0], s);
                if (e == hipSuccess) e = sfmm::launch_kernel(op, k, d_items, b.size(), src, dst, s);
                if (e == hipSuccess) e = hipEventRecord(ev[1], s);
                if (e == hipSuccess) e = hipEventSynchronize(ev[1]);
                float t = 0.0f;
                if (e == hipSuccess) e = hipEventElapsedTime(&t, ev[0], ev[1]);
                ms = std::min(ms, t);
            }
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
    if (d_items) (void)hipFree(d_items);
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

int sfmm_translate_host(sfmm_context *ctx, int op, const sfmm_item *items, size_t n, const double *src_pool,
                        size_t src_len, double *dst_pool, size_t dst_len)
{
    if (!ctx) return SFMM_ERR_LOGIC;
    // Bounds of every block against the host pools, before any copy.
    for (size_t i = 0; i < n; ++i) {
        const sfmm_item &it = items[i];
        if (it.src + uint64_t(it.p_in) * (it.p_in + 1) > src_len ||
            it.dst + uint64_t(it.p_out) * (it.p_out + 1) > dst_len)
            return fail(ctx, SFMM_ERR_LOGIC, "solidfmm: item outside its pool");
    }
    sfmm_plan *plan = nullptr;
    int rc = sfmm_plan_create(ctx, op, items, n, &plan);
    if (rc != SFMM_OK) return rc;
    if (plan->n_items == 0) {
        sfmm_plan_destroy(plan);
        return SFMM_OK;
    }
    double *d_src = nullptr, *d_dst = nullptr;
    hipError_t e = hipMalloc(&d_src, std::max<size_t>(src_len, 2) * sizeof(double));
    if (e == hipSuccess) e = hipMalloc(&d_dst, std::max<size_t>(dst_len, 2) * sizeof(double));
    if (e != hipSuccess) {
        if (d_src) (void)hipFree(d_src);
        sfmm_plan_destroy(plan);
        return fail(ctx, SFMM_ERR_ALLOC, std::string("hipMalloc: ") + hipGetErrorString(e));
    }
    e = hipMemcpy(d_src, src_pool, src_len * sizeof(double), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(d_dst, dst_pool, dst_len * sizeof(double), hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        rc = sfmm_plan_execute(plan, d_src, src_len, d_dst, dst_len, nullptr);
        if (rc == SFMM_OK) {
            e = hipMemcpy(dst_pool, d_dst, dst_len * sizeof(double), hipMemcpyDeviceToHost);
            if (e != hipSuccess) rc = hip_fail(ctx, e, "hipMemcpy");
        }
    } else {
        rc = hip_fail(ctx, e, "hipMemcpy");
    }
    (void)hipFree(d_src);
    (void)hipFree(d_dst);
    sfmm_plan_destroy(plan);
    return rc;
}

int sfmm_synchronize(sfmm_context *ctx, void *stream)
{
    if (!ctx) return SFMM_ERR_LOGIC;
    hipError_t e = hipStreamSynchronize(static_cast<hipStream_t>(stream));
    if (e != hipSuccess) return hip_fail(ctx, e, "hipStreamSynchronize");
    return SFMM_OK;
}

}  // extern "C"
