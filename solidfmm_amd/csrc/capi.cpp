// C ABI of the translation engine (include/solidfmm_gpu.h): contexts,
// validated plans, launches and the host-pool convenience path.
//
// Validation follows the reference's checked entry points
// (translations.cpp:634-1001): every item is checked before any work is
// enqueued, so a failing call writes nothing.  Targets that repeat within a
// list are applied in list order: item k of a target goes to "round" k, and
// rounds run one after another on the stream, each round a set of launches
// whose targets are pairwise distinct.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/solidfmm_gpu.h"
#include "translate.hpp"

struct sfmm_context {
    int device = 0;
    size_t P = 0;
    int precision = SFMM_FP64;
    int family = SFMM_KERNELS_AUTO;
    std::string err;
    std::mutex mu;
};

struct sfmm_plan {
    struct Launch {
        uint32_t p_in, p_out;
        size_t offset, count;
    };
    sfmm_context *ctx = nullptr;
    int op = 0;
    int family = SFMM_KERNELS_AUTO;
    sfmm_item *d_items = nullptr;
    size_t n_items = 0;
    size_t src_extent = 0, dst_extent = 0;
    std::vector<Launch> launches;
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

// Validation and grouping.  Items with an order of zero contribute nothing
// (the reference pads such lanes with zeros and writes them to its trash
// buffer, microkernel_generic.cpp:340-380) and are dropped.
int build_groups(sfmm_context *ctx, int op, const sfmm_item *items, size_t n, std::vector<sfmm_item> &ordered,
                 std::vector<sfmm_plan::Launch> &launches)
{
    if (op < SFMM_M2M || op > SFMM_L2L) return fail(ctx, SFMM_ERR_LOGIC, "solidfmm: unknown operator");
    if (n && !items) return fail(ctx, SFMM_ERR_LOGIC, "solidfmm: null item list");
    struct Key {
        uint32_t round, p_in, p_out;
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
        if (!sfmm::have_kernel(op, it.p_in, it.p_out, ctx->family))
            return fail(ctx, SFMM_ERR_UNSUPPORTED,
                        std::string("solidfmm: no kernel for ") + op_name(op) + " with p_in=" +
                            std::to_string(it.p_in) + ", p_out=" + std::to_string(it.p_out));
        uint32_t round = seen[it.dst]++;
        keys.push_back({round, it.p_in, it.p_out, i});
    }
    std::stable_sort(keys.begin(), keys.end(), [](const Key &a, const Key &b) {
        if (a.round != b.round) return a.round < b.round;
        if (a.p_in != b.p_in) return a.p_in < b.p_in;
        return a.p_out < b.p_out;
    });
    ordered.clear();
    ordered.reserve(keys.size());
    launches.clear();
    for (size_t k = 0; k < keys.size(); ++k) {
        const Key &key = keys[k];
        if (launches.empty() || k == 0 || keys[k - 1].round != key.round || keys[k - 1].p_in != key.p_in ||
            keys[k - 1].p_out != key.p_out)
            launches.push_back({key.p_in, key.p_out, ordered.size(), 0});
        ordered.push_back(items[key.index]);
        launches.back().count++;
    }
    return SFMM_OK;
}

}  // namespace

extern "C" {

size_t sfmm_max_order(void) { return sfmm::max_kernel_order(); }

int sfmm_create(sfmm_context **out, int device, size_t P, int precision)
{
    if (!out) return SFMM_ERR_LOGIC;
    *out = nullptr;
    if (P == 0) return SFMM_ERR_LOGIC;
    if (P > sfmm::max_kernel_order()) return SFMM_ERR_OVERFLOW;
    if (precision != SFMM_FP64) return SFMM_ERR_UNSUPPORTED;
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
    plan->family = ctx->family;
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
    if (plan->d_items) (void)hipFree(plan->d_items);
    delete plan;
}

size_t sfmm_plan_launches(const sfmm_plan *plan) { return plan ? plan->launches.size() : 0; }

void sfmm_plan_extents(const sfmm_plan *plan, size_t *src_len, size_t *dst_len)
{
    if (src_len) *src_len = plan ? plan->src_extent : 0;
    if (dst_len) *dst_len = plan ? plan->dst_extent : 0;
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
    for (const auto &L : plan->launches) {
        e = sfmm::launch_translate(plan->op, L.p_in, L.p_out, plan->d_items + L.offset, L.count, src_pool, dst_pool,
                                   static_cast<hipStream_t>(stream), plan->family);
        if (e != hipSuccess) return hip_fail(ctx, e, "kernel launch");
    }
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
    if (!sfmm::have_kernel(op, p_in, p_out, ctx->family))
        return fail(ctx, SFMM_ERR_UNSUPPORTED, "solidfmm: no kernel for orders");
    hipError_t e = hipSetDevice(ctx->device);
    if (e != hipSuccess) return hip_fail(ctx, e, "hipSetDevice");
    e = sfmm::launch_translate(op, p_in, p_out, dev_items, n, src_pool, dst_pool, static_cast<hipStream_t>(stream),
                               ctx->family);
    if (e != hipSuccess) return hip_fail(ctx, e, "kernel launch");
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
