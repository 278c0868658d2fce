/*
 * solidfmm_gpu.h -- C ABI of the MI355X translation engine.
 *
 * Batched M2M / M2L / L2L translations of solid-harmonic expansions, run by
 * hand-written HIP kernels on one AMD Instinct MI355X (gfx950).  The C++
 * drop-in headers under include/solidfmm/ (the reference's public API) are
 * implemented on top of these entry points; other languages bind them
 * directly.
 *
 * Reference interfaces replaced (file:line in vorticle/solidfmm 1.5):
 *   solid<real> coefficient layout        solidfmm/solid.hpp:57-60
 *   translation_info<real>                solidfmm/translations.hpp:39-45
 *   operator_handle / buffer_handle       solidfmm/handles.hpp:33-110
 *   m2m / m2l / l2l (+ _unchecked)        solidfmm/translations.hpp:103-170
 *   multithreaded_translator<real>        solidfmm/translations.hpp:47-87
 *
 * Conventions
 *   - Order P means degrees n = 0..P-1; an expansion holds P(P+1) reals:
 *     Re C_n^m at n(n+1)+m, Im C_n^m at (n+1)^2+m (solid.hpp:59-60).
 *     Vector-valued expansions repeat that block per dimension.
 *   - A pool is a device array of reals holding many expansions; an item
 *     names its source and target by offset (in reals) into the source and
 *     target pools.  Offsets must be even (16-byte aligned blocks), which
 *     every packed pool of whole expansions satisfies.
 *   - Every operator accumulates: target += T(source).  The shift (x,y,z)
 *     is new centre minus old centre (doc/solidfmm.tex:1954).
 *   - Entry points return SFMM_OK or a negative SFMM_ERR_* code; the text
 *     of the last error on a context is sfmm_last_error().  No C++
 *     exception crosses this boundary.
 *   - Calls on one context are serialised internally; launches are
 *     asynchronous on the caller's HIP stream (NULL = the null stream).
 */
#ifndef SOLIDFMM_GPU_H
#define SOLIDFMM_GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Operators. */
enum { SFMM_M2M = 0, SFMM_M2L = 1, SFMM_L2L = 2 };

/* Arithmetic of the translation kernels (sfmm_create).  Pools are FP64 in
 * every mode; geometry (Euler angles, their powers, the r^n scale tables)
 * and the accumulation into the target are FP64.  The reduced modes run the
 * bulk arithmetic (rotations, swaps, z-translation) in FP32 or in
 * double-word FP32 (hi/lo float pairs, about 48 bits) and support orders up
 * to 24 (SFMM_ERR_OVERFLOW above). */
enum {
    SFMM_FP64 = 0,   /* FP64 throughout (the reference's double path) */
    SFMM_FP32 = 1,   /* FP32 bulk arithmetic */
    SFMM_SPLIT32 = 2 /* double-word FP32 bulk arithmetic */
};

/* Kernel families (sfmm_set_kernels, sfmm_set_tuning).  Lane and MFMA
 * kernels exist per order bound P; lists whose items have other orders
 * p_in, p_out <= P run their per-item-order variants. */
enum {
    SFMM_KERNELS_AUTO = 0, /* per order bucket: the autotuned family, else the default */
    SFMM_KERNELS_LANE = 1, /* one lane per translation, FP64 VALU (orders <= 12) */
    SFMM_KERNELS_MFMA = 2, /* 16 translations per wave, swaps on MFMA (orders <= 24; FP32 contexts: square M2L) */
    SFMM_KERNELS_GENERIC = 3 /* one workgroup per translation, any orders <= 86 */
};

/* Error codes (mirror the reference's exception types). */
enum {
    SFMM_OK = 0,
    SFMM_ERR_LOGIC = -1,       /* std::logic_error: bad argument, dimension mismatch, P == 0 */
    SFMM_ERR_RANGE = -2,       /* std::out_of_range: order above the context's order */
    SFMM_ERR_OVERFLOW = -3,    /* std::overflow_error: order too large for the type */
    SFMM_ERR_ALLOC = -4,       /* std::bad_alloc: host or device allocation failed */
    SFMM_ERR_HIP = -5,         /* std::runtime_error: a HIP runtime call failed */
    SFMM_ERR_UNSUPPORTED = -6  /* valid request this library has no kernel for */
};

/* One translation.  48 bytes, naturally aligned.  Replaces
 * translation_info<real> (translations.hpp:39-45): pointers become pool
 * offsets, and the orders travel with the item. */
typedef struct sfmm_item {
    uint64_t src;   /* offset (reals) of the source block in the source pool */
    uint64_t dst;   /* offset (reals) of the target block in the target pool */
    uint32_t p_in;  /* order of the source expansion */
    uint32_t p_out; /* order of the target expansion */
    double x, y, z; /* shift: new centre minus old centre */
} sfmm_item;

typedef struct sfmm_context sfmm_context;
typedef struct sfmm_plan sfmm_plan;

/* Highest order the library supports (FP64): 86, the reference's limit. */
size_t sfmm_max_order(void);

/* Context on HIP device `device` for orders up to P, with the arithmetic
 * `precision` (SFMM_FP64, SFMM_FP32 or SFMM_SPLIT32).  Replaces
 * operator_handle<real>(P) (handles.hpp:40, handles.cpp:42-45): P == 0 is
 * SFMM_ERR_LOGIC, P above sfmm_max_order() (24 for the reduced modes) is
 * SFMM_ERR_OVERFLOW.  Reduced-precision contexts run square M2L of orders
 * <= 16 (FP32) or <= 10 (split) in a one-lane-per-translation kernel,
 * FP32 square M2L of orders 17..24 on the matrix cores (FP32 MFMA, the
 * kernel family SFMM_KERNELS_MFMA selects at any order up to 24), and
 * everything else in a one-workgroup-per-translation kernel (no
 * autotuning). */
int sfmm_create(sfmm_context **ctx, int device, size_t P, int precision);
void sfmm_destroy(sfmm_context *ctx);
size_t sfmm_order(const sfmm_context *ctx);
const char *sfmm_last_error(const sfmm_context *ctx);

/* Kernel family used by plans created afterwards on this context (a tuning
 * knob with no counterpart in the reference; results agree to rounding). */
int sfmm_set_kernels(sfmm_context *ctx, int family);

/* Plan: a validated, device-resident translation list.  Replaces the
 * validation and queueing of m2m/m2l/l2l (translations.cpp:634-1001): every
 * item is checked before anything is written (orders <= context order,
 * else SFMM_ERR_RANGE; odd offsets SFMM_ERR_LOGIC).  Items are bucketed by
 * max(p_in, p_out) and the family chosen for that bucket (sfmm_set_kernels,
 * else the autotuned family, else the default); each bucket is one launch,
 * its items sorted by (p_in, p_out).  Items sharing a target: if no target
 * occurs more than 4 times they are applied in list order by rounds (the
 * k-th item of a target runs in round k); otherwise (e.g. M2L interaction
 * lists) every item writes its own scratch block and a segmented reduction
 * adds each target's blocks, in list order, into the target once.  Results
 * are bitwise reproducible either way. */
int sfmm_plan_create(sfmm_context *ctx, int op, const sfmm_item *host_items, size_t n, sfmm_plan **plan);
/* Two-step plan creation for request loops that carve device memory out of
 * one arena under an HBM budget (no counterpart in the reference).
 * sfmm_plan_prepare validates and groups like sfmm_plan_create but touches
 * no device and holds the context's lock only to read its settings, so
 * several host threads can prepare plans while kernels run.
 * sfmm_plan_device_bytes is the device memory the plan needs (its items,
 * plus reduction tables and scratch when targets repeat).  sfmm_plan_bind
 * then copies the plan's tables into caller device memory `dev_mem` (at
 * least sfmm_plan_device_bytes bytes, 256-byte aligned) asynchronously on
 * `stream`; the plan allocates and frees no device memory, so nothing
 * waits for the device.  Execute a bound plan on the same stream (or after
 * the copy), and keep `dev_mem` to the plan until its last launch has
 * completed. */
int sfmm_plan_prepare(sfmm_context *ctx, int op, const sfmm_item *host_items, size_t n, sfmm_plan **plan);
size_t sfmm_plan_device_bytes(const sfmm_plan *plan);
int sfmm_plan_bind(sfmm_plan *plan, void *dev_mem, size_t dev_bytes, void *stream);
/* As sfmm_plan_bind, but the tables go through caller-owned pinned host
 * memory `host_mem` (at least sfmm_plan_host_bytes bytes): they are copied
 * there on the host and from there to the device asynchronously, so the
 * call returns at once (a copy from the plan's own pageable memory makes
 * the HIP runtime stage it synchronously).  host_mem must stay unchanged
 * until the stream has passed the copy. */
size_t sfmm_plan_host_bytes(const sfmm_plan *plan);
int sfmm_plan_bind_staged(sfmm_plan *plan, void *dev_mem, size_t dev_bytes, void *host_mem, size_t host_bytes,
                          void *stream);
/* The two halves of sfmm_plan_bind_staged, for a preparation thread and a
 * launching thread: copy the tables into pinned host memory (host only),
 * then bind from there (an asynchronous copy on `stream`; host_tables must
 * stay unchanged until the stream has passed it). */
int sfmm_plan_copy_tables(const sfmm_plan *plan, void *host_mem, size_t host_bytes);
int sfmm_plan_bind_pinned(sfmm_plan *plan, void *dev_mem, size_t dev_bytes, const void *host_tables, void *stream);
void sfmm_plan_destroy(sfmm_plan *plan);
size_t sfmm_plan_launches(const sfmm_plan *plan);

/* Reals of each pool the plan reads or writes: the largest offset+size. */
void sfmm_plan_extents(const sfmm_plan *plan, size_t *src_len, size_t *dst_len);

/* Launch i of a plan: kernel family, kernel order bound, whether items
 * carry their own orders (1) or all have the bound as orders (0), and the
 * number of items. */
int sfmm_plan_launch_info(const sfmm_plan *plan, size_t i, int *family, uint32_t *order, int *per_item_orders,
                          size_t *count);

/* Run a plan: target pool += T(source pool), asynchronously on `stream`.
 * Pools are device pointers of src_len / dst_len reals; a pool shorter than
 * the plan's extent is SFMM_ERR_LOGIC and nothing is launched.  Launches
 * that write pairwise distinct targets (the order buckets of one round) run
 * on `stream` and up to three streams of the context forked from it and
 * joined back by events, so one bucket's tail overlaps the next's start;
 * everything enqueued on `stream` afterwards sees the whole plan done.  Replaces
 * multithreaded_translator<real>::m2m/m2l/l2l (translations.hpp:75-81) for
 * device-resident pools. */
int sfmm_plan_execute(sfmm_plan *plan, const double *src_pool, size_t src_len, double *dst_pool, size_t dst_len,
                      void *stream);

/* Capture the plan's launches on these pools into a HIP graph kept with the
 * plan (replacing an earlier capture); sfmm_plan_replay then enqueues the
 * whole plan as one graph launch on any stream.  The buckets of a round
 * become parallel branches of the graph, as in sfmm_plan_execute.  The pools are bound at
 * capture: capture again to run on other pools.  No counterpart in the
 * reference (a launch-overhead tool for lists that are replayed, such as
 * the same interaction lists every time step). */
int sfmm_plan_capture(sfmm_plan *plan, const double *src_pool, size_t src_len, double *dst_pool, size_t dst_len);
int sfmm_plan_replay(sfmm_plan *plan, void *stream);

/* Autotuner: for each order bucket max(p_in, p_out) present in the list,
 * time the plan every kernel family that covers it would get for a sample
 * of the bucket's items (up to 262,144 in list order, with the list's
 * repeated targets, so runs and reductions are timed as they will run) and
 * keep the fastest for plans created afterwards on this context.  The
 * sample runs in scratch pools the call allocates and frees (two of the
 * largest sample x P(P+1) doubles over the buckets: 1.8 GB for a full
 * sample at order 20), on `stream`; it blocks until done.  sfmm_get_tuning reports the
 * family used per order bucket (index = order, entry 0 unused);
 * sfmm_set_tuning sets one entry (SFMM_KERNELS_AUTO restores the default).
 * Plays the role of the reference's schedule_by_measuring_throughput
 * (translations.hpp:72-73): a measured choice of how to run a list. */
int sfmm_autotune(sfmm_context *ctx, int op, const sfmm_item *host_items, size_t n, void *stream);
int sfmm_get_tuning(sfmm_context *ctx, int op, int *family_by_order, size_t len);
int sfmm_set_tuning(sfmm_context *ctx, int op, uint32_t order, int family);

/* Direct launch over device-resident items the caller has already grouped:
 * all items have orders (p_in, p_out) and pairwise distinct targets.
 * Replaces m2m/m2l/l2l_unchecked (translations.hpp:154-170): nothing is
 * validated. */
int sfmm_launch(sfmm_context *ctx, int op, uint32_t p_in, uint32_t p_out, const sfmm_item *dev_items,
                size_t n, const double *src_pool, double *dst_pool, void *stream);

/* Allocation-free translation of a list over device pools.  No counterpart
 * in the reference: for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  It runs what sfmm_plan_create + sfmm_plan_execute would run
 * on the same list (same validation, kernels and results), but the plan's
 * device tables and, when targets repeat, its per-item scratch blocks live
 * in a static device array of the library, SFMM_NOALLOC_SCRATCH_BYTES
 * (64 MiB) per GPU, set up by sfmm_create and held until the process ends;
 * the call itself allocates no device memory.  A list that needs more
 * scratch than that is SFMM_ERR_LOGIC and nothing is enqueued (a list of
 * distinct targets needs 48 bytes per item plus 256-byte alignment; one
 * with repeated targets about P(P+1) 8 + 64 bytes per item more).  The
 * tables are copied from host memory on `stream` before the kernels.  Every
 * call on one GPU shares that GPU's array: calls that overlap on one GPU
 * (other streams, other threads, other contexts) race and must be
 * serialised by the caller.  The array keeps the last call's contents;
 * nothing in it needs to start at zero. */
#define SFMM_NOALLOC_SCRATCH_BYTES ((size_t)64 << 20)
int sfmm_translate_noalloc(sfmm_context *ctx, int op, const sfmm_item *host_items, size_t n, const double *src_pool,
                           size_t src_len, double *dst_pool, size_t dst_len, void *stream);

/* Host-pool convenience: copies the pools to the device, runs the list and
 * copies the target pool back (blocking).  The drop-in C++ API and the
 * classic single-thread entry points use this path. */
int sfmm_translate_host(sfmm_context *ctx, int op, const sfmm_item *items, size_t n, const double *src_pool,
                        size_t src_len, double *dst_pool, size_t dst_len);

/* Diagnostics: the folded swap block of degree n (part 0: real, 1:
 * imaginary; operator_data.cpp:47-147) as the kernels use it, row-major
 * (n+1) x (n+1) into out.  Host only; needs no device. */
int sfmm_swap_block(int n, int part, double *out);

/* Particle-side operators, batched over cells (the steps on either side of
 * the translations; in the reference user code over harmonics::R / dR and
 * solid::fmadd / dot, harmonics.cpp:31-174, solid.cpp:40-105).  A cell
 * names its expansion centre, its points [first, first + count) and its
 * expansion block (offset in reals, order) in a pool.  48 bytes.
 * Vector-valued expansions (dim D = 1..4, as solid<real>(P, dim),
 * solid.hpp:28-66): the block holds D consecutive P(P+1)-real blocks, and
 * every point is a row of 3 + D doubles (x, y, z, q_0 .. q_{D-1}); D = 1
 * gives (x, y, z, q) quadruples.  All cells of a set share D. */
typedef struct sfmm_cell {
    double x, y, z;   /* expansion centre */
    uint64_t offset;  /* block offset (reals) in the pool, even */
    uint32_t first;   /* first point */
    uint32_t count;   /* number of points */
    uint32_t order;   /* expansion order P (degrees 0..P-1) */
    uint32_t dim;     /* D, 0 meaning 1 */
} sfmm_cell;

typedef struct sfmm_cellset sfmm_cellset;

/* Validated, device-resident cell list (like a plan, reusable every step):
 * orders <= the context's order (SFMM_ERR_RANGE), point ranges within
 * n_points, even offsets and one dim in 1..4 for all cells
 * (SFMM_ERR_LOGIC). */
int sfmm_cellset_create(sfmm_context *ctx, const sfmm_cell *host_cells, size_t n_cells, size_t n_points,
                        sfmm_cellset **set);
void sfmm_cellset_destroy(sfmm_cellset *set);

/* P2M: block d of cell c += sum_i q_{i,d} R(P_c, p_i - centre_c)
 * (tests/test_m2l.cpp:70 forms M = R(P, x_s - x_a) for a unit charge).
 * n_points counts points (rows of 3 + D doubles). */
int sfmm_p2m(sfmm_cellset *set, const double *dev_points, size_t n_points, double *dev_pool, size_t pool_len,
             void *stream);

/* L2P: phi[i D + d] += dot(L_{c,d}, R(P_c, p_i - centre_c)) and, if
 * dev_grad is not NULL, grad[3 (i D + d) + k] += dot(L_{c,d}, d/dx_k R(...))
 * (tests/test_m2l.cpp:112-116).  phi holds D n_points reals and grad
 * 3 D n_points. */
int sfmm_l2p(sfmm_cellset *set, const double *dev_points, size_t n_points, const double *dev_pool, size_t pool_len,
             double *dev_phi, double *dev_grad, void *stream);

/* Diagnostics with no counterpart in the reference: the device's issue
 * rate (flop/s, a multiply-add counted as 2) of the arithmetic the
 * precision modes trade against, operands in registers, eight waves per
 * SIMD.  kind 0: FP64 FMA (VALU), 1: FP32 FMA (VALU, not packed), 2: the
 * double-word FP32 multiply-add of SFMM_SPLIT32, 3: FP64 MFMA 16x16x4,
 * 4: FP32 MFMA 16x16x4.  Allocates and frees a small output buffer. */
int sfmm_probe_arith(int device, int kind, double *flops_per_s);

/* HIP device of a context (-1 for NULL). */
int sfmm_device(const sfmm_context *ctx);

/* Continuous-batching server for translation requests (native loop; the
 * reference has no counterpart: multithreaded_translator,
 * translations.hpp:47-87, serialises whole batches with a mutex).  The
 * server keeps the multipole pool (M2M source and target, M2L source) and
 * the local pool (M2L target, L2L source and target) of `ncells` cells and
 * runs requests on `slots` streams of its own:
 *   - sfmm_server_submit copies a request (operator, items, the cells it
 *     reads and writes) and queues it in arrival order; `workers` host
 *     threads prepare plans (sfmm_plan_prepare) while the GPU runs others,
 *     at a lower CPU priority (niceness 5) than the server's scheduling
 *     thread, which should have a core to itself;
 *   - the oldest waiting request is admitted when it is prepared, a slot is
 *     free, its device memory fits in the server's arena of `arena_bytes`,
 *     and it has no hazard with a request in flight (it writes no cell one
 *     reads or writes; it reads no cell one writes).  Admission is in
 *     arrival order, so the pools end as if the requests ran one by one;
 *   - admitted requests are bound into the arena and launched between two
 *     events; a scheduler thread polls completions.
 * Times: sfmm_server_request_info fills times[5]: arrival and admission
 * (host clock) and completion (GPU clock), in ms since the server's
 * creation, the request's own GPU time, and when the loop noticed the
 * completion (host clock; -1 for a request that never ran); rc is its status (a request whose preparation
 * failed, or that can never fit in the arena, finishes with an error code
 * and does not run).  counters: max in flight, arena peak and size, polls
 * blocked on preparation, slot, arena, hazard.  Destroy drains first. */
typedef struct sfmm_server sfmm_server;
int sfmm_server_create(sfmm_context *ctx, double *m_pool, size_t m_len, double *l_pool, size_t l_len, size_t ncells,
                       int slots, size_t arena_bytes, int workers, sfmm_server **sv);
int sfmm_server_submit(sfmm_server *sv, int op, const sfmm_item *items, size_t n, const uint32_t *src_cells,
                       size_t n_src, const uint32_t *dst_cells, size_t n_dst, uint64_t *rid);
/* As sfmm_server_submit, but the request arrives at `arrival_ms` on the
 * server's clock (sfmm_server_clock_ms): it is held, not prepared and not
 * admitted before then.  A time in the past means now.  For replaying
 * arrival processes without a host thread timing the submissions. */
int sfmm_server_submit_at(sfmm_server *sv, int op, const sfmm_item *items, size_t n, const uint32_t *src_cells,
                          size_t n_src, const uint32_t *dst_cells, size_t n_dst, double arrival_ms, uint64_t *rid);
int sfmm_server_drain(sfmm_server *sv);
int sfmm_server_request_info(sfmm_server *sv, uint64_t rid, double *times, int *rc);
int sfmm_server_counters(sfmm_server *sv, uint64_t *out);
double sfmm_server_clock_ms(sfmm_server *sv);
void sfmm_server_destroy(sfmm_server *sv);

/* Host-pool staging for callers whose expansions live in host memory (the
 * drop-in's solid* lists, FFI callers with host arrays).  A staging object
 * owns pinned host pools, device pools, plan memory and a stream, all
 * grow-only and reused across calls, so a call allocates nothing once
 * warm.  Use: sfmm_staging_reserve returns the pinned source and target
 * pools (contents undefined); the caller packs its blocks into them (the
 * solid.hpp layout, offsets as in sfmm_item); sfmm_staging_translate then
 * runs dst += T(src) over them -- copies in, plan, kernels, copy out -- and
 * returns with the result in the pinned target pool.  The host-side plan
 * preparation overlaps the copies in.  One staging object per calling
 * thread; objects on one context may run concurrently (disjoint pools).
 * Replaces the per-thread workspace of the reference
 * (threadlocal_buffer.hpp:30-70) for the host-pointer path. */
typedef struct sfmm_staging sfmm_staging;
int sfmm_staging_create(sfmm_context *ctx, sfmm_staging **st);
void sfmm_staging_destroy(sfmm_staging *st);
int sfmm_staging_reserve(sfmm_staging *st, size_t src_len, size_t dst_len, double **h_src, double **h_dst);
int sfmm_staging_translate(sfmm_staging *st, int op, const sfmm_item *items, size_t n, size_t src_len,
                           size_t dst_len);

/* Smallest enclosing ball (host code, no context): of n points given as
 * coordinate arrays, or of n balls given as (x, y, z, r) quadruples;
 * out = (x, y, z, r).  n == 0 or a negative radius is SFMM_ERR_LOGIC.
 * Replaces solidfmm::bounding_sphere (sphere.hpp:38-42, sphere.cpp:546-627). */
int sfmm_bounding_sphere(size_t n, const double *x, const double *y, const double *z, double *out);
int sfmm_bounding_sphere_of_spheres(size_t n, const double *spheres, double *out);

/* Batched form for trees: the smallest ball of points[first[s] ..
 * first[s] + count[s]) for every set s, points being rows of `stride`
 * doubles whose first three are x, y, z; out holds n_sets quadruples.
 * Runs on `threads` host threads (0: all cores).  Empty sets and ranges
 * beyond n_points are SFMM_ERR_LOGIC.  No counterpart in the reference
 * (its users loop over the single-set form). */
int sfmm_bounding_spheres_of_ranges(const double *points, size_t n_points, size_t stride, const uint64_t *first,
                                    const uint64_t *count, size_t n_sets, double *out, int threads);

/* Blocks until all work the context issued on `stream` is done. */
int sfmm_synchronize(sfmm_context *ctx, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* SOLIDFMM_GPU_H */
