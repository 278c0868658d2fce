// C entry points over the compiled reference library (vorticle/solidfmm 1.5)
// -- TEST INFRASTRUCTURE ONLY.
//
// Built by oracle/Makefile, together with the reference's own sources taken
// where they lie, into oracle/_ref/libsolidfmm_ref.so.  Used to write golden
// vectors (tests/golden/gen_golden.py), to pin the oracle restatement, and as
// the CPU baseline timed by bench.py.  The product library never loads it.
//
// Reference interfaces called (file:line in the reference):
//   operator_handle / buffer_handle          handles.hpp:33-110
//   m2l / m2m / l2l (translation_info form)  translations.hpp:122-170
//   swap_matrix<double>::operator()          swap_matrix.hpp:48
//   harmonics::R / S                         harmonics.hpp:30-48
//
// Reference defect (microkernel_avx512f.cpp:2182): the AVX-512 double
// solid2buf stores r20 where r21 belongs, so on AVX-512 hosts lanes 12-15 of
// every 16-lane batch read Im C_n^m (m = 2 mod 4, n >= 3) of lane l-4.  The
// `chunk` argument below issues the list in calls of at most `chunk` items;
// chunk <= 12 keeps every item in lanes 0-11, where results are correct.
//
// multithreaded_translator is not built: it needs hwloc, which this image
// lacks.  translate() with nthreads > 1 runs the reference's single-thread
// entry point on N std::threads instead, each owning the targets of one hash
// range, with its own buffer_handle.  ref_pool_translate_f64() follows the
// batch protocol of multithreaded_translator (multithreaded_translator.cpp:
// 200-330, 375-654) and of its benchmark (tests/benchmark_multithreaded_
// translator.cpp:64-88) more closely: persistent workers, one pinned to each
// physical core of the process's affinity mask (sched_setaffinity in place
// of hwloc), created before any timing and released per call by barriers;
// every worker scans the whole list and runs the items whose target hash
// falls in its equal share of the hash range, in chunks of 1337 through
// the *_unchecked entry points; a warm-up call on the first 1024 items, then
// timed calls over the full list.
#include <solidfmm/sphere.hpp>
#include <solidfmm/handles.hpp>
#include <solidfmm/harmonics.hpp>
#include <solidfmm/solid.hpp>
#include <solidfmm/swap_matrix.hpp>
#include <solidfmm/translations.hpp>

#include <solidfmm/microkernel.hpp>
#include <solidfmm/microkernel_avx.hpp>
#include <solidfmm/microkernel_avx512f.hpp>

#include <atomic>
#include <barrier>
#include <chrono>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <exception>
#include <fstream>
#include <map>
#include <memory>
#include <set>
#include <string>
#include <thread>
#include <vector>

#include <pthread.h>
#include <sched.h>
#include <unistd.h>

namespace {

struct shim_item {
    uint64_t src, dst;
    uint32_t p_in, p_out;
    double x, y, z;
};

template <typename real>
struct staged {
    std::map<uint64_t, size_t> src_idx, dst_idx;
    std::vector<solidfmm::solid<real>> srcs, dsts;
    std::vector<solidfmm::translation_info<real>> list;
};

template <typename real>
void stage(staged<real> &s, const shim_item *items, size_t n, const real *src_pool, const real *dst_pool)
{
    // Distinct offsets become distinct solids; reserve first so the
    // pointers placed in translation_info stay valid.
    for (size_t i = 0; i < n; ++i) {
        if (!s.src_idx.count(items[i].src)) {
            size_t k = s.src_idx.size();
            s.src_idx[items[i].src] = k;
        }
        if (!s.dst_idx.count(items[i].dst)) {
            size_t k = s.dst_idx.size();
            s.dst_idx[items[i].dst] = k;
        }
    }
    s.srcs.resize(s.src_idx.size());
    s.dsts.resize(s.dst_idx.size());
    std::vector<uint32_t> pin(s.srcs.size(), 0), pout(s.dsts.size(), 0);
    for (size_t i = 0; i < n; ++i) {
        pin[s.src_idx[items[i].src]] = items[i].p_in;
        pout[s.dst_idx[items[i].dst]] = items[i].p_out;
    }
    for (auto &kv : s.src_idx) {
        size_t P = pin[kv.second];
        s.srcs[kv.second] = solidfmm::solid<real>(P);
        if (P) std::memcpy(s.srcs[kv.second].memptr(), src_pool + kv.first, sizeof(real) * P * (P + 1));
    }
    for (auto &kv : s.dst_idx) {
        size_t P = pout[kv.second];
        s.dsts[kv.second] = solidfmm::solid<real>(P);
        if (P) std::memcpy(s.dsts[kv.second].memptr(), dst_pool + kv.first, sizeof(real) * P * (P + 1));
    }
    s.list.resize(n);
    for (size_t i = 0; i < n; ++i) {
        s.list[i].source = &s.srcs[s.src_idx[items[i].src]];
        s.list[i].target = &s.dsts[s.dst_idx[items[i].dst]];
        s.list[i].x = static_cast<real>(items[i].x);
        s.list[i].y = static_cast<real>(items[i].y);
        s.list[i].z = static_cast<real>(items[i].z);
    }
}

template <typename real>
void unstage(const staged<real> &s, real *dst_pool)
{
    for (auto &kv : s.dst_idx) {
        const auto &d = s.dsts[kv.second];
        size_t P = d.order();
        if (P) std::memcpy(dst_pool + kv.first, d.memptr(), sizeof(real) * P * (P + 1));
    }
}

template <typename real>
void call(int op, const solidfmm::operator_handle<real> &oph, solidfmm::buffer_handle<real> &bh,
          const solidfmm::translation_info<real> *b, const solidfmm::translation_info<real> *e, bool checked)
{
    const solidfmm::operator_data<real> &od = oph;
    solidfmm::threadlocal_buffer<real> &tb = bh;
    if (checked) {
        if (op == 0) solidfmm::m2m(od, tb, b, e);
        else if (op == 1) solidfmm::m2l(od, tb, b, e);
        else solidfmm::l2l(od, tb, b, e);
    } else {
        if (op == 0) solidfmm::m2m_unchecked(od, tb, b, e);
        else if (op == 1) solidfmm::m2l_unchecked(od, tb, b, e);
        else solidfmm::l2l_unchecked(od, tb, b, e);
    }
}

uint64_t target_hash(uint64_t v)
{
    v ^= v >> 33;
    v *= 0xff51afd7ed558ccdULL;
    v ^= v >> 33;
    v *= 0xc4ceb9fe1a85ec53ULL;
    v ^= v >> 33;
    return v;
}

template <typename real>
void call_chunked(int op, const solidfmm::operator_handle<real> &oph, solidfmm::buffer_handle<real> &bh,
                  const solidfmm::translation_info<real> *b, const solidfmm::translation_info<real> *e,
                  bool checked, size_t chunk)
{
    if (chunk == 0) {
        call<real>(op, oph, bh, b, e, checked);
        return;
    }
    for (const solidfmm::translation_info<real> *p = b; p < e; p += chunk)
        call<real>(op, oph, bh, p, (e - p > (ptrdiff_t)chunk) ? p + chunk : e, checked);
}

template <typename real>
int translate(int op, size_t P, const shim_item *items, size_t n, const real *src_pool, real *dst_pool,
              int nthreads, int repeats, size_t chunk, double *seconds)
{
    try {
        if (op < 0 || op > 2) return -1;
        staged<real> s;
        stage(s, items, n, src_pool, dst_pool);
        solidfmm::operator_handle<real> oph(P);
        double best = 0;
        if (nthreads <= 1) {
            solidfmm::buffer_handle<real> bh = oph.make_buffer();
            for (int r = 0; r < (repeats > 0 ? repeats : 1); ++r) {
                if (r > 0)  // re-time on fresh targets
                    for (auto &d : s.dsts) d.zeros();
                auto t0 = std::chrono::steady_clock::now();
                call_chunked<real>(op, oph, bh, s.list.data(), s.list.data() + n, true, chunk);
                auto t1 = std::chrono::steady_clock::now();
                double t = std::chrono::duration<double>(t1 - t0).count();
                if (r == 0 || t < best) best = t;
            }
        } else {
            // Partition by target hash so no two threads write one target.
            std::vector<std::vector<solidfmm::translation_info<real>>> part(nthreads);
            for (size_t i = 0; i < n; ++i) {
                uint64_t h = target_hash(reinterpret_cast<uintptr_t>(s.list[i].target));
                part[h % nthreads].push_back(s.list[i]);
            }
            std::vector<solidfmm::buffer_handle<real>> bufs;
            for (int t = 0; t < nthreads; ++t) bufs.push_back(oph.make_buffer());
            for (int r = 0; r < (repeats > 0 ? repeats : 1); ++r) {
                if (r > 0)
                    for (auto &d : s.dsts) d.zeros();
                auto t0 = std::chrono::steady_clock::now();
                std::vector<std::thread> th;
                for (int t = 0; t < nthreads; ++t)
                    th.emplace_back([&, t] {
                        call_chunked<real>(op, oph, bufs[t], part[t].data(), part[t].data() + part[t].size(),
                                           false, chunk);
                    });
                for (auto &x : th) x.join();
                auto t1 = std::chrono::steady_clock::now();
                double t = std::chrono::duration<double>(t1 - t0).count();
                if (r == 0 || t < best) best = t;
            }
        }
        if (seconds) *seconds = best;
        unstage(s, dst_pool);
        return 0;
    } catch (const std::exception &) {
        return -2;
    }
}


// ---------------------------------------------------------------------------
// Pinned worker pool over the reference's single-thread entry points.
// ---------------------------------------------------------------------------

// One CPU per physical core of the affinity mask (the first CPU of each
// (package, core) pair), from sysfs topology; every CPU in the mask when the
// topology is unreadable.
std::vector<int> physical_core_cpus(int *nmask)
{
    cpu_set_t mask;
    CPU_ZERO(&mask);
    sched_getaffinity(0, sizeof(mask), &mask);
    std::set<std::pair<int, int>> seen;
    std::vector<int> cpus;
    int count = 0;
    for (int c = 0; c < CPU_SETSIZE; ++c) {
        if (!CPU_ISSET(c, &mask)) continue;
        ++count;
        char path[128];
        int core = -1, pkg = -1;
        std::snprintf(path, sizeof path, "/sys/devices/system/cpu/cpu%d/topology/core_id", c);
        std::ifstream(path) >> core;
        std::snprintf(path, sizeof path, "/sys/devices/system/cpu/cpu%d/topology/physical_package_id", c);
        std::ifstream(path) >> pkg;
        if (core < 0 || pkg < 0) { cpus.push_back(c); continue; }
        if (seen.insert({pkg, core}).second) cpus.push_back(c);
    }
    if (nmask) *nmask = count;
    return cpus;
}

// Equal shares of the hash range, as schedule_equally_among_threads.
inline int owner(uint64_t h, int nthreads)
{
    return static_cast<int>((static_cast<unsigned __int128>(h) * static_cast<unsigned>(nthreads)) >> 64);
}

template <typename real>
struct pool {
    using item = solidfmm::translation_info<real>;
    const solidfmm::operator_handle<real> &oph;
    int op = 1, n = 0;
    const item *begin = nullptr, *end = nullptr;
    bool die = false;
    std::barrier<> start, done;
    std::vector<std::thread> th;
    std::vector<int> cpus;
    int pinned = 0;

    pool(const solidfmm::operator_handle<real> &o, std::vector<int> c)
        : oph(o), n(static_cast<int>(c.size())), start(n + 1), done(n + 1), cpus(std::move(c))
    {
        std::barrier<> ready(n + 1);
        std::atomic<int> ok{0};
        for (int t = 0; t < n; ++t)
            th.emplace_back([this, t, &ready, &ok] {
                cpu_set_t one;
                CPU_ZERO(&one);
                CPU_SET(cpus[t], &one);
                if (pthread_setaffinity_np(pthread_self(), sizeof(one), &one) == 0) ok.fetch_add(1);
                solidfmm::buffer_handle<real> bh = oph.make_buffer();
                std::vector<item> buf(1337);
                ready.arrive_and_wait();
                while (true) {
                    start.arrive_and_wait();
                    if (die) return;
                    size_t nb = 0;
                    for (const item *i = begin; i != end; ++i) {
                        if (owner(target_hash(reinterpret_cast<uintptr_t>(i->target)), n) != t) continue;
                        buf[nb++] = *i;
                        if (nb == buf.size()) {
                            call<real>(op, oph, bh, buf.data(), buf.data() + nb, false);
                            nb = 0;
                        }
                    }
                    if (nb) call<real>(op, oph, bh, buf.data(), buf.data() + nb, false);
                    done.arrive_and_wait();
                }
            });
        ready.arrive_and_wait();
        pinned = ok.load();
    }

    double run(int o, const item *b, const item *e)
    {
        op = o;
        begin = b;
        end = e;
        auto t0 = std::chrono::steady_clock::now();
        start.arrive_and_wait();
        done.arrive_and_wait();
        auto t1 = std::chrono::steady_clock::now();
        return std::chrono::duration<double>(t1 - t0).count();
    }

    ~pool()
    {
        die = true;
        start.arrive_and_wait();
        for (auto &x : th) x.join();
    }
};

const char *microkernel_name()
{
    std::unique_ptr<solidfmm::microkernel<double>> k(solidfmm::get_microkernel<double>());
    if (dynamic_cast<solidfmm::microkernel_double_avx512f *>(k.get())) return "microkernel_double_avx512f";
    if (dynamic_cast<solidfmm::microkernel_double_avx *>(k.get())) return "microkernel_double_avx";
    return "microkernel_double_generic";
}

}  // namespace

extern "C" {

// Host description for the CPU baseline: cpu model (from /proc/cpuinfo),
// online CPUs, CPUs in the affinity mask, physical cores in the mask, the
// microkernel the reference selects on this host, and its tile.
int ref_host_info(char *model, size_t model_len, int *nproc, int *nmask, int *physical_cores, char *kernel,
                  size_t kernel_len, int *rows, int *cols)
{
    std::string m = "unknown";
    std::ifstream f("/proc/cpuinfo");
    for (std::string line; std::getline(f, line);)
        if (line.rfind("model name", 0) == 0) {
            auto p = line.find(':');
            if (p != std::string::npos) m = line.substr(p + 2);
            break;
        }
    std::snprintf(model, model_len, "%s", m.c_str());
    *nproc = static_cast<int>(sysconf(_SC_NPROCESSORS_ONLN));
    *physical_cores = static_cast<int>(physical_core_cpus(nmask).size());
    std::snprintf(kernel, kernel_len, "%s", microkernel_name());
    std::unique_ptr<solidfmm::microkernel<double>> k(solidfmm::get_microkernel<double>());
    *rows = static_cast<int>(k->rows);
    *cols = static_cast<int>(k->cols);
    return 0;
}

// Times the list on the pinned pool: `threads` workers (0: one per physical
// core of the affinity mask), a warm-up call on the first `warmup` items,
// then `repeats` timed calls over the full list, each on zeroed targets;
// seconds[r] receives each timed call's wall time.  The targets hold one
// application of the list at return (lanes 12-15 of its 16-wide batches
// carry the AVX-512 defect noted above; timing is unaffected).  *used
// receives the worker count, *pinned how many took their CPU.
int ref_pool_translate_f64(int op, size_t P, const void *items, size_t n, const double *src_pool,
                           double *dst_pool, int threads, size_t warmup, int repeats, double *seconds, int *used,
                           int *pinned)
{
    try {
        if (op < 0 || op > 2 || repeats < 1) return -1;
        staged<double> s;
        stage(s, static_cast<const shim_item *>(items), n, src_pool, dst_pool);
        solidfmm::operator_handle<double> oph(P);
        std::vector<int> cpus = physical_core_cpus(nullptr);
        if (threads > 0 && static_cast<size_t>(threads) < cpus.size()) cpus.resize(threads);
        pool<double> pl(oph, cpus);
        *used = pl.n;
        *pinned = pl.pinned;
        const auto *b = s.list.data();
        pl.run(op, b, b + std::min(n, warmup));
        for (int r = 0; r < repeats; ++r) {
            for (auto &d : s.dsts) d.zeros();
            seconds[r] = pl.run(op, b, b + n);
        }
        unstage(s, dst_pool);
        return 0;
    } catch (const std::exception &) {
        return -2;
    }
}

}  // extern "C"

namespace {
}  // namespace

extern "C" {

int ref_translate_f64(int op, size_t P, const void *items, size_t n, const double *src_pool, double *dst_pool,
                      int nthreads, int repeats, size_t chunk, double *seconds)
{
    return translate<double>(op, P, static_cast<const shim_item *>(items), n, src_pool, dst_pool, nthreads,
                             repeats, chunk, seconds);
}

int ref_translate_f32(int op, size_t P, const void *items, size_t n, const float *src_pool, float *dst_pool,
                      int nthreads, int repeats, size_t chunk, double *seconds)
{
    return translate<float>(op, P, static_cast<const shim_item *>(items), n, src_pool, dst_pool, nthreads,
                            repeats, chunk, seconds);
}

int ref_swap_entry(int n, int m, int l, double *out)
{
    try {
        solidfmm::swap_matrix<double> B(n + 1);
        *out = B(n, m, l);
        return 0;
    } catch (const std::exception &) {
        return -2;
    }
}

// kind 0: R, 1: S.
int ref_harmonic(int kind, size_t P, double x, double y, double z, double *out)
{
    try {
        solidfmm::solid<double> s = (kind == 0)   ? solidfmm::harmonics::R<double>(P, x, y, z)
                                    : (kind == 1) ? solidfmm::harmonics::S<double>(P, x, y, z)
                                                  : solidfmm::harmonics::dR<double>(P, x, y, z);
        const size_t dims = (kind == 2) ? 3 : 1;
        if (P) std::memcpy(out, s.memptr(), sizeof(double) * dims * P * (P + 1));
        return 0;
    } catch (const std::exception &) {
        return -2;
    }
}

// Exercises the reference's validation: returns 0 if the call went through,
// 1 for std::logic_error, 2 for std::out_of_range, 3 for std::overflow_error,
// 4 for any other exception.
int ref_check_errors(int op, size_t P_op, size_t P_in, size_t dim_in, size_t P_out, size_t dim_out)
{
    try {
        solidfmm::operator_handle<double> oph(P_op);
        solidfmm::buffer_handle<double> bh = oph.make_buffer();
        solidfmm::solid<double> a(P_in, dim_in), b(P_out, dim_out);
        a.zeros();
        b.zeros();
        solidfmm::translation_info<double> it{&a, &b, 1.0, 1.0, 1.0};
        call<double>(op, oph, bh, &it, &it + 1, true);
        return 0;
    } catch (const std::out_of_range &) {
        return 2;
    } catch (const std::overflow_error &) {
        return 3;
    } catch (const std::logic_error &) {
        return 1;
    } catch (...) {
        return 4;
    }
}

// Smallest enclosing ball through the reference (sphere.cpp:546-627):
// kind 0 of n points (x, y, z arrays in data[0..3n)), kind 1 of n balls
// (data = n quadruples x, y, z, r).  out = x, y, z, r.  1 on a C++ exception.
int ref_bounding_sphere(int kind, size_t n, const double *data, double *out)
{
    try {
        solidfmm::sphere<double> s;
        if (kind == 0) {
            s = solidfmm::bounding_sphere(n, data, data + n, data + 2 * n);
        } else {
            std::vector<solidfmm::sphere<double>> b(n);
            for (size_t i = 0; i < n; ++i) b[i] = {data[4 * i], data[4 * i + 1], data[4 * i + 2], data[4 * i + 3]};
            s = solidfmm::bounding_sphere(b.data(), b.data() + n);
        }
        out[0] = s.x;
        out[1] = s.y;
        out[2] = s.z;
        out[3] = s.r;
        return 0;
    } catch (const std::exception &) {
        return 1;
    }
}

}  // extern "C"
