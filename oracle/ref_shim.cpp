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
// lacks.  ref_translate_threads() runs the reference's single-thread entry
// point on N std::threads instead, each owning the targets of one hash range
// (the partitioning multithreaded_translator.cpp:298-319 uses), with its own
// buffer_handle.
#include <solidfmm/sphere.hpp>
#include <solidfmm/handles.hpp>
#include <solidfmm/harmonics.hpp>
#include <solidfmm/solid.hpp>
#include <solidfmm/swap_matrix.hpp>
#include <solidfmm/translations.hpp>

#include <chrono>
#include <cstdint>
#include <cstring>
#include <exception>
#include <map>
#include <thread>
#include <vector>

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
