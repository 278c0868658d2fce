// Runs the public translation API of solidfmm (translations.hpp:47-170,
// handles.hpp:33-110, solid.hpp:28-66) on seeded inputs -- TEST
// INFRASTRUCTURE ONLY.
//
// oracle/Makefile builds it twice from this one source: against the
// reference library (compiled from the reference's own sources) and against
// the drop-in headers over libsolidfmm_gpu.so.  tests/test_gpu_dropin.py
// runs both and compares what they write.
//
// Forms exercised, for double and float, M2M / M2L / L2L, dim 1 and 3,
// lists with repeated targets, mixed orders and nonzero initial targets:
//   translation_info  m2l(op, buf, begin, end)             translations.hpp:122-135
//   unchecked         m2l_unchecked(op, buf, begin, end)    translations.hpp:154-170
//   classic           m2l(op, buf, howmany, M, L, x, y, z)   translations.hpp:103-121
//   mt, mt_unchecked  multithreaded_translator<real>(P).m2l(begin, end) and
//                     .m2l_unchecked(...)                    translations.hpp:47-87
// The reference's multithreaded_translator needs hwloc, which this image
// lacks, so the reference build (SFMM_WITH_MT=0) writes its
// translation_info result under the mt names: the drop-in's batch class is
// compared with the reference's single-thread result, which is the same
// mathematics (multithreaded_translator.cpp:298-319 calls m2l_unchecked).
//
// Error cases print "err <name> <code> <targets unchanged>" with code 0 no
// exception, 1 std::logic_error, 2 std::out_of_range, 3 std::overflow_error,
// 4 anything else (translations.cpp:634-1001, operator_data.cpp:151-170).
//
// L2L lists keep p_in <= p_out + 1: above that the reference reads Euler
// tables it did not compute (translations.cpp:414-417 vs 198-224).
//
// Usage: dropin_check <output file>   (records: u32 name length, name,
// u64 count, count doubles).  Per scenario: <base>_orders (the targets'
// orders), <base>_initial (the targets before the call), then the targets
// after each form; targets are concatenated, each dim blocks of P(P+1).
#include <solidfmm/handles.hpp>
#include <solidfmm/solid.hpp>
#include <solidfmm/translations.hpp>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#ifndef SFMM_WITH_MT
#define SFMM_WITH_MT 0
#endif

// The reference build issues every list in calls of at most 12 / dim items
// (each item fills dim SIMD lanes, translations.cpp:654-658): its AVX-512
// solid2buf stores a register into the wrong lane of 16-wide batches
// (microkernel_avx512f.cpp:2182), corrupting lanes 12-15.  Targets still
// accumulate in list order.  The drop-in build issues whole lists.
#if SFMM_WITH_MT
#define CHUNK 0
#else
#define CHUNK 12
#endif

using namespace solidfmm;

namespace {

FILE *out_file;
uint64_t rng_state;

double urand()  // SplitMix64, 53-bit uniforms in [0, 1)
{
    uint64_t z = (rng_state += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return static_cast<double>(z >> 11) * 0x1.0p-53;
}

void record(const std::string &name, const std::vector<double> &v)
{
    uint32_t len = static_cast<uint32_t>(name.size());
    uint64_t n = v.size();
    std::fwrite(&len, sizeof len, 1, out_file);
    std::fwrite(name.data(), 1, len, out_file);
    std::fwrite(&n, sizeof n, 1, out_file);
    std::fwrite(v.data(), sizeof(double), n, out_file);
}

// Inputs are float values (exact in double too), so the double run on the
// same seed is the accurate result for judging both float paths.
float frand(double lo, double hi) { return static_cast<float>(lo + (hi - lo) * urand()); }

template <typename real>
void fill(solid<real> &s)
{
    for (size_t d = 0; d < s.dimension(); ++d)
        for (size_t n = 0; n < s.order(); ++n)
            for (size_t m = 0; m <= n; ++m) {
                s.re(n, m, d) = frand(-1, 1);
                s.im(n, m, d) = m ? frand(-1, 1) : real(0);
            }
}

template <typename real>
struct scenario {
    int op = 1;
    std::vector<solid<real>> src, dst;
    std::vector<translation_info<real>> list;

    // `count` items over `ntargets` targets (repeated, shuffled), orders in
    // 1..P, the item's shift a direction times U[2,4] (M2L) or an octant
    // offset (M2M, L2L; one zero shift).
    scenario(int o, size_t P, size_t dim, size_t count, size_t ntargets, uint64_t seed) : op(o)
    {
        rng_state = seed;
        dst.resize(ntargets);
        for (auto &t : dst) {
            t = solid<real>(1 + static_cast<size_t>(urand() * P), dim);
            fill(t);
        }
        src.resize(count);
        list.resize(count);
        for (size_t i = 0; i < count; ++i) {
            size_t t = static_cast<size_t>(urand() * ntargets);
            size_t pin = 1 + static_cast<size_t>(urand() * P);
            if (op == 2 && pin > dst[t].order() + 1) pin = dst[t].order() + 1;
            src[i] = solid<real>(pin, dim);
            fill(src[i]);
            double x, y, z;
            if (op == 1) {
                double cz = 2 * urand() - 1, ph = 2 * M_PI * urand(), r = 2 + 2 * urand();
                double rho = std::sqrt(std::max(0.0, 1 - cz * cz));
                x = r * rho * std::cos(ph);
                y = r * rho * std::sin(ph);
                z = r * cz;
            } else {
                double s = 0.5 + 0.5 * urand();
                x = (urand() < 0.5 ? -0.5 : 0.5) * s;
                y = (urand() < 0.5 ? -0.5 : 0.5) * s;
                z = (urand() < 0.5 ? -0.5 : 0.5) * s;
                if (i == 3) x = y = z = 0;
            }
            list[i] = translation_info<real>{&src[i], &dst[t], static_cast<float>(x), static_cast<float>(y),
                                             static_cast<float>(z)};
        }
    }

    std::vector<double> targets() const
    {
        std::vector<double> v;
        for (const auto &t : dst)
            for (size_t i = 0; i < t.dimension() * t.order() * (t.order() + 1); ++i) v.push_back(t.memptr()[i]);
        return v;
    }
};

template <typename real>
void run_ti(int op, const operator_data<real> &od, threadlocal_buffer<real> &tb, const translation_info<real> *b,
            const translation_info<real> *e, bool checked)
{
    const ptrdiff_t chunk = (b != e) ? CHUNK / static_cast<ptrdiff_t>(b->source->dimension()) : 0;
    if (chunk && e - b > chunk) {
        for (; b < e; b += chunk) run_ti(op, od, tb, b, e - b > chunk ? b + chunk : e, checked);
        return;
    }
    if (checked) {
        if (op == 0) m2m(od, tb, b, e);
        else if (op == 1) m2l(od, tb, b, e);
        else l2l(od, tb, b, e);
    } else {
        if (op == 0) m2m_unchecked(od, tb, b, e);
        else if (op == 1) m2l_unchecked(od, tb, b, e);
        else l2l_unchecked(od, tb, b, e);
    }
}

template <typename real>
void run_classic(int op, const operator_data<real> &od, threadlocal_buffer<real> &tb, scenario<real> &s)
{
    std::vector<const solid<real> *> in;
    std::vector<solid<real> *> outp;
    std::vector<real> x, y, z;
    for (auto &it : s.list) {
        in.push_back(it.source);
        outp.push_back(it.target);
        x.push_back(it.x);
        y.push_back(it.y);
        z.push_back(it.z);
    }
    size_t n = s.list.size(), step = CHUNK ? CHUNK / s.src[0].dimension() : n;
    for (size_t i = 0; i < n; i += step) {
        size_t k = std::min(step, n - i);
        if (op == 0) m2m(od, tb, k, in.data() + i, outp.data() + i, x.data() + i, y.data() + i, z.data() + i);
        else if (op == 1) m2l(od, tb, k, in.data() + i, outp.data() + i, x.data() + i, y.data() + i, z.data() + i);
        else l2l(od, tb, k, in.data() + i, outp.data() + i, x.data() + i, y.data() + i, z.data() + i);
    }
}

#if SFMM_WITH_MT
template <typename real>
void run_mt(int op, const multithreaded_translator<real> &mt, const translation_info<real> *b,
            const translation_info<real> *e, bool checked)
{
    if (checked) {
        if (op == 0) mt.m2m(b, e);
        else if (op == 1) mt.m2l(b, e);
        else mt.l2l(b, e);
    } else {
        if (op == 0) mt.m2m_unchecked(b, e);
        else if (op == 1) mt.m2l_unchecked(b, e);
        else mt.l2l_unchecked(b, e);
    }
}
#endif

const char *op_name(int op) { return op == 0 ? "m2m" : op == 1 ? "m2l" : "l2l"; }

template <typename real>
void forms(const char *tname, size_t P, size_t dim, uint64_t seed)
{
    operator_handle<real> oph(P);
    buffer_handle<real> bh = oph.make_buffer();
    const operator_data<real> &od = oph;
    threadlocal_buffer<real> &tb = bh;
#if SFMM_WITH_MT
    // the rest of the class's surface (translations.hpp:47-87): the
    // topology constructor, both scheduling policies, move construction and
    // move assignment; the object that results runs the mt forms below
    multithreaded_translator<real> first(P, nullptr, nullptr, nullptr);
    first.schedule_by_measuring_throughput();
    multithreaded_translator<real> moved(std::move(first));
    multithreaded_translator<real> mt(P);
    mt = std::move(moved);
    mt.schedule_equally_among_threads();
#endif
    for (int op = 0; op < 3; ++op) {
        std::string base = std::string(tname) + "_" + op_name(op) + "_d" + std::to_string(dim);
        uint64_t sd = seed + 101 * op + dim;
        {
            scenario<real> s(op, P, dim, 48, 12, sd);
            std::vector<double> layout;
            for (const auto &t : s.dst) layout.push_back(static_cast<double>(t.order()));
            record(base + "_orders", layout);
            record(base + "_initial", s.targets());
            run_ti(op, od, tb, s.list.data(), s.list.data() + s.list.size(), true);
            record(base + "_ti", s.targets());
        }
        {
            scenario<real> s(op, P, dim, 48, 12, sd);
            run_ti(op, od, tb, s.list.data(), s.list.data() + s.list.size(), false);
            record(base + "_unchecked", s.targets());
        }
        {
            scenario<real> s(op, P, dim, 48, 12, sd);
            run_classic(op, od, tb, s);
            record(base + "_classic", s.targets());
        }
        for (int checked = 1; checked >= 0; --checked) {
            scenario<real> s(op, P, dim, 48, 12, sd);
#if SFMM_WITH_MT
            run_mt(op, mt, s.list.data(), s.list.data() + s.list.size(), checked);
#else
            run_ti(op, od, tb, s.list.data(), s.list.data() + s.list.size(), checked);
#endif
            record(base + (checked ? "_mt" : "_mt_unchecked"), s.targets());
        }
    }
}

template <typename F>
int catch_code(F f)
{
    try {
        f();
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

// A list whose last item is invalid: a checked call must throw before any
// write (the reference validates the whole list first).
template <typename real, typename Call>
void error_case(const char *name, size_t P, size_t bad_pin, size_t bad_pout, size_t bad_dim_in, size_t bad_dim_out,
                Call call)
{
    std::vector<solid<real>> src(3), dst(3);
    rng_state = 77;
    for (int i = 0; i < 3; ++i) {
        bool bad = (i == 2);
        src[i] = solid<real>(bad ? bad_pin : P, bad ? bad_dim_in : 1);
        dst[i] = solid<real>(bad ? bad_pout : P, bad ? bad_dim_out : 1);
        fill(src[i]);
        fill(dst[i]);
    }
    std::vector<translation_info<real>> list;
    for (int i = 0; i < 3; ++i) list.push_back({&src[i], &dst[i], real(1), real(-1), real(0.5)});
    std::vector<std::vector<real>> before;
    for (auto &d : dst)
        before.emplace_back(d.memptr(), d.memptr() + d.dimension() * d.order() * (d.order() + 1));
    int code = catch_code([&] { call(list.data(), list.data() + list.size()); });
    bool unchanged = true;
    for (size_t i = 0; i < dst.size(); ++i)
        unchanged = unchanged && std::memcmp(before[i].data(), dst[i].memptr(), before[i].size() * sizeof(real)) == 0;
    std::printf("err %s %d %d\n", name, code, unchanged ? 1 : 0);
}

template <typename real>
void errors(const char *tname, size_t Pmax_type)
{
    const size_t P = 8;
    operator_handle<real> oph(P);
    buffer_handle<real> bh = oph.make_buffer();
    const operator_data<real> &od = oph;
    threadlocal_buffer<real> &tb = bh;
    auto ti = [&](const translation_info<real> *b, const translation_info<real> *e) { m2l(od, tb, b, e); };
    auto ti_l2l = [&](const translation_info<real> *b, const translation_info<real> *e) { l2l(od, tb, b, e); };
    std::string t = tname;
    error_case<real>((t + "_source_order_above_P").c_str(), P, P + 1, P, 1, 1, ti);
    error_case<real>((t + "_target_order_above_P").c_str(), P, P, P + 1, 1, 1, ti_l2l);
    error_case<real>((t + "_dimension_mismatch").c_str(), P, P, P, 1, 3, ti);
#if SFMM_WITH_MT
    multithreaded_translator<real> mt(P);
    auto mtc = [&](const translation_info<real> *b, const translation_info<real> *e) { mt.m2l(b, e); };
    auto mtc_m2m = [&](const translation_info<real> *b, const translation_info<real> *e) { mt.m2m(b, e); };
#else
    auto mtc = ti;
    auto mtc_m2m = [&](const translation_info<real> *b, const translation_info<real> *e) { m2m(od, tb, b, e); };
#endif
    error_case<real>((t + "_mt_source_order_above_P").c_str(), P, P + 1, P, 1, 1, mtc);
    error_case<real>((t + "_mt_dimension_mismatch").c_str(), P, P, P, 3, 1, mtc_m2m);
    std::printf("err %s_handle_P0 %d 1\n", tname, catch_code([] { operator_handle<real> h(0); }));
    std::printf("err %s_handle_above_max %d 1\n", tname,
                catch_code([&] { operator_handle<real> h(Pmax_type + 1); }));
    std::printf("err %s_handle_at_max %d 1\n", tname, catch_code([&] { operator_handle<real> h(Pmax_type); }));
    // a buffer made for another operator order
    operator_handle<real> other(P + 4);
    buffer_handle<real> obh = other.make_buffer();
    threadlocal_buffer<real> &otb = obh;
    auto wrong_buf = [&](const translation_info<real> *b, const translation_info<real> *e) { m2l(od, otb, b, e); };
    error_case<real>((t + "_incompatible_buffer").c_str(), P, P, P, 1, 1, wrong_buf);
#if SFMM_WITH_MT
    std::printf("err %s_mt_P0 %d 1\n", tname, catch_code([] { multithreaded_translator<real> m(0); }));
#else
    std::printf("err %s_mt_P0 %d 1\n", tname, catch_code([] { operator_handle<real> h(0); }));
#endif
}

}  // namespace

int main(int argc, char **argv)
{
    if (argc < 2) {
        std::fprintf(stderr, "usage: %s <output file>\n", argv[0]);
        return 2;
    }
    out_file = std::fopen(argv[1], "wb");
    if (!out_file) return 2;
    forms<double>("f64", 12, 1, 1000);
    forms<double>("f64", 12, 3, 2000);
    forms<double>("f64p10", 10, 1, 3000);  // the f32 inputs, in double
    forms<double>("f64p10", 10, 3, 4000);
    forms<float>("f32", 10, 1, 3000);
    forms<float>("f32", 10, 3, 4000);
    errors<double>("f64", 86);
    errors<float>("f32", 18);
    std::fclose(out_file);
    std::printf("done\n");
    return 0;
}
