// Implementation of the drop-in C++ API (include/solidfmm/*.hpp) over the
// C ABI of solidfmm_gpu.h.
//
//   solid<real>, dot, fmadd          solid.hpp:28-79, solid.cpp:40-201
//   harmonics R, S, dR, dS           harmonics.hpp:30-48 (Dehnen's recurrences,
//                                    doc/solidfmm.tex:1076-1106)
//   operator_handle / buffer_handle  handles.hpp:33-110, handles.cpp:30-103
//   m2m / m2l / l2l (+ _unchecked)   translations.hpp:103-170
//   multithreaded_translator         translations.hpp:47-87
//
// Translations run on the GPU: the distinct solids of a call are packed into
// the pinned host pools of a staging object (FP64; float solids are
// widened), translated by sfmm_staging_translate (device pools, one plan,
// list order per target) and written back.  Vector-valued solids become one item per dimension, as the
// reference queues them (translations.cpp:654-658).
#include <solidfmm/handles.hpp>
#include <solidfmm/harmonics.hpp>
#include <solidfmm/operator_data.hpp>
#include <solidfmm/solid.hpp>
#include <solidfmm/threadlocal_buffer.hpp>
#include <solidfmm/translations.hpp>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <new>
#include <stdexcept>
#include <string>
#include <thread>
#include <unordered_map>
#include <utility>
#include <vector>

#include "../../include/solidfmm_gpu.h"

namespace solidfmm
{

// ---------------------------------------------------------------------------
// solid<real>
// ---------------------------------------------------------------------------

template <typename real>
solid<real>::solid( size_t arg_P, size_t arg_dim )
{
    resize(arg_P, arg_dim);
    zeros();
}

template <typename real>
solid<real>::solid( const solid &rhs )
{
    resize(rhs.P, rhs.dim);
    if (data) std::memcpy(data, rhs.data, sizeof(real) * P * (P + 1) * dim);
}

template <typename real>
solid<real>::solid( solid &&rhs ) noexcept: P { rhs.P }, dim { rhs.dim }, data { rhs.data }
{
    rhs.P = 0;
    rhs.dim = 1;
    rhs.data = nullptr;
}

template <typename real>
solid<real>& solid<real>::operator=( const solid &rhs )
{
    if (this != &rhs) {
        resize(rhs.P, rhs.dim);
        if (data) std::memcpy(data, rhs.data, sizeof(real) * P * (P + 1) * dim);
    }
    return *this;
}

template <typename real>
solid<real>& solid<real>::operator=( solid &&rhs ) noexcept
{
    std::swap(P, rhs.P);
    std::swap(dim, rhs.dim);
    std::swap(data, rhs.data);
    return *this;
}

template <typename real>
solid<real>::~solid() { std::free(data); }

template <typename real>
void solid<real>::resize( size_t arg_P, size_t arg_dim )
{
    const size_t want = sizeof(real) * arg_P * (arg_P + 1) * arg_dim;
    const size_t have = sizeof(real) * P * (P + 1) * dim;
    if (want != have) {
        if (want) {
            void *tmp = std::realloc(data, want);
            if (!tmp) throw std::bad_alloc {};
            data = static_cast<real *>(tmp);
        } else {
            std::free(data);
            data = nullptr;
        }
    }
    P = arg_P;
    dim = arg_dim;
}

template <typename real>
void solid<real>::reinit( size_t arg_P, size_t arg_dim )
{
    resize(arg_P, arg_dim);
    zeros();
}

template <typename real>
void solid<real>::zeros() noexcept
{
    if (data) std::memset(data, 0, sizeof(real) * P * (P + 1) * dim);
}

template class solid<float>;
template class solid<double>;

namespace
{

// Re sum_{n, m=-n..n} conj(A_n^m) B_n^m over the common degrees, for every
// pair of dimensions (solid.cpp:40-67).
template <typename real>
void dot_impl( const solid<real> &A, const solid<real> &B, real *result ) noexcept
{
    const size_t P = std::min(A.order(), B.order());
    const size_t N = P * (P + 1);
    const size_t sa = A.order() * (A.order() + 1), sb = B.order() * (B.order() + 1);
    for (size_t da = 0; da < A.dimension(); ++da)
        for (size_t db = 0; db < B.dimension(); ++db) {
            const real *a = A.memptr() + sa * da;
            const real *b = B.memptr() + sb * db;
            real all = 0, m0 = 0;
            for (size_t k = 0; k < N; ++k) all += a[k] * b[k];
            for (size_t n = 0; n < P; ++n) m0 += a[n * (n + 1)] * b[n * (n + 1)];
            *result++ = real(2) * all - m0;
        }
}

template <typename real>
void fmadd_impl( const real *fac, bool per_dim, const solid<real> &A, solid<real> &B )
{
    if (A.dimension() != B.dimension()) throw std::logic_error { "solidfmm::fmadd(): dimension mismatch." };
    if (A.order() != B.order()) throw std::logic_error { "solidfmm::fmadd(): order mismatch." };
    const size_t len = A.order() * (A.order() + 1);
    for (size_t d = 0; d < A.dimension(); ++d) {
        const real f = per_dim ? fac[d] : fac[0];
        const real *in = A.memptr() + d * len;
        real *out = B.memptr() + d * len;
        for (size_t i = 0; i < len; ++i) out[i] += f * in[i];
    }
}

}  // namespace

void dot( const solid<float > &A, const solid<float > &B, float  *r ) noexcept { dot_impl(A, B, r); }
void dot( const solid<double> &A, const solid<double> &B, double *r ) noexcept { dot_impl(A, B, r); }
void fmadd( const float f, const solid<float> &A, solid<float> &B ) { fmadd_impl(&f, false, A, B); }
void fmadd( const float *f, const solid<float> &A, solid<float> &B ) { fmadd_impl(f, true, A, B); }
void fmadd( const double f, const solid<double> &A, solid<double> &B ) { fmadd_impl(&f, false, A, B); }
void fmadd( const double *f, const solid<double> &A, solid<double> &B ) { fmadd_impl(f, true, A, B); }

// ---------------------------------------------------------------------------
// Solid harmonics.  R and S by Dehnen's recurrences (doc/solidfmm.tex:
// 1076-1106); the gradients by differentiating the same recurrences
// (forward-mode, one dimension per Cartesian direction).
// ---------------------------------------------------------------------------

namespace
{

// Values and, when grad != nullptr, d/dx, d/dy, d/dz of R (kind 0) or S (kind 1).
template <typename real>
void harmonic_impl( int kind, size_t P, real x, real y, real z, real *val, real *grad )
{
    const size_t L = P * (P + 1);
    std::fill(val, val + L, real(0));
    if (grad) std::fill(grad, grad + 3 * L, real(0));
    if (P == 0) return;
    auto RE = [](size_t n, size_t m) { return n * (n + 1) + m; };
    auto IM = [](size_t n, size_t m) { return (n + 1) * (n + 1) + m; };
    // dual numbers: value plus three partials, kept side by side
    struct D {
        real v, dx, dy, dz;
    };
    auto mul = [](D a, D b) { return D { a.v * b.v, a.dx * b.v + a.v * b.dx, a.dy * b.v + a.v * b.dy, a.dz * b.v + a.v * b.dz }; };
    auto add = [](D a, D b) { return D { a.v + b.v, a.dx + b.dx, a.dy + b.dy, a.dz + b.dz }; };
    auto sub = [](D a, D b) { return D { a.v - b.v, a.dx - b.dx, a.dy - b.dy, a.dz - b.dz }; };
    auto scl = [](real s, D a) { return D { s * a.v, s * a.dx, s * a.dy, s * a.dz }; };
    const D X { x, 1, 0, 0 }, Y { y, 0, 1, 0 }, Z { z, 0, 0, 1 };
    const D r2 = add(add(mul(X, X), mul(Y, Y)), mul(Z, Z));
    std::vector<D> re(L), im(L);
    if (kind == 0) {
        re[0] = D { 1, 0, 0, 0 };
        for (size_t m = 1; m < P; ++m) {  // R_m^m = (x+iy)/(2m) R_{m-1}^{m-1}
            const D a = re[RE(m - 1, m - 1)], b = im[IM(m - 1, m - 1)];
            const real f = real(1) / real(2 * m);
            re[RE(m, m)] = scl(f, sub(mul(a, X), mul(b, Y)));
            im[IM(m, m)] = scl(f, add(mul(a, Y), mul(b, X)));
        }
        for (size_t m = 0; m < P; ++m)
            for (size_t n = m + 1; n < P; ++n) {  // (n^2-m^2) R_n^m = (2n-1) z R_{n-1}^m - r^2 R_{n-2}^m
                const real f = real(1) / (real(n + m) * real(n - m));
                D pr = scl(real(2 * n - 1), mul(Z, re[RE(n - 1, m)]));
                D pi = scl(real(2 * n - 1), mul(Z, im[IM(n - 1, m)]));
                if (n >= m + 2) {
                    pr = sub(pr, mul(r2, re[RE(n - 2, m)]));
                    pi = sub(pi, mul(r2, im[IM(n - 2, m)]));
                }
                re[RE(n, m)] = scl(f, pr);
                im[IM(n, m)] = scl(f, pi);
            }
    } else {
        const real rr = x * x + y * y + z * z;
        const real rinv = real(1) / std::sqrt(rr);
        // S_0^0 = 1/|x|: d/dx = -x/|x|^3
        const real r3 = rinv * rinv * rinv;
        re[0] = D { rinv, -x * r3, -y * r3, -z * r3 };
        const real q = real(1) / rr;
        const D qd { q, -2 * x * q * q, -2 * y * q * q, -2 * z * q * q };
        for (size_t m = 1; m < P; ++m) {  // S_m^m = (2m-1)(x+iy)/|x|^2 S_{m-1}^{m-1}
            const D a = re[RE(m - 1, m - 1)], b = im[IM(m - 1, m - 1)];
            const D f = scl(real(2 * m - 1), qd);
            re[RE(m, m)] = mul(f, sub(mul(a, X), mul(b, Y)));
            im[IM(m, m)] = mul(f, add(mul(a, Y), mul(b, X)));
        }
        for (size_t m = 0; m < P; ++m)
            for (size_t n = m + 1; n < P; ++n) {  // |x|^2 S_n^m = (2n-1) z S_{n-1}^m - ((n-1)^2-m^2) S_{n-2}^m
                D pr = scl(real(2 * n - 1), mul(Z, re[RE(n - 1, m)]));
                D pi = scl(real(2 * n - 1), mul(Z, im[IM(n - 1, m)]));
                if (n >= m + 2) {
                    const real c = real(n - 1 - m) * real(n - 1 + m);
                    pr = sub(pr, scl(c, re[RE(n - 2, m)]));
                    pi = sub(pi, scl(c, im[IM(n - 2, m)]));
                }
                re[RE(n, m)] = mul(qd, pr);
                im[IM(n, m)] = mul(qd, pi);
            }
    }
    for (size_t n = 0; n < P; ++n)
        for (size_t m = 0; m <= n; ++m) {
            const D a = re[RE(n, m)], b = im[IM(n, m)];
            val[RE(n, m)] = a.v;
            val[IM(n, m)] = b.v;
            if (grad) {
                grad[0 * L + RE(n, m)] = a.dx;
                grad[0 * L + IM(n, m)] = b.dx;
                grad[1 * L + RE(n, m)] = a.dy;
                grad[1 * L + IM(n, m)] = b.dy;
                grad[2 * L + RE(n, m)] = a.dz;
                grad[2 * L + IM(n, m)] = b.dz;
            }
        }
}

template <typename real>
solid<real> value( int kind, size_t P, real x, real y, real z )
{
    solid<real> s(P);
    if (P) harmonic_impl<real>(kind, P, x, y, z, s.memptr(), nullptr);
    return s;
}

template <typename real>
solid<real> gradient( int kind, size_t P, real x, real y, real z )
{
    solid<real> g(P, 3);
    if (P) {
        std::vector<real> v(P * (P + 1));
        harmonic_impl<real>(kind, P, x, y, z, v.data(), g.memptr());
    }
    return g;
}

}  // namespace

namespace harmonics
{
template <> solid<float > R( size_t P, float  x, float  y, float  z ) { return value<float >(0, P, x, y, z); }
template <> solid<float > S( size_t P, float  x, float  y, float  z ) { return value<float >(1, P, x, y, z); }
template <> solid<float > dR( size_t P, float  x, float  y, float  z ) { return gradient<float >(0, P, x, y, z); }
template <> solid<float > dS( size_t P, float  x, float  y, float  z ) { return gradient<float >(1, P, x, y, z); }
template <> solid<double> R( size_t P, double x, double y, double z ) { return value<double>(0, P, x, y, z); }
template <> solid<double> S( size_t P, double x, double y, double z ) { return value<double>(1, P, x, y, z); }
template <> solid<double> dR( size_t P, double x, double y, double z ) { return gradient<double>(0, P, x, y, z); }
template <> solid<double> dS( size_t P, double x, double y, double z ) { return gradient<double>(1, P, x, y, z); }
}  // namespace harmonics

// ---------------------------------------------------------------------------
// Operator data and buffers.
// ---------------------------------------------------------------------------

template <typename real> constexpr size_t type_limit();
template <> constexpr size_t type_limit<float >() { return 18; }  // (2P-2)! must fit (operator_data.cpp:35-45)
template <> constexpr size_t type_limit<double>() { return 86; }

template <typename real>
operator_data<real>::operator_data( size_t arg_P ): P { arg_P }, ctx { nullptr }
{
    if (P == 0) throw std::logic_error { "solidfmm::operator_data(): Requested order P = 0." };
    if (P > type_limit<real>() || P > sfmm_max_order())
        throw std::overflow_error { "solidfmm::operator_data(): "
                                    "Requested order is too large for the chosen floating point type." };
    int rc = sfmm_create(&ctx, 0, P, SFMM_FP64);
    if (rc == SFMM_ERR_ALLOC) throw std::bad_alloc {};
    if (rc != SFMM_OK) throw std::runtime_error { "solidfmm::operator_data(): no usable GPU (HIP error)" };
}

template <typename real>
operator_data<real>::~operator_data() { sfmm_destroy(ctx); }

template class operator_data<float>;
template class operator_data<double>;

// Host staging reused across calls: the item list, the distinct blocks of
// a call, and a C-ABI staging object (pinned host pools, device pools,
// stream), created on first use with the operator data's context.
template <typename real>
struct threadlocal_buffer<real>::staging {
    struct block {
        real *host;     // the solid's data
        uint64_t off;   // offset in the staging pool (reals)
        uint64_t len;   // reals
    };
    std::vector<sfmm_item> items;
    std::unordered_map<const void *, uint64_t> src_off, dst_off;
    std::vector<block> src_blocks, dst_blocks;
    sfmm_staging *st = nullptr;
    ~staging() { sfmm_staging_destroy(st); }
};

template <typename real>
threadlocal_buffer<real>::threadlocal_buffer() noexcept {}

template <typename real>
threadlocal_buffer<real>::threadlocal_buffer( const operator_data<real> &op ): P { op.order() }, s { new staging }
{
}

template <typename real>
threadlocal_buffer<real>::threadlocal_buffer( const threadlocal_buffer &rhs ):
    P { rhs.P }, s { rhs.s ? new staging : nullptr }
{
}

template <typename real>
threadlocal_buffer<real>::threadlocal_buffer( threadlocal_buffer &&rhs ) noexcept: P { rhs.P }, s { rhs.s }
{
    rhs.P = 0;
    rhs.s = nullptr;
}

template <typename real>
threadlocal_buffer<real> &threadlocal_buffer<real>::operator=( const threadlocal_buffer &rhs )
{
    if (this != &rhs) {
        threadlocal_buffer tmp { rhs };
        swap(tmp);
    }
    return *this;
}

template <typename real>
threadlocal_buffer<real> &threadlocal_buffer<real>::operator=( threadlocal_buffer &&rhs ) noexcept
{
    swap(rhs);
    return *this;
}

template <typename real>
threadlocal_buffer<real>::~threadlocal_buffer() { delete s; }

template <typename real>
void threadlocal_buffer<real>::swap( threadlocal_buffer &rhs ) noexcept
{
    std::swap(P, rhs.P);
    std::swap(s, rhs.s);
}

template <typename real>
void threadlocal_buffer<real>::initialise( const operator_data<real> &op )
{
    threadlocal_buffer tmp { op };
    swap(tmp);
}

template <typename real>
bool threadlocal_buffer<real>::compatible( const operator_data<real> &op ) const noexcept
{
    return s != nullptr && P >= op.order();
}

template class threadlocal_buffer<float>;
template class threadlocal_buffer<double>;

template <typename real> constexpr size_t default_order();
template <> constexpr size_t default_order<float >() { return 18; }  // handles.cpp:32-33
template <> constexpr size_t default_order<double>() { return 30; }

template <typename real>
operator_handle<real>::operator_handle():
    p { new operator_data<real>(std::min(default_order<real>(), size_t(sfmm_max_order()))) } {}

template <typename real>
operator_handle<real>::operator_handle( size_t P ): p { new operator_data<real>(P) } {}

template <typename real>
operator_handle<real>::~operator_handle() { delete p; }

template <typename real>
operator_handle<real>::operator_handle( const operator_handle &rhs ): p { new operator_data<real>(rhs.p->order()) } {}

template <typename real>
operator_handle<real> &operator_handle<real>::operator=( const operator_handle &rhs )
{
    if (&rhs != this) {
        pointer tmp = new operator_data<real>(rhs.p->order());
        delete p;
        p = tmp;
    }
    return *this;
}

template <typename real>
buffer_handle<real> operator_handle<real>::make_buffer() const { return buffer_handle<real> { *p }; }

template <typename real>
buffer_handle<real>::buffer_handle(): p { new threadlocal_buffer<real> {} } {}

template <typename real>
buffer_handle<real>::buffer_handle( const operator_data<real> &op ): p { new threadlocal_buffer<real> { op } } {}

template <typename real>
buffer_handle<real>::buffer_handle( const buffer_handle &rhs ): p { new threadlocal_buffer<real> { *rhs.p } } {}

template <typename real>
buffer_handle<real>::~buffer_handle() { delete p; }

template <typename real>
buffer_handle<real> &buffer_handle<real>::operator=( const buffer_handle &rhs )
{
    *p = *rhs.p;
    return *this;
}

template class operator_handle<float>;
template class operator_handle<double>;
template class buffer_handle<float>;
template class buffer_handle<double>;

// ---------------------------------------------------------------------------
// Translations.
// ---------------------------------------------------------------------------

namespace
{

const char *op_name( int op ) { return op == SFMM_M2M ? "m2m" : op == SFMM_M2L ? "m2l" : "l2l"; }
template <typename real> const char *type_name();
template <> const char *type_name<float >() { return "float"; }
template <> const char *type_name<double>() { return "double"; }

template <typename real>
void check( int op, const operator_data<real> &od, const threadlocal_buffer<real> &buf,
            const translation_info<real> *b, const translation_info<real> *e, bool check_buffer )
{
    const std::string who = std::string("solidfmm::") + op_name(op) + "<" + type_name<real>() + ">(): ";
    if (check_buffer && !buf.compatible(od))
        throw std::logic_error { who + "operator data and buffer incompatible" };
    for (const translation_info<real> *i = b; i != e; ++i) {
        if (i->source->dimension() != i->target->dimension()) throw std::logic_error { who + "dimension mismatch." };
        if (i->source->order() > od.order() || i->target->order() > od.order())
            throw std::out_of_range { who + "orders exceeding operator data." };
    }
}

// Run f(b, e) over [0, n) on up to 16 threads when the blocks move at
// least `bytes` in total; otherwise on the calling thread.
template <class F>
void parallel_blocks( size_t n, size_t bytes, F f )
{
    const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
    const unsigned nt = unsigned(std::min<size_t>({ size_t(16), size_t(hw), bytes / (2u << 20) + 1, n }));
    if (nt <= 1) {
        f(size_t(0), n);
        return;
    }
    std::vector<std::thread> pool;
    const size_t chunk = (n + nt - 1) / nt;
    for (unsigned t = 1; t < nt; ++t) {
        const size_t b = t * chunk, e = std::min(n, b + chunk);
        if (b < e) pool.emplace_back(f, b, e);
    }
    f(size_t(0), std::min(n, chunk));
    for (auto &th : pool) th.join();
}

// Translate on the GPU; returns an SFMM_* code.  The distinct source and
// target blocks are packed (in parallel, widened to FP64) straight into the
// staging object's pinned pools, translated there, and the targets are
// written back from the pinned target pool.
template <typename real>
int run( int op, const operator_data<real> &od, threadlocal_buffer<real> &buf, const translation_info<real> *b,
         const translation_info<real> *e )
{
    typename threadlocal_buffer<real>::staging &st = *buf.stage();
    st.items.clear();
    st.src_off.clear();
    st.dst_off.clear();
    st.src_blocks.clear();
    st.dst_blocks.clear();
    uint64_t src_len = 0, dst_len = 0;
    for (const translation_info<real> *i = b; i != e; ++i) {
        const solid<real> *s = i->source;
        solid<real> *t = i->target;
        const size_t pin = s->order(), pout = t->order(), dim = s->dimension();
        if (pin == 0 || pout == 0) continue;
        auto si = st.src_off.find(s);
        uint64_t so;
        if (si == st.src_off.end()) {
            so = src_len;
            st.src_off.emplace(s, so);
            st.src_blocks.push_back({ const_cast<real *>(s->memptr()), so, pin * (pin + 1) * dim });
            src_len += pin * (pin + 1) * dim;
        } else so = si->second;
        auto ti = st.dst_off.find(t);
        uint64_t to;
        if (ti == st.dst_off.end()) {
            to = dst_len;
            st.dst_off.emplace(t, to);
            st.dst_blocks.push_back({ t->memptr(), to, pout * (pout + 1) * dim });
            dst_len += pout * (pout + 1) * dim;
        } else to = ti->second;
        for (size_t d = 0; d < dim; ++d) {
            sfmm_item it {};
            it.src = so + d * pin * (pin + 1);
            it.dst = to + d * pout * (pout + 1);
            it.p_in = static_cast<uint32_t>(pin);
            it.p_out = static_cast<uint32_t>(pout);
            it.x = i->x;
            it.y = i->y;
            it.z = i->z;
            st.items.push_back(it);
        }
    }
    if (st.items.empty()) return SFMM_OK;
    if (!st.st) {
        const int rc = sfmm_staging_create(od.context(), &st.st);
        if (rc != SFMM_OK) return rc;
    }
    double *hs = nullptr, *hd = nullptr;
    int rc = sfmm_staging_reserve(st.st, src_len, dst_len, &hs, &hd);
    if (rc != SFMM_OK) return rc;
    auto pack = [&](const std::vector<typename threadlocal_buffer<real>::staging::block> &blocks, double *pool,
                    uint64_t len) {
        parallel_blocks(blocks.size(), len * sizeof(double), [&](size_t bb, size_t be) {
            for (size_t k = bb; k < be; ++k) {
                const auto &blk = blocks[k];
                double *q = pool + blk.off;
                for (uint64_t j = 0; j < blk.len; ++j) q[j] = static_cast<double>(blk.host[j]);
            }
        });
    };
    pack(st.src_blocks, hs, src_len);
    pack(st.dst_blocks, hd, dst_len);
    rc = sfmm_staging_translate(st.st, op, st.items.data(), st.items.size(), src_len, dst_len);
    if (rc != SFMM_OK) return rc;
    parallel_blocks(st.dst_blocks.size(), dst_len * sizeof(double), [&](size_t bb, size_t be) {
        for (size_t k = bb; k < be; ++k) {
            const auto &blk = st.dst_blocks[k];
            const double *q = hd + blk.off;
            for (uint64_t j = 0; j < blk.len; ++j) blk.host[j] = static_cast<real>(q[j]);
        }
    });
    return SFMM_OK;
}

template <typename real>
void raise( int rc, const operator_data<real> &od, int op )
{
    const std::string who = std::string("solidfmm::") + op_name(op) + "<" + type_name<real>() + ">(): ";
    const std::string msg = who + sfmm_last_error(od.context());
    switch (rc) {
    case SFMM_OK: return;
    case SFMM_ERR_ALLOC: throw std::bad_alloc {};
    case SFMM_ERR_RANGE: throw std::out_of_range { msg };
    case SFMM_ERR_OVERFLOW: throw std::overflow_error { msg };
    case SFMM_ERR_LOGIC: throw std::logic_error { msg };
    default: throw std::runtime_error { msg };
    }
}

template <typename real>
void checked( int op, const operator_data<real> &od, threadlocal_buffer<real> &buf, const translation_info<real> *b,
              const translation_info<real> *e )
{
    check(op, od, buf, b, e, true);
    raise(run(op, od, buf, b, e), od, op);
}

// The unchecked entry points cannot report errors; a failing GPU call ends
// the process with a message instead of writing partial results silently.
template <typename real>
void unchecked( int op, const operator_data<real> &od, threadlocal_buffer<real> &buf, const translation_info<real> *b,
                const translation_info<real> *e ) noexcept
{
    const int rc = run(op, od, buf, b, e);
    if (rc != SFMM_OK) {
        std::fprintf(stderr, "solidfmm::%s_unchecked: %s\n", op_name(op), sfmm_last_error(od.context()));
        std::abort();
    }
}

// The classic (howmany, arrays) form: validate every entry first, then run
// the list (translations.cpp:663-690).
template <typename real>
void classic( int op, const operator_data<real> &od, threadlocal_buffer<real> &buf, size_t howmany,
              const solid<real> *const *in, solid<real> *const *out, const real *x, const real *y, const real *z )
{
    std::vector<translation_info<real>> list(howmany);
    for (size_t i = 0; i < howmany; ++i) list[i] = translation_info<real> { in[i], out[i], x[i], y[i], z[i] };
    checked(op, od, buf, list.data(), list.data() + howmany);
}

}  // namespace

#define SFMM_DROPIN_OPS(NAME, OP)                                                                                   \
    void NAME( const operator_data<float> &op, threadlocal_buffer<float> &buf, size_t howmany,                     \
               const solid<float> *const *const a, solid<float> *const *const b, const float *x, const float *y, \
               const float *z )                                                                                    \
    {                                                                                                              \
        classic<float>(OP, op, buf, howmany, a, b, x, y, z);                                                       \
    }                                                                                                              \
    void NAME( const operator_data<double> &op, threadlocal_buffer<double> &buf, size_t howmany,                   \
               const solid<double> *const *const a, solid<double> *const *const b, const double *x,               \
               const double *y, const double *z )                                                                  \
    {                                                                                                              \
        classic<double>(OP, op, buf, howmany, a, b, x, y, z);                                                      \
    }                                                                                                              \
    void NAME( const operator_data<float> &op, threadlocal_buffer<float> &buf, const translation_info<float> *b,   \
               const translation_info<float> *e )                                                                  \
    {                                                                                                              \
        checked<float>(OP, op, buf, b, e);                                                                         \
    }                                                                                                              \
    void NAME( const operator_data<double> &op, threadlocal_buffer<double> &buf,                                   \
               const translation_info<double> *b, const translation_info<double> *e )                             \
    {                                                                                                              \
        checked<double>(OP, op, buf, b, e);                                                                        \
    }                                                                                                              \
    void NAME##_unchecked( const operator_data<float> &op, threadlocal_buffer<float> &buf,                        \
                           const translation_info<float> *b, const translation_info<float> *e ) noexcept           \
    {                                                                                                              \
        unchecked<float>(OP, op, buf, b, e);                                                                       \
    }                                                                                                              \
    void NAME##_unchecked( const operator_data<double> &op, threadlocal_buffer<double> &buf,                      \
                           const translation_info<double> *b, const translation_info<double> *e ) noexcept         \
    {                                                                                                              \
        unchecked<double>(OP, op, buf, b, e);                                                                      \
    }

SFMM_DROPIN_OPS(m2m, SFMM_M2M)
SFMM_DROPIN_OPS(m2l, SFMM_M2L)
SFMM_DROPIN_OPS(l2l, SFMM_L2L)
#undef SFMM_DROPIN_OPS

// ---------------------------------------------------------------------------
// multithreaded_translator: one GPU context; calls are serialised.
// ---------------------------------------------------------------------------

template <typename real>
class multithreaded_translator<real>::impl
{
public:
    explicit impl( size_t P ): op { P }, buf { op } {}
    std::mutex m;
    operator_data<real> op;
    threadlocal_buffer<real> buf;
};

template <typename real>
multithreaded_translator<real>::multithreaded_translator( size_t P ): p { new impl(P) } {}

template <typename real>
multithreaded_translator<real>::multithreaded_translator( size_t P, hwloc_topology *, cpuset_t *, cpuset_t * ):
    p { new impl(P) } {}

template <typename real>
multithreaded_translator<real>::multithreaded_translator( multithreaded_translator &&rhs ) noexcept: p { rhs.p }
{
    rhs.p = nullptr;
}

template <typename real>
multithreaded_translator<real> &multithreaded_translator<real>::operator=( multithreaded_translator &&rhs ) noexcept
{
    std::swap(p, rhs.p);
    return *this;
}

template <typename real>
multithreaded_translator<real>::~multithreaded_translator() { delete p; }

template <typename real>
void multithreaded_translator<real>::schedule_equally_among_threads() noexcept {}

template <typename real>
void multithreaded_translator<real>::schedule_by_measuring_throughput() {}

#define SFMM_MT_OP(NAME, OP)                                                                    \
    template <typename real>                                                                    \
    void multithreaded_translator<real>::NAME( const item *b, const item *e ) const            \
    {                                                                                           \
        std::lock_guard<std::mutex> lock(p->m);                                                 \
        check(OP, p->op, p->buf, b, e, false);                                                  \
        raise(run(OP, p->op, p->buf, b, e), p->op, OP);                                         \
    }                                                                                           \
    template <typename real>                                                                    \
    void multithreaded_translator<real>::NAME##_unchecked( const item *b, const item *e ) const noexcept \
    {                                                                                           \
        std::lock_guard<std::mutex> lock(p->m);                                                 \
        unchecked(OP, p->op, p->buf, b, e);                                                     \
    }

SFMM_MT_OP(m2m, SFMM_M2M)
SFMM_MT_OP(m2l, SFMM_M2L)
SFMM_MT_OP(l2l, SFMM_L2L)
#undef SFMM_MT_OP

template class multithreaded_translator<float>;
template class multithreaded_translator<double>;

}  // namespace solidfmm
