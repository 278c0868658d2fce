/*
 * Extended-precision truth for the translation operators (M2M, M2L, L2L)
 * -- TEST INFRASTRUCTURE ONLY.
 *
 * This file is part of the oracle.  Only the test-suite, the smoke check and the
 * CPU-baseline leg of the benchmark may load it; the product library never
 * links, calls or falls back to it.
 *
 * It evaluates the operators in IEEE binary128 (__float128: 113-bit
 * significand, unit roundoff 9.6e-35, libquadmath) from FP64 inputs, which
 * are exact in binary128, by two independent routes:
 *
 *   method 0, naive: the O(P^4) double sums of doc/solidfmm.tex:1404-1433
 *     (M2M sum over l = -k..k) over the solid harmonics R and S, evaluated by
 *     the recurrences of harmonics.cpp:31-128 (the diagonal first, then
 *     down each column m).
 *   method 1, rotation: the O(P^3) algorithm of translations.cpp:242-471 --
 *     Euler data by repeated complex products (microkernel_generic.cpp:36-97),
 *     z rotations (:99-140), the folded real/imaginary swap blocks
 *     (operator_data.cpp:47-147) built from the Dehnen recurrence
 *     (swap_matrix.cpp:148-171), the z translations (:216-283), and back.
 *
 * The two routes share only the input coefficients and the definition of the
 * operators, so their agreement to ~1e-28 (tests/test_truth.py) pins both.
 * Results come back as double-double pairs (hi, lo) with hi + lo equal to the
 * binary128 value rounded to ~106 bits: the error of an FP64 result against
 * the truth is then (got - hi) - lo, measured far below FP64's own rounding.
 *
 * L2L with P_in >= P_out + 2 uses the mathematically correct Euler tables
 * (as far as they are read), as the FP64 oracle does: the reference computes
 * them only to min(P_in, P_out) (translations.cpp:414-417) and reads them to
 * P_in (translations.cpp:198-224).
 */
#include <pthread.h>
#include <quadmath.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef __float128 qr;

#define TQ_PMAX 100

typedef struct tq_item {
    uint64_t src;
    uint64_t dst;
    uint32_t p_in;
    uint32_t p_out;
    double x, y, z;
} tq_item;

enum { OP_M2M = 0, OP_M2L = 1, OP_L2L = 2 };

#define RE(n, m) ((n) * ((n) + 1) + (m))
#define IM(n, m) (((n) + 1) * ((n) + 1) + (m))

/* ------------------------------------------------------------------------ */
/* Route 0: solid harmonics and the naive double sums.                      */
/* ------------------------------------------------------------------------ */

/* R_n^m(x), n < P, in the solid layout (harmonics.cpp:31-79). */
static void harm_R(int P, qr x, qr y, qr z, qr *o)
{
    for (int i = 0; i < P * (P + 1); ++i) o[i] = 0;
    if (P == 0) return;
    o[RE(0, 0)] = 1;
    if (P == 1) return;
    o[RE(1, 0)] = z;
    o[RE(1, 1)] = x / 2;
    o[IM(1, 1)] = y / 2;
    for (int m = 2; m < P; ++m) {
        qr a = o[RE(m - 1, m - 1)], b = o[IM(m - 1, m - 1)];
        o[RE(m, m)] = (a * x - b * y) / (2 * m);
        o[IM(m, m)] = (a * y + b * x) / (2 * m);
    }
    for (int m = 1; m < P - 1; ++m) {
        o[RE(m + 1, m)] = z * o[RE(m, m)];
        o[IM(m + 1, m)] = z * o[IM(m, m)];
    }
    qr r2 = x * x + y * y + z * z;
    for (int m = 0; m < P - 2; ++m)
        for (int n = m + 2; n < P; ++n) {
            qr f = (qr)((n + m) * (n - m));
            o[RE(n, m)] = ((2 * n - 1) * z * o[RE(n - 1, m)] - r2 * o[RE(n - 2, m)]) / f;
            o[IM(n, m)] = ((2 * n - 1) * z * o[IM(n - 1, m)] - r2 * o[IM(n - 2, m)]) / f;
        }
}

/* S_n^m(x), n < P (harmonics.cpp:81-128). */
static void harm_S(int P, qr x, qr y, qr z, qr *o)
{
    for (int i = 0; i < P * (P + 1); ++i) o[i] = 0;
    if (P == 0) return;
    qr r2inv = 1 / (x * x + y * y + z * z);
    o[RE(0, 0)] = sqrtq(r2inv);
    if (P == 1) return;
    o[RE(1, 0)] = z * r2inv * o[RE(0, 0)];
    o[RE(1, 1)] = x * r2inv * o[RE(0, 0)];
    o[IM(1, 1)] = y * r2inv * o[RE(0, 0)];
    for (int m = 2; m < P; ++m) {
        qr a = o[RE(m - 1, m - 1)], b = o[IM(m - 1, m - 1)], f = (2 * m - 1) * r2inv;
        o[RE(m, m)] = (a * x - b * y) * f;
        o[IM(m, m)] = (a * y + b * x) * f;
    }
    for (int m = 1; m < P - 1; ++m) {
        qr f = (2 * m + 1) * z * r2inv;
        o[RE(m + 1, m)] = f * o[RE(m, m)];
        o[IM(m + 1, m)] = f * o[IM(m, m)];
    }
    for (int m = 0; m < P - 2; ++m)
        for (int n = m + 2; n < P; ++n) {
            qr c = (qr)((n + m - 1) * (n - m - 1));
            o[RE(n, m)] = ((2 * n - 1) * z * o[RE(n - 1, m)] - c * o[RE(n - 2, m)]) * r2inv;
            o[IM(n, m)] = ((2 * n - 1) * z * o[IM(n - 1, m)] - c * o[IM(n - 2, m)]) * r2inv;
        }
}

/* C_n^m for any m: C_n^{-m} = (-1)^m conj(C_n^m) (doc/solidfmm.tex:1096-1106);
 * zero outside |m| <= n < P. */
static void coef_q(const qr *c, int P, int n, int m, qr *re, qr *im)
{
    if (n < 0 || n >= P || m > n || -m > n) { *re = 0; *im = 0; return; }
    if (m >= 0) { *re = c[RE(n, m)]; *im = c[IM(n, m)]; return; }
    qr s = ((-m) & 1) ? -1 : 1;
    *re = s * c[RE(n, -m)];
    *im = -s * c[IM(n, -m)];
}

static void coef_d(const double *c, int P, int n, int m, qr *re, qr *im)
{
    if (n < 0 || n >= P || m > n || -m > n) { *re = 0; *im = 0; return; }
    if (m >= 0) { *re = c[RE(n, m)]; *im = c[IM(n, m)]; return; }
    qr s = ((-m) & 1) ? -1 : 1;
    *re = s * (qr)c[RE(n, -m)];
    *im = -s * (qr)c[IM(n, -m)];
}

static int naive_one(int op, int Pi, int Po, qr x, qr y, qr z, const double *src, qr *dst)
{
    if (Pi == 0 || Po == 0) return 0;
    int Ph = Pi + Po;
    qr *h = (qr *)malloc(sizeof(qr) * (size_t)Ph * (Ph + 1));
    if (!h) return -1;
    if (op == OP_M2L) harm_S(Ph, x, y, z, h);
    else if (op == OP_M2M) harm_R(Ph, -x, -y, -z, h);
    else harm_R(Ph, x, y, z, h);
    for (int n = 0; n < Po; ++n)
        for (int m = 0; m <= n; ++m) {
            qr ar = 0, ai = 0;
            for (int k = (op == OP_L2L ? n : 0); k < Pi; ++k)
                for (int l = -k; l <= k; ++l) {
                    qr cr, ci, hr, hi;
                    coef_d(src, Pi, k, l, &cr, &ci);
                    if (op == OP_M2L) { /* (-1)^n conj(M_k^l) S_{n+k}^{m+l} */
                        coef_q(h, Ph, n + k, m + l, &hr, &hi);
                        ar += cr * hr + ci * hi;
                        ai += cr * hi - ci * hr;
                    } else if (op == OP_M2M) { /* M_k^l R_{n-k}^{m-l}(-r) */
                        coef_q(h, Ph, n - k, m - l, &hr, &hi);
                        ar += cr * hr - ci * hi;
                        ai += cr * hi + ci * hr;
                    } else { /* L_k^l conj(R_{k-n}^{l-m}(r)) */
                        coef_q(h, Ph, k - n, l - m, &hr, &hi);
                        ar += cr * hr + ci * hi;
                        ai += ci * hr - cr * hi;
                    }
                }
            if (op == OP_M2L && (n & 1)) { ar = -ar; ai = -ai; }
            dst[RE(n, m)] += ar;
            dst[IM(n, m)] += ai;
        }
    free(h);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Route 1: the rotation-based algorithm.                                   */
/* ------------------------------------------------------------------------ */

/* Full swap matrices B_n^{m,l}, -n <= m,l <= n, n < TQ_PMAX, by the Dehnen
 * recurrence (swap_matrix.cpp:148-171).  Entries are dyadic rationals; in
 * binary128 they are exact far beyond the orders used here. */
static qr *swap_tab;
static qr *fold_tab[4][TQ_PMAX];
static qr fact[2 * TQ_PMAX], invfact[2 * TQ_PMAX];
static pthread_once_t tables_once = PTHREAD_ONCE_INIT;
static int tables_rc = -1;

static size_t swap_base(int n) { return (size_t)n * (4 * (size_t)n * n - 1) / 3; }

static qr B(int n, int m, int l)
{
    if (n < 0 || l < -n || l > n || m < -n || m > n) return 0;
    return swap_tab[swap_base(n) + (size_t)(m + n) + (size_t)(l + n) * (2 * n + 1)];
}

static void Bset(int n, int m, int l, qr v) { swap_tab[swap_base(n) + (size_t)(m + n) + (size_t)(l + n) * (2 * n + 1)] = v; }

/* Folded block acting on the stored m >= 0 coefficients (operator_data.cpp:
 * 47-147).  kind 0: real part, 1: imaginary part, 2, 3: their transposes
 * (used by the L-type transforms).  Entry [m][l] vanishes unless (l+m+n) is
 * even (real) or odd (imaginary). */
static qr folded(int kind, int n, int m, int l)
{
    int real_part = (kind == 0 || kind == 2), tr = (kind >= 2);
    int par = (l + m + n) & 1;
    if (real_part ? par : !par) return 0;
    qr bp = tr ? B(n, m, l) : B(n, l, m);
    qr bm = tr ? B(n, m, -l) : B(n, -l, m);
    if (real_part) {
        if (l == 0) return tr ? B(n, m, 0) : B(n, 0, m);
        return (l & 1) ? bp - bm : bp + bm;
    }
    return (l & 1) ? bp + bm : bp - bm;
}

static void build_tables(void)
{
    swap_tab = (qr *)malloc(sizeof(qr) * swap_base(TQ_PMAX));
    if (!swap_tab) return;
    for (int n = 0; n < TQ_PMAX; ++n) {
        if (n == 0) { swap_tab[0] = 1; continue; }
        for (int l = -n; l <= n; ++l) Bset(n, 0, l, (B(n - 1, 0, l - 1) - B(n - 1, 0, l + 1)) / 2);
        for (int m = 0; m < n; ++m)
            for (int l = -n; l <= n; ++l) {
                Bset(n, m + 1, l, (B(n - 1, m, l - 1) + B(n - 1, m, l + 1)) / 2 + B(n - 1, m, l));
                Bset(n, -m - 1, l, (B(n - 1, -m, l - 1) + B(n - 1, -m, l + 1)) / 2 - B(n - 1, -m, l));
            }
    }
    for (int k = 0; k < 4; ++k)
        for (int n = 0; n < TQ_PMAX; ++n) {
            qr *a = (qr *)malloc(sizeof(qr) * (size_t)(n + 1) * (n + 1));
            if (!a) return;
            for (int m = 0; m <= n; ++m)
                for (int l = 0; l <= n; ++l) a[m * (n + 1) + l] = folded(k, n, m, l);
            fold_tab[k][n] = a;
        }
    fact[0] = 1;
    for (int i = 1; i < 2 * TQ_PMAX; ++i) fact[i] = i * fact[i - 1];
    for (int i = 0; i < 2 * TQ_PMAX; ++i) invfact[i] = 1 / fact[i];
    tables_rc = 0;
}

static int tables_ready(void)
{
    pthread_once(&tables_once, build_tables);
    return tables_rc;
}

typedef struct euler_q {
    qr r[TQ_PMAX + 2], rinv[TQ_PMAX + 2], ca[TQ_PMAX + 2], sa[TQ_PMAX + 2], cb[TQ_PMAX + 2], sb[TQ_PMAX + 2];
} euler_q;

/* cos(alpha) = y/rho, sin(alpha) = x/rho, cos(beta) = z/r, sin(beta) =
 * -rho/r; integer powers by repeated complex products (doc/solidfmm.tex:
 * 2493-2562). */
static void euler(qr x, qr y, qr z, int k, euler_q *e)
{
    e->r[0] = e->rinv[0] = e->ca[0] = e->cb[0] = 1;
    e->sa[0] = e->sb[0] = 0;
    qr rho = sqrtq(x * x + y * y), r = sqrtq(x * x + y * y + z * z);
    e->r[1] = r;
    e->rinv[1] = 1 / r;
    e->ca[1] = rho == 0 ? 1 : y / rho;
    e->sa[1] = rho == 0 ? 0 : x / rho;
    e->cb[1] = z / r;
    e->sb[1] = -rho / r;
    for (int i = 2; i < k; ++i) {
        e->r[i] = e->r[1] * e->r[i - 1];
        e->rinv[i] = e->rinv[1] * e->rinv[i - 1];
        e->ca[i] = e->ca[i - 1] * e->ca[1] - e->sa[i - 1] * e->sa[1];
        e->sa[i] = e->sa[i - 1] * e->ca[1] + e->ca[i - 1] * e->sa[1];
        e->cb[i] = e->cb[i - 1] * e->cb[1] - e->sb[i - 1] * e->sb[1];
        e->sb[i] = e->sb[i - 1] * e->cb[1] + e->cb[i - 1] * e->sb[1];
    }
}

/* (re, im) <- scale * e^{+-i m theta} (re, im), m = 0..k-1. */
static void rotscale(const qr *c, const qr *s, qr scale, qr *re, qr *im, int k, int forward, int mag)
{
    for (int i = 0; i < k; ++i) {
        qr a = re[i], b = im[i];
        if (mag) { /* |z| propagated through |rotation| */
            qr ca = fabsq(c[i]), sa = fabsq(s[i]);
            re[i] = fabsq(scale) * (ca * a + sa * b);
            im[i] = fabsq(scale) * (sa * a + ca * b);
        } else if (forward) {
            re[i] = scale * (c[i] * a - s[i] * b);
            im[i] = scale * (s[i] * a + c[i] * b);
        } else {
            re[i] = scale * (c[i] * a + s[i] * b);
            im[i] = scale * (c[i] * b - s[i] * a);
        }
    }
}

static void swap_apply(const qr *A, int n, qr *v, int mag)
{
    qr out[TQ_PMAX];
    for (int m = 0; m <= n; ++m) {
        qr acc = 0;
        for (int l = 0; l <= n; ++l) acc += (mag ? fabsq(A[m * (n + 1) + l]) : A[m * (n + 1) + l]) * v[l];
        out[m] = acc;
    }
    memcpy(v, out, sizeof(qr) * (size_t)(n + 1));
}

static void swap_pair(int transposed, int n, qr *re, qr *im, int mag)
{
    swap_apply(fold_tab[transposed ? 2 : 0][n], n, re, mag);
    swap_apply(fold_tab[transposed ? 3 : 1][n], n, im, mag);
}

#define TRI(m, n) ((size_t)(m) * TQ_PMAX + (size_t)(n))

typedef struct work_q {
    qr tin_re[TQ_PMAX * TQ_PMAX], tin_im[TQ_PMAX * TQ_PMAX];
    qr tout_re[TQ_PMAX * TQ_PMAX], tout_im[TQ_PMAX * TQ_PMAX];
    euler_q e;
} work_q;

/* mag = 1: the same steps on magnitudes, |A| |v| for every linear map A
 * (inputs |src|, |cos|, |sin|, all signs dropped).  The result M bounds
 * every intermediate sum's terms, and u * M is the first-order scale of the
 * rounding error of any FP64 evaluation of the algorithm, whatever its
 * summation order. */
static void rotation_one(int op, int Pi, int Po, qr x, qr y, qr z, const double *src, qr *dst, work_q *w, int mag)
{
    if (Pi == 0 || Po == 0) return;
    if (op != OP_M2L && x == 0 && y == 0 && z == 0) {
        /* zero shift: plain copy-add (translations.cpp:539-547, 596-604) */
        int P = Pi < Po ? Pi : Po;
        for (int i = 0; i < P * (P + 1); ++i) dst[i] += mag ? fabsq((qr)src[i]) : (qr)src[i];
        return;
    }
    if (op == OP_M2M) { x = -x; y = -y; z = -z; }
    int Pend = Pi < Po ? Pi : Po, Pmax = Pi > Po ? Pi : Po;
    euler(x, y, z, (op == OP_L2L ? (Pi > Pend + 1 ? Pi : Pend + 1) : Pmax + 1), &w->e);
    const euler_q *e = &w->e;
    qr re[TQ_PMAX], im[TQ_PMAX];

    for (int n = 0; n < Pi; ++n) {
        for (int m = 0; m <= n; ++m) {
            re[m] = mag ? fabsq((qr)src[RE(n, m)]) : (qr)src[RE(n, m)];
            im[m] = mag ? fabsq((qr)src[IM(n, m)]) : (qr)src[IM(n, m)];
        }
        int tr = (op == OP_L2L);
        rotscale(e->ca, e->sa, op == OP_L2L ? e->r[n] : e->rinv[n + 1], re, im, n + 1, 1, mag);
        swap_pair(tr, n, re, im, mag);
        rotscale(e->cb, e->sb, 1, re, im, n + 1, 1, mag);
        swap_pair(tr, n, re, im, mag);
        for (int m = 0; m <= n; ++m) {
            w->tin_re[TRI(m, n)] = re[m];
            w->tin_im[TRI(m, n)] = im[m];
        }
    }

    int Nend = (op == OP_L2L) ? Pend : Po;
    for (int m = 0; m < Pend; ++m)
        for (int N = m; N < Nend; ++N) {
            qr ar = 0, ai = 0;
            if (op == OP_M2L) {
                for (int K = m; K < Pi; ++K) {
                    ar += fact[N + K] * w->tin_re[TRI(m, K)];
                    ai += fact[N + K] * w->tin_im[TRI(m, K)];
                }
                if (((N + m) & 1) && !mag) { ar = -ar; ai = -ai; }
            } else if (op == OP_M2M) {
                for (int K = m; K <= N && K < Pi; ++K) {
                    ar += invfact[N - K] * w->tin_re[TRI(m, K)];
                    ai += invfact[N - K] * w->tin_im[TRI(m, K)];
                }
            } else {
                for (int K = N; K < Pi; ++K) {
                    ar += invfact[K - N] * w->tin_re[TRI(m, K)];
                    ai += invfact[K - N] * w->tin_im[TRI(m, K)];
                }
            }
            w->tout_re[TRI(m, N)] = ar;
            w->tout_im[TRI(m, N)] = ai;
        }

    for (int N = 0; N < Nend; ++N) {
        for (int m = 0; m <= N; ++m) {
            int have = (op == OP_L2L) ? (m < Pend) : (m < Pi);
            re[m] = have ? w->tout_re[TRI(m, N)] : 0;
            im[m] = have ? w->tout_im[TRI(m, N)] : 0;
        }
        int tr = (op != OP_M2M);
        swap_pair(tr, N, re, im, mag);
        rotscale(e->cb, e->sb, 1, re, im, N + 1, 0, mag);
        swap_pair(tr, N, re, im, mag);
        rotscale(e->ca, e->sa, op == OP_M2M ? e->r[N + 1] : e->rinv[N], re, im, N + 1, 0, mag);
        for (int m = 0; m <= N; ++m) {
            dst[RE(N, m)] += re[m];
            dst[IM(N, m)] += im[m];
        }
    }
}

/* ------------------------------------------------------------------------ */
/* Driver: items in list order per target; threads own disjoint targets.    */
/* ------------------------------------------------------------------------ */

typedef struct job {
    int op, method, nthreads, tid, rc;
    const tq_item *items;
    size_t n;
    const double *src;
    qr *acc, *mag;
} job;

static void *run_job(void *arg)
{
    job *j = (job *)arg;
    work_q *w = (work_q *)malloc(sizeof(work_q));
    if (!w) { j->rc = -3; return NULL; }
    for (size_t i = 0; i < j->n; ++i) {
        const tq_item *it = &j->items[i];
        /* targets partitioned by a hash of the target offset, so each
         * target's items run on one thread in list order */
        uint64_t h = it->dst * 0x9E3779B97F4A7C15ull;
        if ((int)((h >> 32) % (uint64_t)j->nthreads) != j->tid) continue;
        if (it->p_in > TQ_PMAX - 2 || it->p_out > TQ_PMAX - 2) { j->rc = -2; break; }
        if (j->method == 0) {
            if (naive_one(j->op, (int)it->p_in, (int)it->p_out, it->x, it->y, it->z, j->src + it->src,
                          j->acc + it->dst) != 0) { j->rc = -3; break; }
        } else {
            rotation_one(j->op, (int)it->p_in, (int)it->p_out, it->x, it->y, it->z, j->src + it->src,
                         j->acc + it->dst, w, 0);
        }
        if (j->mag)
            rotation_one(j->op, (int)it->p_in, (int)it->p_out, it->x, it->y, it->z, j->src + it->src,
                         j->mag + it->dst, w, 1);
    }
    free(w);
    return NULL;
}

/* dst = dst0 + T(src) for every item in list order, in binary128; the
 * result is returned as hi + lo (each binary128 value split into two
 * doubles).  dst0 may be NULL (zero targets).  method 0: naive sums,
 * 1: rotations.  mag, if not NULL, receives |dst0| + the sum over items of
 * the magnitude propagation of the rotation route (rotation_one, mag = 1).
 * Returns 0, or -1 bad argument, -2 order too large, -3 out of memory. */
int truth_translate(int op, const tq_item *items, size_t n, const double *src, const double *dst0, size_t dst_len,
                    double *hi, double *lo, double *mag, int method, int threads)
{
    if (op < 0 || op > 2 || method < 0 || method > 1) return -1;
    if (tables_ready() != 0) return -3;
    for (size_t i = 0; i < n; ++i)
        if (items[i].dst + (size_t)items[i].p_out * (items[i].p_out + 1) > dst_len) return -1;
    qr *acc = (qr *)malloc(sizeof(qr) * (dst_len ? dst_len : 1));
    if (!acc) return -3;
    for (size_t i = 0; i < dst_len; ++i) acc[i] = dst0 ? (qr)dst0[i] : 0;
    qr *macc = NULL;
    if (mag) {
        macc = (qr *)malloc(sizeof(qr) * (dst_len ? dst_len : 1));
        if (!macc) { free(acc); return -3; }
        for (size_t i = 0; i < dst_len; ++i) macc[i] = dst0 ? fabsq((qr)dst0[i]) : 0;
    }
    if (threads < 1) threads = 1;
    if (threads > 256) threads = 256;
    job jobs[256];
    pthread_t tids[256];
    int rc = 0;
    for (int t = 0; t < threads; ++t) {
        jobs[t] = (job){op, method, threads, t, 0, items, n, src, acc, macc};
        if (threads == 1) run_job(&jobs[t]);
        else if (pthread_create(&tids[t], NULL, run_job, &jobs[t]) != 0) { rc = -3; threads = t; break; }
    }
    if (threads > 1)
        for (int t = 0; t < threads; ++t) pthread_join(tids[t], NULL);
    for (int t = 0; t < threads; ++t)
        if (jobs[t].rc && !rc) rc = jobs[t].rc;
    for (size_t i = 0; i < dst_len; ++i) {
        double h = (double)acc[i];
        hi[i] = h;
        lo[i] = (double)(acc[i] - (qr)h);
        if (mag) mag[i] = (double)macc[i];
    }
    free(acc);
    free(macc);
    return rc;
}

/* Harmonics in binary128 as hi/lo pairs (kind 0: R, 1: S), for the
 * closed-form checks. */
int truth_harmonic(int kind, int P, double x, double y, double z, double *hi, double *lo)
{
    if (P < 0 || P > 2 * TQ_PMAX || kind < 0 || kind > 1) return -1;
    qr *o = (qr *)malloc(sizeof(qr) * (size_t)(P * (P + 1) + 1));
    if (!o) return -3;
    if (kind == 0) harm_R(P, x, y, z, o);
    else harm_S(P, x, y, z, o);
    for (int i = 0; i < P * (P + 1); ++i) {
        hi[i] = (double)o[i];
        lo[i] = (double)(o[i] - (qr)hi[i]);
    }
    free(o);
    return 0;
}
