/*
 * CPU oracle for the translation operators (M2M, M2L, L2L) on solid
 * harmonics -- TEST INFRASTRUCTURE ONLY.
 *
 * This file is the oracle of the HIP product path.  Only the test-suite,
 * the smoke check and the CPU-baseline leg of the benchmark may load it.  The
 * product library never links, calls or falls back to it.
 *
 * It restates, one translation at a time and in plain scalar FP64, the
 * algorithm of the reference library vorticle/solidfmm 1.5.  Citations are
 * file:line in that reference:
 *
 *   swap matrices B_n^{m,l}        swap_matrix.cpp:148-171 (Dehnen recurrence)
 *   folded swap matrices           operator_data.cpp:47-147
 *   Euler data                     microkernel_generic.cpp:36-97
 *   rotscale                       microkernel_generic.cpp:99-140
 *   swap                           microkernel_generic.cpp:142-214
 *   zm2l / zm2m                    microkernel_generic.cpp:216-283
 *   forward/backward transforms    translations.cpp:36-239
 *   m2l_impl / m2m_impl / l2l_impl translations.cpp:242-471
 *   zero-shift shortcut            translations.cpp:539-547, 596-604
 *   harmonics R, S                 harmonics.cpp:31-128
 *   naive O(P^4) formulas          doc/solidfmm.tex:1404-1433
 *
 * The reference batches `cols` translations per call (SIMD lanes) and pads
 * short lanes with zeros; a lane's result does not depend on its neighbours,
 * so the per-translation restatement computes the same numbers up to the
 * rounding of the summation order.
 *
 * Deliberate deviation: for L2L with P_in >= P_out + 2 the reference computes
 * the Euler tables only up to index min(P_in,P_out) (translations.cpp:414-417)
 * but forward_transform_L reads r^n and the rotation powers for every input
 * degree n < P_in (translations.cpp:198-224), i.e. zero-initialised or stale
 * buffer contents.  The oracle computes the tables as far as they are read,
 * which is the mathematically correct L2L (checked against the naive formula).
 *
 * Pinning: see tests/test_oracle.py (B_4 known answer, golden vectors from
 * the compiled reference, naive O(P^4) formulas, physical convergence,
 * L2L exactness and composition, rotation identities).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define ORC_PMAX 88

typedef struct orc_item {
    uint64_t src;
    uint64_t dst;
    uint32_t p_in;
    uint32_t p_out;
    double x, y, z;
} orc_item;

/* ------------------------------------------------------------------------ */
/* Swap matrices B_n^{m,l}, n < nmax, stored densely per n as (2n+1)^2.      */
/* swap_matrix.cpp:148-171: B_0 = 1; row m=0 from 0.5*(b(l-1) - b(l+1));    */
/* rows m+1 and -m-1 from rows m and -m of the previous degree.             */
/* ------------------------------------------------------------------------ */

static double *swap_tab = NULL;
static int swap_nmax = 0;

static size_t swap_base(int n) { return (size_t)n * (4 * (size_t)n * n - 1) / 3; }

static double swap_get(int n, int m, int l)
{
    if (n < 0 || l < -n || l > n || m < -n || m > n) return 0.0;
    return swap_tab[swap_base(n) + (size_t)(m + n) + (size_t)(l + n) * (2 * n + 1)];
}

static void swap_set(int n, int m, int l, double v)
{
    swap_tab[swap_base(n) + (size_t)(m + n) + (size_t)(l + n) * (2 * n + 1)] = v;
}

static int swap_reserve(int nmax)
{
    if (nmax <= swap_nmax) return 0;
    double *t = (double *)realloc(swap_tab, sizeof(double) * swap_base(nmax));
    if (!t) return -1;
    swap_tab = t;
    for (int n = swap_nmax; n < nmax; ++n) {
        if (n == 0) { swap_tab[0] = 1.0; continue; }
        for (int l = -n; l <= n; ++l)
            swap_set(n, 0, l, 0.5 * (swap_get(n - 1, 0, l - 1) - swap_get(n - 1, 0, l + 1)));
        for (int m = 0; m < n; ++m)
            for (int l = -n; l <= n; ++l) {
                swap_set(n, m + 1, l,
                         0.5 * (swap_get(n - 1, m, l - 1) + swap_get(n - 1, m, l + 1)) + swap_get(n - 1, m, l));
                swap_set(n, -m - 1, l,
                         0.5 * (swap_get(n - 1, -m, l - 1) + swap_get(n - 1, -m, l + 1)) - swap_get(n - 1, -m, l));
            }
    }
    swap_nmax = nmax;
    return 0;
}

/* Folded (n+1)x(n+1) matrices acting on the stored m >= 0 coefficients,
 * operator_data.cpp:47-147.  kind 0: real part, 1: imaginary part,
 * 2: real part transposed, 3: imaginary part transposed.  Entry [m][l] is
 * non-zero only where (l+m+n) is even (real) or odd (imaginary). */
static double folded(int kind, int n, int m, int l)
{
    int real_part = (kind == 0 || kind == 2);
    int tr = (kind >= 2);
    int par = (l + m + n) & 1;
    if (real_part ? par : !par) return 0.0;
    double bp = tr ? swap_get(n, m, l) : swap_get(n, l, m);
    double bm = tr ? swap_get(n, m, -l) : swap_get(n, -l, m);
    if (real_part) {
        if (l == 0) return tr ? swap_get(n, m, 0) : swap_get(n, 0, m);
        return (l & 1) ? bp - bm : bp + bm;
    }
    return (l & 1) ? bp + bm : bp - bm;
}

/* Dense folded matrices cached per degree. */
static double *fold_tab[4][ORC_PMAX];

static int fold_reserve(int P)
{
    if (swap_reserve(P) != 0) return -1;
    for (int k = 0; k < 4; ++k)
        for (int n = 0; n < P; ++n) {
            if (fold_tab[k][n]) continue;
            double *a = (double *)malloc(sizeof(double) * (size_t)(n + 1) * (n + 1));
            if (!a) return -1;
            for (int m = 0; m <= n; ++m)
                for (int l = 0; l <= n; ++l) a[m * (n + 1) + l] = folded(k, n, m, l);
            fold_tab[k][n] = a;
        }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Primitives (microkernel_generic.cpp semantics, one translation).         */
/* ------------------------------------------------------------------------ */

typedef struct euler_t {
    double r[ORC_PMAX + 2], rinv[ORC_PMAX + 2];
    double ca[ORC_PMAX + 2], sa[ORC_PMAX + 2];
    double cb[ORC_PMAX + 2], sb[ORC_PMAX + 2];
} euler_t;

/* microkernel_generic.cpp:36-97.  cos(alpha)=y/rho, sin(alpha)=x/rho,
 * cos(beta)=z/r, sin(beta)=-rho/r; powers by repeated complex product. */
static void euler(double x, double y, double z, int k, euler_t *e)
{
    if (k >= 1) {
        e->r[0] = e->rinv[0] = e->ca[0] = e->cb[0] = 1.0;
        e->sa[0] = e->sb[0] = 0.0;
    }
    if (k > 1) {
        double rxy = hypot(x, y);
        double rxyz = hypot(rxy, z);
        e->r[1] = rxyz;
        e->rinv[1] = 1.0 / rxyz;
        e->ca[1] = (rxy == 0) ? 1.0 : y / rxy;
        e->sa[1] = (rxy == 0) ? 0.0 : x / rxy;
        e->cb[1] = z * e->rinv[1];
        e->sb[1] = -rxy * e->rinv[1];
    }
    for (int i = 2; i < k; ++i) {
        e->r[i] = e->r[1] * e->r[i - 1];
        e->rinv[i] = e->rinv[1] * e->rinv[i - 1];
        double c = e->ca[1], s = e->sa[1], cn = e->ca[i - 1], sn = e->sa[i - 1];
        e->ca[i] = cn * c - sn * s;
        e->sa[i] = sn * c + cn * s;
        c = e->cb[1]; s = e->sb[1]; cn = e->cb[i - 1]; sn = e->sb[i - 1];
        e->cb[i] = cn * c - sn * s;
        e->sb[i] = sn * c + cn * s;
    }
}

/* microkernel_generic.cpp:99-140: (re,im) <- scale * e^{+-i m theta} (re,im). */
static void rotscale(const double *c, const double *s, double scale, double *re, double *im, int k, int forward)
{
    for (int i = 0; i < k; ++i) {
        double a = re[i], b = im[i];
        if (forward) {
            re[i] = scale * (c[i] * a - s[i] * b);
            im[i] = scale * (s[i] * a + c[i] * b);
        } else {
            re[i] = scale * (c[i] * a + s[i] * b);
            im[i] = scale * (-s[i] * a + c[i] * b);
        }
    }
}

/* microkernel_generic.cpp:142-214: out = A * in with the folded matrix A. */
static void swap_apply(const double *A, int n, double *v)
{
    double out[ORC_PMAX];
    for (int m = 0; m <= n; ++m) {
        double acc = 0.0;
        for (int l = 0; l <= n; ++l) acc += A[m * (n + 1) + l] * v[l];
        out[m] = acc;
    }
    memcpy(v, out, sizeof(double) * (size_t)(n + 1));
}

static void swap_pair(int transposed, int n, double *re, double *im)
{
    swap_apply(fold_tab[transposed ? 2 : 0][n], n, re);
    swap_apply(fold_tab[transposed ? 3 : 1][n], n, im);
}

/* Triangles indexed [m][n], m <= n. */
#define TRI(m, n) ((size_t)(m) * ORC_PMAX + (size_t)(n))

static double factorial_tab[2 * ORC_PMAX];
static double invfac_tab[2 * ORC_PMAX];
static int fac_ready = 0;

static void fac_init(void)
{
    if (fac_ready) return;
    factorial_tab[0] = 1.0;
    for (int i = 1; i < 2 * ORC_PMAX; ++i) factorial_tab[i] = (double)i * factorial_tab[i - 1];
    for (int i = 0; i < 2 * ORC_PMAX; ++i) invfac_tab[i] = 1.0 / factorial_tab[i];
    fac_ready = 1;
}

enum { OP_M2M = 0, OP_M2L = 1, OP_L2L = 2 };

/* One translation; target += T(source).  translations.cpp:242-471. */
static void translate_one(int op, int Pi, int Po, double x, double y, double z, const double *src, double *dst,
                          double *tin_re, double *tin_im, double *tout_re, double *tout_im)
{
    if (Pi == 0 || Po == 0) return;

    if (op != OP_M2L) {
        /* Zero shift: plain copy-add over the common degrees,
         * translations.cpp:539-547 and 596-604. */
        double r2 = x * x + y * y + z * z;
        if (r2 == 0) {
            int P = Pi < Po ? Pi : Po;
            for (int i = 0; i < P * (P + 1); ++i) dst[i] += src[i];
            return;
        }
    }
    if (op == OP_M2M) { x = -x; y = -y; z = -z; } /* translations.cpp:331-333 */

    int Pend = Pi < Po ? Pi : Po;
    int Pmax = Pi > Po ? Pi : Po;
    int k;
    if (op == OP_L2L) k = (Pi > Pend + 1) ? Pi : Pend + 1; /* see header note */
    else k = Pmax + 1;

    euler_t e;
    euler(x, y, z, k, &e);

    double re[ORC_PMAX], im[ORC_PMAX];

    /* Forward transforms (translations.cpp:36-85, 189-239). */
    for (int n = 0; n < Pi; ++n) {
        for (int m = 0; m <= n; ++m) {
            re[m] = src[n * (n + 1) + m];
            im[m] = src[(n + 1) * (n + 1) + m];
        }
        int tr = (op == OP_L2L);
        double scale = (op == OP_L2L) ? e.r[n] : e.rinv[n + 1];
        rotscale(e.ca, e.sa, scale, re, im, n + 1, 1);
        swap_pair(tr, n, re, im);
        rotscale(e.cb, e.sb, 1.0, re, im, n + 1, 1);
        swap_pair(tr, n, re, im);
        for (int m = 0; m <= n; ++m) {
            tin_re[TRI(m, n)] = re[m];
            tin_im[TRI(m, n)] = im[m];
        }
    }

    /* z-translation (translations.cpp:283-296, 363-377, 440-451). */
    int Nend = (op == OP_L2L) ? Pend : Po;
    for (int m = 0; m < Pend; ++m)
        for (int N = m; N < Nend; ++N) {
            double ar = 0.0, ai = 0.0;
            if (op == OP_M2L) {
                for (int K = m; K < Pi; ++K) {
                    ar += factorial_tab[N + K] * tin_re[TRI(m, K)];
                    ai += factorial_tab[N + K] * tin_im[TRI(m, K)];
                }
                if ((N + m) & 1) { ar = -ar; ai = -ai; }
            } else if (op == OP_M2M) {
                int Kend = (N + 1 < Pi) ? N + 1 : Pi;
                for (int K = m; K < Kend; ++K) {
                    ar += invfac_tab[N - K] * tin_re[TRI(m, K)];
                    ai += invfac_tab[N - K] * tin_im[TRI(m, K)];
                }
            } else {
                for (int K = N; K < Pi; ++K) {
                    ar += invfac_tab[K - N] * tin_re[TRI(m, K)];
                    ai += invfac_tab[K - N] * tin_im[TRI(m, K)];
                }
            }
            tout_re[TRI(m, N)] = ar;
            tout_im[TRI(m, N)] = ai;
        }

    /* Backward transforms (translations.cpp:87-187) and accumulation
     * into the target (microkernel_generic.cpp:361-380). */
    for (int N = 0; N < Nend; ++N) {
        for (int m = 0; m <= N; ++m) {
            int have = (op == OP_L2L) ? (m < Pend) : (m < Pi);
            re[m] = have ? tout_re[TRI(m, N)] : 0.0;
            im[m] = have ? tout_im[TRI(m, N)] : 0.0;
        }
        int tr = (op != OP_M2M);
        double scale = (op == OP_M2M) ? e.r[N + 1] : e.rinv[N];
        swap_pair(tr, N, re, im);
        rotscale(e.cb, e.sb, 1.0, re, im, N + 1, 0);
        swap_pair(tr, N, re, im);
        rotscale(e.ca, e.sa, scale, re, im, N + 1, 0);
        for (int m = 0; m <= N; ++m) {
            dst[N * (N + 1) + m] += re[m];
            dst[(N + 1) * (N + 1) + m] += im[m];
        }
    }
}

/* ------------------------------------------------------------------------ */
/* Solid harmonics R and S (harmonics.cpp:31-128), Dehnen's convention.     */
/* out holds P(P+1) reals in the solid layout.                              */
/* ------------------------------------------------------------------------ */

#define RE(n, m) ((n) * ((n) + 1) + (m))
#define IM(n, m) (((n) + 1) * ((n) + 1) + (m))

static void harm_R(int P, double x, double y, double z, double *o)
{
    memset(o, 0, sizeof(double) * (size_t)P * (P + 1));
    if (P == 0) return;
    o[RE(0, 0)] = 1;
    if (P == 1) return;
    o[RE(1, 0)] = z;
    o[RE(1, 1)] = x / 2;
    o[IM(1, 1)] = y / 2;
    for (int m = 2; m < P; ++m) {
        double a = o[RE(m - 1, m - 1)], b = o[IM(m - 1, m - 1)], f = 1.0 / (2.0 * m);
        o[RE(m, m)] = (a * x - b * y) * f;
        o[IM(m, m)] = (a * y + b * x) * f;
    }
    for (int m = 1; m < P - 1; ++m) {
        o[RE(m + 1, m)] = z * o[RE(m, m)];
        o[IM(m + 1, m)] = z * o[IM(m, m)];
    }
    double r2 = x * x + y * y + z * z;
    for (int m = 0; m < P - 2; ++m)
        for (int n = m + 2; n < P; ++n) {
            double f = 1.0 / ((double)(n + m) * (double)(n - m));
            o[RE(n, m)] = (o[RE(n - 1, m)] * z * (2 * n - 1) - r2 * o[RE(n - 2, m)]) * f;
            o[IM(n, m)] = (o[IM(n - 1, m)] * z * (2 * n - 1) - r2 * o[IM(n - 2, m)]) * f;
        }
}

static void harm_S(int P, double x, double y, double z, double *o)
{
    memset(o, 0, sizeof(double) * (size_t)P * (P + 1));
    if (P == 0) return;
    double r2inv = 1.0 / (x * x + y * y + z * z);
    o[RE(0, 0)] = sqrt(r2inv);
    if (P == 1) return;
    o[RE(1, 0)] = z * r2inv * o[RE(0, 0)];
    o[RE(1, 1)] = x * r2inv * o[RE(0, 0)];
    o[IM(1, 1)] = y * r2inv * o[RE(0, 0)];
    for (int m = 2; m < P; ++m) {
        double a = o[RE(m - 1, m - 1)], b = o[IM(m - 1, m - 1)], f = (2.0 * m - 1) * r2inv;
        o[RE(m, m)] = (a * x - b * y) * f;
        o[IM(m, m)] = (a * y + b * x) * f;
    }
    for (int m = 1; m < P - 1; ++m) {
        double f = (2.0 * m + 1) * z * r2inv;
        o[RE(m + 1, m)] = f * o[RE(m, m)];
        o[IM(m + 1, m)] = f * o[IM(m, m)];
    }
    for (int m = 0; m < P - 2; ++m)
        for (int n = m + 2; n < P; ++n) {
            double c = (double)(n + m - 1) * (double)(n - m - 1);
            o[RE(n, m)] = (o[RE(n - 1, m)] * z * (2 * n - 1) - c * o[RE(n - 2, m)]) * r2inv;
            o[IM(n, m)] = (o[IM(n - 1, m)] * z * (2 * n - 1) - c * o[IM(n - 2, m)]) * r2inv;
        }
}

/* Complex coefficient C_n^m for any m, using C_n^{-m} = (-1)^m conj(C_n^m)
 * (doc/solidfmm.tex:1096-1106); zero for |m| > n or n < 0. */
static void coef(const double *c, int P, int n, int m, double *re, double *im)
{
    if (n < 0 || n >= P || m > n || -m > n) { *re = 0; *im = 0; return; }
    if (m >= 0) { *re = c[RE(n, m)]; *im = c[IM(n, m)]; return; }
    int am = -m;
    double s = (am & 1) ? -1.0 : 1.0;
    *re = s * c[RE(n, am)];
    *im = -s * c[IM(n, am)];
}

/* Naive O(P^4) translations, doc/solidfmm.tex:1404-1433 (the M2M sum over l
 * runs over -k..k; terms outside vanish). */
static int translate_naive_one(int op, int Pi, int Po, double x, double y, double z, const double *src, double *dst)
{
    if (Pi == 0 || Po == 0) return 0;
    int Ph = Pi + Po; /* harmonic degrees needed: up to Pi+Po-2 */
    double *h = (double *)malloc(sizeof(double) * (size_t)Ph * (Ph + 1));
    if (!h) return -1;
    if (op == OP_M2L) harm_S(Ph, x, y, z, h);
    else if (op == OP_M2M) harm_R(Ph, -x, -y, -z, h);
    else harm_R(Ph, x, y, z, h);

    for (int n = 0; n < Po; ++n)
        for (int m = 0; m <= n; ++m) {
            double ar = 0, ai = 0;
            for (int k = 0; k < Pi; ++k) {
                if (op == OP_L2L && k < n) continue;
                for (int l = -k; l <= k; ++l) {
                    double cr, ci, hr, hi;
                    coef(src, Pi, k, l, &cr, &ci);
                    if (op == OP_M2L) {
                        /* (-1)^n conj(M_k^l) S_{n+k}^{m+l}(r) */
                        coef(h, Ph, n + k, m + l, &hr, &hi);
                        double pr = cr * hr + ci * hi;
                        double pi = cr * hi - ci * hr;
                        if (n & 1) { pr = -pr; pi = -pi; }
                        ar += pr; ai += pi;
                    } else if (op == OP_M2M) {
                        /* M_k^l R_{n-k}^{m-l}(-r) */
                        coef(h, Ph, n - k, m - l, &hr, &hi);
                        ar += cr * hr - ci * hi;
                        ai += cr * hi + ci * hr;
                    } else {
                        /* L_k^l conj(R_{k-n}^{l-m}(r)) */
                        coef(h, Ph, k - n, l - m, &hr, &hi);
                        ar += cr * hr + ci * hi;
                        ai += ci * hr - cr * hi;
                    }
                }
            }
            dst[RE(n, m)] += ar;
            dst[IM(n, m)] += ai;
        }
    free(h);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Exported C ABI (ctypes).                                                 */
/* ------------------------------------------------------------------------ */

int orc_prepare(int P)
{
    if (P < 0 || P > ORC_PMAX - 1) return -1;
    fac_init();
    return fold_reserve(P > 0 ? P : 1);
}

int orc_swap_entry(int n, int m, int l, double *out)
{
    if (n < 0 || n >= ORC_PMAX - 1) return -1;
    if (swap_reserve(n + 1) != 0) return -1;
    *out = swap_get(n, m, l);
    return 0;
}

int orc_folded_matrix(int kind, int n, double *out)
{
    if (kind < 0 || kind > 3 || n < 0 || n >= ORC_PMAX - 1) return -1;
    if (orc_prepare(n + 1) != 0) return -1;
    memcpy(out, fold_tab[kind][n], sizeof(double) * (size_t)(n + 1) * (n + 1));
    return 0;
}

int orc_euler(double x, double y, double z, int k, double *out6k)
{
    if (k < 0 || k > ORC_PMAX + 1) return -1;
    euler_t e;
    euler(x, y, z, k, &e);
    for (int i = 0; i < k; ++i) {
        out6k[0 * k + i] = e.r[i];
        out6k[1 * k + i] = e.rinv[i];
        out6k[2 * k + i] = e.ca[i];
        out6k[3 * k + i] = e.sa[i];
        out6k[4 * k + i] = e.cb[i];
        out6k[5 * k + i] = e.sb[i];
    }
    return 0;
}

/* Items are processed in list order; repeated targets accumulate in that
 * order, as the reference's single-thread entry points do. */
int orc_translate(int op, const orc_item *items, size_t n, const double *src_pool, double *dst_pool)
{
    if (op < 0 || op > 2) return -1;
    int Pmax = 1;
    for (size_t i = 0; i < n; ++i) {
        if ((int)items[i].p_in > Pmax) Pmax = (int)items[i].p_in;
        if ((int)items[i].p_out > Pmax) Pmax = (int)items[i].p_out;
    }
    if (Pmax > ORC_PMAX - 2) return -2;
    if (orc_prepare(Pmax) != 0) return -3;
    size_t tri = (size_t)ORC_PMAX * ORC_PMAX;
    double *w = (double *)malloc(sizeof(double) * 4 * tri);
    if (!w) return -3;
    for (size_t i = 0; i < n; ++i) {
        const orc_item *it = &items[i];
        translate_one(op, (int)it->p_in, (int)it->p_out, it->x, it->y, it->z, src_pool + it->src,
                      dst_pool + it->dst, w, w + tri, w + 2 * tri, w + 3 * tri);
    }
    free(w);
    return 0;
}

int orc_translate_naive(int op, const orc_item *items, size_t n, const double *src_pool, double *dst_pool)
{
    if (op < 0 || op > 2) return -1;
    for (size_t i = 0; i < n; ++i) {
        const orc_item *it = &items[i];
        if (translate_naive_one(op, (int)it->p_in, (int)it->p_out, it->x, it->y, it->z, src_pool + it->src,
                                dst_pool + it->dst) != 0)
            return -3;
    }
    return 0;
}

/* kind 0: R, 1: S. */
/* Gradient of R: three blocks of P(P+1) reals (d/dx, d/dy, d/dz), by the
 * same recurrences as harm_R differentiated term by term (forward automatic
 * differentiation, harmonics.cpp:81-174). */
static void harm_dR(int P, double x, double y, double z, double *g)
{
    const size_t B = (size_t)P * (P + 1);
    memset(g, 0, sizeof(double) * 3 * B);
    if (P < 2) return;
    double *v = (double *)calloc(B, sizeof(double));
    if (!v) return;
#define G(d, k) g[(size_t)(d) * B + (k)]
    v[RE(0, 0)] = 1;
    v[RE(1, 0)] = z;     G(2, RE(1, 0)) = 1;
    v[RE(1, 1)] = x / 2; G(0, RE(1, 1)) = 0.5;
    v[IM(1, 1)] = y / 2; G(1, IM(1, 1)) = 0.5;
    for (int m = 2; m < P; ++m) {
        const double a = v[RE(m - 1, m - 1)], b = v[IM(m - 1, m - 1)], f = 1.0 / (2.0 * m);
        double ga[3], gb[3];
        for (int d = 0; d < 3; ++d) {
            ga[d] = G(d, RE(m - 1, m - 1));
            gb[d] = G(d, IM(m - 1, m - 1));
        }
        v[RE(m, m)] = (a * x - b * y) * f;
        G(0, RE(m, m)) = (ga[0] * x + a - gb[0] * y) * f;
        G(1, RE(m, m)) = (ga[1] * x - gb[1] * y - b) * f;
        G(2, RE(m, m)) = (ga[2] * x - gb[2] * y) * f;
        v[IM(m, m)] = (a * y + b * x) * f;
        G(0, IM(m, m)) = (ga[0] * y + gb[0] * x + b) * f;
        G(1, IM(m, m)) = (ga[1] * y + a + gb[1] * x) * f;
        G(2, IM(m, m)) = (ga[2] * y + gb[2] * x) * f;
    }
    for (int m = 1; m < P - 1; ++m) {
        for (int part = 0; part < 2; ++part) {
            const int src = part ? IM(m, m) : RE(m, m), dst = part ? IM(m + 1, m) : RE(m + 1, m);
            v[dst] = z * v[src];
            G(0, dst) = z * G(0, src);
            G(1, dst) = z * G(1, src);
            G(2, dst) = v[src] + z * G(2, src);
        }
    }
    const double r2 = x * x + y * y + z * z, r2d[3] = {2 * x, 2 * y, 2 * z};
    for (int m = 0; m < P - 2; ++m)
        for (int n = m + 2; n < P; ++n) {
            const double f = 1.0 / ((double)(n + m) * (double)(n - m));
            for (int part = 0; part < 2; ++part) {
                const int k = part ? IM(n, m) : RE(n, m), k1 = part ? IM(n - 1, m) : RE(n - 1, m),
                          k2 = part ? IM(n - 2, m) : RE(n - 2, m);
                const double a = v[k1], b = v[k2];
                v[k] = (a * z * (2 * n - 1) - r2 * b) * f;
                for (int d = 0; d < 3; ++d) {
                    double t = G(d, k1) * z * (2 * n - 1) - r2d[d] * b - r2 * G(d, k2);
                    if (d == 2) t += a * (2 * n - 1);
                    G(d, k) = t * f;
                }
            }
        }
#undef G
    free(v);
}

int orc_harmonic(int kind, int P, double x, double y, double z, double *out)
{
    if (P < 0 || P > 4 * ORC_PMAX) return -1;
    if (kind == 0) harm_R(P, x, y, z, out);
    else if (kind == 1) harm_S(P, x, y, z, out);
    else if (kind == 2) harm_dR(P, x, y, z, out);
    else return -1;
    return 0;
}

/* Evaluation pairing dot(A,B) = Re sum_{n,m=-n..n} conj(A_n^m) B_n^m over the
 * common degrees (solid.cpp:40-67). */
double orc_dot(const double *a, int PA, const double *b, int PB)
{
    int P = PA < PB ? PA : PB;
    double t = 0, t0 = 0;
    for (int k = 0; k < P * (P + 1); ++k) t += a[k] * b[k];
    for (int k = 0; k < P; ++k) t0 += a[k * (k + 1)] * b[k * (k + 1)];
    return 2 * t - t0;
}
