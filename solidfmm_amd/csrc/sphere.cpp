// Smallest enclosing balls of points and of balls (host code), for the
// multipole acceptance criterion of a tree: a cell's expansion centre and
// radius are the smallest ball around its points.  The reference offers the
// same two functions (sphere.hpp:38-42, sphere.cpp:546-627); this file also
// implements them for the drop-in header include/solidfmm/sphere.hpp.
//
// Method.  Both problems are LP-type with combinatorial dimension 4 in three
// dimensions: the smallest ball is fixed by a support set of at most four
// of the inputs, which touch its boundary from inside (Fischer and Gaertner,
// "The smallest enclosing ball of balls", 2004).  We run the pivoting scheme
// the reference also uses: keep a support set S and its smallest ball;
// find the input that sticks out farthest; if it sticks out, the new
// support set is a subset of S plus that input.  The new basis is found
// here by brute force over the at most 15 subsets that contain the new
// input: for each, the ball tangent from inside to all its members with the
// centre in their affine hull (a 3x3 linear solve and one quadratic); the
// smallest such ball that encloses all of S plus the new input is the
// smallest ball of that set.  The radius grows with every pivot, so the
// loop ends; a pivot that does not grow it ends the loop too (round-off).
#include <algorithm>
#include <array>
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <thread>
#include <vector>

#include <solidfmm/sphere.hpp>

#include "../../include/solidfmm_gpu.h"

namespace sfmm {
namespace {

struct Ball {
    double x = 0, y = 0, z = 0, r = 0;
};

inline double dist(const Ball &a, double x, double y, double z)
{
    const double dx = x - a.x, dy = y - a.y, dz = z - a.z;
    return std::sqrt(dx * dx + dy * dy + dz * dz);
}

// how far input b reaches beyond the centre of B
inline double reach(const Ball &B, const Ball &b) { return dist(B, b.x, b.y, b.z) + b.r; }

// Solve the k x k system G l = rhs (k <= 3) by Gaussian elimination with
// partial pivoting; false when G is (numerically) singular.
bool solve(int k, double G[3][3], double rhs[3], double out[3])
{
    double A[3][4];
    double scale = 0;
    for (int i = 0; i < k; ++i) {
        for (int j = 0; j < k; ++j) {
            A[i][j] = G[i][j];
            scale = std::max(scale, std::fabs(G[i][j]));
        }
        A[i][k] = rhs[i];
    }
    if (scale == 0) return false;
    for (int c = 0; c < k; ++c) {
        int p = c;
        for (int i = c + 1; i < k; ++i)
            if (std::fabs(A[i][c]) > std::fabs(A[p][c])) p = i;
        if (std::fabs(A[p][c]) <= 1e-13 * scale) return false;
        if (p != c)
            for (int j = 0; j <= k; ++j) std::swap(A[p][j], A[c][j]);
        for (int i = c + 1; i < k; ++i) {
            const double f = A[i][c] / A[c][c];
            for (int j = c; j <= k; ++j) A[i][j] -= f * A[c][j];
        }
    }
    for (int i = k - 1; i >= 0; --i) {
        double s = A[i][k];
        for (int j = i + 1; j < k; ++j) s -= A[i][j] * out[j];
        out[i] = s / A[i][i];
    }
    return true;
}

// Ball tangent from inside to all of u[0..k) (k <= 4), centre in the affine
// hull of their centres.  With d_i = c_i - c_0 and c = c_0 + sum l_j d_j:
// |c - c_i| = R - r_i for all i gives, after subtracting the i = 0 case,
// sum_j 2 d_i.d_j l_j = |d_i|^2 - r_i^2 + r_0^2 + 2 R (r_i - r_0), linear in
// (l, R); the i = 0 case |sum l_j d_j| = R - r_0 then fixes R.
int tangent_balls(const Ball *u, int k, Ball out[2])
{
    if (k == 1) {
        out[0] = u[0];
        return 1;
    }
    const int m = k - 1;
    double d[3][3];
    for (int i = 0; i < m; ++i) {
        d[i][0] = u[i + 1].x - u[0].x;
        d[i][1] = u[i + 1].y - u[0].y;
        d[i][2] = u[i + 1].z - u[0].z;
    }
    double G[3][3], a[3], b[3];
    for (int i = 0; i < m; ++i) {
        for (int j = 0; j < m; ++j) G[i][j] = 2 * (d[i][0] * d[j][0] + d[i][1] * d[j][1] + d[i][2] * d[j][2]);
        const double dd = d[i][0] * d[i][0] + d[i][1] * d[i][1] + d[i][2] * d[i][2];
        a[i] = dd - u[i + 1].r * u[i + 1].r + u[0].r * u[0].r;
        b[i] = 2 * (u[i + 1].r - u[0].r);
    }
    double la[3] = {0, 0, 0}, lb[3] = {0, 0, 0};
    if (!solve(m, G, a, la) || !solve(m, G, b, lb)) return 0;
    double uu[3] = {0, 0, 0}, vv[3] = {0, 0, 0};
    for (int j = 0; j < m; ++j)
        for (int c = 0; c < 3; ++c) {
            uu[c] += la[j] * d[j][c];
            vv[c] += lb[j] * d[j][c];
        }
    const double r0 = u[0].r;
    const double A = vv[0] * vv[0] + vv[1] * vv[1] + vv[2] * vv[2] - 1.0;
    const double B = 2 * (uu[0] * vv[0] + uu[1] * vv[1] + uu[2] * vv[2] + r0);
    const double C = uu[0] * uu[0] + uu[1] * uu[1] + uu[2] * uu[2] - r0 * r0;
    double roots[2];
    int nr = 0;
    if (std::fabs(A) < 1e-14) {
        if (B == 0) return 0;
        roots[nr++] = -C / B;
    } else {
        const double disc = B * B - 4 * A * C;
        if (disc < 0) return 0;
        const double sq = std::sqrt(disc);
        // stable pair of roots
        const double q = -0.5 * (B + (B >= 0 ? sq : -sq));
        if (q != 0) roots[nr++] = C / q;
        roots[nr++] = q / A;
    }
    double rmax = 0;
    for (int i = 0; i < k; ++i) rmax = std::max(rmax, u[i].r);
    int found = 0;
    for (int t = 0; t < nr; ++t) {
        const double R = roots[t];
        if (!(R >= rmax * (1 - 1e-12)) || !std::isfinite(R)) continue;
        out[found++] = Ball{u[0].x + uu[0] + R * vv[0], u[0].y + uu[1] + R * vv[1], u[0].z + uu[2] + R * vv[2], R};
    }
    return found;
}

// Smallest ball of `s` (size <= 4) plus `p` with p on its boundary.
bool new_basis(const std::vector<Ball> &s, const Ball &p, std::vector<Ball> &basis, Ball &ball)
{
    const int ns = int(s.size());
    // enclosure tolerance: relative to the radius, plus the rounding of
    // distances between coordinates of this magnitude
    double coord = std::fabs(p.x) + std::fabs(p.y) + std::fabs(p.z);
    for (const Ball &b : s) coord = std::max(coord, std::fabs(b.x) + std::fabs(b.y) + std::fabs(b.z));
    bool found = false;
    for (int mask = 0; mask < (1 << ns); ++mask) {
        Ball u[4];
        int k = 0;
        u[k++] = p;
        for (int i = 0; i < ns; ++i)
            if (mask & (1 << i)) {
                if (k == 4) {
                    k = 5;
                    break;
                }
                u[k++] = s[i];
            }
        if (k > 4) continue;
        Ball cand[2];
        const int nc = tangent_balls(u, k, cand);
        for (int t = 0; t < nc; ++t) {
            const Ball &c = cand[t];
            if (found && c.r >= ball.r) continue;
            const double tol = 1e-12 * c.r + 1e-14 * coord;
            bool encloses = reach(c, p) <= c.r + tol;
            for (int i = 0; i < ns && encloses; ++i) encloses = reach(c, s[i]) <= c.r + tol;
            if (!encloses) continue;
            ball = c;
            basis.assign(u, u + k);
            found = true;
        }
    }
    return found;
}

template <class Get>
Ball smallest_ball(size_t n, Get get)
{
    std::vector<Ball> basis{get(0)};
    Ball B = basis[0];
    for (int pivots = 0; pivots < 1000; ++pivots) {
        double far = reach(B, get(0));
        size_t fi = 0;
        for (size_t i = 1; i < n; ++i) {
            const double d = reach(B, get(i));
            if (d > far) {
                far = d;
                fi = i;
            }
        }
        const Ball p = get(fi);
        // distances between coordinates of magnitude |c| carry ~ulp(|c|) of
        // rounding; closer than that counts as on the ball
        const double coord = std::max({std::fabs(B.x), std::fabs(B.y), std::fabs(B.z)});
        if (far <= B.r * (1 + 1e-15) + 4e-16 * coord) break;
        std::vector<Ball> nb;
        Ball nball;
        if (!new_basis(basis, p, nb, nball) || !(nball.r > B.r)) break;
        basis.swap(nb);
        B = nball;
    }
    // the radius covers every input exactly (up to the last ulp)
    double R = B.r;
    for (size_t i = 0; i < n; ++i) R = std::max(R, reach(B, get(i)));
    B.r = R;
    return B;
}

Ball ball_of_points(size_t n, const double *x, const double *y, const double *z, size_t stride = 1)
{
    return smallest_ball(n, [&](size_t i) { return Ball{x[i * stride], y[i * stride], z[i * stride], 0.0}; });
}

}  // namespace
}  // namespace sfmm

extern "C" {

int sfmm_bounding_sphere(size_t n, const double *x, const double *y, const double *z, double *out)
{
    if (n == 0 || !x || !y || !z || !out) return SFMM_ERR_LOGIC;
    const sfmm::Ball b = sfmm::ball_of_points(n, x, y, z);
    out[0] = b.x;
    out[1] = b.y;
    out[2] = b.z;
    out[3] = b.r;
    return SFMM_OK;
}

int sfmm_bounding_sphere_of_spheres(size_t n, const double *spheres, double *out)
{
    if (n == 0 || !spheres || !out) return SFMM_ERR_LOGIC;
    for (size_t i = 0; i < n; ++i)
        if (!(spheres[4 * i + 3] >= 0)) return SFMM_ERR_LOGIC;
    const sfmm::Ball b = sfmm::smallest_ball(n, [&](size_t i) {
        return sfmm::Ball{spheres[4 * i], spheres[4 * i + 1], spheres[4 * i + 2], spheres[4 * i + 3]};
    });
    out[0] = b.x;
    out[1] = b.y;
    out[2] = b.z;
    out[3] = b.r;
    return SFMM_OK;
}

int sfmm_bounding_spheres_of_ranges(const double *points, size_t n_points, size_t stride, const uint64_t *first,
                                    const uint64_t *count, size_t n_sets, double *out, int threads)
{
    if (n_sets == 0) return SFMM_OK;
    if (!points || !first || !count || !out || stride < 3) return SFMM_ERR_LOGIC;
    for (size_t s = 0; s < n_sets; ++s)
        if (count[s] == 0 || first[s] > n_points || count[s] > n_points - first[s]) return SFMM_ERR_LOGIC;
    unsigned nt = threads > 0 ? unsigned(threads) : std::max(1u, std::thread::hardware_concurrency());
    nt = std::min<unsigned>(nt, 64);
    // largest sets first, handed out round-robin, so threads end together
    std::vector<size_t> order(n_sets);
    for (size_t s = 0; s < n_sets; ++s) order[s] = s;
    std::sort(order.begin(), order.end(), [&](size_t a, size_t b) { return count[a] > count[b]; });
    auto work = [&](unsigned t) {
        for (size_t k = t; k < n_sets; k += nt) {
            const size_t s = order[k];
            const double *p = points + first[s] * stride;
            const sfmm::Ball b = sfmm::ball_of_points(count[s], p, p + 1, p + 2, stride);
            out[4 * s] = b.x;
            out[4 * s + 1] = b.y;
            out[4 * s + 2] = b.z;
            out[4 * s + 3] = b.r;
        }
    };
    std::vector<std::thread> pool;
    for (unsigned t = 1; t < nt; ++t) pool.emplace_back(work, t);
    work(0);
    for (auto &th : pool) th.join();
    return SFMM_OK;
}

}  // extern "C"

// ---- drop-in: include/solidfmm/sphere.hpp --------------------------------
namespace solidfmm {

template struct sphere<float>;
template struct sphere<double>;

namespace {

// float results: centre rounded, radius recomputed from the rounded centre
// and rounded up, so the float ball still encloses every input
template <class Get>
sphere<float> to_float(const sfmm::Ball &b, size_t n, Get get)
{
    sphere<float> s;
    s.x = float(b.x);
    s.y = float(b.y);
    s.z = float(b.z);
    double R = 0;
    for (size_t i = 0; i < n; ++i) {
        const sfmm::Ball q = get(i);
        R = std::max(R, sfmm::dist(sfmm::Ball{s.x, s.y, s.z, 0}, q.x, q.y, q.z) + q.r);
    }
    float r = float(R);
    if (double(r) < R) r = std::nextafter(r, INFINITY);
    s.r = r;
    return s;
}

void need_some(size_t n)
{
    if (n == 0)
        throw std::logic_error{"solidfmm::bounding_sphere: Cannot compute bounding sphere of an empty set."};
}

}  // namespace

sphere<double> bounding_sphere(size_t n, const double *x, const double *y, const double *z)
{
    need_some(n);
    const sfmm::Ball b = sfmm::ball_of_points(n, x, y, z);
    return sphere<double>{b.x, b.y, b.z, b.r};
}

sphere<float> bounding_sphere(size_t n, const float *x, const float *y, const float *z)
{
    need_some(n);
    auto get = [&](size_t i) { return sfmm::Ball{x[i], y[i], z[i], 0.0}; };
    return to_float(sfmm::smallest_ball(n, get), n, get);
}

sphere<double> bounding_sphere(const sphere<double> *begin, const sphere<double> *end)
{
    const size_t n = size_t(end - begin);
    need_some(n);
    const sfmm::Ball b =
        sfmm::smallest_ball(n, [&](size_t i) { return sfmm::Ball{begin[i].x, begin[i].y, begin[i].z, begin[i].r}; });
    return sphere<double>{b.x, b.y, b.z, b.r};
}

sphere<float> bounding_sphere(const sphere<float> *begin, const sphere<float> *end)
{
    const size_t n = size_t(end - begin);
    need_some(n);
    auto get = [&](size_t i) { return sfmm::Ball{begin[i].x, begin[i].y, begin[i].z, begin[i].r}; };
    return to_float(sfmm::smallest_ball(n, get), n, get);
}

}  // namespace solidfmm
