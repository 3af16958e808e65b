// symmetry.cpp -- the host side of the symmetry-reduced driver
// (symmetry.f90:16-246, moments.f90:15-240): equivalent-nuclei reduction of
// a set of spin-1/2 couplings, and the blocks (total-spin multiplets of the
// reduced sets) with their weights.  The propagation of each block runs on
// the device through spc_system_set_block and the trace-sampling entry
// points.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <numeric>
#include <vector>

#include "../../include/spinchem.h"

namespace {

// Eigenvalues (ascending) and the first components of the orthonormal
// eigenvectors of the symmetric tridiagonal matrix with diagonal d[0..n-1]
// and off-diagonal e[0..n-2]: implicit QL iterations with Wilkinson shifts
// (Golub and Van Loan, Matrix Computations, 4th ed., sec. 8.3), carrying only
// the first row of the eigenvector matrix, which is all the quadrature needs.
bool tridiag_eigen(std::vector<double> d, std::vector<double> e, std::vector<double> &lam, std::vector<double> &z0)
{
    const int n = (int)d.size();
    e.resize(n, 0.0);
    std::vector<double> z(n, 0.0); // first row of the accumulated rotations
    if (n > 0) z[0] = 1.0;
    for (int l = 0; l < n; ++l) {
        int iter = 0;
        for (;;) {
            int mm = l;
            for (; mm < n - 1; ++mm) {
                const double dd = std::fabs(d[mm]) + std::fabs(d[mm + 1]);
                if (std::fabs(e[mm]) <= std::numeric_limits<double>::epsilon() * dd) break;
            }
            if (mm == l) break;
            if (++iter > 60) return false;
            // Wilkinson shift from the leading 2x2 block at l
            double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
            double r = std::hypot(g, 1.0);
            g = d[mm] - d[l] + e[l] / (g + (g >= 0.0 ? r : -r));
            double s = 1.0, c = 1.0, p = 0.0;
            int i = mm - 1;
            for (; i >= l; --i) {
                double f = s * e[i];
                const double b = c * e[i];
                r = std::hypot(f, g);
                e[i + 1] = r;
                if (r == 0.0) { // deflation inside the sweep
                    d[i + 1] -= p;
                    e[mm] = 0.0;
                    break;
                }
                s = f / r;
                c = g / r;
                g = d[i + 1] - p;
                r = (d[i] - g) * s + 2.0 * c * b;
                p = s * r;
                d[i + 1] = g + p;
                g = c * r - b;
                // the rotation in the plane (i, i+1), first row only
                f = z[i + 1];
                z[i + 1] = s * z[i] + c * f;
                z[i] = c * z[i] - s * f;
            }
            if (r == 0.0 && i >= l) continue;
            d[l] -= p;
            e[l] = g;
            e[mm] = 0.0;
        }
    }
    std::vector<int> idx(n);
    std::iota(idx.begin(), idx.end(), 0);
    std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) { return d[a] < d[b]; });
    lam.resize(n);
    z0.resize(n);
    for (int i = 0; i < n; ++i) {
        lam[i] = d[idx[i]];
        z0[i] = z[idx[i]];
    }
    return true;
}

// quad_rule (moments.f90:95-150): the m-point Gauss rule of the discrete
// measure sum_i w_i delta(x - x_i), from the Stieltjes (Lanczos) recurrence
// on diag(x) started at sqrt(w): nodes = eigenvalues of the tridiagonal
// matrix, weights = ||sqrt(w)||^2 (first eigenvector component)^2.
bool quad_rule(const std::vector<double> &wi, const std::vector<double> &xi, int m, std::vector<double> &w,
               std::vector<double> &x)
{
    const int np = (int)xi.size();
    std::vector<std::vector<double>> P(m, std::vector<double>(np));
    std::vector<double> q(np), alpha(m), beta(m);
    for (int i = 0; i < np; ++i) q[i] = std::sqrt(wi[i]);
    for (int k = 0; k < m; ++k) {
        double q2 = 0.0;
        for (double v : q) q2 += v * v;
        q2 = std::sqrt(q2);
        if (q2 == 0.0) return false; // 'quad_rule 1'
        for (int i = 0; i < np; ++i) P[k][i] = q[i] / q2;
        for (int i = 0; i < np; ++i) q[i] = xi[i] * P[k][i];
        for (int j = k; j >= 0; --j) {
            double pq = 0.0;
            for (int i = 0; i < np; ++i) pq += P[j][i] * q[i];
            if (j == k) {
                beta[k] = q2;
                alpha[k] = pq;
            }
            for (int i = 0; i < np; ++i) q[i] -= pq * P[j][i];
        }
    }
    std::vector<double> off(beta.begin() + 1, beta.end());
    std::vector<double> z0;
    if (!tridiag_eigen(alpha, off, x, z0)) return false;
    w.resize(m);
    for (int j = 0; j < m; ++j) w[j] = (beta[0] * z0[j]) * (beta[0] * z0[j]);
    return true;
}

// Solves G q = f (m x m, row-major) by Gaussian elimination with partial
// pivoting; false if G is singular.
bool solve(std::vector<double> G, std::vector<double> f, int m, std::vector<double> &q)
{
    for (int c = 0; c < m; ++c) {
        int p = c;
        for (int r = c + 1; r < m; ++r)
            if (std::fabs(G[r * m + c]) > std::fabs(G[p * m + c])) p = r;
        if (G[p * m + c] == 0.0) return false;
        if (p != c) {
            for (int k = 0; k < m; ++k) std::swap(G[c * m + k], G[p * m + k]);
            std::swap(f[c], f[p]);
        }
        for (int r = c + 1; r < m; ++r) {
            const double t = G[r * m + c] / G[c * m + c];
            for (int k = c; k < m; ++k) G[r * m + k] -= t * G[c * m + k];
            f[r] -= t * f[c];
        }
    }
    q.assign(m, 0.0);
    for (int r = m - 1; r >= 0; --r) {
        double s = f[r];
        for (int k = r + 1; k < m; ++k) s -= G[r * m + k] * q[k];
        q[r] = s / G[r * m + r];
    }
    return true;
}

// newton (moments.f90:184-239): the m couplings a whose integer-weighted
// moments sum_j n_j a_j^k equal the quadrature's sum_j w_j aw_j^k, k = 1..m,
// from the start a; icode 0 on success (converged, or stopped when the step
// stops shrinking, keeping the previous iterate), 1 if a step exceeds |a|,
// 2 after 30 iterations.
int newton(const std::vector<double> &aw, const std::vector<double> &w, std::vector<double> &a,
           const std::vector<int> &n, int m)
{
    double dap = 0.0;
    for (int iter = 1; iter <= 30; ++iter) {
        std::vector<double> f(m), G((size_t)m * m), q;
        for (int k = 1; k <= m; ++k) {
            double fa = 0.0, fw = 0.0;
            for (int j = 0; j < m; ++j) {
                fa += n[j] * std::pow(a[j], k);
                fw += w[j] * std::pow(aw[j], k);
            }
            f[k - 1] = fa - fw;
            for (int j = 0; j < m; ++j) G[(size_t)(k - 1) * m + j] = k * n[j] * std::pow(a[j], k - 1);
        }
        if (!solve(G, f, m, q)) return 1; // singular (a set of size 0): stdlib's solve_lu stops there
        double a_abs = 0.0, d_abs = 0.0;
        for (int j = 0; j < m; ++j) {
            a_abs += std::fabs(a[j]);
            d_abs += std::fabs(q[j]);
        }
        const std::vector<double> prev = a;
        for (int j = 0; j < m; ++j) a[j] -= q[j];
        const double da = d_abs / a_abs;
        if (da > 1.0) return 1;
        if (1.0 + da == 1.0) return 0;
        if (iter > 1 && da >= dap) {
            a = prev;
            return 0;
        }
        dap = da;
    }
    return 2;
}

// C(n, r) in 64-bit integers, as nCr of symmetry.f90:228-246.
int64_t n_choose_r(int n, int r)
{
    int64_t num = 1, den = 1;
    for (int64_t i = 1; i <= std::min(n - r, r); ++i) {
        num *= (int64_t)n - i + 1;
        den *= i;
    }
    return num / den;
}

} // namespace

extern "C" {

int spc_shrink(int32_t n, const double *a_iso, int32_t *M, double *a_bar, int32_t *n_bar)
{
    if (!M || n < 0) return SPC_ERR_INVALID;
    if (n == 0) { // no nuclei on this electron: no sets (moments.f90:50-56)
        *M = 0;
        return SPC_OK;
    }
    const int m = *M;
    if (!a_iso || !a_bar || !n_bar || m < 1) return SPC_ERR_INVALID;
    if (m >= n || m > 62) return SPC_ERR_INVALID; // 'Error: Symetry reduction requires M<N'
    const std::vector<double> a(a_iso, a_iso + n), ones(n, 1.0);
    std::vector<double> wb, awb;
    if (!quad_rule(ones, a, m, wb, awb)) return SPC_ERR_BREAKDOWN;
    // integer weights: the floors plus one on ndiff of the sets, the choice
    // whose first m+1 power sums come closest (moments.f90:67-90)
    std::vector<int> nfloor(m);
    int nftot = 0;
    for (int i = 0; i < m; ++i) {
        nfloor[i] = (int)wb[i];
        nftot += nfloor[i];
    }
    const int ndiff = n - nftot;
    double dmom = std::numeric_limits<double>::max();
    bool found = false;
    for (uint64_t ib = 0; ib < ((uint64_t)1 << m); ++ib) {
        if (__builtin_popcountll(ib) != ndiff) continue;
        std::vector<int> nt(m);
        for (int i = 0; i < m; ++i) nt[i] = nfloor[i] + (int)((ib >> i) & 1);
        std::vector<double> at = awb;
        if (newton(awb, wb, at, nt, m) != 0) continue;
        double emom = 0.0;
        for (int k = 1; k <= m + 1; ++k) {
            double exact = 0.0, approx = 0.0;
            for (double v : a) exact += std::pow(v, k);
            // as moments.f90:81: the power sums of the m couplings themselves
            for (double v : at) approx += std::pow(v, k);
            emom += std::fabs(approx - exact);
        }
        if (emom < dmom) {
            dmom = emom;
            found = true;
            for (int i = 0; i < m; ++i) {
                n_bar[i] = nt[i];
                a_bar[i] = at[i];
            }
        }
    }
    return found ? SPC_OK : SPC_ERR_BREAKDOWN;
}

int spc_symmetry_blocks(int32_t n_sets, const int32_t *n_bar, int64_t capacity, int64_t *n_blocks, int32_t *mult,
                        int64_t *Z_block, int64_t *w_block)
{
    if (n_sets < 0 || !n_blocks || (n_sets > 0 && !n_bar)) return SPC_ERR_INVALID;
    // cartesian_product (symmetry.f90:187-225): set i takes the multiplicities
    // 1 + n_i mod 2, 3 + n_i mod 2, .., n_i + 1 (total spins of n_i spin-1/2),
    // the first set varying fastest
    int64_t count = 1;
    for (int i = 0; i < n_sets; ++i) {
        if (n_bar[i] < 1) return SPC_ERR_INVALID;
        count *= n_bar[i] / 2 + 1;
        if (count > ((int64_t)1 << 40)) return SPC_ERR_UNSUPPORTED;
    }
    *n_blocks = count;
    if (!mult) return SPC_OK; // size query
    if (capacity < count || !Z_block || !w_block) return SPC_ERR_INVALID;
    std::vector<int32_t> K((size_t)count * std::max(n_sets, 1));
    for (int64_t c = 0; c < count; ++c) {
        int64_t r = c;
        for (int i = 0; i < n_sets; ++i) {
            const int opts = n_bar[i] / 2 + 1;
            const int j = (int)(r % opts);
            r /= opts;
            K[(size_t)c * n_sets + i] = 2 * j + 1 + n_bar[i] % 2;
        }
    }
    // sort_col_prod (utils.f90:25-50): columns in ascending order of their
    // products, stably
    std::vector<int64_t> prod((size_t)count);
    for (int64_t c = 0; c < count; ++c) {
        int64_t p = 1;
        for (int i = 0; i < n_sets; ++i) p *= K[(size_t)c * n_sets + i];
        prod[c] = p;
    }
    std::vector<int64_t> idx((size_t)count);
    std::iota(idx.begin(), idx.end(), 0);
    std::stable_sort(idx.begin(), idx.end(), [&](int64_t a, int64_t b) { return prod[a] < prod[b]; });
    for (int64_t c = 0; c < count; ++c) {
        const int64_t src = idx[c];
        int64_t w = 1;
        for (int i = 0; i < n_sets; ++i) {
            const int g = K[(size_t)src * n_sets + i], ni = n_bar[i];
            mult[(size_t)c * n_sets + i] = g;
            // weight (symmetry.f90:211-226): copies of the spin-(g-1)/2 multiplet
            // among n_i spin-1/2, C(n, (n+g-1)/2) 2g / (n+g+1)
            const int64_t wi = n_choose_r(ni, (ni + g - 1) / 2) * (int64_t)(2 * g) / (int64_t)(ni + g + 1);
            w *= wi;
        }
        Z_block[c] = prod[src];
        w_block[c] = w;
    }
    return SPC_OK;
}

} // extern "C"
