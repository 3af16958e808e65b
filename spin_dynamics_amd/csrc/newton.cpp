// newton.cpp -- host side of the Newton-basis Krylov block: shifts for the
// chain y_{j+1} = (H - sig_j) y_j from the spectrum of H's Hermitian part.
//
// Newton bases with Leja-ordered Ritz values keep the block of chain vectors
// far better conditioned than the monomial chain [q, Hq, H^2 q, ..]
// (Bai, Hu and Reichel, IMA J. Numer. Anal. 14 (1994) 563; Hoemmen, PhD
// thesis, UC Berkeley 2010, sec. 7.3).  The spectrum of H does not change
// between steps, so the shifts are computed once per Hamiltonian: a short
// Lanczos run (Paige's form, no reorthogonalisation: extreme Ritz values
// converge first and spurious copies only repeat a point) on
// A = (H + H^H)/2, the tridiagonal's eigenvalues by Sturm-sequence bisection,
// then Leja ordering (each point maximises the product of its distances to
// the points before it; the largest in magnitude first).
#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdint>
#include <vector>

#include "internal.hpp"

namespace spc {

namespace {

using cplx = std::complex<double>;

// Eigenvalues of the symmetric tridiagonal matrix with diagonal a and
// off-diagonal b (b[i] couples i and i + 1), ascending, by bisection on the
// Sturm count (number of eigenvalues below x).
std::vector<double> tridiagonal_eigenvalues(const std::vector<double> &a, const std::vector<double> &b)
{
    const int n = (int)a.size();
    std::vector<double> b2(n > 0 ? n - 1 : 0);
    double lo = 0.0, hi = 0.0;
    for (int i = 0; i < n; ++i) {
        const double r = (i > 0 ? std::fabs(b[i - 1]) : 0.0) + (i + 1 < n ? std::fabs(b[i]) : 0.0);
        lo = i == 0 ? a[i] - r : std::min(lo, a[i] - r);
        hi = i == 0 ? a[i] + r : std::max(hi, a[i] + r);
        if (i + 1 < n) b2[i] = b[i] * b[i];
    }
    const double scale = std::max(std::fabs(lo), std::fabs(hi));
    const double tiny = std::max(scale, 1e-300) * 1e-300;
    auto below = [&](double x) {
        int cnt = 0;
        double d = 1.0;
        for (int i = 0; i < n; ++i) {
            d = (a[i] - x) - (i > 0 ? b2[i - 1] / d : 0.0);
            if (d == 0.0) d = -tiny;
            if (d < 0.0) ++cnt;
        }
        return cnt;
    };
    std::vector<double> ev(n);
    for (int i = 0; i < n; ++i) {
        double l = lo, h = hi;
        for (int it = 0; it < 200 && h - l > 1e-15 * std::max(scale, 1e-300); ++it) {
            const double mid = 0.5 * (l + h);
            if (below(mid) > i)
                h = mid;
            else
                l = mid;
        }
        ev[i] = 0.5 * (l + h);
    }
    return ev;
}

} // namespace

int newton_shifts_csr(int64_t N, const int32_t *rowptr, const int32_t *col, const double2 *val, int count,
                      double *out)
{
    if (N < 1 || count < 1 || !rowptr || !col || !val || !out) return -1;
    const int steps = (int)std::min<int64_t>(N, 40);
    std::vector<cplx> q(N), qp(N, 0.0), w(N);
    // fixed start vector: unit-modulus entries with scattered phases
    uint64_t st = 0x9E3779B97F4A7C15ull;
    const double inv = 1.0 / std::sqrt((double)N);
    for (int64_t i = 0; i < N; ++i) {
        st = st * 6364136223846793005ull + 1442695040888963407ull;
        const double ph = 6.283185307179586 * (double)(st >> 11) * 0x1.0p-53;
        q[i] = cplx(std::cos(ph), std::sin(ph)) * inv;
    }
    std::vector<double> alpha, beta;
    double bprev = 0.0, anorm = 0.0;
    for (int j = 0; j < steps; ++j) {
        // w = A q = (H q + H^H q) / 2 in one pass over the rows
        std::fill(w.begin(), w.end(), cplx(0.0, 0.0));
        for (int64_t i = 0; i < N; ++i) {
            cplx acc = 0.0;
            const cplx qi = q[i];
            for (int32_t k = rowptr[i]; k < rowptr[i + 1]; ++k) {
                const cplx v(val[k].x, val[k].y);
                acc += v * q[col[k]];
                w[col[k]] += std::conj(v) * qi;
            }
            w[i] += acc;
        }
        double a = 0.0;
        for (int64_t i = 0; i < N; ++i) {
            w[i] *= 0.5;
            a += std::real(std::conj(q[i]) * w[i]);
        }
        double bn = 0.0;
        for (int64_t i = 0; i < N; ++i) {
            w[i] -= a * q[i] + bprev * qp[i];
            bn += std::norm(w[i]);
        }
        bn = std::sqrt(bn);
        alpha.push_back(a);
        anorm = std::max(anorm, std::fabs(a) + bn + bprev);
        if (j + 1 == steps || !(bn > 1e-12 * anorm)) break; // invariant subspace: the Ritz values are exact
        beta.push_back(bn);
        for (int64_t i = 0; i < N; ++i) {
            qp[i] = q[i];
            q[i] = w[i] / bn;
        }
        bprev = bn;
    }
    const std::vector<double> ritz = tridiagonal_eigenvalues(alpha, beta);
    if (ritz.empty()) return -1;
    // Leja order over the Ritz values, log distances (no overflow)
    std::vector<double> leja;
    std::vector<char> used(ritz.size(), 0);
    std::vector<double> logd(ritz.size(), 0.0);
    for (size_t t = 0; t < ritz.size(); ++t) {
        size_t best = ritz.size();
        for (size_t i = 0; i < ritz.size(); ++i) {
            if (used[i]) continue;
            const double key = t == 0 ? std::fabs(ritz[i]) : logd[i];
            const double bkey = best == ritz.size() ? -INFINITY : (t == 0 ? std::fabs(ritz[best]) : logd[best]);
            if (best == ritz.size() || key > bkey) best = i;
        }
        if (best == ritz.size() || (t > 0 && !(logd[best] > -INFINITY))) break; // only repeats of chosen points left
        used[best] = 1;
        leja.push_back(ritz[best]);
        for (size_t i = 0; i < ritz.size(); ++i)
            if (!used[i]) logd[i] += std::log(std::fabs(ritz[i] - ritz[best]));
    }
    for (int j = 0; j < count; ++j) out[j] = leja[(size_t)j % leja.size()];
    return 0;
}

} // namespace spc
