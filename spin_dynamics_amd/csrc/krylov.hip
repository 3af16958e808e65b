// krylov.hip -- restarted Krylov propagation of M trace vectors at once.
//
// Replaces get_krylov / expm / convert_to_krylov and the inner step loop of
// qm_dynamics.f90:340-366.  All vectors share H, so the M trace vectors are the
// lane dimension: every basis vector is an [N][Mp] block, lane r = trace vector
// r, one 1 KiB row segment per wavefront.  The sparse product gathers whole
// rows (one scalar-loaded H entry, 64 lanes of x), and the Gram-Schmidt dot
// products need no cross-lane reduction: each lane sums its own vector over the
// rows it visits, waves of a workgroup combine through LDS, workgroups through
// a deterministic second pass.
//
// Orthogonalisation: the reference uses modified Gram-Schmidt (i dependent
// reductions per column).  Here each column costs two streaming passes and one
// reduction, with classical Gram-Schmidt plus a delayed re-orthogonalisation
// (the "DCGS2" form of Swirydowicz et al., 2020):
//   pass 1 (column i): v = H u_i  and  a = Q^H u_i, b = Q^H v, c = u_i^H v,
//                      d = u_i^H u_i
//   coefficients:      beta_i^2 = d - |a|^2, column i-1 corrected by a,
//                      w_i = H q_i = (v - Q g - g_i q_i)/beta_i via the
//                      Arnoldi relation (g = Hbar a), t_i = Q^H w_i from a,b,c
//   pass 2:            q_i = (u_i - Q a)/beta_i,  u_{i+1} = w_i - Q t_i
// A last dots pass corrects column m-1.  The result is the reference's
// Hessenberg matrix A (qm_dynamics.f90:432-467) to rounding.
#include <algorithm>
#include <utility>

#include "common.hpp"
#include "internal.hpp"

namespace spc {

namespace {

constexpr int kBlock = 256;
constexpr int kWaves = kBlock / kWave;

// --------------------------------------------------------------- pass 1
// NQ = i previous basis vectors.  SPMM: v = H u plus dots (a, b, c, d);
// otherwise dots a = Q^H u only (final correction of the last column).
template <int NQ, bool SPMM>
__global__ __launch_bounds__(kBlock) void k_pass1(const int32_t *__restrict__ rowptr,
                                                   const int32_t *__restrict__ colidx,
                                                   const double2 *__restrict__ hval,
                                                   const double2 *__restrict__ Q,
                                                   const double2 *__restrict__ u, double2 *__restrict__ v,
                                                   double2 *__restrict__ part, int64_t N, int Mp, int rpw)
{
    constexpr int NACC = SPMM ? 2 * NQ + 2 : NQ;
    constexpr int NACC1 = NACC > 0 ? NACC : 1;
    const int nch = Mp / kWave;
    const int lid = xcd_remap(blockIdx.x, gridDim.x);
    const int ch = lid % nch, rb = lid / nch;
    const int lane = threadIdx.x & (kWave - 1);
    const int wave = wave_uniform(threadIdx.x / kWave);
    const int r = ch * kWave + lane;
    const int64_t NMp = N * (int64_t)Mp;

    double2 acc[NACC1];
#pragma unroll
    for (int j = 0; j < NACC1; ++j) acc[j] = cz(0, 0);

    const int64_t row0 = ((int64_t)rb * kWaves + wave) * rpw;
    const int64_t row1 = row0 + rpw < N ? row0 + rpw : N;
    for (int64_t row = row0; row < row1; ++row) {
        const int64_t off = row * Mp + r;
        const double2 uu = u[off];
        double2 vv = cz(0, 0);
        if constexpr (SPMM) {
            const int e0 = rowptr[row], e1 = rowptr[row + 1];
            int e = e0;
            for (; e + 4 <= e1; e += 4) {
                const int c0 = colidx[e], c1 = colidx[e + 1], c2 = colidx[e + 2], c3 = colidx[e + 3];
                const double2 x0 = u[(int64_t)c0 * Mp + r], x1 = u[(int64_t)c1 * Mp + r];
                const double2 x2 = u[(int64_t)c2 * Mp + r], x3 = u[(int64_t)c3 * Mp + r];
                vv = cfma(hval[e], x0, vv);
                vv = cfma(hval[e + 1], x1, vv);
                vv = cfma(hval[e + 2], x2, vv);
                vv = cfma(hval[e + 3], x3, vv);
            }
            for (; e < e1; ++e) vv = cfma(hval[e], u[(int64_t)colidx[e] * Mp + r], vv);
            v[off] = vv;
        }
#pragma unroll
        for (int k = 0; k < NQ; ++k) {
            const double2 q = Q[k * NMp + off];
            acc[k] = cfma_conj(q, uu, acc[k]);
            if constexpr (SPMM) acc[NQ + k] = cfma_conj(q, vv, acc[NQ + k]);
        }
        if constexpr (SPMM) {
            acc[2 * NQ] = cfma_conj(uu, vv, acc[2 * NQ]);
            acc[2 * NQ + 1].x += cabs2(uu);
        }
    }

    if constexpr (NACC > 0) {
        __shared__ double2 red[kWaves - 1][8][kWave];
#pragma unroll
        for (int base = 0; base < NACC; base += 8) {
            if (wave > 0) {
#pragma unroll
                for (int j = 0; j < 8; ++j)
                    if (base + j < NACC) red[wave - 1][j][lane] = acc[base + j];
            }
            __syncthreads();
            if (wave == 0) {
#pragma unroll
                for (int j = 0; j < 8; ++j)
                    if (base + j < NACC) {
                        double2 s = acc[base + j];
#pragma unroll
                        for (int w = 0; w < kWaves - 1; ++w) s = cadd(s, red[w][j][lane]);
                        part[((int64_t)rb * NACC + base + j) * Mp + r] = s;
                    }
            }
            __syncthreads();
        }
    }
}

// Deterministic reduction of the workgroup partials: dots[j][r] = sum_rb part[rb][j][r].
__global__ void k_reduce(const double2 *__restrict__ part, double2 *__restrict__ dots, int nacc, int nrb, int Mp)
{
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (int64_t)nacc * Mp) return;
    const int j = (int)(idx / Mp), r = (int)(idx % Mp);
    double2 s = cz(0, 0);
    for (int rb = 0; rb < nrb; ++rb) s = cadd(s, part[((int64_t)rb * nacc + j) * Mp + r]);
    dots[idx] = s;
}

// ------------------------------------------------------------ coefficients
// One thread per lane; i = index of the basis vector being finalised.
__global__ void k_coeffs(int i, int m, int M, int Mp, const double2 *__restrict__ dots, double2 *__restrict__ coef,
                         double2 *__restrict__ tprev, double2 *__restrict__ hess, double *__restrict__ beta0,
                         int32_t *__restrict__ flags)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= Mp) return;
    const int ncoef = 2 * i + 3;
    if (r >= M) {
        for (int j = 0; j < ncoef; ++j) coef[(int64_t)j * Mp + r] = cz(0, 0);
        return;
    }
    double2 a[kMaxKrylov + 1], b[kMaxKrylov + 1], g[kMaxKrylov + 1], t[kMaxKrylov + 1];
    for (int k = 0; k < i; ++k) {
        a[k] = dots[(int64_t)k * Mp + r];
        b[k] = dots[(int64_t)(i + k) * Mp + r];
    }
    const double2 c = dots[(int64_t)(2 * i) * Mp + r];
    double beta2 = dots[(int64_t)(2 * i + 1) * Mp + r].x;
    for (int k = 0; k < i; ++k) beta2 -= cabs2(a[k]);
    if (!(beta2 > 0.0) || !isfinite(beta2)) { // Krylov breakdown (reference: division by 0)
        flags[0] = 1;
        for (int j = 0; j < ncoef; ++j) coef[(int64_t)j * Mp + r] = cz(0, 0);
        return;
    }
    const double beta = sqrt(beta2), ib = 1.0 / beta;
    auto H = [&](int row, int col) -> double2 & { return hess[((int64_t)col * m + row) * Mp + r]; };
    if (i == 0) {
        beta0[r] = beta;
    } else {
        for (int k = 0; k < i; ++k) H(k, i - 1) = cadd(tprev[(int64_t)k * Mp + r], a[k]);
        H(i, i - 1) = cz(beta, 0.0);
    }
    for (int l = 0; l <= i; ++l) { // g = Hbar a
        double2 s = cz(0, 0);
        for (int k = l > 0 ? l - 1 : 0; k < i; ++k) s = cfma(H(l, k), a[k], s);
        g[l] = s;
    }
    double2 ab = cz(0, 0);
    for (int k = 0; k < i; ++k) {
        t[k] = cscale(ib, csub(b[k], g[k]));
        ab = cfma_conj(a[k], b[k], ab);
    }
    t[i] = cscale(ib, csub(cscale(ib, csub(c, ab)), g[i]));
    const double2 s = cadd(cscale(ib, g[i]), t[i]);
    for (int k = 0; k < i; ++k) {
        coef[(int64_t)k * Mp + r] = cscale(-ib, a[k]);
        double2 x = cadd(cscale(ib, g[k]), t[k]);
        x = csub(x, cscale(ib, cmul(s, a[k])));
        coef[(int64_t)(i + k) * Mp + r] = cscale(-1.0, x);
    }
    coef[(int64_t)(2 * i) * Mp + r] = cz(ib, 0.0);
    coef[(int64_t)(2 * i + 1) * Mp + r] = cscale(-ib, s);
    coef[(int64_t)(2 * i + 2) * Mp + r] = cz(ib, 0.0);
    for (int k = 0; k <= i; ++k) tprev[(int64_t)k * Mp + r] = t[k];
}

// Last column: h(k, m-1) = t_{m-1}[k] + a_k.
__global__ void k_finalize(int m, int M, int Mp, const double2 *__restrict__ dots, const double2 *__restrict__ tprev,
                           double2 *__restrict__ hess)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= Mp) return;
    for (int k = 0; k < m; ++k) {
        double2 h = r < M ? cadd(tprev[(int64_t)k * Mp + r], dots[(int64_t)k * Mp + r]) : cz(0, 0);
        hess[((int64_t)(m - 1) * m + k) * Mp + r] = h;
    }
}

// ------------------------------------------------------------------ pass 2
template <int NQ>
__global__ __launch_bounds__(kBlock) void k_pass2(const double2 *__restrict__ Q, double2 *__restrict__ qout,
                                                   double2 *__restrict__ u, const double2 *__restrict__ v,
                                                   const double2 *__restrict__ coef, int64_t N, int Mp, int rpw)
{
    const int nch = Mp / kWave;
    const int lid = xcd_remap(blockIdx.x, gridDim.x);
    const int ch = lid % nch, rb = lid / nch;
    const int lane = threadIdx.x & (kWave - 1);
    const int wave = wave_uniform(threadIdx.x / kWave);
    const int r = ch * kWave + lane;
    const int64_t NMp = N * (int64_t)Mp;
    double2 cq[NQ > 0 ? NQ : 1], cu[NQ > 0 ? NQ : 1];
#pragma unroll
    for (int k = 0; k < NQ; ++k) {
        cq[k] = coef[(int64_t)k * Mp + r];
        cu[k] = coef[(int64_t)(NQ + k) * Mp + r];
    }
    const double2 cqu = coef[(int64_t)(2 * NQ) * Mp + r];
    const double2 cuu = coef[(int64_t)(2 * NQ + 1) * Mp + r];
    const double2 cuv = coef[(int64_t)(2 * NQ + 2) * Mp + r];
    const int64_t row0 = ((int64_t)rb * kWaves + wave) * rpw;
    const int64_t row1 = row0 + rpw < N ? row0 + rpw : N;
    for (int64_t row = row0; row < row1; ++row) {
        const int64_t off = row * Mp + r;
        const double2 uu = u[off], vv = v[off];
        double2 qi = cmul(cqu, uu);
        double2 un = cfma(cuv, vv, cmul(cuu, uu));
#pragma unroll
        for (int k = 0; k < NQ; ++k) {
            const double2 q = Q[k * NMp + off];
            qi = cfma(cq[k], q, qi);
            un = cfma(cu[k], q, un);
        }
        qout[off] = qi;
        u[off] = un;
    }
}

// ----------------------------------------------------------------- psi = Q c
// Visits nuclear rows n' and the four electron rows e*Z + n'; with OBS the
// singlet/triplet populations of the new state are summed on the fly:
// P_Tp = |psi_aa|^2, P_Tm = |psi_bb|^2, P_S = |psi_ab - psi_ba|^2/2,
// P_T0 = |psi_ab + psi_ba|^2/2 (equal to Re(c^H P_X c) of
// convert_to_krylov, qm_dynamics.f90:483-496 and 356-359).
template <int MK, bool OBS>
__global__ __launch_bounds__(kBlock) void k_psi(const double2 *__restrict__ Q, const double2 *__restrict__ cvec,
                                                 double2 *__restrict__ psi, double2 *__restrict__ part, int64_t Z,
                                                 int Mp, int rpw)
{
    const int nch = Mp / kWave;
    const int lid = xcd_remap(blockIdx.x, gridDim.x);
    const int ch = lid % nch, rb = lid / nch;
    const int lane = threadIdx.x & (kWave - 1);
    const int wave = wave_uniform(threadIdx.x / kWave);
    const int r = ch * kWave + lane;
    const int64_t N = 4 * Z;
    const int64_t NMp = N * (int64_t)Mp;
    double2 cc[MK];
#pragma unroll
    for (int k = 0; k < MK; ++k) cc[k] = cvec[(int64_t)k * Mp + r];
    double pS = 0, pT0 = 0, pTp = 0, pTm = 0;
    const int64_t n0 = ((int64_t)rb * kWaves + wave) * rpw;
    const int64_t n1 = n0 + rpw < Z ? n0 + rpw : Z;
    for (int64_t n = n0; n < n1; ++n) {
        double2 p[4];
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const int64_t off = (e * Z + n) * Mp + r;
            double2 s = cz(0, 0);
#pragma unroll
            for (int k = 0; k < MK; ++k) s = cfma(Q[k * NMp + off], cc[k], s);
            p[e] = s;
            psi[off] = s;
        }
        if constexpr (OBS) {
            pTp += cabs2(p[0]);
            pTm += cabs2(p[3]);
            pS += 0.5 * cabs2(csub(p[1], p[2]));
            pT0 += 0.5 * cabs2(cadd(p[1], p[2]));
        }
    }
    if constexpr (OBS) {
        __shared__ double2 red[kWaves - 1][2][kWave];
        double2 a0 = cz(pS, pT0), a1 = cz(pTp, pTm);
        if (wave > 0) {
            red[wave - 1][0][lane] = a0;
            red[wave - 1][1][lane] = a1;
        }
        __syncthreads();
        if (wave == 0) {
            for (int w = 0; w < kWaves - 1; ++w) {
                a0 = cadd(a0, red[w][0][lane]);
                a1 = cadd(a1, red[w][1][lane]);
            }
            part[((int64_t)rb * 2 + 0) * Mp + r] = a0;
            part[((int64_t)rb * 2 + 1) * Mp + r] = a1;
        }
    }
}

// Writes the reduced populations of each lane at its current output index.
__global__ void k_obs_store(int M, int Mp, const double2 *__restrict__ dots, const int32_t *__restrict__ lane_k,
                            double *__restrict__ obs, int64_t n_out)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= M) return;
    const int64_t k = lane_k[r];
    if (k <= 0 || k >= n_out) return;
    const double2 a0 = dots[r], a1 = dots[Mp + r];
    obs[(k * 4 + 0) * Mp + r] = a0.x; // P_S
    obs[(k * 4 + 1) * Mp + r] = a0.y; // P_T0
    obs[(k * 4 + 2) * Mp + r] = a1.x; // P_Tp
    obs[(k * 4 + 3) * Mp + r] = a1.y; // P_Tm
}

// ------------------------------------------------------------------- expm
// One workgroup per lane: Pade(6,6) with scaling and squaring exactly as
// matrix_exponential.f90:14-87 (s from the L-inf norm, q = 6 coefficients,
// E <- D^{-1} E, s squarings), the solve done by Gaussian elimination with
// partial pivoting on [D | E] in LDS.
// C = A B for m x m column-major LDS matrices; C may alias A or B (results are
// staged in registers: m*m <= 1024 = 4 per thread).
__device__ void lds_matmul(int m, const double2 *A, const double2 *B, double2 *C)
{
    double2 tmp[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int idx = threadIdx.x + q * kBlock;
        if (idx < m * m) {
            const int i = idx % m, j = idx / m;
            double2 s = cz(0, 0);
            for (int k = 0; k < m; ++k) s = cfma(A[i + m * k], B[k + m * j], s);
            tmp[q] = s;
        }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int idx = threadIdx.x + q * kBlock;
        if (idx < m * m) C[idx] = tmp[q];
    }
    __syncthreads();
}

__device__ void expm_block(int m, double2 *X, double2 *P, double2 *E, double2 *D)
{
    __shared__ int ipiv[kMaxKrylov + 1];
    __shared__ double2 mult[kMaxKrylov];
    const int tid = threadIdx.x;
    // L-inf norm (mat_norm_linf, matrix_exponential.f90:89-104); D as scratch
    double *red = (double *)D;
    if (tid < m) {
        double rs = 0.0;
        for (int j = 0; j < m; ++j) rs += hypot(X[tid + m * j].x, X[tid + m * j].y);
        red[tid] = rs;
    }
    __syncthreads();
    if (tid == 0) {
        double nrm = 0.0;
        for (int i = 0; i < m; ++i) nrm = fmax(nrm, red[i]);
        int s = 0;
        if (nrm > 0.0) { // log_2 and the s rule, matrix_exponential.f90:48-50
            int ee = (int)(log(fabs(nrm)) / log(2.0)) + 1;
            s = ee + 1 > 0 ? ee + 1 : 0;
        }
        ipiv[kMaxKrylov] = s;
    }
    __syncthreads();
    const int s = ipiv[kMaxKrylov];
    const double sc = 1.0 / ldexp(1.0, s);
    double c = 0.5;
    for (int idx = tid; idx < m * m; idx += blockDim.x) {
        double2 x = cz(X[idx].x * sc, X[idx].y * sc);
        X[idx] = x;
        P[idx] = x;
        const double id = (idx % (m + 1) == 0) ? 1.0 : 0.0;
        E[idx] = cz(id + c * x.x, c * x.y);
        D[idx] = cz(id - c * x.x, -c * x.y);
    }
    __syncthreads();
    bool pos = true;
    for (int k = 2; k <= 6; ++k) {
        c = c * (double)(6 - k + 1) / (double)(k * (2 * 6 - k + 1));
        lds_matmul(m, X, P, P); // X^k (x and X are one variable in the reference)
        for (int idx = tid; idx < m * m; idx += blockDim.x) {
            const double2 p = P[idx];
            E[idx] = cadd(E[idx], cscale(c, p));
            D[idx] = pos ? cadd(D[idx], cscale(c, p)) : csub(D[idx], cscale(c, p));
        }
        pos = !pos;
        __syncthreads();
    }
    // Solve D Y = E (the reference forms inv(D) then inv(D) E), result in E.
    for (int k = 0; k < m; ++k) {
        if (tid == 0) {
            int p = k;
            double best = fabs(D[k + m * k].x) + fabs(D[k + m * k].y);
            for (int i = k + 1; i < m; ++i) {
                double v = fabs(D[i + m * k].x) + fabs(D[i + m * k].y);
                if (v > best) {
                    best = v;
                    p = i;
                }
            }
            ipiv[k] = p;
        }
        __syncthreads();
        const int p = ipiv[k];
        if (p != k)
            for (int j = tid; j < 2 * m; j += blockDim.x) {
                double2 *Mx = j < m ? D : E;
                const int jj = j < m ? j : j - m;
                double2 t0 = Mx[k + m * jj];
                Mx[k + m * jj] = Mx[p + m * jj];
                Mx[p + m * jj] = t0;
            }
        __syncthreads();
        if (tid > k && tid < m) mult[tid] = cdiv(D[tid + m * k], D[k + m * k]);
        __syncthreads();
        for (int idx = tid; idx < (m - k - 1) * 2 * m; idx += blockDim.x) {
            const int i = k + 1 + idx % (m - k - 1), j = idx / (m - k - 1);
            if (j < m) {
                if (j > k) D[i + m * j] = csub(D[i + m * j], cmul(mult[i], D[k + m * j]));
            } else {
                const int jj = j - m;
                E[i + m * jj] = csub(E[i + m * jj], cmul(mult[i], E[k + m * jj]));
            }
        }
        __syncthreads();
    }
    for (int k = m - 1; k >= 0; --k) {
        const double2 piv = D[k + m * k];
        for (int j = tid; j < m; j += blockDim.x) E[k + m * j] = cdiv(E[k + m * j], piv);
        __syncthreads();
        for (int idx = tid; idx < k * m; idx += blockDim.x) {
            const int i = idx % k, j = idx / k;
            E[i + m * j] = csub(E[i + m * j], cmul(D[i + m * k], E[k + m * j]));
        }
        __syncthreads();
    }
    for (int q = 0; q < s; ++q) lds_matmul(m, E, E, E);
}

// mode 0: X = -i dt hess (lane-strided input), output E lane-strided.
// mode 1: single matrix, X = A (column-major), output column-major.
__global__ __launch_bounds__(kBlock) void k_expm(int m, int Mp, int M, double dt, const double2 *__restrict__ in,
                                                  double2 *__restrict__ out, int mode)
{
    extern __shared__ double2 lds[];
    const int r = blockIdx.x;
    const int mm = m * m;
    double2 *X = lds, *P = X + mm, *E = P + mm, *D = E + mm;
    if (mode == 0 && r >= M) {
        for (int idx = threadIdx.x; idx < mm; idx += blockDim.x) out[(int64_t)idx * Mp + r] = cz(0, 0);
        return;
    }
    for (int idx = threadIdx.x; idx < mm; idx += blockDim.x) {
        if (mode == 0) {
            const double2 h = in[(int64_t)idx * Mp + r];
            X[idx] = cz(dt * h.y, -dt * h.x); // (0,-1)*dt*h
        } else {
            X[idx] = in[idx];
        }
    }
    __syncthreads();
    expm_block(m, X, P, E, D);
    for (int idx = threadIdx.x; idx < mm; idx += blockDim.x) {
        if (mode == 0)
            out[(int64_t)idx * Mp + r] = E[idx];
        else
            out[idx] = E[idx];
    }
}

} // namespace

size_t expm_lds_bytes(int m) { return 4 * (size_t)m * m * sizeof(double2); }

namespace {

// --------------------------------------------------------------- step loop
// qm_dynamics.f90:348-364 for each lane: c <- E c while |c_m|^2 <= tol |c|^2
// and t < t_end, the step that exceeds tol kept.  With gram != 0 the
// populations of every step are Re(c^H G_X c); otherwise they come from the
// psi update and a lane must take exactly one step (flag 1 otherwise).
__global__ void k_step(int m, int M, int Mp, double dt, double t_end, double tol, int gram,
                       const double2 *__restrict__ E, const double *__restrict__ beta0, double2 *__restrict__ cout,
                       double *__restrict__ lane_t, int32_t *__restrict__ lane_k, double *__restrict__ obs,
                       int64_t n_out, const double2 *__restrict__ G, int32_t *__restrict__ flags)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= M) return;
    double t = lane_t[r];
    int64_t k = lane_k[r];
    double2 c[kMaxKrylov], cn[kMaxKrylov];
    for (int j = 0; j < m; ++j) c[j] = cz(0, 0);
    c[0] = cz(beta0[r], 0.0);
    if (t < t_end) {
        double c2 = 1.0, cm2 = 0.0;
        int steps = 0;
        while (cm2 <= tol * c2 && t < t_end) {
            if (!gram && steps == 1) {
                flags[1] = 1;
                break;
            }
            for (int i = 0; i < m; ++i) {
                double2 s = cz(0, 0);
                for (int j = 0; j < m; ++j) s = cfma(E[((int64_t)j * m + i) * Mp + r], c[j], s);
                cn[i] = s;
            }
            for (int i = 0; i < m; ++i) c[i] = cn[i];
            k = k + 1;
            if (k >= n_out) {
                flags[3] = 1;
                break;
            }
            if (gram) {
                for (int x = 0; x < 4; ++x) {
                    const double2 *Gx = G + (int64_t)x * m * m * Mp;
                    double acc = 0.0;
                    for (int i = 0; i < m; ++i) {
                        double2 s = cz(0, 0);
                        for (int j = 0; j < m; ++j) s = cfma(Gx[((int64_t)i * m + j) * Mp + r], c[j], s);
                        acc += c[i].x * s.x + c[i].y * s.y; // Re(conj(c_i) s)
                    }
                    obs[(k * 4 + x) * Mp + r] = acc;
                }
            }
            c2 = 0.0;
            for (int i = 0; i < m; ++i) c2 += cabs2(c[i]);
            cm2 = cabs2(c[m - 1]);
            t = t + dt;
            ++steps;
        }
    }
    for (int j = 0; j < m; ++j) cout[(int64_t)j * Mp + r] = c[j];
    lane_t[r] = t;
    lane_k[r] = (int32_t)k;
}

// ----------------------------------------------------- projector Grams
// G_X(i,j) = <x_i, x_j> over the nuclear rows for x = singlet combination
// (ab - ba)/sqrt2, T0 combination (ab + ba)/sqrt2, Tp block, Tm block.
// Re(c^H G_X c) equals Re(c^H P_X c) of convert_to_krylov
// (qm_dynamics.f90:469-498), whose P_S/P_T0 differ from G_X only by
// anti-Hermitian parts.  Used when tol > 0 (several steps per rebuild).
__global__ __launch_bounds__(kWave) void k_gram(const double2 *__restrict__ Q, double2 *__restrict__ part, int m,
                                                 int64_t Z, int Mp, int64_t rows_per_block)
{
    const int nch = Mp / kWave;
    const int ch = blockIdx.x % nch;
    const int rb = blockIdx.x / nch;
    const int xi = blockIdx.y; // x*m + i
    const int x = xi / m, i = xi % m;
    const int r = ch * kWave + threadIdx.x;
    const int64_t NMp = 4 * Z * (int64_t)Mp;
    double2 acc[kMaxKrylov];
    for (int j = 0; j < m; ++j) acc[j] = cz(0, 0);
    const double h = 0.70710678118654752440;
    auto load = [&](int j, int64_t n) -> double2 {
        // x follows the output order: 0 P_S, 1 P_T0, 2 P_Tp, 3 P_Tm
        const double2 *q = Q + (int64_t)j * NMp;
        if (x == 2) return q[n * Mp + r];
        if (x == 3) return q[(3 * Z + n) * Mp + r];
        const double2 ab = q[(Z + n) * Mp + r], ba = q[(2 * Z + n) * Mp + r];
        return x == 0 ? cscale(h, csub(ab, ba)) : cscale(h, cadd(ab, ba));
    };
    const int64_t n0 = rb * rows_per_block;
    const int64_t n1 = n0 + rows_per_block < Z ? n0 + rows_per_block : Z;
    for (int64_t n = n0; n < n1; ++n) {
        const double2 xi_ = load(i, n);
        for (int j = 0; j < m; ++j) acc[j] = cfma_conj(xi_, load(j, n), acc[j]);
    }
    for (int j = 0; j < m; ++j) part[(((int64_t)rb * 4 * m + xi) * m + j) * Mp + r] = acc[j];
}

// ---------------------------------------------------------- dispatch tables
using Pass1Fn = void (*)(const int32_t *, const int32_t *, const double2 *, const double2 *, const double2 *,
                         double2 *, double2 *, int64_t, int, int);
using Pass2Fn = void (*)(const double2 *, double2 *, double2 *, const double2 *, const double2 *, int64_t, int, int);
using PsiFn = void (*)(const double2 *, const double2 *, double2 *, double2 *, int64_t, int, int);

template <size_t... I>
constexpr auto make_pass1(std::index_sequence<I...>, std::true_type)
{
    return std::array<Pass1Fn, sizeof...(I)>{&k_pass1<(int)I, true>...};
}
template <size_t... I>
constexpr auto make_pass1(std::index_sequence<I...>, std::false_type)
{
    return std::array<Pass1Fn, sizeof...(I)>{&k_pass1<(int)I, false>...};
}
template <size_t... I>
constexpr auto make_pass2(std::index_sequence<I...>)
{
    return std::array<Pass2Fn, sizeof...(I)>{&k_pass2<(int)I>...};
}
template <size_t... I>
constexpr auto make_psi(std::index_sequence<I...>, std::true_type)
{
    return std::array<PsiFn, sizeof...(I)>{&k_psi<(int)I + 1, true>...};
}
template <size_t... I>
constexpr auto make_psi(std::index_sequence<I...>, std::false_type)
{
    return std::array<PsiFn, sizeof...(I)>{&k_psi<(int)I + 1, false>...};
}

const auto kPass1Spmm = make_pass1(std::make_index_sequence<kMaxKrylov + 1>{}, std::true_type{});
const auto kPass1Dots = make_pass1(std::make_index_sequence<kMaxKrylov + 1>{}, std::false_type{});
const auto kPass2 = make_pass2(std::make_index_sequence<kMaxKrylov + 1>{});
const auto kPsiObs = make_psi(std::make_index_sequence<kMaxKrylov>{}, std::true_type{});
const auto kPsiPlain = make_psi(std::make_index_sequence<kMaxKrylov>{}, std::false_type{});

inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

// Algorithmic bytes: one complex128 pass over the M useful lanes, and the CSR.
inline double vec_bytes(const KrylovWork &w) { return 16.0 * (double)w.N * (double)w.M; }
inline double csr_bytes(const KrylovWork &w) { return 20.0 * (double)w.nnz + 4.0 * (double)(w.N + 1); }

struct ProfScope {
    const KrylovWork &w;
    hipStream_t s;
    bool on;
    ProfScope(const KrylovWork &w_, int fam, double bytes, hipStream_t s_) : w(w_), s(s_), on(w_.prof && w_.prof->on)
    {
        if (on) w.prof->begin(fam, bytes, s);
    }
    ~ProfScope()
    {
        if (on) w.prof->end(s);
    }
};

inline int nrb_for(int64_t rows, int rpw) { return (int)ceil_div(rows, (int64_t)kWaves * rpw); }

hipError_t reduce(const KrylovWork &w, int nacc, int nrb, hipStream_t s)
{
    const int64_t n = (int64_t)nacc * w.Mp;
    if (n == 0) return hipSuccess;
    ProfScope ps(w, PF_REDUCE, 16.0 * (double)n * (nrb + 1), s);
    hipLaunchKernelGGL(k_reduce, dim3((unsigned)ceil_div(n, 256)), dim3(256), 0, s, w.part, w.dots, nacc, nrb, w.Mp);
    return hipGetLastError();
}

} // namespace

size_t partials_needed(const KrylovWork &w)
{
    const size_t nrb1 = nrb_for(w.N, w.rpw1), nrbp = nrb_for(w.Z, w.rpwp);
    size_t need = nrb1 * (size_t)(2 * w.m + 2) * w.Mp;
    need = std::max(need, nrbp * (size_t)2 * w.Mp);
    return need;
}

hipError_t arnoldi(KrylovWork &w, hipStream_t s)
{
    const int nch = w.Mp / kWave;
    const int nrb1 = nrb_for(w.N, w.rpw1), nrb2 = nrb_for(w.N, w.rpw2);
    const int cblocks = (int)ceil_div(w.Mp, 64);
    hipError_t e = hipSuccess;
    // u_0 = psi: w.u aliases w.psi (pass 2 overwrites it in place; psi is
    // rebuilt from Q c after the step)
    for (int i = 0; i < w.m; ++i) {
        {
        ProfScope ps(w, PF_PASS1, (i + 2) * vec_bytes(w) + csr_bytes(w), s);
        hipLaunchKernelGGL(kPass1Spmm[i], dim3(nrb1 * nch), dim3(kBlock), 0, s, w.rowptr, w.col, w.val, w.Q, w.u,
                           w.v, w.part, w.N, w.Mp, w.rpw1);
        }
        if ((e = hipGetLastError()) != hipSuccess) return e;
        if ((e = reduce(w, 2 * i + 2, nrb1, s)) != hipSuccess) return e;
        {
        ProfScope ps(w, PF_COEFFS, 0.0, s);
        hipLaunchKernelGGL(k_coeffs, dim3(cblocks), dim3(64), 0, s, i, w.m, w.M, w.Mp, w.dots, w.coef, w.tprev,
                           w.hess, w.beta0, w.flags);
        }
        if ((e = hipGetLastError()) != hipSuccess) return e;
        {
        ProfScope ps(w, PF_PASS2, (i + 4) * vec_bytes(w), s);
        hipLaunchKernelGGL(kPass2[i], dim3(nrb2 * nch), dim3(kBlock), 0, s, w.Q, w.Q + (int64_t)i * w.N * w.Mp, w.u,
                           w.v, w.coef, w.N, w.Mp, w.rpw2);
        }
        if ((e = hipGetLastError()) != hipSuccess) return e;
    }
    {
    ProfScope ps(w, PF_DOTS, (w.m + 1) * vec_bytes(w), s);
    hipLaunchKernelGGL(kPass1Dots[w.m], dim3(nrb1 * nch), dim3(kBlock), 0, s, w.rowptr, w.col, w.val, w.Q, w.u, w.v,
                       w.part, w.N, w.Mp, w.rpw1);
    }
    if ((e = hipGetLastError()) != hipSuccess) return e;
    if ((e = reduce(w, w.m, nrb1, s)) != hipSuccess) return e;
    hipLaunchKernelGGL(k_finalize, dim3(cblocks), dim3(64), 0, s, w.m, w.M, w.Mp, w.dots, w.tprev, w.hess);
    return hipGetLastError();
}

hipError_t spmm_only(KrylovWork &w, hipStream_t s)
{
    const int nch = w.Mp / kWave;
    const int nrb1 = nrb_for(w.N, w.rpw1);
    hipLaunchKernelGGL(kPass1Spmm[0], dim3(nrb1 * nch), dim3(kBlock), 0, s, w.rowptr, w.col, w.val, w.Q, w.u, w.v,
                       w.part, w.N, w.Mp, w.rpw1);
    return hipGetLastError();
}

hipError_t expm_lanes(KrylovWork &w, double dt, hipStream_t s)
{
    const size_t lds = expm_lds_bytes(w.m);
    ProfScope ps(w, PF_EXPM, 0.0, s);
    hipLaunchKernelGGL(k_expm, dim3(w.Mp), dim3(kBlock), lds, s, w.m, w.Mp, w.M, dt, w.hess, w.E, 0);
    return hipGetLastError();
}

hipError_t expm_single(int m, const double2 *A, double2 *E, hipStream_t s)
{
    const size_t lds = expm_lds_bytes(m);
    hipLaunchKernelGGL(k_expm, dim3(1), dim3(kBlock), lds, s, m, 1, 1, 0.0, A, E, 1);
    return hipGetLastError();
}

hipError_t gram_projectors(KrylovWork &w, hipStream_t s)
{
    const int nch = w.Mp / kWave;
    // row blocks sized so the partials fit the shared buffer
    int64_t per = (int64_t)4 * w.m * w.m * w.Mp;
    int nrb = (int)std::max<int64_t>(1, std::min<int64_t>(64, (int64_t)(w.part_capacity / per)));
    int64_t rows = ceil_div(w.Z, nrb);
    nrb = (int)ceil_div(w.Z, rows);
    ProfScope ps(w, PF_GRAM, 0.0, s);
    hipLaunchKernelGGL(k_gram, dim3(nrb * nch, 4 * w.m), dim3(kWave), 0, s, w.Q, w.part, w.m, w.Z, w.Mp, rows);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    const int nacc = 4 * w.m * w.m;
    const int64_t n = (int64_t)nacc * w.Mp;
    hipLaunchKernelGGL(k_reduce, dim3((unsigned)ceil_div(n, 256)), dim3(256), 0, s, w.part, w.gram, nacc, nrb, w.Mp);
    return hipGetLastError();
}

hipError_t step_lanes(KrylovWork &w, double dt, double t_end, double tol, bool gram, hipStream_t s)
{
    ProfScope ps(w, PF_STEP, 0.0, s);
    hipLaunchKernelGGL(k_step, dim3((unsigned)ceil_div(w.Mp, 64)), dim3(64), 0, s, w.m, w.M, w.Mp, dt, t_end, tol,
                       gram ? 1 : 0, w.E, w.beta0, w.c, w.lane_t, w.lane_k, w.obs, w.n_out, w.gram, w.flags);
    return hipGetLastError();
}

hipError_t psi_update(KrylovWork &w, bool observe, hipStream_t s)
{
    const int nch = w.Mp / kWave;
    const int nrbp = nrb_for(w.Z, w.rpwp);
    const PsiFn fn = observe ? kPsiObs[w.m - 1] : kPsiPlain[w.m - 1];
    {
    ProfScope ps(w, PF_PSI, (w.m + 1) * vec_bytes(w), s);
    hipLaunchKernelGGL(fn, dim3(nrbp * nch), dim3(kBlock), 0, s, w.Q, w.c, w.psi, w.part, w.Z, w.Mp, w.rpwp);
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess || !observe) return e;
    if ((e = reduce(w, 2, nrbp, s)) != hipSuccess) return e;
    hipLaunchKernelGGL(k_obs_store, dim3((unsigned)ceil_div(w.M, 64)), dim3(64), 0, s, w.M, w.Mp, w.dots, w.lane_k,
                       w.obs, w.n_out);
    return hipGetLastError();
}

namespace {
// acc[x*n_out + k] += sum_{r<M} obs[(k*4 + x)*Mp + r], lanes in order.
__global__ void k_obs_lanes(const double *__restrict__ obs, int64_t n_out, int M, int Mp, double *__restrict__ acc)
{
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= 4 * n_out) return;
    const int x = (int)(idx / n_out);
    const int64_t k = idx % n_out;
    double s = acc[idx];
    for (int r = 0; r < M; ++r) s += obs[(k * 4 + x) * Mp + r];
    acc[idx] = s;
}
} // namespace

hipError_t obs_reduce_lanes(const double *obs, int64_t n_out, int M, int Mp, double *acc, hipStream_t s)
{
    hipLaunchKernelGGL(k_obs_lanes, dim3((unsigned)ceil_div(4 * n_out, 256)), dim3(256), 0, s, obs, n_out, M, Mp, acc);
    return hipGetLastError();
}

} // namespace spc
