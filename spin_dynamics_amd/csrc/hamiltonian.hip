// hamiltonian.hip -- on-device CSR assembly of the radical-pair Hamiltonian.
//
// Replaces getHiso_sp + coo_to_csr (hamiltonian.f90:17-205, utils.f90:52-89).
// The reference builds H from Kronecker products of spin-operator factors
// (spinop.f90:58-95, kroenecker.f90:56-200); here each row is generated from
// the digits of its basis index: a term acting on a set of spins connects the
// row to the column whose digits differ on exactly that set.  The device basis
// keeps the electron pair as the fastest index (row = 4*nuc + e, the
// reference's e*Z + nuc permuted), so electron flips stay within one 4-row
// group and the gathers of the sparse product stay local.
//
// H = -B n.(S1+S2) + sum_k [S_e(k).A_k.I_k + zeta_k n.I_k] + S1.T.S2 - i kbar/2
// with T = (2J + 2i dk) 1 + D, dk = (kS - kT)/4, kbar = (kS + 3kT)/4; this is
// the reference's Zeeman + hyperfine + exchange/dipolar - iK/2 with
// K = kS P_S + kT P_T.  Exact zeros off the diagonal are not stored, which
// gives the reference's sparsity pattern for isotropic couplings (nnz formula
// hamiltonian.f90:62-73).
#include <hipcub/hipcub.hpp>

#include <algorithm>

#include "common.hpp"
#include "internal.hpp"

namespace spc {

namespace {

// Transverse parts are written with ladder operators, S_x = (S+ + S-)/2,
// S_y = (S+ - S-)/(2i):
//   f_x S_x + f_y S_y = S+ (f_x - i f_y)/2 + S- (f_x + i f_y)/2
//   sum_{a,b in x,y} X_ab S_a I_b = sum_{s,t = +,-} C_st S_s I_t with
//   C_++ = [(X_xx - X_yy) - i(X_xy + X_yx)]/4,  C_+- = [(X_xx + X_yy) + i(X_xy - X_yx)]/4,
//   C_-+ = [(X_xx + X_yy) - i(X_xy - X_yx)]/4,  C_-- = [(X_xx - X_yy) + i(X_xy + X_yx)]/4.
// Couplings that vanish by symmetry (isotropic tensors, field along z) then
// give exact zeros, as the reference's explicit S+I- / S-I+ terms do.

// Coefficient of S_s (s = +1 raising, -1 lowering) in f_x S_x + f_y S_y.
__device__ __forceinline__ double2 ladder1(double2 fx, double2 fy, int s)
{
    // (f_x -/+ i f_y)/2
    return s > 0 ? cz(0.5 * (fx.x + fy.y), 0.5 * (fx.y - fy.x)) : cz(0.5 * (fx.x - fy.y), 0.5 * (fx.y + fy.x));
}

// C_st for a complex 3x3 coupling X(a,b) given as (re, im) column-major arrays.
__device__ __forceinline__ double2 ladder2(const double *re, const double *im, int s, int t)
{
    const double xx_r = re[0], xx_i = im ? im[0] : 0.0, yy_r = re[4], yy_i = im ? im[4] : 0.0;
    const double xy_r = re[3], xy_i = im ? im[3] : 0.0, yx_r = re[1], yx_i = im ? im[1] : 0.0;
    double2 d, o; // d = X_xx -/+ X_yy, o = X_xy +/- X_yx
    if (s == t) {
        d = cz(xx_r - yy_r, xx_i - yy_i);
        o = cz(xy_r + yx_r, xy_i + yx_i);
    } else {
        d = cz(xx_r + yy_r, xx_i + yy_i);
        o = cz(xy_r - yx_r, xy_i - yx_i);
    }
    // sign of i*o: ++ : -, +- : +, -+ : -, -- : +
    const double sg = (s > 0) == (t > 0) ? (s > 0 ? -1.0 : 1.0) : (s > 0 ? 1.0 : -1.0);
    // d + sg * i * o
    return cz(0.25 * (d.x - sg * o.y), 0.25 * (d.y + sg * o.x));
}

// Ladder matrix elements <d|I+|d+1> and <d|I-|d-1> for multiplicity g.
__device__ __forceinline__ double nuc_ladder(int g, int d, int dc)
{
    const double S = 0.5 * (g - 1), mc = S - dc;
    if (dc == d + 1) return sqrt(S * (S + 1.0) - mc * (mc + 1.0)); // I+ raises the column state
    return sqrt(S * (S + 1.0) - mc * (mc - 1.0));                  // I- lowers it
}

__device__ __forceinline__ double2 Tab(const DevSystem &S, int a, int b)
{
    return cz(S.Tre[3 * b + a], S.Tim[3 * b + a]);
}

__device__ __forceinline__ bool nonzero(double2 v) { return v.x != 0.0 || v.y != 0.0; }

// Nuclear flip k (electron unchanged): g_b = zeta n_b + A(z,b) m_e.
__device__ double2 nuc_value(const DevSystem &S, int k, int d, int dc, double me)
{
    const double gx = S.zeta[k] * S.dir[0] + S.A[k][3 * 0 + 2] * me;
    const double gy = S.zeta[k] * S.dir[1] + S.A[k][3 * 1 + 2] * me;
    const double2 c = ladder1(cz(gx, 0.0), cz(gy, 0.0), dc == d + 1 ? 1 : -1);
    return cscale(nuc_ladder(S.mult[k], d, dc), c);
}

// Electron e(k) and nucleus k flip together: C_st(A_k) <ladders>.
__device__ double2 enuc_value(const DevSystem &S, int k, int se, int d, int dc)
{
    const int s = se == 0 ? 1 : -1; // row alpha <- column beta is S+
    const double2 c = ladder2(S.A[k], nullptr, s, dc == d + 1 ? 1 : -1);
    return cscale(nuc_ladder(S.mult[k], d, dc), c);
}

// Emits the row's entries in ascending column order.  Device basis order is
// electron-fastest: row = 4*nuc + e with e = 2*s1 + s2 (0 = alpha), so the
// entries of a row come as nuclear offsets nuc - s_k (k ascending), nuc,
// nuc + s_k (k descending), and within one nuclear column by electron state.
template <class Emit>
__device__ void gen_row(const DevSystem &S, int64_t row, Emit &&emit)
{
    const int e = (int)(row & 3);
    const int64_t nuc = row >> 2;
    const int s1 = e >> 1, s2 = e & 1;
    const double m1 = 0.5 - s1, m2 = 0.5 - s2;
    int d[kMaxNuclei];
    double mk[kMaxNuclei];
    for (int k = 0; k < S.nn; ++k) {
        d[k] = (int)((nuc / S.stride[k]) % S.mult[k]);
        mk[k] = 0.5 * (S.mult[k] - 1) - d[k];
    }
    const int e1f = 2 * (1 - s1) + s2, e2f = 2 * s1 + (1 - s2);

    // nuclear flip of nucleus k to digit dc: the nucleus alone (electron e) and
    // together with its electron (state ef), emitted in electron order
    auto flip_pair = [&](int k, int dc, int64_t nuc_c) {
        const int el = k < S.n1 ? 0 : 1;
        const int se = el == 0 ? s1 : s2;
        const int ef = el == 0 ? e1f : e2f;
        const double2 vn = nuc_value(S, k, d[k], dc, el == 0 ? m1 : m2);
        const double2 ve = enuc_value(S, k, se, d[k], dc);
        const int64_t b = 4 * nuc_c;
        if (e < ef) {
            if (nonzero(vn)) emit(b + e, vn);
            if (nonzero(ve)) emit(b + ef, ve);
        } else {
            if (nonzero(ve)) emit(b + ef, ve);
            if (nonzero(vn)) emit(b + e, vn);
        }
    };

    for (int k = 0; k < S.nn; ++k) // nuc - stride_k, descending stride
        if (d[k] >= 1) flip_pair(k, d[k] - 1, nuc - S.stride[k]);

    for (int ec = 0; ec < 4; ++ec) { // same nuclear index
        const int64_t col = 4 * nuc + ec;
        if (ec == e) {
            double2 v = cz(0.0, -0.5 * S.kbar);
            v = cadd(v, cz(S.w[2] * (m1 + m2), 0.0));
            v = cadd(v, cscale(m1 * m2, Tab(S, 2, 2)));
            for (int k = 0; k < S.nn; ++k) {
                double me = k < S.n1 ? m1 : m2;
                v.x = fma(S.A[k][8] * me + S.zeta[k] * S.dir[2], mk[k], v.x);
            }
            emit(col, v); // the diagonal is always stored
        } else if (ec == e1f || ec == e2f) {
            const bool first = ec == e1f;
            const int sr = first ? s1 : s2;
            const double mo = first ? m2 : m1;
            const int k0 = first ? 0 : S.n1, k1 = first ? S.n1 : S.nn;
            double2 f[2];
            for (int a = 0; a < 2; ++a) {
                f[a] = cz(S.w[a], 0.0);
                f[a] = cadd(f[a], cscale(mo, first ? Tab(S, a, 2) : Tab(S, 2, a)));
                for (int k = k0; k < k1; ++k) f[a].x = fma(S.A[k][3 * 2 + a], mk[k], f[a].x);
            }
            const double2 v = ladder1(f[0], f[1], sr == 0 ? 1 : -1);
            if (nonzero(v)) emit(col, v);
        } else {
            const double2 v = ladder2(S.Tre, S.Tim, s1 == 0 ? 1 : -1, s2 == 0 ? 1 : -1);
            if (nonzero(v)) emit(col, v);
        }
    }

    for (int k = S.nn - 1; k >= 0; --k) // nuc + stride_k, ascending stride
        if (d[k] + 1 < S.mult[k]) flip_pair(k, d[k] + 1, nuc + S.stride[k]);
}

__global__ void k_h_count(const DevSystem *__restrict__ sys, int64_t N, int32_t *__restrict__ counts)
{
    int64_t row = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (row >= N) {
        if (row == N) counts[N] = 0;
        return;
    }
    int c = 0;
    gen_row(*sys, row, [&](int64_t, double2) { ++c; });
    counts[row] = c;
}

__global__ void k_h_fill(const DevSystem *__restrict__ sys, int64_t N, const int32_t *__restrict__ rowptr,
                         int32_t *__restrict__ col, double2 *__restrict__ val)
{
    int64_t row = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (row >= N) return;
    int32_t p = rowptr[row];
    gen_row(*sys, row, [&](int64_t c, double2 v) {
        col[p] = (int32_t)c;
        val[p] = v;
        ++p;
    });
}

// Block row n merges CSR rows 4n..4n+3 (each sorted by column, so by nuclear
// column nc = col >> 2): the union of their nc is the block columns.
__global__ void k_block_count(const int32_t *__restrict__ rowptr, const int32_t *__restrict__ col, int64_t Z,
                              int32_t *__restrict__ bcounts)
{
    const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (n > Z) return;
    if (n == Z) {
        bcounts[Z] = 0;
        return;
    }
    int p[4], pe[4];
    for (int e = 0; e < 4; ++e) {
        p[e] = rowptr[4 * n + e];
        pe[e] = rowptr[4 * n + e + 1];
    }
    int cnt = 0;
    for (;;) {
        int best = INT32_MAX;
        for (int e = 0; e < 4; ++e)
            if (p[e] < pe[e]) best = min(best, col[p[e]] >> 2);
        if (best == INT32_MAX) break;
        for (int e = 0; e < 4; ++e)
            while (p[e] < pe[e] && (col[p[e]] >> 2) == best) ++p[e];
        ++cnt;
    }
    bcounts[n] = cnt;
}

__global__ void k_block_fill(const int32_t *__restrict__ rowptr, const int32_t *__restrict__ col,
                             const double2 *__restrict__ val, int64_t Z, const int32_t *__restrict__ browptr,
                             int32_t *__restrict__ bcol, double2 *__restrict__ bval)
{
    const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= Z) return;
    int p[4], pe[4];
    for (int e = 0; e < 4; ++e) {
        p[e] = rowptr[4 * n + e];
        pe[e] = rowptr[4 * n + e + 1];
    }
    int b = browptr[n];
    for (;;) {
        int best = INT32_MAX;
        for (int e = 0; e < 4; ++e)
            if (p[e] < pe[e]) best = min(best, col[p[e]] >> 2);
        if (best == INT32_MAX) break;
        bcol[b] = best;
        double2 *B = bval + (int64_t)b * 16;
        for (int q = 0; q < 16; ++q) B[q] = cz(0, 0);
        for (int e = 0; e < 4; ++e)
            while (p[e] < pe[e] && (col[p[e]] >> 2) == best) {
                B[4 * e + (col[p[e]] & 3)] = val[p[e]];
                ++p[e];
            }
        ++b;
    }
}

// Values of a CSR into an existing 4x4 block pattern (browptr, bcol) whose
// blocks cover every CSR entry; pattern positions without an entry get 0.
__global__ void k_block_fill_pattern(const int32_t *__restrict__ rowptr, const int32_t *__restrict__ col,
                                     const double2 *__restrict__ val, int64_t Z, const int32_t *__restrict__ browptr,
                                     const int32_t *__restrict__ bcol, double2 *__restrict__ bval,
                                     int32_t *__restrict__ missing)
{
    const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= Z) return;
    const int b0 = browptr[n], b1 = browptr[n + 1];
    for (int b = b0; b < b1; ++b)
        for (int q = 0; q < 16; ++q) bval[(int64_t)b * 16 + q] = cz(0, 0);
    for (int e = 0; e < 4; ++e)
        for (int p = rowptr[4 * n + e]; p < rowptr[4 * n + e + 1]; ++p) {
            const int c = col[p], cb = c >> 2;
            int b = b0;
            while (b < b1 && bcol[b] != cb) ++b;
            if (b == b1) {
                missing[0] = 1;
                continue;
            }
            bval[(int64_t)b * 16 + 4 * e + (c & 3)] = val[p];
        }
}

// out[o][i] = v0[i] + sum_a f[o][a] za[i] over the nb*16 block values.
__global__ void k_field_values(const double2 *__restrict__ v0, const double2 *__restrict__ zx,
                               const double2 *__restrict__ zy, const double2 *__restrict__ zz,
                               const double *__restrict__ f, int64_t count, int n_fields, double2 *__restrict__ out)
{
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < count * n_fields; idx += stride) {
        const int64_t o = idx / count, i = idx % count;
        const double fx = f[3 * o], fy = f[3 * o + 1], fz = f[3 * o + 2];
        double2 v = v0[i];
        const double2 x = zx[i], y = zy[i], z = zz[i];
        v.x = fma(fx, x.x, v.x);
        v.y = fma(fx, x.y, v.y);
        v.x = fma(fy, y.x, v.x);
        v.y = fma(fy, y.y, v.y);
        v.x = fma(fz, z.x, v.x);
        v.y = fma(fz, z.y, v.y);
        out[idx] = v;
    }
}

// ------------------------------------------------------- Kronecker tables
__device__ int find_block(const int32_t *__restrict__ browptr, const int32_t *__restrict__ bcol, int64_t n, int64_t c)
{
    for (int b = browptr[n]; b < browptr[n + 1]; ++b)
        if (bcol[b] == c) return b;
    return -1;
}

// the 2x2 electron matrix of a flip block B of the nucleus on electron el
__device__ void kron_2x2(const double2 *B, int el, double2 *M)
{
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j) M[2 * i + j] = el == 0 ? B[4 * (2 * i) + 2 * j] : B[4 * i + j];
}

__global__ void k_kron_diag(const int32_t *__restrict__ browptr, const int32_t *__restrict__ bcol,
                            const double2 *__restrict__ bval, int64_t nb, int nsets, int64_t Z,
                            double2 *__restrict__ diag, int32_t *__restrict__ bad)
{
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= Z * nsets) return;
    const int64_t set = idx / Z, n = idx % Z;
    const int b = find_block(browptr, bcol, n, n);
    double2 *D = diag + (set * Z + n) * 16;
    for (int q = 0; q < 16; ++q) D[q] = b >= 0 ? bval[(set * nb + b) * 16 + q] : cz(0, 0);
    if (b < 0) bad[0] = 1;
}

__global__ void k_kron_gen(const KronDev *__restrict__ kd, const int32_t *__restrict__ browptr,
                           const int32_t *__restrict__ bcol, const double2 *__restrict__ bval, int64_t nb, int nsets,
                           double2 *__restrict__ gen)
{
    const int nn = kd->nn;
    const int idx = blockIdx.x * blockDim.x + threadIdx.x; // (set, k, dir, d)
    if (idx >= nsets * nn * 6) return;
    const int set = idx / (nn * 6), k = (idx / 6) % nn, dir = (idx / 3) % 2, d = idx % 3;
    double2 M[4] = {cz(0, 0), cz(0, 0), cz(0, 0), cz(0, 0)};
    const int g = kd->mult[k];
    const int64_t st = kd->stride[k];
    const bool valid = dir == 0 ? d < g - 1 : (d > 0 && d < g);
    if (valid) {
        const int64_t n = d * st, c = dir == 0 ? n + st : n - st;
        const int b = find_block(browptr, bcol, n, c);
        if (b >= 0) kron_2x2(bval + (set * nb + b) * 16, kd->elec[k], M);
    }
    double2 *out = gen + (((int64_t)set * nn + k) * 6 + dir * 3 + d) * 4;
    for (int q = 0; q < 4; ++q) out[q] = M[q];
}

// Every block of every row must be the diagonal block or c M (x) 1 of one
// nucleus as tabulated, and every tabulated nonzero flip must be present.
__global__ void k_kron_check(const KronDev *__restrict__ kd, const int32_t *__restrict__ browptr,
                             const int32_t *__restrict__ bcol, const double2 *__restrict__ bval, int64_t nb,
                             int nsets, int64_t Z, const double2 *__restrict__ gen, int32_t *__restrict__ bad)
{
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= Z * nsets) return;
    const int64_t set = idx / Z, n = idx % Z;
    const int nn = kd->nn;
    const double2 *G = gen + (int64_t)set * nn * 24;
    for (int b = browptr[n]; b < browptr[n + 1]; ++b) {
        const int64_t c = bcol[b];
        if (c == n) continue;
        int kk = -1, dir = 0, d = 0;
        for (int k = 0; k < nn; ++k) {
            const int64_t st = kd->stride[k];
            const int dn = (int)((n / st) % kd->mult[k]);
            if (c == n + st && dn < kd->mult[k] - 1) kk = k, dir = 0, d = dn;
            if (c == n - st && dn > 0) kk = k, dir = 1, d = dn;
        }
        if (kk < 0) {
            bad[0] = 1;
            return;
        }
        const double2 *B = bval + (set * nb + b) * 16;
        const double2 *M = G + ((int64_t)kk * 6 + dir * 3 + d) * 4;
        const int el = kd->elec[kk];
        for (int e = 0; e < 4; ++e)
            for (int f = 0; f < 4; ++f) {
                const int a = e >> 1, bb = e & 1, a2 = f >> 1, b2 = f & 1;
                double2 want = cz(0, 0);
                if (el == 0 && bb == b2) want = M[2 * a + a2];
                if (el == 1 && a == a2) want = M[2 * bb + b2];
                const double2 got = B[4 * e + f];
                if (got.x != want.x || got.y != want.y) {
                    bad[0] = 1;
                    return;
                }
            }
    }
    for (int k = 0; k < nn; ++k) {
        const int64_t st = kd->stride[k];
        const int g = kd->mult[k];
        const int dn = (int)((n / st) % g);
        for (int dir = 0; dir < 2; ++dir) {
            const bool valid = dir == 0 ? dn < g - 1 : dn > 0;
            if (!valid) continue;
            const double2 *M = G + ((int64_t)k * 6 + dir * 3 + dn) * 4;
            bool nz = false;
            for (int q = 0; q < 4; ++q) nz = nz || M[q].x != 0.0 || M[q].y != 0.0;
            if (nz && find_block(browptr, bcol, n, dir == 0 ? n + st : n - st) < 0) {
                bad[0] = 1;
                return;
            }
        }
    }
}

} // namespace

hipError_t h_block_count(const int32_t *rowptr, const int32_t *col, int64_t Z, int32_t *bcounts, hipStream_t s)
{
    hipLaunchKernelGGL(k_block_count, dim3((unsigned)((Z + 1 + 255) / 256)), dim3(256), 0, s, rowptr, col, Z, bcounts);
    return hipGetLastError();
}

hipError_t h_block_fill(const int32_t *rowptr, const int32_t *col, const double2 *val, int64_t Z,
                        const int32_t *browptr, int32_t *bcol, double2 *bval, hipStream_t s)
{
    hipLaunchKernelGGL(k_block_fill, dim3((unsigned)((Z + 255) / 256)), dim3(256), 0, s, rowptr, col, val, Z, browptr,
                       bcol, bval);
    return hipGetLastError();
}

hipError_t h_count(const DevSystem *d_sys, int64_t N, int32_t *counts, hipStream_t s)
{
    int64_t threads = N + 1;
    int blocks = (int)((threads + 255) / 256);
    hipLaunchKernelGGL(k_h_count, dim3(blocks), dim3(256), 0, s, d_sys, N, counts);
    return hipGetLastError();
}

hipError_t h_scan(const int32_t *counts, int32_t *rowptr, int64_t N, void *tmp, size_t &tmp_bytes, hipStream_t s)
{
    return hipcub::DeviceScan::ExclusiveSum(tmp, tmp_bytes, counts, rowptr, (int)(N + 1), s);
}

hipError_t h_fill(const DevSystem *d_sys, int64_t N, const int32_t *rowptr, int32_t *col, double2 *val,
                  hipStream_t s)
{
    int blocks = (int)((N + 255) / 256);
    hipLaunchKernelGGL(k_h_fill, dim3(blocks), dim3(256), 0, s, d_sys, N, rowptr, col, val);
    return hipGetLastError();
}

hipError_t h_block_fill_pattern(const int32_t *rowptr, const int32_t *col, const double2 *val, int64_t Z,
                                const int32_t *browptr, const int32_t *bcol, double2 *bval, int32_t *missing,
                                hipStream_t s)
{
    hipLaunchKernelGGL(k_block_fill_pattern, dim3((unsigned)((Z + 255) / 256)), dim3(256), 0, s, rowptr, col, val, Z,
                       browptr, bcol, bval, missing);
    return hipGetLastError();
}

hipError_t h_field_values(const double2 *v0, const double2 *zx, const double2 *zy, const double2 *zz,
                          const double *fields, int64_t count, int n_fields, double2 *out, hipStream_t s)
{
    const int64_t total = count * n_fields;
    const unsigned blocks = (unsigned)std::min<int64_t>((total + 255) / 256, 8192);
    hipLaunchKernelGGL(k_field_values, dim3(blocks), dim3(256), 0, s, v0, zx, zy, zz, fields, count, n_fields, out);
    return hipGetLastError();
}

// Neighbour masks per nuclear index n over all nuclei k: minus (digit d_k > 0),
// plus (d_k < mult_k - 1), one (d_k = 1); out[3n + 0, 1, 2].
__global__ void k_kron_masks(const KronDev *__restrict__ kd, int64_t Z, uint64_t *__restrict__ out)
{
    const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= Z) return;
    uint64_t mi = 0, pl = 0, on = 0;
    for (int k = 0; k < kd->nn; ++k) {
        const int mu = kd->mult[k];
        const int d = (int)((n / kd->stride[k]) % mu);
        if (d > 0) mi |= 1ull << k;
        if (d < mu - 1) pl |= 1ull << k;
        if (d == 1) on |= 1ull << k;
    }
    out[3 * n] = mi;
    out[3 * n + 1] = pl;
    out[3 * n + 2] = on;
}

hipError_t kron_masks(const KronDev *d_kron, int64_t Z, uint64_t *out, hipStream_t s)
{
    hipLaunchKernelGGL(k_kron_masks, dim3((unsigned)((Z + 255) / 256)), dim3(256), 0, s, d_kron, Z, out);
    return hipGetLastError();
}

hipError_t kron_tables(const KronDev *d_kron, const KronDev &h_kron, const int32_t *browptr, const int32_t *bcol,
                       const double2 *bval, int64_t nb, int nsets, int64_t Z, double2 *diag, double2 *gen,
                       int32_t *bad, hipStream_t s)
{
    const int64_t rows = Z * nsets;
    hipLaunchKernelGGL(k_kron_diag, dim3((unsigned)((rows + 255) / 256)), dim3(256), 0, s, browptr, bcol, bval, nb,
                       nsets, Z, diag, bad);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    const int ng = nsets * h_kron.nn * 6;
    if (ng > 0) {
        hipLaunchKernelGGL(k_kron_gen, dim3((unsigned)((ng + 255) / 256)), dim3(256), 0, s, d_kron, browptr, bcol,
                           bval, nb, nsets, gen);
        if ((e = hipGetLastError()) != hipSuccess) return e;
    }
    hipLaunchKernelGGL(k_kron_check, dim3((unsigned)((rows + 255) / 256)), dim3(256), 0, s, d_kron, browptr, bcol,
                       bval, nb, nsets, Z, gen, bad);
    return hipGetLastError();
}

} // namespace spc
