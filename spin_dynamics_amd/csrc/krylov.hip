// krylov.hip -- restarted Krylov propagation of M trace vectors at once.
//
// Replaces get_krylov / expm / convert_to_krylov and the inner step loop of
// qm_dynamics.f90:340-366.  All vectors share H, so the M trace vectors are the
// lane dimension: every basis vector is an [N][Mp] block in the device basis
// order (row = 4*nuc + e), lane r = trace vector r, one 1 KiB row segment per
// wavefront.  The sparse product gathers whole rows (one scalar-loaded H entry,
// 64 lanes of x), and the Gram-Schmidt dot products need no cross-lane
// reduction: each lane sums its own vector over the rows it visits, waves of a
// workgroup combine through LDS, workgroups through a deterministic tree pass.
//
// Orthogonalisation.  The reference uses modified Gram-Schmidt (i dependent
// reductions per column).  By default the basis is built a block of S columns
// at a time (arnoldi_sstep below: a chain of products, one dots pass and one
// forming pass per block).  With SPC_SSTEP=0 (and for m < 8) each column costs
// one sparse product and one streaming pass over the basis, with classical
// Gram-Schmidt plus a delayed re-orthogonalisation (the "DCGS2" form of
// Swirydowicz et al., 2020):
//   product  (column i): v = H u_i,  c = u_i^H v,  d = u_i^H u_i
//   coefficients:        a = Q^H u_i and s = <P_S Q, P_S u_i> come from the
//                        previous update pass; beta_i^2 = d - |a|^2; column i-1
//                        corrected by a; b = Q^H v from the Arnoldi relation
//                        H^H = H + iK, K = kT + (kS - kT) P_S (Haberkorn):
//                        b_k = sum_l conj(h_lk) a_l - i(kT a_k + (kS-kT) s_k);
//                        w_i = H q_i = (v - Q g - g_i q_i)/beta_i (g = Hbar a),
//                        t_i = Q^H w_i from a, b, c
//   update:              q_i = (u_i - Q a)/beta_i,  u_{i+1} = w_i - Q t_i,
//                        and the dots a, s of u_{i+1} against q_0..q_i
// The last update's dots correct column m-1.  The result is the reference's
// Hessenberg matrix A (qm_dynamics.f90:432-467) to rounding.
#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <array>
#include <utility>

#include "common.hpp"
#include "internal.hpp"

namespace spc {

namespace {

constexpr int kBlock = 256;
constexpr int kWaves = kBlock / kWave;

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

// Workgroup -> (row block, lane chunk).  Workgroups are dealt round-robin to
// the 8 XCDs; each XCD takes one contiguous eighth of the row blocks and walks
// it chunk after chunk, so the rows an XCD gathers from stay in its L2.
__device__ __forceinline__ void block_map(int bid, int nrb, int nch, int &rb, int &ch)
{
    if ((nrb & 7) == 0) {
        const int x = bid & 7, j = bid >> 3, per = nrb >> 3;
        ch = j / per;
        rb = x * per + j % per;
    } else {
        ch = bid / nrb;
        rb = bid % nrb;
    }
}

// Sums NACC per-lane accumulators over the 4 waves of the block into
// part[rb][j][r]; red must hold (kWaves-1)*8*kWave double2.
template <int NACC>
__device__ __forceinline__ void block_reduce(double2 (&acc)[NACC > 0 ? NACC : 1], double2 *red, double2 *part,
                                             int rb, int r, int lane, int wave, int Mp)
{
#pragma unroll
    for (int base = 0; base < NACC; base += 8) {
        if (wave > 0) {
#pragma unroll
            for (int j = 0; j < 8; ++j)
                if (base + j < NACC) red[((wave - 1) * 8 + j) * kWave + lane] = acc[base + j];
        }
        __syncthreads();
        if (wave == 0) {
#pragma unroll
            for (int j = 0; j < 8; ++j)
                if (base + j < NACC) {
                    double2 s = acc[base + j];
#pragma unroll
                    for (int w = 0; w < kWaves - 1; ++w) s = cadd(s, red[(w * 8 + j) * kWave + lane]);
                    part[((int64_t)rb * NACC + base + j) * Mp + r] = s;
                }
        }
        __syncthreads();
    }
}

// ------------------------------------------------------- sparse product
// The stored product row: (H u)_i - sig u_i (sig: the shift of a Newton chain
// product; every product kernel applies it last, in this form, so they stay
// bit-identical; sig = 0 stores H u unchanged).
__device__ __forceinline__ double2 shift_sub(double2 y, double sig, double2 x)
{
    return sig == 0.0 ? y : cz(fma(-sig, x.x, y.x), fma(-sig, x.y, y.y));
}

// v = H u with H in 4x4 blocks over nuclear indices (block row n = rows
// 4n..4n+3, block column nc = rows 4nc..4nc+3): a thread computes the 4 rows of
// its nuclear index, loading each neighbouring nuclear index's 4 rows once.
// Block values are wave-uniform (scalar loads).  Also c = u^H v, d = |u|^2.
__global__ __launch_bounds__(kBlock) void k_spmm(const int32_t *__restrict__ browptr, const int32_t *__restrict__ bcol,
                                                  const double2 *__restrict__ bvals, const double2 *__restrict__ u,
                                                  double2 *__restrict__ v, double2 *__restrict__ part, int64_t Z,
                                                  int Mp, int rpw, int nrb, int64_t bval_stride, int chunks_per_bval,
                                                  double sig)
{
    __shared__ double2 red[(kWaves - 1) * 8 * kWave];
    int rb, ch;
    block_map(blockIdx.x, nrb, Mp / kWave, rb, ch);
    // lane chunks may carry their own H (orientations batched in lanes)
    const double2 *__restrict__ bval = bvals + (int64_t)(ch / chunks_per_bval) * bval_stride;
    const int lane = threadIdx.x & (kWave - 1);
    const int wave = wave_uniform(threadIdx.x / kWave);
    const int r = ch * kWave + lane;
    double2 acc[2] = {cz(0, 0), cz(0, 0)};
    const int64_t n0 = ((int64_t)rb * kWaves + wave) * rpw;
    const int64_t n1 = n0 + rpw < Z ? n0 + rpw : Z;
    for (int64_t n = n0; n < n1; ++n) {
        const int b0 = browptr[n], b1 = browptr[n + 1];
        double2 y[4] = {cz(0, 0), cz(0, 0), cz(0, 0), cz(0, 0)};
        int b = b0;
        for (; b + 2 <= b1; b += 2) {
            double2 x[2][4];
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const int64_t cb = (int64_t)bcol[b + h] * 4 * Mp + r;
#pragma unroll
                for (int e = 0; e < 4; ++e) x[h][e] = u[cb + e * Mp];
            }
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const double2 *B = bval + (int64_t)(b + h) * 16;
#pragma unroll
                for (int e = 0; e < 4; ++e)
#pragma unroll
                    for (int f = 0; f < 4; ++f) y[e] = cfma(B[4 * e + f], x[h][f], y[e]);
            }
        }
        if (b < b1) {
            const int64_t cb = (int64_t)bcol[b] * 4 * Mp + r;
            double2 x[4];
#pragma unroll
            for (int e = 0; e < 4; ++e) x[e] = u[cb + e * Mp];
            const double2 *B = bval + (int64_t)b * 16;
#pragma unroll
            for (int e = 0; e < 4; ++e)
#pragma unroll
                for (int f = 0; f < 4; ++f) y[e] = cfma(B[4 * e + f], x[f], y[e]);
        }
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const int64_t off = (4 * n + e) * Mp + r;
            const double2 uu = u[off];
            y[e] = shift_sub(y[e], sig, uu);
            v[off] = y[e];
            acc[0] = cfma_conj(uu, y[e], acc[0]);
            acc[1].x += cabs2(uu);
        }
    }
    block_reduce<2>(acc, red, part, rb, r, lane, wave, Mp);
}

// The same product from the Kronecker form of H: per nuclear index n the 4x4
// diagonal block and, for each nucleus k, the flips to n -+ stride_k, each a
// 2x2 matrix on one electron's spin (half the FMAs of dense 4x4 blocks, and
// the tables are a few KB: scalar cache).  Columns are taken in ascending
// order, n - stride_0 < ... < n < ... < n + stride_0, and within a block in
// ascending order, so the sums equal the block kernel's bit for bit.
__device__ __forceinline__ void kron_apply(const double2 *__restrict__ M, int el, const double2 (&x)[4], double2 (&y)[4])
{
    if (el == 0) { // M on s1: y[2a+b] += M[a][0] x[b] + M[a][1] x[2+b]
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int b = 0; b < 2; ++b) {
                y[2 * a + b] = cfma(M[2 * a], x[b], y[2 * a + b]);
                y[2 * a + b] = cfma(M[2 * a + 1], x[2 + b], y[2 * a + b]);
            }
    } else { // M on s2: y[2a+b] += M[b][0] x[2a] + M[b][1] x[2a+1]
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int b = 0; b < 2; ++b) {
                y[2 * a + b] = cfma(M[2 * b], x[2 * a], y[2 * a + b]);
                y[2 * a + b] = cfma(M[2 * b + 1], x[2 * a + 1], y[2 * a + b]);
            }
    }
}

// Existing neighbours of nuclear index n as bit masks over the nuclei k <
// kmax: minus (digit d_k > 0), plus (d_k < mult_k - 1), one (d_k = 1; with
// the direction it fixes d_k for spin-1/2 and spin-1 nuclei).  n is uniform.
struct NbrMasks {
    uint64_t minus, plus, one;
};
__device__ __forceinline__ NbrMasks nbr_masks(const KronDev *__restrict__ kd, int n, int kmax)
{
    NbrMasks m{0, 0, 0};
    for (int k = 0; k < kmax; ++k) {
        const int mu = kd->mult[k];
        const int d = (n / kd->stride[k]) % mu;
        if (d > 0) m.minus |= 1ull << k;
        if (d < mu - 1) m.plus |= 1ull << k;
        if (d == 1) m.one |= 1ull << k;
    }
    return m;
}

// The masks of n over the nuclei k < kmax, from the precomputed table when there is one.
__device__ __forceinline__ NbrMasks masks_at(const uint64_t *__restrict__ kmask, const KronDev *__restrict__ kd, int n,
                                             int kmax)
{
    if (!kmask) return nbr_masks(kd, n, kmax);
    const uint64_t keep = kmax >= 64 ? ~0ull : (1ull << kmax) - 1;
    return NbrMasks{kmask[3 * (int64_t)n] & keep, kmask[3 * (int64_t)n + 1] & keep, kmask[3 * (int64_t)n + 2] & keep};
}

// The masks of n from the table (which the caller has).
__device__ __forceinline__ NbrMasks masks_table(const uint64_t *__restrict__ kmask, int n, int kmax)
{
    const uint64_t keep = kmax >= 64 ? ~0ull : (1ull << kmax) - 1;
    return NbrMasks{kmask[3 * (int64_t)n] & keep, kmask[3 * (int64_t)n + 1] & keep, kmask[3 * (int64_t)n + 2] & keep};
}

// y += the flips of the existing neighbours in `mask`, G at a time: minus
// neighbours n - stride_k in ascending k, plus neighbours n + stride_k in
// descending k (ascending columns either way).  The G rows of a batch are
// loaded before any is used and with no branch around the loads (a slot past
// the last neighbour reads row n itself and is not applied): loads under
// branches made the compiler wait for each before issuing the next.  base is
// the offset of this thread's element of row 4n; rows of index n + s are at
// base + s * M4.
// EL: the electron of every nucleus in `mask` (a compile-time choice of the
// 2x2 layout: no branch per flip).
template <int G, bool MINUS, int EL>
__device__ __forceinline__ void kron_walk(const KronDev *__restrict__ kd, const double2 *__restrict__ gen,
                                          const double2 *__restrict__ u, int64_t base, int64_t M4, int Mp,
                                          uint64_t mask, uint64_t one, double2 (&y)[4])
{
    while (mask) {
        double2 xn[G][4];
        int ks[G], sh[G];
#pragma unroll
        for (int j = 0; j < G; ++j) {
            ks[j] = -1;
            sh[j] = 0;
            if (mask) {
                const int k = MINUS ? __builtin_ctzll(mask) : 63 - __builtin_clzll(mask);
                mask &= ~(1ull << k);
                ks[j] = k;
                sh[j] = MINUS ? -kd->stride[k] : kd->stride[k];
            }
        }
#pragma unroll
        for (int j = 0; j < G; ++j) {
            const int64_t cb = base + (int64_t)sh[j] * M4;
#pragma unroll
            for (int f = 0; f < 4; ++f) xn[j][f] = u[cb + f * Mp];
        }
        __asm__ volatile("" ::: "memory");
#pragma unroll
        for (int j = 0; j < G; ++j)
            if (ks[j] >= 0) {
                const int k = ks[j];
                const int d = ((one >> k) & 1) ? 1 : (MINUS ? 2 : 0);
                kron_apply(gen + ((int64_t)k * 6 + (MINUS ? 3 : 0) + d) * 4, EL, xn[j], y);
            }
    }
}

// The nuclei of electron 1 precede those of electron 2 (lower k), so walking
// electron 1's minus neighbours before electron 2's (and electron 2's plus
// neighbours before electron 1's) keeps the ascending column order.
template <int G, bool MINUS>
__device__ __forceinline__ void kron_neighbours(const KronDev *__restrict__ kd, const double2 *__restrict__ gen,
                                                const double2 *__restrict__ u, int64_t base, int64_t M4, int Mp,
                                                uint64_t mask, uint64_t one, double2 (&y)[4])
{
    const uint64_t m2 = mask & kd->e2mask, m1 = mask & ~kd->e2mask;
    if (MINUS) {
        kron_walk<G, true, 0>(kd, gen, u, base, M4, Mp, m1, one, y);
        kron_walk<G, true, 1>(kd, gen, u, base, M4, Mp, m2, one, y);
    } else {
        kron_walk<G, false, 1>(kd, gen, u, base, M4, Mp, m2, one, y);
        kron_walk<G, false, 0>(kd, gen, u, base, M4, Mp, m1, one, y);
    }
}

template <int G>
__global__ __launch_bounds__(kBlock) void k_spmm_kron(const KronDev *__restrict__ kd,
                                                       const uint64_t *__restrict__ kmask,
                                                       const double2 *__restrict__ diags,
                                                       const double2 *__restrict__ gens, int64_t dstride,
                                                       int64_t gstride, int chunks_per_set,
                                                       const double2 *__restrict__ u, double2 *__restrict__ v,
                                                       double2 *__restrict__ part, int64_t Z, int Mp, int rpw, int nrb,
                                                       double sig)
{
    __shared__ double2 red[(kWaves - 1) * 8 * kWave];
    int rb, ch;
    block_map(blockIdx.x, nrb, Mp / kWave, rb, ch);
    const int lane = threadIdx.x & (kWave - 1);
    const int wave = wave_uniform(threadIdx.x / kWave);
    const int r = ch * kWave + lane;
    const int set = ch / chunks_per_set;
    const double2 *__restrict__ diag = diags + (int64_t)set * dstride;
    const double2 *__restrict__ gen = gens + (int64_t)set * gstride;
    const int nn = kd->nn;
    double2 acc[2] = {cz(0, 0), cz(0, 0)};
    const int64_t n0 = ((int64_t)rb * kWaves + wave) * rpw;
    const int64_t n1 = n0 + rpw < Z ? n0 + rpw : Z;
    for (int64_t n = n0; n < n1; ++n) {
        const int nu = (int)n;
        double2 x0[4];
#pragma unroll
        for (int f = 0; f < 4; ++f) x0[f] = u[(4 * n + f) * Mp + r];
        double2 y[4] = {cz(0, 0), cz(0, 0), cz(0, 0), cz(0, 0)};
        const NbrMasks nm = masks_at(kmask, kd, nu, nn);
        const int64_t base = 4 * n * Mp + r, M4 = 4 * (int64_t)Mp;
        kron_neighbours<G, true>(kd, gen, u, base, M4, Mp, nm.minus, nm.one, y);
        // the diagonal block
        {
            const double2 *D = diag + n * 16;
#pragma unroll
            for (int e = 0; e < 4; ++e)
#pragma unroll
                for (int f = 0; f < 4; ++f) y[e] = cfma(D[4 * e + f], x0[f], y[e]);
        }
        kron_neighbours<G, false>(kd, gen, u, base, M4, Mp, nm.plus, nm.one, y);
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const int64_t off = (4 * n + e) * Mp + r;
            y[e] = shift_sub(y[e], sig, x0[e]);
            v[off] = y[e];
            acc[0] = cfma_conj(x0[e], y[e], acc[0]);
            acc[1].x += cabs2(x0[e]);
        }
    }
    block_reduce<2>(acc, red, part, rb, r, lane, wave, Mp);
}

// The same Kronecker product on lane slices of W = 64 >> Q trace vectors, for
// an H whose last Q nuclei are spin-1/2 (strides 1, 2, .., 2^(Q-1)).  A wave
// takes the 2^Q consecutive nuclear indices n = (qi << Q) + g (g = lane group
// of W threads) for its W lanes: the digits of the other nuclei are the same
// for the whole wave (scalar), the last Q nuclei flip between the wave's own
// indices, and the wave's diagonal blocks (contiguous) are staged in LDS.
// When Mp / W is a multiple of 8, XCD x takes the slices 8s + x, s = 0, 1, ..,
// each over all rows: its L2 then holds the neighbours of 4 MiB / (64 W)
// nuclear indices instead of a quarter of that at W = 64.  Sums are taken in
// the order of k_spmm_kron (bit-identical products).
template <int G, int Q>
__global__ __launch_bounds__(kBlock) void k_spmm_kron_slices(const KronDev *__restrict__ kd,
                                                              const uint64_t *__restrict__ kmask,
                                                              const double2 *__restrict__ diags,
                                                              const double2 *__restrict__ gens, int64_t dstride,
                                                              int64_t gstride, int chunks_per_set,
                                                              const double2 *__restrict__ u, double2 *__restrict__ v,
                                                              double2 *__restrict__ part, int64_t Z, int Mp, int rpw,
                                                              int nrb, double sig)
{
    constexpr int W = kWave >> Q;
    constexpr int NG = 1 << Q;
    __shared__ double2 red[kWaves - 1][2][W];
    __shared__ double2 dl[kWaves][NG * 16];
    __shared__ double2 xl[kWaves][NG][4][W]; // the wave's own rows, for the in-wave flips
    const int nsl = Mp / W;
    int rb, sl;
    if ((nsl & 7) == 0) {
        const int x = blockIdx.x & 7, j = blockIdx.x >> 3;
        sl = (j / nrb) * 8 + x;
        rb = j % nrb;
    } else {
        sl = blockIdx.x / nrb;
        rb = blockIdx.x % nrb;
    }
    const int t = threadIdx.x & (kWave - 1);
    const int lane = t & (W - 1), g = t >> (6 - Q);
    const int wave = wave_uniform(threadIdx.x / kWave);
    const int r = sl * W + lane;
    const int set = (sl * W / kWave) / chunks_per_set;
    const double2 *__restrict__ diag = diags + (int64_t)set * dstride;
    const double2 *__restrict__ gen = gens + (int64_t)set * gstride;
    const int nn = kd->nn, nuni = nn - Q; // nuclei 0 .. nuni-1 have wave-uniform digits
    const int64_t M4 = 4 * (int64_t)Mp;
    const int64_t toff = (int64_t)g * M4 + r; // this thread's offset from row 4 (qi << Q)
    double2 acc[2] = {cz(0, 0), cz(0, 0)};
    const int64_t Zq = Z >> Q;
    const int64_t q0 = ((int64_t)rb * kWaves + wave) * rpw;
    const int64_t q1 = q0 + rpw < Zq ? q0 + rpw : Zq;
    for (int64_t qi = q0; qi < q1; ++qi) {
        const int nb = (int)(qi << Q);
        const int64_t rowb = (int64_t)nb * M4;
#pragma unroll
        for (int i = 0; i < (NG * 16 + kWave - 1) / kWave; ++i)
            if (i * kWave + t < NG * 16) dl[wave][i * kWave + t] = diag[(int64_t)nb * 16 + i * kWave + t];
        double2 x0[4];
#pragma unroll
        for (int f = 0; f < 4; ++f) x0[f] = u[rowb + toff + f * Mp];
        double2 y[4] = {cz(0, 0), cz(0, 0), cz(0, 0), cz(0, 0)};
        // neighbours n - stride_k of the uniform nuclei, k ascending
        const NbrMasks nm = masks_at(kmask, kd, nb, nuni);
        kron_neighbours<G, true>(kd, gen, u, rowb + toff, M4, Mp, nm.minus, nm.one, y);
        // the in-wave nuclei: every index has exactly one neighbour per nucleus,
        // g ^ 2^j (below it when bit j of g is set, above it otherwise), whose
        // rows the partner lane group holds: exchanged through LDS
#pragma unroll
        for (int f = 0; f < 4; ++f) xl[wave][g][f][lane] = x0[f];
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
        for (int j = Q - 1; j >= 0; --j) { // k = nn-1-j ascending
            const int k = nn - 1 - j;
            if ((g >> j) & 1) {
                double2 xp[4];
#pragma unroll
                for (int f = 0; f < 4; ++f) xp[f] = xl[wave][g ^ (1 << j)][f][lane];
                kron_apply(gen + ((int64_t)k * 6 + 4) * 4, kd->elec[k], xp, y);
            }
        }
        {
            // one block row at a time (4 LDS reads in registers, not 16)
            const double2 *D = &dl[wave][g * 16];
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                double2 de[4];
#pragma unroll
                for (int f = 0; f < 4; ++f) de[f] = D[4 * e + f];
#pragma unroll
                for (int f = 0; f < 4; ++f) y[e] = cfma(de[f], x0[f], y[e]);
                __asm__ volatile("" ::: "memory");
            }
        }
#pragma unroll
        for (int j = 0; j < Q; ++j) { // k = nn-1-j descending
            const int k = nn - 1 - j;
            if (!((g >> j) & 1)) {
                double2 xp[4];
#pragma unroll
                for (int f = 0; f < 4; ++f) xp[f] = xl[wave][g ^ (1 << j)][f][lane];
                kron_apply(gen + ((int64_t)k * 6) * 4, kd->elec[k], xp, y);
            }
        }
        // neighbours n + stride_k of the uniform nuclei, k descending
        kron_neighbours<G, false>(kd, gen, u, rowb + toff, M4, Mp, nm.plus, nm.one, y);
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            y[e] = shift_sub(y[e], sig, x0[e]);
            ntstore(&v[rowb + toff + e * Mp], y[e]); // streamed past L2, which keeps u's rows
            acc[0] = cfma_conj(x0[e], y[e], acc[0]);
            acc[1].x += cabs2(x0[e]);
        }
        // the next index's diagonal blocks overwrite dl only after every lane read it
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
    // sum over the wave's lane groups (a + b = b + a: every group holds the same sum), then the waves
#pragma unroll
    for (int s = W; s < kWave; s <<= 1)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            const double2 o = cz(__shfl_xor(acc[j].x, s), __shfl_xor(acc[j].y, s));
            acc[j] = cadd(acc[j], o);
        }
    if (wave > 0 && g == 0) {
        red[wave - 1][0][lane] = acc[0];
        red[wave - 1][1][lane] = acc[1];
    }
    __syncthreads();
    if (wave == 0 && g == 0) {
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            double2 s = acc[j];
#pragma unroll
            for (int w = 0; w < kWaves - 1; ++w) s = cadd(s, red[w][j][lane]);
            part[((int64_t)rb * 2 + j) * Mp + r] = s;
        }
    }
}

// The lane-slice product with NW waves per workgroup on NW consecutive quads
// (tiles of NW 2^Q consecutive nuclear indices): the BD = log2(NW) nuclei above
// the in-wave ones have wave-uniform digits (the wave's bits), and their flips
// pair the workgroup's own waves, so those neighbour rows also come from LDS
// instead of L2.  Needs the last Q + BD nuclei spin-1/2.  Same summation order
// as k_spmm_kron (bit-identical products).  rpg = tiles per workgroup.
// DOTS = false: v only (the block Arnoldi's chain products need no dots).
// MPC > 0: the lane stride Mp as a compile-time constant (it must equal the
// argument), so that the row offsets of a gathered index become immediate
// offsets of its loads and the neighbour shifts single scalar shifts.
// NE2 >= 0 (with the neighbour-mask table): the last NE2 of the in-tile nuclei (counted from the last) are
// electron 2's, the others electron 1's (electron 1's nuclei come first), so
// every in-tile flip has a compile-time 2x2 layout; and the in-wave flips run
// without divergence: a lane group whose flip falls on the other side of the
// diagonal block reads a row of zeros (adding M 0 changes at most the sign of
// a zero), where a branch on the lane group had made the compiler copy y at
// every merge; and the thread's own rows go to LDS as soon as they arrive and
// are read back there, so no registers hold them over the neighbour walks
// (76 VGPRs instead of 88: 6 waves per SIMD instead of 5).
template <int G, int Q, int NW, bool DOTS = true, int MPC = 0, int NE2 = -1>
__global__ __launch_bounds__(64 * NW) void k_spmm_kron_tiles(const KronDev *__restrict__ kd,
                                                            const uint64_t *__restrict__ kmask,
                                                            const double2 *__restrict__ diags,
                                                            const double2 *__restrict__ gens, int64_t dstride,
                                                            int64_t gstride, int chunks_per_set,
                                                            const double2 *__restrict__ u, double2 *__restrict__ v,
                                                            double2 *__restrict__ part, int64_t Z, int Mp_arg, int rpg,
                                                            int nrb, int lslow, double sig)
{
    const int Mp = MPC > 0 ? MPC : Mp_arg;
    constexpr int W = kWave >> Q;
    constexpr int NG = 1 << Q;
    constexpr int BD = NW == 16 ? 4 : NW == 8 ? 3 : NW == 4 ? 2 : NW == 2 ? 1 : 0;
    static_assert((1 << BD) == NW, "NW a power of two up to 16");
    __shared__ double2 red[NW - 1][2][W];
    __shared__ double2 dl[NW][NG * 16];
    __shared__ double2 xl[NW][NG][4][W]; // the workgroup's own rows, for the in-tile flips
    __shared__ double2 xz[4][W];         // zeros (NE2 >= 0)
    if (NE2 >= 0 && threadIdx.x < 4 * W) (&xz[0][0])[threadIdx.x] = cz(0, 0); // read after the tile's barrier
    // the electron of the in-tile nucleus k at position pos from the last
    auto elec_of = [&](int pos, int k) { return NE2 >= 0 ? (pos < NE2 ? 1 : 0) : kd->elec[k]; };
    const int nsl = Mp / W;
    int rb, sl;
    if ((nsl & 7) == 0) {
        const int x = blockIdx.x & 7, j = blockIdx.x >> 3;
        sl = (j / nrb) * 8 + x;
        rb = j % nrb;
    } else {
        sl = blockIdx.x / nrb;
        rb = blockIdx.x % nrb;
    }
    const int t = threadIdx.x & (kWave - 1);
    const int lane = t & (W - 1), g = t >> (6 - Q);
    const int wave = wave_uniform(threadIdx.x / kWave);
    const int r = sl * W + lane;
    const int set = (sl * W / kWave) / chunks_per_set;
    const double2 *__restrict__ diag = diags + (int64_t)set * dstride;
    const double2 *__restrict__ gen = gens + (int64_t)set * gstride;
    const int nn = kd->nn, nuni = nn - Q - BD; // nuclei 0 .. nuni-1: digits uniform over the tile
    const int64_t M4 = 4 * (int64_t)Mp;
    const int64_t toff = (int64_t)g * M4 + r;
    double2 acc[2] = {cz(0, 0), cz(0, 0)};
    const int64_t Zt = Z >> (Q + BD);
    const int64_t t0 = (int64_t)rb * rpg;
    const int64_t t1 = t0 + rpg < Zt ? t0 + rpg : Zt;
    const int64_t zs = Zt / lslow; // tiles per value of the slowest digits
    for (int64_t tj = t0; tj < t1; ++tj) {
        // work order: the slowest digits (lslow combinations) vary fastest
        const int64_t ti = (tj % lslow) * zs + tj / lslow;
        const int nb = (int)((((ti << BD) + wave)) << Q); // this wave's quad
        const int64_t rowb = (int64_t)nb * M4;
        // the neighbour masks (scalar loads) first: their latency then overlaps the rows' loads
        const NbrMasks nm = NE2 >= 0 ? masks_table(kmask, nb, nuni) : masks_at(kmask, kd, nb, nuni);
#pragma unroll
        for (int i = 0; i < (NG * 16 + kWave - 1) / kWave; ++i)
            if (i * kWave + t < NG * 16) dl[wave][i * kWave + t] = diag[(int64_t)nb * 16 + i * kWave + t];
        double2 x0[4];
#pragma unroll
        for (int f = 0; f < 4; ++f) x0[f] = u[rowb + toff + f * Mp];
        if constexpr (NE2 >= 0) // this thread's rows go to LDS at once and are read back there: no registers
#pragma unroll           // hold them over the neighbour walks
            for (int f = 0; f < 4; ++f) xl[wave][g][f][lane] = x0[f];
        double2 y[4] = {cz(0, 0), cz(0, 0), cz(0, 0), cz(0, 0)};
        // neighbours n - stride_k of the uniform nuclei, k ascending
        kron_neighbours<G, true>(kd, gen, u, rowb + toff, M4, Mp, nm.minus, nm.one, y);
        if constexpr (NE2 < 0)
#pragma unroll
            for (int f = 0; f < 4; ++f) xl[wave][g][f][lane] = x0[f];
        __syncthreads();
        // the tile nuclei (bit b of the wave: partner wave ^ 2^b), k ascending
#pragma unroll
        for (int b = BD - 1; b >= 0; --b) {
            const int k = nn - 1 - Q - b;
            if ((wave >> b) & 1) {
                double2 xp[4];
#pragma unroll
                for (int f = 0; f < 4; ++f) xp[f] = xl[wave ^ (1 << b)][g][f][lane];
                kron_apply(gen + ((int64_t)k * 6 + 4) * 4, elec_of(Q + b, k), xp, y);
            }
        }
        // the in-wave nuclei (bit j of g: partner lane group g ^ 2^j), k ascending
#pragma unroll
        for (int j = Q - 1; j >= 0; --j) {
            const int k = nn - 1 - j;
            if constexpr (NE2 >= 0) {
                const double2 *src = ((g >> j) & 1) ? &xl[wave][g ^ (1 << j)][0][lane] : &xz[0][lane];
                double2 xp[4];
#pragma unroll
                for (int f = 0; f < 4; ++f) xp[f] = src[f * W];
                kron_apply(gen + ((int64_t)k * 6 + 4) * 4, elec_of(j, k), xp, y);
            } else if ((g >> j) & 1) {
                double2 xp[4];
#pragma unroll
                for (int f = 0; f < 4; ++f) xp[f] = xl[wave][g ^ (1 << j)][f][lane];
                kron_apply(gen + ((int64_t)k * 6 + 4) * 4, kd->elec[k], xp, y);
            }
        }
        {
            const double2 *D = &dl[wave][g * 16];
            double2 xd[4];
#pragma unroll
            for (int f = 0; f < 4; ++f) xd[f] = NE2 >= 0 ? xl[wave][g][f][lane] : x0[f];
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                double2 de[4];
#pragma unroll
                for (int f = 0; f < 4; ++f) de[f] = D[4 * e + f];
#pragma unroll
                for (int f = 0; f < 4; ++f) y[e] = cfma(de[f], xd[f], y[e]);
                __asm__ volatile("" ::: "memory");
            }
        }
#pragma unroll
        for (int j = 0; j < Q; ++j) { // in-wave, k descending
            const int k = nn - 1 - j;
            if constexpr (NE2 >= 0) {
                const double2 *src = ((g >> j) & 1) ? &xz[0][lane] : &xl[wave][g ^ (1 << j)][0][lane];
                double2 xp[4];
#pragma unroll
                for (int f = 0; f < 4; ++f) xp[f] = src[f * W];
                kron_apply(gen + ((int64_t)k * 6) * 4, elec_of(j, k), xp, y);
            } else if (!((g >> j) & 1)) {
                double2 xp[4];
#pragma unroll
                for (int f = 0; f < 4; ++f) xp[f] = xl[wave][g ^ (1 << j)][f][lane];
                kron_apply(gen + ((int64_t)k * 6) * 4, kd->elec[k], xp, y);
            }
        }
#pragma unroll
        for (int b = 0; b < BD; ++b) { // tile nuclei, k descending
            const int k = nn - 1 - Q - b;
            if (!((wave >> b) & 1)) {
                double2 xp[4];
#pragma unroll
                for (int f = 0; f < 4; ++f) xp[f] = xl[wave ^ (1 << b)][g][f][lane];
                kron_apply(gen + ((int64_t)k * 6) * 4, elec_of(Q + b, k), xp, y);
            }
        }
        // neighbours n + stride_k of the uniform nuclei, k descending
        kron_neighbours<G, false>(kd, gen, u, rowb + toff, M4, Mp, nm.plus, nm.one, y);
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            y[e] = shift_sub(y[e], sig, NE2 >= 0 ? xl[wave][g][e][lane] : x0[e]);
            // the column form's v is streamed past L2, which keeps u's rows; a
            // chain product's output is the next product's input (cached:
            // 36.5 -> 32.7 us per 64-lane product at N = 36 864)
            if constexpr (DOTS)
                ntstore(&v[rowb + toff + e * Mp], y[e]);
            else
                v[rowb + toff + e * Mp] = y[e];
            if constexpr (DOTS) {
                acc[0] = cfma_conj(x0[e], y[e], acc[0]);
                acc[1].x += cabs2(x0[e]);
            }
        }
        __syncthreads(); // xl and dl are rewritten for the next tile
    }
    if constexpr (!DOTS) return;
#pragma unroll
    for (int s = W; s < kWave; s <<= 1)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            const double2 o = cz(__shfl_xor(acc[j].x, s), __shfl_xor(acc[j].y, s));
            acc[j] = cadd(acc[j], o);
        }
    if (wave > 0 && g == 0) {
        red[wave - 1][0][lane] = acc[0];
        red[wave - 1][1][lane] = acc[1];
    }
    __syncthreads();
    if (wave == 0 && g == 0) {
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            double2 s = acc[j];
#pragma unroll
            for (int w = 0; w < NW - 1; ++w) s = cadd(s, red[w][j][lane]);
            part[((int64_t)rb * 2 + j) * Mp + r] = s;
        }
    }
}

// Deterministic tree reduction: out[j][r] = sum_rb part[rb][j][r].  The
// partials were just written (they sit in L2/MALL), so the pass is bound by
// load latency: 16 waves per (j, 64 lanes), each with 4 loads in flight, take
// 64 rows per round trip; the order of the sums is fixed.
constexpr int kRedWaves = 16;
__global__ __launch_bounds__(kRedWaves *kWave) void k_reduce(const double2 *__restrict__ part,
                                                             double2 *__restrict__ out, int nacc, int nrb, int Mp)
{
    __shared__ double2 red[kRedWaves][kWave];
    const int j = blockIdx.x, ch = blockIdx.y;
    const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
    const int r = ch * kWave + lane;
    const int64_t step = (int64_t)nacc * Mp;
    const double2 *p = part + (int64_t)j * Mp + r;
    double2 s0 = cz(0, 0), s1 = cz(0, 0), s2 = cz(0, 0), s3 = cz(0, 0);
    int rb = w;
    for (; rb + 3 * kRedWaves < nrb; rb += 4 * kRedWaves) {
        const double2 a0 = p[rb * step], a1 = p[(rb + kRedWaves) * step];
        const double2 a2 = p[(rb + 2 * kRedWaves) * step], a3 = p[(rb + 3 * kRedWaves) * step];
        s0 = cadd(s0, a0);
        s1 = cadd(s1, a1);
        s2 = cadd(s2, a2);
        s3 = cadd(s3, a3);
    }
    for (; rb < nrb; rb += kRedWaves) s0 = cadd(s0, p[rb * step]);
    red[w][lane] = cadd(cadd(s0, s1), cadd(s2, s3));
    __syncthreads();
    if (w == 0) {
        double2 t = red[0][lane];
        for (int q = 1; q < kRedWaves; ++q) t = cadd(t, red[q][lane]);
        out[(int64_t)j * Mp + r] = t;
    }
}

// ------------------------------------------------------------ coefficients
// Column i of the Arnoldi process for kCoefLanes lanes per block, 32 threads
// per lane (thread q takes the index k = q; i < kMaxKrylov = 32).  The
// per-lane dots and the Hessenberg columns 0..i-2 are staged in LDS first
// (columns packed, rows <= col + 1), so the O(i^2) algebra reads LDS rather
// than chains of global loads.
//   d1: [c, d] of u_i from the product;  d2: [a_0..a_{i-1}, s_0..s_{i-1}]
//   from the previous update.  Writes coef = [cq_0..cq_{i-1}, cu_0..cu_{i-1},
//   cqu, cuu, cuv], tprev = t_i, Hessenberg column i-1, beta0 (i = 0).
constexpr int kCoefLanes = 8;
// (at most 4 waves per SIMD: the compiler would otherwise spill the staging
// registers to reach 8)
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(1, 4))) void k_coeffs(
    int i, int m, int M, int Mp, double kS, double kT,
                                                    const double2 *__restrict__ d1, const double2 *__restrict__ d2,
                                                    double2 *__restrict__ coef, double2 *__restrict__ tprev,
                                                    double2 *__restrict__ hess, double *__restrict__ beta0,
                                                    int32_t *__restrict__ flags)
{
    constexpr int NL = kCoefLanes, TP = kBlock / kCoefLanes;
    static_assert(TP >= kMaxKrylov, "one thread per index k");
    // packed Hessenberg columns of the NL lanes, lane-major: NL * coef_pack(m)
    // values of dynamic LDS (sized by the Krylov order at launch)
    extern __shared__ double2 shd[];
    const int pack = m * (m + 3) / 2;
    __shared__ double2 sa[NL][kMaxKrylov], sst[NL][kMaxKrylov], sb[NL][kMaxKrylov];
    __shared__ double red[TP][NL];
    __shared__ double2 redc[TP][NL];
    __shared__ double sbeta[NL];
    __shared__ double2 sgi[NL], ss[NL];
    const int l = threadIdx.x % NL, q = threadIdx.x / NL;
    const int r = blockIdx.x * NL + l;
    const bool active = r < M;
    const int64_t ld = Mp;
    const int lbase = l * pack; // this lane's packed columns in shd
    auto off = [](int col) { return col * (col + 3) / 2; };

    if (q < i) {
        sa[l][q] = d2[(int64_t)q * ld + r];
        sst[l][q] = d2[(int64_t)(i + q) * ld + r];
    }
    const int ne = i >= 2 ? off(i - 1) : 0; // packed entries of columns 0..i-2
    // six loads in flight per thread; every slot loads (clamped index): loads
    // under branches are issued one at a time
    constexpr int kBatch = 6;
    for (int u0 = 0; u0 * TP < ne; u0 += kBatch) {
        double2 hv[kBatch];
#pragma unroll
        for (int u = 0; u < kBatch; ++u) {
            const int e = min(q + TP * (u0 + u), ne - 1);
            int col = (int)((sqrt(9.0 + 8.0 * e) - 3.0) * 0.5); // then one rounding correction
            col += off(col + 1) <= e ? 1 : 0;
            col -= off(col) > e ? 1 : 0;
            hv[u] = hess[((int64_t)col * m + (e - off(col))) * ld + r];
        }
        // (slots past the end rewrite the last entry with its own value)
#pragma unroll
        for (int u = 0; u < kBatch; ++u) shd[lbase + min(q + TP * (u0 + u), ne - 1)] = hv[u];
    }
    __syncthreads();

    // beta_i^2 = d - |a|^2
    red[q][l] = active && q < i ? cabs2(sa[l][q]) : 0.0;
    __syncthreads();
    if (q == 0) {
        double beta2 = active ? d1[ld + r].x : 1.0;
        for (int w = 0; w < TP; ++w) beta2 -= red[w][l];
        double beta = 0.0;
        if (active) {
            if (!(beta2 > 0.0) || !isfinite(beta2))
                flags[0] = 1; // Krylov breakdown (the reference divides by zero here)
            else
                beta = sqrt(beta2);
            if (i == 0) beta0[r] = beta;
        }
        sbeta[l] = beta;
    }
    __syncthreads();
    const double beta = sbeta[l];
    const bool ok = active && beta > 0.0;
    const double ib = ok ? 1.0 / beta : 0.0;
    // finalise column i-1: h(k,i-1) = t_{i-1}[k] + a_k, h(i,i-1) = beta
    if (ok && i > 0) {
        if (q < i) {
            const double2 h = cadd(tprev[(int64_t)q * ld + r], sa[l][q]);
            hess[((int64_t)(i - 1) * m + q) * ld + r] = h;
            shd[lbase + off(i - 1) + q] = h;
        }
        if (q == 0) {
            hess[((int64_t)(i - 1) * m + i) * ld + r] = cz(beta, 0.0);
            shd[lbase + off(i - 1) + i] = cz(beta, 0.0);
        }
    }
    __syncthreads();
    // b_k and g_k for k = q; g_i kept for the last coefficient
    double2 abp = cz(0, 0);
    if (ok && q <= i) {
        const int k = q;
        double2 g = cz(0, 0); // g_k = sum_{j >= k-1} h(k,j) a_j
        for (int j = k > 0 ? k - 1 : 0; j < i; ++j) g = cfma(shd[lbase + off(j) + k], sa[l][j], g);
        if (k == i) {
            sgi[l] = g;
        } else {
            double2 b = cz(0, 0); // (H q_k)^H u_i
            const int lmax = k + 1 < i ? k + 1 : i - 1;
            for (int ll = 0; ll <= lmax; ++ll) b = cfma_conj(shd[lbase + off(k) + ll], sa[l][ll], b);
            if (k == i - 1) b.x += beta * beta; // h(i,i-1) q_i^H u_i = beta^2
            const double2 a = sa[l][k], sv = sst[l][k];
            // - i (kT a_k + (kS - kT) s_k)
            const double2 ka = cz(kT * a.x + (kS - kT) * sv.x, kT * a.y + (kS - kT) * sv.y);
            b = cadd(b, cz(ka.y, -ka.x));
            sb[l][k] = b;
            tprev[(int64_t)k * ld + r] = cscale(ib, csub(b, g));
            coef[(int64_t)k * ld + r] = cscale(-ib, a);
            abp = cfma_conj(a, b, abp);
        }
    }
    redc[q][l] = abp;
    __syncthreads();
    if (q == 0) {
        double2 sv = cz(0, 0);
        if (ok) {
            double2 ab = cz(0, 0);
            for (int w = 0; w < TP; ++w) ab = cadd(ab, redc[w][l]);
            const double2 c = d1[r];
            const double2 gi = sgi[l];
            const double2 ti = cscale(ib, csub(cscale(ib, csub(c, ab)), gi));
            tprev[(int64_t)i * ld + r] = ti;
            sv = cadd(cscale(ib, gi), ti);
        }
        ss[l] = sv;
        coef[(int64_t)(2 * i) * ld + r] = cz(ib, 0.0);
        coef[(int64_t)(2 * i + 1) * ld + r] = cscale(-ib, sv);
        coef[(int64_t)(2 * i + 2) * ld + r] = cz(ib, 0.0);
    }
    __syncthreads();
    const double2 sv = ss[l];
    if (q < i) { // cu_k = -(b_k - s a_k)/beta
        const int k = q;
        double2 cu = cz(0, 0);
        if (ok) cu = cscale(-ib, csub(sb[l][k], cmul(sv, sa[l][k])));
        coef[(int64_t)(i + k) * ld + r] = cu;
        if (!ok) coef[(int64_t)k * ld + r] = cz(0, 0);
    }
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

// ---------------------------------------------------------------- update
// NQ = i.  The 4 waves of a block take the 4 electron rows e = wave of one
// nuclear index n at a time (row 4n + e): q_i = cqu u + sum cq_k q_k and
// u_{i+1} = cuv v + cuu u + sum cu_k q_k.  The new rows go through LDS, and
// wave w then forms, for k = w mod 4, the dots of u_{i+1} with q_k over the 4
// rows (a) and of the singlet parts, rows ab - ba, (s).  LAST: column m-1 only
// needs a, and u_m is not stored.
// From this many basis vectors on, the update's dot phase exchanges rows
// through LDS instead of re-reading them (see k_update).
#ifndef SPC_UPDATE_LDS_DOTS
#define SPC_UPDATE_LDS_DOTS 13
#endif
constexpr int kUpdateLdsDots = SPC_UPDATE_LDS_DOTS;

template <int NQ, bool LAST>
__global__ __launch_bounds__(kBlock) void k_update(const double2 *__restrict__ Q, const double2 *__restrict__ Qre,
                                                    double2 *__restrict__ qout, double2 *__restrict__ u,
                                                    const double2 *__restrict__ v, const double2 *__restrict__ coef,
                                                    double2 *__restrict__ part, int64_t Z, int Mp, int npb, int nrb,
                                                    int64_t qs)
{
    constexpr int NC = 2 * NQ + 3;
    constexpr int NA = NQ + 1;
    constexpr int NACC = LAST ? NA : 2 * NA;
    constexpr int NK = (NA + kWaves - 1) / kWaves; // dot indices per wave
    __shared__ double2 sc[NC * kWave];
    __shared__ double2 sun[kWaves * kWave], sqi[kWaves * kWave];
    __shared__ double2 sx[NQ >= kUpdateLdsDots ? kWaves : 1][kWaves][kWave];
    int rb, ch;
    block_map(blockIdx.x, nrb, Mp / kWave, rb, ch);
    const int lane = threadIdx.x & (kWave - 1);
    const int wave = wave_uniform(threadIdx.x / kWave);
    const int r = ch * kWave + lane;
    const int64_t NMp = qs; // basis vector stride
    for (int idx = threadIdx.x; idx < NC * kWave; idx += kBlock)
        sc[idx] = coef[(int64_t)(idx / kWave) * Mp + ch * kWave + idx % kWave];
    __syncthreads();
    const double2 cqu = sc[(2 * NQ) * kWave + lane], cuu = sc[(2 * NQ + 1) * kWave + lane];
    const double2 cuv = sc[(2 * NQ + 2) * kWave + lane];
    double2 aa[NK], as[NK];
#pragma unroll
    for (int j = 0; j < NK; ++j) aa[j] = as[j] = cz(0, 0);

    const int64_t n0 = (int64_t)rb * npb;
    const int64_t n1 = n0 + npb < Z ? n0 + npb : Z;
    for (int64_t n = n0; n < n1; ++n) {
        // compiler barrier: coefficients are re-read from LDS per row instead
        // of being held in registers across the loop
        __asm__ volatile("" ::: "memory");
        const int64_t off = (4 * n + wave) * Mp + r;
        // u and v are read once and the new rows written once: non-temporal,
        // so L2 keeps the basis rows the dot phase below reads again
        const double2 uu = ntload(u + off), vv = ntload(v + off);
        // every load of the row is issued before the first use (the compiler
        // otherwise may trade the loads in flight for occupancy and wait on
        // each one in turn, which starves this bandwidth-bound pass)
        double2 xq[NQ > 0 ? NQ : 1];
#pragma unroll
        for (int k = 0; k < NQ; ++k) xq[k] = Q[k * NMp + off];
        __asm__ volatile("" ::: "memory");
        double2 qi = cmul(cqu, uu);
        double2 un = cfma(cuv, vv, cmul(cuu, uu));
#pragma unroll
        for (int k = 0; k < NQ; ++k) {
            qi = cfma(sc[k * kWave + lane], xq[k], qi);
            un = cfma(sc[(NQ + k) * kWave + lane], xq[k], un);
        }
        ntstore(qout + off, qi);
        if constexpr (!LAST) ntstore(u + off, un);
        sun[wave * kWave + lane] = un;
        sqi[wave * kWave + lane] = qi;
        __syncthreads();
        double2 w4[4];
#pragma unroll
        for (int e = 0; e < 4; ++e) w4[e] = sun[e * kWave + lane];
        const double2 dun = csub(w4[1], w4[2]);
        if constexpr (NQ >= kUpdateLdsDots) {
            // the dot phase's rows come from the waves that loaded them, through
            // LDS, four basis indices (one per wave) at a time: re-reading them
            // from L2 misses once 64 blocks per XCD hold NQ * 4 KiB each
#pragma unroll
            for (int c0 = 0; c0 < NA; c0 += kWaves) {
#pragma unroll
                for (int kk = 0; kk < kWaves; ++kk) {
                    if (c0 + kk < NQ)
                        sx[kk][wave][lane] = xq[c0 + kk < NQ ? c0 + kk : 0];
                    else if (c0 + kk == NQ)
                        sx[kk][wave][lane] = qi;
                }
                __syncthreads();
                const int k = c0 + wave;
                if (k <= NQ) {
                    double2 x[4];
#pragma unroll
                    for (int e = 0; e < 4; ++e) x[e] = sx[wave][e][lane];
                    double2 t = aa[c0 / kWaves];
#pragma unroll
                    for (int e = 0; e < 4; ++e) t = cfma_conj(x[e], w4[e], t);
                    aa[c0 / kWaves] = t;
                    if constexpr (!LAST)
                        as[c0 / kWaves] = cfma_conj(cscale(0.5, csub(x[1], x[2])), dun, as[c0 / kWaves]);
                }
                __syncthreads(); // sx (and, last, sun/sqi) are rewritten next
            }
        } else {
            const int64_t base = (4 * n) * Mp + r;
            double2 xd[NK][4]; // the dot phase's rows, all loads issued first as above
#pragma unroll
            for (int j = 0; j < NK; ++j) {
                const int k = wave + kWaves * j;
#pragma unroll
                for (int e = 0; e < 4; ++e)
                    xd[j][e] = k < NQ ? ntload(Qre + k * NMp + base + e * Mp) : sqi[e * kWave + lane];
            }
            __asm__ volatile("" ::: "memory");
#pragma unroll
            for (int j = 0; j < NK; ++j) {
                const int k = wave + kWaves * j;
                if (k > NQ) break;
                const double2 *x = xd[j];
                double2 t = aa[j];
#pragma unroll
                for (int e = 0; e < 4; ++e) t = cfma_conj(x[e], w4[e], t);
                aa[j] = t;
                if constexpr (!LAST) as[j] = cfma_conj(cscale(0.5, csub(x[1], x[2])), dun, as[j]);
            }
            __syncthreads(); // sun/sqi are rewritten for the next nuclear index
        }
    }
#pragma unroll
    for (int j = 0; j < NK; ++j) {
        const int k = wave + kWaves * j;
        if (k > NQ) break;
        part[((int64_t)rb * NACC + k) * Mp + r] = aa[j];
        if constexpr (!LAST) part[((int64_t)rb * NACC + NA + k) * Mp + r] = as[j];
    }
}

// ----------------------------------------------------------------- psi = Q c
// Per nuclear index: psi rows 4n + e = sum_k c_k q_k, and with OBS the
// populations of the new state: P_S = |psi_ab - psi_ba|^2/2, P_T0 =
// |psi_ab + psi_ba|^2/2, P_Tp = |psi_aa|^2, P_Tm = |psi_bb|^2 (equal to
// Re(c^H P_X c) of convert_to_krylov, qm_dynamics.f90:483-496 and 356-359).
// Vector 0 is read from q0: Q itself, or psi, updated in place (an element of
// q0 is read only by the thread that writes the same element of psi, before
// it writes it).
template <int MK, bool OBS>
__global__ __launch_bounds__(kBlock) void k_psi(const double2 *__restrict__ Q, const double2 *q0,
                                                 const double2 *__restrict__ cvec, double2 *__restrict__ psi,
                                                 double2 *__restrict__ part, int64_t Z, int Mp, int rpw, int nrb,
                                                 int64_t qs)
{
    __shared__ double2 red[(kWaves - 1) * 8 * kWave];
    int rb, ch;
    block_map(blockIdx.x, nrb, Mp / kWave, rb, ch);
    const int lane = threadIdx.x & (kWave - 1);
    const int wave = wave_uniform(threadIdx.x / kWave);
    const int r = ch * kWave + lane;
    const int64_t NMp = qs;
    double2 cc[MK];
#pragma unroll
    for (int k = 0; k < MK; ++k) cc[k] = cvec[(int64_t)k * Mp + r];
    double2 acc[2] = {cz(0, 0), cz(0, 0)};
    const int64_t n0 = ((int64_t)rb * kWaves + wave) * rpw;
    const int64_t n1 = n0 + rpw < Z ? n0 + rpw : Z;
    for (int64_t n = n0; n < n1; ++n) {
        double2 p[4];
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const int64_t off = (4 * n + e) * Mp + r;
            double2 s = cfma(q0[off], cc[0], cz(0, 0));
#pragma unroll
            for (int k = 1; k < MK; ++k) s = cfma(Q[k * NMp + off], cc[k], s);
            p[e] = s;
            psi[off] = s;
        }
        if constexpr (OBS) {
            acc[0].x += 0.5 * cabs2(csub(p[1], p[2])); // P_S
            acc[0].y += 0.5 * cabs2(cadd(p[1], p[2])); // P_T0
            acc[1].x += cabs2(p[0]);                   // P_Tp
            acc[1].y += cabs2(p[3]);                   // P_Tm
        }
    }
    if constexpr (OBS) block_reduce<2>(acc, red, part, rb, r, lane, wave, Mp);
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
        if (tid < kWave) { // the first row of largest |re| + |im| in column k at or below k (m <= 64)
            double v = tid >= k && tid < m ? fabs(D[tid + m * k].x) + fabs(D[tid + m * k].y) : -1.0;
            int p = tid;
#pragma unroll
            for (int sft = kWave / 2; sft >= 1; sft >>= 1) {
                const double ov = __shfl_xor(v, sft);
                const int op = __shfl_xor(p, sft);
                if (ov > v || (ov == v && op < p)) {
                    v = ov;
                    p = op;
                }
            }
            if (tid == 0) ipiv[k] = p;
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
    { // every load in flight before the first store (m*m <= 4 blockDim.x)
        double2 hv[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int64_t idx = min((int)threadIdx.x + kBlock * u, mm - 1);
            hv[u] = in[mode == 0 ? idx * Mp + r : idx];
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int idx = threadIdx.x + kBlock * u;
            if (idx < mm) X[idx] = mode == 0 ? cz(dt * hv[u].y, -dt * hv[u].x) : hv[u]; // (0,-1)*dt*h
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

// C[o][k][r] = scale_o beta_r (exp(-i tau_o A_r))_{k,0}: the Krylov coefficients
// of exp(-i tau_o H) psi_r = Q_r exp(-i tau_o A_r) e_1 beta_r for several times
// tau_o at once (tau = 0 gives beta e_1).  Grid (Mp, ntau).
__global__ __launch_bounds__(kBlock) void k_expm_coeffs(int m, int Mp, int M, TauSet ts,
                                                         const double2 *__restrict__ hess,
                                                         const double *__restrict__ beta0, double2 *__restrict__ C)
{
    extern __shared__ double2 lds[];
    const int r = blockIdx.x, o = blockIdx.y;
    const int mm = m * m;
    double2 *X = lds, *P = X + mm, *E = P + mm, *D = E + mm;
    double2 *out = C + (int64_t)o * m * Mp;
    if (r >= M) {
        for (int k = threadIdx.x; k < m; k += blockDim.x) out[(int64_t)k * Mp + r] = cz(0, 0);
        return;
    }
    const double tau = ts.tau[o];
    for (int idx = threadIdx.x; idx < mm; idx += blockDim.x) {
        const double2 h = hess[(int64_t)idx * Mp + r];
        X[idx] = cz(tau * h.y, -tau * h.x); // (0,-1)*tau*h
    }
    __syncthreads();
    expm_block(m, X, P, E, D);
    const double f = ts.scale[o] * beta0[r];
    for (int k = threadIdx.x; k < m; k += blockDim.x) out[(int64_t)k * Mp + r] = cscale(f, E[k]);
}

} // namespace

size_t expm_lds_bytes(int m) { return 4 * (size_t)m * m * sizeof(double2); }

namespace {

// Position of G(x, y), x <= y < nd, in the packed upper triangle (row by row).
__host__ __device__ __forceinline__ int packed_index(int x, int y, int nd) { return x * nd - x * (x - 1) / 2 + y - x; }

// --------------------------------------------------------------- step loop
// qm_dynamics.f90:348-364 for each lane when a basis serves one step (the
// populations then come from the psi update): one step from c = beta e_1,
// i.e. c = beta E(:,0) (the general loop's sum adds only E(i,j) * 0 beyond
// j = 0: the same bits), with the loop's bookkeeping; a lane whose tol rule
// asks for a second step raises flag 1.
__global__ void k_step(int m, int M, int Mp, double dt, double t_end, double tol, const double2 *__restrict__ E,
                       const double *__restrict__ beta0, double2 *__restrict__ cout, double *__restrict__ lane_t,
                       int32_t *__restrict__ lane_k, int64_t n_out, int32_t *__restrict__ flags)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= M) return;
    double t = lane_t[r];
    int64_t k = lane_k[r];
    const double b0 = beta0[r];
    if (t < t_end) {
        if (k + 1 >= n_out) {
            for (int i = 0; i < m; ++i) cout[(int64_t)i * Mp + r] = cfma(E[(int64_t)i * Mp + r], cz(b0, 0.0), cz(0, 0));
            flags[3] = 1;
            lane_k[r] = (int32_t)(k + 1);
            return;
        }
        double c2 = 0.0, cm2 = 0.0;
        for (int i = 0; i < m; ++i) {
            const double2 ci = cfma(E[(int64_t)i * Mp + r], cz(b0, 0.0), cz(0, 0));
            cout[(int64_t)i * Mp + r] = ci;
            c2 += cabs2(ci);
            if (i == m - 1) cm2 = cabs2(ci);
        }
        k = k + 1;
        t = t + dt;
        if (cm2 <= tol * c2 && t < t_end) flags[1] = 1;
    } else {
        for (int j = 0; j < m; ++j) cout[(int64_t)j * Mp + r] = j == 0 ? cz(b0, 0.0) : cz(0, 0);
    }
    lane_t[r] = t;
    lane_k[r] = (int32_t)k;
}

// The same loop with projector Grams (tol > 0: several steps per basis), one
// wave per lane.  The populations are Re(d^H G_X d) for four Hermitian G_X
// (X-th at G + X*gxs; entry (a, b), a <= b < m, at packed_index(a, b, nd)
// when nd > 0, else at a*m + b of a full matrix), d = R^-1 c when Ri is given
// (a Newton block: G_X = Y^H P_X Y over its chain vectors, coef layout
// Ri[(x*m + j)*Mp + r] = (R^-1)[x][j]), else d = c (G_X over an orthonormal
// basis).  E(i,j) at E[(j*m + i)*Mp + r].  E and R^-1 sit in LDS (column by
// column: thread q reads row q conflict-free), each thread keeps its share of
// the triangle entries of the four G_X in registers, and the sums over the
// lane's entries are wave reductions.
constexpr int kStepEntries = (kMaxKrylov * (kMaxKrylov + 1) / 2 + kWave - 1) / kWave;
__global__ __launch_bounds__(kWave) void k_step_proj(int m, int M, int Mp, double dt, double t_end, double tol,
                                                     const double2 *__restrict__ E, const double *__restrict__ beta0,
                                                     double2 *__restrict__ cout, double *__restrict__ lane_t,
                                                     int32_t *__restrict__ lane_k, double *__restrict__ obs,
                                                     int64_t n_out, const double2 *__restrict__ G, int64_t gxs, int nd,
                                                     const double2 *__restrict__ Ri, int32_t *__restrict__ flags)
{
    constexpr int MX = kMaxKrylov, NE = kStepEntries;
    __shared__ double2 Es[MX * MX], Rs[MX * MX], cs[2][MX], ds[MX];
    const int r = blockIdx.x, q = threadIdx.x;
    if (r >= M) return; // uniform over the workgroup
    { // every load in flight before the first store (clamped indices, no branch around the loads)
        constexpr int LPT = MX * MX / kWave;
        const int mm = m * m;
        double2 te[LPT], tr[LPT];
#pragma unroll
        for (int u = 0; u < LPT; ++u) {
            const int64_t idx = min(q + kWave * u, mm - 1);
            te[u] = E[idx * Mp + r]; // (j*m + i)
            tr[u] = Ri ? Ri[idx * Mp + r] : cz(0, 0);
        }
#pragma unroll
        for (int u = 0; u < LPT; ++u) {
            const int idx = q + kWave * u;
            if (idx < mm) {
                Es[idx] = te[u];
                // (R^-1)[i][j] from coef[(i*m + j)], stored at j*m + i
                const int i = idx / m, j = idx % m;
                if (Ri) Rs[j * m + i] = j >= i ? tr[u] : cz(0, 0);
            }
        }
    }
    // this thread's triangle entries e = q + 64k (row-major over the leading m x m)
    const int np = m * (m + 1) / 2;
    int ea[NE], eb[NE];
    double wt[NE];
    double2 g[4][NE];
#pragma unroll
    for (int k = 0; k < NE; ++k) {
        const int e = q + kWave * k;
        int a = 0, rem = e < np ? e : 0;
        while (rem >= m - a) {
            rem -= m - a;
            ++a;
        }
        const int b = a + rem;
        ea[k] = a;
        eb[k] = b;
        wt[k] = e < np ? (a == b ? 1.0 : 2.0) : 0.0;
        const int64_t off = nd > 0 ? (int64_t)packed_index(a, b, nd) : (int64_t)a * m + b;
#pragma unroll
        for (int x = 0; x < 4; ++x) g[x][k] = G[(x * gxs) + off * Mp + r]; // (0, 0) when e >= np: weight 0
    }
    const double b0 = beta0[r];
    if (q < m) cs[0][q] = q == 0 ? cz(b0, 0.0) : cz(0, 0);
    double t = lane_t[r];
    int64_t k = lane_k[r];
    __syncthreads();
    double c2 = 1.0, cm2 = 0.0;
    int cur = 0;
    while (cm2 <= tol * c2 && t < t_end) {
        double2 cq = cz(0, 0);
        if (q < m) {
            double2 c1 = cz(0, 0); // two chains: half the FMA latency
            int j = 0;
#pragma unroll 4
            for (; j + 1 < m; j += 2) {
                cq = cfma(Es[j * m + q], cs[cur][j], cq);
                c1 = cfma(Es[(j + 1) * m + q], cs[cur][j + 1], c1);
            }
            if (j < m) cq = cfma(Es[j * m + q], cs[cur][j], cq);
            cq = cadd(cq, c1);
            cs[cur ^ 1][q] = cq;
        }
        cur ^= 1;
        k = k + 1;
        if (k >= n_out) {
            if (q == 0) flags[3] = 1;
            break;
        }
        __syncthreads();
        if (Ri) {
            if (q < m) {
                double2 dq = cz(0, 0), d1 = cz(0, 0);
                int j = q;
#pragma unroll 4
                for (; j + 1 < m; j += 2) {
                    dq = cfma(Rs[j * m + q], cs[cur][j], dq);
                    d1 = cfma(Rs[(j + 1) * m + q], cs[cur][j + 1], d1);
                }
                if (j < m) dq = cfma(Rs[j * m + q], cs[cur][j], dq);
                ds[q] = cadd(dq, d1);
            }
            __syncthreads();
        }
        const double2 *dv = Ri ? ds : cs[cur];
        double p[5] = {0.0, 0.0, 0.0, 0.0, q < m ? cabs2(cq) : 0.0};
#pragma unroll
        for (int e = 0; e < NE; ++e) {
            const double2 da = dv[ea[e]], db = dv[eb[e]];
#pragma unroll
            for (int x = 0; x < 4; ++x) {
                const double2 gb = cmul(g[x][e], db);
                p[x] += wt[e] * (da.x * gb.x + da.y * gb.y); // Re(conj(d_a) G_ab d_b)
            }
        }
#pragma unroll
        for (int sft = kWave / 2; sft >= 1; sft >>= 1)
#pragma unroll
            for (int x = 0; x < 5; ++x) p[x] += __shfl_xor(p[x], sft);
        if (q == 0)
            for (int x = 0; x < 4; ++x) obs[(k * 4 + x) * Mp + r] = p[x];
        c2 = p[4];
        cm2 = cabs2(cs[cur][m - 1]);
        t = t + dt;
        __syncthreads(); // the next step rewrites ds and cs[cur ^ 1]
    }
    if (q < m) cout[(int64_t)q * Mp + r] = cs[cur][q];
    if (q == 0) {
        lane_t[r] = t;
        lane_k[r] = (int32_t)k;
    }
}

// ---------------------------------------------------------- dispatch tables
using UpdateFn = void (*)(const double2 *, const double2 *, double2 *, double2 *, const double2 *, const double2 *,
                          double2 *, int64_t, int, int, int, int64_t);
using PsiFn = void (*)(const double2 *, const double2 *, const double2 *, double2 *, double2 *, int64_t, int, int,
                       int, int64_t);

template <size_t... I>
constexpr auto make_update(std::index_sequence<I...>, std::false_type)
{
    return std::array<UpdateFn, sizeof...(I)>{&k_update<(int)I, false>...};
}
template <size_t... I>
constexpr auto make_update(std::index_sequence<I...>, std::true_type)
{
    return std::array<UpdateFn, sizeof...(I)>{&k_update<(int)I, true>...};
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

const auto kUpdate = make_update(std::make_index_sequence<kMaxKrylov>{}, std::false_type{});
const auto kUpdateLast = make_update(std::make_index_sequence<kMaxKrylov>{}, std::true_type{});
const auto kPsiObs = make_psi(std::make_index_sequence<kMaxKrylov>{}, std::true_type{});
const auto kPsiPlain = make_psi(std::make_index_sequence<kMaxKrylov>{}, std::false_type{});

inline int nrb_for(int64_t rows, int rpw) { return (int)ceil_div(rows, (int64_t)kWaves * rpw); }

hipError_t reduce(const KrylovWork &w, int nacc, int nrb, double2 *out, hipStream_t s)
{
    if (nacc == 0) return hipSuccess;
    ProfScope ps(w, PF_REDUCE, 16.0 * (double)nacc * w.Mp * (nrb + 1), s);
    hipLaunchKernelGGL(k_reduce, dim3(nacc, w.Mp / kWave), dim3(kRedWaves * kWave), 0, s, w.part, out, nacc, nrb,
                       w.Mp);
    return hipGetLastError();
}

using KronFn = void (*)(const KronDev *, const uint64_t *, const double2 *, const double2 *, int64_t, int64_t, int,
                        const double2 *, double2 *, double2 *, int64_t, int, int, int, double);

// Lane-slice depth of the Kronecker product: 2^Q nuclear indices per wave on
// 64 >> Q lanes.  Default 2 where the system allows it (N = 36864, M = 256 on
// MI355X: 160 us per product at Q = 2, 167 at 1, 202 at 3, 212 at 0), 3 for
// a chunk of 64 lanes (the 4 x 64 lane split of M = 256 steps in 14.7 ms at
// Q = 3, 15.0 at Q = 2); SPC_KRON_Q overrides it (read at every launch),
// capped by the system.
int kron_slices(const KrylovWork &w)
{
    if (!w.kron) return 0;
    int q = w.Mp <= kWave ? 3 : 2;
    if (const char *e = std::getenv("SPC_KRON_Q")) q = std::min(3, std::max(0, std::atoi(e))); // kernels for 0..3
    return std::min(q, w.kq);
}

// Waves per workgroup of the tiled product (k_spmm_kron_tiles; 0: the plain
// lane slices).  At Q = 2: 8, 16 from 24576 nuclear indices on (MI355X: 147 us
// per product at N = 36864, M = 256 against 162 with 16 and 158 without tiles;
// 423 us at N = 147456, M = 128 against 447 with 8).  At Q = 3 (64-lane
// chunks): 4 (17.0 / 51.6 / 189 us at N = 12288 / 49152 / 147456 against
// 18.6 / 55.3 / 188 with 8).  Halved while the system has fewer than
// Q + log2(NW) trailing spin-1/2 nuclei; SPC_KRON_TILE = 0, 4, 8 or 16
// overrides it (read at every launch).
int kron_tile(const KrylovWork &w)
{
    const int q = kron_slices(w);
    if (q != 2 && q != 3) return 0;
    int nw = q == 3 ? 4 : w.Z >= 24576 ? 16 : 8;
    if (const char *e = std::getenv("SPC_KRON_TILE")) nw = std::atoi(e);
    if (nw != 4 && nw != 8 && nw != 16) return 0;
    for (; nw >= 4; nw /= 2)
        if (w.kq >= q + (nw == 16 ? 4 : nw == 8 ? 3 : 2)) return nw;
    return 0;
}

// Work order of the tiles: the slowest nucleus's digits vary fastest when its
// flip partners lie further apart than an XCD's L2 holds of the XCD's lane
// slice (stride x 4 rows x W lanes x 16 B > 4 MiB), so that both partners of
// a flip are read while in L2; returns the number of its digits (1: natural
// order).  N = 147456 (stride 12288): 155 against 184 us per 64-lane product;
// smaller strides gain nothing.  SPC_KRON_ORDER=0 keeps the natural order.
int tile_order(const KrylovWork &w, int fast, int lanes)
{
    if (const char *e = std::getenv("SPC_KRON_ORDER"))
        if (std::atoi(e) == 0) return 1;
    const int l = w.slow_mult;
    if (l < 2 || w.slow_stride * 4 * lanes * (int64_t)sizeof(double2) <= (int64_t)4 << 20) return 1;
    return ((w.Z >> fast) % l) == 0 ? l : 1;
}

// Tiles per workgroup of the tiled product: about 4096 workgroups in all
// (SPC_KRON_RPG sets it; read at every launch).
inline int tile_rpg(const KrylovWork &w, int nw)
{
    if (const char *e = std::getenv("SPC_KRON_RPG")) return std::max(1, std::atoi(e));
    const int q = kron_slices(w);
    const int64_t tiles = w.Z >> (q + (nw == 16 ? 4 : nw == 8 ? 3 : 2));
    const int64_t nsl = (w.Mp << q) / kWave;
    return (int)std::max<int64_t>(1, ceil_div(tiles * nsl, 4096 / std::max(1, nw / 4)));
}

// Row blocks of the product's partial sums (the reduce after it needs them).
int spmm_nrb(const KrylovWork &w)
{
    const int q = kron_slices(w);
    if (const int nw = kron_tile(w)) {
        const int64_t tiles = w.Z >> (q + (nw == 16 ? 4 : nw == 8 ? 3 : 2));
        return (int)ceil_div(tiles, tile_rpg(w, nw));
    }
    return nrb_for(w.Z >> q, w.rpw1);
}

hipError_t spmm_launch(KrylovWork &w, hipStream_t s)
{
    const int nrb1 = nrb_for(w.Z, w.rpw1);
    ProfScope ps(w, PF_PASS1, 2 * vec_bytes(w) + csr_bytes(w), s);
    if (w.kron) {
        static int group = -1;
        if (group < 0) {
            const char *e = std::getenv("SPC_KRON_GROUP");
            group = e ? std::min(4, std::max(2, std::atoi(e))) : 2; // 2: fastest at Q = 2 (160 us vs 165 for 3)
        }
        const int q = kron_slices(w);
        if (const int nw = kron_tile(w)) {
            const int nrb = spmm_nrb(w), nsl = (w.Mp << q) / kWave, rpg = tile_rpg(w, nw);
            using TileFn = void (*)(const KronDev *, const uint64_t *, const double2 *, const double2 *, int64_t,
                                    int64_t, int, const double2 *, double2 *, double2 *, int64_t, int, int, int, int,
                                    double);
            static const TileFn tfn[2][2][3][2] = {
                {{{k_spmm_kron_tiles<2, 2, 4>, k_spmm_kron_tiles<3, 2, 4>},
                  {k_spmm_kron_tiles<2, 2, 8>, k_spmm_kron_tiles<3, 2, 8>},
                  {k_spmm_kron_tiles<2, 2, 16>, k_spmm_kron_tiles<3, 2, 16>}},
                 {{k_spmm_kron_tiles<2, 3, 4>, k_spmm_kron_tiles<3, 3, 4>},
                  {k_spmm_kron_tiles<2, 3, 8>, k_spmm_kron_tiles<3, 3, 8>},
                  {k_spmm_kron_tiles<2, 3, 16>, k_spmm_kron_tiles<3, 3, 16>}}},
                {{{k_spmm_kron_tiles<2, 2, 4, false>, k_spmm_kron_tiles<3, 2, 4, false>},
                  {k_spmm_kron_tiles<2, 2, 8, false>, k_spmm_kron_tiles<3, 2, 8, false>},
                  {k_spmm_kron_tiles<2, 2, 16, false>, k_spmm_kron_tiles<3, 2, 16, false>}},
                 {{k_spmm_kron_tiles<2, 3, 4, false>, k_spmm_kron_tiles<3, 3, 4, false>},
                  {k_spmm_kron_tiles<2, 3, 8, false>, k_spmm_kron_tiles<3, 3, 8, false>},
                  {k_spmm_kron_tiles<2, 3, 16, false>, k_spmm_kron_tiles<3, 3, 16, false>}}}};
            const int ti = nw == 4 ? 0 : nw == 8 ? 1 : 2;
            TileFn fn = tfn[w.no_dots ? 1 : 0][q - 2][ti][std::min(group, 3) - 2];
            if (w.no_dots && w.Mp == 64 && q == 3 && nw == 4 && group == 2) {
                // the chain product at 64 lanes; with the electrons of the 5 in-tile nuclei known (kne2p)
                static const TileFn c64[7] = {k_spmm_kron_tiles<2, 3, 4, false, 64>,
                                              k_spmm_kron_tiles<2, 3, 4, false, 64, 0>,
                                              k_spmm_kron_tiles<2, 3, 4, false, 64, 1>,
                                              k_spmm_kron_tiles<2, 3, 4, false, 64, 2>,
                                              k_spmm_kron_tiles<2, 3, 4, false, 64, 3>,
                                              k_spmm_kron_tiles<2, 3, 4, false, 64, 4>,
                                              k_spmm_kron_tiles<2, 3, 4, false, 64, 5>};
                fn = c64[w.kne2p > 0 && w.kmask ? 1 + std::min(w.kne2p - 1, 5) : 0];
            }
            hipLaunchKernelGGL(fn, dim3(nrb * nsl), dim3(64 * nw),
                               0, s, w.kron,
                               w.kmask, w.kdiag, w.kgen,
                               w.kdiag_stride, w.kgen_stride, std::max(w.chunks_per_bval, 1), w.u, w.v, w.part, w.Z,
                               w.Mp, rpg, nrb, tile_order(w, q + (nw == 16 ? 4 : nw == 8 ? 3 : 2), kWave >> q),
                               w.sigma);
            return hipGetLastError();
        }
        if (q > 0) {
            static const KronFn fns[3][3] = {
                {k_spmm_kron_slices<2, 1>, k_spmm_kron_slices<3, 1>, k_spmm_kron_slices<4, 1>},
                {k_spmm_kron_slices<2, 2>, k_spmm_kron_slices<3, 2>, k_spmm_kron_slices<4, 2>},
                {k_spmm_kron_slices<2, 3>, k_spmm_kron_slices<3, 3>, k_spmm_kron_slices<4, 3>}};
            const int nrb = spmm_nrb(w), nsl = (w.Mp << q) / kWave;
            hipLaunchKernelGGL(fns[q - 1][group - 2], dim3(nrb * nsl), dim3(kBlock), 0, s, w.kron, w.kmask, w.kdiag, w.kgen,
                               w.kdiag_stride, w.kgen_stride, std::max(w.chunks_per_bval, 1), w.u, w.v, w.part, w.Z,
                               w.Mp, w.rpw1, nrb, w.sigma);
            return hipGetLastError();
        }
        auto fn = group == 2 ? k_spmm_kron<2> : group == 3 ? k_spmm_kron<3> : k_spmm_kron<4>;
        hipLaunchKernelGGL(fn, dim3(nrb1 * (w.Mp / kWave)), dim3(kBlock), 0, s, w.kron, w.kmask, w.kdiag, w.kgen,
                           w.kdiag_stride, w.kgen_stride, std::max(w.chunks_per_bval, 1), w.u, w.v, w.part, w.Z,
                           w.Mp, w.rpw1, nrb1, w.sigma);
        return hipGetLastError();
    }
    hipLaunchKernelGGL(k_spmm, dim3(nrb1 * (w.Mp / kWave)), dim3(kBlock), 0, s, w.browptr, w.bcol, w.bval, w.u, w.v,
                       w.part, w.Z, w.Mp, w.rpw1, nrb1, w.bval_stride, std::max(w.chunks_per_bval, 1), w.sigma);
    return hipGetLastError();
}


// ------------------------------------------------------- s-step Arnoldi
// The same Krylov basis and Hessenberg matrix built a block of S columns at a
// time (Hoemmen 2010, ch. 3; block classical Gram-Schmidt with the Pythagorean
// Gram matrix and Cholesky QR: BCGS-PIP, Carson, Lund, Rozloznik and Thomas,
// 2021).  With K orthonormal vectors q_0..q_{K-1}, the chain
// y_t = H^{t+1} q_{K-1} (t < n, n <= S) takes n sparse products; one pass over
// [Q_K, Y] gives the dots C = Q_K^H Y and G = Y^H Y; per lane
// R^H R = G - C^H C (Cholesky, columns scaled to unit diagonal first); one
// more pass writes the new vectors q_{K+t} = (Y R^-1 - Q_K C R^-1)_t in place
// of Y.  The first block takes Y = [psi, H psi, .., H^(S-1) psi] (K = 0,
// beta_0 = R_00).  The Hessenberg columns follow from H y_{t-1} = y_t written
// in the new basis (y_t = Q_K C_t + Q_new R_t); the last block carries one
// more chain vector, whose projection on the basis gives the last column.
// Per block the basis is read twice instead of once per column; a lane whose
// Gram matrix is not positive definite raises the breakdown flag.
constexpr int kSMaxCols = 40;  // K + n columns of one pass (at most m + 1)
constexpr int kSMaxNY = 9;     // chain vectors per block (S, and the last block's extra one)
// dot columns per workgroup of the dots pass (5 for 8 and 9 chain vectors: registers)
constexpr int sgroup(int ny) { return ny >= 8 ? 5 : 6; }
constexpr int kSLanes = 16;    // lanes per workgroup of the forming pass
constexpr int kSCoefLanes = 4; // lanes per workgroup of the coefficient kernel
struct ColPtrs {
    const double2 *p[kSMaxCols];
};
struct OutPtrs {
    double2 *p[kSMaxNY];
};

// part[rb][x*NY + j][r] = sum over the block's rows of conj(x_x) y_j, for the
// sgroup(NY) columns x of group g; y_j = x_{K+j}.  Workgroup -> (g, row block,
// lane chunk) with the groups of one row block consecutive on one XCD (they
// read the same rows of Y).
template <int NY>
__global__ __launch_bounds__(kBlock) void k_sdots(ColPtrs x, int nx, int K, double2 *__restrict__ part, int64_t N,
                                                  int Mp, int64_t rpb, int nrb, int ngroups)
{
    constexpr int KG = sgroup(NY);
    __shared__ double2 red[kWaves - 1][KG][kWave];
    const int nb = nrb * (Mp / kWave) * ngroups;
    int bid = blockIdx.x;
    if ((nb & 7) == 0) bid = (bid & 7) * (nb >> 3) + (bid >> 3);
    const int g = bid % ngroups, t = bid / ngroups;
    const int rb = t % nrb, ch = t / nrb;
    const int lane = threadIdx.x & (kWave - 1);
    const int wave = wave_uniform(threadIdx.x / kWave);
    const int r = ch * kWave + lane;
    const int x0 = g * KG;
    const int nxg = min(KG, nx - x0);
    const double2 *xp[KG], *yp[NY];
#pragma unroll
    for (int i = 0; i < KG; ++i) xp[i] = x.p[min(x0 + i, nx - 1)];
#pragma unroll
    for (int j = 0; j < NY; ++j) yp[j] = x.p[K + j];
    double2 acc[KG][NY];
#pragma unroll
    for (int i = 0; i < KG; ++i)
#pragma unroll
        for (int j = 0; j < NY; ++j) acc[i][j] = cz(0, 0);
    const int64_t row0 = (int64_t)rb * rpb;
    const int64_t row1 = row0 + rpb < N ? row0 + rpb : N;
    for (int64_t row = row0 + wave; row < row1; row += kWaves) {
        const int64_t off = row * Mp + r;
        double2 yv[NY], xv[KG];
#pragma unroll
        for (int j = 0; j < NY; ++j) yv[j] = yp[j][off];
#pragma unroll
        for (int i = 0; i < KG; ++i) xv[i] = xp[i][off];
#pragma unroll
        for (int i = 0; i < KG; ++i)
#pragma unroll
            for (int j = 0; j < NY; ++j) acc[i][j] = cfma_conj(xv[i], yv[j], acc[i][j]);
    }
    const int ndots = nx * NY;
#pragma unroll
    for (int j = 0; j < NY; ++j) {
        if (wave > 0)
#pragma unroll
            for (int i = 0; i < KG; ++i) red[wave - 1][i][lane] = acc[i][j];
        __syncthreads();
        if (wave == 0)
#pragma unroll
            for (int i = 0; i < KG; ++i) {
                if (i < nxg) {
                    double2 s = acc[i][j];
#pragma unroll
                    for (int w = 0; w < kWaves - 1; ++w) s = cadd(s, red[w][i][lane]);
                    part[((int64_t)rb * ndots + (x0 + i) * NY + j) * Mp + r] = s;
                }
            }
        __syncthreads();
    }
}

// o_j = sum_x coef[x*NY + j] x_x over the rows of the block, for kSLanes lanes
// (their coefficients in LDS); the outputs may be inputs (each thread reads
// all of its row before writing it).
template <int NY>
__global__ __launch_bounds__(kBlock) void k_sform(ColPtrs x, int nx, OutPtrs o, const double2 *__restrict__ coef,
                                                  int64_t N, int Mp, int64_t rpb)
{
    extern __shared__ double2 sc[]; // [nx * NY][kSLanes]
    constexpr int RS = kBlock / kSLanes;
    const int ngl = Mp / kSLanes;
    const int lg = blockIdx.x % ngl, rb = blockIdx.x / ngl;
    const int ln = threadIdx.x % kSLanes, rs = threadIdx.x / kSLanes;
    const int r = lg * kSLanes + ln;
    for (int idx = threadIdx.x; idx < nx * NY * kSLanes; idx += kBlock)
        sc[idx] = coef[(int64_t)(idx / kSLanes) * Mp + lg * kSLanes + idx % kSLanes];
    __syncthreads();
    const int64_t row0 = (int64_t)rb * rpb;
    const int64_t row1 = row0 + rpb < N ? row0 + rpb : N;
    for (int64_t row = row0 + rs; row < row1; row += RS) {
        const int64_t off = row * Mp + r;
        double2 out[NY];
#pragma unroll
        for (int j = 0; j < NY; ++j) out[j] = cz(0, 0);
        for (int c0 = 0; c0 < nx; c0 += 8) {
            double2 v[8];
#pragma unroll
            for (int i = 0; i < 8; ++i) v[i] = x.p[min(c0 + i, nx - 1)][off];
#pragma unroll
            for (int i = 0; i < 8; ++i) {
                if (c0 + i < nx) {
#pragma unroll
                    for (int j = 0; j < NY; ++j) out[j] = cfma(sc[((c0 + i) * NY + j) * kSLanes + ln], v[i], out[j]);
                }
            }
        }
#pragma unroll
        for (int j = 0; j < NY; ++j) ntstore(o.p[j] + off, out[j]);
    }
}

// Per lane: R from G - C^H C, the forming coefficients and the Hessenberg
// columns.  dots[x*n + t] (x < K + n, t < n): C = Q_K^H Y for x < K, G = Y^H Y
// for x >= K.  The first nf = n - extra chain vectors are formed into basis
// vectors (coef[x*nf + j] as k_sform reads it) and give the columns K-1 ..
// K+nf-2 (first block, K = 0: columns 0 .. nf-2 and beta_0).  With extra = 1
// (the last block) chain vector nf is not formed: its projection on the basis,
// [C_nf; R^-H (G_{:nf,nf} - C^H C_nf)], gives the last column K+nf-1 = m-1.
__global__ __launch_bounds__(kBlock) void k_scoef(int K, int n, int extra, int m, int M, int Mp,
                                                  const double2 *__restrict__ dots, double2 *__restrict__ coef,
                                                  double2 *__restrict__ hess, double *__restrict__ beta0,
                                                  int32_t *__restrict__ flags, unsigned long long *__restrict__ cond)
{
    constexpr int NL = kSCoefLanes, TP = kBlock / NL;
    static_assert(TP >= kMaxKrylov, "one thread per Hessenberg row");
    __shared__ double2 sD[NL][kSMaxCols * kSMaxNY];
    __shared__ double2 sR[NL][kSMaxNY * kSMaxNY], sRi[NL][kSMaxNY * kSMaxNY], sW[NL][2 * kSMaxNY];
    __shared__ double sd[NL][kSMaxNY], sinv[NL];
    __shared__ int sok[NL], sfail[NL];
    const int l = threadIdx.x % NL, q = threadIdx.x / NL;
    const int r = blockIdx.x * NL + l;
    const bool active = r < M;
    const int nd = (K + n) * n;
    const int nf = n - extra;
    if (q == 0) sfail[l] = 0;
    for (int base = 0; base < nd; base += 8 * TP) { // eight loads in flight per thread
        double2 v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) v[u] = active ? dots[(int64_t)min(base + q + u * TP, nd - 1) * Mp + r] : cz(0, 0);
#pragma unroll
        for (int u = 0; u < 8; ++u)
            if (base + q + u * TP < nd) sD[l][base + q + u * TP] = v[u];
    }
    __syncthreads();
    const double2 *D = sD[l]; // C(c, t) = D[c*n + t], G(a, t) = D[(K + a)*n + t]
    double2 *R = sR[l], *Ri = sRi[l], *W = sW[l];
    // G - C^H C on rows a < nf, columns b < n (column nf: the extra vector's)
    for (int idx = q; idx < nf * n; idx += TP) {
        const int a = idx / n, b = idx % n;
        double2 g = D[(K + a) * n + b];
        for (int c = 0; c < K; ++c) g = csub(g, cmul(cconj(D[c * n + a]), D[c * n + b]));
        if (b < nf)
            R[a * nf + b] = g;
        else
            W[a] = g;
    }
    __syncthreads();
    // scaled Cholesky (right-looking, the lane's 32 threads over the entries),
    // then R^-1 and the extra vector's coordinates R^-H w, one column per thread
    if (q == 0) sok[l] = active ? 1 : 0;
    if (q < nf) {
        const double x = R[q * nf + q].x;
        const bool good = x > 0.0 && isfinite(x);
        sd[l][q] = good ? sqrt(x) : 1.0;
        if (!good) sfail[l] = 1;
    }
    __syncthreads();
    for (int idx = q; idx < nf * nf; idx += TP)
        R[idx] = cscale(1.0 / (sd[l][idx / nf] * sd[l][idx % nf]), R[idx]);
    __syncthreads();
    double pmin = 1.0; // smallest pivot of the unit-diagonal factor (thread q = 0)
    for (int i = 0; i < nf; ++i) {
        if (q == 0) {
            double x = R[i * nf + i].x;
            if (!(x > 0.0) || !isfinite(x)) {
                sfail[l] = 1;
                x = 1.0;
            }
            const double rii = sqrt(x);
            pmin = fmin(pmin, rii);
            R[i * nf + i] = cz(rii, 0.0);
            sinv[l] = 1.0 / rii;
        }
        __syncthreads();
        if (q > i && q < nf) R[i * nf + q] = cscale(sinv[l], R[i * nf + q]);
        __syncthreads();
        for (int idx = q; idx < nf * nf; idx += TP) {
            const int a = idx / nf, b = idx % nf;
            if (a > i && b >= a) R[idx] = csub(R[idx], cmul(cconj(R[i * nf + a]), R[i * nf + b]));
        }
        __syncthreads();
    }
    const bool ok = sok[l] != 0 && sfail[l] == 0;
    for (int idx = q; idx < nf * nf; idx += TP) {
        const int a = idx / nf, b = idx % nf;
        R[idx] = (b < a || !ok) ? cz(0, 0) : cscale(sd[l][b], R[idx]);
    }
    __syncthreads();
    if (q < nf) {
        const int j = q;
        if (ok) {
            Ri[j * nf + j] = cz(1.0 / R[j * nf + j].x, 0.0);
            for (int i = j - 1; i >= 0; --i) {
                double2 s = cz(0, 0);
                for (int k = i + 1; k <= j; ++k) s = cfma(R[i * nf + k], Ri[k * nf + j], s);
                Ri[i * nf + j] = cscale(-1.0 / R[i * nf + i].x, s);
            }
        }
        for (int i = ok ? j + 1 : 0; i < nf; ++i) Ri[i * nf + j] = cz(0, 0);
        double2 pj = cz(0, 0);
        if (ok && extra)
            for (int a = 0; a <= j; ++a) pj = cfma_conj(Ri[a * nf + j], W[a], pj);
        W[kSMaxNY + j] = pj;
    }
    if (q == 0 && active && !ok) flags[0] = 1; // Krylov breakdown (the block's vectors are dependent)
    if (q == 0 && active && ok && cond) atomicMax(cond, (unsigned long long)__double_as_longlong(1.0 / pmin));
    __syncthreads();
    // forming coefficients: x < K: -(C R^-1)[x][j]; x = K + t: (R^-1)[t][j]
    for (int idx = q; idx < (K + nf) * nf; idx += TP) {
        const int xx = idx / nf, j = idx % nf;
        double2 cf = cz(0, 0);
        if (xx < K) {
            for (int t = 0; t <= j; ++t) cf = csub(cf, cmul(D[xx * n + t], Ri[t * nf + j]));
        } else {
            cf = Ri[(xx - K) * nf + j];
        }
        coef[(int64_t)idx * Mp + r] = cf;
    }
    if (q >= m) return;
    double2 acc[kSMaxNY];
    if (K == 0) {
        // H y_j = y_{j+1}, y_j = Q R_j: column j = (R_{j+1} - sum_{l<j} col_l R[l][j]) / R[j][j]
#pragma unroll
        for (int j = 0; j < kSMaxNY - 1; ++j) {
            acc[j] = cz(0, 0);
            if (ok && j < nf - 1) {
                double2 a = q < nf ? R[q * nf + j + 1] : cz(0, 0);
#pragma unroll
                for (int k = 0; k < j; ++k) a = csub(a, cmul(acc[k], R[k * nf + j]));
                acc[j] = cscale(1.0 / R[j * nf + j].x, a);
            }
            if (j < nf - 1) hess[((int64_t)j * m + q) * Mp + r] = acc[j];
        }
        if (q == 0) beta0[r] = ok ? R[0].x : 0.0;
        return;
    }
    // y_t = Q_K C_t + Q_new R_t (t < nf; t = nf: the projection of the extra
    // vector); H q_{K-1} = y_0 and H y_{t-1} = y_t
    const int ncol = nf + extra;
#pragma unroll
    for (int t = 0; t < kSMaxNY; ++t) {
        double2 a = cz(0, 0);
        if (ok && t < ncol) {
            if (q < K)
                a = D[q * n + t];
            else if (q - K < nf)
                a = t < nf ? R[(q - K) * nf + t] : W[kSMaxNY + q - K];
        }
        acc[t] = a;
    }
    if (ok) {
        // the old columns' entries of row q (zero below the subdiagonal), eight
        // loads in flight at a time
        const int c_lo = q > 0 ? q - 1 : 0, c_hi = K - 1;
        for (int c0 = c_lo; c0 < c_hi; c0 += 8) {
            double2 h[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) h[u] = hess[((int64_t)min(c0 + u, c_hi - 1) * m + q) * Mp + r];
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                if (c0 + u < c_hi) {
#pragma unroll
                    for (int t = 1; t < kSMaxNY; ++t)
                        if (t < ncol) acc[t] = csub(acc[t], cmul(h[u], D[(c0 + u) * n + t - 1]));
                }
            }
        }
#pragma unroll
        for (int t = 1; t < kSMaxNY; ++t) {
            if (t < ncol) {
                double2 a = csub(acc[t], cmul(acc[0], D[(K - 1) * n + t - 1]));
#pragma unroll
                for (int k = 0; k < t - 1; ++k) a = csub(a, cmul(acc[k + 1], R[k * nf + t - 1]));
                acc[t] = cscale(1.0 / R[(t - 1) * nf + t - 1].x, a);
            }
        }
    }
#pragma unroll
    for (int t = 0; t < kSMaxNY; ++t)
        if (t < ncol) hess[((int64_t)(K - 1 + t) * m + q) * Mp + r] = acc[t];
}

using SDotsFn = void (*)(ColPtrs, int, int, double2 *, int64_t, int, int64_t, int, int);
using SFormFn = void (*)(ColPtrs, int, OutPtrs, const double2 *, int64_t, int, int64_t);
const SDotsFn kSDots[kSMaxNY] = {k_sdots<1>, k_sdots<2>, k_sdots<3>, k_sdots<4>, k_sdots<5>,
                                 k_sdots<6>, k_sdots<7>, k_sdots<8>, k_sdots<9>};
const SFormFn kSForm[kSMaxNY] = {k_sform<1>, k_sform<2>, k_sform<3>, k_sform<4>, k_sform<5>,
                                 k_sform<6>, k_sform<7>, k_sform<8>, k_sform<9>};

// Row blocks of the dots pass: a multiple of 8 (XCD mapping), as many as the
// partials buffer holds for the largest pass, at most sstep_row_blocks(N).
int sstep_nrb(const KrylovWork &w, int S)
{
    const size_t worst = (size_t)(w.m + 1) * (size_t)std::min(S + 1, kSMaxNY);
    const size_t fit = w.part_capacity / (worst * (size_t)w.Mp);
    return (int)std::min<size_t>((size_t)sstep_row_blocks(w.N), fit & ~(size_t)7);
}

hipError_t sdots_pass(KrylovWork &w, const ColPtrs &x, int nx, int K, int n, int nrb, hipStream_t s)
{
    const int ngroups = (nx + sgroup(n) - 1) / sgroup(n);
    const int64_t rpb = ceil_div(w.N, nrb);
    {
        ProfScope ps(w, PF_SDOTS, nx * vec_bytes(w), s);
        hipLaunchKernelGGL(kSDots[n - 1], dim3(nrb * (w.Mp / kWave) * ngroups), dim3(kBlock), 0, s, x, nx, K, w.part,
                           w.N, w.Mp, rpb, nrb, ngroups);
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    return reduce(w, nx * n, nrb, w.sbuf, s);
}

hipError_t arnoldi_sstep(KrylovWork &w, int S, int nrb, hipStream_t s)
{
    const int m = w.m;
    const int64_t qs = qstride(w);
    double2 *const u0 = w.u, *const v0 = w.v;
    double2 *const sdots = w.sbuf, *const scoef = w.sbuf + (size_t)kSMaxCols * kSMaxNY * w.Mp;
    auto slot = [&](int k) { return w.Q + (int64_t)k * qs; };
    auto product = [&](double2 *in, double2 *out) {
        w.u = in;
        w.v = out;
        w.no_dots = true;
        const hipError_t e = spmm_launch(w, s);
        w.u = u0;
        w.v = v0;
        w.no_dots = false;
        return e;
    };
    hipError_t e = hipSuccess;
    const int64_t rpf = 128; // rows per workgroup of the forming pass
    const int nbf = (int)ceil_div(w.N, rpf) * (w.Mp / kSLanes);
    for (int K = 0; K < m;) {
        ColPtrs x{};
        OutPtrs o{};
        int n, extra = 0;
        if (K == 0) { // [psi, H psi, .., H^(n-1) psi]: blocks of S columns throughout (m = 24, S = 6: 6, 6, 6, 6)
            n = std::min(S, m - 1);
            x.p[0] = u0;
            o.p[0] = slot(0);
            for (int j = 1; j < n; ++j) {
                if ((e = product(j == 1 ? u0 : slot(j - 1), slot(j))) != hipSuccess) return e;
                x.p[j] = o.p[j] = slot(j);
            }
        } else {
            n = std::min(S, m - K);
            for (int c = 0; c < K; ++c) x.p[c] = slot(c);
            for (int t = 0; t < n; ++t) {
                if ((e = product(slot(K - 1 + t), slot(K + t))) != hipSuccess) return e;
                x.p[K + t] = o.p[t] = slot(K + t);
            }
            if (K + n == m) { // one more chain vector, H q_{m-1}'s direction, gives the last column
                extra = 1;
                if ((e = product(slot(m - 1), v0)) != hipSuccess) return e;
                x.p[m] = v0;
            }
        }
        const int nd = n + extra;
        if ((e = sdots_pass(w, x, K + nd, K, nd, nrb, s)) != hipSuccess) return e;
        {
            ProfScope ps(w, PF_COEFFS, 0.0, s);
            hipLaunchKernelGGL(k_scoef, dim3((unsigned)ceil_div(w.Mp, kSCoefLanes)), dim3(kBlock), 0, s, K, nd, extra,
                               m, w.M, w.Mp, sdots, scoef, w.hess, w.beta0, w.flags, w.cond);
        }
        if ((e = hipGetLastError()) != hipSuccess) return e;
        if (extra && w.fold_psi) { // psi_update forms psi from [Q_K, Y] directly
            w.fold_K = K;
            w.fold_nf = n;
            w.fold_coef = scoef;
            w.fold_q0 = nullptr;
            break;
        }
        {
            const int nx = K + n;
            ProfScope ps(w, PF_SFORM, (nx + n) * vec_bytes(w), s);
            const size_t lds = sizeof(double2) * (size_t)nx * n * kSLanes;
            if (lds > 65536) return hipErrorInvalidValue; // m <= 32, n <= 8: at most 64 KiB
            hipLaunchKernelGGL(kSForm[n - 1], dim3(nbf), dim3(kBlock), lds, s, x, nx, o, scoef, w.N, w.Mp, rpf);
        }
        if ((e = hipGetLastError()) != hipSuccess) return e;
        K += n;
    }
    return hipSuccess;
}

// ------------------------------------------------- Newton-basis Arnoldi
// A basis that serves one step, built as one block in the Newton basis (Bai,
// Hu and Reichel, IMA J. Numer. Anal. 14 (1994) 563; Hoemmen 2010, ch. 7):
// y_0 = psi, y_{j+1} = (H - sig_j) y_j for j < m with Leja-ordered Ritz
// values sig_j of H's Hermitian part as shifts, one pass for the Gram matrix
// of [y_0 .. y_m] (upper triangle, packed), and per lane its Cholesky factor
// Y = Q R (columns scaled to unit diagonal first), R^-1 and the Hessenberg
// matrix.  Since span(y_0 .. y_j) is the Krylov space of order j + 1 for any
// shifts, Q and the Hessenberg matrix are those of the column form up to
// rounding.  No basis vector is written: psi_update takes psi's coordinates
// over y_0 .. y_{m-1} from R^-1.  Per step m products, one pass over m + 1
// vectors and the psi pass over m, against 2m, 49 and 25 + 40 passes of
// the blocks of 8 at m = 24.
constexpr int kNTile = 5;        // Gram tile edge of the dots pass (25 complex accumulators per thread)
constexpr int kNRowBlocks = 64;  // row blocks of the dots pass (at most)
constexpr int kNPacked = (kMaxKrylov + 1) * (kMaxKrylov + 2) / 2; // packed Gram entries at m = kMaxKrylov
struct NShifts {
    double s[kMaxKrylov];
};

// part[rb][packed_index(x, y)][r] = sum over the block's rows of conj(x_x) x_y
// for the tile pair (a, b), a <= b, of T-column tiles (x in tile a, y in tile
// b, x <= y).  Workgroup -> (pair, row block, lane chunk), the pairs of one
// row block consecutive on one XCD (they read the same rows).
template <int T>
__global__ __launch_bounds__(kBlock) void k_ngram(ColPtrs x, int nd, double2 *__restrict__ part, int64_t N, int Mp,
                                                  int64_t rpb, int nrb, int nt)
{
    __shared__ double2 red[kWaves - 1][T][kWave];
    const int npairs = nt * (nt + 1) / 2, npk = nd * (nd + 1) / 2;
    const int nb = nrb * (Mp / kWave) * npairs;
    int bid = blockIdx.x;
    if ((nb & 7) == 0) bid = (bid & 7) * (nb >> 3) + (bid >> 3);
    const int pr = bid % npairs, t = bid / npairs;
    const int rb = t % nrb, ch = t / nrb;
    int a = 0, rem = pr;
    while (rem >= nt - a) {
        rem -= nt - a;
        ++a;
    }
    const int b = a + rem;
    const int lane = threadIdx.x & (kWave - 1);
    const int wave = wave_uniform(threadIdx.x / kWave);
    const int r = ch * kWave + lane;
    const int x0 = a * T, y0 = b * T;
    const double2 *xp[T], *yp[T];
#pragma unroll
    for (int i = 0; i < T; ++i) {
        xp[i] = x.p[min(x0 + i, nd - 1)];
        yp[i] = x.p[min(y0 + i, nd - 1)];
    }
    double2 acc[T][T];
#pragma unroll
    for (int i = 0; i < T; ++i)
#pragma unroll
        for (int j = 0; j < T; ++j) acc[i][j] = cz(0, 0);
    const int64_t row0 = (int64_t)rb * rpb;
    const int64_t row1 = row0 + rpb < N ? row0 + rpb : N;
    if (a == b) { // diagonal tile: the upper triangle of one tile's vectors
        for (int64_t row = row0 + wave; row < row1; row += kWaves) {
            const int64_t off = row * Mp + r;
            double2 xv[T];
#pragma unroll
            for (int i = 0; i < T; ++i) xv[i] = xp[i][off];
#pragma unroll
            for (int i = 0; i < T; ++i)
#pragma unroll
                for (int j = i; j < T; ++j) acc[i][j] = cfma_conj(xv[i], xv[j], acc[i][j]);
        }
    } else {
        for (int64_t row = row0 + wave; row < row1; row += kWaves) {
            const int64_t off = row * Mp + r;
            double2 xv[T], yv[T];
#pragma unroll
            for (int j = 0; j < T; ++j) yv[j] = yp[j][off];
#pragma unroll
            for (int i = 0; i < T; ++i) xv[i] = xp[i][off];
#pragma unroll
            for (int i = 0; i < T; ++i)
#pragma unroll
                for (int j = 0; j < T; ++j) acc[i][j] = cfma_conj(xv[i], yv[j], acc[i][j]);
        }
    }
#pragma unroll
    for (int j = 0; j < T; ++j) {
        if (wave > 0)
#pragma unroll
            for (int i = 0; i < T; ++i) red[wave - 1][i][lane] = acc[i][j];
        __syncthreads();
        if (wave == 0)
#pragma unroll
            for (int i = 0; i < T; ++i) {
                const int xi = x0 + i, yj = y0 + j;
                if (xi < nd && yj < nd && xi <= yj) {
                    double2 s = acc[i][j];
#pragma unroll
                    for (int w = 0; w < kWaves - 1; ++w) s = cadd(s, red[w][i][lane]);
                    part[((int64_t)rb * npk + packed_index(xi, yj, nd)) * Mp + r] = s;
                }
            }
        __syncthreads();
    }
}

// The same partials with each row of the chain read once per lane: a wave
// takes L lanes and its 64 / L threads per lane the tile pairs (a, b) of that
// lane's Gram matrix (T-column tiles, npairs <= 64 / L), so the chain's rows
// for its lanes are loaded once (the 64 threads load nd x L elements
// together, coalesced in runs of L lanes) and shared through LDS, with D rows
// in flight while the previous D are used.  The kWaves waves of a workgroup
// take every kWaves-th row of one row block for the same L lanes and add
// their sums through LDS at the end.  Vectors: y_0 = u0, y_v = Q + v qs for
// 0 < v < nd - 1, y_{nd-1} = ylast.  Tiles on the diagonal compute all T x T
// products (the lower ones are not stored).
// BIN = kBinAll sums over all N rows.  BIN = 0 .. 3 sums over the Z nuclear
// indices n the electron-state components of the projector P_X: 0 singlet
// (y[4n+1] - y[4n+2])/sqrt2, 1 T0 (y[4n+1] + y[4n+2])/sqrt2, 2 T+ y[4n],
// 3 T- y[4n+3] (N is then Z), so the four bins add up to the plain Gram
// matrix and bin X is Y^H P_X Y.  Partial sums go to
// part[(rb * pks + pko + packed) * Mp + r].
constexpr int kBinAll = 4;
template <int T, int L, int D, int BIN>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(1, 2))) void k_ngram_rows(
    const double2 *__restrict__ u0, const double2 *__restrict__ Q, int64_t qs, const double2 *__restrict__ ylast,
    int nd, double2 *__restrict__ part, int64_t N, int Mp, int64_t rpb, int nrb, int nt, int pks, int pko)
{
    // largest nd: npairs <= 64 / L tile pairs (25 at L = 4, 33 at L = 2)
    constexpr int NDX = L == 4 ? 5 * T : kMaxKrylov + 1;
    constexpr int EPT = (NDX * L + kWave - 1) / kWave; // elements per thread per row
    __shared__ double2 buf[kWaves][2][D][NDX * L];
    static_assert(kWaves * 2 * D * NDX * L >= (kWaves - 1) * T * kWave, "the sums reuse the row buffers");
    double2(*red)[T][kWave] = reinterpret_cast<double2(*)[T][kWave]>(&buf[0][0][0][0]);
    const int t = threadIdx.x & (kWave - 1);
    const int wave = wave_uniform(threadIdx.x / kWave);
    const int ngr = Mp / L;
    const int nb = nrb * ngr;
    int bid = blockIdx.x;
    if ((nb & 7) == 0) bid = (bid & 7) * (nb >> 3) + (bid >> 3); // neighbouring lane groups on one XCD
    const int lg = bid % ngr, rb = bid / ngr;
    const int r0 = lg * L; // the workgroup's first lane
    const int ne = nd * L;
    // loading: element e = t + 64k is lane e % L of vector e / L
    const double2 *lp[EPT];
    int lc[EPT];
#pragma unroll
    for (int k = 0; k < EPT; ++k) {
        const int e = min(t + kWave * k, ne - 1), v = e / L;
        lp[k] = v == 0 ? u0 : v == nd - 1 ? ylast : Q + (int64_t)v * qs;
        lc[k] = r0 + e % L;
    }
    // computing: thread t takes lane l = t % L and tile pair p = t / L
    const int l = t % L, p = t / L;
    int a = 0, rem = min(p, nt * (nt + 1) / 2 - 1);
    while (rem >= nt - a) {
        rem -= nt - a;
        ++a;
    }
    const int x0 = a * T, y0 = (a + rem) * T;
    double2 acc[T][T];
#pragma unroll
    for (int i = 0; i < T; ++i)
#pragma unroll
        for (int j = 0; j < T; ++j) acc[i][j] = cz(0, 0);
    const int64_t rbeg = (int64_t)rb * rpb;
    const int64_t rend = rbeg + rpb < N ? rbeg + rpb : N;
    // this wave's rows: rbeg + wave + kWaves i, i < nr
    const int64_t nr = rend - rbeg > wave ? (rend - rbeg - wave + kWaves - 1) / kWaves : 0;
    // two components per element for the singlet and T0 bins (combined when
    // stashed, so the loads stay in flight over the products)
    constexpr int D2 = BIN == 0 || BIN == 1 ? D : 1;
    constexpr int ROW0 = BIN == 2 ? 0 : BIN == 3 ? 3 : 1;
    double2 pf[D][EPT], pf2[D2][EPT];
    int64_t fi0 = 0; // first row index of the rows in pf
    auto fetch = [&](int64_t i0) {
        fi0 = i0;
#pragma unroll
        for (int d = 0; d < D; ++d) {
            const int64_t row = rbeg + wave + kWaves * min(i0 + d, nr - 1);
#pragma unroll
            for (int k = 0; k < EPT; ++k) {
                if constexpr (BIN == kBinAll) {
                    pf[d][k] = lp[k][row * Mp + lc[k]];
                } else {
                    pf[d][k] = lp[k][(4 * row + ROW0) * Mp + lc[k]];
                    if constexpr (BIN == 0 || BIN == 1) pf2[d][k] = lp[k][(4 * row + 2) * Mp + lc[k]];
                }
            }
        }
    };
    auto stash = [&](int bb) { // rows past the wave's last are stored as zeros (they add nothing)
        constexpr double h = 0.70710678118654752440;
#pragma unroll
        for (int d = 0; d < D; ++d)
#pragma unroll
            for (int k = 0; k < EPT; ++k)
                if (t + kWave * k < ne) {
                    double2 v = pf[d][k];
                    if constexpr (BIN == 0) v = cscale(h, csub(pf[d][k], pf2[d][k]));
                    if constexpr (BIN == 1) v = cscale(h, cadd(pf[d][k], pf2[d][k]));
                    buf[wave][bb][d][t + kWave * k] = fi0 + d < nr ? v : cz(0, 0);
                }
    };
    auto wave_sync = [] {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    };
    if (nr > 0) {
        fetch(0);
        stash(0);
    }
    wave_sync();
    int bb = 0;
    for (int64_t i0 = 0; i0 < nr; i0 += D, bb ^= 1) {
        const bool more = i0 + D < nr;
        if (more) fetch(i0 + D);
#pragma unroll
        for (int d = 0; d < D; ++d) {
            const double2 *B = buf[wave][bb][d];
            double2 yv[T];
#pragma unroll
            for (int j = 0; j < T; ++j) yv[j] = B[min(y0 + j, nd - 1) * L + l];
#pragma unroll
            for (int i = 0; i < T; ++i) {
                const double2 xi = B[min(x0 + i, nd - 1) * L + l];
#pragma unroll
                for (int j = 0; j < T; ++j) acc[i][j] = cfma_conj(xi, yv[j], acc[i][j]);
            }
        }
        wave_sync(); // every lane has read buffer bb ^ 1's rows of the previous round
        if (more) stash(bb ^ 1);
        wave_sync();
    }
    const bool writer = p < nt * (nt + 1) / 2;
    const int r = r0 + l;
    __syncthreads(); // every wave is done with the row buffers
#pragma unroll
    for (int j = 0; j < T; ++j) {
        if (wave > 0)
#pragma unroll
            for (int i = 0; i < T; ++i) red[wave - 1][i][t] = acc[i][j];
        __syncthreads();
        if (wave == 0 && writer)
#pragma unroll
            for (int i = 0; i < T; ++i) {
                const int xi = x0 + i, yj = y0 + j;
                if (xi < nd && yj < nd && xi <= yj) {
                    double2 s = acc[i][j];
#pragma unroll
                    for (int w = 0; w < kWaves - 1; ++w) s = cadd(s, red[w][i][t]);
                    part[((int64_t)rb * pks + pko + packed_index(xi, yj, nd)) * Mp + r] = s;
                }
            }
        __syncthreads();
    }
}

// The Gram matrix as the sum of its four projector bins: G[e][r] =
// sum_x bins[x][e][r] over the npk packed entries (one thread per entry and lane).
__global__ void k_sum_bins(int64_t n, int64_t bstride, const double2 *__restrict__ bins, double2 *__restrict__ G)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double2 a = bins[i], b = bins[i + bstride], c = bins[i + 2 * bstride], d = bins[i + 3 * bstride];
    G[i] = cadd(cadd(a, b), cadd(c, d));
}

// Per lane (one wave each): from the packed Gram matrix G of [y_0 .. y_m]
// (nd = m + 1), R^H R = G[:m, :m] (Cholesky with the columns scaled to unit
// diagonal first), coef[x*m + j] = (R^-1)[x][j] (q_j = sum_x coef[x*m + j] y_x),
// the extra vector's coordinates p = R^-H G[:m, m] = Q^H y_m, the Hessenberg
// columns h_j = (r_{j+1} + sig_j R_j - sum_{k<j} R[k][j] h_k) / R[j][j]
// (from H y_j = y_{j+1} + sig_j y_j in the basis; r_{j+1} = R_{j+1}, r_m = p),
// beta_0 = R_00 = |psi|, and cond[1] = max(cond[1], m ||L^-1||_F^2), L the
// scaled factor (a bound on the scaled Gram matrix's condition number).
__global__ __launch_bounds__(kWave) void k_ncoef(int m, int M, int Mp, NShifts sh, const double2 *__restrict__ dots,
                                                 double2 *__restrict__ coef, double2 *__restrict__ hess,
                                                 double *__restrict__ beta0, int32_t *__restrict__ flags,
                                                 unsigned long long *__restrict__ cond)
{
    constexpr int MX = kMaxKrylov;
    __shared__ double2 R[MX * MX], Ri[MX * MX], Hc[MX * MX];
    __shared__ double2 W[MX], P[MX];
    __shared__ double sd[MX], colk[MX], sinv;
    __shared__ int sfail;
    const int r = blockIdx.x, q = threadIdx.x;
    const bool active = r < M;
    const int nd = m + 1;
    if (q == 0) sfail = active ? 0 : 1;
    { // G rows 0 .. m-1, upper triangle: every load in flight before the first store
        constexpr int EPT = (MX * (MX + 1) + kWave - 1) / kWave;
        const int ne = m * nd;
        double2 gv[EPT];
#pragma unroll
        for (int u = 0; u < EPT; ++u) {
            const int e = min(q + kWave * u, ne - 1), a = e / nd, b = max(e % nd, a);
            gv[u] = dots[(int64_t)packed_index(a, b, nd) * Mp + r];
        }
#pragma unroll
        for (int u = 0; u < EPT; ++u) {
            const int e = q + kWave * u, a = e / nd, b = e % nd;
            if (e >= ne || b < a) continue;
            if (b < m)
                R[a * MX + b] = gv[u];
            else
                W[a] = gv[u];
        }
    }
    __syncthreads();
    if (q < m) {
        const double x = R[q * MX + q].x;
        const bool good = x > 0.0 && isfinite(x);
        sd[q] = good ? sqrt(x) : 1.0;
        if (!good) sfail = 1;
    }
    __syncthreads();
    for (int idx = q; idx < m * m; idx += kWave) {
        const int a = idx / m, b = idx % m;
        if (b >= a) R[a * MX + b] = cscale(1.0 / (sd[a] * sd[b]), R[a * MX + b]);
    }
    __syncthreads();
    for (int i = 0; i < m; ++i) { // right-looking Cholesky of the unit-diagonal matrix
        if (q == 0) {
            double x = R[i * MX + i].x;
            if (!(x > 0.0) || !isfinite(x)) {
                sfail = 1;
                x = 1.0;
            }
            const double rii = sqrt(x);
            R[i * MX + i] = cz(rii, 0.0);
            sinv = 1.0 / rii;
        }
        __syncthreads();
        if (q > i && q < m) R[i * MX + q] = cscale(sinv, R[i * MX + q]);
        __syncthreads();
        for (int idx = q; idx < m * m; idx += kWave) {
            const int a = idx / m, b = idx % m;
            if (a > i && b >= a) R[a * MX + b] = csub(R[a * MX + b], cmul(cconj(R[i * MX + a]), R[i * MX + b]));
        }
        __syncthreads();
    }
    const bool ok = sfail == 0;
    for (int idx = q; idx < m * m; idx += kWave) { // R = L^H diag(sd)
        const int a = idx / m, b = idx % m;
        R[a * MX + b] = (b < a || !ok) ? cz(0, 0) : cscale(sd[b], R[a * MX + b]);
    }
    __syncthreads();
    if (q < m) { // column j of R^-1, the extra vector's coordinate p_j and column j's share of ||L^-1||_F^2
        const int j = q;
        if (ok) {
            Ri[j * MX + j] = cz(1.0 / R[j * MX + j].x, 0.0);
            for (int i = j - 1; i >= 0; --i) {
                double2 acc = cz(0, 0);
                for (int k = i + 1; k <= j; ++k) acc = cfma(R[i * MX + k], Ri[k * MX + j], acc);
                Ri[i * MX + j] = cscale(-1.0 / R[i * MX + i].x, acc);
            }
        }
        for (int i = ok ? j + 1 : 0; i < m; ++i) Ri[i * MX + j] = cz(0, 0);
        double2 pj = cz(0, 0);
        double kc = 0.0;
        if (ok)
            for (int a = 0; a <= j; ++a) {
                pj = cfma_conj(Ri[a * MX + j], W[a], pj);
                kc += cabs2(cscale(sd[a], Ri[a * MX + j]));
            }
        P[j] = pj;
        colk[j] = kc;
    }
    __syncthreads();
    if (q == 0 && active) {
        if (!ok) {
            flags[0] = 1; // Krylov breakdown (the chain vectors are dependent)
        } else if (cond) {
            double kap = 0.0;
            for (int j = 0; j < m; ++j) kap += colk[j];
            kap *= m;
            atomicMax(cond + 1, (unsigned long long)__double_as_longlong(kap));
        }
    }
    for (int idx = q; idx < m * m; idx += kWave) coef[(int64_t)idx * Mp + r] = Ri[(idx / m) * MX + idx % m];
    if (q == 0) beta0[r] = ok ? R[0].x : 0.0;
    if (q >= m) return;
    for (int j = 0; j < m; ++j) { // row q of the Hessenberg columns (thread q's own entries of Hc)
        double2 a = cz(0, 0);
        if (ok) {
            a = j + 1 < m ? R[q * MX + j + 1] : P[q];
            a = cadd(a, cscale(sh.s[j], R[q * MX + j]));
            for (int k = 0; k < j; ++k) a = csub(a, cmul(Hc[k * MX + q], R[k * MX + j]));
            a = cscale(1.0 / R[j * MX + j].x, a);
        }
        Hc[j * MX + q] = a;
        hess[((int64_t)j * m + q) * Mp + r] = a;
    }
}

// k_fold for the coefficient sets C[o][m][Mp], o = blockIdx.z: D[o][x] =
// [x < K] C[o][x] + sum_j coef[x*nf + j] C[o][K+j].  Grid (lanes / 64, K + nf, sets).
__global__ void k_fold_sets(int K, int nf, int m, int Mp, const double2 *__restrict__ C,
                            const double2 *__restrict__ coef, double2 *__restrict__ D)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x, x = blockIdx.y;
    if (r >= Mp) return;
    const double2 *Co = C + (int64_t)blockIdx.z * m * Mp;
    double2 *Do = D + (int64_t)blockIdx.z * m * Mp;
    double2 acc = x < K ? Co[(int64_t)x * Mp + r] : cz(0, 0);
    for (int j = 0; j < nf; ++j) acc = cfma(coef[((int64_t)x * nf + j) * Mp + r], Co[(int64_t)(K + j) * Mp + r], acc);
    Do[(int64_t)x * Mp + r] = acc;
}

// psi's coordinates when the last block (basis columns K .. K+nf-1) was left
// as its chain vectors y_0 .. y_{nf-1} in those slots: q_{K+j} = sum_x
// coef[x*nf + j] x_x over x = [q_0 .. q_{K-1}, y_0 .. y_{nf-1}] (k_sform's
// coefficients), so sum_k c_k q_k = sum_x d_x x_x with d_x = [x < K] c_x +
// sum_j coef[x*nf + j] c_{K+j}.
// Grid (lanes / 64, K + nf): one thread per lane and coordinate x.
__global__ void k_fold(int K, int nf, int Mp, const double2 *__restrict__ c, const double2 *__restrict__ coef,
                       double2 *__restrict__ d)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x, x = blockIdx.y;
    if (r >= Mp) return;
    double2 acc = x < K ? c[(int64_t)x * Mp + r] : cz(0, 0);
    for (int j = 0; j < nf; ++j) acc = cfma(coef[((int64_t)x * nf + j) * Mp + r], c[(int64_t)(K + j) * Mp + r], acc);
    d[(int64_t)x * Mp + r] = acc;
}

// The Newton block (SPC_NEWTON=0 keeps the monomial blocks; read at every
// rebuild).
bool newton_enabled()
{
    const char *e = std::getenv("SPC_NEWTON");
    return !(e && std::strcmp(e, "0") == 0);
}

// The Newton dots pass: SPC_NGRAM = tiles (a workgroup per tile pair,
// k_ngram), rows2 (default) or rows3 (each row read once per lane, 2 or 3
// rows in flight, k_ngram_rows); read at every rebuild.  N = 36 864, 64-lane
// chunk, m = 24: 483, 291 and 333 us (rows3 holds 256 VGPRs: one wave per
// SIMD); 5.47, 4.58 and 4.61 ms per step at M = 256.
int ngram_form()
{
    const char *e = std::getenv("SPC_NGRAM");
    if (e && std::strcmp(e, "tiles") == 0) return 0;
    if (e && std::strcmp(e, "rows3") == 0) return 3;
    return 2;
}

// Row blocks of the Newton dots pass: a multiple of 8, at most kNRowBlocks and
// sstep_row_blocks(N), as many as the partials buffer holds; 0: none fit.
int newton_nrb(const KrylovWork &w)
{
    if (w.m > kMaxKrylov) return 0;
    const size_t npk = (size_t)(w.m + 1) * (size_t)(w.m + 2) / 2 * (w.gbins ? 4 : 1);
    const size_t fit = w.part_capacity / (npk * (size_t)w.Mp);
    return (int)std::min<size_t>((size_t)std::min(sstep_row_blocks(w.N), kNRowBlocks), fit & ~(size_t)7);
}

hipError_t arnoldi_newton(KrylovWork &w, int nrb, hipStream_t s)
{
    const int m = w.m, nd = m + 1, npk = nd * (nd + 1) / 2;
    const int64_t qs = qstride(w);
    double2 *const u0 = w.u, *const v0 = w.v;
    ColPtrs x{};
    x.p[0] = u0;
    hipError_t e = hipSuccess;
    for (int j = 0; j < m; ++j) { // y_1 .. y_{m-1} in slots 1 .. m-1, y_m in v
        double2 *const out = j == m - 1 ? v0 : w.Q + (int64_t)(j + 1) * qs;
        w.u = j == 0 ? u0 : w.Q + (int64_t)j * qs;
        w.v = out;
        w.no_dots = true;
        w.sigma = w.shifts[j];
        e = spmm_launch(w, s);
        w.u = u0;
        w.v = v0;
        w.no_dots = false;
        w.sigma = 0.0;
        if (e != hipSuccess) return e;
        x.p[j + 1] = out;
    }
    const int nt = (nd + kNTile - 1) / kNTile, npairs = nt * (nt + 1) / 2;
    using RowsFn = void (*)(const double2 *, const double2 *, int64_t, const double2 *, int, double2 *, int64_t, int,
                            int64_t, int, int, int, int);
    NShifts sh{};
    for (int j = 0; j < m; ++j) sh.s[j] = w.shifts[j];
    double2 *const coef = w.sbuf + (size_t)kNPacked * w.Mp;
    if (w.gbins) { // several steps per basis: the four projector bins, their sum the Gram matrix
        const int L = npairs <= 16 ? 4 : 2;
        if (npairs > 32 || w.Mp % L) return hipErrorInvalidValue;
        const RowsFn fns[2][4] = {{k_ngram_rows<kNTile, 4, 2, 0>, k_ngram_rows<kNTile, 4, 2, 1>,
                                   k_ngram_rows<kNTile, 4, 2, 2>, k_ngram_rows<kNTile, 4, 2, 3>},
                                  {k_ngram_rows<kNTile, 2, 2, 0>, k_ngram_rows<kNTile, 2, 2, 1>,
                                   k_ngram_rows<kNTile, 2, 2, 2>, k_ngram_rows<kNTile, 2, 2, 3>}};
        {
            // bins 0, 1 read rows 4n+1, 4n+2 of each vector, bins 2, 3 one row each: 1.5 passes
            ProfScope ps(w, PF_GRAM, 1.5 * nd * vec_bytes(w), s);
            const int64_t rpb = ceil_div(w.Z, nrb);
            for (int x = 0; x < 4; ++x)
                hipLaunchKernelGGL(fns[L == 4 ? 0 : 1][x], dim3(nrb * (w.Mp / L)), dim3(kBlock), 0, s, u0, w.Q, qs, v0,
                                   nd, w.part, w.Z, w.Mp, rpb, nrb, nt, 4 * npk, x * npk);
        }
        if ((e = hipGetLastError()) != hipSuccess) return e;
        if ((e = reduce(w, 4 * npk, nrb, w.gbins, s)) != hipSuccess) return e;
        {
            ProfScope ps(w, PF_COEFFS, 0.0, s);
            const int64_t n = (int64_t)npk * w.Mp;
            hipLaunchKernelGGL(k_sum_bins, dim3((unsigned)ceil_div(n, 256)), dim3(256), 0, s, n, n, w.gbins, w.sbuf);
            hipLaunchKernelGGL(k_ncoef, dim3(w.Mp), dim3(kWave), 0, s, m, w.M, w.Mp, sh, w.sbuf, coef, w.hess,
                               w.beta0, w.flags, w.cond);
        }
        if ((e = hipGetLastError()) != hipSuccess) return e;
        // the step loop evaluates Re(d^H (Y^H P_X Y) d), d = R^-1 c, on the bins
        w.gram_ready = true;
        w.step_G = w.gbins;
        w.step_gxs = (int64_t)npk * w.Mp;
        w.step_nd = nd;
        w.step_Ri = coef;
        w.fold_K = 0;
        w.fold_nf = m;
        w.fold_coef = coef;
        w.fold_q0 = u0;
        return hipSuccess;
    }
    {
        ProfScope ps(w, PF_SDOTS, nd * vec_bytes(w), s);
        const int64_t rpb = ceil_div(w.N, nrb);
        const int form = ngram_form();
        if (form == 0) {
            hipLaunchKernelGGL(k_ngram<kNTile>, dim3(nrb * (w.Mp / kWave) * npairs), dim3(kBlock), 0, s, x, nd, w.part,
                               w.N, w.Mp, rpb, nrb, nt);
        } else {
            const int L = npairs <= 16 ? 4 : 2; // tile pairs per lane: 64 / L threads
            if (npairs > 32 || w.Mp % L) return hipErrorInvalidValue;
            const RowsFn fn = L == 4 ? (form == 2 ? k_ngram_rows<kNTile, 4, 2, kBinAll> : k_ngram_rows<kNTile, 4, 3, kBinAll>)
                                     : (form == 2 ? k_ngram_rows<kNTile, 2, 2, kBinAll> : k_ngram_rows<kNTile, 2, 3, kBinAll>);
            hipLaunchKernelGGL(fn, dim3(nrb * (w.Mp / L)), dim3(kBlock), 0, s, u0, w.Q, qs, v0, nd, w.part, w.N, w.Mp,
                               rpb, nrb, nt, npk, 0);
        }
    }
    if ((e = hipGetLastError()) != hipSuccess) return e;
    if ((e = reduce(w, npk, nrb, w.sbuf, s)) != hipSuccess) return e;
    {
        ProfScope ps(w, PF_COEFFS, 0.0, s);
        hipLaunchKernelGGL(k_ncoef, dim3(w.Mp), dim3(kWave), 0, s, m, w.M, w.Mp, sh, w.sbuf, coef, w.hess,
                           w.beta0, w.flags, w.cond);
    }
    if ((e = hipGetLastError()) != hipSuccess) return e;
    w.fold_K = 0;
    w.fold_nf = m;
    w.fold_coef = coef;
    w.fold_q0 = u0;
    return hipSuccess;
}

} // namespace

int sstep_row_blocks(int64_t N) { return (int)std::clamp<int64_t>((N / 128 + 7) & ~(int64_t)7, 8, 256); }

size_t sstep_partials(int64_t N, int m, int S, int Mp)
{
    const size_t mono = (size_t)sstep_row_blocks(N) * (size_t)(m + 1) * (size_t)std::min(S + 1, kSMaxNY);
    const size_t newton = (size_t)std::min(sstep_row_blocks(N), kNRowBlocks) * (size_t)((m + 1) * (m + 2) / 2);
    return std::max(mono, newton) * (size_t)Mp;
}

// Elements between consecutive basis vectors: N Mp plus a pad of 256 bytes
// (SPC_BASIS_PAD_BYTES sets it; read once).  Without it the vectors of the
// 64-lane workload lie 9 x 4 MiB apart, so the same row of every vector falls
// on the same HBM channels, and the passes that read 25 vectors in 64-byte
// runs (the Newton Gram pass) queue on them: 295 -> 240 us per Gram pass and
// 4.33-4.39 -> 4.21-4.27 ms per step at N = 36 864, M = 256 (MI355X; pads of
// 128-768 bytes alike, 1-3 KiB 255 us).
int64_t basis_stride(int64_t N, int Mp)
{
    static int64_t pad = -1;
    if (pad < 0) {
        const char *e = std::getenv("SPC_BASIS_PAD_BYTES");
        pad = e ? std::max<long long>(0, std::atoll(e)) : 256;
    }
    return N * (int64_t)Mp + (pad + 15) / 16;
}

size_t partials_needed(const KrylovWork &w)
{
    const size_t nrb1 = nrb_for(w.Z, w.rpw1), nrb2 = ceil_div(w.Z, w.rpw2), nrbp = nrb_for(w.Z, w.rpwp);
    size_t need = nrb1 * 2 * (size_t)w.Mp;
    need = std::max(need, nrb2 * (size_t)(2 * w.m) * w.Mp);
    need = std::max(need, nrbp * (size_t)2 * w.Mp);
    return need;
}

hipError_t arnoldi(KrylovWork &w, hipStream_t s)
{
    static_assert(2 * kSMaxCols * kSMaxNY <= kSStepBuf, "s-step buffer size");
    static_assert(kNPacked + kMaxKrylov * kMaxKrylov <= kSStepBuf, "Newton buffer size");
    w.fold_K = w.fold_nf = 0;
    w.fold_coef = w.fold_q0 = nullptr;
    w.gram_ready = false;
    // a Newton block where a basis serves one step, or several with the
    // projector bins (the Grams then come from the chain)
    if ((w.fold_psi || w.gbins) && w.shifts && w.sstep >= 2 && w.sbuf && w.m >= 2 && newton_enabled()) {
        const int nrb = newton_nrb(w);
        if (nrb >= 8) return arnoldi_newton(w, nrb, s);
    }
    if (w.sstep >= 2 && w.sbuf && w.m >= 8) {
        const int S = std::min(w.sstep, kSMaxNY - 1);
        const int nrb = sstep_nrb(w, S);
        if (nrb >= 8) return arnoldi_sstep(w, S, nrb, s);
    }
    // u_0 = psi: w.u aliases w.psi (the update overwrites it in place; psi is
    // rebuilt from Q c after the step)
    const int nch = w.Mp / kWave;
    const int nrb1 = spmm_nrb(w), nrb2 = (int)ceil_div(w.Z, w.rpw2);
    hipError_t e = hipSuccess;
    for (int i = 0; i < w.m; ++i) {
        if ((e = spmm_launch(w, s)) != hipSuccess) return e;
        if ((e = reduce(w, 2, nrb1, w.dots1, s)) != hipSuccess) return e;
        {
            ProfScope ps(w, PF_COEFFS, 0.0, s);
            const size_t lds = sizeof(double2) * kCoefLanes * (size_t)(w.m * (w.m + 3) / 2);
            hipLaunchKernelGGL(k_coeffs, dim3(w.Mp / kCoefLanes), dim3(kBlock), lds, s, i, w.m, w.M, w.Mp, w.kS, w.kT, w.dots1,
                               w.dots, w.coef, w.tprev, w.hess, w.beta0, w.flags);
        }
        if ((e = hipGetLastError()) != hipSuccess) return e;
        const bool last = i == w.m - 1;
        {
            ProfScope ps(w, PF_PASS2, (i + (last ? 3 : 4)) * vec_bytes(w), s);
            hipLaunchKernelGGL(last ? kUpdateLast[i] : kUpdate[i], dim3(nrb2 * nch), dim3(kBlock), 0, s, w.Q, w.Q,
                               w.Q + (int64_t)i * qstride(w), w.u, w.v, w.coef, w.part, w.Z, w.Mp, w.rpw2, nrb2,
                               qstride(w));
        }
        if ((e = hipGetLastError()) != hipSuccess) return e;
        if ((e = reduce(w, last ? i + 1 : 2 * (i + 1), nrb2, w.dots, s)) != hipSuccess) return e;
    }
    hipLaunchKernelGGL(k_finalize, dim3((unsigned)ceil_div(w.Mp, 64)), dim3(64), 0, s, w.m, w.M, w.Mp, w.dots,
                       w.tprev, w.hess);
    return hipGetLastError();
}

hipError_t spmm_only(KrylovWork &w, hipStream_t s) { return spmm_launch(w, s); }

hipError_t fold_sets(KrylovWork &w, const double2 *C, int nout, const double2 **Cuse, const double2 **q0,
                     hipStream_t s)
{
    *Cuse = C;
    *q0 = w.Q;
    if (w.fold_nf == 0) return hipSuccess;
    if (w.fold_K + w.fold_nf != w.m || !w.fold_coef || nout < 1 || (size_t)nout * w.m > (size_t)kNPacked)
        return hipErrorInvalidValue;
    double2 *const D = w.sbuf; // the dots area (psi_update's own fold comes after the last use)
    hipLaunchKernelGGL(k_fold_sets, dim3((unsigned)ceil_div(w.Mp, 64), w.m, nout), dim3(64), 0, s, w.fold_K, w.fold_nf,
                       w.m, w.Mp, C, w.fold_coef, D);
    *Cuse = D;
    if (w.fold_q0) *q0 = w.fold_q0;
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

hipError_t expm_coeffs(KrylovWork &w, const TauSet &ts, double2 *C, hipStream_t s)
{
    const size_t lds = expm_lds_bytes(w.m);
    ProfScope ps(w, PF_EXPM, 0.0, s);
    hipLaunchKernelGGL(k_expm_coeffs, dim3(w.Mp, ts.n), dim3(kBlock), lds, s, w.m, w.Mp, w.M, ts, w.hess, w.beta0,
                       C);
    return hipGetLastError();
}

// Projector Grams of a formed basis q_0 .. q_{m-1} (column and monomial
// forms): the four bins of k_ngram_rows over the m basis vectors, G_X =
// Q^H P_X Q packed (convert_to_krylov, qm_dynamics.f90:469-498; Re(c^H G_X c)
// equals Re(c^H P_X c) there, whose P_S, P_T0 differ from G_X only by
// anti-Hermitian parts).  After a Newton block its own bins are used.
hipError_t gram_projectors(KrylovWork &w, hipStream_t s)
{
    if (w.gram_ready) { // the Newton block left its bins and R^-1 for the step loop
        w.gram_ready = false;
        return hipSuccess;
    }
    const int m = w.m, nd = m, npk = nd * (nd + 1) / 2;
    const int nt = (nd + kNTile - 1) / kNTile, npairs = nt * (nt + 1) / 2;
    const int L = npairs <= 16 ? 4 : 2;
    if (!w.gbins || npairs > 32 || w.Mp % L || m < 1) return hipErrorInvalidValue;
    const size_t per = (size_t)4 * npk * w.Mp;
    const int nrb = (int)std::min<size_t>({(size_t)std::min(sstep_row_blocks(w.N), kNRowBlocks), w.part_capacity / per});
    if (nrb < 1) return hipErrorInvalidValue;
    const int64_t qs = qstride(w);
    using RowsFn = void (*)(const double2 *, const double2 *, int64_t, const double2 *, int, double2 *, int64_t, int,
                            int64_t, int, int, int, int);
    const RowsFn fns[2][4] = {{k_ngram_rows<kNTile, 4, 2, 0>, k_ngram_rows<kNTile, 4, 2, 1>,
                               k_ngram_rows<kNTile, 4, 2, 2>, k_ngram_rows<kNTile, 4, 2, 3>},
                              {k_ngram_rows<kNTile, 2, 2, 0>, k_ngram_rows<kNTile, 2, 2, 1>,
                               k_ngram_rows<kNTile, 2, 2, 2>, k_ngram_rows<kNTile, 2, 2, 3>}};
    {
        ProfScope ps(w, PF_GRAM, 1.5 * nd * vec_bytes(w), s);
        const int64_t rpb = ceil_div(w.Z, nrb);
        for (int x = 0; x < 4; ++x)
            hipLaunchKernelGGL(fns[L == 4 ? 0 : 1][x], dim3(nrb * (w.Mp / L)), dim3(kBlock), 0, s, w.Q, w.Q, qs,
                               w.Q + (int64_t)(m - 1) * qs, nd, w.part, w.Z, w.Mp, rpb, nrb, nt, 4 * npk, x * npk);
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    if ((e = reduce(w, 4 * npk, nrb, w.gbins, s)) != hipSuccess) return e;
    w.step_G = w.gbins;
    w.step_gxs = (int64_t)npk * w.Mp;
    w.step_nd = nd;
    w.step_Ri = nullptr;
    return hipSuccess;
}

hipError_t step_lanes(KrylovWork &w, double dt, double t_end, double tol, bool gram, hipStream_t s)
{
    ProfScope ps(w, PF_STEP, 0.0, s);
    if (gram) {
        if (!w.step_G) return hipErrorInvalidValue;
        hipLaunchKernelGGL(k_step_proj, dim3(w.Mp), dim3(kWave), 0, s, w.m, w.M, w.Mp, dt, t_end, tol, w.E, w.beta0,
                           w.c, w.lane_t, w.lane_k, w.obs, w.n_out, w.step_G, w.step_gxs, w.step_nd, w.step_Ri,
                           w.flags);
    } else {
        hipLaunchKernelGGL(k_step, dim3((unsigned)ceil_div(w.Mp, 64)), dim3(64), 0, s, w.m, w.M, w.Mp, dt, t_end, tol,
                           w.E, w.beta0, w.c, w.lane_t, w.lane_k, w.n_out, w.flags);
    }
    return hipGetLastError();
}

hipError_t psi_update(KrylovWork &w, bool observe, hipStream_t s)
{
    const int nch = w.Mp / kWave;
    const int nrbp = nrb_for(w.Z, w.rpwp);
    const PsiFn fn = observe ? kPsiObs[w.m - 1] : kPsiPlain[w.m - 1];
    const double2 *cvec = w.c;
    hipError_t e = hipSuccess;
    const double2 *q0 = w.Q;
    if (w.fold_nf > 0) { // the last block unformed: psi's coordinates over [Q_K, Y] (the dots area is free)
        if (w.fold_K + w.fold_nf != w.m || !w.fold_coef) return hipErrorInvalidValue;
        double2 *const d = w.sbuf;
        hipLaunchKernelGGL(k_fold, dim3((unsigned)ceil_div(w.Mp, 64), w.m), dim3(64), 0, s, w.fold_K, w.fold_nf, w.Mp,
                           w.c, w.fold_coef, d);
        if ((e = hipGetLastError()) != hipSuccess) return e;
        cvec = d;
        if (w.fold_q0) q0 = w.fold_q0;
        w.fold_K = w.fold_nf = 0;
        w.fold_coef = w.fold_q0 = nullptr;
    }
    {
        ProfScope ps(w, PF_PSI, (w.m + 1) * vec_bytes(w), s);
        hipLaunchKernelGGL(fn, dim3(nrbp * nch), dim3(kBlock), 0, s, w.Q, q0, cvec, w.psi, w.part, w.Z, w.Mp, w.rpwp,
                           nrbp, qstride(w));
    }
    e = hipGetLastError();
    if (e != hipSuccess || !observe) return e;
    if ((e = reduce(w, 2, nrbp, w.dots1, s)) != hipSuccess) return e;
    hipLaunchKernelGGL(k_obs_store, dim3((unsigned)ceil_div(w.M, 64)), dim3(64), 0, s, w.M, w.Mp, w.dots1, w.lane_k,
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

namespace {
__global__ void k_obs_groups(const double *__restrict__ obs, int64_t n_out, int groups, int group, int M_g, int Mp,
                             double *__restrict__ acc)
{
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= 4 * n_out * groups) return;
    const int g = (int)(idx / (4 * n_out));
    const int64_t xk = idx % (4 * n_out);
    const int x = (int)(xk / n_out);
    const int64_t k = xk % n_out;
    double s = acc[idx];
    for (int r = 0; r < M_g; ++r) s += obs[(k * 4 + x) * Mp + (int64_t)g * group + r];
    acc[idx] = s;
}
} // namespace

hipError_t obs_reduce_groups(const double *obs, int64_t n_out, int groups, int group, int M_g, int Mp, double *acc,
                             hipStream_t s)
{
    hipLaunchKernelGGL(k_obs_groups, dim3((unsigned)ceil_div(4 * n_out * groups, 256)), dim3(256), 0, s, obs, n_out,
                       groups, group, M_g, Mp, acc);
    return hipGetLastError();
}

hipError_t obs_reduce_lanes(const double *obs, int64_t n_out, int M, int Mp, double *acc, hipStream_t s)
{
    hipLaunchKernelGGL(k_obs_lanes, dim3((unsigned)ceil_div(4 * n_out, 256)), dim3(256), 0, s, obs, n_out, M, Mp, acc);
    return hipGetLastError();
}

} // namespace spc
