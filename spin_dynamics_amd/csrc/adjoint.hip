// adjoint.hip -- kernels of the reverse-mode gradient of the singlet yield.
//
// The forward map is psi_{n+1} = U psi_n, U = exp(X), X = -i dt H, for every
// trace vector (lane), and F = sum_n omega_n psi_n^H P_S psi_n (the Simpson
// yield of the averaged singlet population).  Its adjoint lambda_n =
// omega_n P_S psi_n + U^H lambda_{n+1} runs backwards through the same Krylov
// machinery on H' = -H^H (U^H = exp(-i dt H')), and
//   dF/dtheta = 2 dt Im sum_n int_0^1 b_n(s)^H (dH/dtheta) a_n(s) ds,
//   a_n(s) = exp(s X) psi_n,  b_n(s) = exp((1-s) X^H) lambda_{n+1},
// with the s integral by Gauss-Legendre: both a and b come from one Krylov
// basis each (Q exp(-i tau A) e_1 beta), so they cost one streaming pass over
// the basis (combine).  The bilinear forms are gathered as a sampled dense
// product over the nuclear neighbour pattern of H (grad_density), accumulated
// over all steps; the contraction with dH/dtheta runs once on the host.
#include <algorithm>
#include <array>
#include <utility>

#include "common.hpp"
#include "internal.hpp"

namespace spc {

namespace {

constexpr int kBlock = 256;
constexpr int kGroup = 16; // lanes per row segment in the combine pass

inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

// A wave covers 4 device rows (one nuclear index) x 16 lanes; the block's 4
// waves take nuclear indices base + wave, base + wave + 4, ...  The per-lane
// coefficients of the 16 lanes sit in LDS ([o][k][lane]), read as broadcasts.
// Vector 0 is read from q0 (Q itself, or an output updated in place: an
// element of q0 is read only by the thread that writes it, before it does).
template <int MK>
__global__ __launch_bounds__(kBlock) void k_combine(const double2 *__restrict__ Q, const double2 *q0, int64_t qs,
                                                     const double2 *__restrict__ C, CombineOut out,
                                                     const double2 *__restrict__ src, double omega, int64_t Z, int M,
                                                     int Mp, int64_t rows_per_block)
{
    __shared__ double2 cs[kCombineOut][MK][kGroup];
    const int ngroups = Mp / kGroup;
    const int lg = blockIdx.x % ngroups;
    const int64_t rb = blockIdx.x / ngroups;
    const int lr = threadIdx.x & (kGroup - 1), e = (threadIdx.x >> 4) & 3, wave = threadIdx.x >> 6;
    const int r = lg * kGroup + lr;
    const int nout = out.n;
    for (int idx = threadIdx.x; idx < nout * MK * kGroup; idx += kBlock) {
        const int o = idx / (MK * kGroup), k = (idx / kGroup) % MK, l = idx % kGroup;
        cs[o][k][l] = C[((int64_t)o * MK + k) * Mp + lg * kGroup + l];
    }
    __syncthreads();
    const int64_t NMp = qs; // basis vector stride
    const int64_t n0 = rb * rows_per_block, n1 = n0 + rows_per_block < Z ? n0 + rows_per_block : Z;
    const bool live = r < M;
    for (int64_t n = n0 + wave; n < n1; n += 4) {
        const int64_t off = (4 * n + e) * (int64_t)Mp + r;
        double2 acc[kCombineOut];
#pragma unroll
        for (int o = 0; o < kCombineOut; ++o) acc[o] = cz(0, 0);
#pragma unroll
        for (int k = 0; k < MK; ++k) {
            const double2 q = k == 0 ? q0[off] : Q[k * NMp + off];
#pragma unroll
            for (int o = 0; o < kCombineOut; ++o)
                if (o < nout) acc[o] = cfma(cs[o][k][lr], q, acc[o]);
        }
        if (src && (e == 1 || e == 2)) {
            const double2 d = out.src_diff ? src[n * (int64_t)Mp + r]
                                           : csub(src[(4 * n + 1) * (int64_t)Mp + r], src[(4 * n + 2) * (int64_t)Mp + r]);
            const double f = (e == 1 ? 0.5 : -0.5) * omega;
            acc[0] = cadd(acc[0], cscale(f, d));
        }
        if (out.diff0) { // row 4n+2's value (lane group e = 2) to the row 4n+1 thread (e = 1) of the same lane
            const double2 o2 = cz(__shfl_down(acc[0].x, kGroup), __shfl_down(acc[0].y, kGroup));
            if (e == 1 && out.y[0]) out.y[0][n * (int64_t)Mp + r] = live ? csub(acc[0], o2) : cz(0, 0);
        }
#pragma unroll
        for (int o = 0; o < kCombineOut; ++o)
            if (o < nout && out.y[o] && !(o == 0 && out.diff0)) out.y[o][off] = live ? acc[o] : cz(0, 0);
    }
}

__global__ void k_singlet_source(const double2 *__restrict__ psi, double omega, double2 *__restrict__ lam, int64_t Z,
                                 int Mp)
{
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= Z * Mp) return;
    const int64_t n = idx / Mp, r = idx % Mp;
    const double2 d = csub(psi[(4 * n + 1) * Mp + r], psi[(4 * n + 2) * Mp + r]);
    const double2 x = cscale(0.5 * omega, d);
    lam[(4 * n + 0) * Mp + r] = cz(0, 0);
    lam[(4 * n + 1) * Mp + r] = x;
    lam[(4 * n + 2) * Mp + r] = cz(-x.x, -x.y);
    lam[(4 * n + 3) * Mp + r] = cz(0, 0);
}

// One wave per nuclear index n, lanes = trace vectors (chunks of 64).  For
// each neighbour slot the 4x4 electron block sum_j sum_lanes conj(b_j) a_j is
// formed per lane, then summed over the 64 lanes through LDS (16 values at a
// time, each summed by 4 lanes of 16 in a fixed order) and added to S.  The
// slot's old values of S are loaded before its products, and with one lane
// chunk and all JB nodes (the gradient's case) the rows of every node are
// loaded at once: two memory round trips per slot instead of JB + 1 (the
// 5000-step gradient at N = 36 864, M = 64: 19.24 -> 18.83 s of device time;
// loading the next slot's rows during this slot's reduction, at 256 VGPRs,
// was slower: 19.84 s).
__global__ __launch_bounds__(kBlock) void k_grad_density(const double2 *__restrict__ a, const double2 *__restrict__ b,
                                                          int J, int64_t NMp, const DevSystem *__restrict__ sys, int M,
                                                          int Mp, double2 *__restrict__ S)
{
    // the lane sums' transpose: lane l's value d at red[d][17 (l / 16) + l % 16] (rows of 68, runs of 16 padded
    // to 17), so that neither the writes nor the reads of 16 consecutive values by 4 lanes per row share a bank
    // (unpadded, the 16 rows d lay 512 bytes apart, on the same banks)
    __shared__ double red[kBlock / kWave][16][4 * 17];
    const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x >> 6;
    const int64_t Z = sys->Z;
    const int64_t n = (int64_t)blockIdx.x * (kBlock / kWave) + wave;
    if (n >= Z) return; // whole wave
    const int nn = sys->nn;
    const int nslot = 1 + 2 * nn;
    const int nch = Mp / kWave;
    // b's rows at n for the first lane chunk, kept over the slots
    constexpr int JB = 4;
    double2 b0[JB][4];
    const bool b0_ok = nch == 1 && J <= JB;
    if (b0_ok && lane < M) {
#pragma unroll
        for (int j = 0; j < JB; ++j)
#pragma unroll
            for (int e = 0; e < 4; ++e)
                b0[j][e] = j < J ? b[j * NMp + (4 * n + e) * (int64_t)Mp + lane] : cz(0, 0);
    }
    for (int slot = 0; slot < nslot; ++slot) {
        int64_t ns = n;
        if (slot > 0) {
            const int k = (slot - 1) >> 1;
            const int64_t st = sys->stride[k];
            const int g = sys->mult[k];
            const int d = (int)((uint32_t)n / (uint32_t)st) % g; // n < 2^31 (wave-uniform)
            if (slot & 1) {
                if (d >= g - 1) continue;
                ns = n + st;
            } else {
                if (d == 0) continue;
                ns = n - st;
            }
        }
        double *Sd = reinterpret_cast<double *>(S + ((int64_t)n * nslot + slot) * 16);
        // the slot's old sums, loaded while the products run
        const int dl = lane >> 2, hl = lane & 3;
        double sold[2] = {0.0, 0.0};
        if (hl == 0) {
            sold[0] = Sd[dl];
            sold[1] = Sd[16 + dl];
        }
        double2 P[16];
#pragma unroll
        for (int q = 0; q < 16; ++q) P[q] = cz(0, 0);
        if (b0_ok && J == JB) {
            if (lane < M) { // the rows of all JB nodes loaded before the first product
                double2 aa[JB][4];
#pragma unroll
                for (int j = 0; j < JB; ++j)
#pragma unroll
                    for (int e = 0; e < 4; ++e) aa[j][e] = a[j * NMp + (4 * ns + e) * (int64_t)Mp + lane];
#pragma unroll
                for (int j = 0; j < JB; ++j)
#pragma unroll
                    for (int ep = 0; ep < 4; ++ep)
#pragma unroll
                        for (int e = 0; e < 4; ++e) P[4 * ep + e] = cfma_conj(b0[j][ep], aa[j][e], P[4 * ep + e]);
            }
        } else if (b0_ok) {
            if (lane < M) {
#pragma unroll
                for (int j = 0; j < JB; ++j) {
                    if (j >= J) break;
                    double2 aa[4];
#pragma unroll
                    for (int e = 0; e < 4; ++e) aa[e] = a[j * NMp + (4 * ns + e) * (int64_t)Mp + lane];
#pragma unroll
                    for (int ep = 0; ep < 4; ++ep)
#pragma unroll
                        for (int e = 0; e < 4; ++e) P[4 * ep + e] = cfma_conj(b0[j][ep], aa[e], P[4 * ep + e]);
                }
            }
        } else {
            for (int ch = 0; ch < nch; ++ch) {
                const int r = ch * kWave + lane;
                if (r >= M) continue;
                for (int j = 0; j < J; ++j) {
                    double2 bb[4], aa[4];
#pragma unroll
                    for (int e = 0; e < 4; ++e) {
                        bb[e] = b[j * NMp + (4 * n + e) * (int64_t)Mp + r];
                        aa[e] = a[j * NMp + (4 * ns + e) * (int64_t)Mp + r];
                    }
#pragma unroll
                    for (int ep = 0; ep < 4; ++ep)
#pragma unroll
                        for (int e = 0; e < 4; ++e) P[4 * ep + e] = cfma_conj(bb[ep], aa[e], P[4 * ep + e]);
                }
            }
        }
#pragma unroll
        for (int round = 0; round < 2; ++round) {
#pragma unroll
            for (int d = 0; d < 16; ++d) {
                const int q = (round * 16 + d) >> 1;
                red[wave][d][17 * (lane >> 4) + (lane & 15)] = ((round * 16 + d) & 1) ? P[q].y : P[q].x;
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            const int d = lane >> 2, h = lane & 3;
            double sum = 0.0;
#pragma unroll
            for (int t = 0; t < 16; ++t) sum += red[wave][d][17 * h + t];
            const double o1 = __shfl_xor(sum, 1);
            const double s01 = (h & 1) ? o1 + sum : sum + o1;
            const double o2 = __shfl_xor(s01, 2);
            const double tot = (h & 2) ? o2 + s01 : s01 + o2;
            if (h == 0) Sd[round * 16 + d] = sold[round] + tot;
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        }
    }
}

using CombineFn = void (*)(const double2 *, const double2 *, int64_t, const double2 *, CombineOut, const double2 *,
                           double, int64_t, int, int, int64_t);

template <size_t... I>
constexpr auto make_combine(std::index_sequence<I...>)
{
    return std::array<CombineFn, sizeof...(I)>{&k_combine<(int)I + 1>...};
}
const auto kCombine = make_combine(std::make_index_sequence<kMaxKrylov>{});

} // namespace

hipError_t combine(const double2 *Q, int64_t qs, int m, const double2 *C, const CombineOut &out, const double2 *src,
                   double omega, int64_t Z, int M, int Mp, hipStream_t s, const double2 *q0)
{
    if (m < 1 || m > kMaxKrylov || out.n < 1 || out.n > kCombineOut || Mp % kGroup) return hipErrorInvalidValue;
    const int ngroups = Mp / kGroup;
    // about 2048 blocks: rows per block a multiple of the 4 waves
    int64_t rpb = std::max<int64_t>(4, ceil_div(Z * ngroups, 2048) / 4 * 4);
    const int64_t nrb = ceil_div(Z, rpb);
    hipLaunchKernelGGL(kCombine[m - 1], dim3((unsigned)(nrb * ngroups)), dim3(kBlock), 0, s, Q, q0 ? q0 : Q, qs, C,
                       out, src, omega, Z, M, Mp, rpb);
    return hipGetLastError();
}

hipError_t singlet_source(const double2 *psi, double omega, double2 *lam, int64_t Z, int Mp, hipStream_t s)
{
    hipLaunchKernelGGL(k_singlet_source, dim3((unsigned)ceil_div(Z * Mp, 256)), dim3(256), 0, s, psi, omega, lam, Z,
                       Mp);
    return hipGetLastError();
}

hipError_t grad_density(const double2 *a, const double2 *b, int J, int64_t NMp, const DevSystem *d_sys,
                        const DevSystem &h_sys, int M, int Mp, double2 *S, hipStream_t s)
{
    const int64_t nb = ceil_div(h_sys.Z, kBlock / kWave);
    hipLaunchKernelGGL(k_grad_density, dim3((unsigned)nb), dim3(kBlock), 0, s, a, b, J, NMp, d_sys, M, Mp, S);
    return hipGetLastError();
}

} // namespace spc
