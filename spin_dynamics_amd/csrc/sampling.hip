// sampling.hip -- trace vectors psi_r = theta (x) z_r, written lane-fastest in
// the device basis order (psi[(4*n + e)*Mp + r]).
//
// Kinds: SU(Z) coherent states (sample_SUZ, qm_dynamics.f90:416-430: complex
// normals re then im from the Marsaglia polar method, normalised), random
// phase ((x + i y)/sqrt(Z) with (x, y) = rng%circle(1), m_random.f90:280-297),
// caller-injected vectors, and nuclear basis states (exact_dynamics,
// qm_dynamics.f90:82-86).  Each lane owns one xoroshiro128+ stream seeded with
// set_seed(seed_r) (m_random.f90:90-98), as the OpenMP loop does per sample
// (qm_dynamics.f90:326-327).
#include "common.hpp"
#include "internal.hpp"

namespace spc {

namespace {

// Sampled vectors (kinds 0, 1): one thread per lane, as its stream is
// sequential (the polar method rejects pairs).  Padding lanes r >= M are
// zero.  Kind 0 draws the stream twice, once for the norm and once for the
// normalised values (the same numbers in the same order), writing every
// element once.  Kind 1 keeps only the sequential part here: the accepted
// pair (a, b) of each element goes to its alpha-alpha row, and
// k_init_phases turns the pairs into the vectors in parallel (a wave alone
// on its SIMD runs a dependent chain of FP64 divisions at full latency).
__global__ void k_init_sampled(VecInit vi, int M, int Mp, int64_t Z, double2 *__restrict__ psi)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= Mp) return;
    const int64_t ld = Mp;
    if (r >= M) {
        for (int64_t n = 0; n < 4 * Z; ++n) psi[n * ld + r] = cz(0, 0);
        return;
    }
    Rng g;
    if (vi.kind == 1) {
        // one try per iteration, each lane at its own element: the wave runs
        // as many iterations as its slowest lane needs tries in all, not the
        // sum over elements of the slowest lane's tries
        rng_set_seed(g, vi.seeds[2 * r], vi.seeds[2 * r + 1]);
        int64_t n = 0;
        while (n < Z) {
            const double a = 2.0 * rng_unif01(g) - 1.0;
            const double b = 2.0 * rng_unif01(g) - 1.0;
            if (a * a + b * b <= 1.0) {
                psi[(4 * n) * ld + r] = cz(a, b);
                ++n;
            }
        }
        return;
    }
    rng_set_seed(g, vi.seeds[2 * r], vi.seeds[2 * r + 1]);
    double nrm2 = 0.0;
    for (int64_t n = 0; n < Z; ++n) {
        const double x = rng_normal(g);
        const double y = rng_normal(g);
        nrm2 += x * x + y * y;
    }
    const double nrm = sqrt(nrm2);
    rng_set_seed(g, vi.seeds[2 * r], vi.seeds[2 * r + 1]);
    for (int64_t n = 0; n < Z; ++n) {
        const double x = rng_normal(g);
        const double y = rng_normal(g);
        const double2 z = cz(x / nrm, y / nrm);
        for (int e = 0; e < 4; ++e) psi[(4 * n + e) * ld + r] = cmul(vi.theta[e], z);
    }
}

// Random-phase vectors from the accepted pairs (kind 1): z = (x + i y)/sqrt(Z)
// with (x, y) = rng%circle's point for the pair (a, b); one thread per
// (nuclear index, lane) of the M sampled lanes.
__global__ void k_init_phases(VecInit vi, int M, int Mp, int64_t Z, double2 *__restrict__ psi)
{
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= Z * Mp) return;
    const int64_t n = idx / Mp;
    const int r = (int)(idx % Mp);
    if (r >= M) return;
    const double2 ab = psi[(4 * n) * Mp + r];
    double x, y;
    circle_point(ab.x, ab.y, x, y);
    const double f = 1.0 / sqrt((double)Z);
    const double2 z = cz(x * f, y * f);
    for (int e = 0; e < 4; ++e) psi[(4 * n + e) * Mp + r] = cmul(vi.theta[e], z);
}

// Injected vectors and basis states (kinds 2, 3): one thread per (nuclear
// index, lane), lanes fastest; padding lanes are zero.
__global__ void k_init_given(VecInit vi, int M, int Mp, int64_t Z, double2 *__restrict__ psi)
{
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= Z * Mp) return;
    const int64_t n = idx / Mp;
    const int r = (int)(idx % Mp);
    double2 z = cz(0.0, 0.0);
    if (r < M) {
        if (vi.kind == 2)
            z = vi.inj[(int64_t)r * Z + n];
        else if (n == vi.basis_offset + r)
            z = cz(1.0, 0.0);
    }
    for (int e = 0; e < 4; ++e) psi[(4 * n + e) * Mp + r] = cmul(vi.theta[e], z);
}

} // namespace

hipError_t init_vectors(const VecInit &vi, int M, int Mp, int64_t Z, double2 *psi, hipStream_t s)
{
    if (vi.kind == 0 || vi.kind == 1) {
        const int threads = 64;
        hipLaunchKernelGGL(k_init_sampled, dim3((Mp + threads - 1) / threads), dim3(threads), 0, s, vi, M, Mp, Z,
                           psi);
        if (vi.kind == 1)
            hipLaunchKernelGGL(k_init_phases, dim3((unsigned)((Z * Mp + 255) / 256)), dim3(256), 0, s, vi, M, Mp, Z,
                               psi);
    } else {
        const int threads = 256;
        hipLaunchKernelGGL(k_init_given, dim3((unsigned)((Z * Mp + threads - 1) / threads)), dim3(threads), 0, s, vi,
                           M, Mp, Z, psi);
    }
    return hipGetLastError();
}

} // namespace spc
