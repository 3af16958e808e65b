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

__global__ void k_init_vectors(VecInit vi, int M, int Mp, int64_t Z, double2 *__restrict__ psi)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= Mp) return;
    const int64_t ld = Mp;
    if (r >= M) {
        for (int64_t n = 0; n < 4 * Z; ++n) psi[n * ld + r] = cz(0, 0);
        return;
    }
    if (vi.kind == 0 || vi.kind == 1) {
        Rng g;
        rng_set_seed(g, vi.seeds[2 * r], vi.seeds[2 * r + 1]);
        double nrm2 = 0.0;
        const double f = 1.0 / sqrt((double)Z);
        for (int64_t n = 0; n < Z; ++n) {
            double2 z;
            if (vi.kind == 0) {
                z.x = rng_normal(g);
                z.y = rng_normal(g);
                nrm2 += z.x * z.x + z.y * z.y;
            } else {
                double x, y;
                rng_circle(g, x, y);
                z = cz(x * f, y * f);
            }
            psi[(4 * n) * ld + r] = z; // staged in the alpha-alpha row
        }
        const double nrm = vi.kind == 0 ? sqrt(nrm2) : 1.0;
        for (int64_t n = 0; n < Z; ++n) {
            double2 z = psi[(4 * n) * ld + r];
            if (vi.kind == 0) z = cz(z.x / nrm, z.y / nrm);
            for (int e = 0; e < 4; ++e) psi[(4 * n + e) * ld + r] = cmul(vi.theta[e], z);
        }
        return;
    }
    for (int64_t n = 0; n < Z; ++n) {
        double2 z;
        if (vi.kind == 2)
            z = vi.inj[(int64_t)r * Z + n];
        else
            z = (n == vi.basis_offset + r) ? cz(1.0, 0.0) : cz(0.0, 0.0);
        for (int e = 0; e < 4; ++e) psi[(4 * n + e) * ld + r] = cmul(vi.theta[e], z);
    }
}

} // namespace

hipError_t init_vectors(const VecInit &vi, int M, int Mp, int64_t Z, double2 *psi, hipStream_t s)
{
    int threads = 64;
    int blocks = (Mp + threads - 1) / threads;
    hipLaunchKernelGGL(k_init_vectors, dim3(blocks), dim3(threads), 0, s, vi, M, Mp, Z, psi);
    return hipGetLastError();
}

} // namespace spc
