// semiclassical.hip -- semiclassical (Schulten-Wolynes / Manolopoulos-Hore)
// singlet correlation function, one sample per group of 2, 4 or 8 lanes.
//
// Semantics of the reference's src/sc (sc_dynamics.f90:17-100,
// class_radical_pair.f90:147-251), with its defects removed: electron state
// x = [S1, S2, T12 (column-major), 1] (16 values); per step: electron half step
// (x <- exp(dt/2 (K+J)) x, then S1 and the T12 columns precess about
// w1 = -B n + sum_{k in e1} a_k I_k, S2 and the T12 rows about w2), nuclear full
// step (I_k precesses about a_k S_e normalised to sqrt(3)/2), electron half
// step.  P_S = x_16/4 - tr T12 and P_S(t) = 4 <P_S(0) P_S(t)> over samples.
//
// Mapping to CDNA4 (FP64 VALU-bound, all state in VGPRs):
// - A sample's nuclei are split over a group of L lanes (L = 2; 4 or 8 only
//   where 2 lanes cannot hold the nuclei: every lane repeats the electron
//   work, which costs more than the registers a wider split saves -- 0.79 s
//   against 1.09 s and 1.41 s for 2^20 samples of 48 nuclei, 4001 steps, on
//   MI355X); lane h of the group holds the
//   e1 nuclei j = L q + h in slots q < nl1 and the e2 nuclei j = L (q - nl1) + h
//   in slots q >= nl1 (nl1 = ceil(n1/L) is wave-uniform, so the electron a
//   slot couples to is a uniform branch).  The group combines its partial
//   hyperfine fields with a butterfly of shuffles whose additions run in a
//   fixed order, so every lane holds a bit-identical electron state.
// - Nuclear spins are stored scaled, I'_k = a_k I_k: the field is the plain
//   sum of the I'_k, and a rotation commutes with the scaling.  The nuclear
//   precession angle a_k sqrt(3)/2 dt is constant, so its cos/sin are loaded
//   once per lane.
// - exp(dt/2 (K+J)) is block diagonal with four invariant 4x4 blocks, and K+J
//   restricted to (S1x,S2x,T32,T23), (S1y,S2y,T13,T31) and (S1z,S2z,T21,T12)
//   is one and the same 4x4 matrix: 64 FMAs per half step instead of 256, with
//   32 wave-uniform coefficients held in scalar registers.
// - The fields after the nuclear step serve both the second half of this
//   step and the first half of the next, so one field sum and one pair of
//   sincos per step.
// - P_0 P_S(t) is summed over each wave's samples with shuffles and added to
//   part[st][wave] by lane 0: no barriers in the step loop; the grid is
//   persistent and fixed, so the sums are deterministic.
#include <algorithm>

#include "common.hpp"
#include "internal.hpp"

namespace spc {

namespace {

// Positions in x of the three spin blocks (one shared 4x4 generator) and of
// the block (T11, T22, T33, 1).
__device__ constexpr int kSpinBlk[3][4] = {{0, 3, 11, 13}, {1, 4, 12, 8}, {2, 5, 7, 9}};
__device__ constexpr int kDiagBlk[4] = {6, 10, 14, 15};

__device__ __forceinline__ void rotate_cs(double *v, const double *h, double c, double s)
{
    // v_par + c v_perp + s (h x v), h a unit vector
    const double dp = v[0] * h[0] + v[1] * h[1] + v[2] * h[2];
    const double cx = h[1] * v[2] - h[2] * v[1];
    const double cy = h[2] * v[0] - h[0] * v[2];
    const double cz_ = h[0] * v[1] - h[1] * v[0];
    const double r0 = dp * h[0] + c * (v[0] - dp * h[0]) + s * cx;
    const double r1 = dp * h[1] + c * (v[1] - dp * h[1]) + s * cy;
    const double r2 = dp * h[2] + c * (v[2] - dp * h[2]) + s * cz_;
    v[0] = r0;
    v[1] = r1;
    v[2] = r2;
}

// axis and sincos of a precession about field f for time t (identity for f = 0)
struct Rot {
    double h[3], c, s;
};
__device__ __forceinline__ Rot make_rot(const double *f, double t)
{
    Rot r;
    const double w = sqrt(f[0] * f[0] + f[1] * f[1] + f[2] * f[2]);
    const bool z = !(w > 0.0);
    const double iw = 1.0 / (z ? 1.0 : w); // one division for the three components
    r.h[0] = z ? 0.0 : f[0] * iw;
    r.h[1] = z ? 0.0 : f[1] * iw;
    r.h[2] = z ? 1.0 : f[2] * iw;
    double sn, cs;
    sincos(z ? 0.0 : w * t, &sn, &cs);
    r.c = cs;
    r.s = sn;
    return r;
}

template <int NL, int L>
__device__ __forceinline__ void fields(const double *wz, int nl1, const double (&I)[NL][3], double *w1, double *w2)
{
    double f1[3] = {0, 0, 0}, f2[3] = {0, 0, 0};
#pragma unroll
    for (int q = 0; q < NL; ++q) {
        if (q < nl1) {
            f1[0] += I[q][0];
            f1[1] += I[q][1];
            f1[2] += I[q][2];
        } else {
            f2[0] += I[q][0];
            f2[1] += I[q][1];
            f2[2] += I[q][2];
        }
    }
    // butterfly over the group; at each level the lower half's partial is
    // added first, so both partners form the same sum
#pragma unroll
    for (int m = 1; m < L; m <<= 1) {
        const bool low = (threadIdx.x & m) == 0;
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            const double o1 = __shfl_xor(f1[d], m), o2 = __shfl_xor(f2[d], m);
            f1[d] = low ? f1[d] + o1 : o1 + f1[d];
            f2[d] = low ? f2[d] + o2 : o2 + f2[d];
        }
    }
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        w1[d] = wz[d] + f1[d];
        w2[d] = wz[d] + f2[d];
    }
}

// x <- blockdiag(KJ) x, then the precessions of S1, S2 and T12.  kA: the
// spin-block coefficients, kB: the (T11, T22, T33, 1) block, both row-major.
__device__ __forceinline__ void electron_half(double *x, const double *kA, const double *kB, const Rot &r1,
                                              const Rot &r2)
{
#pragma unroll
    for (int b = 0; b < 3; ++b) {
        double y[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            double s = 0.0;
#pragma unroll
            for (int j = 0; j < 4; ++j) s = fma(kA[4 * i + j], x[kSpinBlk[b][j]], s);
            y[i] = s;
        }
#pragma unroll
        for (int i = 0; i < 4; ++i) x[kSpinBlk[b][i]] = y[i];
    }
    {
        double y[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            double s = 0.0;
#pragma unroll
            for (int j = 0; j < 4; ++j) s = fma(kB[4 * i + j], x[kDiagBlk[j]], s);
            y[i] = s;
        }
#pragma unroll
        for (int i = 0; i < 4; ++i) x[kDiagBlk[i]] = y[i];
    }
    rotate_cs(x + 0, r1.h, r1.c, r1.s);
    rotate_cs(x + 3, r2.h, r2.c, r2.s);
#pragma unroll
    for (int b = 0; b < 3; ++b) rotate_cs(x + 6 + 3 * b, r1.h, r1.c, r1.s); // T(:,b) rotates with S1
#pragma unroll
    for (int a = 0; a < 3; ++a) {                                             // T(a,:) rotates with S2
        double v[3] = {x[6 + a], x[9 + a], x[12 + a]};
        rotate_cs(v, r2.h, r2.c, r2.s);
        x[6 + a] = v[0];
        x[9 + a] = v[1];
        x[12 + a] = v[2];
    }
}

__device__ __forceinline__ double ps_of(const double *x) { return 0.25 * x[15] - (x[6] + x[10] + x[14]); }

// ---- forward tangents (d/d eps of the state for a field w(eps) = w + eps dw)
// A rotation by the rotation vector phi (angle |phi| about phi/|phi|) varies
// as d R(phi) = [om]x R(phi), om = J(phi) dphi with the left Jacobian of SO(3)
// J(phi) = 1 + (1 - cos t)/t^2 [phi]x + (t - sin t)/t^3 [phi]x^2, t = |phi|; so
// a rotated vector v' = R v has tangent R dv + om x v'.
__device__ __forceinline__ void cross3(const double *a, const double *b, double *c)
{
    c[0] = a[1] * b[2] - a[2] * b[1];
    c[1] = a[2] * b[0] - a[0] * b[2];
    c[2] = a[0] * b[1] - a[1] * b[0];
}

// om for the electron precession about field w for time t (phi = w t), with
// r = make_rot(w, t); the Jacobian's coefficients by their series below
// |phi| = 0.05 (the closed forms cancel there).
__device__ __forceinline__ void rot_jac(const double *w, const double *dw, double t, const Rot &r, double *om)
{
    const double th2 = (w[0] * w[0] + w[1] * w[1] + w[2] * w[2]) * t * t;
    double A, B;
    if (th2 < 0.0025) {
        A = 0.5 - th2 * (1.0 / 24 - th2 * (1.0 / 720 - th2 * (1.0 / 40320)));
        B = 1.0 / 6 - th2 * (1.0 / 120 - th2 * (1.0 / 5040 - th2 * (1.0 / 362880)));
    } else {
        const double th = sqrt(th2);
        A = (1.0 - r.c) / th2;
        B = (th - r.s) / (th2 * th);
    }
    double c1[3], c2[3];
    cross3(w, dw, c1);
    cross3(w, c1, c2);
    const double a = A * t * t, b = B * t * t * t;
#pragma unroll
    for (int d = 0; d < 3; ++d) om[d] = t * dw[d] + a * c1[d] + b * c2[d];
}

// v <- R v and dv <- R dv + om x (R v)
__device__ __forceinline__ void rotate_tan(double *v, double *dv, const Rot &r, const double *om)
{
    rotate_cs(v, r.h, r.c, r.s);
    rotate_cs(dv, r.h, r.c, r.s);
    double c[3];
    cross3(om, v, c);
    dv[0] += c[0];
    dv[1] += c[1];
    dv[2] += c[2];
}

// electron_half on (x, dx): the linear step acts on both; every rotation
// carries its tangent rule, in electron_half's order
__device__ __forceinline__ void electron_half_tan(double *x, double *dx, const double *kA, const double *kB,
                                                  const Rot &r1, const Rot &r2, const double *om1,
                                                  const double *om2)
{
#pragma unroll
    for (int pass = 0; pass < 2; ++pass) {
        double *z = pass ? dx : x;
#pragma unroll
        for (int b = 0; b < 3; ++b) {
            double y[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                double s = 0.0;
#pragma unroll
                for (int j = 0; j < 4; ++j) s = fma(kA[4 * i + j], z[kSpinBlk[b][j]], s);
                y[i] = s;
            }
#pragma unroll
            for (int i = 0; i < 4; ++i) z[kSpinBlk[b][i]] = y[i];
        }
        double y[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            double s = 0.0;
#pragma unroll
            for (int j = 0; j < 4; ++j) s = fma(kB[4 * i + j], z[kDiagBlk[j]], s);
            y[i] = s;
        }
#pragma unroll
        for (int i = 0; i < 4; ++i) z[kDiagBlk[i]] = y[i];
    }
    rotate_tan(x + 0, dx + 0, r1, om1);
    rotate_tan(x + 3, dx + 3, r2, om2);
#pragma unroll
    for (int b = 0; b < 3; ++b) rotate_tan(x + 6 + 3 * b, dx + 6 + 3 * b, r1, om1);
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        double v[3] = {x[6 + a], x[9 + a], x[12 + a]}, dv[3] = {dx[6 + a], dx[9 + a], dx[12 + a]};
        rotate_tan(v, dv, r2, om2);
        x[6 + a] = v[0], x[9 + a] = v[1], x[12 + a] = v[2];
        dx[6 + a] = dv[0], dx[9 + a] = dv[1], dx[12 + a] = dv[2];
    }
}

// Tangent of the nuclear axis h = S/|S| (n = |S|): dh = (dS - h (h.dS)) / n;
// u = dh and v = h x dh give a nucleus rotated by angle a about h the
// om = sin(a) u + (1 - cos(a)) v (J(a h) a dh with dh orthogonal to h).
__device__ __forceinline__ void axis_tan(const double *S, const double *dS, const double *h, double n, bool z,
                                         double *u, double *v)
{
    const double hd = h[0] * dS[0] + h[1] * dS[1] + h[2] * dS[2];
    const double inv = z ? 0.0 : 1.0 / n;
#pragma unroll
    for (int d = 0; d < 3; ++d) u[d] = (dS[d] - h[d] * hd) * inv;
    cross3(h, u, v);
}

// part[st][wave] += sums over the wave's samples of (P_S(0) P_S(t_st),
// P_S(0) x_16(t_st)), for every sample group the wave processes (kScThreads/L
// samples per group).
// With TAN the kernel also carries the tangent (dx, dI) of the state for the
// field direction p.dw and adds (P_S(0) dP_S(t), P_S(0) dx_16(t)) to
// partd[st][wave].  With BAL the electron-2 slots start at NL/2 (p.nl1 =
// NL/2), known at compile time: the unrolled slot loops then pick each
// slot's electron without selects.
constexpr int kRing = 8; // steps per flush of a wave's partial sums (8 steps x 2 or 4 values = 16 or 32 sums)
template <int NL, int L, bool TAN, bool BAL>
__global__ __launch_bounds__(kScThreads) void k_sc(ScParams p, const int64_t *__restrict__ seeds,
                                                   double *__restrict__ part, double *__restrict__ partd,
                                                   int64_t n_samples, int64_t n_steps, int64_t n_groups)
{
    const int h = threadIdx.x & (L - 1);
    double kA[16], kB[16]; // wave-uniform: scalar registers
#pragma unroll
    for (int i = 0; i < 16; ++i) kA[i] = p.kjb[i], kB[i] = p.kjb[16 + i];
    double cq[NL], sq[NL]; // this lane's nuclear precession cos/sin per slot
#pragma unroll
    for (int q = 0; q < NL; ++q) cq[q] = p.cs[2 * (L * q + h)], sq[q] = p.cs[2 * (L * q + h) + 1];
    const int lane = threadIdx.x & 63;
    const int64_t nwaves = (int64_t)gridDim.x * (kScThreads / kWave);
    const int64_t gwave = (int64_t)blockIdx.x * (kScThreads / kWave) + (threadIdx.x >> 6);
    const double shalf = 0.86602540378443864676; // sqrt(3/4)
    const double wz[3] = {p.w[0], p.w[1], p.w[2]};
    const double dwz[3] = {p.dw[0], p.dw[1], p.dw[2]};
    const double th = 0.5 * p.dt;
    const int nl1 = BAL ? NL / 2 : p.nl1;
    // per wave: the lane values of the last kRing steps (NV values per step)
    constexpr int NV = TAN ? 4 : 2, NP = 64 / (kRing * NV);
    static_assert(NP >= 1 && kRing * NV * NP == 64, "one sum per lane group");
    __shared__ double ring[kScThreads / kWave][kRing][NV][kWave];
    const int wv = threadIdx.x >> 6;
    auto ring_sync = [] {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    };

    for (int64_t grp = blockIdx.x; grp < n_groups; grp += gridDim.x) {
        const int64_t sample = grp * (kScThreads / L) + threadIdx.x / L;
        const bool active = sample < n_samples;
        double x[16];
        double I[NL][3];
#pragma unroll
        for (int q = 0; q < NL; ++q) I[q][0] = I[q][1] = I[q][2] = 0.0;
        {
            Rng g;
            const int64_t sidx = active ? sample : 0;
            rng_set_seed(g, seeds[2 * sidx], seeds[2 * sidx + 1]);
            rng_sphere(g, shalf, x + 0);
            rng_sphere(g, shalf, x + 3);
            double tmp[3];
            for (int k = 0; k < p.nn; ++k) { // the same stream in both lanes; each keeps its own nuclei
                rng_sphere(g, p.nuc[64 + k], tmp);
                const double a = p.nuc[k];
                const bool e2 = k >= p.n1;
                const int j = e2 ? k - p.n1 : k;
                const int slot = j / L + (e2 ? nl1 : 0);
                const bool mine = (j & (L - 1)) == h;
#pragma unroll
                for (int q = 0; q < NL; ++q)
                    if (mine && q == slot) {
                        I[q][0] = a * tmp[0];
                        I[q][1] = a * tmp[1];
                        I[q][2] = a * tmp[2];
                    }
            }
        }
#pragma unroll
        for (int a = 0; a < 3; ++a)
#pragma unroll
            for (int b = 0; b < 3; ++b) x[6 + a + 3 * b] = x[a] * x[3 + b];
        x[15] = 1.0;
        const double p0 = ps_of(x);
        const bool counts = active && h == 0;
        double w1[3], w2[3];
        fields<NL, L>(wz, nl1, I, w1, w2);
        Rot r1 = make_rot(w1, th), r2 = make_rot(w2, th);
        // tangents: zero at t = 0 (the initial state does not depend on the field)
        double dx[TAN ? 16 : 1], dI[TAN ? NL : 1][3], dw1[3], dw2[3], om1[3], om2[3];
        if constexpr (TAN) {
#pragma unroll
            for (int i = 0; i < 16; ++i) dx[i] = 0.0;
#pragma unroll
            for (int q = 0; q < NL; ++q) dI[q][0] = dI[q][1] = dI[q][2] = 0.0;
#pragma unroll
            for (int d = 0; d < 3; ++d) dw1[d] = dw2[d] = dwz[d];
            rot_jac(w1, dw1, th, r1, om1);
            rot_jac(w2, dw2, th, r2, om2);
        }

        for (int64_t st = 0; st <= n_steps; ++st) {
            if (st > 0) {
                if constexpr (TAN)
                    electron_half_tan(x, dx, kA, kB, r1, r2, om1, om2);
                else
                    electron_half(x, kA, kB, r1, r2);
                // nuclear full step about S_e / |S_e| (angle a_k sqrt(3)/2 dt)
                const double n1 = sqrt(x[0] * x[0] + x[1] * x[1] + x[2] * x[2]);
                const double n2 = sqrt(x[3] * x[3] + x[4] * x[4] + x[5] * x[5]);
                const bool z1 = !(n1 > 0.0), z2 = !(n2 > 0.0);
                const double i1 = 1.0 / (z1 ? 1.0 : n1), i2 = 1.0 / (z2 ? 1.0 : n2);
                const double h1[3] = {x[0] * i1, x[1] * i1, x[2] * i1};
                const double h2[3] = {x[3] * i2, x[4] * i2, x[5] * i2};
                if constexpr (TAN) {
                    double u1[3], v1[3], u2[3], v2[3];
                    axis_tan(x + 0, dx + 0, h1, n1, z1, u1, v1);
                    axis_tan(x + 3, dx + 3, h2, n2, z2, u2, v2);
#pragma unroll
                    for (int q = 0; q < NL; ++q) {
                        const bool e1 = q < nl1;
                        const bool zz = e1 ? z1 : z2;
                        const double c = zz ? 1.0 : cq[q], sn = zz ? 0.0 : sq[q];
                        const double *hh = e1 ? h1 : h2, *uu = e1 ? u1 : u2, *vv = e1 ? v1 : v2;
                        rotate_cs(I[q], hh, c, sn);
                        rotate_cs(dI[q], hh, c, sn);
                        double om[3], cr[3];
#pragma unroll
                        for (int d = 0; d < 3; ++d) om[d] = sn * uu[d] + (1.0 - c) * vv[d];
                        cross3(om, I[q], cr);
#pragma unroll
                        for (int d = 0; d < 3; ++d) dI[q][d] += cr[d];
                    }
                } else {
#pragma unroll
                    for (int q = 0; q < NL; ++q) {
                        if (q < nl1)
                            rotate_cs(I[q], h1, z1 ? 1.0 : cq[q], z1 ? 0.0 : sq[q]);
                        else
                            rotate_cs(I[q], h2, z2 ? 1.0 : cq[q], z2 ? 0.0 : sq[q]);
                    }
                }
                fields<NL, L>(wz, nl1, I, w1, w2);
                r1 = make_rot(w1, th);
                r2 = make_rot(w2, th);
                if constexpr (TAN) {
                    fields<NL, L>(dwz, nl1, dI, dw1, dw2);
                    rot_jac(w1, dw1, th, r1, om1);
                    rot_jac(w2, dw2, th, r2, om2);
                    electron_half_tan(x, dx, kA, kB, r1, r2, om1, om2);
                } else {
                    electron_half(x, kA, kB, r1, r2);
                }
            }
            // this step's lane values into the wave's ring; every kRing steps (and at
            // the end) the wave's sums go to part/partd, one read-modify-write per
            // sum, so a load's latency is paid once per kRing steps
            const int slot = (int)(st % kRing);
            ring[wv][slot][0][lane] = counts ? p0 * ps_of(x) : 0.0;
            ring[wv][slot][1][lane] = counts ? p0 * x[15] : 0.0;
            if constexpr (TAN) {
                ring[wv][slot][2][lane] = counts ? p0 * ps_of(dx) : 0.0;
                ring[wv][slot][3][lane] = counts ? p0 * dx[15] : 0.0;
            }
            if (slot == kRing - 1 || st == n_steps) {
                ring_sync();
                // sum q = lane / NP (step q / NV, value q % NV) over the NP parts of 64 / NP lanes
                const int q = lane / NP, part_i = lane % NP, j = q / NV, c = q % NV;
                double acc = 0.0;
#pragma unroll
                for (int i = 0; i < 64 / NP; ++i) acc += ring[wv][j][c][part_i * (64 / NP) + i];
#pragma unroll
                for (int o = 1; o < NP; o <<= 1) acc += __shfl_xor(acc, o);
                const int64_t st0 = st - slot;
                if (part_i == 0 && j <= slot) {
                    double *base = (c < 2 ? part : partd) + 2 * ((st0 + j) * nwaves + gwave) + (c & 1);
                    *base += acc;
                }
                ring_sync(); // the ring is rewritten from the next step on
            }
        }
    }
}

// ps[st] = 4/n_samples * sum_w part[st][w].x, pt[st] = 4/n_samples * sum_w part[st][w].y - ps[st]
__global__ void k_sc_reduce(const double2 *__restrict__ part, int64_t nwaves, int64_t n_steps, int64_t n_samples,
                            double *__restrict__ ps, double *__restrict__ pt)
{
    const int64_t st = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (st > n_steps) return;
    double a = 0.0, b = 0.0;
    for (int64_t w = 0; w < nwaves; ++w) {
        const double2 v = part[st * nwaves + w];
        a += v.x;
        b += v.y;
    }
    const double f = 4.0 / (double)n_samples;
    ps[st] = f * a;
    pt[st] = f * b - f * a;
}

__global__ void k_fill_zero(double *__restrict__ a, int64_t n)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        a[i] = 0.0;
}

template <int NL, int L, bool TAN>
hipError_t launch(const ScParams &p, const int64_t *seeds, double *part, double *partd, int nblocks,
                  int64_t n_samples, int64_t n_steps, hipStream_t s)
{
    const int64_t n_groups = (n_samples + kScThreads / L - 1) / (kScThreads / L);
    const bool bal = 2 * p.nl1 == NL;
    if constexpr (TAN && NL > 8) { // only the balanced layout fits the registers (sc_slots)
        if (!bal) return hipErrorInvalidValue;
        hipLaunchKernelGGL((k_sc<NL, L, TAN, true>), dim3(nblocks), dim3(kScThreads), 0, s, p, seeds, part, partd,
                           n_samples, n_steps, n_groups);
    } else {
        if (bal)
            hipLaunchKernelGGL((k_sc<NL, L, TAN, true>), dim3(nblocks), dim3(kScThreads), 0, s, p, seeds, part,
                               partd, n_samples, n_steps, n_groups);
        else
            hipLaunchKernelGGL((k_sc<NL, L, TAN, false>), dim3(nblocks), dim3(kScThreads), 0, s, p, seeds, part,
                               partd, n_samples, n_steps, n_groups);
    }
    return hipGetLastError();
}

} // namespace

int sc_slots(int n1, int n2, int lanes, bool tangent)
{
    const int need1 = (n1 + lanes - 1) / lanes, need2 = (n2 + lanes - 1) / lanes, need = need1 + need2;
    for (int nl : {2, 4, 6, 8, 12, 16, 24, 32}) {
        if (need > nl) continue;
        // the tangent kernel carries twice the state: up to 8 slots, or 12 in
        // the balanced layout (each electron's nuclei in one half)
        const bool fits = tangent ? (nl <= 8 || (nl == 12 && need1 <= 6 && need2 <= 6)) : (lanes == 2 || nl <= 12);
        if (fits) return nl;
    }
    return -1;
}

hipError_t sc_run(const ScParams &p, const int64_t *seeds, double *part, int nblocks, int64_t n_samples,
                  int64_t n_steps, double *ps, double *pt, hipStream_t s, double *partd, double *dps, double *dpt)
{
    const int64_t nwaves = (int64_t)nblocks * (kScThreads / kWave);
    const int64_t npart = 2 * (n_steps + 1) * nwaves;
    const bool tan = partd != nullptr;
    hipLaunchKernelGGL(k_fill_zero, dim3((unsigned)std::min<int64_t>((npart + 255) / 256, 4096)), dim3(256), 0, s,
                       part, npart);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    if (tan) {
        hipLaunchKernelGGL(k_fill_zero, dim3((unsigned)std::min<int64_t>((npart + 255) / 256, 4096)), dim3(256), 0,
                           s, partd, npart);
        if ((e = hipGetLastError()) != hipSuccess) return e;
    }
#define SPC_SC_CASE(NL, L)                                                                                         \
    if (!tan && p.slots == NL && p.lanes == L)                                                                     \
        e = launch<NL, L, false>(p, seeds, part, partd, nblocks, n_samples, n_steps, s);                          \
    else
#define SPC_SC_TCASE(NL, L)                                                                                        \
    if (tan && p.slots == NL && p.lanes == L)                                                                      \
        e = launch<NL, L, true>(p, seeds, part, partd, nblocks, n_samples, n_steps, s);                           \
    else
    SPC_SC_CASE(2, 2) SPC_SC_CASE(4, 2) SPC_SC_CASE(6, 2) SPC_SC_CASE(8, 2) SPC_SC_CASE(12, 2) SPC_SC_CASE(16, 2)
    SPC_SC_CASE(24, 2) SPC_SC_CASE(32, 2)
    SPC_SC_CASE(2, 4) SPC_SC_CASE(4, 4) SPC_SC_CASE(6, 4) SPC_SC_CASE(8, 4) SPC_SC_CASE(12, 4)
    SPC_SC_CASE(2, 8) SPC_SC_CASE(4, 8) SPC_SC_CASE(6, 8) SPC_SC_CASE(8, 8) SPC_SC_CASE(12, 8)
    SPC_SC_TCASE(2, 2) SPC_SC_TCASE(4, 2) SPC_SC_TCASE(6, 2) SPC_SC_TCASE(8, 2) SPC_SC_TCASE(12, 2)
    SPC_SC_TCASE(2, 4) SPC_SC_TCASE(4, 4) SPC_SC_TCASE(6, 4) SPC_SC_TCASE(8, 4) SPC_SC_TCASE(12, 4)
    SPC_SC_TCASE(2, 8) SPC_SC_TCASE(4, 8) SPC_SC_TCASE(6, 8) SPC_SC_TCASE(8, 8) SPC_SC_TCASE(12, 8)
    return hipErrorInvalidValue;
#undef SPC_SC_CASE
#undef SPC_SC_TCASE
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_sc_reduce, dim3((unsigned)((n_steps + 1 + 255) / 256)), dim3(256), 0, s,
                       reinterpret_cast<const double2 *>(part), nwaves, n_steps, n_samples, ps, pt);
    if ((e = hipGetLastError()) != hipSuccess || !tan) return e;
    hipLaunchKernelGGL(k_sc_reduce, dim3((unsigned)((n_steps + 1 + 255) / 256)), dim3(256), 0, s,
                       reinterpret_cast<const double2 *>(partd), nwaves, n_steps, n_samples, dps, dpt);
    return hipGetLastError();
}

} // namespace spc
