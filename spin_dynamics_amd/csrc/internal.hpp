// internal.hpp -- host-side declarations shared by the HIP translation units
// and the C-ABI layer.  Not installed; the public interface is spinchem.h.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

namespace spc {

// Spin system as the device sees it (built once per field value).
struct DevSystem {
    int nn;                  // number of nuclei
    int n1;                  // nuclei on electron 1 (the first n1 entries)
    int64_t Z;               // nuclear dimension
    int mult[64];
    int64_t stride[64];      // digit stride in the nuclear index
    double A[64][9];         // A(a,b) at A[k][3*b + a]
    double zeta[64];         // nuclear Zeeman B*gamma_k/gamma_e
    double gamma[64];        // gamma_k (rad s^-1 T^-1 / 1e9, the reference's units)
    double w[3];             // electron Zeeman vector, -B n
    double dir[3];           // field direction n
    double Tre[9], Tim[9];   // S1.T.S2 coupling T(a,b) at [3*b + a], T = (2J + 2i dk) 1 + D
    double kbar;             // (kS + 3 kT)/4
};

// Nuclear digits of the Kronecker product form: H rows 4n + e couple n only to
// itself (4x4 block) and to n +- stride_k, through c(m) M (x) 1 on the
// electron of nucleus k (M a 2x2 matrix, c the ladder element).
struct KronDev {
    int nn;
    int32_t stride[64];
    int32_t mult[64];
    int32_t elec[64]; // 0: electron 1 (spin s1 of e = 2 s1 + s2), 1: electron 2
    uint64_t e2mask;  // bit k: nucleus k on electron 2 (they follow electron 1's)
};

// How many of the last nuclei (at most 8) are spin-1/2 with strides 1, 2, 4,
// ...: the product can then give a wave (and a workgroup) 2^q consecutive
// nuclear indices.
inline int kron_inwave(const KronDev &k)
{
    int q = 0;
    while (q < 8 && q < k.nn && k.mult[k.nn - 1 - q] == 2 && k.stride[k.nn - 1 - q] == (1 << q)) ++q;
    return q;
}

// ---------------------------------------------------------- hamiltonian.hip
// counts[N]: entries per row; then rowptr by exclusive scan; then fill.
hipError_t h_count(const DevSystem *d_sys, int64_t N, int32_t *counts, hipStream_t s);
hipError_t h_scan(const int32_t *counts, int32_t *rowptr, int64_t N, void *tmp, size_t &tmp_bytes,
                  hipStream_t s);
hipError_t h_fill(const DevSystem *d_sys, int64_t N, const int32_t *rowptr, int32_t *col, double2 *val,
                  hipStream_t s);
// 4x4 blocks over nuclear indices from the CSR: bcounts[Z+1], then browptr by
// the same scan, then bcol[nb] and bval[nb][16].
hipError_t h_block_count(const int32_t *rowptr, const int32_t *col, int64_t Z, int32_t *bcounts, hipStream_t s);
hipError_t h_block_fill(const int32_t *rowptr, const int32_t *col, const double2 *val, int64_t Z,
                        const int32_t *browptr, int32_t *bcol, double2 *bval, hipStream_t s);
// The values of another CSR on an existing block pattern (missing[0] set to 1
// if an entry has no block).
hipError_t h_block_fill_pattern(const int32_t *rowptr, const int32_t *col, const double2 *val, int64_t Z,
                                const int32_t *browptr, const int32_t *bcol, double2 *bval, int32_t *missing,
                                hipStream_t s);
// Kronecker tables from the block form: diag[n] = block (n, n); gen[k][dir][d]
// = the 2x2 matrix of the block from row n (digit d) to n + stride_k (dir 0)
// or n - stride_k (dir 1), taken at the row whose other digits are 0.  bad[0]
// is set if any block of H differs from that form (then the tables are not
// used).  nsets value sets (stride nb*16 in bval) give nsets table sets.
hipError_t kron_tables(const KronDev *d_kron, const KronDev &h_kron, const int32_t *browptr, const int32_t *bcol,
                       const double2 *bval, int64_t nb, int nsets, int64_t Z, double2 *diag, double2 *gen,
                       int32_t *bad, hipStream_t s);
// Per nuclear index the neighbour masks (minus, plus, one) over all nuclei: out[3 Z].
hipError_t kron_masks(const KronDev *d_kron, int64_t Z, uint64_t *out, hipStream_t s);
// out[o] = v0 + sum_a fields[o][a] z_a for n_fields field vectors (count values each).
hipError_t h_field_values(const double2 *v0, const double2 *zx, const double2 *zy, const double2 *zz,
                          const double *fields, int64_t count, int n_fields, double2 *out, hipStream_t s);

// ---------------------------------------------------------------- newton.cpp
// count Newton shifts of H (a host copy of its CSR) in Leja order into out:
// the Ritz values of a 40-step Lanczos run on (H + H^H)/2, repeated
// cyclically when there are fewer; 0 on success.
int newton_shifts_csr(int64_t N, const int32_t *rowptr, const int32_t *col, const double2 *val, int count,
                      double *out);

// ------------------------------------------------------------- sampling.hip
struct VecInit {
    int kind;            // 0 SU(Z), 1 phase, 2 injected, 3 basis
    const int64_t *seeds; // [M][2] per-lane seeds (kinds 0, 1)
    const double2 *inj;   // [M][Z] injected nuclear vectors (kind 2)
    int64_t basis_offset; // first basis index (kind 3)
    double2 theta[4];     // electron state
};
hipError_t init_vectors(const VecInit &vi, int M, int Mp, int64_t Z, double2 *psi, hipStream_t s);

// -------------------------------------------------------- semiclassical.hip
constexpr int kScThreads = 256; // threads per block of the semiclassical kernel
struct ScParams {
    int nn, n1;          // nuclei, of which the first n1 couple to electron 1
    int lanes;           // lanes per sample: 2, 4 or 8
    int nl1;             // first electron-2 slot: slots/2 when both electrons fit a half, else ceil(n1 / lanes)
    int slots;           // slots per lane (kernel instantiation), from sc_slots
    const double *nuc;   // device [2][64]: a_k (mT), sqrt(I(I+1)) per nucleus in draw order
    const double *cs;    // device [slots][lanes][cos, sin] of a_k sqrt(3)/2 dt (1, 0 for empty slots)
    const double *kjb;   // device [2][4][4]: exp(dt/2 (K+J)) on (S1x,S2x,T32,T23) and (T11,T22,T33,1), row-major
    double w[3];         // electron Zeeman field -B n
    double dw[3];        // its tangent (the tangent kernel only)
    double dt;
};
// Slots per lane (2 .. 32 for 2 lanes, 2 .. 12 for 4 or 8; for the tangent
// kernel, which carries twice the state, 2 .. 8, or 12 when each electron's
// nuclei fit half of them) for n1 + n2 nuclei split over `lanes` lanes; -1 if
// more are needed.
int sc_slots(int n1, int n2, int lanes, bool tangent = false);
// part: 2 (n_steps + 1) nblocks kScThreads/64 doubles (zeroed here); ps, pt:
// n_steps + 1 values of P_S(t), P_T(t).  The summation order depends on nblocks.
// With partd (same size) the tangent kernel also gives dps, dpt: their
// derivatives for the field moving along p.dw.
hipError_t sc_run(const ScParams &p, const int64_t *seeds, double *part, int nblocks, int64_t n_samples,
                  int64_t n_steps, double *ps, double *pt, hipStream_t s, double *partd = nullptr,
                  double *dps = nullptr, double *dpt = nullptr);

// --------------------------------------------------------------- krylov.hip
// Kernel families timed by the optional event profiler (spc_profile_*).
enum ProfFamily {
    PF_PASS1 = 0, PF_PASS2, PF_DOTS, PF_PSI, PF_REDUCE, PF_COEFFS, PF_EXPM, PF_STEP, PF_GRAM, PF_SDOTS, PF_SFORM,
    PF_COUNT
};

// Event pairs around launches on the work stream; read after a synchronize.
struct Prof {
    bool on = false;
    struct Rec {
        int fam;
        double bytes;
        hipEvent_t a, b;
    };
    std::vector<Rec> recs;
    std::vector<hipEvent_t> pool;
    size_t used = 0;
    hipEvent_t get()
    {
        if (used == pool.size()) {
            hipEvent_t e;
            if (hipEventCreate(&e) != hipSuccess) return nullptr;
            pool.push_back(e);
        }
        return pool[used++];
    }
    void begin(int fam, double bytes, hipStream_t s)
    {
        Rec r{fam, bytes, get(), get()};
        if (r.a) hipEventRecord(r.a, s);
        recs.push_back(r);
    }
    void end(hipStream_t s)
    {
        if (!recs.empty() && recs.back().b) hipEventRecord(recs.back().b, s);
    }
    void clear()
    {
        recs.clear();
        used = 0;
    }
    ~Prof()
    {
        for (auto e : pool) hipEventDestroy(e);
    }
};

// Device workspace of one propagation chunk; all arrays lane-fastest.
struct KrylovWork {
    int64_t N, Z;
    int M, Mp, m;
    const int32_t *rowptr; // CSR of H (device basis order)
    const int32_t *col;
    const double2 *val;
    const int32_t *browptr; // the same H in 4x4 blocks over nuclear indices
    const int32_t *bcol;
    const double2 *bval;     // [nblocks][4][4]
    int64_t bval_stride;     // values per lane group (0: one H for all lanes)
    int chunks_per_bval;     // 64-lane chunks sharing one value set
    // Kronecker form of the same H (used by the product when set): per nuclear
    // index the 4x4 diagonal block, per nucleus the 2x2 electron matrices of
    // its flips (see kron_tables); strides per lane group as bval_stride
    const struct KronDev *kron;
    const double2 *kdiag;    // [Z][16]
    const double2 *kgen;     // [nn][2][3][4]
    int64_t kdiag_stride, kgen_stride;
    int kq;                  // trailing spin-1/2 nuclei of strides 1, 2, 4, .. (kron_inwave, 0..8)
    int kne2p;               // 1 + the number of electron-2 nuclei (they are the last); 0: not known
    int slow_mult;           // multiplicity and stride (nuclear indices) of the slowest nucleus
    int64_t slow_stride;
    const uint64_t *kmask;   // [Z][3] neighbour masks (kron_masks), or null
    double2 *Q;      // [m] basis vectors [N][Mp], qs elements apart (qstride)
    int64_t qs;      // basis vector stride in elements (0: N*Mp)
    double2 *u;      // [N][Mp]
    double2 *v;      // [N][Mp]
    double2 *psi;    // [N][Mp]
    double2 *part;   // partial sums [nrb][nacc][Mp]
    double2 *dots;   // reduced update-pass sums [2*kMaxKrylov + 2][Mp]
    double2 *dots1;  // reduced product-pass sums [2][Mp] (also the populations)
    double2 *coef;   // pass-2 coefficients [2*kMaxKrylov + 4][Mp]
    double2 *tprev;  // tentative column [kMaxKrylov + 1][Mp]
    double2 *hess;   // Hessenberg [m][m][Mp] (col-major per lane)
    double2 *E;      // expm [m][m][Mp]
    double2 *c;      // Krylov coefficients [m][Mp]
    double *beta0;   // [Mp]
    double *lane_t;  // [Mp]
    int32_t *lane_k; // [Mp] next output index
    int32_t *flags;  // [4]: 0 breakdown, 1 direct-mode violation, 2 non-finite
    double *obs;     // [n_out][4][Mp]
    int64_t n_out;
    size_t part_capacity; // elements of part
    int rpw1;              // nuclear indices per wave of the sparse product
    int rpw2;              // nuclear indices per block of the update pass
    int rpwp;              // nuclear indices per wave of the psi pass
    int64_t nnz;           // Hamiltonian entries (algorithmic bytes of the product)
    double kS, kT;         // recombination rates (mT units), for Q^H H u via H^H = H + iK
    Prof *prof;            // optional kernel timing
    int sstep = 0;            // block size of the s-step Arnoldi (0 or 1: one column per pass)
    bool no_dots = false;     // products may skip their u^H v, u^H u partials (tiled product only)
    double sigma = 0.0;       // products store H u - sigma u (the Newton chain's shift)
    double2 *sbuf = nullptr;  // its dots and forming coefficients, kSStepBuf x Mp
    // the conditioning monitor, two words since the last reset (bits of
    // positive doubles; 0 = none yet): [0] the largest 1/p, p the smallest
    // pivot of a monomial block's unit-diagonal Cholesky factor, [1] the
    // largest condition estimate of a Newton block's scaled Gram matrix; the
    // block basis loses orthogonality as they grow
    unsigned long long *cond = nullptr;
    // Newton shifts of H (host, at least m values in Leja order; null: none):
    // with them a basis that serves one step is built as one block
    const double *shifts = nullptr;
    // Set by a caller that reads nothing of the basis after psi_update (one
    // step per rebuild): the block Arnoldi then leaves the last block's chain
    // vectors unformed (fold_K, fold_nf > 0 after arnoldi: psi's coordinates
    // over them follow from fold_coef, and slot 0's vector is fold_q0) and
    // psi_update folds the forming coefficients into c instead.
    bool fold_psi = false;
    // Packed projector bins [4][(m+1)(m+2)/2][Mp] (null: none; tol > 0):
    // a Newton block leaves Y^H P_X Y of its chain there (gram_ready until
    // gram_projectors consumes the flag), gram_projectors Q^H P_X Q of a
    // formed basis.
    double2 *gbins = nullptr;
    bool gram_ready = false;
    // what the step loop reads for the populations (set by gram_projectors or
    // the Newton block): the four Hermitian G_X step_gxs apart, packed with
    // row length step_nd (0: full m x m), over R^-1 c when step_Ri is set
    const double2 *step_G = nullptr;
    int64_t step_gxs = 0;
    int step_nd = 0;
    const double2 *step_Ri = nullptr;
    int fold_K = 0, fold_nf = 0;
    const double2 *fold_coef = nullptr, *fold_q0 = nullptr;
};
// Words of KrylovWork::cond.
constexpr int kCondWords = 2;
// Elements per lane of KrylovWork::sbuf.
constexpr int kSStepBuf = 1600;
// Row blocks of the s-step dots pass for dimension N (8 .. 256, N/128 rounded
// up to a multiple of 8), and the partials it needs for order m, block size S.
int sstep_row_blocks(int64_t N);
size_t sstep_partials(int64_t N, int m, int S, int Mp);

inline int64_t qstride(const KrylovWork &w) { return w.qs > 0 ? w.qs : w.N * (int64_t)w.Mp; }
// Stride between the basis vectors of an N x Mp basis: N*Mp plus a pad that
// staggers the vectors over HBM channels and banks (vectors exactly 2^k-ish
// apart put every stream of the update pass on the same bank at once).
int64_t basis_stride(int64_t N, int Mp);
size_t partials_needed(const KrylovWork &w);
hipError_t arnoldi(KrylovWork &w, hipStream_t s);               // psi -> Q, hess, beta0
hipError_t expm_lanes(KrylovWork &w, double dt, hipStream_t s);   // hess -> E
hipError_t gram_projectors(KrylovWork &w, hipStream_t s);         // Q -> gram
hipError_t step_lanes(KrylovWork &w, double dt, double t_end, double tol, bool gram, hipStream_t s);
hipError_t psi_update(KrylovWork &w, bool observe, hipStream_t s); // psi = Q c (+ observables)
hipError_t obs_reduce_lanes(const double *obs, int64_t n_out, int M, int Mp, double *acc, hipStream_t s);
// acc[g][x*n_out + k] += sum over lanes g*group .. g*group+M_g-1 (M_g = M valid lanes of each group)
hipError_t obs_reduce_groups(const double *obs, int64_t n_out, int groups, int group, int M_g, int Mp, double *acc,
                             hipStream_t s);
hipError_t spmm_only(KrylovWork &w, hipStream_t s);              // v = H u
hipError_t expm_single(int m, const double2 *A, double2 *E, hipStream_t s);
// Times (and scales) for expm_coeffs: C[o] = scale_o beta exp(-i tau_o A) e_1.
constexpr int kMaxTaus = 8;
struct TauSet {
    int n;
    double tau[kMaxTaus];
    double scale[kMaxTaus];
};
hipError_t expm_coeffs(KrylovWork &w, const TauSet &ts, double2 *C, hipStream_t s); // hess, beta0 -> C[n][m][Mp]
// The coefficient sets C[nout][m][Mp] over the basis as coefficients over what
// arnoldi left in w (the last block's chain vectors, vector 0 in fold_q0):
// *Cuse (C itself, or their fold in w.sbuf) and *q0 for combine().  Keeps the
// fold for psi_update.
hipError_t fold_sets(KrylovWork &w, const double2 *C, int nout, const double2 **Cuse, const double2 **q0,
                     hipStream_t s);

// ---------------------------------------------------------------- adjoint.hip
// Up to kCombineOut outputs y_o = sum_k C[o][k] q_k (+ omega P_S src for o = 0)
// from one read of the basis Q ([m][N][Mp]; q_0 from q0 when given, which may
// be one of the outputs); outputs with y[o] == nullptr skipped; lanes r >= M
// written as zeros.  diff0: output 0 is stored as its singlet difference
// y[4n+1] - y[4n+2] only ([Z][Mp]).  src_diff: src holds such a difference
// (the singlet source term then reads it instead of src's two rows).
constexpr int kCombineOut = 6;
struct CombineOut {
    double2 *y[kCombineOut];
    int n;
    bool diff0 = false;
    bool src_diff = false;
};
hipError_t combine(const double2 *Q, int64_t qs, int m, const double2 *C, const CombineOut &out, const double2 *src,
                   double omega, int64_t Z, int M, int Mp, hipStream_t s, const double2 *q0 = nullptr);
// lam = omega P_S psi (P_S the singlet projector on each electron block).
hipError_t singlet_source(const double2 *psi, double omega, double2 *lam, int64_t Z, int Mp, hipStream_t s);
// S[n][slot][e'][e] += sum_j sum_lanes conj(b_j[4n+e']) a_j[4n_s+e] over the
// nuclear neighbours n_s of n (slot 0: n; 1+2k: n + stride_k; 2+2k: n - stride_k,
// where the digit allows), for J vectors a_j, b_j (strided by NMp).
hipError_t grad_density(const double2 *a, const double2 *b, int J, int64_t NMp, const DevSystem *d_sys,
                        const DevSystem &h_sys, int M, int Mp, double2 *S, hipStream_t s);

} // namespace spc
