/*
 * spinchem.h -- C-ABI of libspinchem_hip.so, the MI355X (gfx950) path for
 * trace-sampling singlet-yield propagation.
 *
 * Conventions
 *   - Complex numbers are interleaved doubles (re, im), the layout of C99
 *     double _Complex and Fortran complex(c_double_complex).
 *   - Host arrays are caller-allocated; device memory is owned by the context.
 *   - Every call returns 0 (SPC_OK) or an SPC_ERR_* code; spc_last_error() gives
 *     the message.  A context is used from one host thread at a time.
 *   - Units are those of the reference after read_inp (spinsys.f90:60-135):
 *     couplings and field in mT, rates divided by gamma_e, times multiplied by
 *     gamma_e/1000 (ns -> rad/mT).
 *
 * Each entry point names the reference interface it replaces.
 */
#ifndef SPINCHEM_H
#define SPINCHEM_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SPC_OK 0
#define SPC_ERR_INVALID 1     /* bad argument or input (reference: stop 'Error: ...') */
#define SPC_ERR_UNSUPPORTED 2 /* multiplicity, Krylov order or size outside the device path */
#define SPC_ERR_OOM 3         /* device allocation failed */
#define SPC_ERR_HIP 4         /* HIP runtime error */
#define SPC_ERR_BREAKDOWN 5   /* Krylov breakdown (h_{i+1,i} = 0) or non-finite result */
#define SPC_ERR_NO_DEVICE 6   /* no HIP device */

#define SPC_MAX_KRYLOV 32   /* largest N_krylov on the device path */
#define SPC_MAX_NUCLEI 64

/* Initial radical-pair state (qm_dynamics.f90:386-414 init_el_state). */
#define SPC_STATE_SINGLET 0
#define SPC_STATE_T0 1
#define SPC_STATE_TP 2
#define SPC_STATE_TM 3

/* Trace-vector kinds. */
#define SPC_VEC_SUZ 0      /* sample_SUZ: complex normals, normalised (qm_dynamics.f90:416-430) */
#define SPC_VEC_PHASE 1    /* random phase: (x + i y)/sqrt(Z), (x, y) = rng%circle(1) */
#define SPC_VEC_INJECTED 2 /* caller-provided nuclear vectors [M][Z] */

typedef struct spc_ctx spc_ctx;

/* Flattened sys_param (spinsys.f90:21-38) plus the field of main.f90:55-56.
 * Tensors are column-major per nucleus as in the .ini (spinsys.f90:231):
 * A(a,b) = hfc[9k + 3b + a].  With iso_rule = 1 the reference's isotropy test
 * (A11 == A22 for every nucleus of both electrons, spinsys.f90:237-254) decides
 * whether trace(A)/3 replaces each tensor, as getHiso_sp does. */
typedef struct spc_spin_system {
    int32_t n_nuc1, n_nuc2;    /* nuclei on electron 1 and 2                 */
    const int32_t *mult;       /* n_nuc1 + n_nuc2 multiplicities (2 or 3)     */
    const double *hfc;         /* 9 * (n_nuc1 + n_nuc2), mT                   */
    double J;                  /* exchange, mT                                */
    double D[9];               /* dipolar tensor, column-major, mT            */
    double kS, kT;             /* recombination rates / gamma_e               */
    double B;                  /* field magnitude, mT                         */
    double field_dir[3];       /* unit direction; (0, 0, 1) is the reference  */
    int32_t iso_rule;          /* 1: reference isotropy rule, 0: full tensors */
} spc_spin_system;

/* sim_param subset used by the propagation (spinsys.f90:40-56). */
typedef struct spc_sim_params {
    int32_t n_samples;   /* M trace vectors                                    */
    int32_t n_krylov;    /* m (1..SPC_MAX_KRYLOV)                              */
    double t_end;        /* simulation_time, converted                         */
    double dt;           /* dt, converted                                      */
    double tol;          /* integrator_tolerance (>= 0)                        */
    int32_t init_state;  /* SPC_STATE_*                                        */
    int32_t vec_kind;    /* SPC_VEC_*                                          */
} spc_sim_params;

/* ---------------------------------------------------------------- context */
/* No reference counterpart: owns the device, a stream and device buffers. */
int spc_ctx_create(int device, spc_ctx **out);
void spc_ctx_destroy(spc_ctx *ctx);
const char *spc_last_error(const spc_ctx *ctx);
/* Library version string. */
const char *spc_version(void);

/* ------------------------------------------------------------ Hamiltonian */
/* Replaces getHiso_sp + coo_to_csr (hamiltonian.f90:17-205, utils.f90:52-89):
 * assembles the CSR of H on the device from the spin-operator factors.  Also
 * covers anisotropic tensors, tilted fields and full D, which the reference
 * stops on (qm_dynamics.f90:300). */
int spc_system_set(spc_ctx *ctx, const spc_spin_system *sys);
/* The system of one symmetry block of symmetrised_dynamics: replaces
 * reduce_system + getHiso_sp (symmetry.f90:146-178, hamiltonian.f90:17-205)
 * for a reduced system.  Each nucleus stands for a set of equivalent spin-1/2
 * nuclei in one total-spin multiplet: any multiplicity >= 1 (1, a singlet
 * set, drops out; getHiso_sp stops above 3), isotropic couplings (the
 * isotropy rule as iso_rule = 1), the nuclear Zeeman term of 1H for every
 * set.  Otherwise as spc_system_set. */
int spc_system_set_block(spc_ctx *ctx, const spc_spin_system *sys);

/* shrink (moments.f90:15-93): the M sets of equivalent nuclei (couplings
 * a_bar[M], counts n_bar[M], sum n_bar = n) whose first M power sums match
 * those of the n couplings a_iso, via the M-point Gauss rule of the discrete
 * measure on a_iso (quad_rule, moments.f90:95-150), integer counts from its
 * weights and Newton's method on the moment equations (moments.f90:184-239).
 * *M is set to 0 when n = 0.  SPC_ERR_INVALID if M >= n ('Error: Symetry
 * reduction requires M<N'); SPC_ERR_BREAKDOWN if no integer choice converges.
 * Host only: needs no context or device. */
int spc_shrink(int32_t n, const double *a_iso, int32_t *M, double *a_bar, int32_t *n_bar);

/* The blocks of symmetrised_dynamics for n_sets sets of n_bar[i] equivalent
 * spin-1/2 nuclei (cartesian_product, sort_col_prod and weight of
 * symmetry.f90:61-65, 187-226, utils.f90:25-50): every combination of one
 * total-spin multiplicity per set (1 + n mod 2, 3 + n mod 2, .., n + 1, the
 * first set varying fastest), sorted stably by nuclear dimension.  Out, for
 * block b: mult[b][n_sets], Z_block[b] = the product of its multiplicities,
 * w_block[b] = the number of copies of that multiplet combination (its weight
 * in the full trace is w_block Z_block / Z).  With mult == NULL only
 * *n_blocks is set.  Host only. */
int spc_symmetry_blocks(int32_t n_sets, const int32_t *n_bar, int64_t capacity, int64_t *n_blocks,
                        int32_t *mult, int64_t *Z_block, int64_t *w_block);

/* Dimensions of the current Hamiltonian; nnz = the entries
 * spc_hamiltonian_copy writes. */
int spc_system_info(const spc_ctx *ctx, int64_t *N, int64_t *Z, int64_t *nnz);
/* Copies H as CSR in the reference basis order (0-based rowptr[N+1],
 * col[nnz], val[2*nnz]), the output of getHiso_sp + coo_to_csr
 * (hamiltonian.f90:17-205, utils.f90:52-89).  When H has getHiso_sp's form
 * (isotropic tensors, field along z, axial coupling) the rows hold that
 * routine's structure, zeros it stores included (nnz = N + N/2 +
 * sum_k (g_k-1) N/g_k, hamiltonian.f90:62-73), in coo_to_csr's order: reverse
 * COO order, i.e. the last nucleus's S-I+ first and the diagonal last.  Other
 * Hamiltonians (which the reference does not build) give their stored
 * entries, ordered by the same term classes and by column within a class. */
int spc_hamiltonian_copy(spc_ctx *ctx, int32_t *rowptr, int32_t *col, double *val);

/* ---------------------------------------------------------------- dynamics */
/* ceiling((t_end+dt)/dt) (qm_dynamics.f90:305); outputs hold this + 1 values. */
int64_t spc_n_steps(double t_end, double dt);

/* Replaces trace_sampling_para (qm_dynamics.f90:257-384): per-sample seeds
 * from seed (main.f90:50, qm_dynamics.f90:315-320), trace vectors, restarted
 * Krylov propagation with the tol rule, ordered sum over samples and 1/M.
 * trace_vectors: [M][Z] complex when vec_kind == SPC_VEC_INJECTED, else NULL.
 * P_*: n_out = spc_n_steps()+1 values each.  Phi_S/Phi_T: NULL or n_out/2
 * cumulative Simpson yields (observables.f90:74-101) of the averaged traces. */
int spc_trace_sample(spc_ctx *ctx, const spc_sim_params *sim, const int64_t seed[2],
                     const double *trace_vectors, double *P_S, double *P_T0, double *P_Tp,
                     double *P_Tm, int64_t n_out, double *Phi_S, double *Phi_T);

/* Size of the scratch area of spc_trace_sample_noalloc, per GPU. */
#define SPC_SCRATCH_BYTES (64u << 20)

/* spc_trace_sample without device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (serving
 * loops that must not call hipMalloc, code around graph capture).  Same
 * arguments, results and errors as spc_trace_sample; every device buffer of
 * the call (Krylov basis, vectors, partial sums, populations, staging) is
 * carved from a scratch area of SPC_SCRATCH_BYTES that the library holds on
 * each GPU for the life of the process, and a workspace beyond it returns
 * SPC_ERR_UNSUPPORTED (about 20 MiB for N = 192, M = 32, m = 20 over 5001
 * steps).  The lanes run as one chunk on the context stream: the call creates
 * no streams or events and allocates nothing on the device; the event
 * profiler must be off.  Every call on a GPU shares that GPU's scratch area:
 * calls that overlap on one GPU (two contexts or threads) race. */
int spc_trace_sample_noalloc(spc_ctx *ctx, const spc_sim_params *sim, const int64_t seed[2],
                             const double *trace_vectors, double *P_S, double *P_T0, double *P_Tp,
                             double *P_Tm, int64_t n_out, double *Phi_S, double *Phi_T);

/* As spc_trace_sample, with the per-sample seeds given ([M][2]): the pairs
 * trace_sampling_para draws from the caller's generator with two next() calls
 * per sample (qm_dynamics.f90:315-320) and passes to set_seed (:326). */
int spc_trace_sample_seeds(spc_ctx *ctx, const spc_sim_params *sim, const int64_t *sample_seeds,
                           double *P_S, double *P_T0, double *P_Tp, double *P_Tm, int64_t n_out,
                           double *Phi_S, double *Phi_T);

/* Replaces exact_dynamics (qm_dynamics.f90:24-137): the Z nuclear basis
 * states as trace vectors (sim->n_samples and vec_kind are ignored). */
int spc_exact_trace(spc_ctx *ctx, const spc_sim_params *sim, double *P_S, double *P_T0,
                    double *P_Tp, double *P_Tm, int64_t n_out, double *Phi_S, double *Phi_T);

/* Replaces expm (matrix_exponential.f90:14-117) for one m x m matrix
 * (column-major, m <= SPC_MAX_KRYLOV), on the device. */
int spc_expm_small(spc_ctx *ctx, int32_t m, const double *A, double *E);

/* Replaces get_krylov (qm_dynamics.f90:432-467) for M vectors at once: psi
 * [M][N] in, Q [M][m][N] (vector-major), A [M][m][m] (column-major) and
 * c0 = ||psi|| [M] out.  The product path keeps these on the device; this
 * entry point exists to check the Arnoldi kernels: the one-column-per-pass
 * form, or the block s-step form when SPC_SSTEP is set (2..8). */
int spc_krylov_basis(spc_ctx *ctx, int32_t M, int32_t m, const double *psi, double *Q, double *A,
                     double *beta);

/* y = H x for M vectors ([M][N] complex each), the spmv of qm_dynamics.f90:452. */
int spc_spmm(spc_ctx *ctx, int32_t M, const double *x, double *y);

/* Reverse-mode gradient of the final singlet yield Phi_S(T) (the last Simpson
 * value of observables.f90:96-99 over the sample-averaged P_S of
 * spc_trace_sample_seeds) with respect to every parameter of H: no
 * counterpart in the reference.  tol must be 0 (one Krylov step per
 * rebuild).  Vectors as spc_trace_sample_seeds (vec_kind SUZ or PHASE with
 * sample_seeds [M][2]) or injected (vec_kind INJECTED, trace_vectors [M][Z]).
 * The adjoint runs backwards on -H^H.  The per-step vectors of the last steps
 * are kept from the forward pass as far as 80% of the free device memory
 * allows (environment SPC_ADJOINT_KEEP caps their number); earlier steps are
 * recomputed from checkpoints every checkpoint_every steps (0: 64).  Outputs (each may be NULL but
 * Phi_S): P_S [n_out] (n_out = ceiling((t_end+dt)/dt)+1), Phi_S (scalar),
 * grad_field [3] = d/d(B n), grad_hfc [9 per nucleus, column-major like hfc],
 * grad_J, grad_D [9, column-major], grad_kS, grad_kT (kS, kT in the units of
 * spc_spin_system). */
int spc_trace_sample_grad(spc_ctx *ctx, const spc_sim_params *sim, const int64_t *sample_seeds,
                          const double *trace_vectors, int32_t checkpoint_every, double *P_S, int64_t n_out,
                          double *Phi_S, double *grad_field, double *grad_hfc, double *grad_J, double *grad_D,
                          double *grad_kS, double *grad_kT);

/* Many field orientations of one radical pair at once: no counterpart in the
 * reference, which runs one field per call (app/main.f90:48-79).  The system
 * of spc_system_set (its field magnitude B; its direction is replaced) at
 * n_orient directions dirs [n][3], with the same M trace vectors (sample_seeds
 * [M][2], as spc_trace_sample_seeds) for every orientation; M a multiple of
 * 64.  Orientations share one block pattern of H and go into the lane
 * dimension, orient_per_chunk at a time (0: as many as fit); up to 4 chunks
 * run at once on their own streams when memory allows (SPC_LANE_STREAMS caps
 * them); with use_graph each chunk's step is captured once in a hipGraph and
 * replayed.  tol > 0 applies the reference's rebuild rule
 * (qm_dynamics.f90:348-364) per lane: a chunk then runs in rounds (a basis
 * per lane, then each lane's steps until |c_m|^2 > tol |c|^2, populations
 * from projector Grams), one captured round replayed per round, with a host
 * check of the lane clocks between rounds.  Outputs:
 * Phi_S [n_orient] final singlet yields; P_S [n_orient][n_out] or NULL. */
int spc_orientation_batch(spc_ctx *ctx, const spc_sim_params *sim, const int64_t *sample_seeds, int32_t n_orient,
                          const double *dirs, int32_t orient_per_chunk, int32_t use_graph, double *Phi_S, double *P_S,
                          int64_t n_out);

/* Autotuner for spc_orientation_batch (no counterpart in the reference):
 * the smallest Krylov order (down from sim->n_krylov in steps of 2, then the
 * odd order between the last that passed and the first that failed) whose
 * pilot yields (64 steps, up to 4 orientations) stay within yield_tol of the
 * requested order's, scaled to the pilot's share of the run, and the
 * orientations per chunk (power of two) with the least device time per
 * orientation-step over a short timed run, with as many chunks in flight as
 * spc_orientation_batch runs.  Every run takes one step per basis (sim->tol
 * is not used). */
int spc_batch_tune(spc_ctx *ctx, const spc_sim_params *sim, const int64_t *sample_seeds, int32_t n_orient,
                   const double *dirs, double yield_tol, int32_t *orient_per_chunk, int32_t *n_krylov,
                   double *ms_per_orient_step);

/* Autotuner for spc_orientation_batch with several steps per basis (no
 * counterpart in the reference): over the orders 16, 20, 24, 28, the loosest
 * tolerance of 1e-12, 1e-16, .., 1e-28 (the reference's rebuild rule,
 * qm_dynamics.f90:348-364) whose pilot yields (256 steps, up to 4
 * orientations) stay within a tenth of yield_tol of the requested order at
 * tol = 0, scaled to the pilot's share of the run; of those the (order, tol) with the
 * least pilot time per orientation-step, or the requested order at tol = 0
 * (*tol = 0) if that is faster or none passes; then the orientations per
 * chunk as spc_batch_tune. */
int spc_batch_tune_adaptive(spc_ctx *ctx, const spc_sim_params *sim, const int64_t *sample_seeds, int32_t n_orient,
                            const double *dirs, double yield_tol, int32_t *orient_per_chunk, int32_t *n_krylov,
                            double *tol, double *ms_per_orient_step);

/* Many radical pairs of uneven Hilbert-space dimension, each at the same
 * n_orient field directions dirs [n][3], in one call: no counterpart in the
 * reference, which runs one system and one field per call
 * (app/main.f90:48-79).  Pair p is systems[p] (its field magnitude B; its
 * direction is replaced) with the same M trace vectors (sample_seeds [M][2])
 * for every pair and orientation; n_krylov[p], tol[p] and
 * orient_per_chunk[p] per pair (each array may be NULL: sim's order and tol,
 * as many orientations per chunk as fit; from spc_batch_tune_adaptive for
 * the 1e-10 yield tolerance).  Every pair's Hamiltonian, field basis and
 * Newton shifts are resident at once, and the chunks of all pairs (largest
 * dimension first) share up to 4 streams: each stream takes the next chunk of
 * any pair when its current one ends, its step or round replayed from the
 * stream's hipGraph for that pair and width.  Results equal
 * spc_orientation_batch per pair.  The context's own system (spc_system_set)
 * is kept.  Output: Phi_S [n_pairs][n_orient]. */
int spc_pairs_batch(spc_ctx *ctx, int32_t n_pairs, const spc_spin_system *systems, const spc_sim_params *sim,
                    const int64_t *sample_seeds, int32_t n_orient, const double *dirs, const int32_t *n_krylov,
                    const double *tol, const int32_t *orient_per_chunk, double *Phi_S);

/* ------------------------------------------------------- semiclassical */
/* Semiclassical spin system (isotropic couplings, as src/sc uses a_iso). */
typedef struct spc_sc_system {
    int32_t n_nuc1, n_nuc2;  /* nuclei on electron 1 and 2 (n_nuc1 + n_nuc2 <= 64) */
    const int32_t *mult;     /* multiplicities (spin I = (mult-1)/2)               */
    const double *a_iso;     /* isotropic hyperfine couplings, mT                   */
    double B;                /* field magnitude, mT                                 */
    double field_dir[3];     /* unit direction                                      */
    double J;                /* exchange, mT (2J S1.S2)                              */
    double kS, kT;           /* rates / gamma_e                                     */
} spc_sc_system;

/* Replaces semi_classical (sc_dynamics.f90:17-100, class_radical_pair.f90:
 * 12-251; unreachable and non-compiling in the reference): classical electron
 * and nuclear spin vectors, one sample per lane pair.  Samples are seeded with
 * two next() draws each from set_seed(seed).  P_S, P_T: n_steps + 1 values of
 * 4 <P_S(0) P_S(t)> and 4 <P_S(0) P_T(t)> with P_T = x_16 - P_S
 * (class_radical_pair.f90:174-175); P_T, Phi_S, Phi_T may be NULL; Phi_S and
 * Phi_T hold (n_steps + 1)/2 Simpson yields (observables.f90:74-101). */
int spc_semiclassical_run(spc_ctx *ctx, const spc_sc_system *sys, int64_t n_samples, const int64_t seed[2],
                          double dt, int64_t n_steps, double *P_S, double *P_T, double *Phi_S, double *Phi_T);

/* Derivative of the semiclassical run with respect to the field vector, along
 * dfield (mT per unit): no counterpart in the reference.  The same samples as
 * spc_semiclassical_run (same seed), each carrying the forward-mode tangent of
 * its classical state through every precession, so dP_S, dP_T are the exact
 * derivatives of the sample averages P_S, P_T (which are also returned) and
 * dPhi_S, dPhi_T those of the Simpson yields.  dfield = the unit field
 * direction gives d/dB at fixed direction.  P_T, Phi_*, dP_T, dPhi_* may be
 * NULL. */
int spc_semiclassical_tangent(spc_ctx *ctx, const spc_sc_system *sys, int64_t n_samples, const int64_t seed[2],
                              double dt, int64_t n_steps, const double dfield[3], double *P_S, double *P_T,
                              double *dP_S, double *dP_T, double *Phi_S, double *Phi_T, double *dPhi_S,
                              double *dPhi_T);

/* Last-run timing of the propagation loop: milliseconds of device time and
 * number of Krylov rebuilds summed over the lane chunks (with tol = 0, the
 * steps times the chunks run), for benchmarks. */
int spc_last_run_stats(const spc_ctx *ctx, double *device_ms, int64_t *rebuilds, int64_t *steps);

/* Kernel-family event profiler (no reference counterpart; benchmarks only).
 * When enabled, the next propagation records HIP events around every launch
 * on the context stream; spc_profile_read sums one family's device time,
 * launch count and algorithmic bytes (complex128 vector passes over the M
 * trace vectors, plus the CSR for the sparse product). */
#define SPC_PROF_PASS1 0  /* v = H u (with Gram-Schmidt dots: columns)  */
#define SPC_PROF_PASS2 1  /* basis update q_i, u_{i+1}                 */
#define SPC_PROF_DOTS 2   /* last-column correction dots               */
#define SPC_PROF_PSI 3    /* psi = Q c with fused populations          */
#define SPC_PROF_REDUCE 4 /* workgroup-partial reductions              */
#define SPC_PROF_COEFFS 5 /* per-vector Arnoldi (and block) algebra    */
#define SPC_PROF_EXPM 6   /* m x m Pade exponential per vector         */
#define SPC_PROF_STEP 7   /* c <- E c step loop                        */
#define SPC_PROF_GRAM 8   /* projector Gram matrices (tol > 0)         */
#define SPC_PROF_SDOTS 9  /* s-step block dots [Q_K, Y]^H Y            */
#define SPC_PROF_SFORM 10 /* s-step block basis vectors                */
int spc_profile_enable(spc_ctx *ctx, int on);
int spc_profile_read(spc_ctx *ctx, int family, double *total_ms, int64_t *launches, double *bytes);

/* ----------------------------------------------------------- .ini reader */
/* FiNeR file_ini subset used by read_inp (spinsys.f90:60-135): sections,
 * key = value lines, space/comma separated lists, ';' '#' '!' comments. */
typedef struct spc_ini spc_ini;
int spc_ini_open(const char *path, spc_ini **out);
void spc_ini_close(spc_ini *ini);
int spc_ini_num_sections(const spc_ini *ini);
int spc_ini_has_section(const spc_ini *ini, const char *section);
/* Number of values of an option, or -1 if absent. */
int spc_ini_count_values(const spc_ini *ini, const char *section, const char *option);
/* Copies the raw value string (trimmed, quotes removed); 0 or SPC_ERR_INVALID if absent. */
int spc_ini_get_string(const spc_ini *ini, const char *section, const char *option, char *buf,
                       int32_t buflen);
int spc_ini_get_doubles(const spc_ini *ini, const char *section, const char *option, double *out,
                        int32_t n);
int spc_ini_get_int64s(const spc_ini *ini, const char *section, const char *option, int64_t *out,
                       int32_t n);

#ifdef __cplusplus
}
#endif

#endif
