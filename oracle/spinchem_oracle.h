/*
 * spinchem_oracle.h -- CPU oracle for trace-sampling singlet-yield propagation.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the reference
 * algorithm (KenHino/Spin_dynamics, Fortran) used as the checker for the HIP
 * product path and as the timed CPU baseline.  Only tests/, the smoke check and
 * the CPU-baseline leg of the benchmark may load it; the product never does.
 *
 * Each function cites the reference file:line it restates.  Pins (what checks
 * each function independently of this restatement) are listed per function:
 *   - orc_rng_*            : compiled reference m_random.f90 (oracle/_ref) and
 *                            the committed fixture tests/golden/ref_rng.json
 *   - orc_get_kinetics     : compiled reference observables.f90 (oracle/_ref)
 *                            and tests/golden/ref_kinetics.json
 *   - orc_build_iso_csr,
 *     orc_build_general_csr: a numpy dense Kronecker builder written from the
 *                            textbook Hamiltonian (tests/dense_model.py), the
 *                            nnz formula of hamiltonian.f90:62-73
 *   - orc_expm             : scipy.linalg.expm on matrices spanning the
 *                            scaling branches, closed forms
 *   - orc_get_krylov       : Arnoldi relation and orthonormality
 *   - orc_trace_sampling,
 *     orc_exact_dynamics   : dense scipy propagation, kS=kT exact decay,
 *                            no-nuclei closed form
 *   - orc_semiclassical    : no reference output exists; pinned by the uniform
 *                            decay at kS = kT, T12 = S1 S2^T in every sample
 *                            at K = J = 0, conservation of S + I for one
 *                            electron and one nucleus at zero field, and the
 *                            Schulten-Wolynes frozen-nuclei closed form
 */
#ifndef SPINCHEM_ORACLE_H
#define SPINCHEM_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* spinsys.f90:16-19 */
#define ORC_GAMMA_E 176.08596
#define ORC_GAMMA_1H 0.26752218708
#define ORC_GAMMA_14N 0.0193297792

/* ---------------------------------------------------------------- RNG --- */
/* m_random.f90:18-26 (state), 90-98 set_seed, 124-144 jump, 159-167 unif_01,
 * 170-198 normal/two_normals, 280-318 circle/sphere, 321-340 next/rotl. */
typedef struct {
    uint64_t s[2];
    double stored_normal;
    int have_stored_normal;
} orc_rng;

void orc_rng_init_default(orc_rng *r);
void orc_rng_set_seed(orc_rng *r, const int64_t seed[2]);
int64_t orc_rng_next(orc_rng *r);
void orc_rng_jump(orc_rng *r);
double orc_rng_unif01(orc_rng *r);
double orc_rng_normal(orc_rng *r);
void orc_rng_two_normals(orc_rng *r, double out[2]);
void orc_rng_circle(orc_rng *r, double radius, double out[2]);
void orc_rng_sphere(orc_rng *r, double radius, double out[3]);

/* --------------------------------------------------------- spin system --- */
/* Flattened sys_param (spinsys.f90:21-38).  Units as after read_inp: mT for
 * couplings and field, kS/kT already divided by gamma_e (spinsys.f90:88-89). */
typedef struct {
    int n1, n2;          /* nuclei on electron 1 and 2                         */
    const int *mult;     /* n1+n2 multiplicities, electron-1 nuclei first       */
    const double *hfc;   /* 9 per nucleus, column-major: A(a,b) = hfc[9k+3b+a]  */
    double J;            /* exchange, mT                                        */
    double D[9];         /* dipolar tensor, column-major, mT                    */
    double kS, kT;       /* recombination rates in mT units                     */
    double B;            /* field magnitude, mT (w = -B, main.f90:55-56)        */
    double dir[3];       /* unit field direction; (0,0,1) is the reference      */
    int nuclear_1h;      /* 0: 14N's gyromagnetic ratio for multiplicity 3, 1H's
                            otherwise (hamiltonian.f90:132-138); 1: 1H's for
                            every nucleus (the multiplets of a symmetry block) */
} orc_system;

typedef struct {
    int n;          /* rows = cols */
    int64_t nnz;
    int32_t *rowptr; /* n+1, 0-based */
    int32_t *col;    /* nnz, 0-based */
    double *val;     /* 2*nnz, interleaved complex */
} orc_csr;

void orc_csr_free(orc_csr *h);

/* Returns Z1*Z2 (nuclear dimension) or -1 on bad multiplicity. */
int64_t orc_nuclear_dim(const orc_system *s);

/* hamiltonian.f90:17-205 getHiso_sp (COO with the reference entry order, a_iso =
 * trace(A)/3 per nucleus, field along z) followed by utils.f90:52-89 coo_to_csr
 * (within-row order reversed).  Returns 0, or a nonzero code for the reference's
 * stop conditions (multiplicity not in {2,3}: 2, non-axial D: 3). */
int orc_build_iso_csr(const orc_system *s, orc_csr *out);

/* Anisotropic / tilted extension (not in the reference): full hyperfine tensors
 * H_hf = sum_ab A(a,b) S_a I_b (hamiltonian.f90:279-284 convention), electron
 * Zeeman -B n.S, nuclear Zeeman +(B gamma_N/gamma_e) n.I, S1.T.S2 with
 * T = (2J + 2i dk) 1 + D, and -i kbar/2.  Rows hold columns in ascending order;
 * exact zeros off the diagonal are dropped. */
int orc_build_general_csr(const orc_system *s, orc_csr *out);

void orc_csr_spmv(const orc_csr *h, const double *x, double *y);

/* ------------------------------------------------------------- Krylov --- */
/* qm_dynamics.f90:432-467 get_krylov (modified Gram-Schmidt Arnoldi).
 * psi[N], Q[N*m] column-major (vector j at Q + 2*N*j), c[m], A[m*m] col-major. */
void orc_get_krylov(const orc_csr *h, const double *psi, int m,
                    double *Q, double *c, double *A);

/* matrix_exponential.f90:14-117 expm: Pade(6,6) with 2^s scaling from the
 * L-inf norm; D^-1 E through LU with partial pivoting. A, E: m*m col-major. */
void orc_expm(int m, const double *A, double *E);

/* qm_dynamics.f90:469-498 convert_to_krylov. */
void orc_convert_to_krylov(int m, const double *Q, int64_t Z,
                           double *P_S, double *P_T0, double *P_Tp, double *P_Tm);

/* ---------------------------------------------------------- dynamics --- */
typedef struct {
    int n_samples;     /* M                                                    */
    int n_krylov;      /* m                                                    */
    double t_end, dt;  /* already converted: ns * gamma_e / 1000 (spinsys:107) */
    double tol;        /* integrator_tolerance                                 */
    int init_state;    /* 0 singlet, 1 T0, 2 Tp, 3 Tm (qm_dynamics.f90:386)    */
    int vec_kind;      /* 0: SU(Z) normals (sample_SUZ), 1: random phase
                          (unit-circle draws via rng%circle), 2: injected      */
} orc_sim;

/* ceiling((t_end+dt)/dt) (qm_dynamics.f90:305); output arrays hold this + 1. */
int64_t orc_n_steps(double t_end, double dt);
/* Number of steps the float-accumulated time loop takes (t += dt until t >= t_end). */
int64_t orc_steps_taken(double t_end, double dt);

/* qm_dynamics.f90:386-414 init_el_state; theta as 8 doubles (4 complex). */
int orc_init_el_state(int init_state, double theta[8], double init_pop[4]);

/* qm_dynamics.f90:416-430 sample_SUZ (Z complex entries, normals re then im). */
void orc_sample_SUZ(orc_rng *r, int64_t Z, double *zket);
/* Random-phase vector: z_n = (x_n + i y_n)/sqrt(Z), (x_n, y_n) = rng%circle(1)
 * (m_random.f90:280-297), drawn in order n = 1..Z. */
void orc_sample_phase(orc_rng *r, int64_t Z, double *zket);

/* Per-sample seeds: master set_seed(seed) then two next() per sample
 * (main.f90:50, qm_dynamics.f90:315-320).  seeds[2*M]. */
void orc_sample_seeds(const int64_t master_seed[2], int M, int64_t *seeds);

/* Trace sampling with the semantics of qm_dynamics.f90:257-384
 * (trace_sampling_para): per-sample propagation, ordered sum over samples,
 * scale by 1/M.  Outputs have length n_out = orc_n_steps()+1.
 * injected: [M][Z] interleaved complex when vec_kind == 2, else NULL.
 * per_sample_PS: optional [M][n_out] singlet traces per sample (may be NULL).
 * n_threads <= 0 uses the OpenMP default.  Returns 0 or an error code. */
int orc_trace_sampling(const orc_csr *h, int64_t Z, const orc_sim *sim,
                       const int64_t master_seed[2], const double *injected,
                       double *P_S, double *P_T0, double *P_Tp, double *P_Tm,
                       int64_t n_out, double *per_sample_PS, int n_threads);

/* qm_dynamics.f90:24-137 exact_dynamics: the Z nuclear basis states. */
int orc_exact_dynamics(const orc_csr *h, int64_t Z, const orc_sim *sim,
                       double *P_S, double *P_T0, double *P_Tp, double *P_Tm,
                       int64_t n_out, int n_threads);

/* observables.f90:74-101 get_kinetics: cumulative Simpson yields, length n/2. */
void orc_get_kinetics(int64_t n, const double *P_S, const double *P_T0,
                      const double *P_Tp, const double *P_Tm, double dt,
                      double kS, double kT, double *iden, double *Phi_S,
                      double *Phi_T);

/* ---------------------------------------------------- semiclassical --- */
typedef struct {
    int n1, n2;
    const double *a_iso;   /* n1+n2 isotropic couplings, mT                   */
    const int *mult;       /* n1+n2 multiplicities                            */
    double B;              /* field magnitude, mT (electron Zeeman -B n.S)     */
    double dir[3];         /* field direction (normalised here)                */
    double J;              /* exchange, mT (2J S1.S2)                          */
    double kS, kT;         /* mT units                                         */
} orc_sc_system;

/* Semiclassical (Schulten-Wolynes / Manolopoulos-Hore) singlet correlation
 * function.  Semantics of src/sc (sc_dynamics.f90:17-100,
 * class_radical_pair.f90:147-251) with the defects of that code removed.
 * P_S[n_out], n_out = n_steps + 1, and P_T[n_out] (NULL allowed) with
 * P_T = x_16 - P_S per sample (class_radical_pair.f90:174-175).  Samples
 * seeded by orc_sample_seeds. */
int orc_semiclassical(const orc_sc_system *s, int64_t n_samples,
                      const int64_t master_seed[2], double dt, int64_t n_steps,
                      double *P_S, double *P_T, int n_threads);

/* The state of sample `sample` of the same run after every step: xs
 * [n_steps+1][16] (S1, S2, T12 column-major, x_16) and, if non-NULL, Is
 * [n_steps+1][3 (n1+n2)] the nuclear spin vectors. */
int orc_semiclassical_trajectory(const orc_sc_system *s, const int64_t master_seed[2],
                                 int64_t sample, double dt, int64_t n_steps, double *xs,
                                 double *Is);

/* Planted mistakes, for the tests that show the pins above catch them:
 * T12's columns turned with S2's axis and its rows with S1's, the nuclear
 * axis with the wrong sign or twice its length, the electron hyperfine field
 * at half its length.  0 restores the model. */
#define ORC_SC_MUT_NONE 0
#define ORC_SC_MUT_T12_TRANSPOSED 1
#define ORC_SC_MUT_NUCLEAR_SIGN 2
#define ORC_SC_MUT_NUCLEAR_SCALE 3
#define ORC_SC_MUT_FIELD_SCALE 4
void orc_sc_set_mutation(int which);

#ifdef __cplusplus
}
#endif

#endif
