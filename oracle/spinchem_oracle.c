/*
 * spinchem_oracle.c -- CPU oracle (TEST INFRASTRUCTURE ONLY, see the header).
 *
 * Plain-C restatement of the reference algorithm of KenHino/Spin_dynamics:
 * Hamiltonian assembly, xoroshiro128+ sampling, modified Gram-Schmidt Arnoldi,
 * Pade(6,6) expm, Krylov projectors, the tol-controlled inner step loop,
 * ordered sample reduction and Simpson yields.  OpenMP over trace vectors as in
 * qm_dynamics.f90:322.  Citations are to the reference tree.
 */
#include "spinchem_oracle.h"

#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef double complex zd;
#undef I
#define IU _Complex_I

/* ===================================================================== RNG */

/* m_random.f90:20 default state */
void orc_rng_init_default(orc_rng *r)
{
    r->s[0] = 123456789ull;
    r->s[1] = 987654321ull;
    r->stored_normal = 0.0;
    r->have_stored_normal = 0;
}

static inline uint64_t rotl64(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }

/* m_random.f90:321-331: xoroshiro128+ (a=55, b=14, c=36), wrapping add. */
int64_t orc_rng_next(orc_rng *r)
{
    uint64_t s0 = r->s[0], s1 = r->s[1];
    uint64_t res = s0 + s1;
    s1 ^= s0;
    r->s[0] = rotl64(s0, 55) ^ s1 ^ (s1 << 14);
    r->s[1] = rotl64(s1, 36);
    return (int64_t)res;
}

/* m_random.f90:124-144 */
void orc_rng_jump(orc_rng *r)
{
    static const uint64_t jmp[2] = {0xbeac0467eba5facbull, 0xd86b048b86aa9922ull};
    uint64_t t0 = 0, t1 = 0;
    for (int i = 0; i < 2; ++i)
        for (int b = 0; b < 64; ++b) {
            if (jmp[i] & (1ull << b)) {
                t0 ^= r->s[0];
                t1 ^= r->s[1];
            }
            (void)orc_rng_next(r);
        }
    r->s[0] = t0;
    r->s[1] = t1;
}

/* m_random.f90:90-98: the stored normal is deliberately left untouched. */
void orc_rng_set_seed(orc_rng *r, const int64_t seed[2])
{
    r->s[0] = (uint64_t)seed[0];
    r->s[1] = (uint64_t)seed[1];
    orc_rng_jump(r);
}

/* m_random.f90:159-167 */
double orc_rng_unif01(orc_rng *r)
{
    uint64_t x = (uint64_t)orc_rng_next(r);
    x = (1023ull << 52) | (x >> 12);
    double d;
    memcpy(&d, &x, sizeof d);
    return d - 1.0;
}

/* m_random.f90:187-198 (Marsaglia polar) */
void orc_rng_two_normals(orc_rng *r, double out[2])
{
    double a, b, sum_sq;
    for (;;) {
        a = 2.0 * orc_rng_unif01(r) - 1.0;
        b = 2.0 * orc_rng_unif01(r) - 1.0;
        sum_sq = a * a + b * b;
        if (sum_sq < 1.0 && sum_sq > 0.0) break;
    }
    double f = sqrt(-2.0 * log(sum_sq) / sum_sq);
    out[0] = a * f;
    out[1] = b * f;
}

/* m_random.f90:170-183 */
double orc_rng_normal(orc_rng *r)
{
    if (r->have_stored_normal) {
        r->have_stored_normal = 0;
        return r->stored_normal;
    }
    double tn[2];
    orc_rng_two_normals(r, tn);
    r->stored_normal = tn[1];
    r->have_stored_normal = 1;
    return tn[0];
}

/* m_random.f90:280-297 */
void orc_rng_circle(orc_rng *r, double radius, double out[2])
{
    double a, b, sum_sq;
    for (;;) {
        a = 2.0 * orc_rng_unif01(r) - 1.0;
        b = 2.0 * orc_rng_unif01(r) - 1.0;
        sum_sq = a * a + b * b;
        if (sum_sq <= 1.0) break;
    }
    out[0] = (a * a - b * b) / sum_sq * radius;
    out[1] = 2.0 * a * b / sum_sq * radius;
}

/* m_random.f90:300-318 */
void orc_rng_sphere(orc_rng *r, double radius, double out[3])
{
    double a, b, sum_sq;
    for (;;) {
        a = 2.0 * orc_rng_unif01(r) - 1.0;
        b = 2.0 * orc_rng_unif01(r) - 1.0;
        sum_sq = a * a + b * b;
        if (sum_sq <= 1.0) break;
    }
    double tmp = sqrt(1.0 - sum_sq);
    out[0] = 2.0 * a * tmp * radius;
    out[1] = 2.0 * b * tmp * radius;
    out[2] = (1.0 - 2.0 * sum_sq) * radius;
}

/* ============================================================ sparse COO */

/* Minimal COO matrix with the index arithmetic of kroenecker.f90 / utils.f90. */
typedef struct {
    int nrows, ncols;
    int64_t nnz;
    int32_t *row, *col; /* 1-based, as in the reference */
    zd *data;
} coo_t;

static coo_t coo_alloc(int nr, int nc, int64_t nnz)
{
    coo_t c;
    c.nrows = nr;
    c.ncols = nc;
    c.nnz = nnz;
    c.row = (int32_t *)malloc(sizeof(int32_t) * (nnz ? nnz : 1));
    c.col = (int32_t *)malloc(sizeof(int32_t) * (nnz ? nnz : 1));
    c.data = (zd *)malloc(sizeof(zd) * (nnz ? nnz : 1));
    return c;
}

static void coo_free(coo_t *c)
{
    free(c->row);
    free(c->col);
    free(c->data);
    c->row = c->col = NULL;
    c->data = NULL;
}

/* spinop.f90:58-95 get_spinop_sp: Sz (mult entries), S+ and S- (mult-1). */
static void get_spinop_sp(int mult, coo_t *Sz, coo_t *Sp, coo_t *Sm)
{
    double S = ((double)mult - 1.0) / 2.0;
    *Sz = coo_alloc(mult, mult, mult);
    *Sp = coo_alloc(mult, mult, mult - 1);
    *Sm = coo_alloc(mult, mult, mult - 1);
    Sz->data[0] = S;
    Sz->row[0] = Sz->col[0] = 1;
    for (int i = 2; i <= mult; ++i) {
        Sz->data[i - 1] = Sz->data[i - 2] - 1.0;
        Sz->row[i - 1] = Sz->col[i - 1] = i;
        /* sqrt(S(S+1) - Sz(i)*Sz(i-1)) with complex arithmetic (real here) */
        zd v = csqrt(S * (S + 1.0) - Sz->data[i - 1] * Sz->data[i - 2]);
        Sp->data[i - 2] = v;
        Sp->row[i - 2] = i - 1;
        Sp->col[i - 2] = i;
        Sm->data[i - 2] = v;
        Sm->row[i - 2] = i;
        Sm->col[i - 2] = i - 1;
    }
}

/* kroenecker.f90:56-83 kron_sp */
static coo_t kron_sp(const coo_t *A, const coo_t *B)
{
    coo_t C = coo_alloc(A->nrows * B->nrows, A->ncols * B->ncols, A->nnz * B->nnz);
    int64_t k = 0;
    for (int64_t i = 0; i < A->nnz; ++i)
        for (int64_t j = 0; j < B->nnz; ++j) {
            C.row[k] = B->nrows * (A->row[i] - 1) + B->row[j];
            C.col[k] = B->ncols * (A->col[i] - 1) + B->col[j];
            C.data[k] = A->data[i] * B->data[j];
            ++k;
        }
    return C;
}

/* kroenecker.f90:149-172 kron_iden_right_sp: A (x) 1_N */
static coo_t kron_iden_right(const coo_t *A, int N)
{
    coo_t B = coo_alloc(A->nrows * N, A->ncols * N, A->nnz * N);
    for (int64_t i = 0; i < A->nnz; ++i)
        for (int j = 0; j < N; ++j) {
            B.row[i * N + j] = 1 + (A->row[i] - 1) * N + j;
            B.col[i * N + j] = 1 + (A->col[i] - 1) * N + j;
            B.data[i * N + j] = A->data[i];
        }
    return B;
}

/* kroenecker.f90:174-200 kron_iden_left_sp: 1_A (x) N */
static coo_t kron_iden_left(int A, const coo_t *Nm)
{
    coo_t B = coo_alloc(A * Nm->nrows, A * Nm->ncols, A * Nm->nnz);
    for (int i = 0; i < A; ++i)
        for (int64_t j = 0; j < Nm->nnz; ++j) {
            B.row[i * Nm->nnz + j] = Nm->row[j] + i * Nm->nrows;
            B.col[i * Nm->nnz + j] = Nm->col[j] + i * Nm->nrows;
            B.data[i * Nm->nnz + j] = Nm->data[j];
        }
    return B;
}

/* utils.f90:91-113 scale (returns a copy) */
static coo_t coo_scale(zd c, const coo_t *A)
{
    coo_t B = coo_alloc(A->nrows, A->ncols, A->nnz);
    memcpy(B.row, A->row, sizeof(int32_t) * A->nnz);
    memcpy(B.col, A->col, sizeof(int32_t) * A->nnz);
    for (int64_t i = 0; i < A->nnz; ++i) B.data[i] = c * A->data[i];
    return B;
}

/* utils.f90:116-129 add (identical index arrays required) */
static int coo_add_inplace(coo_t *A, const coo_t *B)
{
    if (A->nnz != B->nnz) return 1;
    for (int64_t i = 0; i < A->nnz; ++i)
        if (A->row[i] != B->row[i] || A->col[i] != B->col[i]) return 1;
    for (int64_t i = 0; i < A->nnz; ++i) A->data[i] += B->data[i];
    return 0;
}

/* utils.f90:147-160 eye_cp_sparse */
static coo_t coo_eye(int N)
{
    coo_t I = coo_alloc(N, N, N);
    for (int i = 0; i < N; ++i) {
        I.row[i] = I.col[i] = i + 1;
        I.data[i] = 1.0;
    }
    return I;
}

/* utils.f90:52-89 coo_to_csr: counting sort by row; within a row the entries
 * land in reverse COO order. */
static void coo_to_csr(const coo_t *C, orc_csr *H)
{
    int n = C->nrows;
    H->n = n;
    H->nnz = C->nnz;
    H->rowptr = (int32_t *)calloc((size_t)n + 1, sizeof(int32_t));
    H->col = (int32_t *)malloc(sizeof(int32_t) * (C->nnz ? C->nnz : 1));
    H->val = (double *)malloc(sizeof(double) * 2 * (C->nnz ? C->nnz : 1));
    int32_t *row_elem = (int32_t *)calloc((size_t)n, sizeof(int32_t));
    for (int64_t i = 0; i < C->nnz; ++i) row_elem[C->row[i] - 1]++;
    H->rowptr[0] = 0;
    for (int i = 0; i < n; ++i) H->rowptr[i + 1] = H->rowptr[i] + row_elem[i];
    for (int64_t i = 0; i < C->nnz; ++i) {
        int r = C->row[i] - 1;
        int64_t idx = H->rowptr[r] + row_elem[r] - 1;
        row_elem[r]--;
        H->col[idx] = C->col[i] - 1;
        H->val[2 * idx] = creal(C->data[i]);
        H->val[2 * idx + 1] = cimag(C->data[i]);
    }
    free(row_elem);
}

void orc_csr_free(orc_csr *h)
{
    free(h->rowptr);
    free(h->col);
    free(h->val);
    h->rowptr = NULL;
    h->col = NULL;
    h->val = NULL;
}

int64_t orc_nuclear_dim(const orc_system *s)
{
    int64_t Z = 1;
    for (int k = 0; k < s->n1 + s->n2; ++k) {
        if (s->mult[k] < 2) return -1;
        Z *= s->mult[k];
    }
    return Z;
}

static int64_t prod_range(const int *m, int a, int b) /* product of m[a..b) */
{
    int64_t p = 1;
    for (int i = a; i < b; ++i) p *= m[i];
    return p;
}

/* hamiltonian.f90:17-205 getHiso_sp, then utils.f90:52 coo_to_csr. */
int orc_build_iso_csr(const orc_system *s, orc_csr *out)
{
    const int n1 = s->n1, n2 = s->n2;
    const int *g = s->mult;
    for (int k = 0; k < n1 + n2; ++k)
        if (g[k] != 2 && g[k] != 3) return 2; /* hamiltonian.f90:137 */
    /* D(1,1) = D[0], D(2,2) = D[4], D(3,3) = D[8] (column-major) */
    if (fabs(s->D[0] - s->D[4]) > 1.0e-12) return 3;          /* :93-95 */
    if (fabs(s->D[0] + 0.5 * s->D[8]) > 1.0e-12) return 3;    /* :96-98 */

    const int Z1 = (int)prod_range(g, 0, n1), Z2 = (int)prod_range(g, n1, n1 + n2);
    const int N = 4 * Z1 * Z2;
    const double w = -s->B; /* main.f90:55-56 */
    const double k_bar = (s->kS + 3.0 * s->kT) / 4.0;
    const double delta_k = (s->kS - s->kT) / 4.0;

    coo_t Sz, Sp, Sm;
    get_spinop_sp(2, &Sz, &Sp, &Sm);
    coo_t S1p = kron_iden_right(&Sp, 2), S1m = kron_iden_right(&Sm, 2), S1z = kron_iden_right(&Sz, 2);
    coo_t S2p = kron_iden_left(2, &Sp), S2m = kron_iden_left(2, &Sm), S2z = kron_iden_left(2, &Sz);
    coo_t S1pS2m = kron_sp(&Sp, &Sm), S1mS2p = kron_sp(&Sm, &Sp), S1zS2z = kron_sp(&Sz, &Sz);
    coo_t E = coo_eye(4);

    /* nnz formula, hamiltonian.f90:62-73 */
    int64_t nnz = N + N / 2;
    for (int k = 0; k < n1 + n2; ++k) nnz += (int64_t)(g[k] - 1) * (N / g[k]);

    coo_t H = coo_alloc(N, N, nnz);
    int64_t start = N;

    /* Hel_diag (hamiltonian.f90:79-86) */
    coo_t t1 = coo_scale(w, &S1z), t2 = coo_scale(w, &S2z);
    zd exch = (2.0 * s->J + 2.0 * IU * delta_k) + s->D[8];
    coo_t t3 = coo_scale(exch, &S1zS2z);
    coo_add_inplace(&t1, &t2);
    coo_add_inplace(&t1, &t3);
    for (int64_t i = 0; i < t1.nnz; ++i) t1.data[i] -= 0.5 * IU * k_bar;
    coo_t Ht = kron_iden_right(&t1, Z1 * Z2);
    for (int64_t i = 0; i < N; ++i) {
        H.row[i] = Ht.row[i];
        H.col[i] = Ht.col[i];
        H.data[i] = Ht.data[i];
    }
    coo_free(&Ht);
    coo_free(&t1);
    coo_free(&t2);
    coo_free(&t3);

#define APPEND(src)                                        \
    do {                                                   \
        for (int64_t q = 0; q < (src).nnz; ++q) {          \
            H.row[start + q] = (src).row[q];               \
            H.col[start + q] = (src).col[q];               \
            H.data[start + q] = (src).data[q];             \
        }                                                  \
        start += (src).nnz;                                \
    } while (0)

    /* flip-flop terms (hamiltonian.f90:100-110) */
    {
        coo_t a = coo_scale((s->J + IU * delta_k) + 0.5 * s->D[0], &S1pS2m);
        coo_t b = kron_iden_right(&a, Z1 * Z2);
        APPEND(b);
        coo_free(&a);
        coo_free(&b);
        a = coo_scale((s->J + IU * delta_k) + 0.5 * s->D[4], &S1mS2p);
        b = kron_iden_right(&a, Z1 * Z2);
        APPEND(b);
        coo_free(&a);
        coo_free(&b);
    }

    /* nuclei of both electrons (hamiltonian.f90:112-203) */
    for (int el = 0; el < 2; ++el) {
        int first = el == 0 ? 0 : n1, count = el == 0 ? n1 : n2;
        coo_t *Sez = el == 0 ? &S1z : &S2z, *Sep = el == 0 ? &S1p : &S2p, *Sem = el == 0 ? &S1m : &S2m;
        for (int i = 0; i < count; ++i) {
            int k = first + i;
            int64_t Zl = el == 0 ? prod_range(g, 0, k) : (int64_t)Z1 * prod_range(g, n1, k);
            int64_t Zr = el == 0 ? prod_range(g, k + 1, n1) * Z2 : prod_range(g, k + 1, n1 + n2);
            const double *A = s->hfc + 9 * k;
            double a_iso = (A[0] + A[4] + A[8]) / 3.0; /* spinsys.f90:232 */
            double gam = g[k] == 3 && !s->nuclear_1h ? ORC_GAMMA_14N : ORC_GAMMA_1H;
            double zeeman_factor = -w * gam / ORC_GAMMA_E;
            coo_t Iz, Ip, Im;
            get_spinop_sp(g[k], &Iz, &Ip, &Im);

            /* diagonal: kron(kron_iden(a*Sez + zf*E, Zl), kron_iden(Iz, Zr)) */
            coo_t x1 = coo_scale(a_iso, Sez), x2 = coo_scale(zeeman_factor, &E);
            coo_add_inplace(&x1, &x2);
            coo_t X = kron_iden_right(&x1, (int)Zl);
            coo_t Y = kron_iden_right(&Iz, (int)Zr);
            coo_t T = kron_sp(&X, &Y);
            for (int64_t q = 0; q < N; ++q) H.data[q] += T.data[q];
            coo_free(&x1);
            coo_free(&x2);
            coo_free(&X);
            coo_free(&Y);
            coo_free(&T);

            /* S+ I- then S- I+ */
            for (int pm = 0; pm < 2; ++pm) {
                coo_t xs = coo_scale(0.5 * a_iso, pm == 0 ? Sep : Sem);
                coo_t Xs = kron_iden_right(&xs, (int)Zl);
                coo_t Ys = kron_iden_right(pm == 0 ? &Im : &Ip, (int)Zr);
                coo_t Ts = kron_sp(&Xs, &Ys);
                APPEND(Ts);
                coo_free(&xs);
                coo_free(&Xs);
                coo_free(&Ys);
                coo_free(&Ts);
            }
            coo_free(&Iz);
            coo_free(&Ip);
            coo_free(&Im);
        }
    }
#undef APPEND
    int rc = start == nnz ? 0 : 4;
    if (rc == 0) coo_to_csr(&H, out);
    coo_free(&H);
    coo_t *all[] = {&Sz, &Sp, &Sm, &S1p, &S1m, &S1z, &S2p, &S2m, &S2z, &S1pS2m, &S1mS2p, &S1zS2z, &E};
    for (size_t i = 0; i < sizeof all / sizeof all[0]; ++i) coo_free(all[i]);
    return rc;
}

/* ---------------------------------------------- general (anisotropic) H --- */

/* <d_row| op |d_col> for a spin of multiplicity g, basis m_d = S - d.
 * op: 0 = x, 1 = y, 2 = z. */
static zd spin_elem(int g, int op, int dr, int dc)
{
    double S = 0.5 * (g - 1);
    double mr = S - dr, mc = S - dc;
    if (op == 2) return dr == dc ? mr : 0.0;
    double sp = 0.0, sm = 0.0; /* <r|S+|c>, <r|S-|c> */
    if (dr == dc - 1) sp = sqrt(S * (S + 1.0) - mc * (mc + 1.0));
    if (dr == dc + 1) sm = sqrt(S * (S + 1.0) - mc * (mc - 1.0));
    if (op == 0) return 0.5 * (sp + sm);
    return -0.5 * IU * (sp - sm); /* Sy = (S+ - S-)/(2i) */
}

typedef struct {
    int64_t col;
    zd val;
} entry_t;

static int cmp_entry(const void *a, const void *b)
{
    int64_t x = ((const entry_t *)a)->col, y = ((const entry_t *)b)->col;
    return (x > y) - (x < y);
}

/* Candidate entries of one row, in any order; returns count.  Digits: electron
 * e = 2*s1 + s2 (0 = alpha), nuclear digits d_k (m = I - d), electron-1 nuclei
 * slowest.  stride[k] = product of later multiplicities within the nuclear index. */
static int general_row(const orc_system *s, const int64_t *stride, int64_t Z, int64_t row,
                       entry_t *out)
{
    const int nn = s->n1 + s->n2;
    const double w[3] = {-s->B * s->dir[0], -s->B * s->dir[1], -s->B * s->dir[2]};
    const double dk = (s->kS - s->kT) / 4.0, kbar = (s->kS + 3.0 * s->kT) / 4.0;
    zd T[3][3];
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) T[a][b] = s->D[3 * b + a] + (a == b ? 2.0 * s->J + 2.0 * IU * dk : 0.0);

    int64_t e = row / Z, nuc = row % Z;
    int s1 = (int)(e >> 1), s2 = (int)(e & 1);
    int d[64];
    for (int k = 0; k < nn; ++k) d[k] = (int)((nuc / stride[k]) % s->mult[k]);
    int cnt = 0;

    /* diagonal */
    zd diag = -0.5 * IU * kbar;
    diag += w[2] * (spin_elem(2, 2, s1, s1) + spin_elem(2, 2, s2, s2));
    diag += T[2][2] * spin_elem(2, 2, s1, s1) * spin_elem(2, 2, s2, s2);
    for (int k = 0; k < nn; ++k) {
        const double *A = s->hfc + 9 * k;
        int se = k < s->n1 ? s1 : s2;
        double gam = s->mult[k] == 3 && !s->nuclear_1h ? ORC_GAMMA_14N : ORC_GAMMA_1H;
        double zeta = s->B * gam / ORC_GAMMA_E;
        double mz = 0.5 * (s->mult[k] - 1) - d[k];
        diag += A[8] * spin_elem(2, 2, se, se) * mz + zeta * s->dir[2] * mz;
    }
    out[cnt].col = row;
    out[cnt].val = diag;
    cnt++;

    /* single electron flips */
    for (int el = 0; el < 2; ++el) {
        int sr = el == 0 ? s1 : s2, sc = 1 - sr, so = el == 0 ? s2 : s1;
        zd v = 0.0;
        for (int a = 0; a < 2; ++a) {
            zd ea = spin_elem(2, a, sr, sc);
            zd f = w[a];
            f += (el == 0 ? T[a][2] : T[2][a]) * spin_elem(2, 2, so, so);
            int k0 = el == 0 ? 0 : s->n1, k1 = el == 0 ? s->n1 : nn;
            for (int k = k0; k < k1; ++k) f += s->hfc[9 * k + 3 * 2 + a] * (0.5 * (s->mult[k] - 1) - d[k]);
            v += ea * f;
        }
        int64_t ce = el == 0 ? 2 * sc + s2 : 2 * s1 + sc;
        if (v != 0.0) {
            out[cnt].col = ce * Z + nuc;
            out[cnt].val = v;
            cnt++;
        }
    }
    /* both electrons flip */
    {
        zd v = 0.0;
        for (int a = 0; a < 2; ++a)
            for (int b = 0; b < 2; ++b)
                v += T[a][b] * spin_elem(2, a, s1, 1 - s1) * spin_elem(2, b, s2, 1 - s2);
        if (v != 0.0) {
            out[cnt].col = (int64_t)(2 * (1 - s1) + (1 - s2)) * Z + nuc;
            out[cnt].val = v;
            cnt++;
        }
    }
    /* nuclear flips, with and without the flip of its electron */
    for (int k = 0; k < nn; ++k) {
        const double *A = s->hfc + 9 * k;
        int g = s->mult[k];
        int el = k < s->n1 ? 0 : 1;
        int se = el == 0 ? s1 : s2;
        double gam = g == 3 && !s->nuclear_1h ? ORC_GAMMA_14N : ORC_GAMMA_1H;
        double zeta = s->B * gam / ORC_GAMMA_E;
        for (int dd = -1; dd <= 1; dd += 2) {
            int dc = d[k] + dd;
            if (dc < 0 || dc >= g) continue;
            int64_t cn = nuc + (int64_t)dd * stride[k];
            zd v = 0.0;
            for (int b = 0; b < 2; ++b) {
                zd ib = spin_elem(g, b, d[k], dc);
                v += ib * (zeta * s->dir[b] + A[3 * b + 2] * spin_elem(2, 2, se, se));
            }
            if (v != 0.0) {
                out[cnt].col = e * Z + cn;
                out[cnt].val = v;
                cnt++;
            }
            zd v2 = 0.0;
            for (int a = 0; a < 2; ++a)
                for (int b = 0; b < 2; ++b)
                    v2 += A[3 * b + a] * spin_elem(2, a, se, 1 - se) * spin_elem(g, b, d[k], dc);
            if (v2 != 0.0) {
                int64_t ce = el == 0 ? 2 * (1 - s1) + s2 : 2 * s1 + (1 - s2);
                out[cnt].col = ce * Z + cn;
                out[cnt].val = v2;
                cnt++;
            }
        }
    }
    qsort(out, (size_t)cnt, sizeof(entry_t), cmp_entry);
    return cnt;
}

int orc_build_general_csr(const orc_system *s, orc_csr *out)
{
    const int nn = s->n1 + s->n2;
    if (nn > 64) return 1;
    for (int k = 0; k < nn; ++k)
        if (s->mult[k] < 2) return 2;
    int64_t Z = orc_nuclear_dim(s), N = 4 * Z;
    int64_t stride[64];
    {
        int64_t p = 1;
        for (int k = nn - 1; k >= 0; --k) {
            stride[k] = p;
            p *= s->mult[k];
        }
    }
    int maxc = 4 + 4 * nn;
    entry_t *buf = (entry_t *)malloc(sizeof(entry_t) * maxc);
    int64_t nnz = 0;
    for (int64_t r = 0; r < N; ++r) nnz += general_row(s, stride, Z, r, buf);
    out->n = (int)N;
    out->nnz = nnz;
    out->rowptr = (int32_t *)malloc(sizeof(int32_t) * (N + 1));
    out->col = (int32_t *)malloc(sizeof(int32_t) * (nnz ? nnz : 1));
    out->val = (double *)malloc(sizeof(double) * 2 * (nnz ? nnz : 1));
    int64_t pos = 0;
    for (int64_t r = 0; r < N; ++r) {
        out->rowptr[r] = (int32_t)pos;
        int c = general_row(s, stride, Z, r, buf);
        for (int q = 0; q < c; ++q) {
            out->col[pos] = (int32_t)buf[q].col;
            out->val[2 * pos] = creal(buf[q].val);
            out->val[2 * pos + 1] = cimag(buf[q].val);
            pos++;
        }
    }
    out->rowptr[N] = (int32_t)pos;
    free(buf);
    return 0;
}

/* stdlib spmv semantics: y = H x, accumulated in stored order. */
void orc_csr_spmv(const orc_csr *h, const double *xr, double *yr)
{
    const zd *x = (const zd *)xr;
    zd *y = (zd *)yr;
    const zd *v = (const zd *)h->val;
    for (int i = 0; i < h->n; ++i) {
        zd acc = 0.0;
        for (int32_t k = h->rowptr[i]; k < h->rowptr[i + 1]; ++k) acc += v[k] * x[h->col[k]];
        y[i] = acc;
    }
}

/* =============================================================== Krylov */

static zd zdot(int64_t n, const zd *a, const zd *b) /* Fortran dot_product */
{
    zd s = 0.0;
    for (int64_t i = 0; i < n; ++i) s += conj(a[i]) * b[i];
    return s;
}

/* qm_dynamics.f90:432-467 */
void orc_get_krylov(const orc_csr *h, const double *psir, int m, double *Qr, double *cr, double *Ar)
{
    const int64_t N = h->n;
    const zd *psi = (const zd *)psir;
    zd *Q = (zd *)Qr, *c = (zd *)cr, *A = (zd *)Ar;
    zd *p = (zd *)malloc(sizeof(zd) * N);
    for (int i = 0; i < m; ++i) c[i] = 0.0;
    for (int i = 0; i < m * m; ++i) A[i] = 0.0;
    c[0] = csqrt(zdot(N, psi, psi));
    for (int64_t n = 0; n < N; ++n) Q[n] = psi[n] / c[0];
    for (int i = 0; i < m - 1; ++i) {
        orc_csr_spmv(h, (const double *)(Q + N * i), (double *)p);
        for (int j = 0; j <= i; ++j) {
            zd a = zdot(N, Q + N * j, p);
            A[j + m * i] = a;
            for (int64_t n = 0; n < N; ++n) p[n] -= a * Q[N * j + n];
        }
        zd nrm = csqrt(zdot(N, p, p));
        A[(i + 1) + m * i] = nrm;
        for (int64_t n = 0; n < N; ++n) Q[N * (i + 1) + n] = p[n] / nrm;
    }
    orc_csr_spmv(h, (const double *)(Q + N * (m - 1)), (double *)p);
    for (int j = 0; j < m; ++j) {
        zd a = zdot(N, Q + N * j, p);
        A[j + m * (m - 1)] = a;
        for (int64_t n = 0; n < N; ++n) p[n] -= a * Q[N * j + n];
    }
    free(p);
}

static void matmul(int m, const zd *A, const zd *B, zd *C) /* col-major C = A B */
{
    for (int j = 0; j < m; ++j)
        for (int i = 0; i < m; ++i) {
            zd s = 0.0;
            for (int k = 0; k < m; ++k) s += A[i + m * k] * B[k + m * j];
            C[i + m * j] = s;
        }
}

/* LU with partial pivoting, then D^{-1} via solving D X = I (LAPACK getrf/getri
 * semantics up to rounding). */
static int inv_lu(int m, const zd *D, zd *Dinv)
{
    zd *L = (zd *)malloc(sizeof(zd) * m * m);
    int *piv = (int *)malloc(sizeof(int) * m);
    memcpy(L, D, sizeof(zd) * m * m);
    int rc = 0;
    for (int k = 0; k < m; ++k) {
        int p = k;
        double best = cabs(L[k + m * k]);
        for (int i = k + 1; i < m; ++i)
            if (cabs(L[i + m * k]) > best) {
                best = cabs(L[i + m * k]);
                p = i;
            }
        piv[k] = p;
        if (best == 0.0) rc = 1;
        if (p != k)
            for (int j = 0; j < m; ++j) {
                zd t = L[k + m * j];
                L[k + m * j] = L[p + m * j];
                L[p + m * j] = t;
            }
        for (int i = k + 1; i < m; ++i) {
            L[i + m * k] /= L[k + m * k];
            for (int j = k + 1; j < m; ++j) L[i + m * j] -= L[i + m * k] * L[k + m * j];
        }
    }
    for (int col = 0; col < m; ++col) {
        zd *x = Dinv + m * col;
        for (int i = 0; i < m; ++i) x[i] = i == col ? 1.0 : 0.0;
        for (int k = 0; k < m; ++k)
            if (piv[k] != k) {
                zd t = x[k];
                x[k] = x[piv[k]];
                x[piv[k]] = t;
            }
        for (int i = 0; i < m; ++i)
            for (int k = 0; k < i; ++k) x[i] -= L[i + m * k] * x[k];
        for (int i = m - 1; i >= 0; --i) {
            for (int k = i + 1; k < m; ++k) x[i] -= L[i + m * k] * x[k];
            x[i] /= L[i + m * i];
        }
    }
    free(L);
    free(piv);
    return rc;
}

/* matrix_exponential.f90:14-117 */
void orc_expm(int m, const double *Ar, double *Er)
{
    const zd *A = (const zd *)Ar;
    zd *E = (zd *)Er;
    const int q = 6;
    zd *A2 = (zd *)malloc(sizeof(zd) * m * m);
    zd *X = (zd *)malloc(sizeof(zd) * m * m);
    zd *T = (zd *)malloc(sizeof(zd) * m * m);
    zd *D = (zd *)malloc(sizeof(zd) * m * m);
    memcpy(A2, A, sizeof(zd) * m * m);
    double a_norm = 0.0; /* mat_norm_linf, :89-104 */
    for (int i = 0; i < m; ++i) {
        double rs = 0.0;
        for (int j = 0; j < m; ++j) rs += cabs(A2[i + m * j]);
        if (rs > a_norm) a_norm = rs;
    }
    int s;
    if (a_norm == 0.0) {
        s = 0;
    } else {
        int ee = (int)(log(fabs(a_norm)) / log(2.0)) + 1; /* log_2, :106-117; int truncates */
        s = ee + 1 > 0 ? ee + 1 : 0;
    }
    double scale = pow(2.0, s);
    for (int i = 0; i < m * m; ++i) A2[i] /= scale;
    memcpy(X, A2, sizeof(zd) * m * m);
    double c = 0.5;
    for (int i = 0; i < m * m; ++i) {
        zd id = (i % (m + 1) == 0) ? 1.0 : 0.0;
        E[i] = id + c * A2[i];
        D[i] = id - c * A2[i];
    }
    int p = 1;
    for (int k = 2; k <= q; ++k) {
        c = c * (double)(q - k + 1) / (double)(k * (2 * q - k + 1));
        matmul(m, A2, X, T); /* X = matmul(A2, X): x and X are one variable */
        memcpy(X, T, sizeof(zd) * m * m);
        for (int i = 0; i < m * m; ++i) {
            E[i] += c * X[i];
            if (p)
                D[i] += c * X[i];
            else
                D[i] -= c * X[i];
        }
        p = !p;
    }
    inv_lu(m, D, T);
    matmul(m, T, E, X);
    memcpy(E, X, sizeof(zd) * m * m);
    for (int k = 0; k < s; ++k) {
        matmul(m, E, E, X);
        memcpy(E, X, sizeof(zd) * m * m);
    }
    free(A2);
    free(X);
    free(T);
    free(D);
}

/* qm_dynamics.f90:469-498 */
void orc_convert_to_krylov(int m, const double *Qr, int64_t Z, double *PSr, double *PT0r, double *PTpr,
                           double *PTmr)
{
    const zd *Q = (const zd *)Qr;
    const int64_t N = 4 * Z;
    zd *PS = (zd *)PSr, *PT0 = (zd *)PT0r, *PTp = (zd *)PTpr, *PTm = (zd *)PTmr;
    for (int i = 0; i < m; ++i)
        for (int j = 0; j < m; ++j) {
            const zd *qi = Q + N * i, *qj = Q + N * j;
            PTp[i + m * j] = zdot(Z, qi, qj);
            zd Pab = zdot(Z, qi + Z, qj + Z);
            zd Pba = zdot(Z, qi + 2 * Z, qj + 2 * Z);
            zd ZQC = zdot(Z, qi + 2 * Z, qj + Z);
            PT0[i + m * j] = 0.5 * (Pab + Pba) + ZQC;
            PS[i + m * j] = 0.5 * (Pab + Pba) - ZQC;
            PTm[i + m * j] = zdot(Z, qi + 3 * Z, qj + 3 * Z);
        }
}

/* ============================================================= dynamics */

int64_t orc_n_steps(double t_end, double dt) { return (int64_t)ceil((t_end + dt) / dt); }

int64_t orc_steps_taken(double t_end, double dt)
{
    double t = 0.0;
    int64_t k = 0;
    while (t < t_end) {
        t = t + dt;
        k++;
    }
    return k;
}

/* qm_dynamics.f90:386-414 */
int orc_init_el_state(int st, double th[8], double pop[4])
{
    for (int i = 0; i < 8; ++i) th[i] = 0.0;
    for (int i = 0; i < 4; ++i) pop[i] = 0.0;
    double r = 1.0 / sqrt(2.0);
    switch (st) {
    case 0: th[2] = r; th[4] = -r; pop[0] = 1.0; break;
    case 1: th[2] = r; th[4] = r; pop[1] = 1.0; break;
    case 2: th[0] = 1.0; pop[2] = 1.0; break;
    case 3: th[6] = 1.0; pop[3] = 1.0; break;
    default: return 1;
    }
    return 0;
}

/* qm_dynamics.f90:416-430 */
void orc_sample_SUZ(orc_rng *r, int64_t Z, double *zr)
{
    zd *z = (zd *)zr;
    for (int64_t i = 0; i < Z; ++i) {
        double re = orc_rng_normal(r);
        double im = orc_rng_normal(r);
        z[i] = re + IU * im;
    }
    zd nrm = csqrt(zdot(Z, z, z));
    for (int64_t i = 0; i < Z; ++i) z[i] = z[i] / nrm;
}

void orc_sample_phase(orc_rng *r, int64_t Z, double *zr)
{
    zd *z = (zd *)zr;
    double f = 1.0 / sqrt((double)Z);
    for (int64_t i = 0; i < Z; ++i) {
        double xy[2];
        orc_rng_circle(r, 1.0, xy);
        z[i] = (xy[0] + IU * xy[1]) * f;
    }
}

void orc_sample_seeds(const int64_t master[2], int M, int64_t *seeds)
{
    orc_rng r;
    orc_rng_init_default(&r);
    orc_rng_set_seed(&r, master);
    for (int i = 0; i < M; ++i) {
        seeds[2 * i] = orc_rng_next(&r);
        seeds[2 * i + 1] = orc_rng_next(&r);
    }
}

/* One sample, qm_dynamics.f90:326-366.  Z_ket given; writes P[4][n_out]. */
static int propagate_sample(const orc_csr *h, int64_t Z, const orc_sim *sim, const zd *theta, const double *pop,
                            const zd *zket, double *P, int64_t n_out)
{
    const int m = sim->n_krylov;
    const int64_t N = 4 * Z;
    zd *psi = (zd *)malloc(sizeof(zd) * N);
    zd *Q = (zd *)malloc(sizeof(zd) * N * m);
    zd *A = (zd *)malloc(sizeof(zd) * m * m);
    zd *Ae = (zd *)malloc(sizeof(zd) * m * m);
    zd *X = (zd *)malloc(sizeof(zd) * m * m);
    zd *PX[4];
    for (int i = 0; i < 4; ++i) PX[i] = (zd *)malloc(sizeof(zd) * m * m);
    zd *c = (zd *)malloc(sizeof(zd) * m), *cn = (zd *)malloc(sizeof(zd) * m), *tmp = (zd *)malloc(sizeof(zd) * m);
    for (int e = 0; e < 4; ++e)
        for (int64_t n = 0; n < Z; ++n) psi[e * Z + n] = theta[e] * zket[n]; /* kron_vector */
    for (int i = 0; i < 4; ++i)
        for (int64_t k = 0; k < n_out; ++k) P[i * n_out + k] = 0.0;
    for (int i = 0; i < 4; ++i) P[i * n_out] = pop[i];
    double t = 0.0;
    int64_t k = 1;
    int rc = 0;
    while (t < sim->t_end) {
        orc_get_krylov(h, (const double *)psi, m, (double *)Q, (double *)c, (double *)A);
        for (int i = 0; i < m * m; ++i) X[i] = -IU * sim->dt * A[i];
        orc_expm(m, (const double *)X, (double *)Ae);
        /* order in the reference arrays: P_S, P_T0, P_Tp, P_Tm */
        orc_convert_to_krylov(m, (const double *)Q, Z, (double *)PX[0], (double *)PX[1], (double *)PX[2],
                              (double *)PX[3]);
        double c2 = 1.0, c_m2 = 0.0;
        while (c_m2 <= sim->tol * c2 && t < sim->t_end) {
            for (int i = 0; i < m; ++i) {
                zd s = 0.0;
                for (int j = 0; j < m; ++j) s += Ae[i + m * j] * c[j];
                cn[i] = s;
            }
            memcpy(c, cn, sizeof(zd) * m);
            if (k >= n_out) {
                rc = 5;
                goto done;
            }
            for (int x = 0; x < 4; ++x) {
                for (int i = 0; i < m; ++i) {
                    zd s = 0.0;
                    for (int j = 0; j < m; ++j) s += PX[x][i + m * j] * c[j];
                    tmp[i] = s;
                }
                P[x * n_out + k] = creal(zdot(m, c, tmp));
            }
            c2 = creal(zdot(m, c, c));
            c_m2 = creal(c[m - 1] * conj(c[m - 1]));
            t = t + sim->dt;
            k = k + 1;
        }
        for (int64_t n = 0; n < N; ++n) { /* psi = matmul(Q, c) */
            zd s = 0.0;
            for (int j = 0; j < m; ++j) s += Q[N * j + n] * c[j];
            psi[n] = s;
        }
    }
done:
    free(psi);
    free(Q);
    free(A);
    free(Ae);
    free(X);
    for (int i = 0; i < 4; ++i) free(PX[i]);
    free(c);
    free(cn);
    free(tmp);
    return rc;
}

/* Shared driver: M vectors produced by make_vec(i), ordered reduction. */
static int run_samples(const orc_csr *h, int64_t Z, const orc_sim *sim, int M, const int64_t *seeds,
                       const double *injected, int basis, double *P_S, double *P_T0, double *P_Tp,
                       double *P_Tm, int64_t n_out, double *per_sample_PS, int n_threads)
{
    double th[8], pop[4];
    if (orc_init_el_state(sim->init_state, th, pop)) return 1;
    if (sim->n_krylov < 1 || (int64_t)h->n != 4 * Z) return 1;
    double *Pall = (double *)malloc(sizeof(double) * 4 * n_out * (size_t)M);
    if (!Pall) return 3;
    int rc_all = 0;
#ifdef _OPENMP
    int nt = n_threads > 0 ? n_threads : omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 1) num_threads(nt) reduction(| : rc_all)
#endif
    for (int i = 0; i < M; ++i) {
        zd *zket = (zd *)malloc(sizeof(zd) * Z);
        if (basis) {
            for (int64_t n = 0; n < Z; ++n) zket[n] = n == i ? 1.0 : 0.0;
        } else if (sim->vec_kind == 2) {
            memcpy(zket, injected + 2 * Z * (int64_t)i, sizeof(zd) * Z);
        } else {
            orc_rng r; /* fresh private copy: no stored normal (m_random.f90:24) */
            orc_rng_init_default(&r);
            orc_rng_set_seed(&r, seeds + 2 * i);
            if (sim->vec_kind == 1)
                orc_sample_phase(&r, Z, (double *)zket);
            else
                orc_sample_SUZ(&r, Z, (double *)zket);
        }
        rc_all |= propagate_sample(h, Z, sim, (const zd *)th, pop, zket, Pall + 4 * n_out * (size_t)i, n_out);
        free(zket);
    }
    /* qm_dynamics.f90:373-378: res = sum_i sample_obs(i) in order, then scale */
    double *outs[4] = {P_S, P_T0, P_Tp, P_Tm};
    for (int x = 0; x < 4; ++x)
        for (int64_t k = 0; k < n_out; ++k) {
            double s = 0.0;
            for (int i = 0; i < M; ++i) s += Pall[4 * n_out * (size_t)i + x * n_out + k];
            outs[x][k] = s * (1.0 / (double)M);
        }
    if (per_sample_PS)
        for (int i = 0; i < M; ++i)
            memcpy(per_sample_PS + n_out * (size_t)i, Pall + 4 * n_out * (size_t)i, sizeof(double) * n_out);
    free(Pall);
    return rc_all;
}

int orc_trace_sampling(const orc_csr *h, int64_t Z, const orc_sim *sim, const int64_t master_seed[2],
                       const double *injected, double *P_S, double *P_T0, double *P_Tp, double *P_Tm,
                       int64_t n_out, double *per_sample_PS, int n_threads)
{
    int M = sim->n_samples;
    if (M < 1) return 1;
    if (sim->vec_kind == 2 && !injected) return 1;
    int64_t *seeds = (int64_t *)malloc(sizeof(int64_t) * 2 * M);
    orc_sample_seeds(master_seed, M, seeds);
    int rc = run_samples(h, Z, sim, M, seeds, injected, 0, P_S, P_T0, P_Tp, P_Tm, n_out, per_sample_PS, n_threads);
    free(seeds);
    return rc;
}

int orc_exact_dynamics(const orc_csr *h, int64_t Z, const orc_sim *sim, double *P_S, double *P_T0, double *P_Tp,
                       double *P_Tm, int64_t n_out, int n_threads)
{
    return run_samples(h, Z, sim, (int)Z, NULL, NULL, 1, P_S, P_T0, P_Tp, P_Tm, n_out, NULL, n_threads);
}

/* observables.f90:74-101 */
void orc_get_kinetics(int64_t n, const double *P_S, const double *P_T0, const double *P_Tp, const double *P_Tm,
                      double dt, double kS, double kT, double *iden, double *Phi_S, double *Phi_T)
{
    double *P_T = (double *)malloc(sizeof(double) * (n ? n : 1));
    for (int64_t i = 0; i < n; ++i) {
        P_T[i] = P_Tp[i] + P_T0[i] + P_Tm[i];
        if (iden) iden[i] = P_S[i] + P_T[i];
    }
    int64_t h = n / 2;
    if (h > 0) {
        Phi_S[0] = 0.0;
        Phi_T[0] = 0.0;
    }
    for (int64_t i = 1; i <= h - 1; ++i) { /* 1-based i as in the reference */
        Phi_S[i] = Phi_S[i - 1] + (1.0 / 3.0) * dt * kS * (P_S[2 * i - 2] + 4 * P_S[2 * i - 1] + P_S[2 * i]);
        Phi_T[i] = Phi_T[i - 1] + (1.0 / 3.0) * dt * kT * (P_T[2 * i - 2] + 4 * P_T[2 * i - 1] + P_T[2 * i]);
    }
    free(P_T);
}

/* ======================================================== semiclassical */

/* Rodrigues rotation of v about field f for time t: v_par + cos(|f|t) v_perp +
 * sin(|f|t) (f_hat x v)  (class_radical_pair.f90:218-251, with the zero-field
 * branch the right way round). */
static void sc_rotate(double *v, const double *f, double t)
{
    double w = sqrt(f[0] * f[0] + f[1] * f[1] + f[2] * f[2]);
    if (w == 0.0) return;
    double h[3] = {f[0] / w, f[1] / w, f[2] / w};
    double dp = v[0] * h[0] + v[1] * h[1] + v[2] * h[2];
    double cs = cos(w * t), sn = sin(w * t);
    double cx = h[1] * v[2] - h[2] * v[1], cy = h[2] * v[0] - h[0] * v[2], cz = h[0] * v[1] - h[1] * v[0];
    double r0 = dp * h[0] + cs * (v[0] - dp * h[0]) + sn * cx;
    double r1 = dp * h[1] + cs * (v[1] - dp * h[1]) + sn * cy;
    double r2 = dp * h[2] + cs * (v[2] - dp * h[2]) + sn * cz;
    v[0] = r0;
    v[1] = r1;
    v[2] = r2;
}

/* 16x16 generator K + J (sc_dynamics.f90:102-185, with K(13,9) set). 0-based,
 * row-major G[16*i + j]. */
static void sc_generator(double kS, double kT, double J, double *G)
{
    double kb = (kS + 3.0 * kT) / 4.0, dk = (kS - kT) / 4.0;
    memset(G, 0, sizeof(double) * 256);
#define SG(i, j, v) G[16 * ((i)-1) + ((j)-1)] += (v)
    for (int i = 1; i <= 16; ++i) SG(i, i, -kb);
    for (int i = 1; i <= 3; ++i) {
        SG(i, i + 3, dk);
        SG(i + 3, i, dk);
    }
    SG(8, 10, dk); SG(10, 8, dk); SG(9, 13, dk); SG(13, 9, dk); SG(12, 14, dk); SG(14, 12, dk);
    SG(7, 11, -dk); SG(11, 7, -dk); SG(7, 15, -dk); SG(15, 7, -dk); SG(11, 15, -dk); SG(15, 11, -dk);
    for (int i = 7; i <= 15; i += 4) {
        SG(i, 16, 0.25 * dk);
        SG(16, i, 4.0 * dk);
    }
    double tJ = 2.0 * J, hJ = 0.5 * J;
    SG(1, 12, tJ); SG(1, 14, -tJ); SG(2, 13, tJ); SG(2, 9, -tJ); SG(3, 8, tJ); SG(3, 10, -tJ);
    SG(4, 12, -tJ); SG(4, 14, tJ); SG(5, 13, -tJ); SG(5, 9, tJ); SG(6, 8, -tJ); SG(6, 10, tJ);
    SG(8, 3, -hJ); SG(8, 6, hJ); SG(9, 2, hJ); SG(9, 5, -hJ); SG(10, 3, hJ); SG(10, 6, -hJ);
    SG(12, 1, -hJ); SG(12, 4, hJ); SG(13, 2, -hJ); SG(13, 5, hJ); SG(14, 1, hJ); SG(14, 4, -hJ);
#undef SG
}

static double sc_ps(const double *x) { return 0.25 * x[15] - (x[6] + x[10] + x[14]); }

/* Planted mistakes for the tests that show the semiclassical pins catch them
 * (0 = the model as stated; never set outside those tests). */
static int sc_mutation = 0;
void orc_sc_set_mutation(int which) { sc_mutation = which; }

/* Electron half step: x <- KJ x, then rotate S1, S2 and T12 columns/rows
 * (class_radical_pair.f90:100-118). */
static void sc_electron_half(double *x, const double *KJ, const double *w1, const double *w2, double t)
{
    double y[16];
    for (int i = 0; i < 16; ++i) {
        double s = 0.0;
        for (int j = 0; j < 16; ++j) s += KJ[16 * i + j] * x[j];
        y[i] = s;
    }
    memcpy(x, y, sizeof y);
    sc_rotate(x + 0, w1, t);
    sc_rotate(x + 3, w2, t);
    /* T(:,b) = S1 S2_b rotates with S1, T(a,:) = S1_a S2 with S2 */
    const double *wc = sc_mutation == ORC_SC_MUT_T12_TRANSPOSED ? w2 : w1;
    const double *wr = sc_mutation == ORC_SC_MUT_T12_TRANSPOSED ? w1 : w2;
    for (int col = 0; col < 3; ++col) sc_rotate(x + 6 + 3 * col, wc, t);
    for (int row = 0; row < 3; ++row) {
        double r[3] = {x[6 + row], x[9 + row], x[12 + row]};
        sc_rotate(r, wr, t);
        x[6 + row] = r[0];
        x[9 + row] = r[1];
        x[12 + row] = r[2];
    }
}

/* Shared set-up of a run: KJ = exp(dt/2 (K + J)) (sc_dynamics.f90:102-185),
 * the Zeeman field -B n and the per-sample seeds (two next() draws each). */
typedef struct {
    double KJ[256];
    double wz[3];
    int nn;
} sc_setup_t;

static int sc_setup(const orc_sc_system *s, double dt, sc_setup_t *u)
{
    double G[256];
    sc_generator(s->kS, s->kT, s->J, G);
    double Gz[512], Ez[512];
    for (int i = 0; i < 16; ++i)
        for (int j = 0; j < 16; ++j) { /* complex, column-major, scaled by dt/2 */
            Gz[2 * (i + 16 * j)] = 0.5 * dt * G[16 * i + j];
            Gz[2 * (i + 16 * j) + 1] = 0.0;
        }
    orc_expm(16, Gz, Ez);
    for (int i = 0; i < 16; ++i)
        for (int j = 0; j < 16; ++j) u->KJ[16 * i + j] = Ez[2 * (i + 16 * j)];
    const double dn = sqrt(s->dir[0] * s->dir[0] + s->dir[1] * s->dir[1] + s->dir[2] * s->dir[2]);
    if (!(dn > 0.0)) return 1;
    for (int a = 0; a < 3; ++a) u->wz[a] = -s->B * s->dir[a] / dn;
    u->nn = s->n1 + s->n2;
    return 0;
}

static void sc_master_seeds(const int64_t master_seed[2], int64_t n, int64_t *seeds)
{
    orc_rng r;
    orc_rng_init_default(&r);
    orc_rng_set_seed(&r, master_seed);
    for (int64_t i = 0; i < n; ++i) {
        seeds[2 * i] = orc_rng_next(&r);
        seeds[2 * i + 1] = orc_rng_next(&r);
    }
}

/* One sample: S1, S2 on spheres of radius sqrt(3/4), I_k of radius
 * sqrt(I(I+1)) (m_random.f90:300-318), T12 = S1 S2^T, x_16 = 1, then n_steps
 * Strang steps: electron half step, nuclear full step about
 * a_k S_e sqrt(3/4)/|S_e| (class_radical_pair.f90:123-145), electron half
 * step.  out[st] = P_S(0) P_S(x(st)), out_t[st] = P_S(0) x_16(st); with xs/Is
 * non-NULL the state after every step ([n_steps+1][16], [n_steps+1][3 nn]). */
static void sc_sample(const orc_sc_system *s, const sc_setup_t *u, const int64_t seed[2], double dt, int64_t n_steps,
                      double *out, double *out_t, double *xs, double *Is)
{
    const int nn = u->nn;
    const double shalf = sqrt(0.75);
    double x[16];
    double *Iv = (double *)malloc(sizeof(double) * 3 * (nn ? nn : 1));
    orc_rng r;
    orc_rng_init_default(&r);
    orc_rng_set_seed(&r, seed);
    orc_rng_sphere(&r, shalf, x + 0);
    orc_rng_sphere(&r, shalf, x + 3);
    for (int k = 0; k < nn; ++k) {
        double I = 0.5 * (s->mult[k] - 1);
        orc_rng_sphere(&r, sqrt(I * (I + 1.0)), Iv + 3 * k);
    }
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) x[6 + a + 3 * b] = x[a] * x[3 + b];
    x[15] = 1.0;
    const double p0 = sc_ps(x);
    const double nuc_sign = sc_mutation == ORC_SC_MUT_NUCLEAR_SIGN ? -1.0 : 1.0;
    const double nuc_scale = sc_mutation == ORC_SC_MUT_NUCLEAR_SCALE ? 2.0 : 1.0;
    const double el_scale = sc_mutation == ORC_SC_MUT_FIELD_SCALE ? 0.5 : 1.0;
    out[0] = p0 * p0;
    out_t[0] = p0 * x[15];
    if (xs) memcpy(xs, x, sizeof x);
    if (Is) memcpy(Is, Iv, sizeof(double) * 3 * nn);
    for (int64_t st = 1; st <= n_steps; ++st) {
        for (int half = 0; half < 2; ++half) {
            double w1[3] = {u->wz[0], u->wz[1], u->wz[2]}, w2[3] = {u->wz[0], u->wz[1], u->wz[2]};
            for (int k = 0; k < nn; ++k) {
                double *w = k < s->n1 ? w1 : w2;
                for (int a = 0; a < 3; ++a) w[a] += el_scale * s->a_iso[k] * Iv[3 * k + a];
            }
            sc_electron_half(x, u->KJ, w1, w2, 0.5 * dt);
            if (half == 0) {
                double n1 = sqrt(x[0] * x[0] + x[1] * x[1] + x[2] * x[2]);
                double n2 = sqrt(x[3] * x[3] + x[4] * x[4] + x[5] * x[5]);
                double f1 = n1 > 0.0 ? shalf / n1 : 0.0, f2 = n2 > 0.0 ? shalf / n2 : 0.0;
                for (int k = 0; k < nn; ++k) {
                    const double *S = k < s->n1 ? x : x + 3;
                    double f = nuc_sign * nuc_scale * (k < s->n1 ? f1 : f2) * s->a_iso[k];
                    double fld[3] = {f * S[0], f * S[1], f * S[2]};
                    sc_rotate(Iv + 3 * k, fld, dt);
                }
            }
        }
        out[st] = p0 * sc_ps(x);
        out_t[st] = p0 * x[15];
        if (xs) memcpy(xs + 16 * st, x, sizeof x);
        if (Is) memcpy(Is + 3 * nn * st, Iv, sizeof(double) * 3 * nn);
    }
    free(Iv);
}

int orc_semiclassical(const orc_sc_system *s, int64_t n_samples, const int64_t master_seed[2], double dt,
                      int64_t n_steps, double *P_S, double *P_T, int n_threads)
{
    if (n_samples < 1 || n_steps < 0) return 1;
    sc_setup_t u;
    if (sc_setup(s, dt, &u)) return 1;
    int64_t *seeds = (int64_t *)malloc(sizeof(int64_t) * 2 * n_samples);
    sc_master_seeds(master_seed, n_samples, seeds);
    double *contrib = (double *)malloc(sizeof(double) * (n_steps + 1) * (size_t)n_samples);
    double *contrib_t = (double *)malloc(sizeof(double) * (n_steps + 1) * (size_t)n_samples);
#ifdef _OPENMP
    int nt = n_threads > 0 ? n_threads : omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 64) num_threads(nt)
#endif
    for (int64_t smp = 0; smp < n_samples; ++smp)
        sc_sample(s, &u, seeds + 2 * smp, dt, n_steps, contrib + (n_steps + 1) * (size_t)smp,
                  contrib_t + (n_steps + 1) * (size_t)smp, NULL, NULL);
    /* P_T = x_16 - P_S per sample (class_radical_pair.f90:174-175) */
    for (int64_t st = 0; st <= n_steps; ++st) {
        double acc = 0.0, acc_t = 0.0;
        for (int64_t smp = 0; smp < n_samples; ++smp) {
            acc += contrib[(n_steps + 1) * (size_t)smp + st];
            acc_t += contrib_t[(n_steps + 1) * (size_t)smp + st];
        }
        P_S[st] = 4.0 * acc / (double)n_samples;
        if (P_T) P_T[st] = 4.0 * acc_t / (double)n_samples - P_S[st];
    }
    free(contrib);
    free(contrib_t);
    free(seeds);
    return 0;
}

int orc_semiclassical_trajectory(const orc_sc_system *s, const int64_t master_seed[2], int64_t sample, double dt,
                                 int64_t n_steps, double *xs, double *Is)
{
    if (sample < 0 || n_steps < 0 || !xs) return 1;
    sc_setup_t u;
    if (sc_setup(s, dt, &u)) return 1;
    int64_t *seeds = (int64_t *)malloc(sizeof(int64_t) * 2 * (sample + 1));
    sc_master_seeds(master_seed, sample + 1, seeds);
    double *o = (double *)malloc(sizeof(double) * 2 * (n_steps + 1));
    sc_sample(s, &u, seeds + 2 * sample, dt, n_steps, o, o + n_steps + 1, xs, Is);
    free(o);
    free(seeds);
    return 0;
}
