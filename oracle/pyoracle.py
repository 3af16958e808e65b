"""ctypes wrapper over oracle/liboracle.so.

TEST INFRASTRUCTURE ONLY: the CPU oracle (a plain-C restatement of the
reference algorithm) is used by tests/, by the smoke check and by the CPU
baseline leg of bench.py, always as the checker, never as the product path.
"""
from __future__ import annotations

import ctypes
import math
import os
from dataclasses import dataclass, field
from typing import Optional, Sequence

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_REF_PATH = os.path.join(_HERE, "_ref", "libref_spinchem.so")

GAMMA_E = 176.08596
GAMMA_1H = 0.26752218708
GAMMA_14N = 0.0193297792

_c_int64_p = ctypes.POINTER(ctypes.c_int64)
_c_double_p = ctypes.POINTER(ctypes.c_double)
_c_int32_p = ctypes.POINTER(ctypes.c_int32)
_c_int_p = ctypes.POINTER(ctypes.c_int)


class OrcRng(ctypes.Structure):
    _fields_ = [("s", ctypes.c_uint64 * 2), ("stored_normal", ctypes.c_double),
                ("have_stored_normal", ctypes.c_int)]


class OrcSystem(ctypes.Structure):
    _fields_ = [("n1", ctypes.c_int), ("n2", ctypes.c_int), ("mult", _c_int_p),
                ("hfc", _c_double_p), ("J", ctypes.c_double), ("D", ctypes.c_double * 9),
                ("kS", ctypes.c_double), ("kT", ctypes.c_double), ("B", ctypes.c_double),
                ("dir", ctypes.c_double * 3), ("nuclear_1h", ctypes.c_int)]


class OrcCsr(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int), ("nnz", ctypes.c_int64), ("rowptr", _c_int32_p),
                ("col", _c_int32_p), ("val", _c_double_p)]


class OrcSim(ctypes.Structure):
    _fields_ = [("n_samples", ctypes.c_int), ("n_krylov", ctypes.c_int), ("t_end", ctypes.c_double),
                ("dt", ctypes.c_double), ("tol", ctypes.c_double), ("init_state", ctypes.c_int),
                ("vec_kind", ctypes.c_int)]


class OrcScSystem(ctypes.Structure):
    _fields_ = [("n1", ctypes.c_int), ("n2", ctypes.c_int), ("a_iso", _c_double_p), ("mult", _c_int_p),
                ("B", ctypes.c_double), ("dir", ctypes.c_double * 3), ("J", ctypes.c_double),
                ("kS", ctypes.c_double), ("kT", ctypes.c_double)]


def _load():
    if not os.path.exists(_LIB_PATH):
        raise FileNotFoundError(f"{_LIB_PATH} not built (run `make -C oracle`)")
    lib = ctypes.CDLL(_LIB_PATH)
    lib.orc_rng_next.restype = ctypes.c_int64
    lib.orc_rng_unif01.restype = ctypes.c_double
    lib.orc_rng_normal.restype = ctypes.c_double
    lib.orc_nuclear_dim.restype = ctypes.c_int64
    lib.orc_n_steps.restype = ctypes.c_int64
    lib.orc_n_steps.argtypes = [ctypes.c_double, ctypes.c_double]
    lib.orc_steps_taken.restype = ctypes.c_int64
    lib.orc_steps_taken.argtypes = [ctypes.c_double, ctypes.c_double]
    return lib


_lib = None


def lib():
    global _lib
    if _lib is None:
        _lib = _load()
    return _lib


def ref_lib():
    """The compiled reference modules (m_random.f90, observables.f90), or None."""
    if not os.path.exists(_REF_PATH):
        return None
    return ctypes.CDLL(_REF_PATH)


# ----------------------------------------------------------------- helpers
def _dp(a: np.ndarray):
    return a.ctypes.data_as(_c_double_p)


@dataclass
class System:
    """Flattened sys_param: multiplicities, 3x3 tensors (mT), J, D, kS/kT in mT
    units (rate / gamma_e), field magnitude B (mT) and direction."""
    mult1: Sequence[int] = ()
    mult2: Sequence[int] = ()
    A1: Optional[np.ndarray] = None  # (n1, 3, 3), A[k][a][b]
    A2: Optional[np.ndarray] = None
    J: float = 0.0
    D: np.ndarray = field(default_factory=lambda: np.zeros((3, 3)))
    kS: float = 0.0
    kT: float = 0.0
    B: float = 0.0
    direction: Sequence[float] = (0.0, 0.0, 1.0)
    nuclear_1h: bool = False  # 1H's gyromagnetic ratio for every nucleus (symmetry blocks)

    def tensors(self) -> np.ndarray:
        n1, n2 = len(self.mult1), len(self.mult2)
        A1 = np.zeros((0, 3, 3)) if n1 == 0 else np.asarray(self.A1, dtype=float).reshape(n1, 3, 3)
        A2 = np.zeros((0, 3, 3)) if n2 == 0 else np.asarray(self.A2, dtype=float).reshape(n2, 3, 3)
        return np.concatenate([A1, A2], axis=0)

    @property
    def Z(self) -> int:
        return int(np.prod(list(self.mult1) + list(self.mult2), dtype=np.int64))

    def to_c(self):
        mult = np.ascontiguousarray(list(self.mult1) + list(self.mult2), dtype=np.intc)
        A = self.tensors()
        # column-major per nucleus: hfc[9k + 3b + a] = A[k][a][b]
        hfc = np.ascontiguousarray(np.transpose(A, (0, 2, 1)).reshape(-1), dtype=float)
        s = OrcSystem()
        s.n1, s.n2 = len(self.mult1), len(self.mult2)
        s.mult = mult.ctypes.data_as(_c_int_p)
        s.hfc = hfc.ctypes.data_as(_c_double_p)
        s.J = self.J
        Dc = np.asarray(self.D, dtype=float).reshape(3, 3).T.reshape(-1)
        for i in range(9):
            s.D[i] = Dc[i]
        s.kS, s.kT, s.B = self.kS, self.kT, self.B
        for i in range(3):
            s.dir[i] = float(self.direction[i])
        s.nuclear_1h = int(bool(self.nuclear_1h))
        s._keep = (mult, hfc)
        return s


def _csr_to_scipy(h: OrcCsr):
    import scipy.sparse as sp
    n, nnz = h.n, h.nnz
    rowptr = np.ctypeslib.as_array(h.rowptr, shape=(n + 1,)).copy()
    col = np.ctypeslib.as_array(h.col, shape=(max(nnz, 1),))[:nnz].copy()
    val = np.ctypeslib.as_array(h.val, shape=(max(2 * nnz, 2),))[:2 * nnz].copy().view(np.complex128)
    return sp.csr_matrix((val, col, rowptr), shape=(n, n))


class Csr:
    """Owns an orc_csr; .scipy() gives a scipy.sparse copy."""

    def __init__(self, h: OrcCsr):
        self.h = h

    def __del__(self):
        try:
            lib().orc_csr_free(ctypes.byref(self.h))
        except Exception:
            pass

    def scipy(self):
        return _csr_to_scipy(self.h)

    def arrays(self):
        """(rowptr, col, val) as stored, 0-based, val complex."""
        n, nnz = self.h.n, self.h.nnz
        rowptr = np.ctypeslib.as_array(self.h.rowptr, shape=(n + 1,)).copy()
        col = np.ctypeslib.as_array(self.h.col, shape=(max(nnz, 1),))[:nnz].copy()
        val = np.ctypeslib.as_array(self.h.val, shape=(max(2 * nnz, 2),))[:2 * nnz].copy().view(np.complex128)
        return rowptr, col, val

    @property
    def n(self):
        return self.h.n

    @property
    def nnz(self):
        return self.h.nnz


def build_iso_csr(sys: System) -> Csr:
    s = sys.to_c()
    h = OrcCsr()
    rc = lib().orc_build_iso_csr(ctypes.byref(s), ctypes.byref(h))
    if rc:
        raise ValueError(f"orc_build_iso_csr failed with code {rc}")
    return Csr(h)


def build_general_csr(sys: System) -> Csr:
    s = sys.to_c()
    h = OrcCsr()
    rc = lib().orc_build_general_csr(ctypes.byref(s), ctypes.byref(h))
    if rc:
        raise ValueError(f"orc_build_general_csr failed with code {rc}")
    return Csr(h)


def expm(A: np.ndarray) -> np.ndarray:
    A = np.asfortranarray(A, dtype=np.complex128)
    m = A.shape[0]
    E = np.zeros((m, m), dtype=np.complex128, order="F")
    lib().orc_expm(ctypes.c_int(m), A.ctypes.data_as(_c_double_p), E.ctypes.data_as(_c_double_p))
    return E


def get_krylov(h: Csr, psi: np.ndarray, m: int):
    psi = np.ascontiguousarray(psi, dtype=np.complex128)
    N = h.n
    Q = np.zeros((N, m), dtype=np.complex128, order="F")
    c = np.zeros(m, dtype=np.complex128)
    A = np.zeros((m, m), dtype=np.complex128, order="F")
    lib().orc_get_krylov(ctypes.byref(h.h), psi.ctypes.data_as(_c_double_p), ctypes.c_int(m),
                         Q.ctypes.data_as(_c_double_p), c.ctypes.data_as(_c_double_p),
                         A.ctypes.data_as(_c_double_p))
    return Q, c, A


def n_steps(t_end: float, dt: float) -> int:
    return int(lib().orc_n_steps(t_end, dt))


def steps_taken(t_end: float, dt: float) -> int:
    return int(lib().orc_steps_taken(t_end, dt))


def sample_seeds(master, M: int) -> np.ndarray:
    seeds = np.zeros(2 * M, dtype=np.int64)
    ms = (ctypes.c_int64 * 2)(*master)
    lib().orc_sample_seeds(ms, ctypes.c_int(M), seeds.ctypes.data_as(_c_int64_p))
    return seeds.reshape(M, 2)


def sample_vector(seed2, Z: int, kind: int = 0) -> np.ndarray:
    r = OrcRng()
    lib().orc_rng_init_default(ctypes.byref(r))
    lib().orc_rng_set_seed(ctypes.byref(r), (ctypes.c_int64 * 2)(*[int(x) for x in seed2]))
    z = np.zeros(Z, dtype=np.complex128)
    fn = lib().orc_sample_phase if kind == 1 else lib().orc_sample_SUZ
    fn(ctypes.byref(r), ctypes.c_int64(Z), z.ctypes.data_as(_c_double_p))
    return z


def trace_sampling(h: Csr, Z: int, *, M: int, m: int, t_end: float, dt: float, tol: float = 0.0,
                   init_state: int = 0, vec_kind: int = 0, seed=(42, 99), injected=None,
                   per_sample: bool = False, n_threads: int = 0):
    """Returns dict with P_S, P_T0, P_Tp, P_Tm (length n_steps+1)."""
    n_out = n_steps(t_end, dt) + 1
    sim = OrcSim(M, m, t_end, dt, tol, init_state, vec_kind)
    outs = [np.zeros(n_out) for _ in range(4)]
    ps = np.zeros((M, n_out)) if per_sample else None
    inj = None
    if vec_kind == 2:
        inj = np.ascontiguousarray(injected, dtype=np.complex128).reshape(M, Z)
    rc = lib().orc_trace_sampling(
        ctypes.byref(h.h), ctypes.c_int64(Z), ctypes.byref(sim), (ctypes.c_int64 * 2)(*seed),
        inj.ctypes.data_as(_c_double_p) if inj is not None else None,
        *[_dp(o) for o in outs], ctypes.c_int64(n_out), _dp(ps) if ps is not None else None,
        ctypes.c_int(n_threads))
    if rc:
        raise RuntimeError(f"orc_trace_sampling failed with code {rc}")
    res = dict(zip(["P_S", "P_T0", "P_Tp", "P_Tm"], outs))
    if per_sample:
        res["per_sample_PS"] = ps
    return res


def exact_dynamics(h: Csr, Z: int, *, m: int, t_end: float, dt: float, tol: float = 0.0,
                   init_state: int = 0, n_threads: int = 0):
    n_out = n_steps(t_end, dt) + 1
    sim = OrcSim(Z, m, t_end, dt, tol, init_state, 0)
    outs = [np.zeros(n_out) for _ in range(4)]
    rc = lib().orc_exact_dynamics(ctypes.byref(h.h), ctypes.c_int64(Z), ctypes.byref(sim),
                                  *[_dp(o) for o in outs], ctypes.c_int64(n_out), ctypes.c_int(n_threads))
    if rc:
        raise RuntimeError(f"orc_exact_dynamics failed with code {rc}")
    return dict(zip(["P_S", "P_T0", "P_Tp", "P_Tm"], outs))


def get_kinetics(P_S, P_T0, P_Tp, P_Tm, dt: float, kS: float, kT: float):
    n = len(P_S)
    arrs = [np.ascontiguousarray(a, dtype=float) for a in (P_S, P_T0, P_Tp, P_Tm)]
    iden = np.zeros(n)
    phis = np.zeros(max(n // 2, 1))
    phit = np.zeros(max(n // 2, 1))
    lib().orc_get_kinetics(ctypes.c_int64(n), *[_dp(a) for a in arrs], ctypes.c_double(dt),
                           ctypes.c_double(kS), ctypes.c_double(kT), _dp(iden), _dp(phis), _dp(phit))
    return iden, phis[:n // 2], phit[:n // 2]


def _sc_system(a_iso1, a_iso2, mult1, mult2, B, direction, J, kS, kT):
    a = np.ascontiguousarray(list(a_iso1) + list(a_iso2), dtype=float)
    mult = np.ascontiguousarray(list(mult1) + list(mult2), dtype=np.intc)
    s = OrcScSystem()
    s.n1, s.n2 = len(a_iso1), len(a_iso2)
    s.a_iso = a.ctypes.data_as(_c_double_p)
    s.mult = mult.ctypes.data_as(_c_int_p)
    s.B, s.J, s.kS, s.kT = B, J, kS, kT
    for i in range(3):
        s.dir[i] = direction[i]
    return s, (a, mult)


def semiclassical_trajectory(a_iso1, a_iso2, mult1, mult2, *, B: float, direction=(0, 0, 1), J: float = 0.0,
                             kS: float = 0.0, kT: float = 0.0, dt: float, n_steps_: int, sample: int = 0,
                             seed=(42, 99)):
    """States of one sample of the semiclassical run after every step:
    x [n_steps_+1][16] (S1, S2, T12 column-major, x_16) and the nuclear
    vectors I [n_steps_+1][n1+n2][3]."""
    s, keep = _sc_system(a_iso1, a_iso2, mult1, mult2, B, direction, J, kS, kT)
    nn = len(a_iso1) + len(a_iso2)
    xs = np.zeros((n_steps_ + 1, 16))
    Is = np.zeros((n_steps_ + 1, max(nn, 1), 3))
    rc = lib().orc_semiclassical_trajectory(ctypes.byref(s), (ctypes.c_int64 * 2)(*seed), ctypes.c_int64(sample),
                                            ctypes.c_double(dt), ctypes.c_int64(n_steps_), _dp(xs), _dp(Is))
    if rc:
        raise RuntimeError(f"orc_semiclassical_trajectory failed with code {rc}")
    del keep
    return xs, Is[:, :nn]


SC_MUTATIONS = {"none": 0, "t12_transposed": 1, "nuclear_sign": 2, "nuclear_scale": 3, "field_scale": 4}


class sc_mutation:
    """Context manager planting one of the oracle's semiclassical mistakes
    (tests that show a pin catches it)."""

    def __init__(self, name: str):
        self.which = SC_MUTATIONS[name]

    def __enter__(self):
        lib().orc_sc_set_mutation(ctypes.c_int(self.which))
        return self

    def __exit__(self, *exc):
        lib().orc_sc_set_mutation(ctypes.c_int(0))
        return False


def semiclassical(a_iso1, a_iso2, mult1, mult2, *, B: float, direction=(0, 0, 1), J: float = 0.0,
                  kS: float = 0.0, kT: float = 0.0, n_samples: int, dt: float, n_steps_: int,
                  seed=(42, 99), n_threads: int = 0, with_pt: bool = False):
    """P_S(t) (and P_T(t) with with_pt) of the semiclassical model, n_steps_ + 1 values."""
    s, keep = _sc_system(a_iso1, a_iso2, mult1, mult2, B, direction, J, kS, kT)
    out = np.zeros(n_steps_ + 1)
    out_t = np.zeros(n_steps_ + 1)
    rc = lib().orc_semiclassical(ctypes.byref(s), ctypes.c_int64(n_samples), (ctypes.c_int64 * 2)(*seed),
                                 ctypes.c_double(dt), ctypes.c_int64(n_steps_), _dp(out), _dp(out_t),
                                 ctypes.c_int(n_threads))
    if rc:
        raise RuntimeError(f"orc_semiclassical failed with code {rc}")
    return (out, out_t) if with_pt else out


# ------------------------------------------------------- symmetry reduction
# symmetry.f90:16-246 and moments.f90:15-240, restated with numpy/scipy for
# the tridiagonal eigenproblem (stdlib steqr) and the linear solve (solve_lu).

def quad_rule(w_init, x_init, n):
    """moments.f90:95-150: discrete Stieltjes procedure on diag(x) from
    sqrt(w), then the n-point Gauss rule (nodes ascending, weights)."""
    from scipy.linalg import eigh_tridiagonal
    x_init = np.asarray(x_init, float)
    q = np.sqrt(np.asarray(w_init, float))
    P = np.zeros((len(x_init), n))
    W = np.zeros(n)
    X = np.zeros(n)
    for k in range(n):
        q2 = float(np.linalg.norm(q))
        if q2 == 0.0:
            raise RuntimeError("quad_rule 1")
        P[:, k] = q / q2
        q = x_init * P[:, k]
        for j in range(k, -1, -1):
            pq = float(P[:, j] @ q)
            if j == k:
                W[k], X[k] = q2, pq
            q = q - pq * P[:, j]
    weight = W[0]
    E = np.roll(W, -1)  # cshift(W, 1)
    lam, Zv = eigh_tridiagonal(X, E[:n - 1])
    return (weight * Zv[0, :]) ** 2, lam


def newton(aw_bar, w_bar, a_bar, n_bar, m):
    """moments.f90:184-239; returns (a_bar, icode)."""
    a = np.array(a_bar, float)
    n_bar = np.asarray(n_bar, float)
    dap = 0.0
    for it in range(1, 31):
        f = np.array([n_bar @ a ** k - w_bar @ aw_bar ** k for k in range(1, m + 1)])
        G = np.array([k * n_bar * a ** (k - 1) for k in range(1, m + 1)])
        try:
            q = np.linalg.solve(G, f)
        except np.linalg.LinAlgError:  # a set of size 0: stdlib's solve_lu stops; the candidate is dropped
            return a, 1
        a_abs, d_abs = np.sum(np.abs(a)), np.sum(np.abs(q))
        prev = a.copy()
        a = a - q
        da = d_abs / a_abs
        if da > 1.0:
            return a, 1
        if 1.0 + da == 1.0:
            return a, 0
        if it > 1 and da >= dap:
            return prev, 0
        dap = da
    return a, 2


def shrink(a_iso, M):
    """moments.f90:15-93: (a_bar, n_bar) of M sets of equivalent nuclei."""
    a_iso = np.asarray(a_iso, float)
    if a_iso.size == 0:
        return np.zeros(0), np.zeros(0, int)
    if M >= a_iso.size:
        raise ValueError("Error: Symetry reduction requires M<N")
    w_bar, aw_bar = quad_rule(np.ones(a_iso.size), a_iso, M)
    nfloor = w_bar.astype(int)
    ndiff = a_iso.size - int(nfloor.sum())
    best, dmom = None, np.inf
    for ib in range(1 << M):
        if bin(ib).count("1") != ndiff:
            continue
        nt = nfloor + np.array([(ib >> i) & 1 for i in range(M)])
        at, icode = newton(aw_bar, w_bar, aw_bar.copy(), nt, M)
        if icode != 0:
            continue
        emom = sum(abs(np.sum(at ** k) - np.sum(a_iso ** k)) for k in range(1, M + 2))
        if emom < dmom:
            dmom, best = emom, (at, nt)
    if best is None:
        raise RuntimeError("shrink: no integer weights converged")
    return best


def symmetry_blocks(n_bar):
    """cartesian_product + sort_col_prod + weight (symmetry.f90:61-65,
    187-226, utils.f90:25-50): (K [blocks][sets], Z [blocks], w [blocks])."""
    n_bar = [int(x) for x in n_bar]
    combos = [[]]
    for n in n_bar:  # the first set varies fastest
        opts = [2 * j + 1 + n % 2 for j in range(n // 2 + 1)]
        combos = [c + [g] for g in opts for c in combos]
    Z = [int(np.prod(c, dtype=np.int64)) for c in combos]
    order = sorted(range(len(combos)), key=lambda i: Z[i])  # Python's sort is stable
    K = [combos[i] for i in order]
    Zs = [Z[i] for i in order]
    W = []
    for k in K:
        w = 1
        for n, g in zip(n_bar, k):
            w *= math.comb(n, (n + g - 1) // 2) * (2 * g) // (n + g + 1)
        W.append(w)
    return K, Zs, W


def symmetrised_dynamics(a1, a2, M1, M2, *, J, kS, kT, B, m, t_end, dt, N_samples, master_seed, block_tol=0.0,
                         tol=0.0, vec_kind=0, init_state=0, D=None):
    """symmetry.f90:16-139 with the n_bar copy and multiplicities above 3 as
    the method intends: per block, trace sampling of the reduced system with
    its own master seed (two next() draws per block from master_seed), each
    scaled by w Z_block / Z; returns the summed traces and the block list."""
    a1_bar, n1_bar = shrink(a1, M1)
    a2_bar, n2_bar = shrink(a2, M2)
    m1 = len(a1_bar)
    n_bar = list(n1_bar) + list(n2_bar)
    K, Zb, W = symmetry_blocks(n_bar)
    Zfull = 2 ** (len(a1) + len(a2))
    r = lib()
    rng = OrcRng()
    r.orc_rng_init_default(ctypes.byref(rng))
    r.orc_rng_set_seed(ctypes.byref(rng), (ctypes.c_int64 * 2)(*master_seed))
    seeds = [(r.orc_rng_next(ctypes.byref(rng)), r.orc_rng_next(ctypes.byref(rng))) for _ in K]
    n_out = n_steps(t_end, dt) + 1
    tot = {k: np.zeros(n_out) for k in ("P_S", "P_T0", "P_Tp", "P_Tm")}
    blocks = []
    for i, (k, z, w) in enumerate(zip(K, Zb, W)):
        wk = w * z / Zfull
        blocks.append((k, z, w, wk))
        if wk < block_tol:
            continue
        g1 = [g for g in k[:m1]]
        g2 = [g for g in k[m1:]]
        keep1 = [j for j, g in enumerate(g1) if g > 1]
        keep2 = [j for j, g in enumerate(g2) if g > 1]
        s = System(mult1=[g1[j] for j in keep1], mult2=[g2[j] for j in keep2],
                   A1=np.array([a1_bar[j] * np.eye(3) for j in keep1]) if keep1 else None,
                   A2=np.array([a2_bar[j] * np.eye(3) for j in keep2]) if keep2 else None,
                   J=J, D=np.zeros((3, 3)) if D is None else D, kS=kS, kT=kT, B=B, nuclear_1h=True)
        h = build_general_csr(s)
        res = trace_sampling(h, s.Z, M=N_samples, m=m, t_end=t_end, dt=dt, tol=tol, init_state=init_state,
                             vec_kind=vec_kind, seed=seeds[i])
        for key in tot:
            tot[key] += res[key] * wk
    return tot, blocks, (a1_bar, n1_bar, a2_bar, n2_bar)
