"""Module `dynamics` of the reference (qm_dynamics.f90) on the gfx950 path.

trace_sampling_para, exact_dynamics and expm keep the reference's names and
argument meaning; the work is done by libspinchem_hip.so through its C-ABI.
"""
from __future__ import annotations

from typing import Optional

import numpy as np

from . import _lib
from .observables import Observables
from .spinsys import SimParam, SysParam, to_c_system

_ctx: Optional[_lib.Context] = None


def context(device: int = 0) -> _lib.Context:
    """Process-wide default context on `device`."""
    global _ctx
    if _ctx is None:
        _ctx = _lib.Context(device)
    return _ctx


class Hamiltonian:
    """The device CSR of one spin system at field B (getHiso_sp + coo_to_csr)."""

    def __init__(self, sys: SysParam, B: float, ctx: Optional[_lib.Context] = None, block: bool = False):
        """block: sys is a symmetry block (spc_system_set_block: multiplets of
        equivalent spin-1/2 nuclei of any multiplicity, 1H nuclear Zeeman)."""
        self.ctx = ctx or context()
        s, self._keep = to_c_system(sys, B)
        setter = self.ctx.lib.spc_system_set_block if block else self.ctx.lib.spc_system_set
        self.ctx.check(setter(self.ctx.h, s))
        N, Z, nnz = (np.zeros(1, np.int64) for _ in range(3))
        self.ctx.check(self.ctx.lib.spc_system_info(self.ctx.h, _lib.ptr(N, _lib._i64p), _lib.ptr(Z, _lib._i64p),
                                                    _lib.ptr(nnz, _lib._i64p)))
        self.N, self.Z, self.nnz = int(N[0]), int(Z[0]), int(nnz[0])

    def csr(self):
        """(rowptr, col, val) copied from the device."""
        rowptr = np.zeros(self.N + 1, np.int32)
        col = np.zeros(max(self.nnz, 1), np.int32)
        val = np.zeros(max(self.nnz, 1), np.complex128)
        self.ctx.check(self.ctx.lib.spc_hamiltonian_copy(self.ctx.h, _lib.ptr(rowptr, _lib._i32p),
                                                         _lib.ptr(col, _lib._i32p), _lib.ptr(val)))
        return rowptr, col[:self.nnz], val[:self.nnz]

    def scipy(self):
        import scipy.sparse as sp
        rowptr, col, val = self.csr()
        return sp.csr_matrix((val, col, rowptr), shape=(self.N, self.N))

    def spmm(self, X: np.ndarray) -> np.ndarray:
        """H x for each row of X ([M][N] complex)."""
        X = np.ascontiguousarray(np.atleast_2d(X), dtype=np.complex128)
        Y = np.zeros_like(X)
        self.ctx.check(self.ctx.lib.spc_spmm(self.ctx.h, X.shape[0], _lib.ptr(X), _lib.ptr(Y)))
        return Y

    def krylov(self, psi: np.ndarray, m: int):
        """get_krylov for each row of psi: returns Q [M][m][N], A [M][m][m] (Fortran order), beta [M]."""
        psi = np.ascontiguousarray(np.atleast_2d(psi), dtype=np.complex128)
        M = psi.shape[0]
        Q = np.zeros((M, m, self.N), np.complex128)
        A = np.zeros((M, m, m), np.complex128)  # [r][col][row]
        beta = np.zeros(M)
        self.ctx.check(self.ctx.lib.spc_krylov_basis(self.ctx.h, M, m, _lib.ptr(psi), _lib.ptr(Q), _lib.ptr(A),
                                                     _lib.ptr(beta)))
        return Q, np.transpose(A, (0, 2, 1)), beta


def _sim_c(sim: SimParam, n_samples: int, vec_kind: int) -> _lib.SimParams:
    if sim.init_state not in _lib.STATE:
        raise ValueError("Error: Invalid initial radical pair state entered.")
    return _lib.SimParams(int(n_samples), int(sim.N_krylov), float(sim.t_end), float(sim.dt), float(sim.tol),
                          _lib.STATE[sim.init_state], int(vec_kind))


def n_steps(t_end: float, dt: float) -> int:
    return int(_lib.load().spc_n_steps(t_end, dt))


def _run(fn, ctx, sim, n_out, *args):
    outs = [np.zeros(n_out) for _ in range(4)]
    phis, phit = np.zeros(max(n_out // 2, 1)), np.zeros(max(n_out // 2, 1))
    ctx.check(fn(ctx.h, *args, *[_lib.ptr(o) for o in outs], n_out, _lib.ptr(phis), _lib.ptr(phit)))
    obs = Observables(*outs)
    obs.Phi_S, obs.Phi_T = phis[:n_out // 2], phit[:n_out // 2]
    obs.iden = obs.P_S + obs.P_T0 + obs.P_Tp + obs.P_Tm
    return obs


def trace_sampling_para(sys: SysParam, sim: SimParam, seed=None, out: Optional[str] = None, *,
                        B: Optional[float] = None, trace_vectors: Optional[np.ndarray] = None,
                        ctx: Optional[_lib.Context] = None, block: bool = False) -> Observables:
    """qm_dynamics.f90:257-384.  The field is B (default -sys.e1.w, as main.f90
    sets w = -B); seed is the master seed set before the call (main.f90:50).
    trace_vectors ([M][Z] nuclear vectors) replaces the sampled ones; block:
    sys is a symmetry block (see Hamiltonian)."""
    ctx = ctx or context()
    B = -sys.e1.w if B is None else B
    Hamiltonian(sys, B, ctx, block=block)
    seed = sim.seed if seed is None else seed
    if trace_vectors is not None:
        kind = _lib.VEC_INJECTED
        tv = np.ascontiguousarray(trace_vectors, dtype=np.complex128)
        M = tv.shape[0]
    else:
        kind = _lib.VEC_PHASE if sim.vec_kind == "phase" else _lib.VEC_SUZ
        tv, M = None, sim.N_samples
    n_out = n_steps(sim.t_end, sim.dt) + 1
    sd = np.ascontiguousarray(seed, dtype=np.int64)
    res = _run(ctx.lib.spc_trace_sample, ctx, sim, n_out, _sim_c(sim, M, kind), _lib.ptr(sd, _lib._i64p),
               _lib.ptr(tv))
    if out is not None:
        res.get_kinetics(sim.dt, sys.kS, sys.kT)
        res.output(out)
    return res


def exact_dynamics(sys: SysParam, sim: SimParam, out: Optional[str] = None, *, B: Optional[float] = None,
                   ctx: Optional[_lib.Context] = None, block: bool = False) -> Observables:
    """qm_dynamics.f90:24-137: the Z nuclear basis states as trace vectors
    (block: sys is a symmetry block, see Hamiltonian)."""
    ctx = ctx or context()
    B = -sys.e1.w if B is None else B
    Hamiltonian(sys, B, ctx, block=block)
    n_out = n_steps(sim.t_end, sim.dt) + 1
    res = _run(ctx.lib.spc_exact_trace, ctx, sim, n_out, _sim_c(sim, 0, 0))
    if out is not None:
        res.get_kinetics(sim.dt, sys.kS, sys.kT)
        res.output(out)
    return res


def expm(A: np.ndarray, ctx: Optional[_lib.Context] = None) -> np.ndarray:
    """matrix_exponential.f90:14-87 on the device (one m x m matrix)."""
    ctx = ctx or context()
    A = np.asfortranarray(A, dtype=np.complex128)
    m = A.shape[0]
    E = np.zeros((m, m), np.complex128, order="F")
    ctx.check(ctx.lib.spc_expm_small(ctx.h, m, _lib.ptr(A), _lib.ptr(E)))
    return E


def last_run_stats(ctx: Optional[_lib.Context] = None):
    ctx = ctx or context()
    ms = np.zeros(1)
    rb = np.zeros(1, np.int64)
    st = np.zeros(1, np.int64)
    ctx.check(ctx.lib.spc_last_run_stats(ctx.h, _lib.ptr(ms), _lib.ptr(rb, _lib._i64p), _lib.ptr(st, _lib._i64p)))
    return float(ms[0]), int(rb[0]), int(st[0])
