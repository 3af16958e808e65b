"""Module `symmetry` of the reference (symmetry.f90) on the gfx950 path:
symmetrised_dynamics, the equivalent-nuclei reduction of trace sampling.

The spin-1/2 nuclei of each electron are approximated by M1 (M2) sets of
equivalent nuclei (shrink, moments.f90:15-93); every combination of one
total-spin multiplet per set is a block whose reduced Hamiltonian has one
"nucleus" per set, of the multiplet's multiplicity.  Each block is trace-
sampled on the device with its own master seed and weighted by the share of
the full nuclear space it stands for; the weighted sum gives the traces and
yields (symmetry.f90:16-139).  As the method intends (and unlike the
reference's code), electron 2's set counts are copied into the block list
(symmetry.f90:59 assigns an empty section) and multiplets of any multiplicity
are propagated (getHiso_sp stops above 3, hamiltonian.f90:137); the multiplets
carry 1H's nuclear Zeeman term (getHiso_sp would give multiplicity-3 sets
14N's).  The host algorithms (shrink, the block list) are the library's
spc_shrink and spc_symmetry_blocks.
"""
from __future__ import annotations

import copy
import os
from typing import List, Optional, Tuple

import numpy as np

from . import _lib
from .dynamics import context, n_steps, trace_sampling_para
from .observables import Observables
from .spinsys import Electron, SimParam, SysParam
from .workloads import Xoroshiro


def shrink(a_iso, M: int) -> Tuple[np.ndarray, np.ndarray]:
    """moments.f90:15-93: couplings and counts of M sets of equivalent nuclei."""
    lib = _lib.load()
    a = np.ascontiguousarray(a_iso, dtype=float)
    Mc = np.array([M], np.int32)
    a_bar = np.zeros(max(M, 1))
    n_bar = np.zeros(max(M, 1), np.int32)
    rc = lib.spc_shrink(len(a), _lib.ptr(a), _lib.ptr(Mc, _lib._i32p), _lib.ptr(a_bar), _lib.ptr(n_bar, _lib._i32p))
    if rc == _lib.SPC_ERR_INVALID and len(a) > 0 and M >= len(a):
        raise _lib.SpinchemError(rc, "Error: Symetry reduction requires M<N")
    if rc:
        raise _lib.SpinchemError(rc, "shrink failed")
    m = int(Mc[0])
    return a_bar[:m].copy(), n_bar[:m].astype(int)


def symmetry_blocks(n_bar) -> Tuple[np.ndarray, np.ndarray, np.ndarray]:
    """cartesian_product + sort_col_prod + weight (symmetry.f90:61-65, 187-226):
    mult [blocks][sets], Z [blocks], copies w [blocks]."""
    lib = _lib.load()
    nb = np.ascontiguousarray(n_bar, dtype=np.int32)
    cnt = np.zeros(1, np.int64)
    rc = lib.spc_symmetry_blocks(len(nb), _lib.ptr(nb, _lib._i32p), 0, _lib.ptr(cnt, _lib._i64p), None, None, None)
    if rc:
        raise _lib.SpinchemError(rc, "symmetry_blocks failed")
    n = int(cnt[0])
    mult = np.zeros((n, max(len(nb), 1)), np.int32)
    Z = np.zeros(n, np.int64)
    w = np.zeros(n, np.int64)
    rc = lib.spc_symmetry_blocks(len(nb), _lib.ptr(nb, _lib._i32p), n, _lib.ptr(cnt, _lib._i64p),
                                 _lib.ptr(mult, _lib._i32p), _lib.ptr(Z, _lib._i64p), _lib.ptr(w, _lib._i64p))
    if rc:
        raise _lib.SpinchemError(rc, "symmetry_blocks failed")
    return mult[:, :len(nb)], Z, w


def reduce_system(sys: SysParam, g_I1, a1_iso, g_I2, a2_iso) -> SysParam:
    """symmetry.f90:146-178: the block's system, isotropic tensors a 1."""
    out = copy.deepcopy(sys)
    out.e1 = Electron(g=sys.e1.g, w=sys.e1.w, g_I=[int(g) for g in g_I1],
                      A=np.array([a * np.eye(3) for a in a1_iso]).reshape(-1, 3, 3))
    out.e2 = Electron(g=sys.e2.g, w=sys.e2.w, g_I=[int(g) for g in g_I2],
                      A=np.array([a * np.eye(3) for a in a2_iso]).reshape(-1, 3, 3))
    return out


def symmetrised_dynamics(sys: SysParam, sim: SimParam, seed=None, out: Optional[str] = None, *,
                         B: Optional[float] = None, ctx: Optional[_lib.Context] = None,
                         log: bool = False) -> Observables:
    """symmetry.f90:16-139.  seed is the master seed of the field
    (main.f90:50); each block draws two next() values from it, in block order,
    as its own master seed.  Blocks whose weight is below sim.block_tol are
    skipped (zero traces).  With out, each computed block writes its traces to
    out/hamiltonian_<i> and the weighted sum, with yields, goes to out."""
    ctx = ctx or context()
    B = -sys.e1.w if B is None else B
    for g in list(sys.e1.g_I) + list(sys.e2.g_I):
        if g != 2:
            raise _lib.SpinchemError(_lib.SPC_ERR_UNSUPPORTED, "Method has not been implemented for nuclei with I>1/2.")
    a1_bar, n1_bar = shrink(sys.e1.a_iso, sim.M1)
    a2_bar, n2_bar = shrink(sys.e2.a_iso, sim.M2)
    m1 = len(a1_bar)
    n_bar = list(n1_bar) + list(n2_bar)
    K, Zb, W = symmetry_blocks(n_bar)
    Z = 2 ** (len(sys.e1.g_I) + len(sys.e2.g_I))
    rng = Xoroshiro(tuple(sim.seed if seed is None else seed))
    signed = lambda x: x - (1 << 64) if x >= (1 << 63) else x  # noqa: E731 (integer(i8) values)
    seeds = [(signed(rng.next()), signed(rng.next())) for _ in range(len(K))]
    n_out = n_steps(sim.t_end, sim.dt) + 1
    res = Observables.zeros(n_out)
    start = sim.restart if sim.isRestart else 1
    for i in range(start - 1, len(K)):
        wk = float(W[i] * Zb[i]) / float(Z)
        if wk < sim.block_tol:
            if log:
                print(f"skipped hamiltonian {i + 1}")
            continue
        red = reduce_system(sys, K[i][:m1], a1_bar, K[i][m1:], a2_bar)
        folder = None
        if out is not None:
            folder = os.path.join(out, f"hamiltonian_{i + 1}")
            os.makedirs(folder, exist_ok=True)
        if log:
            print(f"starting trace sampling of hamiltonian {i + 1}/{len(K)}\nZ = {int(Zb[i])}\nweight_{i + 1} = "
                  f"{wk:.16f}")
        r = trace_sampling_para(red, sim, seeds[i], folder, B=B, ctx=ctx, block=True)
        r.scale(wk)
        res.update(r)
    res.get_kinetics(sim.dt, sys.kS, sys.kT)
    if out is not None:
        res.output(out)
    return res


def blocks_of(sys: SysParam, sim: SimParam) -> List[dict]:
    """The block list of symmetrised_dynamics: multiplicities per set, Z,
    copies and weight."""
    a1_bar, n1_bar = shrink(sys.e1.a_iso, sim.M1)
    a2_bar, n2_bar = shrink(sys.e2.a_iso, sim.M2)
    K, Zb, W = symmetry_blocks(list(n1_bar) + list(n2_bar))
    Z = 2 ** (len(sys.e1.g_I) + len(sys.e2.g_I))
    return [dict(mult=list(map(int, K[i])), Z=int(Zb[i]), copies=int(W[i]), weight=float(W[i] * Zb[i]) / Z)
            for i in range(len(K))]
