"""Many field orientations of a radical pair in one lane batch
(spc_orientation_batch), its autotuner (spc_batch_tune), and the orientation
grids and radical-pair sets the batched workloads use.

The reference evaluates one field per run (app/main.f90:48-79, the same
trace vectors for every field); here all orientations of a pair share one
block pattern of H, each 64-lane chunk carries its orientation's values, and
one captured step is replayed from a hipGraph.
"""
from __future__ import annotations

import math
from dataclasses import dataclass
from typing import List, Optional, Sequence

import numpy as np

from . import _lib
from .dynamics import _sim_c, context, n_steps
from .gradients import sample_seeds
from .spinsys import GAMMA_E, SimParam, SysParam, to_c_system
from .workloads import SYSTEM_SEED, Xoroshiro, _electron, _sim, random_tensor

GOLDEN_ANGLE = math.pi * (3.0 - math.sqrt(5.0))


def fibonacci_sphere(n: int) -> np.ndarray:
    """n unit vectors z_i = 1 - (2i+1)/n, azimuth i * golden angle."""
    i = np.arange(n)
    z = 1.0 - (2 * i + 1) / n
    r = np.sqrt(1.0 - z * z)
    ph = i * GOLDEN_ANGLE
    return np.stack([r * np.cos(ph), r * np.sin(ph), z], axis=1)


def _prep(sys: SysParam, sim: SimParam, B, seed, ctx):
    ctx = ctx or context()
    B = -sys.e1.w if B is None else B
    s, keep = to_c_system(sys, B)
    ctx.check(ctx.lib.spc_system_set(ctx.h, s))
    sd = sample_seeds(sim.seed if seed is None else seed, sim.N_samples)
    kind = _lib.VEC_PHASE if sim.vec_kind == "phase" else _lib.VEC_SUZ
    return ctx, _sim_c(sim, sim.N_samples, kind), sd, keep


def orientation_yields(sys: SysParam, sim: SimParam, dirs: np.ndarray, seed=None, *, B: Optional[float] = None,
                       orient_per_chunk: int = 0, graph: bool = True, with_populations: bool = False,
                       ctx: Optional[_lib.Context] = None):
    """Final singlet yield Phi_S for every direction in dirs ([n][3]) at field
    magnitude B (default -sys.e1.w); with_populations also returns P_S [n][n_out]."""
    ctx, simc, sd, keep = _prep(sys, sim, B, seed, ctx)
    dirs = np.ascontiguousarray(dirs, dtype=np.float64).reshape(-1, 3)
    n = dirs.shape[0]
    n_out = n_steps(sim.t_end, sim.dt) + 1
    phi = np.zeros(n)
    ps = np.zeros((n, n_out)) if with_populations else None
    ctx.check(ctx.lib.spc_orientation_batch(ctx.h, simc, _lib.ptr(sd, _lib._i64p), n, _lib.ptr(dirs),
                                            int(orient_per_chunk), 1 if graph else 0, _lib.ptr(phi), _lib.ptr(ps),
                                            n_out))
    del keep
    return (phi, ps) if with_populations else phi


def tune(sys: SysParam, sim: SimParam, dirs: np.ndarray, seed=None, *, B: Optional[float] = None,
         yield_tol: float = 1e-10, ctx: Optional[_lib.Context] = None):
    """(orientations per chunk, Krylov order, ms per orientation-step) chosen by spc_batch_tune."""
    ctx, simc, sd, keep = _prep(sys, sim, B, seed, ctx)
    dirs = np.ascontiguousarray(dirs, dtype=np.float64).reshape(-1, 3)
    opc = np.zeros(1, np.int32)
    m = np.zeros(1, np.int32)
    ms = np.zeros(1)
    ctx.check(ctx.lib.spc_batch_tune(ctx.h, simc, _lib.ptr(sd, _lib._i64p), dirs.shape[0], _lib.ptr(dirs),
                                     float(yield_tol), _lib.ptr(opc, _lib._i32p), _lib.ptr(m, _lib._i32p),
                                     _lib.ptr(ms)))
    del keep
    return int(opc[0]), int(m[0]), float(ms[0])


def tune_adaptive(sys: SysParam, sim: SimParam, dirs: np.ndarray, seed=None, *, B: Optional[float] = None,
                  yield_tol: float = 1e-10, ctx: Optional[_lib.Context] = None):
    """(orientations per chunk, Krylov order, tol, ms per orientation-step) chosen by
    spc_batch_tune_adaptive: several steps per basis under the reference's
    rebuild rule, at yields within yield_tol of sim.N_krylov with one step
    per basis."""
    ctx, simc, sd, keep = _prep(sys, sim, B, seed, ctx)
    dirs = np.ascontiguousarray(dirs, dtype=np.float64).reshape(-1, 3)
    opc = np.zeros(1, np.int32)
    m = np.zeros(1, np.int32)
    tol = np.zeros(1)
    ms = np.zeros(1)
    ctx.check(ctx.lib.spc_batch_tune_adaptive(ctx.h, simc, _lib.ptr(sd, _lib._i64p), dirs.shape[0], _lib.ptr(dirs),
                                              float(yield_tol), _lib.ptr(opc, _lib._i32p), _lib.ptr(m, _lib._i32p),
                                              _lib.ptr(tol), _lib.ptr(ms)))
    del keep
    return int(opc[0]), int(m[0]), float(tol[0]), float(ms[0])


def pairs_yields(systems: Sequence[SysParam], sim: SimParam, dirs: np.ndarray, seed=None, *,
                 B: Optional[Sequence[float]] = None, n_krylov: Optional[Sequence[int]] = None,
                 tol: Optional[Sequence[float]] = None, orient_per_chunk: Optional[Sequence[int]] = None,
                 ctx: Optional[_lib.Context] = None) -> np.ndarray:
    """Final singlet yields [n_pairs][n_orient] of several radical pairs (of any
    dimensions) at the directions dirs, in one spc_pairs_batch call: every
    pair resident at once, the chunks of all pairs sharing the streams.  B[p]
    defaults to -systems[p].e1.w; per-pair order, tolerance and width default
    to sim's and to as many orientations as fit."""
    import ctypes
    ctx = ctx or context()
    n = len(systems)
    Bs = [(-s.e1.w if B is None else B[i]) for i, s in enumerate(systems)]
    arr = (_lib.SpinSystem * n)()
    keep = []
    for i, s in enumerate(systems):
        c, k = to_c_system(s, Bs[i])
        arr[i] = c
        keep.append(k)
    sd = sample_seeds(sim.seed if seed is None else seed, sim.N_samples)
    kind = _lib.VEC_PHASE if sim.vec_kind == "phase" else _lib.VEC_SUZ
    simc = _sim_c(sim, sim.N_samples, kind)
    dirs = np.ascontiguousarray(dirs, dtype=np.float64).reshape(-1, 3)
    mk = None if n_krylov is None else np.ascontiguousarray(n_krylov, dtype=np.int32)
    tl = None if tol is None else np.ascontiguousarray(tol, dtype=np.float64)
    oc = None if orient_per_chunk is None else np.ascontiguousarray(orient_per_chunk, dtype=np.int32)
    phi = np.zeros((n, dirs.shape[0]))
    ctx.check(ctx.lib.spc_pairs_batch(ctx.h, n, ctypes.cast(arr, ctypes.POINTER(_lib.SpinSystem)), simc,
                                      _lib.ptr(sd, _lib._i64p), dirs.shape[0], _lib.ptr(dirs),
                                      _lib.ptr(mk, _lib._i32p), _lib.ptr(tl), _lib.ptr(oc, _lib._i32p),
                                      _lib.ptr(phi)))
    del keep
    return phi


# ------------------------------------------------------------ pair sets
DIMS = (768, 3072, 12288, 49152, 147456)
_NUCLEI = {768: (1, 6), 3072: (1, 8), 12288: (1, 10), 49152: (1, 12), 147456: (2, 12)}  # (14N, 1H)


@dataclass
class Pair:
    name: str
    sys: SysParam
    N: int


def radical_pairs(n_pairs: int = 48, seed=SYSTEM_SEED, J: float = 0.01, kS_us: float = 2.0, kT_us: float = 0.5,
                  half_width: float = 1.5) -> List[Pair]:
    """Radical pairs of uneven dimension: for each pair, in order, a dimension
    class from {768, 3072, 12288, 49152, 147456} (floor(5 unif_01)), the number
    of 1H on radical A (floor((n_1H + 1) unif_01); the 14N all on A, the rest of
    the 1H on B), then the tensors of A's then B's nuclei with random_tensor
    (eigenvalues U[-1.5, 1.5] mT, random rotation).  J, kS, kT as the plain
    36 864-dimensional case; the field is set per orientation."""
    rng = Xoroshiro(seed)
    pairs = []
    for p in range(n_pairs):
        N = DIMS[min(int(5 * rng.unif01()), 4)]
        n14, n1h = _NUCLEI[N]
        a1h = min(int((n1h + 1) * rng.unif01()), n1h)
        m1 = [3] * n14 + [2] * a1h
        m2 = [2] * (n1h - a1h)
        T = [random_tensor(rng, half_width) for _ in range(len(m1) + len(m2))]
        sys = SysParam(e1=_electron(m1, T[:len(m1)]), e2=_electron(m2, T[len(m1):]), J=J, kS=kS_us / GAMMA_E,
                       kT=kT_us / GAMMA_E)
        sys.e1.w = sys.e2.w = -0.05
        pairs.append(Pair(f"pair{p:02d}_N{N}", sys, N))
    return pairs


def orientation_sim(M: int = 64, m: int = 24, T_ns: float = 10000.0) -> SimParam:
    """M = 64 phase vectors, Krylov order 24 (before tuning), dt = 1 ns and the
    plain 36 864-dimensional case's T = 10 us (the batched case states no T)."""
    return _sim(M, m, T_ns, 1.0, vec_kind="phase")
