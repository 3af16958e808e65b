"""Module `sc_dynamics` of the reference (src/sc/sc_dynamics.f90): the
semiclassical singlet/triplet correlation functions on the gfx950 path.

The reference's semi_classical (sc_dynamics.f90:17-100) is unreachable from
its driver (app/main.f90:74-75 prints "method not yet implemented") and its
module does not compile; this keeps its name and observables and runs the
model it describes through spc_semiclassical_run.
"""
from __future__ import annotations

from typing import Optional, Sequence

import numpy as np

from . import _lib
from .dynamics import context, n_steps
from .observables import Observables
from .spinsys import SimParam, SysParam


def sc_system(a_iso1: Sequence[float], a_iso2: Sequence[float], mult1: Sequence[int], mult2: Sequence[int], *,
              B: float, direction=(0.0, 0.0, 1.0), J: float = 0.0, kS: float = 0.0, kT: float = 0.0):
    """spc_sc_system for couplings a_iso (mT) and multiplicities 2I+1 per electron."""
    a = np.ascontiguousarray(list(a_iso1) + list(a_iso2), dtype=np.float64)
    g = np.ascontiguousarray(list(mult1) + list(mult2), dtype=np.int32)
    s = _lib.ScSystem()
    s.n_nuc1, s.n_nuc2 = len(a_iso1), len(a_iso2)
    s.mult = _lib.ptr(g, _lib._i32p)
    s.a_iso = _lib.ptr(a)
    s.B, s.J, s.kS, s.kT = B, J, kS, kT
    for i in range(3):
        s.field_dir[i] = direction[i]
    return s, (a, g)  # the arrays must outlive the call


def run(a_iso1, a_iso2, mult1, mult2, *, B: float, n_samples: int, dt: float, n_steps_: int, direction=(0, 0, 1),
        J: float = 0.0, kS: float = 0.0, kT: float = 0.0, seed=(42, 99), ctx: Optional[_lib.Context] = None):
    """P_S, P_T (n_steps_ + 1 values each) and the Simpson yields Phi_S, Phi_T."""
    ctx = ctx or context()
    s, keep = sc_system(a_iso1, a_iso2, mult1, mult2, B=B, direction=direction, J=J, kS=kS, kT=kT)
    n_out = n_steps_ + 1
    ps, pt = np.zeros(n_out), np.zeros(n_out)
    phis, phit = np.zeros(max(n_out // 2, 1)), np.zeros(max(n_out // 2, 1))
    sd = np.ascontiguousarray(seed, dtype=np.int64)
    ctx.check(ctx.lib.spc_semiclassical_run(ctx.h, s, int(n_samples), _lib.ptr(sd, _lib._i64p), float(dt),
                                            int(n_steps_), _lib.ptr(ps), _lib.ptr(pt), _lib.ptr(phis),
                                            _lib.ptr(phit)))
    del keep
    return ps, pt, phis[:n_out // 2], phit[:n_out // 2]


def semi_classical(sys: SysParam, sim: SimParam, seed=None, out: Optional[str] = None, *, B: Optional[float] = None,
                   ctx: Optional[_lib.Context] = None) -> Observables:
    """sc_dynamics.f90:17-100 with the field B (default -sys.e1.w): N_samples
    classical samples, N_steps = ceiling((t_end+dt)/dt) steps of dt and
    N_steps + 1 values of P_S and P_T; yields and files as res%get_kinetics /
    res%output.  Isotropic couplings a_iso as src/sc uses."""
    B = -sys.e1.w if B is None else B
    nst = n_steps(sim.t_end, sim.dt)
    ps, pt, phis, phit = run(sys.e1.a_iso, sys.e2.a_iso, sys.e1.g_I, sys.e2.g_I, B=B, n_samples=sim.N_samples,
                             dt=sim.dt, n_steps_=nst, direction=sys.field_dir, J=sys.J, kS=sys.kS, kT=sys.kT,
                             seed=sim.seed if seed is None else seed, ctx=ctx)
    z = np.zeros_like(ps)
    res = Observables(ps, z, z.copy(), z.copy(), iden=ps + pt, Phi_S=phis, Phi_T=phit, P_T=pt)
    if out is not None:
        res.output(out)
    return res
