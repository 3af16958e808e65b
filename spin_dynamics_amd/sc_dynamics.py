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


def run_tangent(a_iso1, a_iso2, mult1, mult2, *, B: float, dfield, n_samples: int, dt: float, n_steps_: int,
                direction=(0, 0, 1), J: float = 0.0, kS: float = 0.0, kT: float = 0.0, seed=(42, 99),
                ctx: Optional[_lib.Context] = None):
    """run() together with the derivatives of P_S, P_T, Phi_S, Phi_T for the
    field vector B n moving along dfield (spc_semiclassical_tangent): returns
    (ps, pt, phis, phit), (dps, dpt, dphis, dphit)."""
    ctx = ctx or context()
    s, keep = sc_system(a_iso1, a_iso2, mult1, mult2, B=B, direction=direction, J=J, kS=kS, kT=kT)
    n_out = n_steps_ + 1
    nh = max(n_out // 2, 1)
    ps, pt, dps, dpt = (np.zeros(n_out) for _ in range(4))
    phis, phit, dphis, dphit = (np.zeros(nh) for _ in range(4))
    sd = np.ascontiguousarray(seed, dtype=np.int64)
    df = np.ascontiguousarray(dfield, dtype=np.float64)
    ctx.check(ctx.lib.spc_semiclassical_tangent(ctx.h, s, int(n_samples), _lib.ptr(sd, _lib._i64p), float(dt),
                                                int(n_steps_), _lib.ptr(df), _lib.ptr(ps), _lib.ptr(pt),
                                                _lib.ptr(dps), _lib.ptr(dpt), _lib.ptr(phis), _lib.ptr(phit),
                                                _lib.ptr(dphis), _lib.ptr(dphit)))
    del keep
    h = n_out // 2
    return (ps, pt, phis[:h], phit[:h]), (dps, dpt, dphis[:h], dphit[:h])


def _direction(theta: float, phi: float) -> np.ndarray:
    return np.array([np.sin(theta) * np.cos(phi), np.sin(theta) * np.sin(phi), np.cos(theta)])


class ScYieldProblem:
    """A semiclassical yield Phi_S(T) as a function of the field (|B|, theta,
    phi): the remaining parameters and the samples are fixed (same seed, so
    the estimator is a smooth function of the field)."""

    def __init__(self, a_iso1, a_iso2, mult1, mult2, *, n_samples: int, dt: float, n_steps_: int, J: float = 0.0,
                 kS: float = 0.0, kT: float = 0.0, seed=(42, 99), ctx: Optional[_lib.Context] = None):
        self.args = (list(a_iso1), list(a_iso2), list(mult1), list(mult2))
        self.kw = dict(n_samples=n_samples, dt=dt, n_steps_=n_steps_, J=J, kS=kS, kT=kT, seed=seed)
        self.ctx = ctx

    def value(self, B: float, theta: float, phi: float) -> float:
        _, _, phis, _ = run(*self.args, B=B, direction=_direction(theta, phi), ctx=self.ctx, **self.kw)
        return float(phis[-1])

    def derivative(self, B: float, theta: float, phi: float, dfield) -> tuple:
        """(Phi_S(T), its derivative along the field-vector direction dfield)."""
        (_, _, phis, _), (_, _, dphis, _) = run_tangent(*self.args, B=B, dfield=dfield,
                                                        direction=_direction(theta, phi), ctx=self.ctx, **self.kw)
        return float(phis[-1]), float(dphis[-1])

    def gradient(self, B: float, theta: float, phi: float, need=(True, True, True)) -> tuple:
        """Phi_S(T) and d/d(|B|, theta, phi): one tangent run per component
        asked for (d/dB along n, d/dtheta along B dn/dtheta, d/dphi along
        B dn/dphi)."""
        n = _direction(theta, phi)
        dirs = (n, B * np.array([np.cos(theta) * np.cos(phi), np.cos(theta) * np.sin(phi), -np.sin(theta)]),
                B * np.array([-np.sin(theta) * np.sin(phi), np.sin(theta) * np.cos(phi), 0.0]))
        F, g = None, [0.0, 0.0, 0.0]
        for i in range(3):
            if need[i]:
                F, g[i] = self.derivative(B, theta, phi, dirs[i])
        if F is None:
            F = self.value(B, theta, phi)
        return F, g


def semiclassical_yield_function():
    """A torch.autograd.Function F = Phi_S(T)(B, theta, phi; problem) whose
    backward uses the forward-mode tangents of spc_semiclassical_tangent (one
    run per field component that needs a gradient)."""
    import torch

    class _ScYield(torch.autograd.Function):
        @staticmethod
        def forward(ctx, B, theta, phi, problem):
            vals = [float(t.detach().cpu()) for t in (B, theta, phi)]
            need = ctx.needs_input_grad[:3]
            F, g = problem.gradient(*vals, need=need) if any(need) else (problem.value(*vals), [0.0] * 3)
            ctx.save_for_backward(torch.tensor(g, dtype=torch.float64))
            return torch.tensor(F, dtype=torch.float64)

        @staticmethod
        def backward(ctx, gout):
            (g,) = ctx.saved_tensors
            return gout * g[0], gout * g[1], gout * g[2], None

    return _ScYield


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
