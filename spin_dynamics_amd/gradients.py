"""Reverse-mode gradient of the singlet yield on the gradient path of
libspinchem_hip.so (spc_trace_sample_grad), and a torch.autograd.Function
over it.

The differentiated quantity is the final Simpson singlet yield Phi_S(T) of
trace sampling (observables.f90:74-101 over the sample-averaged P_S of
qm_dynamics.f90:257-384) at integrator_tolerance = 0; the adjoint sweep runs
on the device with checkpoints (every 64 steps by default) and
recomputation.  All quantities are in the package's internal units (mT, rates
divided by gamma_e, times in rad/mT).  Tensors are differentiated as given:
the reference's isotropy shortcut (trace/3 when A11 == A22) is not applied,
so the gradient is that of the anisotropic model.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, replace
from typing import Optional, Sequence

import numpy as np

from . import _lib
from .dynamics import _sim_c, context, n_steps
from .spinsys import Electron, SimParam, SysParam, to_c_system
from .workloads import Xoroshiro


def sample_seeds(seed, M: int) -> np.ndarray:
    """Per-sample seeds [M][2]: two next() of set_seed(seed) per sample
    (qm_dynamics.f90:315-320)."""
    g = Xoroshiro(tuple(seed))
    out = np.empty((M, 2), dtype=np.int64)
    for i in range(M):
        out[i, 0] = np.uint64(g.next()).astype(np.int64)
        out[i, 1] = np.uint64(g.next()).astype(np.int64)
    return out


@dataclass
class YieldGrad:
    Phi_S: float
    P_S: np.ndarray
    field: np.ndarray    # d/d(B n), 3
    hfc: np.ndarray      # d/dA_k(a, b) as [k][a][b]
    J: float
    D: np.ndarray        # d/dD(a, b) as [a][b]
    kS: float
    kT: float
    device_ms: float


def yield_gradient(sys: SysParam, sim: SimParam, seed=None, *, B: Optional[float] = None,
                   trace_vectors: Optional[np.ndarray] = None, checkpoint_every: int = 64,
                   ctx: Optional[_lib.Context] = None) -> YieldGrad:
    """Phi_S(T) of trace_sampling_para (same seeds / vectors) and its gradient
    with respect to the field vector, every hyperfine tensor element, J, D,
    kS and kT."""
    ctx = ctx or context()
    B = -sys.e1.w if B is None else B
    s, keep = to_c_system(sys, B, iso_rule=0)
    ctx.check(ctx.lib.spc_system_set(ctx.h, s))
    nn = len(sys.e1.g_I) + len(sys.e2.g_I)
    if trace_vectors is not None:
        tv = np.ascontiguousarray(trace_vectors, dtype=np.complex128)
        M, kind, sd = tv.shape[0], _lib.VEC_INJECTED, None
    else:
        tv = None
        M = sim.N_samples
        kind = _lib.VEC_PHASE if sim.vec_kind == "phase" else _lib.VEC_SUZ
        sd = sample_seeds(sim.seed if seed is None else seed, M)
    n_out = n_steps(sim.t_end, sim.dt) + 1
    P_S = np.zeros(n_out)
    phi = np.zeros(1)
    gf, gA, gJ, gD, gkS, gkT = np.zeros(3), np.zeros(max(9 * nn, 1)), np.zeros(1), np.zeros(9), np.zeros(1), np.zeros(1)
    ctx.check(ctx.lib.spc_trace_sample_grad(ctx.h, _sim_c(sim, M, kind), _lib.ptr(sd, _lib._i64p), _lib.ptr(tv),
                                            int(checkpoint_every), _lib.ptr(P_S), n_out, _lib.ptr(phi),
                                            _lib.ptr(gf), _lib.ptr(gA), _lib.ptr(gJ), _lib.ptr(gD), _lib.ptr(gkS),
                                            _lib.ptr(gkT)))
    del keep
    ms = np.zeros(1)
    ctx.check(ctx.lib.spc_last_run_stats(ctx.h, _lib.ptr(ms), None, None))
    hfc = gA[:9 * nn].reshape(nn, 3, 3).transpose(0, 2, 1).copy()  # column-major -> [k][a][b]
    return YieldGrad(float(phi[0]), P_S, gf, hfc, float(gJ[0]), gD.reshape(3, 3).T.copy(), float(gkS[0]),
                     float(gkT[0]), float(ms[0]))


# ------------------------------------------------------------------ torch
SYM = [(0, 0), (1, 1), (2, 2), (0, 1), (0, 2), (1, 2)]  # the 6 independent entries of a symmetric tensor


def tensors_from_sym(h6: np.ndarray) -> np.ndarray:
    """(nn, 6) [xx, yy, zz, xy, xz, yz] -> (nn, 3, 3) symmetric."""
    h6 = np.asarray(h6, float).reshape(-1, 6)
    A = np.zeros((h6.shape[0], 3, 3))
    for q, (a, b) in enumerate(SYM):
        A[:, a, b] = h6[:, q]
        A[:, b, a] = h6[:, q]
    return A


def field_direction(theta: float, phi: float) -> np.ndarray:
    return np.array([math.sin(theta) * math.cos(phi), math.sin(theta) * math.sin(phi), math.cos(theta)])


@dataclass
class YieldProblem:
    """What stays fixed while the yield is differentiated: the nuclei, D,
    the simulation and its trace vectors."""
    mult1: Sequence[int]
    mult2: Sequence[int]
    sim: SimParam
    D: np.ndarray = None
    seed: Optional[tuple] = None
    trace_vectors: Optional[np.ndarray] = None
    checkpoint_every: int = 64
    ctx: Optional[_lib.Context] = None

    def system(self, h6, B, theta, phi, J, kS, kT) -> SysParam:
        A = tensors_from_sym(h6)
        n1 = len(self.mult1)
        sys = SysParam(e1=Electron(g_I=list(self.mult1), A=A[:n1]), e2=Electron(g_I=list(self.mult2), A=A[n1:]),
                       J=float(J), D=np.zeros((3, 3)) if self.D is None else np.asarray(self.D, float),
                       kS=float(kS), kT=float(kT), field_dir=tuple(field_direction(theta, phi)))
        sys.e1.w = sys.e2.w = -float(B)
        return sys

    def value(self, h6, B, theta, phi, J, kS, kT) -> float:
        """Phi_S alone: the forward run of yield_gradient (same system, seeds or
        trace vectors, no adjoint)."""
        sys = self.system(h6, B, theta, phi, J, kS, kT)
        ctx = self.ctx or context()
        s, keep = to_c_system(sys, float(B), iso_rule=0)
        ctx.check(ctx.lib.spc_system_set(ctx.h, s))
        sim = self.sim
        n_out = n_steps(sim.t_end, sim.dt) + 1
        outs = [np.zeros(n_out) for _ in range(4)]
        phis, phit = np.zeros(max(n_out // 2, 1)), np.zeros(max(n_out // 2, 1))
        tail = [*[_lib.ptr(o) for o in outs], n_out, _lib.ptr(phis), _lib.ptr(phit)]
        if self.trace_vectors is not None:
            tv = np.ascontiguousarray(self.trace_vectors, dtype=np.complex128)
            seed0 = np.zeros(2, np.int64)
            ctx.check(ctx.lib.spc_trace_sample(ctx.h, _sim_c(sim, tv.shape[0], _lib.VEC_INJECTED),
                                               _lib.ptr(seed0, _lib._i64p), _lib.ptr(tv), *tail))
        else:
            kind = _lib.VEC_PHASE if sim.vec_kind == "phase" else _lib.VEC_SUZ
            sd = sample_seeds(sim.seed if self.seed is None else self.seed, sim.N_samples)
            ctx.check(ctx.lib.spc_trace_sample_seeds(ctx.h, _sim_c(sim, sim.N_samples, kind), _lib.ptr(sd, _lib._i64p),
                                                     *tail))
        del keep
        return float(phis[n_out // 2 - 1]) if n_out // 2 >= 1 else 0.0

    def evaluate(self, h6, B, theta, phi, J, kS, kT):
        """Phi_S and its gradient in the parameterisation (h6, |B|, theta, phi, J, kS, kT)."""
        sys = self.system(h6, B, theta, phi, J, kS, kT)
        g = yield_gradient(sys, self.sim, self.seed, B=float(B), trace_vectors=self.trace_vectors,
                           checkpoint_every=self.checkpoint_every, ctx=self.ctx)
        h6 = np.asarray(h6, float).reshape(-1, 6)
        gh6 = np.zeros_like(h6)
        for q, (a, b) in enumerate(SYM):
            gh6[:, q] = g.hfc[:, a, b] + (g.hfc[:, b, a] if a != b else 0.0)
        n = field_direction(theta, phi)
        dn_dth = np.array([math.cos(theta) * math.cos(phi), math.cos(theta) * math.sin(phi), -math.sin(theta)])
        dn_dph = np.array([-math.sin(theta) * math.sin(phi), math.sin(theta) * math.cos(phi), 0.0])
        grads = dict(h6=gh6, B=float(g.field @ n), theta=float(B * (g.field @ dn_dth)),
                     phi=float(B * (g.field @ dn_dph)), J=g.J, kS=g.kS, kT=g.kT)
        return g.Phi_S, grads, g


def singlet_yield_function():
    """torch.autograd.Function Phi_S(h6, B, theta, phi, J, kS, kT; problem).

    When some input requires a gradient, forward runs the adjoint with the
    forward pass (spc_trace_sample_grad keeps the forward's step vectors for
    it: 2.9x a forward run in all, against 3.9x for a forward now and the
    gradient call in backward) and backward scales the saved gradients;
    otherwise forward runs the forward pass alone."""
    import torch

    class SingletYield(torch.autograd.Function):
        @staticmethod
        def forward(fctx, h6, B, theta, phi, J, kS, kT, problem: YieldProblem):
            vals = [t.detach().cpu().double().numpy() for t in (h6, B, theta, phi, J, kS, kT)]
            args = (vals[0], *[float(v) for v in vals[1:]])
            dev, dt = h6.device, torch.float64
            fctx.with_grad = any(fctx.needs_input_grad[:7])
            if not fctx.with_grad:
                return torch.tensor(problem.value(*args), dtype=dt, device=dev)
            F, g, _ = problem.evaluate(*args)
            fctx.save_for_backward(torch.as_tensor(g["h6"], dtype=dt, device=dev),
                                   *[torch.tensor(g[k], dtype=dt, device=dev)
                                     for k in ("B", "theta", "phi", "J", "kS", "kT")])
            return torch.tensor(F, dtype=dt, device=dev)

        @staticmethod
        def backward(fctx, grad_out):
            if not fctx.with_grad:
                return (None,) * 8
            gh6, gB, gth, gph, gJ, gkS, gkT = fctx.saved_tensors
            return (grad_out * gh6, grad_out * gB, grad_out * gth, grad_out * gph, grad_out * gJ, grad_out * gkS,
                    grad_out * gkT, None)

    return SingletYield
