"""GPU adjoint gradient of the singlet yield (spc_trace_sample_grad).

Pins: central finite differences of the device forward (the same trace
vectors, tol = 0) for every parameter of H; central finite differences of
the oracle's forward (orc_trace_sampling with injected vectors) for a set of
parameters, so the gradient is that of the reference algorithm's yield;
independence of the checkpoint interval and of how many final steps keep
their vectors instead of being recomputed; the overlapped recomputation
equal to the one-stream sweep, bit for bit; the yield itself equal to the
forward run's; and the torch.autograd.Function's chain rule to the
(symmetric tensor, |B|, theta, phi, J, kS, kT) parameterisation, again
against finite differences.  Tolerance: 1e-6 relative to the largest
gradient entry (finite-difference step 1e-6 relative)."""
import math

import numpy as np
import pytest

from helpers import oracle_system
from oracle import pyoracle
from spin_dynamics_amd import dynamics, gradients, workloads
from spin_dynamics_amd.spinsys import GAMMA_E

pytestmark = pytest.mark.gpu


def _small(M=8, T_ns=40.0, m=12):
    w = workloads.small_anisotropic(m1=(3, 2), m2=(2,), M=M, m=m, T_ns=T_ns)
    w.sys.D = np.array([[0.02, 0.01, 0.0], [0.01, -0.03, 0.005], [0.0, 0.005, 0.01]])
    return w


def _forward_yield(sys, sim, B, ctx, trace_vectors=None):
    res = dynamics.trace_sampling_para(sys, sim, B=B, ctx=ctx, trace_vectors=trace_vectors)
    res.get_kinetics(sim.dt, sys.kS, sys.kT)
    return res.Phi_S[-1]


def _perturbed(sys, B, name, idx, h):
    """Copy of (sys, B) with one linear parameter moved by h."""
    import copy
    s = copy.deepcopy(sys)
    if name == "hfc":
        k, a, b = idx
        n1 = len(s.e1.g_I)
        e, kk = (s.e1, k) if k < n1 else (s.e2, k - n1)
        A = np.array(e.A, float)
        A[kk, a, b] += h
        e.A = A
        return s, B
    if name == "field":
        bvec = B * np.asarray(s.field_dir, float) / np.linalg.norm(s.field_dir)
        bvec[idx] += h
        Bn = float(np.linalg.norm(bvec))
        s.field_dir = tuple(bvec / Bn)
        return s, Bn
    if name == "D":
        D = np.array(s.D, float)
        D[idx] += h
        s.D = D
        return s, B
    setattr(s, name, getattr(s, name) + h)
    return s, B


def _params(sys):
    nn = len(sys.e1.g_I) + len(sys.e2.g_I)
    out = [("hfc", (k, a, b)) for k in range(nn) for a in range(3) for b in range(3)]
    out += [("field", a) for a in range(3)] + [("J", None), ("kS", None), ("kT", None)]
    out += [("D", (a, b)) for a in range(3) for b in range(3)]
    return out


def _adj_value(g, name, idx):
    if name == "hfc":
        return g.hfc[idx]
    if name == "field":
        return g.field[idx]
    if name == "D":
        return g.D[idx]
    return getattr(g, name)


def test_yield_matches_forward_and_every_parameter_matches_fd(ctx, monkeypatch):
    w = _small()
    g = gradients.yield_gradient(w.sys, w.sim, B=w.B, ctx=ctx)
    F0 = _forward_yield(w.sys, w.sim, w.B, ctx)
    monkeypatch.setenv("SPC_ADJOINT_KEEP", "0")
    g0 = gradients.yield_gradient(w.sys, w.sim, B=w.B, checkpoint_every=5, ctx=ctx)
    assert np.max(np.abs(g0.hfc - g.hfc)) <= 1e-12 * np.max(np.abs(g.hfc))
    assert abs(g.Phi_S - F0) <= 1e-12 * abs(F0)
    fwd = dynamics.trace_sampling_para(w.sys, w.sim, B=w.B, ctx=ctx)
    assert np.max(np.abs(g.P_S - fwd.P_S)) < 1e-13
    adj, fd = [], []
    for name, idx in _params(w.sys):
        base = _adj_value(g, name, idx)
        if name in ("kS", "kT"):
            h = 1e-6 * getattr(w.sys, name)
        else:
            h = 1e-6
        sp, Bp = _perturbed(w.sys, w.B, name, idx, h)
        sm, Bm = _perturbed(w.sys, w.B, name, idx, -h)
        d = (_forward_yield(sp, w.sim, Bp, ctx) - _forward_yield(sm, w.sim, Bm, ctx)) / (2 * h)
        adj.append(base)
        fd.append(d)
    adj, fd = np.array(adj), np.array(fd)
    scale = np.max(np.abs(fd))
    assert scale > 0
    assert np.max(np.abs(adj - fd)) <= 1e-6 * scale, (adj, fd)


def test_gradient_matches_oracle_finite_differences(ctx):
    w = _small(M=6, T_ns=30.0)
    Z = w.sys.Z1 * w.sys.Z2
    rng = np.random.default_rng(5)
    tv = rng.normal(size=(6, Z)) + 1j * rng.normal(size=(6, Z))
    tv /= np.linalg.norm(tv, axis=1, keepdims=True)
    g = gradients.yield_gradient(w.sys, w.sim, B=w.B, trace_vectors=tv, ctx=ctx)

    def oracle_yield(sys, B):
        s = oracle_system(sys, B)
        h = pyoracle.build_general_csr(s)
        r = pyoracle.trace_sampling(h, Z, M=6, m=w.sim.N_krylov, t_end=w.sim.t_end, dt=w.sim.dt, vec_kind=2,
                                    injected=tv)
        _, phis, _ = pyoracle.get_kinetics(r["P_S"], r["P_T0"], r["P_Tp"], r["P_Tm"], w.sim.dt, sys.kS, sys.kT)
        return phis[-1]

    assert abs(oracle_yield(w.sys, w.B) - g.Phi_S) <= 1e-10 * abs(g.Phi_S)
    picks = [("hfc", (0, 0, 0)), ("hfc", (0, 1, 2)), ("hfc", (2, 2, 1)), ("field", 0), ("field", 2), ("J", None),
             ("kS", None), ("kT", None), ("D", (0, 1))]
    adj, fd = [], []
    for name, idx in picks:
        h = 1e-6 * getattr(w.sys, name) if name in ("kS", "kT") else 1e-6
        sp, Bp = _perturbed(w.sys, w.B, name, idx, h)
        sm, Bm = _perturbed(w.sys, w.B, name, idx, -h)
        fd.append((oracle_yield(sp, Bp) - oracle_yield(sm, Bm)) / (2 * h))
        adj.append(_adj_value(g, name, idx))
    adj, fd = np.array(adj), np.array(fd)
    assert np.max(np.abs(adj - fd)) <= 1e-6 * np.max(np.abs(fd)), (adj, fd)


@pytest.mark.parametrize("every,keep", [(1, "0"), (7, "0"), (64, "0"), (7, "13"), (64, None)])
def test_checkpoint_interval_does_not_change_the_gradient(ctx, monkeypatch, every, keep):
    w = _small(T_ns=50.0)
    ref = gradients.yield_gradient(w.sys, w.sim, B=w.B, checkpoint_every=64, ctx=ctx)  # all steps kept
    if keep is not None:
        monkeypatch.setenv("SPC_ADJOINT_KEEP", keep)  # recompute all but the last `keep` steps
    g = gradients.yield_gradient(w.sys, w.sim, B=w.B, checkpoint_every=every, ctx=ctx)
    s = np.max(np.abs(ref.hfc))
    assert np.max(np.abs(g.hfc - ref.hfc)) <= 1e-12 * s
    assert np.max(np.abs(g.field - ref.field)) <= 1e-12 * s
    assert abs(g.kS - ref.kS) <= 1e-12 * max(abs(ref.kS), s)


@pytest.mark.parametrize("every,keep", [(1, "0"), (7, "0"), (7, "13")])
def test_overlapped_recomputation_equals_one_stream(ctx, monkeypatch, every, keep):
    # segments recomputed on a second stream while the adjoint steps of the
    # next segment run: bit for bit the gradient of one stream, one kernel at
    # a time (SPC_LANE_SERIAL=1), with many segments and both buffer parities
    w = _small(T_ns=50.0)
    monkeypatch.setenv("SPC_ADJOINT_KEEP", keep)
    a = gradients.yield_gradient(w.sys, w.sim, B=w.B, checkpoint_every=every, ctx=ctx)
    monkeypatch.setenv("SPC_LANE_SERIAL", "1")
    b = gradients.yield_gradient(w.sys, w.sim, B=w.B, checkpoint_every=every, ctx=ctx)
    for k in ("hfc", "field", "J", "D", "kS", "kT", "Phi_S"):
        assert np.array_equal(np.asarray(getattr(a, k)), np.asarray(getattr(b, k))), k


def test_odd_and_even_step_counts(ctx):
    # the last Simpson weight falls on the last step or one before it
    for T in (21.0, 22.0):
        w = _small(T_ns=T)
        g = gradients.yield_gradient(w.sys, w.sim, B=w.B, ctx=ctx)
        h = 1e-6
        sp, _ = _perturbed(w.sys, w.B, "hfc", (1, 2, 2), h)
        sm, _ = _perturbed(w.sys, w.B, "hfc", (1, 2, 2), -h)
        fd = (_forward_yield(sp, w.sim, w.B, ctx) - _forward_yield(sm, w.sim, w.B, ctx)) / (2 * h)
        assert abs(g.hfc[1, 2, 2] - fd) <= 1e-6 * max(abs(fd), np.max(np.abs(g.hfc)))


def test_torch_function_chain_rule(ctx):
    torch = pytest.importorskip("torch")
    w = _small(M=4, T_ns=20.0)
    sim = w.sim
    prob = gradients.YieldProblem(mult1=[3, 2], mult2=[2], sim=sim, ctx=ctx)
    rng = np.random.default_rng(3)
    h6 = rng.uniform(-1, 1, size=(3, 6))
    vals = dict(B=0.3, theta=0.6, phi=0.4, J=0.02, kS=2.0 / GAMMA_E, kT=0.5 / GAMMA_E)
    f = gradients.singlet_yield_function()
    th = torch.tensor(h6, dtype=torch.float64, requires_grad=True)
    scal = {k: torch.tensor(v, dtype=torch.float64, requires_grad=True) for k, v in vals.items()}
    F = f.apply(th, scal["B"], scal["theta"], scal["phi"], scal["J"], scal["kS"], scal["kT"], prob)
    F.backward()

    def value(h6_, **kw):
        p = dict(vals)
        p.update(kw)
        return prob.evaluate(h6_, p["B"], p["theta"], p["phi"], p["J"], p["kS"], p["kT"])[0]

    hs = 1e-6
    for q in (0, 3, 5):
        d = np.zeros_like(h6)
        d[1, q] = hs
        fd = (value(h6 + d) - value(h6 - d)) / (2 * hs)
        assert abs(th.grad[1, q].item() - fd) <= 1e-6 * max(abs(fd), 1e-3)
    for k in vals:
        step = hs * max(abs(vals[k]), 1.0)
        fd = (value(h6, **{k: vals[k] + step}) - value(h6, **{k: vals[k] - step})) / (2 * step)
        assert abs(scal[k].grad.item() - fd) <= 1e-6 * max(abs(fd), 1e-3), k
    assert math.isclose(F.item(), value(h6), rel_tol=1e-13)
    # no input needs a gradient: forward alone (the forward run of the gradient
    # call: the same value), and its device time is a forward run's
    F0 = f.apply(th.detach(), *[scal[k].detach() for k in ("B", "theta", "phi", "J", "kS", "kT")], prob)
    assert math.isclose(F0.item(), F.item(), rel_tol=1e-12)
    _, rebuilds, steps = dynamics.last_run_stats(ctx)
    assert rebuilds == steps  # one trace-sampling pass, no adjoint sweep


def test_gradient_independent_of_the_product_kernel(ctx, monkeypatch):
    # five trailing spin-1/2 nuclei: tiled product by default; the 4x4-block
    # product sums the same terms in the same order
    w = workloads.small_anisotropic(m1=(3,), m2=(2, 2, 2, 2, 2), M=8, m=10, T_ns=12.0)
    g = gradients.yield_gradient(w.sys, w.sim, B=w.B, ctx=ctx)
    monkeypatch.setenv("SPC_SPMM", "blocks")
    gb = gradients.yield_gradient(w.sys, w.sim, B=w.B, ctx=ctx)
    s = np.max(np.abs(gb.hfc))
    assert np.max(np.abs(g.hfc - gb.hfc)) <= 1e-11 * s
    assert np.max(np.abs(g.field - gb.field)) <= 1e-11 * s
    assert abs(g.Phi_S - gb.Phi_S) <= 1e-12 * abs(gb.Phi_S)


def test_gradient_on_the_full_size_system_matches_fd(ctx):
    # the N = 36864 system of the gradient workload (2x14N + 6x1H | 4x1H,
    # anisotropic, tilted field), 8 trace vectors over 20 steps
    w = workloads.anisotropic_36864(M=8, m=24, T_ns=20.0)
    g = gradients.yield_gradient(w.sys, w.sim, B=w.B, ctx=ctx)
    picks = [("hfc", (0, 0, 0)), ("hfc", (0, 1, 2)), ("hfc", (9, 2, 2)), ("field", 0), ("field", 2), ("J", None),
             ("kS", None), ("kT", None)]
    adj, fd = [], []
    for name, idx in picks:
        h = 1e-6 * getattr(w.sys, name) if name in ("kS", "kT") else 1e-6
        sp, Bp = _perturbed(w.sys, w.B, name, idx, h)
        sm, Bm = _perturbed(w.sys, w.B, name, idx, -h)
        fd.append((_forward_yield(sp, w.sim, Bp, ctx) - _forward_yield(sm, w.sim, Bm, ctx)) / (2 * h))
        adj.append(_adj_value(g, name, idx))
    adj, fd = np.array(adj), np.array(fd)
    assert np.max(np.abs(adj - fd)) <= 1e-6 * np.max(np.abs(fd)), (adj, fd)


def test_gradient_workload_shape_with_recomputed_segments(ctx, monkeypatch):
    # the differentiation workload's system and width (N = 36864, M = 64,
    # checkpoints every 64 steps) over 192 steps, with only the last 64 steps
    # kept resident: the two earlier segments are recomputed from checkpoints
    w = workloads.anisotropic_36864(M=64, m=24, T_ns=191.5)
    monkeypatch.setenv("SPC_ADJOINT_KEEP", "64")
    g = gradients.yield_gradient(w.sys, w.sim, B=w.B, ctx=ctx, checkpoint_every=64)
    monkeypatch.delenv("SPC_ADJOINT_KEEP")
    picks = [("hfc", (0, 0, 0)), ("hfc", (1, 0, 2)), ("hfc", (7, 2, 1)), ("hfc", (11, 1, 1)), ("field", 0),
             ("field", 2), ("J", None), ("kS", None), ("kT", None)]
    adj, fd = [], []
    for name, idx in picks:
        h = 1e-6 * getattr(w.sys, name) if name in ("kS", "kT") else 1e-6
        sp, Bp = _perturbed(w.sys, w.B, name, idx, h)
        sm, Bm = _perturbed(w.sys, w.B, name, idx, -h)
        fd.append((_forward_yield(sp, w.sim, Bp, ctx) - _forward_yield(sm, w.sim, Bm, ctx)) / (2 * h))
        adj.append(_adj_value(g, name, idx))
    adj, fd = np.array(adj), np.array(fd)
    assert abs(g.Phi_S - _forward_yield(w.sys, w.sim, w.B, ctx)) <= 1e-12 * abs(g.Phi_S)
    assert np.max(np.abs(adj - fd)) <= 1e-6 * np.max(np.abs(fd)), (adj, fd)
