"""GPU derivative of the semiclassical yield with respect to the field
(spc_semiclassical_tangent, the torch Function over (|B|, theta, phi)).

Every sample carries the forward-mode tangent of its classical state, so the
returned derivative is that of the sample average at fixed samples.  Pins:
central finite differences of the device forward (same seed, so the same
samples) along the field direction, a transverse axis and a random vector; the
same against the oracle's forward (orc_semiclassical) on a small case, tying
the derivative to the reference model; an invariance (no nuclei, kS = kT:
P_S(t) = exp(-kt) P_S(0) for every field, so the derivative vanishes); the
forward values equal to spc_semiclassical_run's; and the torch chain rule to
(|B|, theta, phi).  Tolerance 1e-6 relative to the largest derivative
(finite-difference step 1e-6 of the field).  Runs stay short: over long runs
the classical trajectories of this chaotic dynamics separate and pathwise
derivatives grow."""
import numpy as np
import pytest

from helpers import max_rel
from oracle import pyoracle
from spin_dynamics_amd import sc_dynamics, workloads

pytestmark = pytest.mark.gpu

SMALL = ([0.9, -0.4, 0.3], [0.7, 0.2], [2, 3, 2], [2, 2])
KW = dict(J=0.05, kS=0.2, kT=0.05, dt=0.05)


def _unit(v):
    v = np.asarray(v, float)
    return v / np.linalg.norm(v)


def _fd(fwd, Bvec, dfield, h=1e-6):
    """Central difference of fwd(field vector) along dfield."""
    Bvec, dfield = np.asarray(Bvec, float), np.asarray(dfield, float)
    return [(a - b) / (2 * h) for a, b in zip(fwd(Bvec + h * dfield), fwd(Bvec - h * dfield))]


def _device_fwd(args, ctx, **kw):
    def f(vec):
        B = float(np.linalg.norm(vec))
        return sc_dynamics.run(*args, B=B, direction=tuple(vec / B), ctx=ctx, **kw)
    return f


@pytest.mark.parametrize("which", ["along_field", "transverse", "random"])
def test_tangent_matches_device_finite_differences(ctx, which):
    d = _unit([0.3, 0.2, 0.9])
    B = 0.5
    dfield = {"along_field": d, "transverse": _unit(np.cross(d, [1.0, 0, 0])),
              "random": np.array([0.2, -0.7, 0.1])}[which]
    kw = dict(KW, n_samples=2048, n_steps_=200, seed=(42, 99))
    (ps, pt, phis, phit), (dps, dpt, dphis, dphit) = sc_dynamics.run_tangent(*SMALL, B=B, dfield=dfield,
                                                                            direction=tuple(d), ctx=ctx, **kw)
    fps, fpt, fphis, fphit = _fd(_device_fwd(SMALL, ctx, **kw), B * d, dfield)
    for got, ref in ((dps, fps), (dpt, fpt), (dphis, fphis), (dphit, fphit)):
        assert np.max(np.abs(got - ref)) <= 1e-6 * np.max(np.abs(ref)), which
    # the forward values are those of the plain run (same samples; the lane
    # split, hence the summation order, may differ)
    rps, rpt, rphis, _ = sc_dynamics.run(*SMALL, B=B, direction=tuple(d), ctx=ctx, **kw)
    assert max_rel(ps, rps) < 1e-12 and max_rel(pt, rpt) < 1e-12 and max_rel(phis, rphis) < 1e-12


def test_tangent_matches_oracle_finite_differences(ctx):
    d = _unit([0.1, -0.5, 0.8])
    B, dfield = 0.4, np.array([0.3, 0.1, -0.6])
    kw = dict(KW, n_samples=64, n_steps_=40, seed=(7, 8))
    (_, _, phis, _), (dps, dpt, dphis, _) = sc_dynamics.run_tangent(*SMALL, B=B, dfield=dfield, direction=tuple(d),
                                                                   ctx=ctx, **kw)

    def orc(vec):
        Bn = float(np.linalg.norm(vec))
        return pyoracle.semiclassical(*SMALL, B=Bn, direction=tuple(vec / Bn), with_pt=True, **kw)

    fps, fpt = _fd(orc, B * d, dfield)
    assert np.max(np.abs(dps - fps)) <= 1e-6 * np.max(np.abs(fps))
    assert np.max(np.abs(dpt - fpt)) <= 1e-6 * np.max(np.abs(fpt))


def test_large_nuclear_set_uses_the_wide_lane_split(ctx):
    # 24 + 24 nuclei (the 2^20-sample workload's layout): 4 lanes per sample, 12 slots
    # each, electron 2 from slot 6 on (the balanced layout the 12-slot tangent needs)
    w = workloads.semiclassical_48(n_samples=1024)
    d = _unit([0.2, 0.4, 0.9])
    kw = dict(n_samples=1024, dt=w.dt, n_steps_=120, J=0.0, kS=w.kS, kT=w.kT, seed=(3, 4))
    args = (w.a1, w.a2, w.m1, w.m2)
    (_, _, _, _), (dps, _, dphis, _) = sc_dynamics.run_tangent(*args, B=w.B, dfield=d, direction=tuple(d), ctx=ctx,
                                                              **kw)
    fps, _, fphis, _ = _fd(_device_fwd(args, ctx, **kw), w.B * d, d)
    assert np.max(np.abs(dps - fps)) <= 1e-6 * np.max(np.abs(fps))
    assert np.max(np.abs(dphis - fphis)) <= 1e-6 * np.max(np.abs(fphis))


def test_field_independent_case_has_zero_derivative(ctx):
    # no nuclei, kS = kT: both electrons precess together about the field and
    # tr T12 is invariant, so P_S(t) = exp(-kt) P_S(0) whatever the field
    (ps, _, _, _), (dps, dpt, _, _) = sc_dynamics.run_tangent([], [], [], [], B=0.7, dfield=[0.3, -0.2, 0.5],
                                                             direction=(0.0, 0.6, 0.8), J=0.0, kS=0.3, kT=0.3,
                                                             n_samples=512, dt=0.05, n_steps_=100, ctx=ctx)
    assert np.max(np.abs(dps)) < 1e-13 * np.max(np.abs(ps))
    assert np.max(np.abs(dpt)) < 1e-13 * np.max(np.abs(ps))


def test_torch_function_chain_rule(ctx):
    torch = pytest.importorskip("torch")
    prob = sc_dynamics.ScYieldProblem(*SMALL, n_samples=1024, dt=0.05, n_steps_=120, J=0.05, kS=0.2, kT=0.05,
                                      ctx=ctx)
    vals = dict(B=0.45, theta=0.7, phi=0.3)
    t = {k: torch.tensor(v, dtype=torch.float64, requires_grad=True) for k, v in vals.items()}
    F = sc_dynamics.semiclassical_yield_function().apply(t["B"], t["theta"], t["phi"], prob)
    F.backward()
    assert abs(F.item() - prob.value(**vals)) <= 1e-12 * abs(F.item())
    for k in vals:
        h = 1e-6
        up, dn = dict(vals), dict(vals)
        up[k] += h
        dn[k] -= h
        fd = (prob.value(**up) - prob.value(**dn)) / (2 * h)
        assert abs(t[k].grad.item() - fd) <= 1e-6 * max(abs(fd), 1e-3), k


@pytest.mark.parametrize("lanes", ["2", "4", "8"])
def test_every_lane_split_gives_the_same_derivative(ctx, monkeypatch, lanes):
    monkeypatch.setenv("SPC_SC_LANES", lanes)
    a1, a2 = [0.8, -0.3, 0.5, 0.2, -0.9], [0.6, 0.1, -0.4]
    kw = dict(J=0.03, kS=0.1, kT=0.2, dt=0.05, n_samples=700, n_steps_=60, seed=(5, 6))
    d = _unit([0.2, 0.5, 0.8])
    args = (a1, a2, [2, 3, 2, 2, 3], [2, 2, 3])
    (_, _, _, _), (dps, _, _, _) = sc_dynamics.run_tangent(*args, B=0.3, dfield=d, direction=tuple(d), ctx=ctx, **kw)
    fps, _, _, _ = _fd(_device_fwd(args, ctx, **kw), 0.3 * d, d)
    assert np.max(np.abs(dps - fps)) <= 1e-6 * np.max(np.abs(fps))
