"""GPU parity of the semiclassical kernel (spc_semiclassical_run) against the
oracle (orc_semiclassical) and the model's closed forms.

The device and the oracle draw the same samples (same seeds, same xoroshiro
stream) and integrate them with the same steps; they differ in rounding
(block-diagonal generator product, field sums split over a lane group,
device sincos/sqrt).  Over short runs that stays below 1e-10 relative; over
long runs the trajectories of this chaotic dynamics separate, and the two
sample means then agree within Monte Carlo error."""
import math

import numpy as np
import pytest

from helpers import max_rel
from oracle import pyoracle
from spin_dynamics_amd import _lib, sc_dynamics, workloads

pytestmark = pytest.mark.gpu


def _both(a1, a2, m1, m2, ctx, **kw):
    kw.setdefault("seed", (42, 99))
    ps, pt, phis, phit = sc_dynamics.run(a1, a2, m1, m2, ctx=ctx, **kw)
    ops, opt = pyoracle.semiclassical(a1, a2, m1, m2, with_pt=True, **kw)
    return (ps, pt, phis, phit), (ops, opt)


SMALL = ([0.9, -0.4, 0.3], [0.7, 0.2], [2, 3, 2], [2, 2])


@pytest.mark.parametrize("n_samples", [1, 127, 128, 1000, 4096])
def test_short_run_matches_oracle(ctx, n_samples):
    (ps, pt, phis, phit), (ops, opt) = _both(*SMALL, ctx, B=0.5, direction=(0.3, 0.2, 0.9), J=0.05, kS=0.2,
                                              kT=0.05, n_samples=n_samples, dt=0.05, n_steps_=30)
    assert max_rel(ps, ops) < 1e-10
    assert max_rel(pt, opt) < 1e-10
    _, rphis, rphit = pyoracle.get_kinetics(ops, opt, np.zeros_like(ops), np.zeros_like(ops), 0.05, 0.2, 0.05)
    assert max_rel(phis, rphis) < 1e-10 and max_rel(phit, rphit) < 1e-10


@pytest.mark.parametrize("n1,n2", [(0, 0), (1, 0), (0, 3), (4, 4), (7, 2), (15, 1), (12, 12), (24, 24), (40, 20),
                                   (63, 0), (63, 1), (33, 31)])
def test_every_slot_layout_matches_oracle(ctx, n1, n2):
    rng = np.random.default_rng(n1 * 100 + n2)
    a1 = list(rng.uniform(-1, 1, n1))
    a2 = list(rng.uniform(-1, 1, n2))
    m1 = list(rng.choice([2, 3], n1))
    m2 = list(rng.choice([2, 3], n2))
    (ps, pt, _, _), (ops, opt) = _both(a1, a2, m1, m2, ctx, B=0.3, J=0.02, kS=0.1, kT=0.3, n_samples=300,
                                      dt=0.04, n_steps_=12)
    assert max_rel(ps, ops) < 1e-10
    assert max_rel(pt, opt) < 1e-10


@pytest.mark.parametrize("lanes", ["2", "4", "8"])
def test_every_lane_split_matches_oracle(ctx, monkeypatch, lanes):
    monkeypatch.setenv("SPC_SC_LANES", lanes)
    a1, a2 = [0.8, -0.3, 0.5, 0.2, -0.9], [0.6, 0.1, -0.4]
    (ps, pt, _, _), (ops, opt) = _both(a1, a2, [2, 3, 2, 2, 3], [2, 2, 3], ctx, B=0.2, J=0.03, kS=0.1, kT=0.2,
                                      n_samples=700, dt=0.05, n_steps_=25)
    assert max_rel(ps, ops) < 1e-10
    assert max_rel(pt, opt) < 1e-10


def test_large_system_short_run_matches_oracle(ctx):
    w = workloads.semiclassical_48(n_samples=2048)
    kw = {**w.kwargs(), "n_steps_": 20}
    (ps, pt, _, _), (ops, opt) = _both(w.a1, w.a2, w.m1, w.m2, ctx, **kw)
    assert max_rel(ps, ops) < 1e-10
    assert max_rel(pt, opt) < 1e-10


def test_long_run_agrees_within_monte_carlo_error(ctx):
    kw = dict(B=0.05, J=0.0, kS=0.01, kT=0.003, n_samples=20000, dt=0.08, n_steps_=800)
    (ps, pt, _, _), (ops, opt) = _both(*SMALL, ctx, **kw)
    assert np.max(np.abs(ps - ops)) < 0.05
    assert np.max(np.abs(pt - opt)) < 0.05


def test_uniform_decay_closed_form(ctx):
    k, dt, n = 0.3, 0.05, 200
    ps, pt, _, _ = sc_dynamics.run([], [], [], [], B=1.7, direction=(1, 1, 1), kS=k, kT=k, n_samples=3000, dt=dt,
                                   n_steps_=n, ctx=ctx)
    t = dt * np.arange(n + 1)
    assert np.max(np.abs(ps - ps[0] * np.exp(-k * t))) < 1e-12
    ps, pt, _, _ = sc_dynamics.run(*SMALL, B=0.5, kS=k, kT=k, n_samples=3000, dt=dt, n_steps_=n, ctx=ctx)
    tot = ps + pt
    assert np.max(np.abs(tot - tot[0] * np.exp(-k * t))) < 1e-12


def test_deterministic_and_seeded(ctx):
    kw = dict(B=0.5, kS=0.2, kT=0.1, J=0.05, n_samples=5000, dt=0.05, n_steps_=50)
    a = sc_dynamics.run(*SMALL, ctx=ctx, **kw)
    b = sc_dynamics.run(*SMALL, ctx=ctx, **kw)
    c = sc_dynamics.run(*SMALL, ctx=ctx, seed=(7, 8), **kw)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    assert not np.array_equal(a[0], c[0])


def test_rejects_unsupported_layouts(ctx):
    with pytest.raises(_lib.SpinchemError):
        sc_dynamics.run([0.1] * 40, [0.1] * 25, [2] * 40, [2] * 25, B=0.1, n_samples=10, dt=0.1, n_steps_=2,
                        ctx=ctx)
    with pytest.raises(_lib.SpinchemError):
        sc_dynamics.run([0.1], [], [2], [], B=0.1, n_samples=0, dt=0.1, n_steps_=2, ctx=ctx)
    with pytest.raises(_lib.SpinchemError):
        sc_dynamics.run([0.1], [], [1], [], B=0.1, n_samples=4, dt=0.1, n_steps_=2, ctx=ctx)
