"""GPU: batched orientations with several steps per Krylov basis (tol > 0:
the reference's rebuild rule, qm_dynamics.f90:348-364, for every lane), run
in rounds (a basis per lane from its psi, then each lane's steps until its
rule asks for a new one, populations from projector Grams), and the adaptive
autotuner that picks the order and the tolerance.

Pins: each orientation against the oracle at that orientation with the same
trace vectors and tol (the reference's algorithm, MGS bases); against the
plain per-orientation device runs at the same tol; the captured round
replayed from its graph equal to eager launches and results independent of
the chunks in flight, bit for bit, and of the orientations per chunk; the
tuned (order, tol) against the requested order with one step per basis over
the full 10 us for the small dimension classes and 600-1000 steps for the
large ones (1e-10 relative, the yield tolerance)."""
import copy

import numpy as np
import pytest

from helpers import max_rel, oracle_system
from oracle import pyoracle
from spin_dynamics_amd import batch, dynamics, workloads

pytestmark = pytest.mark.gpu


def _case(M=64, T_ns=30.0, m=14, tol=1e-8):
    w = workloads.small_anisotropic(m1=(3, 2), m2=(2, 2), M=M, m=m, T_ns=T_ns)
    w.sim.tol = tol
    return w


def test_rounds_each_orientation_matches_oracle(ctx):
    w = _case(T_ns=60.0)
    dirs = batch.fibonacci_sphere(5)
    phi, ps = batch.orientation_yields(w.sys, w.sim, dirs, B=w.B, orient_per_chunk=4, with_populations=True,
                                       ctx=ctx)
    _, rounds, steps = dynamics.last_run_stats(ctx)
    assert rounds < 2 * steps  # two chunks, several steps per round
    Z = w.sys.Z1 * w.sys.Z2
    for o, d in enumerate(dirs):
        s = oracle_system(w.sys, w.B)
        s.direction = tuple(d)
        r = pyoracle.trace_sampling(pyoracle.build_general_csr(s), Z, M=w.sim.N_samples, m=w.sim.N_krylov,
                                    t_end=w.sim.t_end, dt=w.sim.dt, tol=w.sim.tol, vec_kind=1, seed=w.sim.seed)
        assert max_rel(ps[o], r["P_S"]) < 1e-9
        _, phis, _ = pyoracle.get_kinetics(r["P_S"], r["P_T0"], r["P_Tp"], r["P_Tm"], w.sim.dt, w.sys.kS, w.sys.kT)
        assert abs(phi[o] - phis[-1]) <= 1e-9 * abs(phis[-1])


def test_rounds_match_plain_runs(ctx):
    w = _case(T_ns=80.0, m=20, tol=1e-12)
    dirs = batch.fibonacci_sphere(3)
    phi, ps = batch.orientation_yields(w.sys, w.sim, dirs, B=w.B, with_populations=True, ctx=ctx)
    for o, d in enumerate(dirs):
        s = copy.deepcopy(w.sys)
        s.field_dir = tuple(d)
        res = dynamics.trace_sampling_para(s, w.sim, B=w.B, ctx=ctx)
        assert max_rel(ps[o], res.P_S) < 1e-12
        res.get_kinetics(w.sim.dt, s.kS, s.kT)
        assert abs(phi[o] - res.Phi_S[-1]) <= 1e-12 * abs(res.Phi_S[-1])


def test_rounds_graph_streams_and_chunking(ctx, monkeypatch):
    w = _case(T_ns=40.0)
    dirs = batch.fibonacci_sphere(7)
    monkeypatch.setenv("SPC_LANE_STREAMS", "1")
    ref = batch.orientation_yields(w.sys, w.sim, dirs, B=w.B, orient_per_chunk=2, graph=False,
                                   with_populations=True, ctx=ctx)
    for n in (1, 2, 4):
        monkeypatch.setenv("SPC_LANE_STREAMS", str(n))
        got = batch.orientation_yields(w.sys, w.sim, dirs, B=w.B, orient_per_chunk=2, graph=True,
                                       with_populations=True, ctx=ctx)
        assert np.array_equal(got[0], ref[0]) and np.array_equal(got[1], ref[1]), n
    monkeypatch.delenv("SPC_LANE_STREAMS")
    for opc in (1, 3, 7):
        got = batch.orientation_yields(w.sys, w.sim, dirs, B=w.B, orient_per_chunk=opc, ctx=ctx)
        assert np.max(np.abs(got - ref[0])) <= 1e-12 * np.max(np.abs(ref[0])), opc


@pytest.mark.parametrize("N,T_ns,n_dirs", [(768, 9999.5, 4), (3072, 9999.5, 4), (12288, 999.5, 4), (49152, 999.5, 2),
                                           (147456, 599.5, 2)])
def test_adaptive_tuning_holds_the_yield_over_long_runs(ctx, N, T_ns, n_dirs):
    # the tuner's (order, tol), chosen from 256-step pilots for the 10 us run,
    # against the requested m = 24 with one step per basis on the first pair
    # of each dimension class
    p = [q for q in batch.radical_pairs() if q.N == N][0]
    dirs = batch.fibonacci_sphere(96)
    opc, m, tol, _ = batch.tune_adaptive(p.sys, batch.orientation_sim(), dirs, B=0.05, ctx=ctx)
    assert tol > 0.0 and 16 <= m <= 28
    sel = dirs[np.linspace(0, 95, n_dirs).astype(int)]
    sim = batch.orientation_sim(T_ns=T_ns)
    ref = batch.orientation_yields(p.sys, sim, sel, B=0.05, ctx=ctx)
    s2 = copy.deepcopy(sim)
    s2.N_krylov = m
    s2.tol = tol
    got = batch.orientation_yields(p.sys, s2, sel, B=0.05, orient_per_chunk=min(opc, n_dirs), ctx=ctx)
    err = np.max(np.abs(got - ref) / np.abs(ref))
    print(f"N={N}: m={m} tol={tol:.0e} opc={opc}: max rel yield difference {err:.2e} over {T_ns + 0.5:.0f} ns")
    assert err <= 1e-10
