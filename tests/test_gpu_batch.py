"""GPU batched field orientations (spc_orientation_batch) and the autotuner.

Pins: every orientation's P_S against the oracle run at that orientation with
the same trace vectors (orc_trace_sampling, 1e-10 relative); against the
plain per-orientation device runs; the captured-graph replay equal to the
eager launches, bit for bit; results independent of how orientations are
chunked into lanes and of how many chunks run at once; the autotuned Krylov order keeps the yields within the
requested tolerance."""
import numpy as np
import pytest

from helpers import max_rel, oracle_system
from oracle import pyoracle
from spin_dynamics_amd import _lib, batch, dynamics, gradients, workloads

pytestmark = pytest.mark.gpu


def _case(M=64, T_ns=30.0, m=14):
    w = workloads.small_anisotropic(m1=(3, 2), m2=(2, 2), M=M, m=m, T_ns=T_ns)
    return w


def test_each_orientation_matches_oracle(ctx):
    w = _case()
    dirs = batch.fibonacci_sphere(6)
    phi, ps = batch.orientation_yields(w.sys, w.sim, dirs, B=w.B, orient_per_chunk=4, with_populations=True,
                                       ctx=ctx)
    Z = w.sys.Z1 * w.sys.Z2
    for o, d in enumerate(dirs):
        s = oracle_system(w.sys, w.B)
        s.direction = tuple(d)
        h = pyoracle.build_general_csr(s)
        r = pyoracle.trace_sampling(h, Z, M=w.sim.N_samples, m=w.sim.N_krylov, t_end=w.sim.t_end, dt=w.sim.dt,
                                    vec_kind=1, seed=w.sim.seed)
        assert max_rel(ps[o], r["P_S"]) < 1e-10
        _, phis, _ = pyoracle.get_kinetics(r["P_S"], r["P_T0"], r["P_Tp"], r["P_Tm"], w.sim.dt, w.sys.kS, w.sys.kT)
        assert abs(phi[o] - phis[-1]) <= 1e-10 * abs(phis[-1])


def test_matches_plain_per_orientation_runs(ctx):
    import copy
    w = _case(T_ns=25.0)
    dirs = batch.fibonacci_sphere(5)
    phi = batch.orientation_yields(w.sys, w.sim, dirs, B=w.B, ctx=ctx)
    for o, d in enumerate(dirs):
        s = copy.deepcopy(w.sys)
        s.field_dir = tuple(d)
        res = dynamics.trace_sampling_para(s, w.sim, B=w.B, ctx=ctx)
        res.get_kinetics(w.sim.dt, s.kS, s.kT)
        assert abs(phi[o] - res.Phi_S[-1]) <= 1e-12 * abs(res.Phi_S[-1])


def test_graph_replay_equals_eager_launches(ctx):
    w = _case(T_ns=20.0)
    dirs = batch.fibonacci_sphere(3)
    a, pa = batch.orientation_yields(w.sys, w.sim, dirs, B=w.B, graph=True, with_populations=True, ctx=ctx)
    b, pb = batch.orientation_yields(w.sys, w.sim, dirs, B=w.B, graph=False, with_populations=True, ctx=ctx)
    c, pc = batch.orientation_yields(w.sys, w.sim, dirs, B=w.B, graph=True, with_populations=True, ctx=ctx)
    assert np.array_equal(pa, pb) and np.array_equal(pa, pc)


@pytest.mark.parametrize("opc", [1, 2, 3, 7])
def test_chunking_does_not_change_results(ctx, opc):
    w = _case(T_ns=15.0)
    dirs = batch.fibonacci_sphere(7)
    ref = batch.orientation_yields(w.sys, w.sim, dirs, B=w.B, orient_per_chunk=7, ctx=ctx)
    got = batch.orientation_yields(w.sys, w.sim, dirs, B=w.B, orient_per_chunk=opc, ctx=ctx)
    assert np.max(np.abs(got - ref)) <= 1e-12 * np.max(np.abs(ref))


def test_chunks_in_flight_do_not_change_results(ctx, monkeypatch):
    # 5 chunks of one orientation, then 7 orientations in chunks of 2 (the last
    # one narrower): with 2 or 4 streams a slot takes further chunks, replaying
    # its step graph between their eager set-up work; every chunk writes only
    # its own sums, so each result is bit for bit that of eager launches, one
    # chunk at a time
    w = _case(T_ns=12.0)
    for n_dirs, opc in ((5, 1), (7, 2)):
        dirs = batch.fibonacci_sphere(n_dirs)
        monkeypatch.setenv("SPC_LANE_STREAMS", "1")
        ref = batch.orientation_yields(w.sys, w.sim, dirs, B=w.B, orient_per_chunk=opc, graph=False,
                                       with_populations=True, ctx=ctx)
        for n in (1, 2, 4):
            monkeypatch.setenv("SPC_LANE_STREAMS", str(n))
            got = batch.orientation_yields(w.sys, w.sim, dirs, B=w.B, orient_per_chunk=opc, graph=True,
                                           with_populations=True, ctx=ctx)
            assert np.array_equal(got[0], ref[0]) and np.array_equal(got[1], ref[1]), (n_dirs, opc, n)


def test_more_trace_vectors_and_su_z(ctx):
    w = _case(M=128, T_ns=12.0)
    w.sim.vec_kind = "SUZ"
    dirs = batch.fibonacci_sphere(3)
    phi = batch.orientation_yields(w.sys, w.sim, dirs, B=w.B, ctx=ctx)
    import copy
    for o, d in enumerate(dirs):
        s = copy.deepcopy(w.sys)
        s.field_dir = tuple(d)
        res = dynamics.trace_sampling_para(s, w.sim, B=w.B, ctx=ctx)
        res.get_kinetics(w.sim.dt, s.kS, s.kT)
        assert abs(phi[o] - res.Phi_S[-1]) <= 1e-12 * abs(res.Phi_S[-1])


def test_autotuner_keeps_the_yield(ctx):
    w = _case(T_ns=40.0, m=20)
    dirs = batch.fibonacci_sphere(8)
    opc, m, ms = batch.tune(w.sys, w.sim, dirs, B=w.B, yield_tol=1e-10, ctx=ctx)
    assert 1 <= opc <= 8 and (opc & (opc - 1)) == 0
    assert 6 <= m <= 20 and ms > 0
    ref = batch.orientation_yields(w.sys, w.sim, dirs, B=w.B, ctx=ctx)
    w.sim.N_krylov = m
    got = batch.orientation_yields(w.sys, w.sim, dirs, B=w.B, orient_per_chunk=opc, ctx=ctx)
    assert np.max(np.abs(got - ref) / np.abs(ref)) <= 1e-10


def test_rejects_unsupported_requests(ctx):
    w = _case(M=32)
    with pytest.raises(_lib.SpinchemError):
        batch.orientation_yields(w.sys, w.sim, batch.fibonacci_sphere(2), B=w.B, ctx=ctx)
    w = _case()
    with pytest.raises(_lib.SpinchemError):
        batch.orientation_yields(w.sys, w.sim, np.zeros((2, 3)), B=w.B, ctx=ctx)


def test_tiled_product_with_per_orientation_tables(ctx):
    # five trailing spin-1/2 nuclei: the product runs as 8-wave tiles, each
    # 64-lane chunk with its own orientation's Kronecker tables
    w = workloads.small_anisotropic(m1=(3,), m2=(2, 2, 2, 2, 2), M=64, m=12, T_ns=12.0)
    dirs = batch.fibonacci_sphere(3)
    phi, ps = batch.orientation_yields(w.sys, w.sim, dirs, B=w.B, orient_per_chunk=3, with_populations=True,
                                       ctx=ctx)
    Z = w.sys.Z1 * w.sys.Z2
    for o, d in enumerate(dirs):
        s = oracle_system(w.sys, w.B)
        s.direction = tuple(d)
        r = pyoracle.trace_sampling(pyoracle.build_general_csr(s), Z, M=w.sim.N_samples, m=w.sim.N_krylov,
                                    t_end=w.sim.t_end, dt=w.sim.dt, vec_kind=1, seed=w.sim.seed)
        assert max_rel(ps[o], r["P_S"]) < 1e-10


@pytest.mark.parametrize("N,T_ns,n_dirs", [(768, 9999.5, 4), (3072, 9999.5, 4), (12288, 999.5, 4), (49152, 999.5, 2),
                                           (147456, 599.5, 2)])
def test_tuned_order_holds_the_yield_over_long_runs(ctx, N, T_ns, n_dirs):
    # the tuner's order (chosen from 64-step pilots for the 10 us run) against
    # the requested m = 24 on the first pair of each dimension class: the full
    # 10 us for the small classes, 600-1000 steps for the large ones; the
    # yields agree far inside the 1e-10 tolerance (measured on MI355X with four
    # orientations: 6e-13, 9e-14, 1e-12, 4e-13, 2.8e-12)
    import copy
    p = [q for q in batch.radical_pairs() if q.N == N][0]
    dirs = batch.fibonacci_sphere(96)
    _, m, _ = batch.tune(p.sys, batch.orientation_sim(), dirs, B=0.05, ctx=ctx)
    assert m < 24
    sel = dirs[np.linspace(0, 95, n_dirs).astype(int)]
    sim = batch.orientation_sim(T_ns=T_ns)
    ref = batch.orientation_yields(p.sys, sim, sel, B=0.05, ctx=ctx)
    s2 = copy.deepcopy(sim)
    s2.N_krylov = m
    got = batch.orientation_yields(p.sys, s2, sel, B=0.05, ctx=ctx)
    assert np.max(np.abs(got - ref) / np.abs(ref)) <= 1e-10
