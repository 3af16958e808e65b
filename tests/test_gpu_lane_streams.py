"""GPU: trace sampling with the lanes split into chunks that run concurrently
on their own streams (SPC_LANE_STREAMS = 1..4).

Every split must give the populations of the oracle on the same samples
(1e-10 relative, the yield tolerance), and the splits must agree with one
another to rounding (1e-12): a chunk's lanes are the samples [start,
start + M), and the chunks are added to the sums in sample order.  The cases
leave ragged chunks (M = 65: one lane in the last chunk; M = 200 in four
chunks of 64, 64, 64 and 8 lanes) and cover each way vectors are made
(on-device phase vectors from the per-sample seeds, injected vectors, the
basis states of the exact trace)."""
import numpy as np
import pytest

from helpers import max_rel, oracle_system
from oracle import pyoracle
from spin_dynamics_amd import dynamics, workloads

pytestmark = pytest.mark.gpu

KEYS = ("P_S", "P_T0", "P_Tp", "P_Tm")


def _run(monkeypatch, streams, fn):
    monkeypatch.setenv("SPC_LANE_STREAMS", str(streams))
    return fn()


def test_sampled_vectors_every_split(ctx, monkeypatch):
    w = workloads.small_anisotropic(M=200, T_ns=120.0)
    s = oracle_system(w.sys, w.B)
    ref = pyoracle.trace_sampling(pyoracle.build_general_csr(s), s.Z, M=200, m=w.sim.N_krylov, t_end=w.sim.t_end,
                                  dt=w.sim.dt, vec_kind=1, seed=w.sim.seed)
    runs = {n: _run(monkeypatch, n, lambda: dynamics.trace_sampling_para(w.sys, w.sim, B=w.B, ctx=ctx))
            for n in (1, 2, 3, 4)}
    for n, res in runs.items():
        for k in KEYS:
            assert max_rel(getattr(res, k), ref[k]) < 1e-10, (n, k)
            assert max_rel(getattr(res, k), getattr(runs[1], k)) < 1e-12, (n, k)
        assert abs(res.Phi_S[-1] - runs[1].Phi_S[-1]) <= 1e-12 * abs(runs[1].Phi_S[-1])
    # the same four chunks on one stream, one kernel at a time: bit for bit the
    # concurrent result (the chunks share nothing but the read-only tables)
    monkeypatch.setenv("SPC_LANE_SERIAL", "1")
    serial = _run(monkeypatch, 4, lambda: dynamics.trace_sampling_para(w.sys, w.sim, B=w.B, ctx=ctx))
    for k in KEYS:
        assert np.array_equal(getattr(serial, k), getattr(runs[4], k)), k


@pytest.mark.parametrize("M", [65, 130])
def test_injected_vectors_ragged_chunks(ctx, monkeypatch, M):
    w = workloads.small_anisotropic(m1=(3, 2), m2=(2, 2), M=M, T_ns=80.0)
    s = oracle_system(w.sys, w.B)
    vecs = np.array([pyoracle.sample_vector(sd, s.Z, 1) for sd in pyoracle.sample_seeds((13, 17), M)])
    ref = pyoracle.trace_sampling(pyoracle.build_general_csr(s), s.Z, M=M, m=w.sim.N_krylov, t_end=w.sim.t_end,
                                  dt=w.sim.dt, vec_kind=2, injected=vecs)
    for n in (2, 3):
        res = _run(monkeypatch, n, lambda: dynamics.trace_sampling_para(w.sys, w.sim, B=w.B, trace_vectors=vecs,
                                                                        ctx=ctx))
        for k in KEYS:
            assert max_rel(getattr(res, k), ref[k]) < 1e-10, (n, k)


def test_exact_trace_split_over_basis_states(ctx, monkeypatch):
    w = workloads.small_anisotropic(m1=(3, 3, 3), m2=(2, 3), T_ns=60.0)  # Z = 162: chunks of 64, 64, 34
    s = oracle_system(w.sys, w.B)
    ref = pyoracle.exact_dynamics(pyoracle.build_general_csr(s), s.Z, m=w.sim.N_krylov, t_end=w.sim.t_end,
                                  dt=w.sim.dt)
    for n in (1, 4):
        res = _run(monkeypatch, n, lambda: dynamics.exact_dynamics(w.sys, w.sim, B=w.B, ctx=ctx))
        assert max_rel(res.P_S, ref["P_S"]) < 1e-10, n
