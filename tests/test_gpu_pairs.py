"""GPU: radical pairs of uneven dimension at many field orientations in one
call (spc_pairs_batch): every pair's system resident at once, the chunks of
all pairs sharing the streams.

Pins: each pair's yields equal its own orientation batch (spc_orientation_batch
with the same order, tolerance and trace vectors) at one step per basis and in
rounds (tol > 0), with different orders per pair; a pair that needs the
column form (a strong field) runs again alone and still equals its own batch;
the context's own system is untouched by the call."""
import copy

import numpy as np
import pytest

from spin_dynamics_amd import batch, dynamics, workloads

pytestmark = pytest.mark.gpu


def _pairs():
    ps = []
    for N in (3072, 768, 12288):
        ps.append([q for q in batch.radical_pairs() if q.N == N][0].sys)
    w = workloads.small_anisotropic(m1=(3, 2), m2=(2, 2), M=64)
    ps.append(w.sys)  # N = 96, B = 0.05 mT
    return ps


@pytest.mark.parametrize("mode", ["tol0", "rounds", "mixed"])
def test_pairs_batch_equals_each_pairs_batch(ctx, mode):
    systems = _pairs()
    dirs = batch.fibonacci_sphere(5)
    sim = batch.orientation_sim(m=20, T_ns=40.5)
    if mode == "tol0":
        orders, tols = [20, 16, 20, 12], [0.0, 0.0, 0.0, 0.0]
    elif mode == "rounds":
        orders, tols = [24, 20, 28, 16], [1e-20, 1e-16, 1e-20, 1e-12]
    else:
        orders, tols = [24, 16, 20, 16], [1e-20, 0.0, 0.0, 1e-16]
    opcs = [2, 4, 1, 3]
    got = batch.pairs_yields(systems, sim, dirs, n_krylov=orders, tol=tols, orient_per_chunk=opcs, ctx=ctx)
    for p, s in enumerate(systems):
        sp = copy.deepcopy(sim)
        sp.N_krylov, sp.tol = orders[p], tols[p]
        ref = batch.orientation_yields(s, sp, dirs, orient_per_chunk=opcs[p], ctx=ctx)
        assert np.max(np.abs(got[p] - ref) / np.abs(ref)) <= 1e-12, (mode, p)


def test_pairs_batch_reruns_a_poorly_conditioned_pair(ctx):
    # the N = 384 pair at 25 mT builds a Newton block past its condition limit:
    # it runs again alone, column by column, as its own batch does
    strong = workloads.small_anisotropic(m1=(3, 2, 2), m2=(2, 2, 2), M=64, m=24)
    strong.sys.e1.w = strong.sys.e2.w = -25.0
    systems = [_pairs()[1], strong.sys]
    dirs = batch.fibonacci_sphere(3)
    sim = batch.orientation_sim(m=24, T_ns=20.5)
    got = batch.pairs_yields(systems, sim, dirs, ctx=ctx)
    for p, s in enumerate(systems):
        ref = batch.orientation_yields(s, sim, dirs, ctx=ctx)
        assert np.max(np.abs(got[p] - ref) / np.abs(ref)) <= 1e-12, p


def test_pairs_batch_keeps_the_context_system(ctx):
    w = workloads.small_anisotropic(M=8, T_ns=20.0)
    before = dynamics.trace_sampling_para(w.sys, w.sim, B=w.B, ctx=ctx)
    H = dynamics.Hamiltonian(w.sys, w.B, ctx)
    x = np.random.default_rng(2).normal(size=(3, H.N)) + 0j
    y0 = H.spmm(x)
    batch.pairs_yields(_pairs()[:2], batch.orientation_sim(T_ns=10.5), batch.fibonacci_sphere(2), ctx=ctx)
    assert np.array_equal(H.spmm(x), y0)  # the context's own H, untouched
    after = dynamics.trace_sampling_para(w.sys, w.sim, B=w.B, ctx=ctx)
    assert np.array_equal(before.P_S, after.P_S)
