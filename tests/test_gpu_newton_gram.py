"""GPU: the Newton block where a basis serves several steps (tol > 0, the
reference's rebuild rule, qm_dynamics.f90:348-364).

The chain's Gram matrix is taken as four projector bins Y^H P_X Y (singlet,
T0, T+, T- components of every nuclear index), whose sum is the Gram matrix
that gives the Cholesky factor; G_X = R^-H (Y^H P_X Y) R^-1 are then the
projector Grams of the orthonormal basis (convert_to_krylov,
qm_dynamics.f90:469-498) without forming it.  Pins: the column form (MGS-
equivalent, its Grams from the formed basis) and the monomial blocks on the
same inputs for orders 4..32; the oracle (the reference's algorithm with its
own tol rule) on small systems; serial against concurrent lane chunks (bit
for bit); the N = 36 864 system against its tol = 0 run.  Rebuild decisions
compare |c_m|^2 with tol |c|^2, and a decision that flips at rounding level
moves a yield by about the Krylov truncation error, so the bounds against
other bases are 1e-10 over a few steps and looser over long runs."""
import copy

import numpy as np
import pytest

from helpers import max_rel, oracle_system
from oracle import pyoracle
from spin_dynamics_amd import batch, dynamics, workloads

pytestmark = pytest.mark.gpu

KEYS = ("P_S", "P_T0", "P_Tp", "P_Tm")


def _pair(N):
    return [q for q in batch.radical_pairs() if q.N == N][0]


@pytest.mark.parametrize("m", [4, 8, 12, 20, 24, 32])
def test_newton_grams_match_the_column_form(ctx, monkeypatch, m):
    p = _pair(3072)
    sim = batch.orientation_sim(m=m, T_ns=7.5)
    sim.tol = 1e-8
    s = copy.deepcopy(p.sys)
    s.field_dir = (-0.2, 0.6, 0.77)
    monkeypatch.setenv("SPC_SSTEP", "0")
    col = dynamics.trace_sampling_para(s, sim, B=0.05, ctx=ctx)
    monkeypatch.delenv("SPC_SSTEP")
    nwt = dynamics.trace_sampling_para(s, sim, B=0.05, ctx=ctx)
    _, rebuilds, steps = dynamics.last_run_stats(ctx)
    monkeypatch.setenv("SPC_NEWTON", "0")
    mono = dynamics.trace_sampling_para(s, sim, B=0.05, ctx=ctx)
    for k in KEYS:
        assert max_rel(getattr(nwt, k), getattr(col, k)) < 1e-10, (m, k)
        assert max_rel(getattr(mono, k), getattr(col, k)) < 1e-10, (m, k)
    if m >= 12:  # a basis of order 12 or more serves several 1 ns steps here
        assert rebuilds < steps


@pytest.mark.parametrize("tol", [1e-5, 1e-8])
def test_newton_grams_vs_oracle(ctx, tol):
    # N = 96, m = 20 (the Gram's bins in several tiles); the reference input's tol
    w = workloads.small_anisotropic(m1=(3, 2), m2=(2,), M=8, m=20, T_ns=120.0, tol=tol)
    s = oracle_system(w.sys, w.B)
    vecs = np.array([pyoracle.sample_vector(sd, s.Z, 1) for sd in pyoracle.sample_seeds((3, 9), 8)])
    res = dynamics.trace_sampling_para(w.sys, w.sim, B=w.B, trace_vectors=vecs, ctx=ctx)
    _, rebuilds, steps = dynamics.last_run_stats(ctx)
    ref = pyoracle.trace_sampling(pyoracle.build_general_csr(s), s.Z, M=8, m=w.sim.N_krylov, t_end=w.sim.t_end,
                                  dt=w.sim.dt, tol=tol, vec_kind=2, injected=vecs)
    for k in KEYS:
        assert max_rel(getattr(res, k), ref[k]) < 1e-9, k
    assert rebuilds < steps


def test_newton_grams_concurrent_chunks_bitwise(ctx, monkeypatch):
    # four 64-lane chunks on their own streams (rounds interleaved, one host
    # check per round) against the same chunks one kernel at a time
    p = _pair(3072)
    sim = batch.orientation_sim(m=24, T_ns=40.5, M=256)
    sim.tol = 1e-10
    s = copy.deepcopy(p.sys)
    con = dynamics.trace_sampling_para(s, sim, B=0.05, ctx=ctx)
    monkeypatch.setenv("SPC_LANE_SERIAL", "1")
    ser = dynamics.trace_sampling_para(s, sim, B=0.05, ctx=ctx)
    for k in KEYS:
        assert np.array_equal(getattr(con, k), getattr(ser, k)), k


@pytest.mark.parametrize("tol,T_ns,bound", [(1e-10, 99.5, 1e-6), (1e-20, 1999.5, 1e-10)])
def test_newton_grams_full_size(ctx, tol, T_ns, bound):
    # N = 36 864, M = 64, m = 24 against the same run at tol = 0: the rule's
    # truncation error at 1e-10 (about 1e-7 of the populations) and at 1e-20
    # (within the yield tolerance over 2000 steps)
    w = workloads.anisotropic_36864(M=64, m=24, T_ns=T_ns)
    ref = dynamics.trace_sampling_para(w.sys, w.sim, B=w.B, ctx=ctx)
    sim = copy.deepcopy(w.sim)
    sim.tol = tol
    res = dynamics.trace_sampling_para(w.sys, sim, B=w.B, ctx=ctx)
    _, rebuilds, steps = dynamics.last_run_stats(ctx)
    err = max(max_rel(getattr(res, k), getattr(ref, k)) for k in KEYS)
    derr = abs(res.Phi_S[-1] - ref.Phi_S[-1]) / abs(ref.Phi_S[-1])
    print(f"N=36864 M=64 tol={tol:.0e}: {rebuilds} rebuilds for {steps} steps, max rel population difference "
          f"{err:.2e}, yield {derr:.2e}")
    assert rebuilds < steps
    assert err < bound and derr < bound


def test_high_field_falls_back_to_the_column_form(ctx):
    # N = 384 at B = 25 mT: the Newton block's condition estimate is past its
    # limit, so the chunk runs column by column with the projector Grams of
    # the formed basis (the same binned pass over q_0 .. q_{m-1})
    w = workloads.small_anisotropic(m1=(3, 2, 2), m2=(2, 2, 2), M=8, m=24, T_ns=60.0, tol=1e-8)
    w.B = 25.0
    w.sys.e1.w = w.sys.e2.w = -25.0
    s = oracle_system(w.sys, w.B)
    vecs = np.array([pyoracle.sample_vector(sd, s.Z, 1) for sd in pyoracle.sample_seeds((4, 8), 8)])
    res = dynamics.trace_sampling_para(w.sys, w.sim, B=w.B, trace_vectors=vecs, ctx=ctx)
    ref = pyoracle.trace_sampling(pyoracle.build_general_csr(s), s.Z, M=8, m=w.sim.N_krylov, t_end=w.sim.t_end,
                                  dt=w.sim.dt, tol=1e-8, vec_kind=2, injected=vecs)
    for k in KEYS:
        assert max_rel(getattr(res, k), ref[k]) < 1e-9, k
