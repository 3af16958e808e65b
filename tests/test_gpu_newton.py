"""GPU: the Newton-basis block (a basis that serves one step, built from the
chain y_{j+1} = (H - sig_j) y_j with Leja-ordered Ritz values as shifts, one
Gram pass and a per-lane Cholesky factor; psi formed from the chain vectors
directly).

The chain spans the same Krylov spaces as the column form for any shifts, so
the yields must agree with the one-column-per-pass Arnoldi to rounding and
with the oracle to the yield tolerance.  Pins: orders 2..32 against the
column path on a 3 072-dimensional pair (single-tile and multi-tile Gram
passes, the largest order); the monomial blocks (SPC_NEWTON=0) against the
same; every lane chunk split; a random-phase and an injected start at
N = 36 864 against the oracle over two steps (test_gpu_sstep); the equal-rate
invariant (exact decay of every norm, 1e-12); the batched orientations and
the gradient's forward steps against their column runs; the 64-lane chain
product's compile-time electron layouts of the in-tile nuclei (every count
of electron-2 nuclei among them) against the oracle."""
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


@pytest.mark.parametrize("m", [2, 3, 4, 5, 8, 9, 12, 16, 20, 24, 29, 32])
def test_newton_block_matches_the_column_path(ctx, monkeypatch, m):
    # nd = m + 1 chain vectors: one Gram tile up to nd = 5, several from 6
    # (a partial last tile at 9, 13, 17, 21, 30, 33)
    p = _pair(3072)
    sim = batch.orientation_sim(m=m, T_ns=3.5)  # four steps
    s = copy.deepcopy(p.sys)
    s.field_dir = (-0.2, 0.6, 0.77)
    monkeypatch.setenv("SPC_SSTEP", "0")
    col = dynamics.trace_sampling_para(s, sim, B=0.05, ctx=ctx)
    monkeypatch.delenv("SPC_SSTEP")
    nwt = dynamics.trace_sampling_para(s, sim, B=0.05, ctx=ctx)
    monkeypatch.setenv("SPC_NEWTON", "0")
    mono = dynamics.trace_sampling_para(s, sim, B=0.05, ctx=ctx)
    for k in KEYS:
        assert max_rel(getattr(nwt, k), getattr(col, k)) < 1e-11, (m, k)
        assert max_rel(getattr(mono, k), getattr(col, k)) < 1e-11, (m, k)


def test_newton_block_small_system_matches_the_oracle(ctx):
    # N = 96, m = 20: more chain vectors than a quarter of the dimension
    w = workloads.small_anisotropic(m1=(3, 2), m2=(2,), M=8, m=20, T_ns=30.0)
    s = oracle_system(w.sys, w.B)
    res = dynamics.trace_sampling_para(w.sys, w.sim, B=w.B, ctx=ctx)
    ref = pyoracle.trace_sampling(pyoracle.build_general_csr(s), s.Z, M=w.sim.N_samples, m=w.sim.N_krylov,
                                  t_end=w.sim.t_end, dt=w.sim.dt, vec_kind=1, seed=w.sim.seed)
    for k in KEYS:
        assert max_rel(getattr(res, k), ref[k]) < 1e-10, k


@pytest.mark.parametrize("serial", ["0", "1"])
def test_newton_block_lane_chunks(ctx, monkeypatch, serial):
    # 256 lanes as four concurrent 64-lane chunks (each its own workspace and
    # stream) or on one stream: bit-identical
    p = _pair(3072)
    sim = batch.orientation_sim(M=256, T_ns=3.5)
    s = copy.deepcopy(p.sys)
    monkeypatch.setenv("SPC_LANE_SERIAL", "1")
    ref = dynamics.trace_sampling_para(s, sim, B=0.05, ctx=ctx)
    monkeypatch.setenv("SPC_LANE_SERIAL", serial)
    res = dynamics.trace_sampling_para(s, sim, B=0.05, ctx=ctx)
    for k in KEYS:
        assert np.array_equal(getattr(res, k), getattr(ref, k)), k


def test_newton_block_batched_orientations_match_the_column_path(ctx, monkeypatch):
    p = _pair(768)
    sim = batch.orientation_sim(T_ns=20.0)
    dirs = batch.fibonacci_sphere(4)
    phi = batch.orientation_yields(p.sys, sim, dirs, B=0.05, ctx=ctx)
    monkeypatch.setenv("SPC_SSTEP", "0")
    phi_c = batch.orientation_yields(p.sys, sim, dirs, B=0.05, ctx=ctx)
    assert max_rel(phi, phi_c) < 1e-11


@pytest.mark.parametrize("n2", [0, 1, 2, 3, 4, 5, 6])
def test_chain_product_with_every_electron_split_of_the_tile_nuclei(ctx, n2):
    # seven spin-1/2 nuclei (N = 512), n2 of them (the last ones) on electron 2:
    # a 64-lane chunk takes tiles of 32 nuclear indices, whose 5 in-tile nuclei
    # then hold min(n2, 5) electron-2 nuclei; the chain product's compile-time
    # electron layout for each count against the oracle over five steps
    w = workloads.small_anisotropic(m1=(2,) * (7 - n2), m2=(2,) * n2, M=64, m=16, T_ns=4.5)
    res = dynamics.trace_sampling_para(w.sys, w.sim, B=w.B, ctx=ctx)
    s = oracle_system(w.sys, w.B)
    ref = pyoracle.trace_sampling(pyoracle.build_general_csr(s), s.Z, M=64, m=16, t_end=w.sim.t_end, dt=w.sim.dt,
                                  vec_kind=1, seed=w.sim.seed)
    for k in KEYS:
        assert max_rel(getattr(res, k), ref[k]) < 1e-10, k
