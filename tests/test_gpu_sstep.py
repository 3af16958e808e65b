"""GPU: the s-step Arnoldi (SPC_SSTEP = block size; SPC_NEWTON=0 keeps the
monomial blocks where a basis serves one step) against the oracle and the
one-column-per-pass Arnoldi.

The block form builds the same Krylov space from a chain of products
y_t = H^(t+1) q_(K-1), orthogonalised a block at a time; its basis and
Hessenberg matrix differ from modified Gram-Schmidt by rounding only, so the
yields must agree with the oracle to the yield tolerance (1e-10 relative) and
with the column-by-column path to rounding.  Pins: the first two steps at
N = 36 864, M = 256 against the oracle; every block size 2..7 against the
column path on a 3 072-dimensional pair; the equal-rate invariant (kS = kT = k:
every vector's norm decays as exp(-k t) whatever the basis, 1e-12 over 40
steps); tol > 0 sub-step outputs (they use the basis' projector Grams)."""
import copy

import numpy as np
import pytest

from helpers import max_rel, oracle_system
from oracle import pyoracle
from spin_dynamics_amd import batch, dynamics, workloads
from spin_dynamics_amd.spinsys import GAMMA_E

pytestmark = pytest.mark.gpu

KEYS = ("P_S", "P_T0", "P_Tp", "P_Tm")


def _pair(N):
    return [q for q in batch.radical_pairs() if q.N == N][0]


@pytest.mark.parametrize("S", ["2", "6", "7", "8"])
def test_block_basis_arnoldi_relation(ctx, monkeypatch, S):
    # the block basis and Hessenberg matrix through spc_krylov_basis: the
    # Arnoldi relation H Q[:, :m-1] = Q A[:, :m-1] holds to rounding (an index
    # or sign slip in the change of basis breaks it at O(1)); orthogonality
    # and the MGS Hessenberg within the block form's loss of orthogonality
    # (numpy model of this system: 3e-11 and 1e-11 at S = 6)
    monkeypatch.setenv("SPC_SSTEP", S)
    w = workloads.small_anisotropic(m1=(3, 2, 2), m2=(2, 2))
    s = oracle_system(w.sys, w.B)
    H = dynamics.Hamiltonian(w.sys, w.B, ctx)
    ref_h = pyoracle.build_general_csr(s)
    rng = np.random.default_rng(3)
    M, m = 5, 16
    psi = rng.normal(size=(M, H.N)) + 1j * rng.normal(size=(M, H.N))
    Q, A, beta = H.krylov(psi, m)
    Hs = H.scipy()
    for r in range(M):
        Qr, cr, Ar = pyoracle.get_krylov(ref_h, psi[r], m)
        assert abs(beta[r] - cr[0].real) < 1e-13 * abs(cr[0])
        Qm = Q[r].T
        assert max_rel(Hs @ Qm[:, :m - 1], Qm @ A[r][:, :m - 1]) < 1e-12
        assert np.max(np.abs(Qm.conj().T @ Qm - np.eye(m))) < 1e-9
        assert max_rel(A[r], Ar) < 1e-9
        assert np.all(A[r][np.tril_indices(m, -2)] == 0)  # Hessenberg: zero below the subdiagonal


@pytest.mark.parametrize("S,newton", [("0", "1"), ("6", "0"), ("8", "0"), ("8", "1")])  # 0: the one-column-per-pass form
def test_sstep_first_steps_match_the_oracle(ctx, monkeypatch, S, newton):
    monkeypatch.setenv("SPC_SSTEP", S)
    monkeypatch.setenv("SPC_NEWTON", newton)
    w = workloads.anisotropic_36864(M=256, m=24, T_ns=1.5)  # two steps
    res = dynamics.trace_sampling_para(w.sys, w.sim, B=w.B, ctx=ctx)
    s = oracle_system(w.sys, w.B)
    ref = pyoracle.trace_sampling(pyoracle.build_general_csr(s), s.Z, M=256, m=24, t_end=w.sim.t_end, dt=w.sim.dt,
                                  vec_kind=1, seed=w.sim.seed)
    for k in KEYS:
        assert max_rel(getattr(res, k), ref[k]) < 1e-10, k


@pytest.mark.parametrize("S", [2, 3, 4, 5, 6, 7, 8])
def test_every_block_size_matches_the_column_path(ctx, monkeypatch, S):
    p = _pair(3072)
    sim = batch.orientation_sim(T_ns=7.5)  # eight steps, m = 24, M = 64
    s = copy.deepcopy(p.sys)
    s.field_dir = (0.3, -0.5, 0.81)
    monkeypatch.setenv("SPC_SSTEP", "0")
    col = dynamics.trace_sampling_para(s, sim, B=0.05, ctx=ctx)
    monkeypatch.setenv("SPC_SSTEP", str(S))
    monkeypatch.setenv("SPC_NEWTON", "0")
    blk = dynamics.trace_sampling_para(s, sim, B=0.05, ctx=ctx)
    for k in KEYS:
        assert max_rel(getattr(blk, k), getattr(col, k)) < 1e-11, (S, k)


@pytest.mark.parametrize("m,S", [(8, 6), (9, 6), (13, 6), (12, 2), (29, 7), (32, 6), (32, 7), (9, 8), (24, 8), (32, 8)])
def test_block_schedules_match_the_column_path(ctx, monkeypatch, m, S):
    # the smallest order the block form takes (8), orders that leave a short
    # last block, the largest order (32) and the widest chain (S = 7: a last
    # block of 7 + 1 vectors at m = 29)
    p = _pair(3072)
    sim = batch.orientation_sim(m=m, T_ns=3.5)  # four steps
    s = copy.deepcopy(p.sys)
    s.field_dir = (-0.2, 0.6, 0.77)
    monkeypatch.setenv("SPC_SSTEP", "0")
    col = dynamics.trace_sampling_para(s, sim, B=0.05, ctx=ctx)
    monkeypatch.setenv("SPC_SSTEP", str(S))
    monkeypatch.setenv("SPC_NEWTON", "0")
    blk = dynamics.trace_sampling_para(s, sim, B=0.05, ctx=ctx)
    for k in KEYS:
        assert max_rel(getattr(blk, k), getattr(col, k)) < 1e-11, (m, S, k)


@pytest.mark.parametrize("S,newton", [("6", "0"), ("8", "0"), ("8", "1")])
def test_sstep_equal_rates_decay_exactly(ctx, monkeypatch, S, newton):
    monkeypatch.setenv("SPC_SSTEP", S)
    monkeypatch.setenv("SPC_NEWTON", newton)
    w = workloads.anisotropic_36864(M=256, m=24, T_ns=39.5)  # 40 steps
    k = 1.5 / GAMMA_E
    w.sys.kS = w.sys.kT = k
    res = dynamics.trace_sampling_para(w.sys, w.sim, B=w.B, ctx=ctx)
    _, _, K = dynamics.last_run_stats(ctx)
    tot = sum(np.asarray(getattr(res, x)) for x in KEYS)[:K + 1]
    t = np.zeros(K + 1)
    for i in range(1, K + 1):
        t[i] = t[i - 1] + w.sim.dt
    assert np.max(np.abs(tot - np.exp(-k * t))) < 1e-12


@pytest.mark.parametrize("S", ["6", "8"])
def test_sstep_with_substep_outputs(ctx, monkeypatch, S):
    p = _pair(3072)
    sim = batch.orientation_sim(T_ns=7.5)
    sim.tol = 1e-8
    s = copy.deepcopy(p.sys)
    monkeypatch.setenv("SPC_SSTEP", "0")
    col = dynamics.trace_sampling_para(s, sim, B=0.05, ctx=ctx)
    monkeypatch.setenv("SPC_SSTEP", S)
    blk = dynamics.trace_sampling_para(s, sim, B=0.05, ctx=ctx)
    for k in KEYS:
        assert max_rel(getattr(blk, k), getattr(col, k)) < 1e-10, k


def _high_field_case(B, T_ns=300.0, M=8):
    # N = 384 (3x2x2 | 2x2x2), anisotropic tensors, tilted field, m = 24:
    # at B = 25 mT the spectrum clusters at -B, 0 and +B, where the monomial
    # chain of the block form is ill-conditioned (numpy model of this
    # schedule: smallest Cholesky pivot 8e-4, ||Q^H Q - I|| = 1.6e-6 at
    # S = 8, against 1.2e-2 and 2e-10 at 0.05 mT)
    w = workloads.small_anisotropic(m1=(3, 2, 2), m2=(2, 2, 2), M=M, m=24, T_ns=T_ns)
    w.B = B
    w.sys.e1.w = w.sys.e2.w = -B
    s = oracle_system(w.sys, w.B)
    vecs = np.array([pyoracle.sample_vector(sd, s.Z, 1) for sd in pyoracle.sample_seeds((4, 8), M)])
    return w, s, vecs


@pytest.mark.parametrize("newton", ["0", "1"])
@pytest.mark.parametrize("B", [0.05, 25.0])
def test_high_field_runs_column_by_column(ctx, monkeypatch, B, newton):
    w, s, vecs = _high_field_case(B)
    monkeypatch.setenv("SPC_NEWTON", newton)
    res = dynamics.trace_sampling_para(w.sys, w.sim, B=w.B, trace_vectors=vecs, ctx=ctx)
    monkeypatch.setenv("SPC_SSTEP", "0")
    col = dynamics.trace_sampling_para(w.sys, w.sim, B=w.B, trace_vectors=vecs, ctx=ctx)
    monkeypatch.delenv("SPC_SSTEP")
    ref = pyoracle.trace_sampling(pyoracle.build_general_csr(s), s.Z, M=len(vecs), m=w.sim.N_krylov,
                                  t_end=w.sim.t_end, dt=w.sim.dt, vec_kind=2, injected=vecs)
    same = all(np.array_equal(getattr(res, k), getattr(col, k)) for k in KEYS)
    # 25 mT: the first rebuild's Newton block is past the condition limit, so
    # the chunk starts again column by column and gives the column form's
    # results bit for bit; 0.05 mT keeps the block form (within rounding of
    # the column form)
    assert same == (B > 1.0)
    for k in KEYS:
        assert max_rel(getattr(res, k), ref[k]) < 1e-10, k
        assert max_rel(getattr(res, k), getattr(col, k)) < 1e-11, k


def test_high_field_batch_and_gradient_run_column_by_column(ctx, monkeypatch):
    import copy as cp
    w, s, _ = _high_field_case(25.0, T_ns=40.0, M=64)
    dirs = batch.fibonacci_sphere(3)
    phi = batch.orientation_yields(w.sys, w.sim, dirs, B=w.B, ctx=ctx)
    g = gradients_of(w, ctx)
    monkeypatch.setenv("SPC_SSTEP", "0")
    phi_c = batch.orientation_yields(w.sys, w.sim, dirs, B=w.B, ctx=ctx)
    g_c = gradients_of(w, ctx)
    monkeypatch.delenv("SPC_SSTEP")
    assert np.array_equal(phi, phi_c)
    assert np.array_equal(g.hfc, g_c.hfc) and np.array_equal(g.field, g_c.field) and g.Phi_S == g_c.Phi_S
    for o, d in enumerate(dirs):
        sd = cp.deepcopy(w.sys)
        sd.field_dir = tuple(d)
        r = dynamics.trace_sampling_para(sd, w.sim, B=w.B, ctx=ctx)
        r.get_kinetics(w.sim.dt, sd.kS, sd.kT)
        assert abs(phi[o] - r.Phi_S[-1]) <= 1e-12 * abs(r.Phi_S[-1])


def gradients_of(w, ctx):
    from spin_dynamics_amd import gradients
    return gradients.yield_gradient(w.sys, w.sim, B=w.B, ctx=ctx)
