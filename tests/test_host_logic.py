"""Host-side logic of the batched, gradient and semiclassical paths (CPU only):
orientation grids, pair sets, seed derivation, the symmetric-tensor
parameterisation and the workload definitions."""
import math

import numpy as np
import pytest

from oracle import pyoracle
from spin_dynamics_amd import batch, gradients, workloads
from spin_dynamics_amd.spinsys import GAMMA_E


def test_fibonacci_sphere_is_unit_and_even():
    d = batch.fibonacci_sphere(96)
    assert d.shape == (96, 3)
    assert np.allclose(np.linalg.norm(d, axis=1), 1.0, atol=1e-15)
    # equal-area bands in z, centroid near the origin, no repeated directions
    assert np.allclose(np.diff(d[:, 2]), -2.0 / 96)
    assert np.linalg.norm(d.mean(axis=0)) < 0.02
    dots = d @ d.T - 2 * np.eye(96)
    assert dots.max() < 0.99


def test_radical_pairs_are_deterministic_and_in_the_dimension_set():
    a, b = batch.radical_pairs(), batch.radical_pairs()
    assert len(a) == 48
    for p, q in zip(a, b):
        assert p.N == q.N and p.sys.e1.g_I == q.sys.e1.g_I
        assert np.array_equal(p.sys.e1.A, q.sys.e1.A)
        assert p.N in batch.DIMS
        assert 4 * p.sys.Z1 * p.sys.Z2 == p.N
        for A in list(p.sys.e1.A) + list(p.sys.e2.A):
            assert np.allclose(A, A.T)
            assert np.max(np.abs(np.linalg.eigvalsh(A))) <= 1.5 + 1e-12
    assert len({p.N for p in a}) == 5


def test_sample_seeds_match_the_oracle_stream():
    # the oracle's generator is pinned to the compiled reference m_random.f90
    for master, M in (((42, 99), 5), ((12345, 6789), 3), ((-7, 3), 4)):
        assert np.array_equal(gradients.sample_seeds(master, M), pyoracle.sample_seeds(master, M))


def test_symmetric_tensor_parameterisation_round_trip():
    rng = np.random.default_rng(0)
    h6 = rng.normal(size=(4, 6))
    A = gradients.tensors_from_sym(h6)
    assert np.allclose(A, np.transpose(A, (0, 2, 1)))
    for q, (a, b) in enumerate(gradients.SYM):
        assert np.array_equal(A[:, a, b], h6[:, q])
    n = gradients.field_direction(0.7, 1.9)
    assert math.isclose(np.linalg.norm(n), 1.0, rel_tol=1e-15)
    prob = gradients.YieldProblem(mult1=[3, 2], mult2=[2, 2], sim=workloads.isotropic_192().sim)
    sys = prob.system(h6, 0.3, 0.7, 1.9, 0.01, 2.0 / GAMMA_E, 0.5 / GAMMA_E)
    assert sys.e1.g_I == [3, 2] and sys.e2.g_I == [2, 2]
    assert np.array_equal(sys.e2.A[1], A[3]) and sys.e1.w == -0.3
    assert np.allclose(sys.field_dir, n)


def test_semiclassical_workload():
    w = workloads.semiclassical_48(n_samples=1024)
    assert len(w.a1) == len(w.a2) == 24
    assert w.m1 == [2] * 20 + [3] * 4 and w.m2 == w.m1
    for a in (w.a1, w.a2):
        mags = np.abs(a)
        assert math.isclose(mags.min(), 0.05) and math.isclose(mags.max(), 1.2)
    assert w.n_steps == 4001
    assert math.isclose(w.dt, 0.5 * GAMMA_E / 1000)


def test_orientation_simulation_parameters():
    sim = batch.orientation_sim()
    assert sim.N_samples == 64 and sim.N_krylov == 24 and sim.vec_kind == "phase"
    assert math.isclose(sim.dt, GAMMA_E / 1000) and math.isclose(sim.t_end, 10 * GAMMA_E)
