"""The symmetry-reduced driver's host algorithms (symmetry.f90:16-246,
moments.f90:15-240): the oracle's restatement and the library's
spc_shrink / spc_symmetry_blocks (host functions, no device).

Pins:
- shrink: the M-point Gauss rule of the measure on the couplings reproduces
  their power sums 0..2M-1; the integer set sizes add up to N and their
  couplings solve the moment equations k = 1..M (Newton); couplings that take
  only M values reduce exactly; M >= N is refused.
- blocks: every combination of one multiplet per set, the first set varying
  fastest, sorted stably by dimension; the copies count the multiplets of n
  spin-1/2 (sum over blocks of copies x dimension = 2^N exactly).
- the decomposition itself: for sets of exactly equivalent nuclei, the exact
  trace (dense propagation over every nuclear basis state) of the full
  system equals the weighted sum of the blocks' exact traces, at a nonzero
  field with each multiplet carrying 1H's nuclear Zeeman term (14N's, which
  getHiso_sp would give a multiplicity-3 block, breaks it)."""
import math

import numpy as np
import pytest

import dense_model as dm
from oracle import pyoracle
from spin_dynamics_amd import symmetry

GE = pyoracle.GAMMA_E


@pytest.mark.parametrize("n,M,seed", [(8, 4, 1), (9, 3, 2), (10, 4, 3), (6, 2, 4)])
def test_shrink_moments_and_library(n, M, seed):
    a = np.random.default_rng(seed).uniform(-1.2, 1.5, n)
    w, x = pyoracle.quad_rule(np.ones(n), a, M)
    for k in range(2 * M):  # Gauss rule exact to degree 2M-1
        assert abs(np.sum(w * x ** k) - np.sum(a ** k)) <= 1e-11 * max(1.0, np.sum(np.abs(a) ** k))
    ab, nb = pyoracle.shrink(a, M)
    assert nb.sum() == n and np.all(nb >= 1)
    for k in range(1, M + 1):
        assert abs(np.sum(nb * ab ** k) - np.sum(w * x ** k)) <= 1e-10 * max(1.0, np.sum(np.abs(a) ** k))
    lab, lnb = symmetry.shrink(a, M)
    assert np.array_equal(lnb, nb)
    assert np.max(np.abs(lab - ab)) <= 1e-12 * np.max(np.abs(ab))


def test_shrink_reference_input():
    # input.ini of the reference: eight couplings, M1 = 4; electron 2 has none
    a = [1.0, 0.3, 0.5, 0.2, 0.1, 0.5, 0.5, 0.5]
    ab, nb = symmetry.shrink(a, 4)
    assert list(nb) == [2, 1, 4, 1]
    assert np.allclose(ab, pyoracle.shrink(a, 4)[0], rtol=1e-12, atol=0)
    e, ne = symmetry.shrink([], 1)
    assert len(e) == 0 and len(ne) == 0


def test_shrink_choice_follows_the_reference_criterion():
    # couplings taking two values: the Gauss rule is exact (nodes 0.4, 0.9,
    # weights 3, 2), but its weights round down to (3, 1); of the integer
    # choices (4, 1) and (3, 2), moments.f90:78-82 keeps the one whose
    # unweighted power sums sum_j a_j^k come closer to sum_i a_i^k, which is
    # (4, 1), where the weighted sums sum_j n_j a_j^k would keep the exact
    # (3, 2).  The library follows the reference (on the reference's own
    # input the two criteria give (2, 1, 4, 1) and (1, 2, 4, 1)).
    a = [0.4, 0.9, 0.4, 0.9, 0.4]
    w, x = pyoracle.quad_rule(np.ones(5), np.array(a), 2)
    assert np.allclose(w, [3, 2], rtol=0, atol=1e-12) and np.allclose(x, [0.4, 0.9], rtol=0, atol=1e-12)
    ab, nb = symmetry.shrink(a, 2)
    assert list(nb) == [4, 1]
    for k in (1, 2):  # the chosen sets still solve the moment equations
        assert abs(np.sum(nb * ab ** k) - np.sum(np.array(a) ** k)) < 1e-12


def test_shrink_refuses_M_not_below_N():
    with pytest.raises(Exception) as ei:
        symmetry.shrink([0.1, 0.2, 0.3], 3)
    assert "M<N" in str(ei.value)


@pytest.mark.parametrize("n_bar", [[2, 1, 4, 1], [3], [5, 2], [1, 1, 1], [4, 4, 3]])
def test_blocks(n_bar):
    K, Z, W = symmetry.symmetry_blocks(n_bar)
    oK, oZ, oW = pyoracle.symmetry_blocks(n_bar)
    assert [list(map(int, k)) for k in K] == oK and list(Z) == oZ and list(W) == oW
    assert len(K) == math.prod(n // 2 + 1 for n in n_bar)
    assert sum(int(w) * int(z) for w, z in zip(W, Z)) == 2 ** sum(n_bar)
    assert list(Z) == sorted(Z)
    for k in K:
        for n, g in zip(n_bar, k):
            assert 1 <= g <= n + 1 and (g - 1) % 2 == n % 2


def _exact_trace(H, Z, K, dt):
    """exact_dynamics by dense propagation of all Z basis states at once:
    P_S, P_T0, P_Tp, P_Tm averaged over theta (x) e_n."""
    from scipy.linalg import expm
    U = expm(-1j * dt * H)
    psi = np.kron(dm.theta().reshape(4, 1), np.eye(Z))  # N x Z
    out = np.zeros((4, K + 1))
    for k in range(K + 1):
        b = psi.reshape(4, Z, Z)
        out[:, k] = [0.5 * np.sum(np.abs(b[1] - b[2]) ** 2), 0.5 * np.sum(np.abs(b[1] + b[2]) ** 2),
                     np.sum(np.abs(b[0]) ** 2), np.sum(np.abs(b[3]) ** 2)]
        psi = U @ psi
    return out / Z


@pytest.mark.parametrize("wrong_gamma", [False, True])
def test_block_decomposition_is_exact(wrong_gamma):
    # electron 1: 3 nuclei at 0.8 mT and 2 at -0.45; electron 2: 2 at 0.5
    sets1, sets2 = [(0.8, 3), (-0.45, 2)], [(0.5, 2)]
    a1 = [a for a, n in sets1 for _ in range(n)]
    a2 = [a for a, n in sets2 for _ in range(n)]
    kw = dict(B=0.6, direction=(0, 0, 1), J=0.07, kS=2.0 / GE, kT=0.5 / GE)
    dt, K = 1.0 * GE / 1000.0, 40
    full = dm.hamiltonian([2] * len(a1), [2] * len(a2), [a * np.eye(3) for a in a1 + a2], **kw)
    ref = _exact_trace(full, 2 ** (len(a1) + len(a2)), K, dt)
    n_bar = [n for _, n in sets1 + sets2]
    a_bar = [a for a, _ in sets1 + sets2]
    blocks, Zb, W = pyoracle.symmetry_blocks(n_bar)
    tot = 0.0
    for k, z, w in zip(blocks, Zb, W):
        keep = [j for j, g in enumerate(k) if g > 1]
        m1 = [k[j] for j in keep if j < len(sets1)]
        m2 = [k[j] for j in keep if j >= len(sets1)]
        A = [a_bar[j] * np.eye(3) for j in keep]
        gam = [dm.GAMMA_14N if (wrong_gamma and k[j] == 3) else dm.GAMMA_1H for j in keep]
        H = dm.hamiltonian(m1, m2, A, gammas=gam, **kw)
        tot = tot + _exact_trace(H, z, K, dt) * (w * z / 2 ** sum(n_bar))
    err = float(np.max(np.abs(tot - ref)))
    if wrong_gamma:
        assert err > 1e-6
    else:
        assert err < 1e-12


def test_oracle_symmetrised_dynamics_equal_rates():
    # kS = kT = k: every block's traces sum to exp(-k t) exactly, and the block
    # weights add up to one, so the symmetrised total does too
    k = 1.5 / GE
    a = [1.0, 0.3, 0.5, 0.2, 0.1, 0.5]
    dt = 1.0 * GE / 1000.0
    tot, blocks, _ = pyoracle.symmetrised_dynamics(a, [0.4, 0.2], 2, 1, J=0.0, kS=k, kT=k, B=0.1, m=10,
                                                   t_end=29.5 * dt, dt=dt, N_samples=3, master_seed=(42, 99),
                                                   vec_kind=1)
    assert abs(sum(b[3] for b in blocks) - 1.0) < 1e-15
    s = tot["P_S"] + tot["P_T0"] + tot["P_Tp"] + tot["P_Tm"]
    t = dt * np.arange(len(s))
    n = pyoracle.steps_taken(29.5 * dt, dt) + 1
    assert np.max(np.abs(s[:n] - np.exp(-k * t[:n]))) < 1e-12
