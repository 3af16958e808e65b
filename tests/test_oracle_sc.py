"""The oracle's semiclassical model (orc_semiclassical) pinned by closed forms.

No reference output exists for src/sc (it does not compile and its driver
entry prints "method not yet implemented"), so the pins are what the model's
mathematics fixes:
- kS = kT = k and J = 0 make the generator K + J = -k 1, so every sample
  decays uniformly; with no nuclei both electron spins precess about the same
  field, which conserves S1.S2 and tr T12 only if T12 turns as R T R^T: P_S(t)
  is exactly e^{-k t} P_S(0).
- kS = kT = k with nuclei: x_16 decays as e^{-k t} in every sample, so
  P_S + P_T is exactly e^{-k t} (P_S(0) + P_T(0)).
- With no field and no nuclei, the singlet population decays at kS whatever
  kT and J (an eigenmode of the (T11, T22, T33, 1) block), and the triplet
  stays empty: statistically, within Monte Carlo error.
- kS = kT = 0 and J = 0 make exp(dt/2 (K + J)) the identity, so T12 only
  turns: its columns with S1's axis, its rows with S2's.  Started as the
  outer product S1 S2^T, it stays S1 S2^T in every sample at every step,
  whatever the nuclei do (class_radical_pair.f90:107-118).
- One electron and one nucleus at zero field with K = J = 0:
  dS/dt = a I x S and dI/dt = a S x I (class_radical_pair.f90:123-145), so
  S + I is a constant of the motion.  The Strang splitting keeps S.I (each
  sub-step turns one vector about the other) and so |S + I| to rounding, and
  the vector S + I to O(dt^2).
- Many spin-1/2 nuclei at zero field, short times: the nuclei barely move, and
  the frozen-nuclei (Schulten-Wolynes) closed form holds,
  P_S(t) = 1/4 + 3/4 f1(t) f2(t),
  f(t) = (1 + 2 (1 - s^2 t^2) exp(-s^2 t^2 / 2)) / 3, s^2 = sum_k a_k^2 I_k(I_k+1) / 3,
  within Monte Carlo error.
- Mistakes planted in the oracle itself (T12 turned the transposed way, the
  nuclear axis with the wrong sign or length, the hyperfine field at half its
  length) and in its inputs (wrong rate, time step) fail these.
"""
import math

import numpy as np
import pytest

from oracle import pyoracle


def _sc(a1, a2, m1, m2, **kw):
    kw.setdefault("seed", (42, 99))
    return pyoracle.semiclassical(a1, a2, m1, m2, with_pt=True, **kw)


@pytest.mark.parametrize("direction", [(0, 0, 1), (0.6, 0.0, 0.8), (1, 1, 1)])
def test_uniform_decay_and_common_precession(direction):
    k, dt, n = 0.3, 0.05, 60
    ps, pt = _sc([], [], [], [], B=1.7, direction=direction, kS=k, kT=k, n_samples=500, dt=dt, n_steps_=n)
    t = dt * np.arange(n + 1)
    assert abs(ps[0] - 1.0) < 0.2  # 4 <P_S(0)^2>, a sample estimate of 1
    assert np.max(np.abs(ps - ps[0] * np.exp(-k * t))) < 1e-13
    assert np.max(np.abs(pt - pt[0] * np.exp(-k * t))) < 1e-13


def test_uniform_decay_conserves_total_with_nuclei():
    k, dt, n = 0.2, 0.04, 80
    ps, pt = _sc([0.9, -0.4, 0.3], [0.7, 0.2], [2, 3, 2], [2, 2], B=0.5, direction=(0.3, 0.2, 0.9), kS=k, kT=k,
                 n_samples=400, dt=dt, n_steps_=n)
    t = dt * np.arange(n + 1)
    tot = ps + pt
    assert np.max(np.abs(tot - tot[0] * np.exp(-k * t))) < 1e-12
    # the hyperfine coupling does mix singlet and triplet
    assert np.max(np.abs(ps - ps[0] * np.exp(-k * t))) > 1e-3


def test_singlet_decays_at_kS_without_mixing():
    kS, kT, dt, n = 0.5, 0.1, 0.05, 40
    ps, pt = _sc([], [], [], [], B=0.0, J=0.3, kS=kS, kT=kT, n_samples=40000, dt=dt, n_steps_=n)
    t = dt * np.arange(n + 1)
    assert np.max(np.abs(ps - np.exp(-kS * t))) < 0.03
    assert np.max(np.abs(pt)) < 0.03
    # the wrong rate is far outside that band
    assert np.max(np.abs(ps - np.exp(-kT * t))) > 0.3


def test_time_step_and_indexing_are_pinned():
    k, n = 0.3, 30
    ps1, _ = _sc([], [], [], [], B=0.4, kS=k, kT=k, n_samples=64, dt=0.05, n_steps_=n)
    ps2, _ = _sc([], [], [], [], B=0.4, kS=k, kT=k, n_samples=64, dt=0.1, n_steps_=n // 2)
    assert np.max(np.abs(ps1[::2] - ps2)) < 1e-13
    assert abs(ps1[1] / ps1[0] - math.exp(-k * 0.05)) < 1e-14


def test_samples_depend_only_on_seed():
    kw = dict(B=0.5, kS=0.2, kT=0.1, J=0.05, n_samples=50, dt=0.05, n_steps_=10)
    a = _sc([0.5, 0.2], [0.3], [2, 3], [2], **kw)
    b = _sc([0.5, 0.2], [0.3], [2, 3], [2], n_threads=1, **kw)
    c = _sc([0.5, 0.2], [0.3], [2, 3], [2], **{**kw, "seed": (1, 2)})
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    assert not np.array_equal(a[0], c[0])


def _traj(*args, **kw):
    return pyoracle.semiclassical_trajectory(*args, **kw)


def _t12_defect(xs):
    S1, S2 = xs[:, 0:3], xs[:, 3:6]
    T = xs[:, 6:15].reshape(-1, 3, 3).transpose(0, 2, 1)  # T[a, b] = x[6 + a + 3 b]
    return float(np.max(np.abs(T - S1[:, :, None] * S2[:, None, :])))


_COUPLED = dict(a_iso1=[0.9, -0.4, 0.3], a_iso2=[0.7, 0.2], mult1=[2, 3, 2], mult2=[2, 2])


def _coupled_traj(sample):
    c = _COUPLED
    return _traj(c["a_iso1"], c["a_iso2"], c["mult1"], c["mult2"], B=0.5, direction=(0.3, 0.2, 0.9), dt=0.05,
                 n_steps_=200, sample=sample)


def test_t12_stays_outer_product_without_recombination():
    for smp in range(4):
        xs, Is = _coupled_traj(smp)
        assert _t12_defect(xs) < 1e-12
        assert np.all(xs[:, 15] == 1.0)
        # the hyperfine fields do turn the two electrons differently
        assert abs(float(np.dot(xs[-1, 0:3], xs[-1, 3:6]) - np.dot(xs[0, 0:3], xs[0, 3:6]))) > 1e-2


def test_t12_pin_catches_transposed_rotation():
    with pyoracle.sc_mutation("t12_transposed"):
        xs, _ = _coupled_traj(0)
    assert _t12_defect(xs) > 1e-2


def _one_nucleus(dt, sample=3, T=20.0):
    xs, Is = _traj([1.3], [], [2], [], B=0.0, dt=dt, n_steps_=int(round(T / dt)), sample=sample)
    S, I = xs[:, 0:3], Is[:, 0, :]
    tot = S + I
    drift = float(np.max(np.abs(tot - tot[0])))
    sdi = float(np.max(np.abs(np.sum(S * I, axis=1) - np.dot(S[0], I[0]))))
    return drift, sdi, S


def test_total_spin_conserved_for_one_nucleus():
    d1, s1, S = _one_nucleus(0.02)
    d2, s2, _ = _one_nucleus(0.01)
    assert s1 < 1e-13 and s2 < 1e-13           # S.I exactly, hence |S + I|
    assert d1 < 1e-4 and 3.9 < d1 / d2 < 4.1    # S + I to second order in dt
    assert float(np.max(np.abs(S - S[0]))) > 0.2  # while S itself precesses


@pytest.mark.parametrize("mistake", ["nuclear_sign", "nuclear_scale", "field_scale"])
def test_total_spin_pin_catches_nuclear_step_mistakes(mistake):
    with pyoracle.sc_mutation(mistake):
        d, _, _ = _one_nucleus(0.02)
    assert d > 0.1


def _schulten_wolynes(t, a):
    s2 = float(np.sum(np.asarray(a) ** 2)) * 0.75 / 3.0
    return (1.0 + 2.0 * (1.0 - s2 * t * t) * np.exp(-0.5 * s2 * t * t)) / 3.0


def _sw_run():
    rng = np.random.default_rng(7)
    n = 24
    a1, a2 = list(rng.uniform(0.1, 0.3, n)), list(rng.uniform(0.1, 0.3, n))
    dt, steps = 0.02, 200  # s t up to about 2, past the dip of f at s t = sqrt(3)
    ps = pyoracle.semiclassical(a1, a2, [2] * n, [2] * n, B=0.0, n_samples=40000, dt=dt, n_steps_=steps)
    t = dt * np.arange(steps + 1)
    return ps, 0.25 + 0.75 * _schulten_wolynes(t, a1) * _schulten_wolynes(t, a2)


def test_frozen_nuclei_limit_schulten_wolynes():
    ps, sw = _sw_run()
    assert np.min(sw) < 0.3  # the window reaches the dip
    assert float(np.max(np.abs(ps - sw))) < 0.015


def test_frozen_nuclei_pin_catches_field_scale():
    with pyoracle.sc_mutation("field_scale"):
        ps, sw = _sw_run()
    assert float(np.max(np.abs(ps - sw))) > 0.1
