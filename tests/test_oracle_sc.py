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
- Mutations planted in the oracle's inputs (wrong rate, time step) fail these.
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
