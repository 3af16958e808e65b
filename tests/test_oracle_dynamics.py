"""Oracle trace sampling / exact dynamics (qm_dynamics.f90:24-384) against dense
scipy propagation and closed forms; Simpson yields (observables.f90:74-101)
against the compiled reference fixture."""
import json
import os

import numpy as np
import pytest

import dense_model as dm
from oracle import pyoracle
from spin_dynamics_amd.observables import Observables

GE = pyoracle.GAMMA_E
DT = 1.0 * GE / 1000.0


def _aniso(kS=2 / GE, kT=0.5 / GE, m1=(3, 2), m2=(2,)):
    rng = np.random.default_rng(8)
    A = []
    for _ in range(len(m1) + len(m2)):
        lam = rng.uniform(-1.5, 1.5, 3)
        Q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
        A.append(Q @ np.diag(lam) @ Q.T)
    A = np.array(A)
    return pyoracle.System(mult1=list(m1), mult2=list(m2), A1=A[:len(m1)], A2=A[len(m1):], J=0.01, kS=kS, kT=kT,
                           B=0.05, direction=(np.sin(0.6), 0, np.cos(0.6)))


def test_trace_sampling_matches_dense_propagation():
    s = _aniso()
    h = pyoracle.build_general_csr(s)
    M, T = 5, 150 * DT
    vecs = np.array([pyoracle.sample_vector(sd, s.Z, 0) for sd in pyoracle.sample_seeds((3, 4), M)])
    res = pyoracle.trace_sampling(h, s.Z, M=M, m=14, t_end=T, dt=DT, vec_kind=2, injected=vecs)
    K = pyoracle.steps_taken(T, DT)
    Hd = dm.hamiltonian(s.mult1, s.mult2, s.tensors(), B=s.B, direction=s.direction, J=s.J, kS=s.kS, kT=s.kT)
    ref = sum(dm.propagate_dense(Hd, np.kron(dm.theta(), v), DT, K) for v in vecs) / M
    for x, key in enumerate(["P_S", "P_T0", "P_Tp", "P_Tm"]):
        assert np.max(np.abs(res[key][:K + 1] - ref[x])) < 1e-13
        assert np.all(res[key][K + 1:] == 0.0)  # unfilled trailing entries stay zero
    assert res["P_S"][0] == 1.0 and res["P_T0"][0] == 0.0


@pytest.mark.parametrize("state,code", [("singlet", 0), ("T0", 1), ("Tp", 2), ("Tm", 3)])
def test_every_initial_state_matches_dense_propagation(state, code):
    """init_el_state (qm_dynamics.f90:386-414): theta for S, T0, T+, T- and the
    one-hot initial populations, against dense propagation from the textbook
    states (alpha beta -+ beta alpha)/sqrt2, alpha alpha, beta beta."""
    s = _aniso()
    h = pyoracle.build_general_csr(s)
    M, T = 3, 60 * DT
    vecs = np.array([pyoracle.sample_vector(sd, s.Z, 1) for sd in pyoracle.sample_seeds((9, 2), M)])
    res = pyoracle.trace_sampling(h, s.Z, M=M, m=14, t_end=T, dt=DT, vec_kind=2, injected=vecs, init_state=code)
    K = pyoracle.steps_taken(T, DT)
    Hd = dm.hamiltonian(s.mult1, s.mult2, s.tensors(), B=s.B, direction=s.direction, J=s.J, kS=s.kS, kT=s.kT)
    ref = sum(dm.propagate_dense(Hd, np.kron(dm.theta(state), v), DT, K) for v in vecs) / M
    for x, key in enumerate(["P_S", "P_T0", "P_Tp", "P_Tm"]):
        assert res[key][0] == (1.0 if x == code else 0.0)
        assert np.max(np.abs(res[key][:K + 1] - ref[x])) < 1e-13


def test_equal_rates_give_exact_exponential_decay():
    k = 3 / GE
    s = _aniso(kS=k, kT=k)
    h = pyoracle.build_general_csr(s)
    T = 300 * DT
    res = pyoracle.trace_sampling(h, s.Z, M=4, m=16, t_end=T, dt=DT, vec_kind=1)
    K = pyoracle.steps_taken(T, DT)
    tot = sum(res[x] for x in ("P_S", "P_T0", "P_Tp", "P_Tm"))[:K + 1]
    t = np.zeros(K + 1)
    for i in range(1, K + 1):
        t[i] = t[i - 1] + DT
    assert np.max(np.abs(tot - np.exp(-k * t))) < 1e-13


def test_no_nuclei_singlet_is_an_eigenstate():
    kS, kT = 2.5 / GE, 0.5 / GE
    s = pyoracle.System(mult1=[], mult2=[], J=0.3, kS=kS, kT=kT, B=1.0)
    h = pyoracle.build_iso_csr(s)
    T = 100 * DT
    res = pyoracle.trace_sampling(h, 1, M=2, m=1, t_end=T, dt=DT)
    K = pyoracle.steps_taken(T, DT)
    t = np.cumsum(np.r_[0.0, np.full(K, DT)])
    assert np.max(np.abs(res["P_S"][:K + 1] - np.exp(-kS * t))) < 1e-13
    assert np.max(np.abs(res["P_T0"][:K + 1])) < 1e-15


def test_exact_dynamics_is_the_dense_trace():
    s = _aniso(m1=(2,), m2=(3,))
    h = pyoracle.build_general_csr(s)
    T = 80 * DT
    res = pyoracle.exact_dynamics(h, s.Z, m=12, t_end=T, dt=DT)
    K = pyoracle.steps_taken(T, DT)
    Hd = dm.hamiltonian(s.mult1, s.mult2, s.tensors(), B=s.B, direction=s.direction, J=s.J, kS=s.kS, kT=s.kT)
    ref = 0
    for i in range(s.Z):
        e = np.zeros(s.Z, complex)
        e[i] = 1
        ref = ref + dm.propagate_dense(Hd, np.kron(dm.theta(), e), DT, K)
    assert np.max(np.abs(res["P_S"][:K + 1] - ref[0] / s.Z)) < 1e-13


def test_trace_sampling_is_unbiased():
    s = _aniso(m1=(2, 2), m2=(2,))
    h = pyoracle.build_general_csr(s)
    T = 60 * DT
    exact = pyoracle.exact_dynamics(h, s.Z, m=12, t_end=T, dt=DT)["P_S"]
    ests = np.array([pyoracle.trace_sampling(h, s.Z, M=16, m=12, t_end=T, dt=DT, seed=(i, 7), vec_kind=0)["P_S"]
                     for i in range(12)])
    err = np.abs(ests.mean(axis=0) - exact)
    sem = ests.std(axis=0, ddof=1) / np.sqrt(len(ests)) + 1e-12
    assert np.max(err / sem) < 5.0


def test_tol_reuses_krylov_spaces_and_stays_accurate():
    s = _aniso()
    h = pyoracle.build_general_csr(s)
    T = 200 * DT
    a = pyoracle.trace_sampling(h, s.Z, M=3, m=16, t_end=T, dt=DT, tol=0.0)
    b = pyoracle.trace_sampling(h, s.Z, M=3, m=16, t_end=T, dt=DT, tol=1e-10)
    assert np.max(np.abs(a["P_S"] - b["P_S"])) < 1e-5


def test_output_length_rule():
    # ceiling((t_end+dt)/dt) + 1 entries; 5 us / 1 ns gives 5002 values
    assert pyoracle.n_steps(5000 * DT, DT) + 1 == 5002
    assert pyoracle.n_steps(10000 * DT, DT) + 1 == 10002


KIN = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "ref_kinetics.json")))


@pytest.mark.parametrize("case", range(len(KIN)))
def test_simpson_yields_match_reference_fixture(case):
    c = KIN[case]
    P = [np.array([float.fromhex(x) for x in p]) for p in c["P"]]
    iden, phs, pht = pyoracle.get_kinetics(*P, c["dt"], c["kS"], c["kT"])
    assert [float(x).hex() for x in iden] == c["iden"]
    assert [float(x).hex() for x in phs] == c["Phi_S"]
    assert [float(x).hex() for x in pht] == c["Phi_T"]
    obs = Observables(*P)
    obs.get_kinetics(c["dt"], c["kS"], c["kT"])
    assert np.array_equal(obs.Phi_S, phs) and np.array_equal(obs.Phi_T, pht)


def test_simpson_yield_closed_form():
    kS, kT = 2.0 / GE, 0.5 / GE
    n = 5002
    t = np.arange(n) * DT
    PS = np.exp(-kS * t)
    z = np.zeros(n)
    _, phs, _ = pyoracle.get_kinetics(PS, z, z, z, DT, kS, kT)
    T = 2 * (n // 2 - 1) * DT
    assert abs(phs[-1] - (1 - np.exp(-kS * T))) < 1e-12
