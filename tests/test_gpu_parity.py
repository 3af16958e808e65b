"""GPU parity: the HIP path against the CPU oracle and the dense model.

Tolerances: Hamiltonian entries and single products to 1e-14 relative;
populations and yields to 1e-10 relative (the yield tolerance of the path)."""
import numpy as np
import pytest

import dense_model as dm
from helpers import max_rel, oracle_system
from oracle import pyoracle
from spin_dynamics_amd import _lib, dynamics, workloads

pytestmark = pytest.mark.gpu


def _dense(h):
    return h.scipy().toarray()


def test_hamiltonian_isotropic_matches_reference_builder(ctx):
    w = workloads.isotropic_192()
    H = dynamics.Hamiltonian(w.sys, w.B, ctx)
    ref = pyoracle.build_iso_csr(oracle_system(w.sys, w.B))
    assert H.N == 192 and H.nnz == ref.nnz == 800
    assert max_rel(H.scipy().toarray(), ref.scipy().toarray()) < 1e-14


@pytest.mark.parametrize("case", ["iso192", "mixed", "exchange"])
def test_hamiltonian_export_is_coo_to_csr_order(ctx, case):
    """getHiso_sp + coo_to_csr (hamiltonian.f90:17-205, utils.f90:52-89): the
    exported rowptr and col equal the oracle's bit for bit (reverse COO order
    within a row, the zeros the reference stores included)."""
    import copy
    from spin_dynamics_amd.spinsys import Electron
    w = workloads.isotropic_192()  # kS = kT, J = 0: its flip-flop entries are stored zeros
    sys = copy.deepcopy(w.sys)
    if case != "iso192":
        m1, m2 = ((3, 2, 2), (2, 3)) if case == "mixed" else ((2, 2), (3,))
        a = np.random.default_rng(3).uniform(-1, 1, len(m1) + len(m2))
        sys.e1 = Electron(g_I=list(m1), A=np.array([x * np.eye(3) for x in a[:len(m1)]]))
        sys.e2 = Electron(g_I=list(m2), A=np.array([x * np.eye(3) for x in a[len(m1):]]))
        if case == "exchange":
            sys.J, sys.kS, sys.kT = 0.2, 2.0 / 176.08596, 0.5 / 176.08596
    H = dynamics.Hamiltonian(sys, w.B, ctx)
    rowptr, col, val = H.csr()
    rr, rc, rv = pyoracle.build_iso_csr(oracle_system(sys, w.B)).arrays()
    assert H.nnz == len(rc)
    assert np.array_equal(rowptr, rr) and np.array_equal(col, rc)
    assert np.max(np.abs(val - rv)) <= 1e-14 * np.max(np.abs(rv))
    if case == "iso192":
        assert np.count_nonzero(rv == 0) >= H.N // 2


def test_hamiltonian_anisotropic_tilted_matches_oracle_and_dense(ctx):
    w = workloads.small_anisotropic(m1=(3, 2, 2), m2=(2, 3))
    s = oracle_system(w.sys, w.B)
    H = dynamics.Hamiltonian(w.sys, w.B, ctx)
    g = pyoracle.build_general_csr(s).scipy()
    h = H.scipy().sorted_indices()  # the export orders a row by term, the general builder by column
    assert np.array_equal(h.indptr, g.indptr) and np.array_equal(h.indices, g.indices)
    assert max_rel(h.data, g.data) < 1e-14
    Hd = dm.hamiltonian(s.mult1, s.mult2, s.tensors(), B=s.B, direction=s.direction, J=s.J, D=s.D, kS=s.kS, kT=s.kT)
    assert max_rel(H.scipy().toarray(), Hd) < 1e-14


@pytest.mark.parametrize("M", [1, 64, 70])
def test_spmm(ctx, M):
    w = workloads.small_anisotropic()
    H = dynamics.Hamiltonian(w.sys, w.B, ctx)
    rng = np.random.default_rng(M)
    X = rng.normal(size=(M, H.N)) + 1j * rng.normal(size=(M, H.N))
    Y = H.spmm(X)
    ref = (H.scipy() @ X.T).T
    assert max_rel(Y, ref) < 1e-14


@pytest.mark.parametrize("m,scale", [(1, 0.3), (5, 0.01), (12, 1.0), (20, 4.0), (24, 40.0), (32, 2.0)])
def test_expm_small(ctx, m, scale):
    from scipy.linalg import expm
    rng = np.random.default_rng(m)
    A = (rng.normal(size=(m, m)) + 1j * rng.normal(size=(m, m))) * scale / m
    E = dynamics.expm(A, ctx)
    assert max_rel(E, pyoracle.expm(A)) < 1e-13
    assert max_rel(E, expm(A)) < 1e-12


def test_krylov_basis_matches_mgs(ctx):
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
        assert max_rel(A[r], Ar) < 1e-11
        Qm = Q[r].T  # N x m
        assert np.max(np.abs(Qm.conj().T @ Qm - np.eye(m))) < 1e-13
        # Arnoldi relation H Q_{m-1} = Q_m A[:, :m-1]
        assert max_rel(Hs @ Qm[:, :m - 1], Qm @ A[r][:, :m - 1]) < 1e-12


@pytest.mark.parametrize("M,tol", [(8, 0.0), (70, 0.0)])
def test_trace_sampling_injected_vs_oracle(ctx, M, tol):
    w = workloads.small_anisotropic(M=M, tol=tol)
    s = oracle_system(w.sys, w.B)
    Z = s.Z
    vecs = np.array([pyoracle.sample_vector(sd, Z, 1) for sd in pyoracle.sample_seeds((7, 11), M)])
    res = dynamics.trace_sampling_para(w.sys, w.sim, B=w.B, trace_vectors=vecs, ctx=ctx)
    ref = pyoracle.trace_sampling(pyoracle.build_general_csr(s), Z, M=M, m=w.sim.N_krylov, t_end=w.sim.t_end,
                                  dt=w.sim.dt, tol=tol, vec_kind=2, injected=vecs)
    for k in ("P_S", "P_T0", "P_Tp", "P_Tm"):
        assert max_rel(getattr(res, k), ref[k]) < 1e-10, k
    _, phs, _ = pyoracle.get_kinetics(ref["P_S"], ref["P_T0"], ref["P_Tp"], ref["P_Tm"], w.sim.dt, s.kS, s.kT)
    assert abs(res.Phi_S[-1] - phs[-1]) < 1e-10 * abs(phs[-1])


@pytest.mark.parametrize("state", ["T0", "Tp", "Tm"])
@pytest.mark.parametrize("tol", [0.0, 1e-8])
def test_triplet_initial_states_vs_oracle(ctx, state, tol):
    """init_el_state's triplet branches (qm_dynamics.f90:395-406) on the device."""
    import copy
    w = workloads.small_anisotropic(M=8, tol=tol)
    sim = copy.deepcopy(w.sim)
    sim.init_state = state
    s = oracle_system(w.sys, w.B)
    vecs = np.array([pyoracle.sample_vector(sd, s.Z, 1) for sd in pyoracle.sample_seeds((7, 13), 8)])
    res = dynamics.trace_sampling_para(w.sys, sim, B=w.B, trace_vectors=vecs, ctx=ctx)
    ref = pyoracle.trace_sampling(pyoracle.build_general_csr(s), s.Z, M=8, m=sim.N_krylov, t_end=sim.t_end,
                                  dt=sim.dt, tol=tol, vec_kind=2, injected=vecs, init_state=_lib.STATE[state])
    bound = 1e-10 if tol == 0.0 else 1e-9
    for k in ("P_S", "P_T0", "P_Tp", "P_Tm"):
        assert max_rel(getattr(res, k), ref[k]) < bound, k
    assert getattr(res, {"T0": "P_T0", "Tp": "P_Tp", "Tm": "P_Tm"}[state])[0] == 1.0


def test_trace_sampling_tol_gram_mode(ctx):
    # tol > 0: several steps per Krylov basis, populations from projector Grams.
    # Rebuild decisions compare |c_m|^2 with tol |c|^2; a rounding-level flip
    # would move results by ~tol, so the bound is loose where a flip is possible.
    w = workloads.small_anisotropic(M=6, tol=1e-8)
    s = oracle_system(w.sys, w.B)
    vecs = np.array([pyoracle.sample_vector(sd, s.Z, 1) for sd in pyoracle.sample_seeds((5, 6), 6)])
    res = dynamics.trace_sampling_para(w.sys, w.sim, B=w.B, trace_vectors=vecs, ctx=ctx)
    ref = pyoracle.trace_sampling(pyoracle.build_general_csr(s), s.Z, M=6, m=w.sim.N_krylov, t_end=w.sim.t_end,
                                  dt=w.sim.dt, tol=1e-8, vec_kind=2, injected=vecs)
    assert max_rel(res.P_S, ref["P_S"]) < 1e-9
    _, rebuilds, steps = dynamics.last_run_stats(ctx)
    assert rebuilds < steps


@pytest.mark.parametrize("kind", ["phase", "SUZ"])
def test_isotropic_192_vs_oracle_fixture(ctx, kind):
    """The isotropic N=192 workload (M=32, m=20, T=5 us) with on-device trace vectors."""
    import os
    w = workloads.isotropic_192(vec_kind=kind)
    ref = np.load(os.path.join(os.path.dirname(__file__), "golden", f"oracle_iso192_{kind}.npz"))
    res = dynamics.trace_sampling_para(w.sys, w.sim, B=w.B, ctx=ctx)
    for k in ("P_S", "P_T0", "P_Tp", "P_Tm"):
        assert max_rel(getattr(res, k), ref[k]) < 1e-10, k
    assert abs(res.Phi_S[-1] - ref["Phi_S"][-1]) < 1e-10 * abs(ref["Phi_S"][-1])


def test_exact_dynamics_vs_oracle(ctx):
    w = workloads.small_anisotropic(m1=(2,), m2=(3,), T_ns=150.0)
    s = oracle_system(w.sys, w.B)
    res = dynamics.exact_dynamics(w.sys, w.sim, B=w.B, ctx=ctx)
    ref = pyoracle.exact_dynamics(pyoracle.build_general_csr(s), s.Z, m=w.sim.N_krylov, t_end=w.sim.t_end,
                                  dt=w.sim.dt)
    assert max_rel(res.P_S, ref["P_S"]) < 1e-10
    # exact trace against dense expm propagation, averaged over basis states
    Hd = dm.hamiltonian(s.mult1, s.mult2, s.tensors(), B=s.B, direction=s.direction, J=s.J, D=s.D, kS=s.kS, kT=s.kT)
    K = pyoracle.steps_taken(w.sim.t_end, w.sim.dt)
    acc = 0
    for i in range(s.Z):
        e = np.zeros(s.Z, complex)
        e[i] = 1
        acc = acc + dm.propagate_dense(Hd, np.kron(dm.theta(), e), w.sim.dt, K)
    assert max_rel(res.P_S[:K + 1], acc[0] / s.Z) < 1e-10


def test_errors_are_reported(ctx):
    w = workloads.small_anisotropic()
    w.sys.e1.g_I = [4, 2]
    with pytest.raises(_lib.SpinchemError) as ei:
        dynamics.Hamiltonian(w.sys, w.B, ctx)
    assert ei.value.code == _lib.SPC_ERR_UNSUPPORTED and "must be 3 or 2" in str(ei.value)
    w = workloads.small_anisotropic()
    w.sim.N_krylov = 33
    with pytest.raises(_lib.SpinchemError) as ei:
        dynamics.trace_sampling_para(w.sys, w.sim, B=w.B, ctx=ctx)
    assert ei.value.code == _lib.SPC_ERR_UNSUPPORTED


@pytest.mark.parametrize("case", ["iso", "aniso_mixed", "trailing1", "aniso_36864_like", "trailing6"])
def test_kronecker_product_equals_block_product(ctx, monkeypatch, case):
    # the Kronecker-form product (diagonal blocks + 2x2 flip tables) and the
    # 4x4-block product sum the same terms in the same order: bitwise equal,
    # also when a wave takes 2^q consecutive nuclear indices on 64 / 2^q lanes
    # (q is capped by the trailing spin-1/2 nuclei: 4 for iso and aniso_36864_like,
    # 0 for aniso_mixed, 1 for trailing1, 6 for trailing6)
    if case == "iso":
        w = workloads.isotropic_192(M=8, T_ns=20.0)
    elif case == "aniso_mixed":
        w = workloads.small_anisotropic(m1=(3, 2, 2), m2=(2, 3), M=8, T_ns=20.0)
    elif case == "trailing1":
        w = workloads.small_anisotropic(m1=(2, 3), m2=(3, 2), M=8, T_ns=20.0)
    elif case == "trailing6":
        w = workloads.small_anisotropic(m1=(3, 2, 2), m2=(2, 2, 2, 2), M=8, T_ns=15.0)
    else:
        w = workloads.small_anisotropic(m1=(3, 3, 2, 2), m2=(2, 2), M=70, m=16, T_ns=15.0)
    H = dynamics.Hamiltonian(w.sys, w.B, ctx)
    rng = np.random.default_rng(1)
    X = rng.normal(size=(5, H.N)) + 1j * rng.normal(size=(5, H.N))
    Xw = rng.normal(size=(130, H.N)) + 1j * rng.normal(size=(130, H.N))  # 192 lanes: 3 chunks
    monkeypatch.setenv("SPC_SPMM", "blocks")
    b, bw = H.spmm(X), H.spmm(Xw)
    rb = dynamics.trace_sampling_para(w.sys, w.sim, B=w.B, ctx=ctx)
    monkeypatch.delenv("SPC_SPMM")
    monkeypatch.setenv("SPC_KRON_TILE", "0")
    for q in ("0", "1", "2", "3"):
        monkeypatch.setenv("SPC_KRON_Q", q)
        a = H.spmm(X)
        assert np.array_equal(a, b), q
        assert np.array_equal(H.spmm(Xw), bw), q
        ra = dynamics.trace_sampling_para(w.sys, w.sim, B=w.B, ctx=ctx)
        # the product's dot products are summed in another order for q > 0
        assert max_rel(ra.P_S, rb.P_S) < 1e-12 and max_rel(ra.P_Tp, rb.P_Tp) < 1e-12, q
        if q == "0":
            assert np.array_equal(ra.P_S, rb.P_S) and np.array_equal(ra.P_Tp, rb.P_Tp)
    # workgroups of 4, 8 or 16 waves on consecutive wave blocks (where the
    # system has q + log2(waves) trailing spin-1/2 nuclei; otherwise fewer
    # waves, then the plain slices)
    for q in ("2", "3"):
        monkeypatch.setenv("SPC_KRON_Q", q)
        for nw in ("0", "4", "8", "16"):
            monkeypatch.setenv("SPC_KRON_TILE", nw)
            assert np.array_equal(H.spmm(X), b), (q, nw)
            assert np.array_equal(H.spmm(Xw), bw), (q, nw)
            ra = dynamics.trace_sampling_para(w.sys, w.sim, B=w.B, ctx=ctx)
            assert max_rel(ra.P_S, rb.P_S) < 1e-12 and max_rel(ra.P_Tp, rb.P_Tp) < 1e-12, (q, nw)
    assert max_rel(b, X @ H.scipy().toarray().T) < 1e-13
