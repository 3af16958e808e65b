"""Oracle Hamiltonian builders against the dense textbook model (numpy
Kronecker products written from the physics), the reference nnz formula
(hamiltonian.f90:62-73), the coo_to_csr row order (utils.f90:79-82) and
structural identities of the Haberkorn term."""
import numpy as np
import pytest
from scipy.linalg import eigvals

import dense_model as dm
from oracle import pyoracle

GE = pyoracle.GAMMA_E


def rand_tensor(rng, w=1.5):
    lam = rng.uniform(-w, w, 3)
    Q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
    return Q @ np.diag(lam) @ Q.T


def iso_system(rng, m1, m2, **kw):
    a = rng.uniform(-1, 1, len(m1) + len(m2))
    A = np.array([x * np.eye(3) for x in a]).reshape(-1, 3, 3)
    return pyoracle.System(mult1=m1, mult2=m2, A1=A[:len(m1)] if m1 else None, A2=A[len(m1):] if m2 else None, **kw)


def dense_of(s):
    return dm.hamiltonian(s.mult1, s.mult2, s.tensors(), B=s.B, direction=s.direction, J=s.J, D=s.D, kS=s.kS,
                          kT=s.kT)


@pytest.mark.parametrize("m1,m2", [([3, 2, 2], [2, 2]), ([2], [3]), ([], [2, 2]), ([3, 3], []), ([], [])])
@pytest.mark.parametrize("J,kS,kT", [(0.0, 1 / GE, 1 / GE), (0.1, 2 / GE, 0.5 / GE)])
def test_iso_builder_matches_dense_and_nnz_formula(m1, m2, J, kS, kT):
    rng = np.random.default_rng(len(m1) * 10 + len(m2))
    s = iso_system(rng, m1, m2, J=J, D=np.diag([0.05, 0.05, -0.1]), kS=kS, kT=kT, B=0.05)
    h = pyoracle.build_iso_csr(s)
    N = 4 * s.Z
    assert h.n == N
    assert h.nnz == N + N // 2 + sum((g - 1) * (N // g) for g in m1 + m2)
    assert np.max(np.abs(h.scipy().toarray() - dense_of(s))) < 1e-15


def test_coo_to_csr_reverses_row_order():
    # getHiso_sp stores the diagonal first; coo_to_csr reverses each row, so the
    # diagonal ends up last in every row.
    s = iso_system(np.random.default_rng(0), [3, 2], [2], J=0.1, kS=1 / GE, kT=0.3 / GE, B=0.2)
    S = pyoracle.build_iso_csr(s).scipy()
    for i in range(S.shape[0]):
        assert S.indices[S.indptr[i + 1] - 1] == i


def test_builder_stop_conditions():
    s = iso_system(np.random.default_rng(1), [4], [2])
    with pytest.raises(ValueError, match="code 2"):
        pyoracle.build_iso_csr(s)
    s = iso_system(np.random.default_rng(1), [2], [2], D=np.diag([0.1, 0.2, 0.3]))
    with pytest.raises(ValueError, match="code 3"):
        pyoracle.build_iso_csr(s)


@pytest.mark.parametrize("theta", [0.0, 37.0, 90.0])
def test_general_builder_anisotropic_tilted_matches_dense(theta):
    rng = np.random.default_rng(int(theta))
    m1, m2 = [3, 2], [2, 3]
    A = np.array([rand_tensor(rng) for _ in range(4)])
    th = np.deg2rad(theta)
    D = rand_tensor(rng, 0.1)
    s = pyoracle.System(mult1=m1, mult2=m2, A1=A[:2], A2=A[2:], J=0.02, D=D, kS=2 / GE, kT=0.5 / GE, B=0.3,
                        direction=(np.sin(th), 0.0, np.cos(th)))
    g = pyoracle.build_general_csr(s).scipy()
    assert np.max(np.abs(g.toarray() - dense_of(s))) < 1e-15
    for i in range(g.shape[0]):  # rows sorted by column
        assert np.all(np.diff(g.indices[g.indptr[i]:g.indptr[i + 1]]) > 0)


def test_general_equals_iso_for_isotropic_z_field():
    s = iso_system(np.random.default_rng(5), [3, 2, 2], [2, 2], J=0.01, kS=2 / GE, kT=0.5 / GE, B=0.05)
    a = pyoracle.build_iso_csr(s).scipy().toarray()
    b = pyoracle.build_general_csr(s).scipy().toarray()
    assert np.max(np.abs(a - b)) < 1e-15


def test_haberkorn_structure():
    rng = np.random.default_rng(9)
    A = np.array([rand_tensor(rng) for _ in range(3)])
    s = pyoracle.System(mult1=[2, 3], mult2=[2], A1=A[:2], A2=A[2:], J=0.05, kS=2 / GE, kT=0.5 / GE, B=0.1,
                        direction=(0.6, 0.0, 0.8))
    H = pyoracle.build_general_csr(s).scipy().toarray()
    herm = 0.5 * (H + H.conj().T)
    anti = 0.5 * (H - H.conj().T)
    assert np.max(np.abs(herm - herm.conj().T)) == 0.0
    # anti-Hermitian part = -iK/2, K = kS P_S + kT P_T
    dims = [2, 2] + s.mult1 + s.mult2
    S1 = [dm.embed(o, 0, dims) for o in dm.spin_ops(2)]
    S2 = [dm.embed(o, 1, dims) for o in dm.spin_ops(2)]
    PS = 0.25 * np.eye(H.shape[0]) - sum(S1[a] @ S2[a] for a in range(3))
    K = s.kS * PS + s.kT * (np.eye(H.shape[0]) - PS)
    assert np.max(np.abs(anti - (-0.5j * K))) < 1e-16
    # kS = kT: anti-Hermitian part is a scalar
    s.kS = s.kT = 1.5 / GE
    H = pyoracle.build_general_csr(s).scipy().toarray()
    assert np.max(np.abs(0.5 * (H - H.conj().T) + 0.5j * s.kS * np.eye(H.shape[0]))) < 1e-17


def test_rotation_invariance_of_spectrum():
    # rotating the field and every hyperfine/dipolar tensor together is a
    # unitary change of spin axes: the (non-Hermitian) spectrum is unchanged
    rng = np.random.default_rng(11)
    A = np.array([rand_tensor(rng) for _ in range(3)])
    D = rand_tensor(rng, 0.05)
    n = np.array([0.3, -0.4, np.sqrt(1 - 0.25)])
    R, _ = np.linalg.qr(rng.normal(size=(3, 3)))
    if np.linalg.det(R) < 0:
        R[:, 0] *= -1
    kw = dict(mult1=[2, 3], mult2=[2], J=0.03, kS=2 / GE, kT=0.5 / GE, B=0.4)
    s0 = pyoracle.System(A1=A[:2], A2=A[2:], D=D, direction=tuple(n), **kw)
    AR = np.array([R @ a @ R.T for a in A])
    s1 = pyoracle.System(A1=AR[:2], A2=AR[2:], D=R @ D @ R.T, direction=tuple(R @ n), **kw)
    e0 = np.sort_complex(eigvals(pyoracle.build_general_csr(s0).scipy().toarray()))
    e1 = np.sort_complex(eigvals(pyoracle.build_general_csr(s1).scipy().toarray()))
    assert np.max(np.abs(e0 - e1)) < 1e-12
