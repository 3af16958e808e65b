"""Oracle expm (matrix_exponential.f90:14-117) against scipy and closed forms;
oracle get_krylov (qm_dynamics.f90:432-467) against the Arnoldi identities."""
import numpy as np
import pytest
from scipy.linalg import expm as sexpm

from oracle import pyoracle

GE = pyoracle.GAMMA_E


@pytest.mark.parametrize("norm", [0.0, 1e-3, 0.2, 0.6, 1.0, 3.7, 20.0, 150.0])
@pytest.mark.parametrize("m", [1, 4, 20, 24])
def test_expm_vs_scipy(norm, m):
    rng = np.random.default_rng(m + int(norm * 10))
    A = rng.normal(size=(m, m)) + 1j * rng.normal(size=(m, m))
    A = A / max(np.max(np.sum(np.abs(A), axis=1)), 1e-300) * norm
    E = pyoracle.expm(A)
    R = sexpm(A)
    assert np.max(np.abs(E - R)) <= 1e-13 * max(1.0, norm) * max(1.0, np.max(np.abs(R)))


def test_expm_closed_forms():
    th = 0.7
    E = pyoracle.expm(np.array([[0, -th], [th, 0]], complex))
    assert np.allclose(E, [[np.cos(th), -np.sin(th)], [np.sin(th), np.cos(th)]], atol=1e-15)
    d = np.array([0.3 + 1j, -2.0, 1e-3j])
    assert np.allclose(pyoracle.expm(np.diag(d)), np.diag(np.exp(d)), atol=1e-14, rtol=1e-14)
    N = np.triu(np.arange(1, 10).reshape(3, 3), 1).astype(complex)
    assert np.allclose(pyoracle.expm(N), np.eye(3) + N + N @ N / 2, atol=1e-13)
    # the propagator of a Hermitian matrix is unitary
    rng = np.random.default_rng(2)
    X = rng.normal(size=(6, 6)) + 1j * rng.normal(size=(6, 6))
    U = pyoracle.expm(-1j * (X + X.conj().T))
    assert np.max(np.abs(U.conj().T @ U - np.eye(6))) < 1e-13


def _system(kS, kT):
    rng = np.random.default_rng(4)
    A = []
    for _ in range(4):
        lam = rng.uniform(-1.5, 1.5, 3)
        Q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
        A.append(Q @ np.diag(lam) @ Q.T)
    return pyoracle.System(mult1=[3, 2], mult2=[2, 2], A1=np.array(A[:2]), A2=np.array(A[2:]), J=0.02,
                           kS=kS, kT=kT, B=0.05, direction=(0.6, 0, 0.8))


@pytest.mark.parametrize("m", [2, 12, 24])
def test_krylov_identities(m):
    s = _system(2 / GE, 0.5 / GE)
    h = pyoracle.build_general_csr(s)
    H = h.scipy()
    rng = np.random.default_rng(m)
    psi = rng.normal(size=h.n) + 1j * rng.normal(size=h.n)
    Q, c, A = pyoracle.get_krylov(h, psi, m)
    assert abs(c[0] - np.linalg.norm(psi)) < 1e-12 and np.all(c[1:] == 0)
    assert np.max(np.abs(Q.conj().T @ Q - np.eye(m))) < 1e-13
    assert np.max(np.abs(np.tril(A, -2))) == 0.0  # upper Hessenberg
    Hn = np.linalg.norm(H.toarray(), 2)
    assert np.max(np.abs(H @ Q[:, :m - 1] - Q @ A[:, :m - 1])) < 1e-12 * Hn
    assert np.max(np.abs(A - Q.conj().T @ (H @ Q))) < 1e-12 * Hn
    assert np.allclose(Q[:, 0], psi / np.linalg.norm(psi), atol=1e-15)


def test_krylov_tridiagonal_when_anti_hermitian_part_is_scalar():
    s = _system(1 / GE, 1 / GE)  # kS = kT: H = Hermitian - i k/2
    h = pyoracle.build_general_csr(s)
    rng = np.random.default_rng(1)
    psi = rng.normal(size=h.n) + 1j * rng.normal(size=h.n)
    _, _, A = pyoracle.get_krylov(h, psi, 16)
    assert np.max(np.abs(np.triu(A, 2))) < 1e-12
    assert np.max(np.abs(np.diag(A).imag + 0.5 / GE)) < 1e-13
