"""Dense textbook radical-pair Hamiltonian built with numpy Kronecker products.

Independent of the oracle and of the product: it is written from the physics,
H = -B n.(S1+S2) + sum_k [S_e(k).A_k.I_k + (B gamma_k/gamma_e) n.I_k]
    + S1.T.S2 - i K/2,   T = 2J 1 + D,   K = kS P_S + kT P_T,
in the basis electron1 (x) electron2 (x) nuclei (electron-1 nuclei first,
m = +I first).  Used to pin the sparse builders and, with scipy's expm, the
propagation.
"""
from __future__ import annotations

import numpy as np

GAMMA_E = 176.08596
GAMMA_1H = 0.26752218708
GAMMA_14N = 0.0193297792


def spin_ops(g: int):
    S = (g - 1) / 2
    m = S - np.arange(g)
    Sz = np.diag(m).astype(complex)
    Sp = np.zeros((g, g), complex)
    for i in range(1, g):
        Sp[i - 1, i] = np.sqrt(S * (S + 1) - m[i] * (m[i] + 1))
    Sm = Sp.conj().T
    Sx = (Sp + Sm) / 2
    Sy = (Sp - Sm) / 2j
    return Sx, Sy, Sz


def embed(op, site, dims):
    out = np.array([[1.0 + 0j]])
    for i, d in enumerate(dims):
        out = np.kron(out, op if i == site else np.eye(d))
    return out


def hamiltonian(mult1, mult2, A, *, B=0.0, direction=(0, 0, 1), J=0.0, D=None, kS=0.0, kT=0.0,
                nuclear_zeeman=True, gammas=None):
    """A: (n1+n2, 3, 3) hyperfine tensors A[k][a][b] (mT); kS/kT in mT units;
    gammas: per-nucleus gyromagnetic ratios (default 14N's for multiplicity 3,
    else 1H's)."""
    mult = list(mult1) + list(mult2)
    dims = [2, 2] + mult
    N = int(np.prod(dims))
    n = np.asarray(direction, float)
    D = np.zeros((3, 3)) if D is None else np.asarray(D, float)
    S1 = [embed(o, 0, dims) for o in spin_ops(2)]
    S2 = [embed(o, 1, dims) for o in spin_ops(2)]
    H = np.zeros((N, N), complex)
    for a in range(3):
        H += -B * n[a] * (S1[a] + S2[a])
    for k, g in enumerate(mult):
        Ik = [embed(o, 2 + k, dims) for o in spin_ops(g)]
        Se = S1 if k < len(mult1) else S2
        for a in range(3):
            for b in range(3):
                H += A[k][a][b] * Se[a] @ Ik[b]
        if nuclear_zeeman:
            gam = gammas[k] if gammas is not None else (GAMMA_14N if g == 3 else GAMMA_1H)
            for b in range(3):
                H += (B * gam / GAMMA_E) * n[b] * Ik[b]
    T = 2 * J * np.eye(3) + D
    for a in range(3):
        for b in range(3):
            H += T[a][b] * S1[a] @ S2[b]
    S1S2 = sum(S1[a] @ S2[a] for a in range(3))
    PS = 0.25 * np.eye(N) - S1S2
    PT = np.eye(N) - PS
    H += -0.5j * (kS * PS + kT * PT)
    return H


def singlet_projector_blocks(psi, Z):
    """P_S, P_T0, P_Tp, P_Tm expectation values of a state (electron-major)."""
    b = psi.reshape(4, Z)
    pTp = np.vdot(b[0], b[0]).real
    pTm = np.vdot(b[3], b[3]).real
    pS = 0.5 * np.vdot(b[1] - b[2], b[1] - b[2]).real
    pT0 = 0.5 * np.vdot(b[1] + b[2], b[1] + b[2]).real
    return pS, pT0, pTp, pTm


def theta(init_state="singlet"):
    r = 1 / np.sqrt(2)
    return {"singlet": np.array([0, r, -r, 0], complex), "T0": np.array([0, r, r, 0], complex),
            "Tp": np.array([1, 0, 0, 0], complex), "Tm": np.array([0, 0, 0, 1], complex)}[init_state]


def propagate_dense(H, psi0, dt, n_steps):
    """Observables of psi_k = expm(-i H dt)^k psi0, k = 0..n_steps."""
    from scipy.linalg import expm
    U = expm(-1j * dt * H)
    Z = H.shape[0] // 4
    out = np.zeros((4, n_steps + 1))
    psi = psi0.copy()
    out[:, 0] = singlet_projector_blocks(psi, Z)
    for k in range(1, n_steps + 1):
        psi = U @ psi
        out[:, k] = singlet_projector_blocks(psi, Z)
    return out
