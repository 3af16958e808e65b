"""GPU: the symmetry-reduced driver (symmetry.f90:16-246) on the device.

Pins: the block Hamiltonian (spc_system_set_block: multiplets of any
multiplicity, 1H nuclear Zeeman) equal to the dense textbook model (1e-14);
symmetrised_dynamics equal to the oracle's per-block trace sampling with the
same block seeds (1e-10 at tol = 0, the yield tolerance; 1e-9 at the
reference input's tol, where a rebuild decision may flip at rounding level);
the exact decomposition of a system of equivalent nuclei into its blocks on
the device (exact traces, 1e-10); the amdflang driver's symmetry_dynamics
output (per-block folders, S_yield.data) equal to the Python mirror."""
import copy
import os
import subprocess

import numpy as np
import pytest

import dense_model as dm
from helpers import max_rel
from oracle import pyoracle
from spin_dynamics_amd import dynamics, symmetry, workloads
from spin_dynamics_amd.spinsys import Electron, GAMMA_E

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "spin_dynamics_amd", "bin", "spinchem")
KEYS = ("P_S", "P_T0", "P_Tp", "P_Tm")


def test_block_hamiltonian_matches_dense(ctx):
    w = workloads.small_anisotropic()
    sys = copy.deepcopy(w.sys)
    m1, m2 = [5, 1, 3], [4, 2]  # the multiplicity-1 set drops out
    a = [0.7, 0.3, -0.4, 0.5, 0.2]
    sys.e1 = Electron(g_I=m1, A=np.array([x * np.eye(3) for x in a[:3]]))
    sys.e2 = Electron(g_I=m2, A=np.array([x * np.eye(3) for x in a[3:]]))
    sys.J, sys.field_dir = 0.05, (0.0, 0.0, 1.0)
    B = 0.6
    H = dynamics.Hamiltonian(sys, B, ctx, block=True)
    keep = [0, 2, 3, 4]
    mults = [5, 3, 4, 2]
    Hd = dm.hamiltonian([5, 3], [4, 2], [a[j] * np.eye(3) for j in keep], B=B, J=sys.J, D=sys.D, kS=sys.kS,
                        kT=sys.kT, gammas=[dm.GAMMA_1H] * 4)
    assert H.N == 4 * int(np.prod(mults))
    assert max_rel(H.scipy().toarray(), Hd) < 1e-14
    # getHiso_sp's structure for any multiplicity (hamiltonian.f90:62-73)
    N = H.N
    assert H.nnz == N + N // 2 + sum((g - 1) * (N // g) for g in mults)


def _oracle_sym(w, T_ns, tol):
    s, sim = w.sys, w.sim
    return pyoracle.symmetrised_dynamics([x for x in s.e1.a_iso], [x for x in s.e2.a_iso], sim.M1, sim.M2, J=s.J,
                                         kS=s.kS, kT=s.kT, B=w.B, m=sim.N_krylov, t_end=T_ns * GAMMA_E / 1000.0,
                                         dt=sim.dt, N_samples=sim.N_samples, master_seed=sim.seed,
                                         block_tol=sim.block_tol, tol=tol, vec_kind=0)


@pytest.mark.parametrize("case", ["reference_input", "two_electrons_field"])
@pytest.mark.parametrize("tol", [0.0, 1e-5])
def test_symmetrised_dynamics_matches_oracle(ctx, case, tol):
    if case == "reference_input":  # input.ini: 8 protons on electron 1, M1 = 4, B = 0
        w = workloads.symmetry_protons(T_ns=30.0, tol=tol)
    else:  # nuclei on both electrons, J, a field: 1H Zeeman on every multiplet
        w = workloads.symmetry_protons(a1=(0.9, 0.4, -0.3, 0.6, 0.2, 0.5), a2=(0.7, -0.2, 0.3, 0.25), M1=3, M2=2,
                                       B=0.8, J=0.05, m=10, T_ns=30.0, tol=tol)
    sim = copy.deepcopy(w.sim)
    sim.t_end = 30.0 * GAMMA_E / 1000.0
    res = symmetry.symmetrised_dynamics(w.sys, sim, B=w.B, ctx=ctx)
    ref, blocks, _ = _oracle_sym(w, 30.0, tol)
    assert abs(sum(b[3] for b in blocks) - 1.0) < 1e-14
    bound = 1e-10 if tol == 0.0 else 1e-9
    for k in KEYS:
        assert max_rel(getattr(res, k), ref[k]) < bound, k
    ob = pyoracle.get_kinetics(ref["P_S"], ref["P_T0"], ref["P_Tp"], ref["P_Tm"], sim.dt, w.sys.kS, w.sys.kT)
    assert abs(res.Phi_S[-1] - ob[1][-1]) <= bound * abs(ob[1][-1])


def test_exact_decomposition_on_the_device(ctx):
    # exactly equivalent sets: the full system's exact trace (Z = 128) equals
    # the weighted sum of its blocks' exact traces, every kernel on the device.
    # The field is tilted: along z, total F_z and the permutation symmetry of
    # each set close the Krylov space of some basis states within 5 vectors
    # (a breakdown); m = 8 is the smallest block's dimension (oracle model of
    # this case: the decomposition holds to 7e-14)
    sets1, sets2 = [(0.8, 3), (-0.45, 2)], [(0.5, 2)]
    a1 = [a for a, n in sets1 for _ in range(n)]
    a2 = [a for a, n in sets2 for _ in range(n)]
    w = workloads.symmetry_protons(a1=a1, a2=a2, M1=2, M2=1, B=0.6, J=0.07, kS_us=2.0, kT_us=0.5, m=8,
                                   T_ns=20.0, dt_ns=0.5, tol=0.0)
    w.sys.field_dir = (0.6, 0.0, 0.8)
    full = dynamics.exact_dynamics(w.sys, w.sim, B=w.B, ctx=ctx)
    n_bar = [n for _, n in sets1 + sets2]
    a_bar = [a for a, _ in sets1 + sets2]
    K, Zb, W = symmetry.symmetry_blocks(n_bar)
    tot = {k: 0.0 for k in KEYS}
    for k, z, c in zip(K, Zb, W):
        red = symmetry.reduce_system(w.sys, k[:2], a_bar[:2], k[2:], a_bar[2:])
        r = dynamics.exact_dynamics(red, w.sim, B=w.B, ctx=ctx, block=True)
        for key in KEYS:
            tot[key] = tot[key] + getattr(r, key) * (float(c * z) / 2 ** 7)
    for key in KEYS:
        assert max_rel(tot[key], getattr(full, key)) < 1e-10, key


def test_driver_symmetry_dynamics(tmp_path):
    w = workloads.symmetry_protons(T_ns=30.0, tol=0.0)
    out = tmp_path / "out"
    out.mkdir()
    ini = tmp_path / "input.ini"
    workloads.write_ini(w, str(ini), str(out), simulation_type="symmetry_dynamics", fields=[0.0])
    r = subprocess.run([DRIVER, str(ini)], capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "starting trace sampling of hamiltonian 1/" in r.stdout
    folder = out / ".000"
    got = {k: np.loadtxt(folder / f"{k}.data") for k in ("S_prob", "T0_prob", "S_yield", "T_yield", "total_prob")}
    blocks = symmetry.blocks_of(w.sys, w.sim)
    for i, b in enumerate(blocks):
        assert (folder / f"hamiltonian_{i + 1}" / "S_prob.data").exists() == (b["weight"] >= w.sim.block_tol)
    sim = copy.deepcopy(w.sim)
    res = symmetry.symmetrised_dynamics(w.sys, sim, B=0.0)
    assert max_rel(got["S_prob"], res.P_S) < 1e-12 and max_rel(got["T0_prob"], res.P_T0) < 1e-12
    assert len(got["S_yield"]) == len(res.P_S) // 2
    assert abs(got["S_yield"][-1] - res.Phi_S[-1]) <= 1e-12 * abs(res.Phi_S[-1])
