"""The amdflang driver `spinchem input.ini` (app/main.f90 semantics) on the
GPU: output files against the oracle."""
import ctypes
import os
import subprocess

import numpy as np
import pytest

from helpers import oracle_system
from oracle import pyoracle
from spin_dynamics_amd import workloads

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "spin_dynamics_amd", "bin", "spinchem")


def run_driver(w, tmp_path, sim_type="trace_sampling", fields=None, output_format=None):
    out = tmp_path / "out"
    out.mkdir(exist_ok=True)
    ini = tmp_path / "input.ini"
    workloads.write_ini(w, str(ini), str(out), simulation_type=sim_type, fields=fields, output_format=output_format)
    r = subprocess.run([DRIVER, str(ini)], capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stdout + r.stderr
    return out


def load(folder):
    return {k: np.loadtxt(os.path.join(folder, f"{k}.data")) for k in
            ("S_prob", "T0_prob", "Tp_prob", "Tm_prob", "S_yield", "T_yield", "total_prob")}


def test_driver_isotropic_192_matches_oracle(tmp_path):
    w = workloads.isotropic_192(vec_kind="SUZ")
    out = run_driver(w, tmp_path, fields=[0.05])
    got = load(out / ".050")  # F0.3 of 0.05, as the reference names the folder
    ref = np.load(os.path.join(ROOT, "tests", "golden", "oracle_iso192_SUZ.npz"))
    assert len(got["S_prob"]) == pyoracle.n_steps(w.sim.t_end, w.sim.dt) + 1
    for a, b in (("S_prob", "P_S"), ("T0_prob", "P_T0"), ("Tp_prob", "P_Tp"), ("Tm_prob", "P_Tm")):
        assert np.max(np.abs(got[a] - ref[b])) < 1e-10, a
    assert abs(got["S_yield"][-1] - ref["Phi_S"][-1]) < 1e-10 * ref["Phi_S"][-1]


def test_driver_field_loop_and_serial_sampling(tmp_path):
    w = workloads.small_anisotropic(M=6, T_ns=120.0, vec_kind="SUZ")
    out = run_driver(w, tmp_path, sim_type="trace_sampling_serial", fields=[0.05, 0.2])
    for B, folder in ((0.05, ".050"), (0.2, ".200")):
        s = oracle_system(w.sys, B)
        # trace_sampling (qm_dynamics.f90:139-255): all vectors from the master
        # stream, set_seed(seed) once per field (main.f90:50)
        r = pyoracle.OrcRng()
        L = pyoracle.lib()
        L.orc_rng_init_default(ctypes.byref(r))
        L.orc_rng_set_seed(ctypes.byref(r), (ctypes.c_int64 * 2)(*w.sim.seed))
        vecs = np.zeros((6, s.Z), complex)
        for i in range(6):
            L.orc_sample_SUZ(ctypes.byref(r), ctypes.c_int64(s.Z), vecs[i].ctypes.data_as(pyoracle._c_double_p))
        ref = pyoracle.trace_sampling(pyoracle.build_general_csr(s), s.Z, M=6, m=w.sim.N_krylov, t_end=w.sim.t_end,
                                      dt=w.sim.dt, vec_kind=2, injected=vecs)
        got = load(out / folder)
        assert np.max(np.abs(got["S_prob"] - ref["P_S"])) < 1e-10


def test_driver_exact_dynamics(tmp_path):
    w = workloads.small_anisotropic(m1=(2,), m2=(3,), T_ns=80.0)
    out = run_driver(w, tmp_path, sim_type="exact_dynamics")
    s = oracle_system(w.sys, w.B)
    ref = pyoracle.exact_dynamics(pyoracle.build_general_csr(s), s.Z, m=w.sim.N_krylov, t_end=w.sim.t_end,
                                  dt=w.sim.dt)
    got = load(out / ".050")
    assert np.max(np.abs(got["S_prob"] - ref["P_S"])) < 1e-10
    assert np.max(np.abs(got["total_prob"] - sum(ref[k] for k in ("P_S", "P_T0", "P_Tp", "P_Tm")))) < 1e-10


def test_driver_semi_classical_matches_oracle(tmp_path):
    # simulation_type = semi_classical (unreachable in the reference driver):
    # isotropic parts of the tensors, N_samples classical samples from the seed
    w = workloads.isotropic_192(M=3000, T_ns=20.0)
    out = run_driver(w, tmp_path, sim_type="semi_classical", fields=[0.05])
    folder = out / ".050"
    got = {k: np.loadtxt(os.path.join(folder, f"{k}.data")) for k in ("S_prob", "T_prob", "S_yield", "T_yield")}
    a1 = [np.trace(a) / 3 for a in w.sys.e1.A]
    a2 = [np.trace(a) / 3 for a in w.sys.e2.A]
    n_steps = pyoracle.n_steps(w.sim.t_end, w.sim.dt)
    ps, pt = pyoracle.semiclassical(a1, a2, w.sys.e1.g_I, w.sys.e2.g_I, B=0.05, J=w.sys.J, kS=w.sys.kS, kT=w.sys.kT,
                                    n_samples=3000, dt=w.sim.dt, n_steps_=n_steps, seed=w.sim.seed, with_pt=True)
    assert len(got["S_prob"]) == n_steps + 1
    assert np.max(np.abs(got["S_prob"] - ps)) < 1e-10 * np.max(np.abs(ps))
    assert np.max(np.abs(got["T_prob"] - pt)) < 1e-10 * max(np.max(np.abs(pt)), 1e-3)
    _, phis, phit = pyoracle.get_kinetics(ps, pt, np.zeros_like(ps), np.zeros_like(ps), w.sim.dt, w.sys.kS, w.sys.kT)
    assert abs(got["S_yield"][-1] - phis[-1]) <= 1e-10 * abs(phis[-1])


def test_driver_gfortran_output_and_parse_output(tmp_path):
    # output_format = gfortran: every line as gfortran's list-directed output of
    # the value it holds, and the reference's reader (example/utils.py:420-456)
    # reads the same traces as from the default output
    from spin_dynamics_amd.observables import gfortran_real
    from test_output_format import parse_output
    w = workloads.small_anisotropic(M=6, T_ns=60.0, vec_kind="SUZ")
    (tmp_path / "a").mkdir()
    (tmp_path / "b").mkdir()
    a = run_driver(w, tmp_path / "a", fields=[0.05])
    b = run_driver(w, tmp_path / "b", fields=[0.05], output_format="gfortran")
    for name in ("S_prob.data", "T0_prob.data", "S_yield.data", "total_prob.data"):
        lines = (b / ".050" / name).read_text().splitlines()
        assert lines and all(line == gfortran_real(float(line)) for line in lines), name
    pa, pb = parse_output(a, ".050", 1e-9), parse_output(b, ".050", 1e-9)
    assert sorted(pa) == sorted(pb) == ["S_prob", "T0_prob", "Tm_prob", "Tp_prob", "time", "total_prob"]
    for k in pa:
        assert np.array_equal(pa[k], pb[k]), k
