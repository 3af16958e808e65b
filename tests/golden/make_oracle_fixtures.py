"""Writes oracle outputs so GPU parity tests compare against them without
re-running the oracle (the oracle itself is pinned by the reference fixtures
and dense checks in tests/):

- iso192: the CPU-runnable workload (isotropic, N = 192, M = 32, m = 20,
  dt = 1 ns, T = 5 us, tol = 0) with both trace-vector kinds;
- aniso36864: the N = 36 864 anisotropic workload (2x14N + 6x1H | 4x1H,
  B = 0.05 mT at 37 deg, J = 0.01 mT, kS = 2, kT = 0.5 us^-1, m = 24,
  dt = 1 ns, tol = 0) for its first 16 random-phase trace vectors (master
  seed (42, 99)) over T = 500 ns (501 steps; about 8 minutes on 8 threads).

    python tests/golden/make_oracle_fixtures.py [iso192] [aniso36864]
"""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))
from oracle import pyoracle  # noqa: E402
from spin_dynamics_amd import workloads  # noqa: E402
from helpers import oracle_system  # noqa: E402


def iso192():
    for kind in ("phase", "SUZ"):
        w = workloads.isotropic_192(vec_kind=kind)
        s = oracle_system(w.sys, w.B)
        h = pyoracle.build_iso_csr(s)
        t0 = time.time()
        res = pyoracle.trace_sampling(h, s.Z, M=w.sim.N_samples, m=w.sim.N_krylov, t_end=w.sim.t_end, dt=w.sim.dt,
                                      tol=0.0, vec_kind=1 if kind == "phase" else 0, seed=w.sim.seed)
        iden, phs, pht = pyoracle.get_kinetics(res["P_S"], res["P_T0"], res["P_Tp"], res["P_Tm"], w.sim.dt,
                                               s.kS, s.kT)
        np.savez_compressed(os.path.join(HERE, f"oracle_iso192_{kind}.npz"), **res, Phi_S=phs, Phi_T=pht)
        print(kind, f"{time.time() - t0:.1f}s", "Phi_S", phs[-1])


def aniso36864(M=16, T_ns=500.0):
    w = workloads.anisotropic_36864(M=M, T_ns=T_ns)
    s = oracle_system(w.sys, w.B)
    h = pyoracle.build_general_csr(s)
    t0 = time.time()
    res = pyoracle.trace_sampling(h, s.Z, M=M, m=w.sim.N_krylov, t_end=w.sim.t_end, dt=w.sim.dt, tol=0.0, vec_kind=1,
                                  seed=w.sim.seed)
    iden, phs, pht = pyoracle.get_kinetics(res["P_S"], res["P_T0"], res["P_Tp"], res["P_Tm"], w.sim.dt, s.kS, s.kT)
    np.savez_compressed(os.path.join(HERE, "oracle_aniso36864_m16_500ns.npz"), **res, Phi_S=phs, Phi_T=pht,
                        M=M, T_ns=T_ns)
    print("aniso36864", f"{time.time() - t0:.1f}s", "Phi_S", phs[-1])


if __name__ == "__main__":
    which = sys.argv[1:] or ["iso192", "aniso36864"]
    for name in which:
        {"iso192": iso192, "aniso36864": aniso36864}[name]()
