"""Writes oracle outputs for the CPU-runnable workload (isotropic, N = 192,
M = 32, m = 20, dt = 1 ns, T = 5 us, tol = 0) with both trace-vector kinds, so
GPU parity tests compare against them without re-running the oracle.  The
oracle itself is pinned by the reference fixtures and dense checks in tests/.

    python tests/golden/make_oracle_fixtures.py
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


def main():
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


if __name__ == "__main__":
    main()
