"""Writes the fixtures pinned by the compiled reference modules.

Runs oracle/_ref/libref_spinchem.so (m_random.f90 and observables.f90 of the
reference, compiled by `make -C oracle ref`) in this container and stores its
outputs; the fixtures travel, the reference does not.

    python tests/golden/make_ref_vectors.py
"""
import ctypes
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
from oracle import pyoracle  # noqa: E402


def main():
    ref = pyoracle.ref_lib()
    if ref is None:
        raise SystemExit("oracle/_ref/libref_spinchem.so is not built (make -C oracle ref)")
    out = {"seeds": [], "next": [], "unif01": [], "normal": [], "circle": [], "sphere": []}
    for seed in [(42, 99), (1, 2), (-4707382666127344949, 123456789), (0, 1)]:
        s = (ctypes.c_int64 * 2)(*seed)
        n = 64
        nx = (ctypes.c_int64 * n)()
        ref.ref_rng_next(s, ctypes.c_int(n), nx)
        out["seeds"].append(list(seed))
        out["next"].append([int(v) for v in nx])
        for kind, key, w in [(0, "unif01", 1), (1, "normal", 1), (2, "circle", 2), (3, "sphere", 3)]:
            buf = (ctypes.c_double * (n * w))()
            ref.ref_rng_draw(s, ctypes.c_int(kind), ctypes.c_int(n), buf)
            out[key].append([float(v).hex() for v in buf])
    with open(os.path.join(HERE, "ref_rng.json"), "w") as f:
        json.dump(out, f)

    rng = np.random.default_rng(7)
    cases = []
    for n in [1, 2, 3, 7, 8, 101, 1000]:
        P = [rng.uniform(0, 1, n) for _ in range(4)]
        dt, kS, kT = 0.17608596, 2.0 / 176.08596, 0.5 / 176.08596
        iden = (ctypes.c_double * n)()
        phs = (ctypes.c_double * max(n // 2, 1))()
        pht = (ctypes.c_double * max(n // 2, 1))()
        arrs = [(ctypes.c_double * n)(*p) for p in P]
        ref.ref_get_kinetics(ctypes.c_int(n), *arrs, ctypes.c_double(dt), ctypes.c_double(kS), ctypes.c_double(kT),
                             iden, phs, pht)
        cases.append({"n": n, "P": [[float(x).hex() for x in p] for p in P], "dt": dt, "kS": kS, "kT": kT,
                      "iden": [float(x).hex() for x in iden], "Phi_S": [float(x).hex() for x in phs[:n // 2]],
                      "Phi_T": [float(x).hex() for x in pht[:n // 2]]})
    with open(os.path.join(HERE, "ref_kinetics.json"), "w") as f:
        json.dump(cases, f)
    print("wrote ref_rng.json, ref_kinetics.json")


if __name__ == "__main__":
    main()
