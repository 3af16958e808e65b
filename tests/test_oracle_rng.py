"""The oracle's xoroshiro128+ stream against the reference generator
(m_random.f90), compiled from the reference sources: committed fixture
tests/golden/ref_rng.json, and the live library when oracle/_ref is built."""
import ctypes
import json
import os

import numpy as np
import pytest

from oracle import pyoracle
from spin_dynamics_amd import workloads

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "ref_rng.json")))


def _rng(seed):
    r = pyoracle.OrcRng()
    pyoracle.lib().orc_rng_init_default(ctypes.byref(r))
    pyoracle.lib().orc_rng_set_seed(ctypes.byref(r), (ctypes.c_int64 * 2)(*seed))
    return r


@pytest.mark.parametrize("case", range(len(GOLD["seeds"])))
def test_stream_matches_reference_fixture(case):
    seed = GOLD["seeds"][case]
    L = pyoracle.lib()
    r = _rng(seed)
    assert [L.orc_rng_next(ctypes.byref(r)) for _ in range(64)] == GOLD["next"][case]
    r = _rng(seed)
    assert [float(L.orc_rng_unif01(ctypes.byref(r))).hex() for _ in range(64)] == GOLD["unif01"][case]
    r = _rng(seed)
    assert [float(L.orc_rng_normal(ctypes.byref(r))).hex() for _ in range(64)] == GOLD["normal"][case]
    for key, fn, w in (("circle", L.orc_rng_circle, 2), ("sphere", L.orc_rng_sphere, 3)):
        r = _rng(seed)
        got = []
        for _ in range(64):
            o = (ctypes.c_double * w)()
            fn(ctypes.byref(r), ctypes.c_double(1.0), o)
            got += [float(x).hex() for x in o]
        assert got == GOLD[key][case]


def test_stream_matches_compiled_reference_live():
    ref = pyoracle.ref_lib()
    if ref is None:
        pytest.skip("oracle/_ref not built (needs the reference sources)")
    for seed in [(42, 99), (7, -3), (2**62, 5)]:
        out = (ctypes.c_int64 * 200)()
        ref.ref_rng_next((ctypes.c_int64 * 2)(*seed), ctypes.c_int(200), out)
        r = _rng(seed)
        assert [pyoracle.lib().orc_rng_next(ctypes.byref(r)) for _ in range(200)] == list(out)


def test_python_generator_matches_oracle():
    x = workloads.Xoroshiro((42, 99))
    r = _rng((42, 99))
    L = pyoracle.lib()
    for _ in range(20):
        assert x.unif01() == L.orc_rng_unif01(ctypes.byref(r))
    for _ in range(20):
        assert x.normal() == L.orc_rng_normal(ctypes.byref(r))


def test_trace_vectors_statistics():
    Z = 2000
    seeds = pyoracle.sample_seeds((42, 99), 8)
    for kind in (0, 1):
        vecs = np.array([pyoracle.sample_vector(s, Z, kind) for s in seeds])
        assert np.allclose(np.linalg.norm(vecs, axis=1), 1.0, atol=1e-14)
        if kind == 1:  # random phase: |z_n|^2 = 1/Z exactly
            assert np.allclose(np.abs(vecs) ** 2, 1.0 / Z, rtol=1e-12)
        else:  # SU(Z): E|z_n|^2 = 1/Z, fluctuating
            m = np.mean(np.abs(vecs) ** 2) * Z
            assert abs(m - 1.0) < 1e-12  # normalised per vector
            assert np.std(np.abs(vecs) ** 2 * Z) > 0.5
    # normals: mean 0, variance 1
    r = _rng((1, 2))
    v = np.array([pyoracle.lib().orc_rng_normal(ctypes.byref(r)) for _ in range(20000)])
    assert abs(v.mean()) < 0.03 and abs(v.var() - 1) < 0.05


def test_sample_seeds_are_master_stream_pairs():
    seeds = pyoracle.sample_seeds((42, 99), 5)
    r = _rng((42, 99))
    L = pyoracle.lib()
    flat = [L.orc_rng_next(ctypes.byref(r)) for _ in range(10)]
    assert seeds.reshape(-1).tolist() == flat
