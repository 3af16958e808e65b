"""GPU: spc_trace_sample_noalloc, the entry point that allocates no device
memory (every buffer from the library's per-GPU scratch area), through the
C-ABI against spc_trace_sample on the same inputs: bit for bit (the same
kernels in the same order on one chunk), on consecutive calls of which the
later ones are smaller (the scratch area keeps the earlier call's contents),
with sampled and injected vectors and with tol > 0; a workspace beyond the
scratch area is refused with SPC_ERR_UNSUPPORTED before anything runs."""
import copy

import numpy as np
import pytest

from oracle import pyoracle
from spin_dynamics_amd import _lib, dynamics, workloads
from spin_dynamics_amd.spinsys import to_c_system

pytestmark = pytest.mark.gpu


def _call(ctx, fn, w, M, T_ns, tol=0.0, vecs=None):
    lib = ctx.lib
    sim = copy.deepcopy(w.sim)
    sim.t_end = T_ns * 176.08596 / 1000.0
    kind = _lib.VEC_INJECTED if vecs is not None else _lib.VEC_PHASE
    c = _lib.SimParams(M, sim.N_krylov, sim.t_end, sim.dt, tol, 0, kind)
    n_out = int(lib.spc_n_steps(sim.t_end, sim.dt)) + 1
    outs = [np.zeros(n_out) for _ in range(4)]
    phis, phit = np.zeros(n_out // 2), np.zeros(n_out // 2)
    seed = np.array(w.sim.seed, dtype=np.int64)
    tv = None if vecs is None else _lib.ptr(np.ascontiguousarray(vecs, dtype=np.complex128))
    rc = fn(ctx.h, c, _lib.ptr(seed, _lib._i64p), tv, *[_lib.ptr(o) for o in outs], n_out, _lib.ptr(phis),
            _lib.ptr(phit))
    return rc, outs + [phis, phit]


def _set(ctx, w):
    s, keep = to_c_system(w.sys, w.B)
    ctx.check(ctx.lib.spc_system_set(ctx.h, s))
    return keep


def _same(ctx, w, M, T_ns, **kw):
    rc, ref = _call(ctx, ctx.lib.spc_trace_sample, w, M, T_ns, **kw)
    assert rc == 0, ctx.lib.spc_last_error(ctx.h)
    rc, got = _call(ctx, ctx.lib.spc_trace_sample_noalloc, w, M, T_ns, **kw)
    assert rc == 0, ctx.lib.spc_last_error(ctx.h)
    for a, b in zip(got, ref):
        assert np.array_equal(a, b)
    assert abs(got[0][0] - 1.0) == 0.0 and got[4][-1] > 0.0


def test_noalloc_equals_trace_sample_on_consecutive_smaller_calls(ctx):
    w = workloads.isotropic_192()  # N = 192: the CPU-runnable workload
    keep = _set(ctx, w)
    _same(ctx, w, 32, 1000.0)           # M = 32, 1001 steps
    _same(ctx, w, 8, 200.0)             # then smaller: M = 8, 201 steps
    _same(ctx, w, 8, 120.0, tol=1e-8)   # several steps per basis (projector Grams)
    ws = workloads.small_anisotropic(M=8)  # a smaller system (N = 48)
    keep2 = _set(ctx, ws)
    vecs = np.array([pyoracle.sample_vector(sd, 12, 1) for sd in pyoracle.sample_seeds((3, 5), 5)])
    _same(ctx, ws, 5, 100.0, vecs=vecs)
    del keep, keep2


def test_noalloc_full_cpu_workload_then_smaller(ctx):
    w = workloads.isotropic_192()  # M = 32, m = 20, T = 5 us: about 20 MiB of workspace
    keep = _set(ctx, w)
    _same(ctx, w, 32, 5000.0)
    _same(ctx, w, 16, 500.0)
    del keep


def test_noalloc_refuses_a_workspace_beyond_the_scratch_area(ctx):
    w = workloads.anisotropic_36864(M=64)
    keep = _set(ctx, w)
    rc, _ = _call(ctx, ctx.lib.spc_trace_sample_noalloc, w, 64, 1.5)
    assert rc == _lib.SPC_ERR_UNSUPPORTED
    assert b"scratch area" in ctx.lib.spc_last_error(ctx.h)
    del keep
