"""The synthetic radical-pair systems the benchmarks and parity tests run.

Every random number comes from the reference generator (xoroshiro128+,
m_random.f90) seeded with set_seed(SYSTEM_SEED) and drawn in the order
documented per function, so a system is fully fixed by its name; write_ini()
turns one into a .ini file the driver reads with the reference's keys.

Units in the returned objects are internal (mT, rates / gamma_e, times in
rad/mT); the .ini holds the reference's input units (mT, us^-1, ns).
"""
from __future__ import annotations

import math
from dataclasses import dataclass
from typing import List, Sequence

import numpy as np

from .spinsys import GAMMA_E, Electron, SimParam, SysParam

SYSTEM_SEED = (12345, 6789)


class Xoroshiro:
    """xoroshiro128+ with set_seed/jump/unif_01/normal (m_random.f90:90-198)."""

    MASK = (1 << 64) - 1

    def __init__(self, seed=SYSTEM_SEED):
        self.s = [seed[0] & self.MASK, seed[1] & self.MASK]
        self.stored = None
        self.jump()

    @staticmethod
    def _rotl(x, k):
        return ((x << k) | (x >> (64 - k))) & Xoroshiro.MASK

    def next(self) -> int:
        a, b = self.s
        r = (a + b) & self.MASK
        b ^= a
        self.s = [self._rotl(a, 55) ^ b ^ ((b << 14) & self.MASK), self._rotl(b, 36)]
        return r

    def jump(self):
        t0 = t1 = 0
        for c in (0xbeac0467eba5facb, 0xd86b048b86aa9922):
            for b in range(64):
                if c & (1 << b):
                    t0 ^= self.s[0]
                    t1 ^= self.s[1]
                self.next()
        self.s = [t0, t1]

    def unif01(self) -> float:
        x = (1023 << 52) | (self.next() >> 12)
        return float(np.array([x], dtype=np.uint64).view(np.float64)[0]) - 1.0

    def normal(self) -> float:
        if self.stored is not None:
            v, self.stored = self.stored, None
            return v
        while True:
            a = 2 * self.unif01() - 1
            b = 2 * self.unif01() - 1
            s = a * a + b * b
            if 0.0 < s < 1.0:
                break
        f = math.sqrt(-2 * math.log(s) / s)
        self.stored = b * f
        return a * f


def random_tensor(rng: Xoroshiro, half_width: float) -> np.ndarray:
    """Symmetric 3x3: eigenvalues U[-w, w] (3 draws), rotation from a unit
    quaternion of 4 normals (uniform on SO(3)), A = R diag(l) R^T."""
    lam = np.array([half_width * (2 * rng.unif01() - 1) for _ in range(3)])
    q = np.array([rng.normal() for _ in range(4)])
    w, x, y, z = q / np.linalg.norm(q)
    R = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                  [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                  [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])
    return R @ np.diag(lam) @ R.T


@dataclass
class Workload:
    name: str
    sys: SysParam
    sim: SimParam
    B: float
    # input-file units for write_ini
    kS_us: float
    kT_us: float
    T_ns: float
    dt_ns: float

    @property
    def N(self) -> int:
        return 4 * self.sys.Z1 * self.sys.Z2


def _electron(mults: Sequence[int], tensors: List[np.ndarray]) -> Electron:
    return Electron(g_I=list(mults), A=np.array(tensors).reshape(-1, 3, 3))


def _sim(M, m, T_ns, dt_ns, tol=0.0, vec_kind="phase", seed=(42, 99)) -> SimParam:
    return SimParam(type="trace_sampling", B=[0.0], init_state="singlet", seed=seed, N_samples=M,
                    t_end=T_ns * GAMMA_E / 1000.0, dt=dt_ns * GAMMA_E / 1000.0, N_krylov=m, tol=tol,
                    vec_kind=vec_kind)


def isotropic_192(M: int = 32, m: int = 20, T_ns: float = 5000.0, vec_kind: str = "phase") -> Workload:
    """Plain trace sampling, CPU-runnable: e1 = 14N + 2x1H, e2 = 2x1H (N = 192);
    isotropic a_k = 2 unif - 1 mT drawn in nucleus order; B = 0.05 mT along z;
    J = D = 0; kS = kT = 1 us^-1; dt = 1 ns; T = 5 us."""
    rng = Xoroshiro()
    m1, m2 = [3, 2, 2], [2, 2]
    a = [2 * rng.unif01() - 1 for _ in range(5)]
    sys = SysParam(e1=_electron(m1, [x * np.eye(3) for x in a[:3]]), e2=_electron(m2, [x * np.eye(3) for x in a[3:]]),
                   J=0.0, kS=1.0 / GAMMA_E, kT=1.0 / GAMMA_E)
    B = 0.05
    sys.e1.w = sys.e2.w = -B
    return Workload("isotropic_192", sys, _sim(M, m, T_ns, 1.0, vec_kind=vec_kind), B, 1.0, 1.0, T_ns, 1.0)


def anisotropic_system(m1: Sequence[int], m2: Sequence[int], *, half_width=1.5, theta_deg=37.0, J=0.01,
                       kS_us=2.0, kT_us=0.5, B=0.05, seed=SYSTEM_SEED) -> SysParam:
    """Tensors drawn with random_tensor in nucleus order (electron 1 first)."""
    rng = Xoroshiro(seed)
    T = [random_tensor(rng, half_width) for _ in range(len(m1) + len(m2))]
    th = math.radians(theta_deg)
    sys = SysParam(e1=_electron(m1, T[:len(m1)]), e2=_electron(m2, T[len(m1):]), J=J, kS=kS_us / GAMMA_E,
                   kT=kT_us / GAMMA_E, field_dir=(math.sin(th), 0.0, math.cos(th)))
    sys.e1.w = sys.e2.w = -B
    return sys


def anisotropic_36864(M: int = 256, m: int = 24, T_ns: float = 10000.0, vec_kind: str = "phase") -> Workload:
    """Plain method at scale: e1 = 2x14N + 6x1H, e2 = 4x1H (N = 36 864); full
    anisotropic tensors (eigenvalues U[-1.5, 1.5] mT, random rotation); B = 0.05
    mT at 37 deg from z in the xz plane; J = 0.01 mT; kS = 2, kT = 0.5 us^-1;
    M = 256, m = 24, dt = 1 ns, T = 10 us, tol = 0."""
    sys = anisotropic_system([3, 3, 2, 2, 2, 2, 2, 2], [2, 2, 2, 2])
    return Workload("anisotropic_36864", sys, _sim(M, m, T_ns, 1.0, vec_kind=vec_kind), 0.05, 2.0, 0.5, T_ns, 1.0)


def small_anisotropic(m1=(3, 2), m2=(2,), M=8, m=12, T_ns=200.0, tol=0.0, vec_kind="phase") -> Workload:
    """Small tilted-field anisotropic system for dense-oracle parity."""
    sys = anisotropic_system(list(m1), list(m2))
    return Workload("small_aniso", sys, _sim(M, m, T_ns, 1.0, tol=tol, vec_kind=vec_kind), 0.05, 2.0, 0.5, T_ns,
                    1.0)


def symmetry_protons(a1=(1.0, 0.3, 0.5, 0.2, 0.1, 0.5, 0.5, 0.5), a2=(), *, M1=4, M2=1, B=0.0, kS_us=20.0,
                     kT_us=5.0, J=0.0, M=10, m=7, T_ns=300.0, dt_ns=0.3, tol=1e-5, block_tol=1e-5,
                     vec_kind="SUZ") -> Workload:
    """The reference's symmetry_dynamics input (input.ini): spin-1/2 nuclei
    with isotropic couplings a1 (electron 1) and a2 (electron 2) in mT,
    reduced to M1 and M2 sets of equivalent nuclei; kS = 20, kT = 5 us^-1,
    N_samples = 10, N_krylov = 7, dt = 0.3 ns, T = 300 ns, tol = 1e-5,
    block_tolerance = 1e-5."""
    e1 = _electron([2] * len(a1), [a * np.eye(3) for a in a1]) if a1 else Electron()
    e2 = _electron([2] * len(a2), [a * np.eye(3) for a in a2]) if a2 else Electron()
    sys = SysParam(e1=e1, e2=e2, J=J, kS=kS_us / GAMMA_E, kT=kT_us / GAMMA_E)
    sys.e1.w = sys.e2.w = -B
    sim = _sim(M, m, T_ns, dt_ns, tol=tol, vec_kind=vec_kind)
    sim.type = "symmetry_dynamics"
    sim.M1, sim.M2, sim.block_tol = M1, M2, block_tol
    return Workload("symmetry_protons", sys, sim, B, kS_us, kT_us, T_ns, dt_ns)


@dataclass
class ScWorkload:
    """A semiclassical system (isotropic couplings) and its run parameters."""
    name: str
    a1: List[float]
    a2: List[float]
    m1: List[int]
    m2: List[int]
    B: float
    direction: tuple
    J: float
    kS: float
    kT: float
    n_samples: int
    dt: float
    n_steps: int

    def kwargs(self) -> dict:
        return dict(B=self.B, direction=self.direction, J=self.J, kS=self.kS, kT=self.kT,
                    n_samples=self.n_samples, dt=self.dt, n_steps_=self.n_steps)


def gamma_like_couplings(rng: Xoroshiro, n: int, lo: float = 0.05, hi: float = 1.2) -> List[float]:
    """n couplings from a gamma(2) spectrum, g = -(ln u1 + ln u2), mapped
    linearly so the smallest is lo and the largest hi (mT); drawn in order."""
    g = [-(math.log(1.0 - rng.unif01()) + math.log(1.0 - rng.unif01())) for _ in range(n)]
    gmin, gmax = min(g), max(g)
    return [lo + (hi - lo) * (x - gmin) / (gmax - gmin) for x in g]


def semiclassical_48(n_samples: int = 1 << 20, T_ns: float = 2000.0, dt_ns: float = 0.5, n_1h: int = 20,
            n_14n: int = 4) -> ScWorkload:
    """Semiclassical Monte Carlo at large nuclear counts: per radical 20 1H
    then 4 14N, couplings from gamma_like_couplings (radical 1 then radical 2);
    1H couplings negative where a further unif_01 < 1/2.  B = 0.05 mT along z,
    J = 0, kS = 2, kT = 0.5 us^-1 (the plain-method rates), dt = 0.5 ns, T = 2 us,
    N_steps = ceiling((T + dt)/dt) as the reference's time loop."""
    rng = Xoroshiro()
    a, m = [], []
    for _ in range(2):
        c = gamma_like_couplings(rng, n_1h + n_14n)
        sign = [(-1.0 if rng.unif01() < 0.5 else 1.0) for _ in range(n_1h)] + [1.0] * n_14n
        a.append([s * x for s, x in zip(sign, c)])
        m.append([2] * n_1h + [3] * n_14n)
    dt = dt_ns * GAMMA_E / 1000.0
    t_end = T_ns * GAMMA_E / 1000.0
    return ScWorkload("semiclassical_48", a[0], a[1], m[0], m[1], 0.05, (0.0, 0.0, 1.0), 0.0, 2.0 / GAMMA_E,
                      0.5 / GAMMA_E, n_samples, dt, math.ceil((t_end + dt) / dt))


def write_ini(w: Workload, path: str, output_folder: str, *, simulation_type: str = "trace_sampling",
              fields: Sequence[float] = None, output_format: str = None):
    """Writes the workload with the reference's keys (spinsys.f90:60-235);
    field_direction, trace_vectors and output_format are extension keys."""
    s, sim = w.sys, w.sim
    fmt = lambda v: " ".join(repr(float(x)) for x in v)
    lines = ["[system variables]", f"J = {s.J!r}", f"D = {fmt(np.asarray(s.D).reshape(3, 3).T.reshape(-1))}",
             f"kS = {w.kS_us!r}", f"kT = {w.kT_us!r}"]
    for n, e in ((1, s.e1), (2, s.e2)):
        lines.append(f"[electron {n}]")
        lines.append("g = 2.0023")
        if e.g_I:
            groups = []
            for g in e.g_I:
                if groups and groups[-1][0] == g:
                    groups[-1][1] += 1
                else:
                    groups.append([g, 1])
            lines.append("I = " + " ".join(str(g) for g, _ in groups))
            lines.append("N_I = " + " ".join(str(c) for _, c in groups))
            for k, A in enumerate(e.A):
                lines.append(f"A{k + 1} = {fmt(np.asarray(A).T.reshape(-1))}")
    lines += ["[simulation parameters]", f"simulation_type = {simulation_type}", f"output_folder = {output_folder}",
              f"seed = {sim.seed[0]} {sim.seed[1]}", f"B = {fmt(fields or [w.B])}", f"initial_state = {sim.init_state}",
              f"N_samples = {sim.N_samples}", f"simulation_time = {w.T_ns!r}", f"dt = {w.dt_ns!r}",
              f"N_krylov = {sim.N_krylov}", f"integrator_tolerance = {sim.tol!r}",
              f"field_direction = {fmt(s.field_dir)}", f"trace_vectors = {sim.vec_kind}"]
    if output_format:
        lines.append(f"output_format = {output_format}")
    if sim.M1 or sim.M2:
        lines += [f"M1 = {sim.M1}", f"M2 = {sim.M2}", f"block_tolerance = {sim.block_tol!r}"]
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")
