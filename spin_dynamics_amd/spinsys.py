"""Spin-system and simulation parameters, and the .ini reader.

Mirrors module `variables` of the reference (spinsys.f90): types electron,
sys_param, sim_param, read_inp with its unit conversions (kS, kT / gamma_e;
times * gamma_e / 1000) and check_inp's section rules.  The .ini parsing
itself is the C-ABI reader (spc_ini_*), shared with the Fortran driver.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

from . import _lib

G_E = 2.002319
GAMMA_E = 176.08596          # spinsys.f90:17
GAMMA_1H = 0.26752218708     # spinsys.f90:18
GAMMA_14N = 0.0193297792     # spinsys.f90:19


class InputError(ValueError):
    """The reference stops with these messages (spinsys.f90:60-163)."""


@dataclass
class Electron:
    """spinsys.f90:21-29: g, Larmor frequency w, nuclear multiplicities and tensors."""
    g: float = G_E
    w: float = 0.0
    g_I: List[int] = field(default_factory=list)
    A: np.ndarray = field(default_factory=lambda: np.zeros((0, 3, 3)))  # A[k][a][b], mT

    @property
    def a_iso(self) -> np.ndarray:
        return np.array([np.trace(a) / 3.0 for a in self.A])

    @property
    def isotropic(self) -> bool:
        """is_isotropic (spinsys.f90:237-254): only A(1,1) == A(2,2) is tested."""
        return all(a[0, 0] == a[1, 1] for a in self.A)


@dataclass
class SysParam:
    """spinsys.f90:31-38; kS/kT in mT units (divided by gamma_e)."""
    e1: Electron = field(default_factory=Electron)
    e2: Electron = field(default_factory=Electron)
    J: float = 0.0
    D: np.ndarray = field(default_factory=lambda: np.zeros((3, 3)))
    kS: float = 0.0
    kT: float = 0.0
    field_dir: tuple = (0.0, 0.0, 1.0)

    @property
    def Z1(self) -> int:
        return int(np.prod(self.e1.g_I, dtype=np.int64))

    @property
    def Z2(self) -> int:
        return int(np.prod(self.e2.g_I, dtype=np.int64))


@dataclass
class SimParam:
    """spinsys.f90:40-56; t_end and dt converted to rad/mT."""
    type: str = "trace_sampling"
    B: List[float] = field(default_factory=lambda: [0.0])
    init_state: str = "singlet"
    seed: tuple = (42, 99)
    N_samples: int = 0
    t_end: float = 0.0
    dt: float = 0.0
    N_krylov: int = 0
    tol: float = 0.0
    M1: int = 0
    M2: int = 0
    block_tol: float = 0.0
    output_folder: str = "."
    restart: Optional[int] = None
    vec_kind: str = "SUZ"   # extension: "SUZ" (reference) or "phase"

    @property
    def isRestart(self) -> bool:
        return self.restart is not None


def to_c_system(sys: SysParam, B: float, iso_rule: int = 1) -> "tuple[_lib.SpinSystem, tuple]":
    """Flattens sys_param for spc_system_set (field magnitude B = -w); iso_rule
    1 applies the reference's isotropy test (spinsys.f90:237-254)."""
    mult = np.ascontiguousarray(list(sys.e1.g_I) + list(sys.e2.g_I), dtype=np.int32)
    A = [np.asarray(a, float).reshape(3, 3) for a in list(sys.e1.A) + list(sys.e2.A)]
    hfc = np.ascontiguousarray(np.array([a.T.reshape(-1) for a in A]).reshape(-1) if A else np.zeros(1),
                               dtype=float)  # column-major per tensor
    s = _lib.SpinSystem()
    s.n_nuc1, s.n_nuc2 = len(sys.e1.g_I), len(sys.e2.g_I)
    s.mult = _lib.ptr(mult, _lib._i32p)
    s.hfc = _lib.ptr(hfc)
    s.J = float(sys.J)
    Dc = np.asarray(sys.D, float).reshape(3, 3).T.reshape(-1)
    for i in range(9):
        s.D[i] = Dc[i]
    s.kS, s.kT, s.B = float(sys.kS), float(sys.kT), float(B)
    for i in range(3):
        s.field_dir[i] = float(sys.field_dir[i])
    s.iso_rule = int(iso_rule)
    return s, (mult, hfc)


class _Ini:
    def __init__(self, path: str):
        self.lib = _lib.load()
        self.h = ctypes.c_void_p()
        if self.lib.spc_ini_open(path.encode(), ctypes.byref(self.h)) != 0:
            raise InputError("Error: Input file does not exist.")

    def __del__(self):
        if getattr(self, "h", None):
            self.lib.spc_ini_close(self.h)

    def count(self, sec, opt) -> int:
        return self.lib.spc_ini_count_values(self.h, sec.encode(), opt.encode())

    def has(self, sec, opt) -> bool:
        return self.count(sec, opt) >= 0

    def string(self, sec, opt) -> str:
        buf = ctypes.create_string_buffer(1024)
        if self.lib.spc_ini_get_string(self.h, sec.encode(), opt.encode(), buf, 1024) != 0:
            raise KeyError(opt)
        return buf.value.decode()

    def doubles(self, sec, opt, n=None) -> np.ndarray:
        n = self.count(sec, opt) if n is None else n
        if n < 0:
            raise KeyError(opt)
        out = np.zeros(max(n, 1))
        if n and self.lib.spc_ini_get_doubles(self.h, sec.encode(), opt.encode(), _lib.ptr(out), n) != 0:
            raise KeyError(opt)
        return out[:n]

    def ints(self, sec, opt, n=None) -> np.ndarray:
        n = self.count(sec, opt) if n is None else n
        if n < 0:
            raise KeyError(opt)
        out = np.zeros(max(n, 1), dtype=np.int64)
        if n and self.lib.spc_ini_get_int64s(self.h, sec.encode(), opt.encode(), _lib.ptr(out, _lib._i64p), n) != 0:
            raise KeyError(opt)
        return out[:n]


SECTIONS = ["system variables", "simulation parameters", "electron 1", "electron 2"]


def check_inp(ini: _Ini):
    """spinsys.f90:137-163."""
    ns = ini.lib.spc_ini_num_sections(ini.h)
    if ns < len(SECTIONS):
        raise InputError("Error: Number of sections in input file is less than expected.")
    if ns > len(SECTIONS):
        raise InputError("Error: Number of sections in input file is more than expected.")
    for s in SECTIONS:
        if not ini.lib.spc_ini_has_section(ini.h, s.encode()):
            raise InputError(f'Error: Section "{s}" is missing from input file.')


def electron_init(ini: _Ini, n: int) -> Electron:
    """spinsys.f90:165-235: N_I copies of each multiplicity in I, tensors A1..An."""
    sec = f"electron {n}"
    e = Electron()
    if ini.has(sec, "g"):
        e.g = float(ini.doubles(sec, "g", 1)[0])
    try:
        mult = [int(x) for x in ini.ints(sec, "I")]
        nI = [int(x) for x in ini.ints(sec, "N_I")]
    except KeyError:
        mult, nI = [], []
    if len(mult) != len(nI):
        raise InputError("Sections N_I and I must have the sam length.")
    g_I = [m for m, k in zip(mult, nI) for _ in range(k)]
    A = []
    for i in range(len(g_I)):
        try:
            v = ini.doubles(sec, f"A{i + 1}", 9)
        except KeyError:
            raise InputError(f"Error: Hyperfine copuling A{i + 1} for electron {n} is missing.")
        A.append(v.reshape(3, 3, order="F"))  # reshape(A_tmp, [3,3]), column-major
    e.g_I = g_I
    e.A = np.array(A).reshape(-1, 3, 3)
    return e


def read_inp(filename: str, *, require_output_folder: bool = True):
    """spinsys.f90:60-135.  Returns (SysParam, SimParam)."""
    if not os.path.exists(filename):
        raise InputError("Error: Input file does not exist.")
    ini = _Ini(filename)
    check_inp(ini)
    sv, sp = "system variables", "simulation parameters"
    sys, sim = SysParam(), SimParam()

    def req(sec, key, msg, kind="d", n=None):
        try:
            return ini.doubles(sec, key, n) if kind == "d" else ini.string(sec, key)
        except KeyError:
            raise InputError(msg)

    sys.J = float(req(sv, "J", "Error: Exchange coupling is missing", n=1)[0])
    sys.D = req(sv, "D", "Error: Dipolar coupling is missing", n=9).reshape(3, 3, order="F")
    sys.kS = float(req(sv, "kS", "Error: Singlet recombination rate is missing", n=1)[0]) / GAMMA_E
    sys.kT = float(req(sv, "kT", "Error: Triplet recombination rate is missing", n=1)[0]) / GAMMA_E
    sim.type = req(sp, "simulation_type", "Error: Simulation type is missing", "s").strip()
    sim.B = [float(x) for x in req(sp, "B", "Error: List of experiment magnetic fields is missing")]
    sim.init_state = req(sp, "initial_state", "Error: Initial state is missing", "s").strip()
    sim.dt = float(req(sp, "dt", "Error: Propagation timestep is missing", n=1)[0])
    if ini.has(sp, "seed"):
        sim.seed = tuple(int(x) for x in ini.ints(sp, "seed", 2))
    opt = lambda key, conv, default: conv(ini.doubles(sp, key, 1)[0]) if ini.has(sp, key) else default
    sim.N_samples = opt("N_samples", int, 0)
    t_ns = opt("simulation_time", float, 0.0)
    sim.t_end = (t_ns * GAMMA_E) / 1000.0
    sim.dt = (sim.dt * GAMMA_E) / 1000.0
    sim.block_tol = opt("block_tolerance", float, 0.0)
    sim.N_krylov = opt("N_krylov", int, 0)
    sim.tol = opt("integrator_tolerance", float, 0.0)
    sim.M1 = opt("M1", int, 0)
    sim.M2 = opt("M2", int, 0)
    sim.restart = opt("restart", int, None)
    if ini.has(sp, "trace_vectors"):
        sim.vec_kind = ini.string(sp, "trace_vectors").strip()
    if ini.has(sp, "field_direction"):
        d = ini.doubles(sp, "field_direction", 3)
        sys.field_dir = tuple(float(x) for x in d)
    sim.output_folder = ini.string(sp, "output_folder").strip() if ini.has(sp, "output_folder") else "."
    if require_output_folder and not os.path.isdir(sim.output_folder):
        raise InputError("Error: Output folder does not exist.")
    sys.e1 = electron_init(ini, 1)
    sys.e2 = electron_init(ini, 2)
    return sys, sim
