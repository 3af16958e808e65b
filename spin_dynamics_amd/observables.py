"""class_observables (observables.f90:9-148): population traces, Simpson
yields and the .data files."""
from __future__ import annotations

import os
from dataclasses import dataclass
from typing import Optional

import numpy as np


def gfortran_real(x: float) -> str:
    """One value as gfortran's list-directed output writes a real(8), the form
    of the reference's .data files from its gfortran builds
    (observables.f90:143): a blank, then 17 significant digits, as G25.17E3
    where F editing applies (0.1 <= |x| < 1e17, and 0) and as 1PE25.16E3
    elsewhere."""
    x = float(x)
    if x == 0.0:
        return "   0.0000000000000000     "
    mant, exp = f"{abs(x):.16e}".split("e")
    e = int(exp)
    if -1 <= e <= 16:  # F editing: 10^(n-1) <= |x| < 10^n, 17 - n decimals
        n = e + 1
        body = f"{x:.{17 - n}f}" + ("." if n == 17 else "")  # F editing keeps the point
        return " " + body.rjust(20) + " " * 5
    body = f"{'-' if x < 0 else ''}{mant}E{'-' if e < 0 else '+'}{abs(e):03d}"
    return " " + body.rjust(25)


@dataclass
class Observables:
    P_S: np.ndarray
    P_T0: np.ndarray
    P_Tp: np.ndarray
    P_Tm: np.ndarray
    iden: Optional[np.ndarray] = None
    Phi_S: Optional[np.ndarray] = None
    Phi_T: Optional[np.ndarray] = None
    P_T: Optional[np.ndarray] = None  # total triplet, when sublevels are not resolved (semiclassical)

    @classmethod
    def zeros(cls, n: int) -> "Observables":
        return cls(*(np.zeros(n) for _ in range(4)))

    def update(self, other: "Observables"):
        for k in ("P_S", "P_T0", "P_Tp", "P_Tm"):
            setattr(self, k, getattr(self, k) + getattr(other, k))

    def scale(self, f: float):
        for k in ("P_S", "P_T0", "P_Tp", "P_Tm"):
            setattr(self, k, getattr(self, k) * f)

    def get_kinetics(self, dt: float, kS: float, kT: float):
        """observables.f90:74-101: cumulative Simpson yields over size/2 pairs."""
        n = len(self.P_S)
        P_T = self.P_Tp + self.P_T0 + self.P_Tm if self.P_T is None else self.P_T
        self.iden = self.P_S + P_T
        h = n // 2
        self.Phi_S = np.zeros(h)
        self.Phi_T = np.zeros(h)
        s = t = 0.0
        for i in range(1, h):
            s = s + (1.0 / 3.0) * dt * kS * (self.P_S[2 * i - 2] + 4 * self.P_S[2 * i - 1] + self.P_S[2 * i])
            t = t + (1.0 / 3.0) * dt * kT * (P_T[2 * i - 2] + 4 * P_T[2 * i - 1] + P_T[2 * i])
            self.Phi_S[i] = s
            self.Phi_T[i] = t

    def output(self, path: str, style: str = "default"):
        """observables.f90:103-148: one value per line; style "gfortran"
        writes each as gfortran's list-directed output does (gfortran_real)."""
        if self.P_T is None:
            pops = [("S_prob.data", self.P_S), ("Tp_prob.data", self.P_Tp), ("T0_prob.data", self.P_T0),
                    ("Tm_prob.data", self.P_Tm)]
        else:
            pops = [("S_prob.data", self.P_S), ("T_prob.data", self.P_T)]
        files = pops + [("total_prob.data", self.iden), ("S_yield.data", self.Phi_S), ("T_yield.data", self.Phi_T)]
        for name, data in files:
            if data is None:
                continue
            with open(os.path.join(path, name), "w") as f:
                for x in data:
                    f.write((gfortran_real(x) if style == "gfortran" else f"   {x!r}") + "\n")
