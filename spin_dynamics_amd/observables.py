"""class_observables (observables.f90:9-148): population traces, Simpson
yields and the .data files."""
from __future__ import annotations

import os
from dataclasses import dataclass
from typing import Optional

import numpy as np


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

    def output(self, path: str):
        """observables.f90:103-148: one value per line."""
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
                    f.write(f"   {x!r}\n")
