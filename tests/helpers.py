"""Shared test helpers: conversion of the package's SysParam to the oracle's
System, and tolerance checks."""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

from oracle import pyoracle  # noqa: E402


def oracle_system(sys_param, B):
    """SysParam at field B -> pyoracle.System (full tensors)."""
    e1, e2 = sys_param.e1, sys_param.e2
    return pyoracle.System(mult1=list(e1.g_I), mult2=list(e2.g_I),
                           A1=np.asarray(e1.A).reshape(-1, 3, 3) if e1.g_I else None,
                           A2=np.asarray(e2.A).reshape(-1, 3, 3) if e2.g_I else None,
                           J=sys_param.J, D=np.asarray(sys_param.D), kS=sys_param.kS, kT=sys_param.kT, B=B,
                           direction=tuple(sys_param.field_dir))


def max_rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))
