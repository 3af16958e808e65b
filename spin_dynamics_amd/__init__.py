"""spin_dynamics_amd -- MI355X (gfx950) trace-sampling singlet-yield propagation.

The compute path is libspinchem_hip.so (hand-written HIP kernels behind a
C-ABI, include/spinchem.h); this package is its host-side mirror of the
reference's Fortran interface (modules variables, dynamics,
matrix_exponential, class_observables, sc_dynamics).
"""
from ._lib import Context, SpinchemError, load  # noqa: F401
from .spinsys import (GAMMA_14N, GAMMA_1H, GAMMA_E, Electron, InputError, SimParam, SysParam,  # noqa: F401
                      read_inp)
from .observables import Observables  # noqa: F401
from .dynamics import (Hamiltonian, exact_dynamics, expm, last_run_stats, n_steps,  # noqa: F401
                       trace_sampling_para)
from .sc_dynamics import semi_classical  # noqa: F401

__all__ = ["Context", "SpinchemError", "load", "Electron", "SysParam", "SimParam", "InputError", "read_inp",
           "Observables", "Hamiltonian", "trace_sampling_para", "exact_dynamics", "expm", "n_steps",
           "last_run_stats", "semi_classical", "GAMMA_E", "GAMMA_1H", "GAMMA_14N"]
