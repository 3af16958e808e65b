"""The .data files (observables.f90:103-148) in gfortran's list-directed form
(the extension key output_format = gfortran) and the reference's own reader,
example/utils.py:420-456 parse_output.

Pins: gfortran writes a real(8) in list-directed output as a blank and 17
significant digits, G25.17E3 where F editing applies and 1PE25.16E3 elsewhere
(these strings are gfortran's documented output for these values; no
gfortran is installed here to run, so the exact spacing is unpinned beyond
them); 17 significant digits identify a double, so a written trace reads
back bit for bit; parse_output's steps (the *_prob.data files of a field
folder, numpy.loadtxt, time = index * dt in microseconds) read the folder."""
import pathlib

import numpy as np
import pytest

from spin_dynamics_amd.observables import Observables, gfortran_real

GFORTRAN = [
    (1.0, "   1.0000000000000000     "),
    (0.5, "  0.50000000000000000     "),
    (0.1, "  0.10000000000000001     "),
    (1e-5, "   1.0000000000000001E-005"),
    (123.456, "   123.45600000000000     "),
    (-0.5, " -0.50000000000000000     "),
    (0.0, "   0.0000000000000000     "),
    (3.14159265358979, "   3.1415926535897900     "),
    (-2.5e-3, "  -2.5000000000000001E-003"),
    (1e17, "   1.0000000000000000E+017"),
]


@pytest.mark.parametrize("x,text", GFORTRAN)
def test_gfortran_list_directed_real(x, text):
    assert gfortran_real(x) == text


def parse_output(output_path, subdir, dt):
    """The steps of example/utils.py:420-456 (without polars): every
    <stem>_prob.data file of the field folder read by numpy.loadtxt, plus a
    time column in microseconds."""
    folder = pathlib.Path(output_path) / subdir
    data = {}
    for f in folder.glob("*.data"):
        if f.stem[-5:] != "_prob":
            continue
        data[f.stem] = np.loadtxt(f)
    if not data:
        raise FileNotFoundError(folder)
    data["time"] = np.arange(len(next(iter(data.values())))) * dt * 1e06
    return data


def test_written_traces_read_back_bit_for_bit(tmp_path):
    rng = np.random.default_rng(4)
    n = 101
    obs = Observables(*(rng.uniform(0, 1, n) * 10.0 ** rng.integers(-25, 3, n) for _ in range(4)))
    obs.get_kinetics(0.176, 0.011, 0.003)
    folder = tmp_path / ".050"
    folder.mkdir()
    obs.output(str(folder), style="gfortran")
    got = parse_output(tmp_path, ".050", 1e-9)
    assert sorted(got) == ["S_prob", "T0_prob", "Tm_prob", "Tp_prob", "time", "total_prob"]
    for k, v in (("S_prob", obs.P_S), ("T0_prob", obs.P_T0), ("Tp_prob", obs.P_Tp), ("Tm_prob", obs.P_Tm)):
        assert np.array_equal(got[k], v)
    assert np.allclose(got["time"], np.arange(n) * 1e-3)
    lines = (folder / "S_prob.data").read_text().splitlines()
    assert all(len(line) == 26 for line in lines)
    assert np.array_equal(np.loadtxt(folder / "S_yield.data"), obs.Phi_S)
