"""CPU checks of the reference's configuration and file formats (host code).

* the `input` reader against the reference's own `input` file (kept verbatim
  in tests/golden/) and its 16-character/atoi/atof semantics;
* xgrid/ygrid/out%04d/divB%04d files read back the way
  data_proc_scripts/read_output.py reads them (struct 'N' + doubles,
  array('d').fromfile with nvar*nx*ny + 1 values);
* the log text of mhd.c:319-338.
"""
import array
import json
import os
import struct

import numpy as np
import pytest

import mhd2dcfdfftw_open_amd as m

HERE = os.path.dirname(os.path.abspath(__file__))
pytestmark = pytest.mark.skipif(not os.path.exists(m.LIB_PATH), reason="libmhd2d.so not built")


def test_reads_reference_input_file():
    want = json.load(open(os.path.join(HERE, "golden", "reference_input.json")))["expected"]
    got = m.read_input(os.path.join(HERE, "golden", "reference_input")).as_dict()
    for k, v in want.items():
        assert got[k] == pytest.approx(v, rel=0, abs=0), k


def test_input_value_is_the_first_16_characters(tmp_path):
    # text after column 16 is ignored, as readLine copies 16 characters
    lines = ["0" + " " * 15 + "7 ignored"] + ["0.3"] + ["1.5"] + ["0.25"] + ["64"] + ["32"]
    lines += ["6.283185307179586"[:16], "3.5", "1.4", "1", "0.05", "0", "0", "0", "1"]
    p = tmp_path / "input"
    p.write_text("\n".join(l.ljust(20) + "; comment" for l in lines) + "\n")
    d = m.read_input(str(p)).as_dict()
    assert d["boundaryTypeX"] == 0 and d["nx"] == 64 and d["ny"] == 32
    assert d["Lx"] == float("6.28318530717958")   # 16 characters of the literal
    assert d["gamma"] == 1.4 and d["hall"] == 1 and d["outdivb"] == 1


def test_input_round_trip(tmp_path):
    kw = dict(boundaryTypeX=0, cfl=0.3, Tmax=0.5, dtOut=0.1, nx=128, ny=64, Lx=6.25, Ly=3.125,
              gamma=5.0 / 3.0, hall=1, di=0.02, resistivity=0.0, viscosity=1e-4,
              divbclean=1, outdivb=1)
    m.write_input(str(tmp_path / "input"), **kw)
    got = m.read_input(str(tmp_path / "input")).as_dict()
    for k, v in kw.items():
        assert got[k] == pytest.approx(v, rel=1e-11), k


def _read_grid(path):
    with open(path, "rb") as f:
        n = struct.unpack("N", f.read(8))[0]
        g = np.array(struct.unpack(f"{n}d", f.read(8 * n)))
        assert f.read() == b""
    return g


def test_output_files_match_the_readers_format(tmp_path):
    L = m.lib()
    d = str(tmp_path).encode()
    nx, ny = 16, 8
    assert L.mhd2d_write_grid(d, nx, ny, 2.0, 4.0) == 0
    np.testing.assert_array_equal(_read_grid(tmp_path / "xgrid.dat"), np.arange(nx) * (2.0 / nx))
    np.testing.assert_array_equal(_read_grid(tmp_path / "ygrid.dat"), np.arange(ny) * (4.0 / ny))
    uu = np.random.default_rng(0).standard_normal((8, nx, ny))
    assert L.mhd2d_write_array(d, 7, 0.125, m._ptr(uu), 8, nx, ny) == 0
    arr = array.array("d")
    with open(tmp_path / "out0007.dat", "rb") as f:
        arr.fromfile(f, 8 * nx * ny + 1)
        assert f.read() == b""
    assert arr[0] == 0.125
    np.testing.assert_array_equal(np.reshape(np.asarray(arr[1:]), [8, nx, ny], order="C"), uu)
    dv = np.random.default_rng(1).standard_normal((nx, ny))
    assert L.mhd2d_write_divb(d, 12, 0.5, m._ptr(dv), nx, ny) == 0
    raw = np.fromfile(tmp_path / "divB0012.dat", dtype=np.float64)
    assert raw[0] == 0.5 and raw.size == nx * ny + 1
    np.testing.assert_array_equal(raw[1:].reshape(nx, ny), dv)
    # restart reader
    t = np.zeros(1)
    back = np.zeros_like(uu)
    assert L.mhd2d_read_array(str(tmp_path / "out0007.dat").encode(), m._ptr(t), m._ptr(back),
                              back.size) == 0
    assert t[0] == 0.125
    np.testing.assert_array_equal(back, uu)


def test_log_text(tmp_path):
    assert m.lib().mhd2d_write_log(str(tmp_path).encode(), 1.5, 2.5e-3, 3725.9, 1234) == 0
    assert (tmp_path / "log").read_text() == (
        "  Simulation time = 1.500000\n"
        "  dt =   2.5000e-03\n"
        "  Real time used (sec) =         3725\n"
        "  Real time used (hh:mm:ss) = (001:02:05)\n"
        "  Iteration =           1234\n")
