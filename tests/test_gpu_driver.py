"""The reference's time-driven run loop (mhd.c:58-230) on the GPU against the
same loop driven by the FP64 oracle: output times and files, step counts, the
dt < 1e-8 stop, both engines (fused context and drop-in entry points, each
with viscosity, resistivity and both div B cleanings) and the mhd2d_run
executable."""
import array
import os
import subprocess

import numpy as np
import pytest

import oracle_lib as orc
from util import assert_fields_close

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def m():
    import mhd2dcfdfftw_open_amd as m
    m.lib()
    return m


def oracle_loop(inp, u0, max_steps=10 ** 9):
    """mhd.c:58-230 with the oracle: [(t, state)] of every output."""
    S = orc.Sim(inp.nx, inp.ny, inp.Lx, inp.Ly, inp.gamma, inp.cfl, inp.hall, inp.di,
                inp.resistivity, inp.viscosity, inp.divbclean)
    S.set_state(u0)
    outs = [(0.0, S.state())]
    t, t_out, steps = 0.0, inp.dtOut, 0
    stopped = False
    while t <= inp.Tmax:
        if t >= t_out:
            outs.append((t, S.state()))
            t_out += inp.dtOut
        if S.dt < 1e-8:
            outs.append((t, S.state()))
            stopped = True
            break
        if steps >= max_steps:
            break
        S.run(1)
        t = S.time
        steps += 1
    if not stopped:
        outs.append((t, S.state()))
    return outs, steps, stopped


def read_outputs(d, nx, ny, count, nvar=8):
    res = []
    for k in range(count):
        a = array.array("d")
        with open(os.path.join(d, f"out{k:04d}.dat"), "rb") as f:
            a.fromfile(f, nvar * nx * ny + 1)
            assert f.read() == b""
        res.append((a[0], np.reshape(np.asarray(a[1:]), [nvar, nx, ny])))
    return res


def _inp(m, **kw):
    base = dict(boundaryTypeX=0, cfl=0.3, Tmax=0.06, dtOut=0.02, nx=64, ny=32,
                Lx=2 * np.pi, Ly=2 * np.pi, gamma=5.0 / 3.0, hall=1, di=0.05,
                resistivity=0.0, viscosity=0.0, divbclean=0, outdivb=1)
    base.update(kw)
    return m.Input(**base)


@pytest.mark.parametrize("engine,extra", [
    ("fused", {}),
    ("fused", dict(viscosity=2e-3, resistivity=1e-3, hall=0)),
    ("fused", dict(divbclean=1)),
    ("fused", dict(divbclean=2, viscosity=2e-3)),
    ("dropin", {}),
    ("dropin", dict(viscosity=2e-3, resistivity=1e-3, hall=0)),
    ("dropin", dict(divbclean=1)),
    ("dropin", dict(divbclean=2)),
])
def test_drive_matches_oracle_loop(m, tmp_path, engine, extra):
    inp = _inp(m, **extra)
    print(f"[driver] {engine} {extra}", flush=True)
    u0 = m.orszag_tang(inp.nx, inp.ny)
    mode = m.DRIVE_FUSED if engine == "fused" else m.DRIVE_DROPIN
    rep = m.drive(inp, u0, outdir=str(tmp_path), mode=mode)
    want, steps, stopped = oracle_loop(inp, u0)
    assert rep["steps"] == steps and not rep["stopped"] and not stopped
    assert rep["outputs"] == len(want)
    got = read_outputs(str(tmp_path), inp.nx, inp.ny, len(want), nvar=want[0][1].shape[0])
    for (tg, ug), (tw, uw) in zip(got, want):
        assert abs(tg - tw) <= 1e-12 * max(tw, 1e-300)
        assert_fields_close(ug, uw, what=f"{engine} output at t={tw:.4f}")
    assert (tmp_path / "log").exists() and (tmp_path / "xgrid.dat").exists()
    assert (tmp_path / f"divB{len(want) - 1:04d}.dat").exists()


@pytest.mark.parametrize("engine", ["fused", "dropin"])
def test_drive_stops_when_dt_below_1e8(m, tmp_path, engine):
    # huge sound speed: dt = cfl * dx / c_s ~ 1e-10 < 1e-8 (mhd.c:131-151)
    inp = _inp(m, hall=0)
    u0 = np.zeros((8, inp.nx, inp.ny))
    u0[0], u0[4], u0[7] = 1e-12, 1.0, 1e4
    mode = m.DRIVE_FUSED if engine == "fused" else m.DRIVE_DROPIN
    rep = m.drive(inp, u0, outdir=str(tmp_path), mode=mode)
    want, steps, stopped = oracle_loop(inp, u0)
    assert stopped and rep["stopped"] == 1 and rep["steps"] == steps == 0
    assert rep["outputs"] == len(want) == 2
    got = read_outputs(str(tmp_path), inp.nx, inp.ny, 2)
    np.testing.assert_array_equal(got[1][1], u0)


def test_mhd2d_run_executable(m, tmp_path):
    exe = os.path.join(os.path.dirname(m.LIB_PATH), "mhd2d_run")
    m.write_input(str(tmp_path / "input"), boundaryTypeX=0, cfl=0.3, Tmax=0.02, dtOut=0.01,
                  nx=32, ny=32, Lx=2 * np.pi, Ly=2 * np.pi, gamma=5.0 / 3.0, hall=0, di=0.0,
                  resistivity=0.0, viscosity=0.0, divbclean=0, outdivb=0)
    r = subprocess.run([exe, "--input", str(tmp_path / "input"), "--out", str(tmp_path),
                        "--ic", "ot", "--quiet"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    inp = m.read_input(str(tmp_path / "input"))
    want, steps, _ = oracle_loop(inp, m.orszag_tang(32, 32))
    got = read_outputs(str(tmp_path), 32, 32, len(want))
    assert f"steps = {steps}," in r.stdout
    assert_fields_close(got[-1][1], want[-1][1], what="mhd2d_run final output")
    # restart from the last output continues the run
    r2 = subprocess.run([exe, "--input", str(tmp_path / "input"), "--out", str(tmp_path),
                         "--ic", str(tmp_path / f"out{len(want) - 1:04d}.dat"), "--quiet",
                         "--max-steps", "2"], capture_output=True, text=True, timeout=120)
    assert r2.returncode == 0, r2.stderr
