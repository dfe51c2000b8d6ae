"""Open x boundary (boundaryTypeX == 1) on one GPU: the reference run on npe
ranks with one-sided closures on the first and last rank and the
characteristic outflow condition on the lines x = 0 and x = Lx, against the
oracle's open mode (pinned on the CPU by polynomial exactness, mirror
symmetry and the eigen-projection of the 1D MHD Jacobian).

* per-field rel-L2 <= 1e-12 every step, npe = 2 … 5, with flow leaving and
  entering the box, sub- and super-fast;
* graph replay = eager; batch = single runs; uniform state is kept;
* FP32 and mixed within their bounds; invalid settings rejected.
"""
import numpy as np
import pytest

import oracle_lib as orc
from util import assert_fields_close, field_errors

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def m():
    import mhd2dcfdfftw_open_amd as m
    m.lib()
    return m


def _state(m, nx, ny, flow, seed=1):
    u = m.turbulence(nx, ny, seed=seed, kmax=min(6, nx // 8, ny // 4), rms=0.1)
    u[1] += flow
    return u


CASES = [
    # nx, ny, npe, background ux, steps
    (64, 32, 2, 0.5, 8),
    (128, 32, 3, -0.4, 6),   # blocks 43, 43, 42
    (128, 64, 4, 0.0, 5),
    (128, 32, 5, 2.5, 6),    # faster than the fast speed: all waves leave at x = Lx
    (256, 32, 2, -0.2, 3),
]


@pytest.mark.parametrize("nx,ny,npe,flow,steps", CASES)
def test_open_boundary_matches_oracle(m, nx, ny, npe, flow, steps):
    print(f"[open] {nx}x{ny} npe={npe} flow={flow}", flush=True)
    u0 = _state(m, nx, ny, flow)
    kw = dict(hall=1, di=0.05)
    ctx = m.Context(nx, ny, npe=npe, open_x=True, **kw)
    ctx.set_state(u0)
    S = orc.Sim(nx, ny, npe=npe, open_x=True, **kw)
    S.set_state(u0)
    for k in range(steps):
        ctx.run(1)
        S.run(1)
        assert_fields_close(ctx.state(), S.state(), what=f"open npe={npe} step {k + 1}")
        assert abs(ctx.scalars()[1] - S.time) <= 1e-12 * S.time


def test_open_graph_and_batch(m):
    nx, ny, npe = 128, 64, 4
    states = [_state(m, nx, ny, f, seed=3 + b) for b, f in enumerate((0.3, -0.3, 0.0))]
    batch = m.Context(nx, ny, npe=npe, open_x=True, hall=1, batch=3)
    for b in range(3):
        batch.set_di(b, 0.02 * b)
        batch.set_state(states[b], inst=b)
    batch.run(5, graph=True)
    for b in range(3):
        one = m.Context(nx, ny, npe=npe, open_x=True, hall=1, di=0.02 * b)
        one.set_state(states[b])
        one.run(5)
        np.testing.assert_array_equal(batch.state(b), one.state())


def test_open_uniform_state_is_kept(m):
    nx, ny = 64, 32
    u0 = np.zeros((8, nx, ny))
    u0[0], u0[1], u0[4], u0[5], u0[7] = 1.0, 0.3, 0.8, 0.2, 0.6
    ctx = m.Context(nx, ny, npe=2, open_x=True)
    ctx.set_state(u0)
    ctx.run(5, graph=True)
    np.testing.assert_allclose(ctx.state(), u0, rtol=0, atol=1e-13)


@pytest.mark.parametrize("prec,tol", [("mixed", 1e-6), ("fp32", 5e-5)])
def test_open_reduced_precision(m, prec, tol):
    nx, ny, npe = 128, 64, 4
    u0 = _state(m, nx, ny, 0.4)
    ref = m.Context(nx, ny, npe=npe, open_x=True, hall=1, di=0.05)
    ref.set_state(u0)
    ref.run(20)
    P = m.MIXED if prec == "mixed" else m.FP32
    c = m.Context(nx, ny, npe=npe, open_x=True, hall=1, di=0.05, precision=P)
    c.set_state(u0)
    c.run(20)
    e = field_errors(c.state(), ref.state())
    print(f"[open] {prec} 20 steps: {e.max():.2e}")
    assert e.max() <= tol, e


def test_open_settings_rejected(m):
    with pytest.raises(Exception):
        m.Context(64, 64, open_x=True)             # one rank: ill-defined in the reference
    with pytest.raises(Exception):
        m.Context(64, 64, npe=16, open_x=True)     # blocks of 4 rows


def test_drive_open_boundary_matches_oracle_loop(m, tmp_path):
    import array
    import os
    inp = m.Input(boundaryTypeX=1, cfl=0.3, Tmax=0.03, dtOut=0.01, nx=64, ny=32,
                  Lx=2 * np.pi, Ly=2 * np.pi, gamma=5.0 / 3.0, hall=1, di=0.05,
                  resistivity=0.0, viscosity=0.0, divbclean=0, outdivb=0)
    u0 = _state(m, 64, 32, 0.3)
    rep = m.drive(inp, u0, outdir=str(tmp_path), mode=m.DRIVE_FUSED, npe=2)
    S = orc.Sim(64, 32, hall=1, di=0.05, npe=2, open_x=True)
    S.set_state(u0)
    S.run(rep["steps"])
    a = array.array("d")
    with open(os.path.join(str(tmp_path), f"out{rep['outputs'] - 1:04d}.dat"), "rb") as f:
        a.fromfile(f, 8 * 64 * 32 + 1)
    assert abs(a[0] - S.time) <= 1e-12 * S.time
    assert_fields_close(np.reshape(np.asarray(a[1:]), [8, 64, 32]), S.state(),
                        what="drive open x final output")
    g = np.fromfile(tmp_path / "xgrid.dat", dtype=np.float64)[1:]
    np.testing.assert_allclose(g, np.arange(64) * 2 * np.pi / 63, rtol=1e-15)
    with pytest.raises(Exception):   # one rank
        m.drive(inp, u0, outdir=str(tmp_path), mode=m.DRIVE_FUSED, npe=1)
