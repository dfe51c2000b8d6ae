"""Parity of the fused device stepping (context API) with the FP64 oracle.

* step-by-step per-field rel-L2 <= 1e-12 and matching dt / time sequences
  (Orszag-Tang with the Hall CFL term, seeded turbulence, linear wave);
* hipGraph replay bitwise equal to eager launches; runs are deterministic;
* FP32 state error against the FP64 oracle stays within its stated bound;
* diagnostics (energy, max div B) match the oracle's;
* at the 2048^2 workload size: finite, deterministic, energy and div B held.
"""
import numpy as np
import pytest

import oracle_lib as orc
from util import FP64_TOL, assert_fields_close, field_errors

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def m():
    import mhd2dcfdfftw_open_amd as m
    m.lib()
    return m


CASES = {
    "ot_hall_64x64": dict(nx=64, ny=64, ic="ot", kw=dict(hall=1, di=0.05), steps=25),
    "turb_128x64": dict(nx=128, ny=64, ic="turb", kw=dict(hall=1, di=0.02), steps=12),
    "turb_32x256": dict(nx=32, ny=256, ic="turb", kw={}, steps=6),
    "wave_64x128": dict(nx=64, ny=128, ic="wave", kw=dict(hall=1, di=0.1), steps=10),
}


def _ic(m, c):
    if c["ic"] == "ot":
        return m.orszag_tang(c["nx"], c["ny"])
    if c["ic"] == "turb":
        return m.turbulence(c["nx"], c["ny"], seed=7, kmax=8)
    return m.linear_wave(c["nx"], c["ny"], seed=5, amp=1e-3)


@pytest.mark.parametrize("name", list(CASES))
def test_fp64_steps_match_oracle(m, name):
    c = CASES[name]
    print(f"[ctx] {name}", flush=True)
    u0 = _ic(m, c)
    ctx = m.Context(c["nx"], c["ny"], **c["kw"])
    ctx.set_state(u0)
    S = orc.Sim(c["nx"], c["ny"], **c["kw"])
    S.set_state(u0)
    worst = 0.0
    for k in range(c["steps"]):
        ctx.run(1)
        S.run(1)
        dt, t, n = ctx.scalars()
        assert n == k + 1
        e = assert_fields_close(ctx.state(), S.state(), what=f"{name} step {k + 1}")
        worst = max(worst, e.max())
        # the dt used in step k+1 and the time after it
        assert abs(t - S.time) <= 1e-12 * S.time
    ctx.sync()
    print(f"[ctx] {name}: worst per-field rel-L2 {worst:.2e}")


def test_dt_sequence_and_hysteresis(m):
    # the +-2 % hysteresis keeps dt fixed for several steps, then resets it
    nx, ny = 64, 64
    u0 = m.orszag_tang(nx, ny)
    ctx = m.Context(nx, ny, hall=1, di=0.05)
    ctx.set_state(u0)
    S = orc.Sim(nx, ny, hall=1, di=0.05)
    S.set_state(u0)
    dts_gpu, dts_orc = [], []
    t_prev_g = t_prev_o = 0.0
    for k in range(30):
        ctx.run(1)
        S.run(1)
        _, t, _ = ctx.scalars()
        dts_gpu.append(t - t_prev_g)
        dts_orc.append(S.time - t_prev_o)
        t_prev_g, t_prev_o = t, S.time
    np.testing.assert_allclose(dts_gpu, dts_orc, rtol=1e-11)
    assert len(set(np.round(dts_orc, 15))) < len(dts_orc)   # hysteresis held dt


def test_graph_replay_bitwise_equal_to_eager(m):
    nx, ny = 128, 128
    u0 = m.turbulence(nx, ny, seed=1, kmax=10)
    a = m.Context(nx, ny)
    b = m.Context(nx, ny)
    a.set_state(u0)
    b.set_state(u0)
    for _ in range(3):          # graph replays separated by host work and syncs
        a.run(4)
        b.run(4, graph=True)
        b.sync()
        np.testing.assert_array_equal(a.state(), b.state())
        assert a.scalars() == b.scalars()


def test_deterministic(m):
    nx, ny = 256, 128
    u0 = m.turbulence(nx, ny, seed=2)
    out = []
    for _ in range(2):
        c = m.Context(nx, ny, hall=1, di=0.02)
        c.set_state(u0)
        c.run(5)
        out.append(c.state())
    np.testing.assert_array_equal(out[0], out[1])


def test_uniform_state_fixed_point(m):
    nx, ny = 64, 32
    uu = np.zeros((8, nx, ny))
    uu[0], uu[1], uu[2], uu[3] = 1.2, 0.1, -0.2, 0.3
    uu[4], uu[5], uu[6], uu[7] = 1.0, -0.4, 0.2, 0.9
    c = m.Context(nx, ny)
    c.set_state(uu)
    c.run(5)
    np.testing.assert_allclose(c.state(), uu, rtol=1e-13, atol=1e-14)


def test_diagnostics_match_oracle(m):
    nx, ny = 64, 128
    u0 = m.turbulence(nx, ny, seed=4, kmax=10)
    c = m.Context(nx, ny)
    c.set_state(u0)
    c.run(3)
    e, dmax = c.diagnostics()
    u = c.state()
    dx, dy = 2 * np.pi / nx, 2 * np.pi / ny
    e_ref = orc.energy(u, dx, dy)
    div = orc.derivX(u[4:5], dx)[0] + orc.derivY(u[5:6], 2 * np.pi)[0]
    assert abs(e - e_ref) <= 1e-12 * e_ref
    assert abs(dmax - np.abs(div).max()) <= 1e-9 * np.abs(div).max() + 1e-14


def test_fp32_error_vs_fp64_oracle(m):
    # FP32 state and arithmetic; stated bound checked after 20 steps
    nx, ny = 128, 128
    u0 = m.turbulence(nx, ny, seed=8)
    c = m.Context(nx, ny, precision=m.FP32)
    c.set_state(u0)
    S = orc.Sim(nx, ny)
    S.set_state(u0)
    c.run(20)
    S.run(20)
    e = field_errors(c.state(), S.state())
    print("[ctx] fp32 per-field rel-L2 after 20 steps:", e)
    assert np.all(e < 1e-4), e


def test_full_size_properties_2048(m):
    nx = ny = 2048
    print("[ctx] 2048^2 turbulence, 4 steps", flush=True)
    u0 = m.turbulence(nx, ny, seed=42, rms=0.1, kmax=16)
    c = m.Context(nx, ny, hall=1, di=0.02)
    c.set_state(u0)
    e0, d0 = c.diagnostics()
    c.run(4)
    c.sync()
    e1, d1 = c.diagnostics()
    u = c.state()
    assert np.isfinite(u).all()
    print(f"[ctx] 2048^2: energy {e0:.12e} -> {e1:.12e}, max|div B| {d0:.3e} -> {d1:.3e}")
    assert abs(e1 - e0) <= 1e-6 * e0          # ideal MHD conserves energy; filters dissipate little
    # |dB/dx| ~ rms * kmax ~ 1.6: div B stays at the scheme's truncation level
    assert d1 <= 1e-4 and d1 <= 100 * max(d0, 1e-9)
    c2 = m.Context(nx, ny, hall=1, di=0.02)
    c2.set_state(u0)
    c2.run(4, graph=True)
    np.testing.assert_array_equal(c2.state(), u)


def test_full_size_properties_4096_mixed(m):
    # the 4096^2 workload of the precision comparison: the mixed form
    # conserves energy like the FP64 run, stays within 1e-6 of it, and its
    # graph replay equals its eager run bit for bit
    n = 4096
    print("[ctx] 4096^2 turbulence, mixed and FP64, 3 steps", flush=True)
    u0 = m.turbulence(n, n, seed=42, rms=0.1, kmin=1, kmax=16)
    ref = m.Context(n, n, hall=1, di=0.02)
    mix = m.Context(n, n, hall=1, di=0.02, precision=m.MIXED)
    for c in (ref, mix):
        c.set_state(u0)
    e0, _ = mix.diagnostics()
    for c in (ref, mix):
        c.run(3)
        c.sync()
    e1, d1 = mix.diagnostics()
    u = mix.state()
    assert np.isfinite(u).all()
    assert abs(e1 - e0) <= 1e-6 * e0
    assert d1 <= 1e-4
    e = field_errors(u, ref.state())
    print(f"[ctx] 4096^2 mixed vs FP64 after 3 steps: {e.max():.2e}")
    assert e.max() <= 1e-6, e
    del ref
    g = m.Context(n, n, hall=1, di=0.02, precision=m.MIXED)
    g.set_state(u0)
    g.run(3, graph=True)
    np.testing.assert_array_equal(g.state(), u)


@pytest.mark.parametrize("nx,ny", [(16, 16), (32, 16), (16, 32)])
def test_small_grids_with_spare_rows(m, nx, ny):
    # short lines put several rows in one y-sweep workgroup (64 for ny = 16),
    # more than the grid has: the spare rows take no part in the CFL
    # reduction, so the run reports no non-finite limit, and it matches the
    # oracle
    print(f"[context] small grid {nx}x{ny}", flush=True)
    u0 = m.turbulence(nx, ny, seed=3, kmax=2)
    ctx = m.Context(nx, ny, hall=1, di=0.05)
    ctx.set_state(u0)
    S = orc.Sim(nx, ny, hall=1, di=0.05)
    S.set_state(u0)
    for k in range(4):
        ctx.run(1)
        ctx.sync()
        S.run(1)
        assert_fields_close(ctx.state(), S.state(), what=f"{nx}x{ny} step {k + 1}")
        assert abs(ctx.scalars()[1] - S.time) <= 1e-12 * S.time
    e, dmax = ctx.diagnostics()
    u = ctx.state()
    dx, dy = 2 * np.pi / nx, 2 * np.pi / ny
    assert abs(e - orc.energy(u, dx, dy)) <= 1e-12 * e
    div = orc.derivX(u[4:5], dx)[0] + orc.derivY(u[5:6], 2 * np.pi)[0]
    assert abs(dmax - np.abs(div).max()) <= 1e-9 * np.abs(div).max() + 1e-14
    np.testing.assert_allclose(ctx.divb(), div, rtol=0, atol=1e-9 * np.abs(div).max() + 1e-14)
