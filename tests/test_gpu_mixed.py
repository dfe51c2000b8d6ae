"""Mixed precision (FP64 state, filters and RK accumulation; FP32 derivative
solves, derivative arrays and R_prev) against FP64.

The FP64 device path is held to the oracle at 1e-12 per field per step
(test_gpu_context.py), so after a 200-step run it stands in for the oracle
where the oracle would take minutes.  Gate: per-field rel-L2 <= 1e-6 after
200 steps (the mixed-form tolerance).  Plain FP32 is measured beside it to
show why the state stays FP64: its error is well above 1e-6.
"""
import numpy as np
import pytest

import oracle_lib as orc
from util import field_errors

pytestmark = pytest.mark.gpu

MIXED_TOL = 1e-6


@pytest.fixture(scope="module")
def m():
    import mhd2dcfdfftw_open_amd as m
    m.lib()
    return m


def _run(m, u0, nx, ny, prec, steps, batch=1, **kw):
    ctx = m.Context(nx, ny, precision=prec, batch=batch, **kw)
    for b in range(batch):
        ctx.set_state(u0, inst=b)
    ctx.run(steps, graph=True)
    ctx.sync()
    return ctx


def test_mixed_matches_oracle_after_20_steps(m):
    nx, ny = 64, 64
    u0 = m.turbulence(nx, ny, seed=4, kmax=8)
    ctx = _run(m, u0, nx, ny, m.MIXED, 20, hall=1, di=0.02)
    S = orc.Sim(nx, ny, hall=1, di=0.02)
    S.set_state(u0)
    S.run(20)
    e = field_errors(ctx.state(), S.state())
    print(f"[mixed] 64^2 20 steps vs oracle: {e.max():.2e}")
    assert e.max() <= MIXED_TOL, e


@pytest.mark.parametrize("nx,ny", [(256, 256), (512, 512), (1024, 256)])
def test_mixed_within_1e6_after_200_steps(m, nx, ny):
    print(f"[mixed] {nx}x{ny} 200 steps", flush=True)
    u0 = m.turbulence(nx, ny, seed=42, kmax=16)
    ref = _run(m, u0, nx, ny, m.FP64, 200, hall=1, di=0.02)
    mix = _run(m, u0, nx, ny, m.MIXED, 200, hall=1, di=0.02)
    # the same dt sequence (dt is reduced in FP64 from the FP64 state)
    assert mix.scalars()[1] == pytest.approx(ref.scalars()[1], rel=1e-9)
    e = field_errors(mix.state(), ref.state())
    f32 = _run(m, u0, nx, ny, m.FP32, 200, hall=1, di=0.02)
    e32 = field_errors(f32.state(), ref.state())
    print(f"[mixed] {nx}x{ny}: mixed max rel-L2 {e.max():.2e}, fp32 {e32.max():.2e}")
    assert e.max() <= MIXED_TOL, e
    assert e32.max() > 10 * e.max()   # why the state is kept in FP64


def test_mixed_batch_equals_single_bitwise(m):
    nx, ny, B = 128, 64, 3
    ctx = m.Context(nx, ny, precision=m.MIXED, hall=1, batch=B)
    states = [m.turbulence(nx, ny, seed=60 + b, kmax=6) for b in range(B)]
    for b in range(B):
        ctx.set_di(b, 0.05 * b)
        ctx.set_state(states[b], inst=b)
    ctx.run(5, graph=True)
    for b in range(B):
        one = m.Context(nx, ny, precision=m.MIXED, hall=1, di=0.05 * b)
        one.set_state(states[b])
        one.run(5)
        np.testing.assert_array_equal(ctx.state(b), one.state())


def test_mixed_diagnostics_in_fp64(m):
    nx, ny = 128, 128
    u0 = m.orszag_tang(nx, ny)
    ctx = m.Context(nx, ny, precision=m.MIXED)
    ctx.set_state(u0)
    e, _ = ctx.diagnostics()
    dx = dy = 2 * np.pi / nx
    assert e == pytest.approx(orc.energy(u0, dx, dy), rel=1e-13)


@pytest.mark.parametrize("seed", [42, 43])
def test_mixed_small_amplitude_wave_within_1e6(m, seed):
    # a 1e-4 linear wave on rho = p = 1, B0 = x: FP32 rounding of the
    # background would swamp the wave's x-derivatives without the constant
    # subtraction of the narrowing x-solve loads
    nx = ny = 256
    print(f"[mixed] linear wave seed {seed}", flush=True)
    u0 = m.linear_wave(nx, ny, seed=seed, amp=1e-4)
    ref = _run(m, u0, nx, ny, m.FP64, 200, hall=1, di=0.1)
    mix = _run(m, u0, nx, ny, m.MIXED, 200, hall=1, di=0.1)
    e = field_errors(mix.state(), ref.state())
    print(f"[mixed] wave: max rel-L2 {e.max():.2e}")
    assert e.max() <= MIXED_TOL, e


@pytest.mark.parametrize("ny,clean", [(2048, 0), (4096, 0), (4096, 2)])
def test_mixed_dissipative_long_lines(m, ny, clean):
    # the y-sweep's mixed Laplacian and GLM instances on 2048- and 4096-point
    # lines against the FP64 device path (held to the oracle per step on
    # the same lines): the 8 MHD fields within the mixed-form tolerance
    nx, steps = 32, 20
    print(f"[mixed] {nx}x{ny} viscous + resistive, cleaning {clean}, {steps} steps", flush=True)
    u0 = m.turbulence(nx, ny, seed=7, kmax=4, rms=0.1)
    if clean == 2:
        x = np.arange(nx) * 2 * np.pi / nx
        u0 = u0.copy()
        u0[4] += 0.01 * np.sin(2 * x)[:, None]
        u0 = np.concatenate([u0, np.zeros((1, nx, ny))])
    kw = dict(hall=1, di=0.05, viscosity=1e-3, resistivity=1e-3, divbclean=clean)
    ref = _run(m, u0, nx, ny, m.FP64, steps, **kw)
    mix = _run(m, u0, nx, ny, m.MIXED, steps, **kw)
    e = field_errors(mix.state()[:8], ref.state()[:8])
    print(f"[mixed] {nx}x{ny} cleaning {clean}: max rel-L2 {e.max():.2e}")
    assert e.max() <= MIXED_TOL, e
