"""Batched contexts: B independent runs of one grid advanced together, each
launch covering every instance (the ensemble workloads: 64 linear waves of
512x512 with d_i drawn per instance).

* every instance of an FP64 batch matches its own oracle run at <= 1e-12 per
  field per step, with its own dt and time (different states and d_i);
* a batch is bitwise equal to the same runs made one context at a time
  (FP64 and FP32), eager and hipGraph;
* one instance reaching the dt < 1e-8 stop freezes while the others go on;
* restarting one instance (set_state_at) leaves the others untouched;
* per-instance diagnostics; the 64 x 512^2 FP32 ensemble shape runs and
  matches single runs bitwise.
"""
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


def _states(m, nx, ny, B):
    out = []
    for b in range(B):
        if b % 3 == 0:
            out.append(m.orszag_tang(nx, ny))
        elif b % 3 == 1:
            out.append(m.turbulence(nx, ny, seed=10 + b, kmax=6))
        else:
            out.append(m.linear_wave(nx, ny, seed=20 + b, amp=1e-3))
    return out


def test_batch_fp64_matches_oracle_per_instance(m):
    nx, ny, B, steps = 64, 32, 4, 6
    print(f"[batch] fp64 {nx}x{ny} B={B}", flush=True)
    u0 = _states(m, nx, ny, B)
    dis = [0.0, 0.05, 0.13, 0.02]
    ctx = m.Context(nx, ny, hall=1, di=0.0, batch=B)
    for b in range(B):
        ctx.set_di(b, dis[b])
        ctx.set_state(u0[b], inst=b)
    sims = []
    for b in range(B):
        S = orc.Sim(nx, ny, hall=1, di=dis[b])
        S.set_state(u0[b])
        sims.append(S)
    for k in range(steps):
        ctx.run(1)
        for b, S in enumerate(sims):
            S.run(1)
            assert_fields_close(ctx.state(b), S.state(), what=f"inst {b} step {k + 1}")
            dt, t, n = ctx.scalars(b)
            assert n == k + 1
            assert abs(t - S.time) <= 1e-12 * S.time
    # the instances really differ in dt (d_i enters the Hall CFL limit)
    assert len({round(ctx.scalars(b)[0], 12) for b in range(B)}) > 1


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("graph", [False, True])
def test_batch_equals_single_runs_bitwise(m, prec, graph):
    nx, ny, B, steps = 128, 64, 3, 5
    P = m.FP64 if prec == "fp64" else m.FP32
    print(f"[batch] bitwise {prec} graph={graph}", flush=True)
    u0 = _states(m, nx, ny, B)
    dis = [0.01, 0.08, 0.2]
    ctx = m.Context(nx, ny, hall=1, precision=P, batch=B)
    for b in range(B):
        ctx.set_di(b, dis[b])
        ctx.set_state(u0[b], inst=b)
    ctx.run(steps, graph=graph)
    for b in range(B):
        one = m.Context(nx, ny, hall=1, di=dis[b], precision=P)
        one.set_state(u0[b])
        one.run(steps, graph=graph)
        np.testing.assert_array_equal(ctx.state(b), one.state())
        assert ctx.scalars(b) == one.scalars()
        assert ctx.diagnostics(b) == one.diagnostics()


def test_one_instance_stops_the_others_continue(m):
    nx, ny, B = 64, 32, 3
    u0 = _states(m, nx, ny, B)
    bad = np.zeros((8, nx, ny))
    bad[0], bad[4], bad[7] = 1e-12, 1.0, 1e4     # dt ~ 1e-10 < 1e-8 (mhd.c:131)
    u0[1] = bad
    ctx = m.Context(nx, ny, batch=B)
    for b in range(B):
        ctx.set_state(u0[b], inst=b)
    ctx.run(4)
    assert ctx.stopped(1)[0] and not ctx.stopped(0)[0] and not ctx.stopped(2)[0]
    np.testing.assert_array_equal(ctx.state(1), bad)
    assert ctx.scalars(1)[2] == 0
    for b in (0, 2):
        S = orc.Sim(nx, ny)
        S.set_state(u0[b])
        S.run(4)
        assert_fields_close(ctx.state(b), S.state(), what=f"inst {b} next to a stopped one")


def test_restart_one_instance(m):
    nx, ny, B = 64, 32, 2
    u0 = _states(m, nx, ny, B)
    ctx = m.Context(nx, ny, batch=B)
    for b in range(B):
        ctx.set_state(u0[b], inst=b)
    ctx.run(3)
    ctx.set_state(u0[1], inst=1)      # instance 1 starts over
    ctx.run(2)
    S0 = orc.Sim(nx, ny)
    S0.set_state(u0[0])
    S0.run(5)
    S1 = orc.Sim(nx, ny)
    S1.set_state(u0[1])
    S1.run(2)
    assert_fields_close(ctx.state(0), S0.state(), what="untouched instance")
    assert_fields_close(ctx.state(1), S1.state(), what="restarted instance")
    assert ctx.scalars(0)[2] == 5 and ctx.scalars(1)[2] == 2


def test_batch_diagnostics_per_instance(m):
    nx, ny, B = 64, 64, 3
    u0 = _states(m, nx, ny, B)
    ctx = m.Context(nx, ny, batch=B)
    for b in range(B):
        ctx.set_state(u0[b], inst=b)
    dx = dy = 2 * np.pi / 64
    for b in range(B):
        e, _ = ctx.diagnostics(b)
        assert e == pytest.approx(orc.energy(u0[b], dx, dy), rel=1e-13)


def test_ensemble_shape_fp32_64x512(m):
    nx = ny = 512
    B, steps = 64, 3
    print(f"[batch] ensemble {B}x{nx}x{ny} fp32", flush=True)
    ctx = m.Context(nx, ny, hall=1, precision=m.FP32, batch=B)
    rng = np.random.default_rng(42)
    dis = rng.uniform(0.0, 0.2, B)
    states = []
    for b in range(B):
        u = m.linear_wave(nx, ny, seed=42 + b, amp=1e-4)
        states.append(u)
        ctx.set_di(b, dis[b])
        ctx.set_state(u, inst=b)
    ctx.run(steps, graph=True)
    ctx.sync()
    for b in (0, 37, 63):
        one = m.Context(nx, ny, hall=1, di=dis[b], precision=m.FP32)
        one.set_state(states[b])
        one.run(steps, graph=True)
        np.testing.assert_array_equal(ctx.state(b), one.state())
        assert np.isfinite(ctx.state(b)).all()
