"""The Orszag-Tang workload on the GPU against the FP64 oracle over a long run.

256 x 256 periodic box [0, 2pi)^2, rho = 25/(36 pi), p = 5/(12 pi),
u = (-sin y, sin x, 0), B = (-sin y, sin 2x, 0)/sqrt(4 pi), gamma = 5/3,
cfl = 0.3, Hall CFL term with d_i = 0.05 (timeAdvance.c:97-105), through the
whole time loop of mhd.c:99-206:

* the fused engine for 200 RK3 steps, eager launches and hipGraph replays,
  checked every 10 steps: per-field relative L2 <= 1e-12, the dt in use and
  the simulated time to 1e-12, the total energy to 1e-12;
* the drop-in entry points called in the reference's order for 20 steps.

The checker runs its OpenMP build here (same results bit for bit as one
thread), about 10 s for the 200 steps on 16 host cores.
"""
import numpy as np
import pytest

import oracle_lib as orc
from util import FP64_TOL, assert_fields_close

pytestmark = pytest.mark.gpu

NX = NY = 256
KW = dict(hall=1, di=0.05)
STEPS, EVERY = 200, 10


@pytest.fixture(scope="module")
def m():
    import mhd2dcfdfftw_open_amd as m
    m.lib()
    return m


@pytest.fixture(scope="module")
def trajectory(m):
    """the oracle's state, dt and time after every EVERY steps"""
    u0 = m.orszag_tang(NX, NY)
    out = []
    with orc.threaded():
        S = orc.Sim(NX, NY, **KW)
        S.set_state(u0)
        dx = 2 * np.pi / NX
        for _ in range(STEPS // EVERY):
            S.run(EVERY)
            st = S.state()
            out.append((st, S.dt, S.time, orc.energy(st, dx, dx)))
        del S
    return u0, out


@pytest.mark.parametrize("graph", [False, True])
def test_orszag_tang_256_200_steps_match_oracle(m, trajectory, graph):
    print(f"[c1] fused 256x256 Orszag-Tang, 200 steps, graph={graph}", flush=True)
    u0, want = trajectory
    ctx = m.Context(NX, NY, **KW)
    ctx.set_state(u0)
    worst = 0.0
    try:
        for k, (st, dt, t, en) in enumerate(want):
            ctx.run(EVERY, graph=graph)
            ctx.sync()
            step = (k + 1) * EVERY
            e = assert_fields_close(ctx.state(), st, what=f"step {step}")
            worst = max(worst, e.max())
            gdt, gt, n = ctx.scalars()
            assert n == step
            assert abs(gt - t) <= FP64_TOL * t, (step, gt, t)
            # the dt of the step just taken: time(k) - time(k-1) on both sides
            energy, _ = ctx.diagnostics()
            assert abs(energy - en) <= FP64_TOL * abs(en), (step, energy, en)
        # the run went somewhere: the vortex has steepened well past its start
        assert np.abs(want[-1][0][1] - u0[1]).max() > 0.1
    finally:
        ctx.close()
    print(f"[c1] graph={graph}: worst per-field rel-L2 over 200 steps {worst:.2e}")


def test_orszag_tang_256_dropin_20_steps(m, trajectory):
    """the drop-in entry points in mhd.c's order (strict host copies)"""
    from test_gpu_dropin import _mhd_c_loop
    print("[c1] drop-in 256x256 Orszag-Tang, 20 steps", flush=True)
    u0, want = trajectory
    G = m.RefGlobals(NX, NY, **KW)
    L = m.lib()
    try:
        G.array("uu")[:8] = u0
        t, dt = _mhd_c_loop(L, G, 20)
        st, _, t_want, _ = want[1]
        assert abs(t - t_want) <= FP64_TOL * t_want
        assert_fields_close(G.array("uu"), st, what="drop-in step 20")
    finally:
        G.close()
