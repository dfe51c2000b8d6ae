"""Instance groups of a batch (mhd2d_set_subbatch): a step runs the batch as
groups of k instances, one whole RK3 step per group.  The instances are
independent, so every group size gives the results of one launch over the
whole batch:

* bitwise equal states, dt, time and step counts for k = 1, 2, 3 of a batch
  of 5 (an uneven last group), eager and hipGraph, in FP64, mixed and FP32,
  ideal, with viscosity / resistivity, diffusive and GLM cleaning, periodic
  and open x on 2 blocks;
* the per-class profile still reports one launch over the batch per stage;
* the autotuner picks a group size (2 or the whole batch of 4 x 1024^2)
  and the tuned run equals the untuned one bit for bit.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def m():
    import mhd2dcfdfftw_open_amd as m
    m.lib()
    return m


def _states(m, nx, ny, B, glm, flow):
    out = []
    x = np.arange(nx) * 2 * np.pi / nx
    for b in range(B):
        u = m.turbulence(nx, ny, seed=30 + b, kmax=min(6, nx // 8, ny // 4), rms=0.1)
        u[1] += flow
        if glm:
            u[4] += 0.01 * np.sin(2 * x)[:, None]   # div B drives psi
            u = np.concatenate([u, np.zeros((1, nx, ny))])
        out.append(u)
    return out


SETTINGS = {
    "ideal_fp64": dict(kw=dict(hall=1), prec="fp64"),
    "ideal_mixed": dict(kw=dict(hall=1), prec="mixed"),
    "ideal_fp32": dict(kw=dict(hall=0), prec="fp32"),
    "dissipative_fp64": dict(kw=dict(hall=1, viscosity=2e-3, resistivity=1e-3), prec="fp64"),
    "diffusive_clean_mixed": dict(kw=dict(hall=1, resistivity=1e-3, divbclean=1),
                                  prec="mixed"),
    "glm_fp64": dict(kw=dict(hall=1, divbclean=2), prec="fp64"),
    "open_glm_fp64": dict(kw=dict(hall=0, divbclean=2, npe=2, open_x=True), prec="fp64"),
    "open_dissipative_mixed": dict(kw=dict(hall=1, viscosity=2e-3, npe=2, open_x=True),
                                   prec="mixed"),
}


def _ctx(m, nx, ny, B, s, states, dis):
    P = {"fp64": m.FP64, "mixed": m.MIXED, "fp32": m.FP32}[s["prec"]]
    c = m.Context(nx, ny, precision=P, batch=B, **s["kw"])
    for b in range(B):
        c.set_di(b, dis[b])
        c.set_state(states[b], inst=b)
    return c


@pytest.mark.parametrize("graph", [False, True])
@pytest.mark.parametrize("name", list(SETTINGS))
def test_groups_equal_whole_batch_bitwise(m, name, graph):
    nx, ny, B, steps = 64, 32, 5, 4
    s = SETTINGS[name]
    glm = s["kw"].get("divbclean") == 2
    print(f"[subbatch] {name} graph={graph}", flush=True)
    states = _states(m, nx, ny, B, glm, 0.3 if s["kw"].get("open_x") else 0.0)
    dis = [0.01, 0.05, 0.0, 0.12, 0.03]
    ref = _ctx(m, nx, ny, B, s, states, dis)
    assert ref.subbatch == B
    ref.run(steps, graph=graph)
    ref.sync()
    for k in (1, 2, 3):
        c = _ctx(m, nx, ny, B, s, states, dis)
        c.set_subbatch(k)
        assert c.subbatch == k
        c.run(steps - 1, graph=graph)
        c.run(1, graph=graph)        # a second replay after host work
        c.sync()
        for b in range(B):
            np.testing.assert_array_equal(c.state(b), ref.state(b), err_msg=f"k={k} inst {b}")
            assert c.scalars(b) == ref.scalars(b)


def test_group_size_bounds_and_profile(m):
    nx, ny, B = 64, 64, 4
    states = _states(m, nx, ny, B, False, 0.0)
    c = _ctx(m, nx, ny, B, SETTINGS["ideal_fp64"], states, [0.02] * B)
    c.set_subbatch(B + 3)          # at least the batch: the whole batch
    assert c.subbatch == B
    c.set_subbatch(0)
    assert c.subbatch == B
    with pytest.raises(m.MHD2DError):
        c.set_subbatch(-1)
    c.set_subbatch(1)
    c.run(1)
    p = c.profile_run(2)
    # one y-sweep, x-derivative (or carry solve), RHS + RK (or the fused
    # chunk kernel) and x-filter per stage, one dt
    xd, rk = ("xcarry", "xrhs_rk") if c.xfuse else ("xderiv", "rhs_rk")
    assert {k: v[1] for k, v in p.items()} == {"ysweep": 6, xd: 6, rk: 6,
                                                "xfilter": 6, "dt_finalize": 2}
    assert all(v[0] > 0 for v in p.values())


def test_autotune_picks_a_group_size_and_results_hold(m):
    n, B = 1024, 4
    print(f"[subbatch] autotune {B} x {n}^2", flush=True)
    states = [m.turbulence(n, n, seed=50 + b, rms=0.1, kmax=16) for b in range(B)]
    s = SETTINGS["ideal_fp64"]
    dis = [0.02] * B
    ref = _ctx(m, n, n, B, s, states, dis)
    ref.run(2)
    ref.sync()
    c = _ctx(m, n, n, B, s, states, dis)
    c.autotune()
    # groups of at least 2^21 cells: the whole batch or pairs
    assert c.subbatch in (2, B)
    print(f"[subbatch] chose {c.subbatch}")
    c.run(2, graph=True)
    c.sync()
    for b in range(B):
        np.testing.assert_array_equal(c.state(b), ref.state(b))
        assert c.scalars(b) == ref.scalars(b)
