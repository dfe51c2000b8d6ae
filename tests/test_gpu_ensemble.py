"""Ensembles of runs on grids of different sizes (the uneven-ensemble
workload: sizes drawn from {256^2, 512x256, 1024^2, 2048x512}, d_i per run).

* runs are grouped by grid shape; every run matches its own oracle run at
  <= 1e-12 per field per step (FP64), eager and as one forked hipGraph;
* an ensemble is bitwise equal to the same runs made one context at a time
  (mixed and FP32);
* the 96-run uneven shape runs (mixed) and matches single runs bitwise;
* one launch per kernel class and stage over all shape groups: per-step
  oracle parity (eager and graph, and after one run is reset), bitwise equal
  to the per-group launches in FP64, mixed and FP32 with and without the
  Laplacian terms, 13 (16) launches a step, refused where it does not fit.
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


RUNS = [dict(nx=64, ny=64, hall=1, di=0.05), dict(nx=128, ny=64, hall=1, di=0.0),
        dict(nx=64, ny=64, hall=1, di=0.11), dict(nx=64, ny=128, hall=1, di=0.02),
        dict(nx=128, ny=64, hall=1, di=0.07)]


def _u0(m, r, k):
    return m.turbulence(r["nx"], r["ny"], seed=100 + k, kmax=6)


@pytest.mark.parametrize("graph", [False, True])
def test_ensemble_fp64_matches_oracle(m, graph):
    print(f"[ens] fp64 graph={graph}", flush=True)
    E = m.Ensemble(RUNS)
    assert E.groups == 3
    sims = []
    for k, r in enumerate(RUNS):
        u0 = _u0(m, r, k)
        E.set_state(k, u0)
        S = orc.Sim(r["nx"], r["ny"], hall=1, di=r["di"])
        S.set_state(u0)
        sims.append(S)
    for step in range(4):
        E.run(1, graph=graph)
        for k, S in enumerate(sims):
            S.run(1)
            assert_fields_close(E.state(k), S.state(), what=f"run {k} step {step + 1}")
            assert abs(E.scalars(k)[1] - S.time) <= 1e-12 * S.time
    E.run(3, graph=graph)
    for k, S in enumerate(sims):
        S.run(3)
        assert_fields_close(E.state(k), S.state(), what=f"run {k} step 7")


@pytest.mark.parametrize("prec", ["mixed", "fp32"])
def test_ensemble_equals_single_runs_bitwise(m, prec):
    P = m.MIXED if prec == "mixed" else m.FP32
    E = m.Ensemble(RUNS, precision=P)
    states = [_u0(m, r, k) for k, r in enumerate(RUNS)]
    for k in range(len(RUNS)):
        E.set_state(k, states[k])
    E.run(5, graph=True)
    for k, r in enumerate(RUNS):
        one = m.Context(r["nx"], r["ny"], hall=1, di=r["di"], precision=P)
        one.set_state(states[k])
        one.run(5, graph=True)
        np.testing.assert_array_equal(E.state(k), one.state())
        assert E.scalars(k) == one.scalars()


def test_uneven_ensemble_shape_96_runs(m):
    rng = np.random.default_rng(42)
    sizes = [(256, 256), (512, 256), (1024, 1024), (2048, 512)]
    runs, states = [], []
    for k in range(96):
        nx, ny = sizes[rng.integers(0, 4)]
        runs.append(dict(nx=nx, ny=ny, hall=1, di=float(rng.uniform(0, 0.1))))
        states.append(m.turbulence(nx, ny, seed=42 + k, rms=float(rng.uniform(0.05, 0.2)),
                                   kmax=16))
    print(f"[ens] 96 runs, {sum(r['nx'] * r['ny'] for r in runs) / 1e6:.1f} M cells", flush=True)
    E = m.Ensemble(runs, precision=m.MIXED)
    assert E.groups == len({(r["nx"], r["ny"]) for r in runs})
    for k in range(96):
        E.set_state(k, states[k])
    E.run(3, graph=True)
    E.sync()
    for k in (0, 47, 95):
        r = runs[k]
        one = m.Context(r["nx"], r["ny"], hall=1, di=r["di"], precision=m.MIXED)
        one.set_state(states[k])
        one.run(3, graph=True)
        np.testing.assert_array_equal(E.state(k), one.state())


# ---------------------------------------------------------------------------
# one launch per kernel class and stage over every shape group
# ---------------------------------------------------------------------------
RUNS_ONE = [dict(nx=256, ny=128, hall=1, di=0.05), dict(nx=512, ny=128, hall=1, di=0.0),
            dict(nx=256, ny=128, hall=1, di=0.11), dict(nx=256, ny=256, hall=1, di=0.02),
            dict(nx=1024, ny=128, hall=1, di=0.07)]


def _u0_one(m, r, k):
    return m.turbulence(r["nx"], r["ny"], seed=300 + k, kmax=8)


@pytest.mark.parametrize("graph", [False, True])
def test_one_launch_ensemble_matches_oracle(m, graph):
    print(f"[ens] one launch, fp64, graph={graph}", flush=True)
    E = m.Ensemble(RUNS_ONE)
    assert E.groups == 4
    assert E.launch_mode == m.ENS_ONE_LAUNCH
    assert E.launches_per_step == 13
    states = [_u0_one(m, r, k) for k, r in enumerate(RUNS_ONE)]
    for k in range(len(RUNS_ONE)):
        E.set_state(k, states[k])
    with orc.threaded():
        sims = []
        for k, r in enumerate(RUNS_ONE):
            S = orc.Sim(r["nx"], r["ny"], hall=1, di=r["di"])
            S.set_state(states[k])
            sims.append(S)
        for step in range(3):
            E.run(1, graph=graph)
            for k, S in enumerate(sims):
                S.run(1)
                assert_fields_close(E.state(k), S.state(), what=f"run {k} step {step + 1}")
                assert abs(E.scalars(k)[1] - S.time) <= 1e-12 * S.time
        # a new state for one run only: its group restarts, the others go on
        u_new = m.turbulence(256, 256, seed=999, kmax=8)
        E.set_state(3, u_new)
        sims[3] = orc.Sim(256, 256, hall=1, di=0.02)
        sims[3].set_state(u_new)
        E.run(3, graph=graph)
        for k, S in enumerate(sims):
            S.run(3)
            assert_fields_close(E.state(k), S.state(), what=f"run {k} after the reset")


@pytest.mark.parametrize("prec", ["fp64", "mixed", "fp32"])
@pytest.mark.parametrize("lap", [False, True])
def test_one_launch_equals_per_group_bitwise(m, prec, lap):
    P = {"fp64": m.FP64, "mixed": m.MIXED, "fp32": m.FP32}[prec]
    runs = [dict(r) for r in RUNS_ONE]
    if lap:
        for r in runs:
            r.update(viscosity=2e-3, resistivity=1e-3)
    print(f"[ens] one launch vs per group, {prec}, lap={lap}", flush=True)
    states = [_u0_one(m, r, k) for k, r in enumerate(runs)]
    out = {}
    for mode in (m.ENS_ONE_LAUNCH, m.ENS_PER_GROUP):
        E = m.Ensemble(runs, precision=P)
        E.set_launch(mode)
        for k in range(len(runs)):
            E.set_state(k, states[k])
        E.run(4, graph=True)
        E.run(2, graph=False)
        out[mode] = ([E.state(k) for k in range(len(runs))],
                     [E.scalars(k) for k in range(len(runs))], E.launches_per_step)
        E.close()
    one, per = out[m.ENS_ONE_LAUNCH], out[m.ENS_PER_GROUP]
    for k in range(len(runs)):
        np.testing.assert_array_equal(one[0][k], per[0][k])
        assert one[1][k] == per[1][k]
    assert one[2] == (16 if lap else 13)
    assert per[2] >= 4 * one[2]


def test_one_launch_refused_where_it_does_not_fit(m):
    E = m.Ensemble(RUNS)            # 64-row lines: per-group launches only
    assert E.launch_mode == m.ENS_PER_GROUP
    with pytest.raises(m.MHD2DError):
        E.set_launch(m.ENS_ONE_LAUNCH)
    E2 = m.Ensemble([dict(nx=256, ny=128, divbclean=1), dict(nx=512, ny=128, divbclean=1)])
    assert E2.launch_mode == m.ENS_PER_GROUP
    with pytest.raises(m.MHD2DError):
        E2.set_launch(m.ENS_ONE_LAUNCH)
