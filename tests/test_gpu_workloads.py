"""The benchmark workloads at their full sizes on the GPU against the FP64
oracle (its OpenMP build; results equal bit for bit to one thread):

* 2048 x 2048 turbulence (seed 42, rms 0.1, 1 <= |k| <= 16, d_i = 0.02),
  the fused FP64 path, 10 steps checked after every step: per-field
  relative L2 <= 1e-12, dt and time;
* 4096 x 4096 turbulence (the same seed): 2 FP64 steps at 1e-12, and the
  mixed form within 1e-6 of the oracle after the same 2 steps;
* the 64 x 512 x 512 batch of linear waves (d_i per instance): instances 0,
  31 and 63 of one batched context after 20 steps;
* the 96-run uneven ensemble (grid sizes drawn from 256 x 256, 512 x 256,
  1024 x 1024, 2048 x 512): the first run of each size after 10 steps.

The smaller grids, longer runs and every setting of the physics are covered
by the other parity tests; these pin the workloads the measurements are
quoted on, at the sizes they run.

With MHD2D_LONG_TESTS=1 the same checks also run over each workload's full
step count (2048^2 for 200 steps, checked every 20; 4096^2 for 100 steps,
checked every 50 in FP64, mixed and the mixed GEMM form; batch instances for
500 steps; the ensemble's runs for 200 steps): many minutes of checker
time, so the default suite skips them.
"""
import os

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


def test_turbulence_2048_every_step(m):
    n, steps = 2048, 10
    print(f"[workload] {n}^2 turbulence, FP64, {steps} steps", flush=True)
    u0 = m.turbulence(n, n, seed=42, rms=0.1, kmin=1, kmax=16)
    ctx = m.Context(n, n, hall=1, di=0.02)
    ctx.set_state(u0)
    worst = 0.0
    try:
        with orc.threaded():
            S = orc.Sim(n, n, hall=1, di=0.02)
            S.set_state(u0)
            for k in range(1, steps + 1):
                ctx.run(1, graph=True)
                S.run(1)
                e = assert_fields_close(ctx.state(), S.state(), what=f"step {k}")
                worst = max(worst, e.max())
                dt, t, nsteps = ctx.scalars()
                assert nsteps == k
                assert abs(dt - S.dt) <= FP64_TOL * S.dt, (k, dt, S.dt)
                assert abs(t - S.time) <= FP64_TOL * S.time, (k, t, S.time)
            del S
    finally:
        ctx.close()
    print(f"[workload] {n}^2: worst per-field rel-L2 over {steps} steps {worst:.2e}")


def test_turbulence_4096_fp64_and_mixed(m):
    n, steps = 4096, 2
    print(f"[workload] {n}^2 turbulence, FP64 and mixed, {steps} steps", flush=True)
    u0 = m.turbulence(n, n, seed=42, rms=0.1, kmin=1, kmax=16)
    with orc.threaded():
        S = orc.Sim(n, n, hall=1, di=0.02)
        S.set_state(u0)
        S.run(steps)
        want = S.state()
        del S
    for prec, tol in (("fp64", FP64_TOL), ("mixed", 1e-6)):
        ctx = m.Context(n, n, hall=1, di=0.02,
                        precision=m.FP64 if prec == "fp64" else m.MIXED)
        try:
            ctx.set_state(u0)
            ctx.run(steps)
            e = assert_fields_close(ctx.state(), want, tol=tol, what=prec)
        finally:
            ctx.close()
        print(f"[workload] {n}^2 {prec}: per-field rel-L2 max {e.max():.2e}")


def test_batch_64x512_instances(m):
    n, B, steps = 512, 64, 20
    print(f"[workload] {B} x {n}^2 linear waves, FP64, {steps} steps", flush=True)
    rng = np.random.default_rng(42)
    dis = rng.uniform(0.0, 0.2, B)
    states = [m.linear_wave(n, n, seed=42 + b, amp=1e-4) for b in range(B)]
    ctx = m.Context(n, n, hall=1, batch=B)
    try:
        for b in range(B):
            ctx.set_di(b, float(dis[b]))
            ctx.set_state(states[b], inst=b)
        ctx.run(steps, graph=True)
        ctx.sync()
        with orc.threaded():
            for b in (0, 31, 63):
                S = orc.Sim(n, n, hall=1, di=float(dis[b]))
                S.set_state(states[b])
                S.run(steps)
                e = assert_fields_close(ctx.state(b), S.state(), what=f"instance {b}")
                print(f"[workload] instance {b}: per-field rel-L2 max {e.max():.2e}")
                del S
    finally:
        ctx.close()


def test_uneven_ensemble_one_run_per_size(m):
    count, steps = 96, 10
    print(f"[workload] {count}-run uneven ensemble, FP64, {steps} steps", flush=True)
    rng = np.random.default_rng(42)
    sizes = [(256, 256), (512, 256), (1024, 1024), (2048, 512)]
    runs, states = [], []
    for k in range(count):
        nx, ny = sizes[rng.integers(0, 4)]
        runs.append(dict(nx=nx, ny=ny, hall=1, di=float(rng.uniform(0, 0.1))))
        states.append(m.turbulence(nx, ny, seed=42 + k, rms=float(rng.uniform(0.05, 0.2)),
                                   kmax=16))
    E = m.Ensemble(runs)
    try:
        assert E.groups == 4
        for k in range(count):
            E.set_state(k, states[k])
        E.run(steps, graph=True)
        E.sync()
        first = {}
        for k, r in enumerate(runs):
            first.setdefault((r["nx"], r["ny"]), k)
        assert len(first) == 4
        with orc.threaded():
            for (nx, ny), k in sorted(first.items()):
                S = orc.Sim(nx, ny, hall=1, di=runs[k]["di"])
                S.set_state(states[k])
                S.run(steps)
                e = field_errors(E.state(k), S.state())
                print(f"[workload] run {k} ({nx}x{ny}): per-field rel-L2 max {e.max():.2e}")
                assert np.all(e <= FP64_TOL), (k, e)
                del S
    finally:
        E.close()


LONG = pytest.mark.skipif(not os.environ.get("MHD2D_LONG_TESTS"),
                          reason="full step counts: minutes of checker time (MHD2D_LONG_TESTS=1)")


@LONG
def test_turbulence_2048_200_steps(m):
    n, steps, every = 2048, 200, 20
    print(f"[workload] {n}^2 turbulence, FP64, {steps} steps", flush=True)
    u0 = m.turbulence(n, n, seed=42, rms=0.1, kmin=1, kmax=16)
    ctx = m.Context(n, n, hall=1, di=0.02)
    ctx.set_state(u0)
    worst = 0.0
    try:
        with orc.threaded():
            S = orc.Sim(n, n, hall=1, di=0.02)
            S.set_state(u0)
            for k in range(every, steps + 1, every):
                ctx.run(every, graph=True)
                S.run(every)
                e = assert_fields_close(ctx.state(), S.state(), what=f"step {k}")
                worst = max(worst, e.max())
                dt, t, nsteps = ctx.scalars()
                assert nsteps == k
                assert abs(dt - S.dt) <= FP64_TOL * S.dt, (k, dt, S.dt)
                assert abs(t - S.time) <= FP64_TOL * S.time, (k, t, S.time)
                print(f"[workload] step {k}: per-field rel-L2 max {e.max():.2e}", flush=True)
            del S
    finally:
        ctx.close()
    print(f"[workload] {n}^2: worst per-field rel-L2 over {steps} steps {worst:.2e}")


@LONG
def test_batch_64x512_500_steps(m):
    n, B, steps = 512, 64, 500
    print(f"[workload] {B} x {n}^2 linear waves, FP64 and mixed, {steps} steps", flush=True)
    rng = np.random.default_rng(42)
    dis = rng.uniform(0.0, 0.2, B)
    states = [m.linear_wave(n, n, seed=42 + b, amp=1e-4) for b in range(B)]
    ctxs = {}
    try:
        for prec in ("fp64", "mixed"):
            c = m.Context(n, n, hall=1, batch=B, precision=m.FP64 if prec == "fp64" else m.MIXED)
            for b in range(B):
                c.set_di(b, float(dis[b]))
                c.set_state(states[b], inst=b)
            c.run(steps, graph=True)
            c.sync()
            ctxs[prec] = c
        with orc.threaded():
            for b in (0, 63):
                S = orc.Sim(n, n, hall=1, di=float(dis[b]))
                S.set_state(states[b])
                S.run(steps)
                want = S.state()
                del S
                e = assert_fields_close(ctxs["fp64"].state(b), want, what=f"fp64 instance {b}")
                em = assert_fields_close(ctxs["mixed"].state(b), want, tol=1e-6,
                                         what=f"mixed instance {b}")
                print(f"[workload] instance {b}: FP64 {e.max():.2e}, mixed {em.max():.2e}")
    finally:
        for c in ctxs.values():
            c.close()


@LONG
def test_uneven_ensemble_200_steps(m):
    count, steps = 96, 200
    print(f"[workload] {count}-run uneven ensemble, FP64 and mixed, {steps} steps", flush=True)
    rng = np.random.default_rng(42)
    sizes = [(256, 256), (512, 256), (1024, 1024), (2048, 512)]
    runs, states = [], []
    for k in range(count):
        nx, ny = sizes[rng.integers(0, 4)]
        runs.append(dict(nx=nx, ny=ny, hall=1, di=float(rng.uniform(0, 0.1))))
        states.append(m.turbulence(nx, ny, seed=42 + k, rms=float(rng.uniform(0.05, 0.2)),
                                   kmax=16))
    first = {}
    for k, r in enumerate(runs):
        first.setdefault((r["nx"], r["ny"]), k)
    ens = {}
    try:
        for prec in ("fp64", "mixed"):
            E = m.Ensemble(runs, precision=m.FP64 if prec == "fp64" else m.MIXED)
            for k in range(count):
                E.set_state(k, states[k])
            E.run(steps, graph=True)
            E.sync()
            ens[prec] = E
        with orc.threaded():
            for (nx, ny), k in sorted(first.items()):
                S = orc.Sim(nx, ny, hall=1, di=runs[k]["di"])
                S.set_state(states[k])
                S.run(steps)
                want = S.state()
                del S
                e = field_errors(ens["fp64"].state(k), want)
                em = field_errors(ens["mixed"].state(k), want)
                print(f"[workload] run {k} ({nx}x{ny}): FP64 {e.max():.2e}, mixed {em.max():.2e}")
                assert np.all(e <= FP64_TOL), (k, e)
                assert np.all(em <= 1e-6), (k, em)
    finally:
        for E in ens.values():
            E.close()


@LONG
def test_turbulence_4096_100_steps_precision_forms(m):
    """the precision comparison's workload: per-field relative L2 against the
    oracle every 50 steps for 100 steps, FP64 (1e-12), mixed with the
    tridiagonal x-derivative and mixed with the 6-product split-BF16 GEMM
    (1e-6 each); the checker takes ~5 s a step at this size on 16 host
    cores, so the 200 steps of the workload do not fit one run's time"""
    n, steps, every = 4096, 100, 50
    print(f"[workload] {n}^2 turbulence, FP64 / mixed / mixed GEMM, {steps} steps", flush=True)
    u0 = m.turbulence(n, n, seed=42, rms=0.1, kmin=1, kmax=16)
    forms = {"fp64": (dict(precision=m.FP64), FP64_TOL),
             "mixed": (dict(precision=m.MIXED), 1e-6),
             "mixed_gemm6": (dict(precision=m.MIXED, xgemm=6), 1e-6)}
    ctxs = {}
    try:
        for name, (kw, _) in forms.items():
            c = m.Context(n, n, hall=1, di=0.02, **kw)
            c.set_state(u0)
            ctxs[name] = c
        with orc.threaded():
            S = orc.Sim(n, n, hall=1, di=0.02)
            S.set_state(u0)
            for k in range(every, steps + 1, every):
                for c in ctxs.values():
                    c.run(every, graph=True)
                for j in range(0, every, 10):
                    S.run(10)
                    print(f"[workload] checker at step {k - every + j + 10}", flush=True)
                want = S.state()
                for name, (_, tol) in forms.items():
                    e = assert_fields_close(ctxs[name].state(), want, tol=tol,
                                            what=f"{name} step {k}")
                    print(f"[workload] step {k} {name}: per-field rel-L2 max {e.max():.2e}",
                          flush=True)
            del S
    finally:
        for c in ctxs.values():
            c.close()
