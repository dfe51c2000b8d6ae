"""Viscosity, resistivity and diffusive div B cleaning in the fused engine
(addSourceTerm, calcRHS.c:498-543): the second derivatives of u and B come
from the y-sweep (filter x d2/dy2 of the spectrum) and an x-solve that adds
d2/dx2; div B of the stage's derivatives gets its own x-solve and y
transform (mhd.c:175-182); all on every cell including the open-boundary
lines.  The GLM cleaning (divBcleaning 2, calcRHS.c:144-158) adds psi as a
ninth field with Ch / Cp per x block from the last cell of the block
(timeAdvance.c:107-112); the open-boundary lines get the psi terms only.

* FP64: per-field rel-L2 <= 1e-12 against the oracle every step, periodic
  and open x (npe 2 and 3), with viscosity only, resistivity only, both;
* graph replay and batches equal eager single runs;
* the mixed form within 1e-6 of FP64 after 40 steps;
* the reference's own `input` (open x, resistivity 1e-4) runs in the fused
  engine of the run loop and matches the oracle;
* GLM: psi itself relatively within 1e-8 of the oracle (its conditioning
  floor is ~1e-9: it is driven by div B, a cancellation of O(1)
  derivatives), the other fields at 1e-12.
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


def _state(m, nx, ny, flow=0.0, seed=3):
    u = m.turbulence(nx, ny, seed=seed, kmax=min(6, nx // 8, ny // 4), rms=0.1)
    u[1] += flow
    return u


CASES = [
    # nx, ny, npe, open, viscosity, resistivity, divbclean, hall, steps
    (64, 32, 1, False, 2e-3, 0.0, 0, 1, 6),
    (64, 64, 1, False, 0.0, 1e-3, 0, 1, 6),
    (128, 32, 1, False, 5e-3, 3e-3, 0, 0, 6),
    (64, 32, 2, True, 2e-3, 1e-3, 0, 1, 6),
    (128, 32, 3, True, 0.0, 2e-3, 0, 0, 5),
    (64, 64, 1, False, 0.0, 0.0, 1, 1, 6),
    (128, 32, 1, False, 2e-3, 1e-3, 1, 0, 5),
    (64, 32, 2, True, 0.0, 1e-3, 1, 1, 6),
    (64, 32, 1, False, 0.0, 0.0, 2, 1, 6),
    (64, 64, 1, False, 2e-3, 1e-3, 2, 0, 6),
    (128, 32, 3, False, 0.0, 0.0, 2, 1, 5),
    (64, 32, 2, True, 0.0, 0.0, 2, 0, 6),
    (128, 32, 3, True, 2e-3, 1e-3, 2, 1, 5),
    # longer y-lines: the y-sweep's Laplacian and GLM instances at 512, 1024,
    # 2048 and 4096 points (one row per workgroup from 2048 on; 16 elements
    # per thread at 4096)
    (64, 512, 1, False, 2e-3, 1e-3, 0, 1, 4),
    (32, 2048, 1, False, 1e-3, 1e-3, 0, 0, 3),
    (32, 1024, 1, False, 0.0, 0.0, 2, 1, 3),
    (32, 2048, 1, False, 1e-3, 1e-3, 2, 0, 3),
    (32, 4096, 1, False, 1e-3, 1e-3, 0, 1, 3),
    (32, 4096, 1, False, 1e-3, 1e-3, 2, 0, 3),
]


def _with_divb(u0, nx, amp=0.01):
    # a div B part (dBx/dx != 0) so that psi is driven by more than round-off
    u = u0.copy()
    x = np.arange(nx) * 2 * np.pi / nx
    u[4] += amp * np.sin(2 * x)[:, None]
    return u


@pytest.mark.parametrize("nx,ny,npe,open_x,nu,eta,clean,hall,steps", CASES)
def test_dissipative_matches_oracle(m, nx, ny, npe, open_x, nu, eta, clean, hall, steps):
    print(f"[dissipative] {nx}x{ny} npe={npe} open={open_x} nu={nu} eta={eta} clean={clean}",
          flush=True)
    u0 = _state(m, nx, ny, 0.3 if open_x else 0.0)
    kw = dict(hall=hall, di=0.05, viscosity=nu, resistivity=eta, divbclean=clean)
    ctx = m.Context(nx, ny, npe=npe, open_x=open_x, **kw)
    S = orc.Sim(nx, ny, npe=npe, open_x=open_x, **kw)
    if clean == 2:
        u0 = np.concatenate([_with_divb(u0, nx), np.zeros((1, nx, ny))])
    ctx.set_state(u0)
    S.set_state(u0)
    for k in range(steps):
        ctx.run(1)
        S.run(1)
        got, want = ctx.state(), S.state()
        assert_fields_close(got, want, what=f"step {k + 1}")
        if clean == 2:
            psi_err = np.linalg.norm(got[8] - want[8]) / np.linalg.norm(want[8])
            assert psi_err <= 1e-8, f"step {k + 1}: psi rel-L2 {psi_err}"
        assert abs(ctx.scalars()[1] - S.time) <= 1e-12 * S.time


@pytest.mark.parametrize("kw", [dict(resistivity=5e-3, viscosity=5e-3), dict(divbclean=1),
                                dict(divbclean=2)])
def test_dissipative_changes_the_state(m, kw):
    # the source terms are really applied: the run departs from the ideal one
    # by far more than round-off (div B of the seeded state is round-off
    # only, so cleaning gets a state with a div B part added)
    u0 = _state(m, 64, 64)
    if kw.get("divbclean"):
        u0 = _with_divb(u0, 64)
    a = m.Context(64, 64, hall=1, di=0.05)
    b = m.Context(64, 64, hall=1, di=0.05, **kw)
    a.set_state(u0)
    b.set_state(np.concatenate([u0, np.zeros((1, 64, 64))]) if kw.get("divbclean") == 2
                else u0)
    for c in (a, b):
        c.run(4)
    assert field_errors(a.state(), b.state()[:8]).max() > 1e-6


def _glm_state(u, clean):
    if clean != 2:
        return u
    return np.concatenate([_with_divb(u, u.shape[1]), np.zeros((1,) + u.shape[1:])])


@pytest.mark.parametrize("clean", [0, 1, 2])
def test_dissipative_graph_and_batch(m, clean):
    u0 = _glm_state(_state(m, 64, 32, seed=4), clean)
    u1 = _glm_state(_state(m, 64, 32, seed=5), clean)
    kw = dict(hall=1, di=0.05, viscosity=2e-3, resistivity=1e-3, divbclean=clean)
    eager = m.Context(64, 32, **kw)
    eager.set_state(u0)
    eager.run(5)
    g = m.Context(64, 32, **kw)
    g.set_state(u0)
    g.run(5, graph=True)
    np.testing.assert_array_equal(g.state(), eager.state())
    b = m.Context(64, 32, batch=2, **kw)
    b.set_state(u1, inst=0)
    b.set_state(u0, inst=1)
    b.run(5)
    np.testing.assert_array_equal(b.state(inst=1), eager.state())


@pytest.mark.parametrize("prec,clean,open_x,steps", [
    ("fp32", 0, False, 20), ("fp32", 1, False, 20), ("fp32", 2, False, 20),
    ("fp32", 1, True, 20), ("fp32", 2, True, 3), ("mixed", 2, True, 20)])
def test_dissipative_reduced_precision_within_bound(m, prec, clean, open_x, steps):
    # the FP32 and mixed instances of the source-term and GLM kernels against
    # the same run in FP64: FP32 within its bound of the ideal run (1e-4;
    # 2e-4 with open ends and cleaning, measured 8.6e-5), mixed within 1e-6.  psi, a cancellation of O(1) derivatives, keeps
    # ~1e-3 of itself in FP32 (measured 3e-4..8e-4 periodic, 5e-3 after 3
    # steps with open ends).  FP32 with GLM and open ends is checked over 3
    # steps only: psi grows at the open lines in the reference's scheme (see
    # the test below) and carries the FP32 difference with it.
    nx, ny = 128, 64
    npe = 2 if open_x else 1
    u0 = _glm_state(_state(m, nx, ny, 0.3 if open_x else 0.0, seed=9), clean)
    kw = dict(hall=1, di=0.05, viscosity=2e-3, resistivity=1e-3, divbclean=clean, npe=npe,
              open_x=open_x)
    ref = m.Context(nx, ny, **kw)
    low = m.Context(nx, ny, precision=m.FP32 if prec == "fp32" else m.MIXED, **kw)
    for c in (ref, low):
        c.set_state(u0)
        c.run(steps)
        c.sync()
    e = field_errors(low.state(), ref.state())
    print(f"[dissipative] {prec} clean={clean} open={open_x} {steps} steps: {e.max():.2e}")
    assert e.max() <= ((2e-4 if open_x else 1e-4) if prec == "fp32" else 1e-6), e
    if clean == 2 and prec == "fp32":
        a, b = low.state()[8], ref.state()[8]
        assert np.linalg.norm(a - b) <= (2e-2 if open_x else 5e-3) * np.linalg.norm(b)


def test_open_glm_psi_growth_tracks_oracle(m):
    # With open x ends the reference gives the boundary lines the psi terms
    # only (no characteristic treatment), and psi grows there: ~1e-3 to ~1
    # in 40 steps on this state (the oracle shows the same).  The device run
    # follows the oracle through that growth, which amplifies round-off.
    nx, ny = 128, 64
    kw = dict(npe=2, open_x=True, divbclean=2, hall=1, di=0.05)
    u0 = _glm_state(_state(m, nx, ny, 0.3, seed=9), 2)
    c = m.Context(nx, ny, **kw)
    c.set_state(u0)
    S = orc.Sim(nx, ny, **kw)
    S.set_state(u0)
    c.run(40)
    S.run(40)
    a, b = c.state(), S.state()
    edge = np.sqrt(np.mean(b[8][:4] ** 2 + b[8][-4:] ** 2) / 2)
    assert edge > 0.1, edge    # the growth this test is about
    assert field_errors(a, b)[:8].max() <= 1e-9
    assert np.linalg.norm(a[8] - b[8]) <= 1e-8 * np.linalg.norm(b[8])


@pytest.mark.parametrize("clean", [0, 1, 2])
def test_dissipative_mixed_precision(m, clean):
    u0 = _glm_state(_state(m, 128, 64, seed=6), clean)
    kw = dict(hall=1, di=0.05, viscosity=2e-3, resistivity=1e-3, divbclean=clean)
    ref = m.Context(128, 64, **kw)
    mix = m.Context(128, 64, precision=m.MIXED, **kw)
    for c in (ref, mix):
        c.set_state(u0)
        c.run(40)
    e = field_errors(mix.state(), ref.state())
    print(f"[dissipative] mixed 40 steps: {e.max():.2e}")
    assert e.max() <= 1e-6, e


@pytest.mark.parametrize("clean", [0, 2])
def test_reference_input_runs_fused(m, tmp_path, clean):
    # the reference's own `input`: open x (npe >= 2), resistivity 1e-4,
    # Hall off; here on a 64 x 32 grid of its Lx = 2, Ly = 4 box (and the
    # same with GLM cleaning, whose output files hold psi as field 9)
    import array
    import os
    inp = m.Input(boundaryTypeX=1, cfl=0.5, Tmax=0.02, dtOut=0.01, nx=64, ny=32,
                  Lx=2.0, Ly=4.0, gamma=1.666666667, hall=0, di=0.002,
                  resistivity=1e-4, viscosity=0.0, divbclean=clean, outdivb=1)
    u0 = _state(m, 64, 32, 0.2)
    rep = m.drive(inp, u0, outdir=str(tmp_path), mode=m.DRIVE_AUTO, npe=2)
    S = orc.Sim(64, 32, 2.0, 4.0, 1.666666667, 0.5, 0, 0.002, 1e-4, 0.0, clean, npe=2,
                open_x=True)
    S.set_state(u0)
    S.run(rep["steps"])
    nv = 9 if clean == 2 else 8
    a = array.array("d")
    with open(os.path.join(str(tmp_path), f"out{rep['outputs'] - 1:04d}.dat"), "rb") as f:
        a.fromfile(f, nv * 64 * 32 + 1)
        assert f.read() == b""
    assert abs(a[0] - S.time) <= 1e-12 * S.time
    assert_fields_close(np.reshape(np.asarray(a[1:]), [nv, 64, 32]), S.state(),
                        what="reference input, final output")


def test_dissipative_ensemble_matches_single_runs(m):
    # an ensemble of grid shapes with the source terms and the diffusive or
    # GLM cleaning equals the same runs made one context at a time
    kw = dict(hall=1, viscosity=2e-3, resistivity=1e-3)
    runs = [dict(nx=64, ny=32, di=0.02, divbclean=1, **kw),
            dict(nx=128, ny=64, di=0.05, divbclean=1, **kw),
            dict(nx=64, ny=32, di=0.04, divbclean=1, **kw),
            dict(nx=64, ny=32, di=0.03, divbclean=2, **kw)]
    states = [_glm_state(_state(m, r["nx"], r["ny"], seed=20 + k), r["divbclean"])
              for k, r in enumerate(runs)]
    ens = m.Ensemble(runs)
    assert ens.groups == 3
    for k, u in enumerate(states):
        ens.set_state(k, u)
    ens.run(4, graph=True)
    for k, r in enumerate(runs):
        c = m.Context(r["nx"], r["ny"], hall=1, di=r["di"], viscosity=2e-3, resistivity=1e-3,
                      divbclean=r["divbclean"])
        c.set_state(states[k])
        c.run(4)
        np.testing.assert_array_equal(ens.state(k), c.state())
