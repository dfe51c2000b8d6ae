"""Results of a reference run on npe MPI ranks, reproduced on one GPU.

The reference splits x into npe blocks (decomposeGrid1D, initialize.c:547-569)
and solves each block's compact systems on its own, with explicit end rows
fed by the 2-row halos of its neighbours (mathlib.c:46-86, parallel.c:4-24),
so its results depend on npe.  The context API's `xblocks` = npe builds that
block structure into the x-solves; the oracle's npe mode (pinned against a
dense block assembly on the CPU) is the check.

* per-field rel-L2 <= 1e-12 every step for even blocks (chunk-aligned) and
  uneven ones (the first nx % npe blocks one row larger, edges inside chunks);
* hipGraph replay bitwise equal to eager launches;
* npe changes the answer at the block edges (the closure is really there);
* FP32 with blocks stays within the FP32 bound of the one-block case;
* the run loop (mhd2d_drive) with npe against the oracle's loop.
"""
import array
import os

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


CASES = [
    # (nx, ny, npe, ic, steps)
    (64, 64, 2, "ot", 12),
    (64, 32, 4, "turb", 8),
    (128, 64, 8, "turb", 6),
    (64, 64, 3, "ot", 10),      # blocks 22, 21, 21
    (128, 32, 5, "turb", 6),    # blocks 26, 26, 26, 25, 25
    (256, 32, 7, "ot", 4),      # C = 8: edges at every residue mod 8
    (2048, 16, 16, "turb", 2),  # the 2048-row x-solve shape (C = 32)
]


def _ic(m, nx, ny, ic):
    if ic == "ot":
        return m.orszag_tang(nx, ny)
    return m.turbulence(nx, ny, seed=3, kmax=min(8, nx // 4, ny // 4))


@pytest.mark.parametrize("nx,ny,npe,ic,steps", CASES)
def test_blocks_match_oracle_npe(m, nx, ny, npe, ic, steps):
    print(f"[npe] {nx}x{ny} npe={npe} {ic}", flush=True)
    u0 = _ic(m, nx, ny, ic)
    kw = dict(hall=1, di=0.05)
    ctx = m.Context(nx, ny, npe=npe, **kw)
    ctx.set_state(u0)
    S = orc.Sim(nx, ny, npe=npe, **kw)
    S.set_state(u0)
    for k in range(steps):
        ctx.run(1)
        S.run(1)
        assert_fields_close(ctx.state(), S.state(), what=f"npe={npe} step {k + 1}")
        _, t, _ = ctx.scalars()
        assert abs(t - S.time) <= 1e-12 * S.time


def test_blocks_graph_replay_equals_eager(m):
    nx, ny, npe = 128, 64, 4
    u0 = _ic(m, nx, ny, "turb")
    a = m.Context(nx, ny, npe=npe, hall=1, di=0.05)
    b = m.Context(nx, ny, npe=npe, hall=1, di=0.05)
    a.set_state(u0)
    b.set_state(u0)
    a.run(6)
    b.run(6, graph=True)
    np.testing.assert_array_equal(a.state(), b.state())


def test_npe_changes_the_result_at_block_edges(m):
    nx, ny, npe = 64, 64, 4
    u0 = _ic(m, nx, ny, "ot")
    one = m.Context(nx, ny)
    four = m.Context(nx, ny, npe=npe)
    one.set_state(u0)
    four.set_state(u0)
    one.run(3)
    four.run(3)
    d = np.abs(one.state() - four.state()).max(axis=(0, 2))   # per x row
    edges = [0, 15, 16, 31, 32, 47, 48, 63]
    assert d[edges].max() > 1e-6
    assert d[edges].max() > 10 * np.median(d)


def test_blocks_fp32_within_bound(m):
    nx, ny, npe = 128, 64, 4
    u0 = _ic(m, nx, ny, "turb")
    ctx = m.Context(nx, ny, npe=npe, precision=m.FP32, hall=1, di=0.05)
    ctx.set_state(u0)
    S = orc.Sim(nx, ny, npe=npe, hall=1, di=0.05)
    S.set_state(u0)
    ctx.run(10)
    S.run(10)
    e = field_errors(ctx.state(), S.state())
    assert e.max() < 5e-5, e


def test_xblocks_out_of_range_is_rejected(m):
    with pytest.raises(Exception):
        m.Context(64, 64, npe=17)   # blocks of fewer than 4 rows


def test_drive_with_npe_matches_oracle_loop(m, tmp_path):
    inp = m.Input(boundaryTypeX=0, cfl=0.3, Tmax=0.04, dtOut=0.02, nx=64, ny=32,
                  Lx=2 * np.pi, Ly=2 * np.pi, gamma=5.0 / 3.0, hall=1, di=0.05,
                  resistivity=0.0, viscosity=0.0, divbclean=0, outdivb=0)
    u0 = m.orszag_tang(64, 32)
    rep = m.drive(inp, u0, outdir=str(tmp_path), mode=m.DRIVE_FUSED, npe=4)
    S = orc.Sim(64, 32, hall=1, di=0.05, npe=4)
    S.set_state(u0)
    S.run(rep["steps"])
    a = array.array("d")
    last = rep["outputs"] - 1
    with open(os.path.join(str(tmp_path), f"out{last:04d}.dat"), "rb") as f:
        a.fromfile(f, 8 * 64 * 32 + 1)
    assert abs(a[0] - S.time) <= 1e-12 * S.time
    assert_fields_close(np.reshape(np.asarray(a[1:]), [8, 64, 32]), S.state(),
                        what="drive npe=4 final output")
    with pytest.raises(Exception):
        m.drive(inp, u0, outdir=str(tmp_path), mode=m.DRIVE_DROPIN, npe=4)
