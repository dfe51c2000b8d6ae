"""Fused x-derivative + RHS + RK (mhd2d_set_xfuse).

Each stage's x-derivative runs as a solve that keeps only the values entering
every chunk of rows (the carries), then one kernel re-solves each chunk from
them in LDS and applies the RHS + RK update; the x-derivative arrays never
reach device memory.  The chunk re-solve repeats the full solve's
recurrences, so the fused path must give the unfused one's results bit for
bit (which the other GPU tests hold to the oracle):

* bitwise equal states, dt, time and step counts after several steps, eager
  and hipGraph, in FP64, mixed and FP32, for every chunk length (8, 16 and
  32 rows; 4096 rows as 128 chunks) and for the narrowest grids a fused
  workgroup tiles;
* a batch with per-instance d_i, run as instance groups;
* one run against the oracle directly (64 x 64 Orszag-Tang, Hall);
* refused (with a message) where it does not apply: source terms, div B
  cleaning, x blocks, open x, the GEMM x-derivative, and a grid whose
  y-lines a fused workgroup does not tile.
"""
import numpy as np
import pytest

from util import assert_fields_close

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def m():
    import mhd2dcfdfftw_open_amd as m
    m.lib()
    return m


def _prec(m, name):
    return {"fp64": m.FP64, "mixed": m.MIXED, "fp32": m.FP32}[name]


def _state(m, nx, ny, seed):
    return m.turbulence(nx, ny, seed=seed, kmax=min(6, nx // 8, ny // 4), rms=0.1)


def _run(m, nx, ny, prec, fuse, steps, graph, B=1, sub=0, dis=None, hall=1):
    c = m.Context(nx, ny, precision=_prec(m, prec), batch=B, hall=hall, di=0.05)
    for b in range(B):
        if dis is not None:
            c.set_di(b, dis[b])
        c.set_state(_state(m, nx, ny, 11 + b), inst=b)
    c.set_xfuse(1 if fuse else 0)
    assert c.xfuse == (1 if fuse else 0)
    if sub:
        c.set_subbatch(sub)
    c.run(steps, graph=graph)
    c.sync()
    out = [(c.state(b), c.scalars(b)) for b in range(B)]
    c.close()
    return out


# (nx, ny, precision): chunk C = nx / 32 clamped to 8..32 rows; a fused
# workgroup covers 512 / C columns (FP64 state) or 1024 / C (FP32)
CASES = [
    (128, 64, "fp64"), (128, 64, "mixed"), (128, 128, "fp32"),
    (1024, 32, "fp64"), (1024, 32, "mixed"), (1024, 64, "fp32"),
    (2048, 16, "fp64"), (2048, 16, "mixed"), (2048, 32, "fp32"),
    (4096, 32, "fp64"), (4096, 32, "mixed"), (4096, 32, "fp32"),
    (256, 512, "fp64"),
    # grids with many row chunks in flight at once: a chunk's update must
    # not reach the rows its neighbours still read as x halos
    (2048, 1024, "fp64"), (1024, 1024, "mixed"), (2048, 512, "fp32"),
]


@pytest.mark.parametrize("graph", [False, True])
@pytest.mark.parametrize("nx,ny,prec", CASES)
def test_fused_equals_unfused_bitwise(m, nx, ny, prec, graph):
    steps = 4
    print(f"[xfuse] {nx}x{ny} {prec} graph={graph}", flush=True)
    ref = _run(m, nx, ny, prec, False, steps, graph)
    got = _run(m, nx, ny, prec, True, steps, graph)
    for (ua, sa), (ub, sb) in zip(ref, got):
        assert sa == sb
        assert np.array_equal(ua, ub), np.max(np.abs(ua - ub))


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_fused_batch_groups_bitwise(m, prec):
    nx, ny, B, steps = 256, 128, 3, 5
    dis = [0.02, 0.0, 0.1]
    print(f"[xfuse] batch {nx}x{ny} B={B} {prec}", flush=True)
    ref = _run(m, nx, ny, prec, False, steps, False, B=B, dis=dis)
    got = _run(m, nx, ny, prec, True, steps, True, B=B, sub=2, dis=dis)
    for (ua, sa), (ub, sb) in zip(ref, got):
        assert sa == sb
        assert np.array_equal(ua, ub)


def test_fused_against_oracle(m):
    from oracle_lib import Sim
    nx = ny = 64
    steps = 20
    print(f"[xfuse] oracle {nx}x{ny} {steps} steps", flush=True)
    u0 = m.orszag_tang(nx, ny)
    c = m.Context(nx, ny, precision=m.FP64, hall=1, di=0.05)
    c.set_state(u0)
    c.set_xfuse(1)
    c.run(steps)
    c.sync()
    got = c.state()
    dt, t, n = c.scalars()
    c.close()
    sim = Sim(nx, ny, hall=1, di=0.05)
    sim.set_state(u0)
    sim.run(steps)
    assert n == steps
    assert abs(t - sim.time) <= 1e-12 * abs(sim.time)
    assert_fields_close(got, sim.state(), what="fused vs oracle")


def test_fused_by_default_in_fp64_and_fp32(m):
    for prec, want in (("fp64", 1), ("fp32", 1), ("mixed", 0)):
        c = m.Context(256, 128, precision=_prec(m, prec), hall=1, di=0.05)
        assert c.xfuse == want, prec
        c.close()
    c = m.Context(256, 128, precision=m.FP64, hall=1, di=0.05, viscosity=1e-3)
    assert c.xfuse == 0
    c.close()


REFUSED = {
    "viscosity": dict(viscosity=1e-3),
    "resistivity": dict(resistivity=1e-3),
    "diffusive_clean": dict(divbclean=1),
    "glm_clean": dict(divbclean=2),
    "x_blocks": dict(npe=2),
    "open_x": dict(npe=2, open_x=True),
    "gemm": dict(xgemm=6, precision="mixed"),
    "untiled_fp32": dict(precision="fp32", nx=2048, ny=16),
}


@pytest.mark.parametrize("name", list(REFUSED))
def test_fused_refused_where_it_does_not_apply(m, name):
    kw = dict(REFUSED[name])
    prec = kw.pop("precision", "fp64")
    nx, ny = kw.pop("nx", 128), kw.pop("ny", 64)
    c = m.Context(nx, ny, precision=_prec(m, prec), hall=1, di=0.05, **kw)
    with pytest.raises(Exception, match="fused x-derivative"):
        c.set_xfuse(1)
    assert c.xfuse == 0
    c.close()
