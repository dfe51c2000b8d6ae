"""The x-solve shape autotuner (mhd2d_autotune / mhd2d_ens_autotune): it
returns valid shapes, leaves the state untouched, and a tuned run equals an
untuned one bit for bit (every y-line is solved the same way whatever the
strip width or workgroup order), eagerly and through the captured graph."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def m():
    import mhd2dcfdfftw_open_amd as m
    m.lib()
    return m


@pytest.mark.parametrize("nx,ny,prec,batch,npe", [
    (256, 128, "fp64", 1, 1),
    (512, 256, "mixed", 3, 1),
    (128, 64, "fp32", 2, 1),
    (128, 64, "fp64", 1, 2),
])
def test_autotune_keeps_results(m, nx, ny, prec, batch, npe):
    print(f"[autotune] {nx}x{ny} {prec} batch={batch} npe={npe}", flush=True)
    p = {"fp64": m.FP64, "mixed": m.MIXED, "fp32": m.FP32}[prec]
    u = [m.turbulence(nx, ny, seed=10 + b, kmax=8) for b in range(batch)]
    ref = m.Context(nx, ny, hall=1, di=0.03, precision=p, batch=batch, npe=npe)
    tun = m.Context(nx, ny, hall=1, di=0.03, precision=p, batch=batch, npe=npe)
    for c in (ref, tun):
        for b in range(batch):
            c.set_state(u[b], inst=b)
    before = [tun.state(inst=b) for b in range(batch)]
    ch = tun.autotune(reps=2)
    for k in ("deriv", "filter"):
        w, r = ch[k]
        assert w in (8, 16) and r in (0, 1), ch
    for b in range(batch):
        np.testing.assert_array_equal(tun.state(inst=b), before[b])
    ref.run(4)
    tun.run(4)
    for b in range(batch):
        np.testing.assert_array_equal(tun.state(inst=b), ref.state(inst=b))
    ref.run(3, graph=True)
    tun.run(3, graph=True)
    for b in range(batch):
        np.testing.assert_array_equal(tun.state(inst=b), ref.state(inst=b))


def test_autotune_4096_has_only_narrow_strips(m):
    # 64-row chunks do not fit 16-column strips
    c = m.Context(4096, 64, hall=1, di=0.02)
    c.set_state(m.turbulence(4096, 64, seed=3, kmax=8))
    ch = c.autotune(reps=1)
    assert ch["deriv"][0] == 8 and ch["filter"][0] == 8


def test_ensemble_autotune(m):
    runs = [dict(nx=256, ny=128, di=0.01), dict(nx=128, ny=128, di=0.02),
            dict(nx=256, ny=128, di=0.05)]
    a = m.Ensemble(runs)
    b = m.Ensemble(runs)
    for e in (a, b):
        for k, r in enumerate(runs):
            e.set_state(k, m.turbulence(r["nx"], r["ny"], seed=k, kmax=6))
    ch = b.autotune(reps=1)
    assert len(ch) == 2
    a.run(3, graph=True)
    b.run(3, graph=True)
    for k in range(len(runs)):
        np.testing.assert_array_equal(a.state(k), b.state(k))


def test_autotune_rejects_gemm_form(m):
    c = m.Context(128, 64, precision=m.MIXED, xgemm=6)
    with pytest.raises(Exception):
        c.autotune()
