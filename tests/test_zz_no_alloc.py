"""mhd2d_run_noalloc (the entry point that allocates nothing) against the
context API it mirrors (mhd2d_create + set_state + run + get_state), through
the C ABI.

Both run the same kernels, so the states must agree bit for bit; the gate
below is the FP64 parity bar (per-field relative L2 <= 1e-12) plus an exact
comparison.  Two calls in a row, the second on a smaller grid, so that the
second call works in the array the first one left dirty; then settings with
the source-term buffers (viscosity, resistivity, diffusive cleaning) and GLM
cleaning; and a grid above the limit is refused with a message.
Device memory for the caller's state comes from hipMalloc through ctypes.
"""
import ctypes

import numpy as np
import pytest

from util import assert_fields_close

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def m():
    import mhd2dcfdfftw_open_amd as m
    L = m.lib()
    f = L.mhd2d_run_noalloc
    f.restype = ctypes.c_int
    f.argtypes = [ctypes.POINTER(m.Params), ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p,
                  ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]
    return m


@pytest.fixture(scope="module")
def hip():
    H = ctypes.CDLL("libamdhip64.so")
    H.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    H.hipFree.argtypes = [ctypes.c_void_p]
    H.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    H.hipDeviceSynchronize.argtypes = []
    return H


H2D, D2H = 1, 2


def _params(m, nx, ny, **kw):
    p = m.Params()
    p.nx, p.ny = nx, ny
    p.Lx = p.Ly = 2 * np.pi
    p.gamma, p.cfl = 5.0 / 3.0, 0.3
    p.hall = kw.get("hall", 0)
    p.di = kw.get("di", 0.0)
    p.resistivity = kw.get("resistivity", 0.0)
    p.viscosity = kw.get("viscosity", 0.0)
    p.divbclean = kw.get("divbclean", 0)
    p.precision = m.FP64
    p.xblocks = kw.get("xblocks", 0)
    return p


def _noalloc(m, hip, u0, steps, **kw):
    """state after `steps` through mhd2d_run_noalloc, dt, time"""
    nv, nx, ny = u0.shape
    host = np.ascontiguousarray(u0, dtype=np.float64)
    nb = host.nbytes
    d = ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(d), nb) == 0
    try:
        assert hip.hipMemcpy(d, host.ctypes.data, nb, H2D) == 0
        dt, t = ctypes.c_double(), ctypes.c_double()
        p = _params(m, nx, ny, **kw)
        rc = m.lib().mhd2d_run_noalloc(ctypes.byref(p), d, steps, None, ctypes.byref(dt),
                                       ctypes.byref(t))
        assert rc == 0, m.lib().mhd2d_last_error()
        out = np.empty_like(host)
        assert hip.hipMemcpy(out.ctypes.data, d, nb, D2H) == 0
        return out, dt.value, t.value
    finally:
        hip.hipFree(d)


def _context(m, u0, steps, **kw):
    nv, nx, ny = u0.shape
    kw = dict(kw)
    if "xblocks" in kw:
        kw["npe"] = kw.pop("xblocks")
    ctx = m.Context(nx, ny, **kw)
    try:
        ctx.set_state(u0)
        ctx.run(steps)
        ctx.sync()
        st = ctx.state()
        dt, t, _ = ctx.scalars()
        return st, dt, t
    finally:
        ctx.close()


def _check(m, hip, u0, steps, what, **kw):
    print(f"[noalloc] {what} {u0.shape} steps={steps} {kw}", flush=True)
    got, dt, t = _noalloc(m, hip, u0, steps, **kw)
    want, dt_w, t_w = _context(m, u0, steps, **kw)
    assert got.shape == want.shape
    assert_fields_close(got, want, what=what)
    np.testing.assert_array_equal(got, want)
    assert dt == dt_w and t == t_w


def test_noalloc_equals_context_two_calls(m, hip):
    # first call on 256 x 128, second on a smaller 64 x 32 grid (the array
    # still holds the first call's data)
    _check(m, hip, m.orszag_tang(256, 128), 12, "first call", hall=1, di=0.05)
    _check(m, hip, m.turbulence(64, 32, seed=5, kmax=6), 9, "second call, smaller")


@pytest.mark.parametrize("kw", [dict(viscosity=2e-3, resistivity=1e-3),
                                dict(divbclean=1), dict(divbclean=2, hall=1, di=0.05),
                                dict(xblocks=3)])
def test_noalloc_settings(m, hip, kw):
    u0 = m.turbulence(128, 64, seed=9, kmax=6)
    if kw.get("divbclean") == 2:
        u0 = np.concatenate([u0, np.zeros((1, 128, 64))])   # psi starts at 0
    _check(m, hip, u0, 6, "settings", **kw)


def test_noalloc_refuses_large_grids(m):
    p = _params(m, 512, 256)     # 131072 cells > 65536
    rc = m.lib().mhd2d_run_noalloc(ctypes.byref(p), ctypes.c_void_p(16), 1, None, None, None)
    assert rc == 1                # MHD2D_EINVAL
    assert b"MHD2D_NOALLOC_MAX_CELLS" in m.lib().mhd2d_last_error()
