"""Pins for the oracle's field operators, FFT conventions and physics.

* r2c / c2r vs numpy.fft (pocketfft, an independent library routine);
* spectral y-derivatives on pure modes (closed forms, Nyquist handling);
* y-filter transfer function from its published form at afy = 0.45;
* x-operators on 2-D fields vs the line operators (layout / halo wrap);
* E = -u x B vs numpy.cross; RHS vs a vector-form numpy restatement of the
  ideal-MHD equations; RK3 vs the Taylor polynomial it must reproduce;
  calcDt vs closed-form wave speeds and the +-2 % hysteresis.
"""
import numpy as np
import pytest

import oracle_lib as orc


@pytest.mark.parametrize("n", [8, 16, 64, 256, 2048])
def test_rfft_matches_numpy(n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n)
    X = orc.rfft(x)
    np.testing.assert_allclose(X, np.fft.rfft(x), rtol=0, atol=1e-12 * np.sqrt(n))
    # c2r is unnormalised and ignores Im of bins 0 and n/2 (FFTW convention)
    Y = X.copy()
    Y[0] += 3j
    Y[-1] += 5j
    np.testing.assert_allclose(orc.irfft(Y, n), n * x, rtol=0, atol=1e-11 * n)


def _grid(nx, ny, L=2 * np.pi):
    x = np.arange(nx) * (L / nx)
    y = np.arange(ny) * (L / ny)
    return np.meshgrid(x, y, indexing="ij")


@pytest.mark.parametrize("m", [0, 1, 7, 15])
def test_derivY_pure_modes(m):
    nx, ny, Ly = 4, 32, 2 * np.pi
    X, Y = _grid(nx, ny)
    f = np.stack([np.sin(m * Y + 0.2 * X), np.cos(m * Y)])
    d = orc.derivY(f, Ly)
    np.testing.assert_allclose(d[0], m * np.cos(m * Y + 0.2 * X), atol=1e-12 * (m + 1))
    np.testing.assert_allclose(d[1], -m * np.sin(m * Y), atol=1e-12 * (m + 1))
    d2 = orc.derivYY(f, Ly)
    np.testing.assert_allclose(d2[0], -m * m * f[0], atol=1e-11 * (m + 1) ** 2)


def test_derivY_nyquist_zeroed_derivYY_kept():
    nx, ny, Ly = 2, 16, 3.0
    j = np.arange(ny)
    f = np.broadcast_to((-1.0) ** j, (1, nx, ny)).copy()
    np.testing.assert_allclose(orc.derivY(f, Ly), 0.0, atol=1e-13)
    k = 2 * np.pi / Ly * (ny // 2)
    np.testing.assert_allclose(orc.derivYY(f, Ly), -k * k * f, rtol=1e-13)


def test_derivY_scales_with_Ly():
    nx, ny = 2, 64
    X, Y = _grid(nx, ny)
    f = np.sin(3 * Y)[None]
    for Ly in (1.0, 2 * np.pi, 7.5):
        d = orc.derivY(f, Ly)
        np.testing.assert_allclose(d[0], 3 * (2 * np.pi / Ly) * np.cos(3 * Y), atol=1e-11)


def test_derivY_matches_numpy_spectral():
    rng = np.random.default_rng(3)
    nx, ny, Ly = 5, 128, 4.0
    f = rng.standard_normal((2, nx, ny))
    k = np.fft.rfftfreq(ny, d=Ly / ny) * 2 * np.pi
    k[-1] = 0.0   # Nyquist dropped by the first derivative
    ref = np.fft.irfft(1j * k * np.fft.rfft(f, axis=-1), n=ny, axis=-1)
    np.testing.assert_allclose(orc.derivY(f, Ly), ref, atol=1e-11)


def test_filtY_published_form():
    ny = 64
    t = orc.filtY_table(ny, 0.45)
    w = 2 * np.pi * np.arange(ny // 2 + 1) / ny
    # Lele (1992) eq. C.2.4 with beta = d = 0 at alpha = 0.45
    a, b, c = 0.9625, 0.95, -0.0125
    np.testing.assert_allclose(t, (a + b * np.cos(w) + c * np.cos(2 * w)) / (1 + 0.9 * np.cos(w)),
                               rtol=1e-14, atol=1e-15)
    assert abs(t[0] - 1.0) < 1e-15 and abs(t[-1]) < 1e-14


def test_filterY_on_modes():
    nx, ny = 3, 64
    X, Y = _grid(nx, ny)
    t = orc.filtY_table(ny)
    for m in (0, 5, 20, 31):
        f = np.cos(m * Y + X)[None]
        np.testing.assert_allclose(orc.filterY(f)[0], t[m] * f[0], atol=1e-13)


@pytest.mark.parametrize("op", ["d1", "d2", "f"])
def test_x_field_ops_match_line_ops(op):
    rng = np.random.default_rng(11)
    nv, nx, ny = 2, 24, 5
    dx = 0.37
    arr = rng.standard_normal((nv, nx, ny))
    if op == "d1":
        out = orc.derivX(arr, dx)
    elif op == "d2":
        out = orc.derivXX(arr, dx)
    else:
        out = orc.filterX(arr)
    for v in range(nv):
        for j in range(ny):
            f = arr[v, :, j]
            line = orc.line_op(op, f, f[-2:], f[:2], DX=dx)
            np.testing.assert_array_equal(out[v, :, j], line)


def test_efield_is_minus_u_cross_b():
    rng = np.random.default_rng(5)
    uu = rng.standard_normal((8, 6, 7))
    E = orc.efield(uu)
    ref = -np.cross(uu[1:4], uu[4:7], axis=0)
    np.testing.assert_allclose(E, ref, rtol=1e-15, atol=1e-15)


def _rhs_vector_form(uu, dudx, dudy, dEdx, dEdy, gamma):
    """Ideal MHD in primitive variables written with vector operations."""
    rho, u, B, p = uu[0], uu[1:4], uu[4:7], uu[7]
    grad = lambda q: np.stack([dudx[q], dudy[q], np.zeros_like(dudx[q])])
    divu = dudx[1] + dudy[2]
    # J = curl B with d/dz = 0
    J = np.stack([dudy[6], -dudx[6], dudx[5] - dudy[4]])
    out = np.zeros_like(uu)
    out[0] = -rho * divu - np.einsum("c...,c...->...", u, grad(0))
    for c in range(3):
        out[1 + c] = -np.einsum("d...,d...->...", u, grad(1 + c))
    out[1:4] += -np.stack([dudx[7], dudy[7], 0 * rho]) / rho + np.cross(J, B, axis=0) / rho
    # dB/dt = -curl E
    curlE = np.stack([dEdy[2], -dEdx[2], dEdx[1] - dEdy[0]])
    out[4:7] = -curlE
    out[7] = -np.einsum("c...,c...->...", u, grad(7)) - gamma * p * divu
    return out


def test_rhs_matches_vector_form():
    rng = np.random.default_rng(8)
    shape = (8, 5, 6)
    uu = rng.standard_normal(shape)
    uu[0] = 1.0 + rng.uniform(0, 1, shape[1:])
    dudx, dudy = rng.standard_normal(shape), rng.standard_normal(shape)
    dEdx, dEdy = rng.standard_normal((3,) + shape[1:]), rng.standard_normal((3,) + shape[1:])
    g = 1.4
    got = orc.rhs(uu, dudx, dudy, dEdx, dEdy, gamma=g)
    np.testing.assert_allclose(got, _rhs_vector_form(uu, dudx, dudy, dEdx, dEdy, g),
                               rtol=1e-13, atol=1e-13)


def test_rhs_source_terms():
    rng = np.random.default_rng(9)
    shape = (8, 4, 4)
    uu = rng.standard_normal(shape)
    uu[0] = 2.0
    z8, z3, z1 = np.zeros(shape), np.zeros((3, 4, 4)), np.zeros((4, 4))
    a, b = rng.standard_normal((3, 4, 4)), rng.standard_normal((3, 4, 4))
    got = orc.rhs(uu, z8, z8, z3, z3, viscosity=0.1, d2udx2=a, d2udy2=b)
    np.testing.assert_allclose(got[1:4], 0.1 / 2.0 * (a + b), rtol=1e-14)
    got = orc.rhs(uu, z8, z8, z3, z3, resistivity=0.2, d2Bdx2=a, d2Bdy2=b)
    np.testing.assert_allclose(got[4:7], 0.2 * (a + b), rtol=1e-14)
    c, d = rng.standard_normal((4, 4)), rng.standard_normal((4, 4))
    got = orc.rhs(uu, z8, z8, z3, z3, divbclean=1, constDiffDivB=0.3, ddivBdx=c, ddivBdy=d)
    np.testing.assert_allclose(got[4], 0.3 * c, rtol=1e-14)
    np.testing.assert_allclose(got[5], 0.3 * d, rtol=1e-14)


def test_rk3_reproduces_taylor_polynomial():
    # Low-storage RK3 of Spalart-Moser-Rogers on u' = lam u gives
    # u1 = (1 + z + z^2/2 + z^3/6) u0 exactly.
    lam = np.array([-0.3, 0.7, -2.0 + 0.5j]).real
    dt = 0.17
    cc = [8 / 15 * dt, 5 / 12 * dt, 3 / 4 * dt]
    dd = [0.0, -17 / 60 * dt, -5 / 12 * dt]
    u = np.ones_like(lam)
    rk = np.zeros_like(lam)
    for s in range(3):
        u, rk = orc.rk_execute(u, lam * u, rk, cc[s], dd[s])
    z = lam * dt
    np.testing.assert_allclose(u, 1 + z + z ** 2 / 2 + z ** 3 / 6, rtol=1e-15)


def _uniform(nx, ny, rho, ux, uy, bx, by, bz, p):
    uu = np.zeros((8, nx, ny))
    uu[0], uu[1], uu[2], uu[4], uu[5], uu[6], uu[7] = rho, ux, uy, bx, by, bz, p
    return uu


def test_calc_dt_closed_form():
    nx, ny, dx, dy = 4, 8, 0.1, 0.05
    g, cfl = 5 / 3, 0.3
    rho, p, b0 = 2.0, 0.7, 1.3
    uu = _uniform(nx, ny, rho, 0.0, 0.0, b0, 0.0, 0.0, p)
    cs, ca = np.sqrt(g * p / rho), b0 / np.sqrt(rho)
    cx = max(cs, ca)                 # along B: fast = max(cs, ca)
    cy = np.sqrt(cs ** 2 + ca ** 2)  # across B: fast = sqrt(cs^2 + ca^2)
    want = cfl * min(dx / cx, dy / cy)
    assert abs(orc.calc_dt(uu, dx, dy, g, cfl) - want) < 1e-15 * want
    # hysteresis keeps dt0 within [0.98, 1.02] x dtmin
    assert orc.calc_dt(uu, dx, dy, g, cfl, dt0=1.019 * want) == 1.019 * want
    assert orc.calc_dt(uu, dx, dy, g, cfl, dt0=0.981 * want) == 0.981 * want
    assert abs(orc.calc_dt(uu, dx, dy, g, cfl, dt0=1.021 * want) - want) < 1e-15 * want
    # flow speed adds along x
    uu2 = _uniform(nx, ny, rho, 0.5, 0.0, b0, 0.0, 0.0, p)
    want2 = cfl * min(dx / (0.5 + cx), dy / cy)
    assert abs(orc.calc_dt(uu2, dx, dy, g, cfl) - want2) < 1e-15 * want2
    # Hall whistler limit: di / min(dx, dy) / rho * max|B|
    ch = 5.0 / min(dx, dy) / rho * b0
    want3 = cfl * min(dx / max(cx, ch), dy / max(cy, ch))
    got3 = orc.calc_dt(uu, dx, dy, g, cfl, hall=1, di=5.0)
    assert abs(got3 - want3) < 1e-15 * want3
    # diffusive limits
    got4 = orc.calc_dt(uu, dx, dy, g, cfl, resistivity=10.0)
    assert abs(got4 - cfl * dy * dy / 10.0) < 1e-15


def test_energy_orszag_tang_closed_form():
    nx = ny = 32
    X, Y = _grid(nx, ny)
    g = 5 / 3
    rho, p = 25 / (36 * np.pi), 5 / (12 * np.pi)
    s = 1 / np.sqrt(4 * np.pi)
    uu = _uniform(nx, ny, rho, 0, 0, 0, 0, 0, p)
    uu[1], uu[2] = -np.sin(Y), np.sin(X)
    uu[4], uu[5] = -np.sin(Y) * s, np.sin(2 * X) * s
    d = 2 * np.pi / nx
    want = 4 * np.pi ** 2 * (rho / 2 + p / (g - 1) + 1 / (8 * np.pi))
    assert abs(orc.energy(uu, d, d, g) - want) < 1e-13 * want
