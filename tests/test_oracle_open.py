"""Pins of the oracle's open x boundary (boundaryTypeX == 1, npe >= 2):
one-sided closures on the first and last rank (mathlib.c:87-159, :219-294,
:355-440; initialize.c:655-861) and the characteristic boundary RHS
(calcRHS.c:163-495).

* every closure row is exact on cubics, so the first- and second-derivative
  solves are exact and the filter preserves cubics on the whole line; the
  one-sided end rows are not exact on quintics (the test can fail);
* the right end is the mirror image of the left end;
* a middle block is untouched by the open ends (equal to the periodic split);
* the characteristic terms equal -R diag(lambda, outgoing only) R^-1 dW/dx
  of the 1D primitive MHD Jacobian, built and diagonalised by numpy, for
  subsonic and supersonic states at both ends; all waves outgoing reproduces
  the full x-derivative terms, all incoming gives zero;
* a uniform state is a fixed point of the open-boundary run.
"""
import numpy as np
import pytest

import oracle_lib as orc

GAMMA = 5.0 / 3.0


def _poly(nx, ny, deg, nv=2, dx=0.1):
    x = np.arange(nx) * dx
    f = np.zeros((nv, nx, ny))
    df = np.zeros_like(f)
    d2f = np.zeros_like(f)
    rng = np.random.default_rng(deg + nx)
    for v in range(nv):
        for j in range(ny):
            c = rng.standard_normal(deg + 1)
            f[v, :, j] = np.polyval(c, x)
            df[v, :, j] = np.polyval(np.polyder(c), x)
            d2f[v, :, j] = np.polyval(np.polyder(c, 2), x) if deg >= 2 else 0.0
    return f, df, d2f


@pytest.mark.parametrize("nx,npe", [(48, 2), (60, 3), (64, 4)])
def test_open_closures_exact_on_cubics(nx, npe):
    dx = 0.1
    f, df, d2f = _poly(nx, 3, 3, dx=dx)
    scale = np.abs(f).max()
    np.testing.assert_allclose(orc.derivX_open(f, dx, npe), df, rtol=0,
                               atol=1e-11 * scale / dx)
    np.testing.assert_allclose(orc.derivXX_open(f, dx, npe), d2f, rtol=0,
                               atol=1e-9 * scale / dx ** 2)
    np.testing.assert_allclose(orc.filterX_open(f, npe), f, rtol=0, atol=1e-12 * scale)


def test_open_end_rows_not_exact_on_quintics():
    # the alpha = 3 end row is the fourth-order one-sided closure
    nx, npe, dx = 48, 2, 0.1
    f, df, _ = _poly(nx, 2, 5, dx=dx)
    err = np.abs(orc.derivX_open(f, dx, npe) - df).max(axis=(0, 2))
    assert err[0] > 1e-9 * np.abs(df).max() and err[-1] > 1e-9 * np.abs(df).max()


@pytest.mark.parametrize("npe", [2, 4])
def test_right_end_mirrors_left_end(npe):
    nx, dx = 64, 0.2
    rng = np.random.default_rng(npe)
    f = rng.standard_normal((2, nx, 5))
    g = f[:, ::-1, :].copy()
    np.testing.assert_allclose(orc.derivX_open(g, dx, npe), -orc.derivX_open(f, dx, npe)[:, ::-1],
                               rtol=0, atol=1e-12)
    np.testing.assert_allclose(orc.derivXX_open(g, dx, npe), orc.derivXX_open(f, dx, npe)[:, ::-1],
                               rtol=0, atol=1e-11)
    np.testing.assert_allclose(orc.filterX_open(g, npe), orc.filterX_open(f, npe)[:, ::-1],
                               rtol=0, atol=1e-13)


def test_middle_block_is_the_periodic_split():
    nx, npe, dx = 60, 3, 0.3
    f = np.random.default_rng(9).standard_normal((2, nx, 4))
    mid = slice(20, 40)
    np.testing.assert_array_equal(orc.derivX_open(f, dx, npe)[:, mid],
                                  orc.derivX(f, dx, npe=npe)[:, mid])
    np.testing.assert_array_equal(orc.filterX_open(f, npe)[:, mid], orc.filterX(f, npe=npe)[:, mid])
    assert np.abs(orc.derivX_open(f, dx, npe)[:, 0] - orc.derivX(f, dx, npe=npe)[:, 0]).max() > 1e-3


def _jacobian(w):
    """1D primitive MHD along x, W = (rho, ux, uy, uz, By, Bz, p), Bx fixed:
    dW/dt + A dW/dx = 0"""
    rho, ux, uy, uz, bx, by, bz, p = w
    A = np.zeros((7, 7))
    A[0, 0], A[0, 1] = ux, rho
    A[1, 1], A[1, 6], A[1, 4], A[1, 5] = ux, 1 / rho, by / rho, bz / rho
    A[2, 2], A[2, 4] = ux, -bx / rho
    A[3, 3], A[3, 5] = ux, -bx / rho
    A[4, 4], A[4, 1], A[4, 2] = ux, by, -bx
    A[5, 5], A[5, 1], A[5, 3] = ux, bz, -bx
    A[6, 6], A[6, 1] = ux, GAMMA * p
    return A


def _outgoing_terms(w, dw, boundary):
    A = _jacobian(w)
    lam, R = np.linalg.eig(A)
    lam, R = lam.real, R.real
    L = np.linalg.inv(R)
    keep = lam < 0 if boundary == 0 else lam > 0
    dW = dw[[0, 1, 2, 3, 5, 6, 7]]
    return -R @ np.diag(lam * keep) @ L @ dW


@pytest.mark.parametrize("boundary", [0, 1])
@pytest.mark.parametrize("regime", ["subsonic_out", "subsonic_in", "all_out", "all_in", "slow"])
def test_characteristics_are_the_outgoing_projection(boundary, regime):
    rng = np.random.default_rng(hash((boundary, regime)) % 2 ** 32)
    for trial in range(6):
        w = np.array([1 + rng.random(), 0.0, rng.normal(), rng.normal(), 0.3 + rng.random(),
                      rng.normal(), rng.normal(), 1 + rng.random()])
        cs2 = GAMMA * w[7] / w[0]
        ca2 = (w[4] ** 2 + w[5] ** 2 + w[6] ** 2) / w[0]
        cf = np.sqrt(cs2 + ca2)
        out_sign = -1 if boundary == 0 else 1    # direction leaving the domain
        ux = {"subsonic_out": 0.3, "subsonic_in": -0.3, "all_out": 1.5, "all_in": -1.5,
              "slow": 0.05}[regime] * cf * out_sign
        w[1] = ux
        dw = rng.normal(size=8)
        uu = np.zeros((8, 3, 3))
        dudx = np.zeros((8, 3, 3))
        uu[:, 1, 1], dudx[:, 1, 1] = w, dw
        got = orc.characteristics_x(uu, dudx, 1, 1, boundary)
        want = _outgoing_terms(w, dw, boundary)
        np.testing.assert_allclose(got, want, rtol=1e-10, atol=1e-11 * np.abs(want).max() + 1e-14)
        if regime == "all_in":
            assert not got.any()


def test_open_run_keeps_a_uniform_state():
    nx, ny = 64, 32
    u0 = np.zeros((8, nx, ny))
    u0[0], u0[1], u0[4], u0[5], u0[7] = 1.0, 0.3, 0.8, 0.2, 0.6
    S = orc.Sim(nx, ny, npe=2, open_x=True, hall=1, di=0.05)
    S.set_state(u0)
    S.run(5)
    np.testing.assert_allclose(S.state(), u0, rtol=0, atol=1e-13)
    # dx of the open grid (initialize.c:125-133) enters the CFL limit
    per = orc.Sim(nx, ny, npe=2)
    per.set_state(u0)
    assert S.dt == pytest.approx(per.dt * (nx / (nx - 1)), rel=1e-12)
