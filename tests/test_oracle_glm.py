"""Pins of the oracle's hyperbolic (GLM) div B cleaning, divBcleaning == 2
(Dedner et al. 2002 eq. 24 as in calcRHS.c:144-158, Ch / Cp from calcDt,
timeAdvance.c:107-112 and divBcleaning.c:4-10).

* Small-amplitude y-modes on a uniform background: to first order the fields
  decouple into a few Fourier amplitudes whose RK3 + y-filter evolution is a
  small linear map written here from the equations of the cleaning system,
  not from the oracle's code.  Two backgrounds cover every cleaning term:
  B0 = (Bx0, 0, Bz0) drives ux, uz through -divB B; B0 = (0, By0, 0) drives uy
  and couples in the pressure wave.  A flipped sign, a dropped Ch^2 or a wrong
  Cp fails them at O(1).
* Ch is the fastest signal speed of the LAST cell of the rank's loop (not a
  maximum), Cp = sqrt(0.18 Ch); per rank with npe > 1.
* With div B = 0 and psi = 0 the cleaning leaves the run unchanged.
"""
import numpy as np
import pytest

import oracle_lib as orc

GAMMA = 5.0 / 3.0


def _filt(m, ny, afy=0.45):
    # published form of the y filter response (initialize.c:166-176)
    w = 2 * np.pi * m / ny
    a, b, c = (5 + 6 * afy) / 8, (1 + 2 * afy) / 2, -(1 - 2 * afy) / 8
    return (a + b * np.cos(w) + c * np.cos(2 * w)) / (1 + 2 * afy * np.cos(w))


def _fast_perp(rho, p, b2):
    return np.sqrt(GAMMA * p / rho + b2 / rho)


def _rk3_filtered(A, x0, dt, f, nsteps):
    """Low-storage RK3 (8/15, 5/12, 3/4; 0, -17/60, -5/12) of x' = A x with the
    linear y filter f applied after every stage."""
    c = [8 / 15, 5 / 12, 3 / 4]
    d = [0.0, -17 / 60, -5 / 12]
    x = x0.copy()
    for _ in range(nsteps):
        rp = np.zeros_like(x)
        for s in range(3):
            r = A @ x
            x = f * (x + c[s] * dt * r + d[s] * dt * rp)
            rp = r
    return x


def _amp(F, y, k, kind):
    basis = np.cos(k * y) if kind == "cos" else np.sin(k * y)
    a = 2.0 / len(y) * (F @ basis)          # one amplitude per x row
    assert np.ptp(a) <= 1e-6 * np.abs(a).max() + 1e-14   # x-independent up to rounding
    return a.mean()


@pytest.mark.parametrize("config", ["bxz", "by"])
def test_glm_linear_modes_match_closed_form(config):
    nx, ny, m, eps = 16, 32, 2, 1e-7
    Ly = 2 * np.pi
    y = np.arange(ny) * Ly / ny
    k = 2 * np.pi * m / Ly
    rho0, p0 = 1.0, 0.6
    u0 = np.zeros((8, nx, ny))
    u0[0], u0[7] = rho0, p0
    if config == "bxz":
        Bx0, By0, Bz0 = 0.8, 0.0, 0.5
    else:
        Bx0, By0, Bz0 = 0.0, 0.9, 0.0
    u0[4], u0[5], u0[6] = Bx0, By0, Bz0
    u0[5] += eps * np.cos(k * y)[None, :]            # div B = -eps k sin(ky)
    S = orc.Sim(nx, ny, Lx=1.0, Ly=Ly, gamma=GAMMA, divbclean=2)
    S.set_state(u0)
    Ch = _fast_perp(rho0, p0, Bx0 ** 2 + By0 ** 2 + Bz0 ** 2)
    ch, cp = S.glm(0)
    assert ch == pytest.approx(Ch, rel=1e-6)          # up to O(eps^2) in the last cell
    assert cp == pytest.approx(np.sqrt(0.18 * ch), rel=1e-15)
    Cp2 = 0.18 * Ch
    dt = S.dt
    nsteps = 25
    S.run_fixed_dt(np.full(nsteps, dt))
    U = S.state()
    f = _filt(m, ny)
    if config == "bxz":
        # x = (By cos, psi sin, ux sin, uz sin)
        A = np.array([[0, -k, 0, 0],
                      [Ch ** 2 * k, -Ch ** 2 / Cp2, 0, 0],
                      [k * Bx0, 0, 0, 0],
                      [k * Bz0, 0, 0, 0]])
        x = _rk3_filtered(A, np.array([eps, 0, 0, 0]), dt, f, nsteps)
        got = [_amp(U[5] - By0, y, k, "cos"), _amp(U[8], y, k, "sin"),
               _amp(U[1], y, k, "sin"), _amp(U[3], y, k, "sin")]
        rest = [U[0] - rho0, U[2], U[7] - p0, U[4] - Bx0, U[6] - Bz0]
    else:
        # x = (By cos, psi sin, uy sin, rho cos, p cos)
        A = np.array([[0, -k, 0, 0, 0],
                      [Ch ** 2 * k, -Ch ** 2 / Cp2, 0, 0, 0],
                      [By0 * k, 0, 0, 0, k / rho0],
                      [0, 0, -rho0 * k, 0, 0],
                      [0, 0, -GAMMA * p0 * k, 0, 0]])
        x = _rk3_filtered(A, np.array([eps, 0, 0, 0, 0]), dt, f, nsteps)
        got = [_amp(U[5] - By0, y, k, "cos"), _amp(U[8], y, k, "sin"),
               _amp(U[2], y, k, "sin"), _amp(U[0] - rho0, y, k, "cos"),
               _amp(U[7] - p0, y, k, "cos")]
        rest = [U[1], U[3], U[4], U[6]]
    scale = np.abs(x).max()
    for g, want in zip(got, x):
        assert abs(g - want) <= 1e-5 * scale, (config, got, x)
    # psi and the driven components really moved (the test has teeth)
    assert abs(x[1]) > 0.05 * scale and abs(x[2]) > 0.05 * scale
    for r in rest:
        assert np.abs(r).max() <= 1e-5 * scale


def test_glm_ch_is_the_last_cells_speed_per_rank():
    nx, ny = 32, 16
    u0 = np.zeros((8, nx, ny))
    u0[0], u0[7], u0[4] = 1.0, 1.0, 0.3
    u0[7, 5, 5] = 50.0                     # the global fastest cell, not used
    u0[:, nx - 1, ny - 1] = [2.0, 0, 0, 0, 0.5, 0.0, 0.2, 3.0]
    u0[:, nx // 2 - 1, ny - 1] = [1.5, 0, 0, 0, 0.0, 0.4, 0.0, 0.7]
    for npe in (1, 2):
        S = orc.Sim(nx, ny, divbclean=2, npe=npe)
        S.set_state(u0)
        last = _fast_perp(2.0, 3.0, 0.25 + 0.04)
        assert S.glm(nx - 1)[0] == pytest.approx(last, rel=1e-14)
        want0 = last if npe == 1 else _fast_perp(1.5, 0.7, 0.16)
        assert S.glm(0)[0] == pytest.approx(want0, rel=1e-14)
        assert S.glm(0)[1] == pytest.approx(np.sqrt(0.18 * want0), rel=1e-15)


def test_glm_leaves_a_divergence_free_run_unchanged():
    # a nonlinear flow that depends on y only with uniform By: div B = dBy/dy
    # stays 0 (dBy/dt = dEz/dx = 0), so psi stays ~0 and the cleaning terms
    # vanish.  (In 2D the x closure rows do not commute with the x filter,
    # which creates div B at the block edges: exactly what the cleaning is for.)
    nx = ny = 32
    y = np.arange(ny) * 2 * np.pi / ny
    Y = np.broadcast_to(y, (nx, ny))
    u0 = np.zeros((8, nx, ny))
    u0[0] = 1.0 + 0.2 * np.cos(Y)
    u0[1], u0[2], u0[3] = 0.3 * np.sin(Y), 0.2 * np.cos(2 * Y), 0.1 * np.sin(3 * Y)
    u0[4], u0[5], u0[6] = 0.5 * np.cos(Y), 0.7, 0.4 * np.sin(2 * Y)
    u0[7] = 0.6 + 0.1 * np.sin(Y)
    A = orc.Sim(nx, ny, divbclean=0)
    B = orc.Sim(nx, ny, divbclean=2)
    A.set_state(u0)
    B.set_state(u0)
    A.run(5)
    B.run(5)
    ua, ub = A.state(), B.state()
    assert ub.shape == (9, nx, ny)
    assert np.abs(ub[8]).max() < 1e-10
    assert np.abs(ua - ub[:8]).max() < 1e-10 * np.abs(ua).max()


def test_rhs_glm_reduces_to_plain_rhs_plus_cleaning_terms():
    # rows unaffected by the cleaning match the plain right-hand side bitwise
    rng = np.random.default_rng(3)
    nx, ny = 8, 8
    uu = rng.standard_normal((9, nx, ny))
    uu[0] = 1.0 + 0.1 * np.abs(uu[0])
    dudx, dudy = rng.standard_normal((9, nx, ny)), rng.standard_normal((9, nx, ny))
    dEdx, dEdy = rng.standard_normal((3, nx, ny)), rng.standard_normal((3, nx, ny))
    g = orc.rhs_glm(uu, dudx, dudy, dEdx, dEdy, ch=1.3, cp=0.5)
    p = orc.rhs(uu[:8], dudx[:8], dudy[:8], dEdx, dEdy)
    for f in (0, 6, 7):                      # rho, Bz, p
        np.testing.assert_array_equal(g[f], p[f])
    divb = dudx[4] + dudy[5]
    for f, b in ((1, 4), (2, 5), (3, 6)):
        np.testing.assert_allclose(g[f] - p[f], -divb * uu[b], rtol=1e-12, atol=1e-13)
    np.testing.assert_allclose(g[4] - p[4], -dudx[8], rtol=1e-12, atol=1e-13)
    np.testing.assert_allclose(g[5] - p[5], -dudy[8], rtol=1e-12, atol=1e-13)
    np.testing.assert_allclose(g[8], -1.69 * (divb + uu[8] / 0.25), rtol=1e-13)
