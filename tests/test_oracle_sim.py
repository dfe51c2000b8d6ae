"""End-to-end pins for the oracle's stage loop (mhd.c:154-206 restated).

The reference ships no tests or fixtures and cannot be built here (no FFTW3),
so the composed trajectory is pinned by known answers and invariants:
* a uniform state is a fixed point;
* a small-amplitude Alfven wave travels at V_A = B0/sqrt(rho) in +x;
* the induction terms keep the discrete div B of the update at round-off
  (Dx and Dy act on different axes and commute);
* runs are deterministic; a fixed-dt replay reproduces the adaptive run.
"""
import numpy as np
import pytest

import oracle_lib as orc


def _grid(nx, ny, L=2 * np.pi):
    x = np.arange(nx) * (L / nx)
    y = np.arange(ny) * (L / ny)
    return np.meshgrid(x, y, indexing="ij")


def orszag_tang(nx, ny):
    X, Y = _grid(nx, ny)
    uu = np.zeros((8, nx, ny))
    s = 1 / np.sqrt(4 * np.pi)
    uu[0] = 25 / (36 * np.pi)
    uu[7] = 5 / (12 * np.pi)
    uu[1], uu[2] = -np.sin(Y), np.sin(X)
    uu[4], uu[5] = -np.sin(Y) * s, np.sin(2 * X) * s
    return uu


def test_uniform_state_is_fixed_point():
    s = orc.Sim(16, 32)
    uu = np.zeros((8, 16, 32))
    uu[0], uu[1], uu[2], uu[3] = 1.3, 0.2, -0.1, 0.05
    uu[4], uu[5], uu[6], uu[7] = 1.0, 0.3, -0.2, 0.8
    s.set_state(uu)
    s.run(5)
    np.testing.assert_allclose(s.state(), uu, rtol=1e-13, atol=1e-14)


def test_alfven_wave_travels_at_va():
    nx, ny = 64, 8
    X, Y = _grid(nx, ny)
    eps, rho, b0 = 1e-5, 1.0, 1.0
    uu = np.zeros((8, nx, ny))
    uu[0], uu[4], uu[7] = rho, b0, 1.0
    uu[6] = eps * np.cos(X)
    uu[3] = -eps * np.cos(X) / np.sqrt(rho)
    s = orc.Sim(nx, ny, cfl=0.3)
    s.set_state(uu)
    s.run(40)
    t = s.time
    va = b0 / np.sqrt(rho)
    bz = s.state()[6]
    want = eps * np.cos(X - va * t)
    err = np.abs(bz - want).max() / eps
    assert t > 0.5
    # residual is the 4th-order end-row closures of the x operators
    assert err < 5e-5, err
    # a wave going the wrong way (sign slip in the induction or Lorentz term)
    # would be off by O(1)
    assert np.abs(bz - eps * np.cos(X + va * t)).max() / eps > 0.1


def test_induction_preserves_discrete_divb():
    nx, ny = 32, 32
    X, Y = _grid(nx, ny)
    uu = orszag_tang(nx, ny)
    uu[3] = 0.1 * np.sin(X + 2 * Y)
    uu[6] = 0.05 * np.cos(3 * X - Y)
    dx = 2 * np.pi / nx
    dudx, dudy = orc.derivX(uu, dx), orc.derivY(uu, 2 * np.pi)
    E = orc.efield(uu)
    R = orc.rhs(uu, dudx, dudy, orc.derivX(E, dx), orc.derivY(E, 2 * np.pi))
    div = orc.derivX(R[4:5], dx)[0] + orc.derivY(R[5:6], 2 * np.pi)[0]
    assert np.abs(div).max() < 1e-12 * np.abs(R[4:6]).max()


def test_orszag_tang_runs_finite_and_deterministic():
    a, b = orc.Sim(32, 32, hall=1, di=0.05), orc.Sim(32, 32, hall=1, di=0.05)
    uu = orszag_tang(32, 32)
    a.set_state(uu)
    b.set_state(uu)
    a.run(20)
    b.run(20)
    assert np.isfinite(a.state()).all()
    np.testing.assert_array_equal(a.state(), b.state())
    assert a.steps == 20 and a.time > 0


def test_fixed_dt_replay_matches_adaptive_run():
    uu = orszag_tang(32, 16)
    a = orc.Sim(32, 16)
    a.set_state(uu)
    dts = []
    for _ in range(6):
        a.run(1)
        dts.append(a.dt)
    b = orc.Sim(32, 16)
    b.set_state(uu)
    b.run_fixed_dt(np.array(dts))
    np.testing.assert_array_equal(a.state(), b.state())


# ---------------------------------------------------------------------------
# The composed stage loop as a linear map.  Around a uniform state the run is
# smooth in the perturbation, so the antisymmetric part of two runs from
# bg + eps v and bg - eps v is eps J v + O(eps^3): J, the Jacobian of K steps
# of the stage loop, is written here from the vector form of linearised MHD
# and the x and y operators assembled independently of the oracle (dense
# block-closure matrices from the published coefficients, numpy FFTs), with
# the stage order of mhd.c:154-206 and filtering.c:98-102:
#   R = L(q);  q <- Fy Fx (q + c_k dt R + d_k dt R_prev);  R_prev <- R.
# A dropped x or y filter, a filter moved before the RK update, a wrong RK
# coefficient or a missing / mis-signed term changes the result at O(1e-3)
# or more of the perturbation, against a 1e-8 tolerance.
# ---------------------------------------------------------------------------
from test_oracle_lines import _dense_operator  # noqa: E402

GAMMA = 5.0 / 3.0
RK_C = (8 / 15, 5 / 12, 3 / 4)        # Spalart, Moser & Rogers (1991)
RK_D = (0.0, -17 / 60, -5 / 12)


def _y_ops(ny, Ly):
    """spectral d/dy (Nyquist bin dropped), d2/dy2 (Nyquist kept) and the
    compact y filter response (alpha_f = 0.45, published 4th-order form) as
    functions on [..., ny] arrays"""
    m = np.fft.fftfreq(ny, 1.0 / ny)              # signed wavenumber index
    k = 2 * np.pi / Ly * m
    k1 = k.copy()
    k1[ny // 2] = 0.0
    w = 2 * np.pi * np.abs(m) / ny
    a = 0.45
    filt = ((5 + 6 * a) / 8 + (1 + 2 * a) / 2 * np.cos(w) - (1 - 2 * a) / 8 * np.cos(2 * w)) / (
        1 + 2 * a * np.cos(w))

    def apply(f, mult):
        return np.real(np.fft.ifft(np.fft.fft(f, axis=-1) * mult, axis=-1))

    return (lambda f: apply(f, 1j * k1), lambda f: apply(f, -k ** 2),
            lambda f: apply(f, filt))


def _linear_stage_map(nx, ny, bg, dt, nsteps, q0, npe=1, visc=0.0, resist=0.0, divbclean=0,
                      variant=None):
    """K steps of the linearised stage loop from perturbation q0 [8, nx, ny];
    variant "no_fx" drops the x filter, "fx_first" filters in x before the RK
    update (the plausible composition slips the test must tell apart)"""
    Lx = Ly = 2 * np.pi
    h = Lx / nx
    sizes = [nx // npe + (1 if r < nx % npe else 0) for r in range(npe)]
    Dx = _dense_operator("d1", nx, h, sizes)
    Dxx = _dense_operator("d2", nx, h, sizes)
    Fx = _dense_operator("f", nx, h, sizes)
    Dy, Dyy, Fy = _y_ops(ny, Ly)
    rho0, p0 = bg[0], bg[7]
    U, B0 = np.array(bg[1:4]), np.array(bg[4:7])
    cdiff = 2 * max(h, Ly / ny) ** 2              # constDiffDivB (initialize.c:17)

    def grad(f):
        return np.stack([np.einsum("ik,kj->ij", Dx, f), Dy(f), np.zeros_like(f)])

    def lap(f):
        return np.einsum("ik,kj->ij", Dxx, f) + Dyy(f)

    def div(v):
        return np.einsum("ik,kj->ij", Dx, v[0]) + Dy(v[1])

    def curl(v):
        gx = lambda f: np.einsum("ik,kj->ij", Dx, f)  # noqa: E731
        return np.stack([Dy(v[2]), -gx(v[2]), gx(v[1]) - Dy(v[0])])

    def adv(f):
        g = grad(f)
        return U[0] * g[0] + U[1] * g[1]

    def L(q):
        drho, du, db, dp = q[0], q[1:4], q[4:7], q[7]
        Bc = B0[:, None, None]
        Uc = U[:, None, None]
        r = np.zeros_like(q)
        r[0] = -adv(drho) - rho0 * div(du)
        j = curl(db)
        r[1:4] = (-np.stack([adv(c) for c in du]) - grad(dp) / rho0
                  + np.cross(j, Bc, axis=0) / rho0 + visc / rho0 * np.stack([lap(c) for c in du]))
        r[4:7] = curl(np.cross(du, Bc, axis=0) + np.cross(Uc, db, axis=0)) + resist * np.stack(
            [lap(c) for c in db])
        if divbclean == 1:
            gdb = grad(div(db))
            r[4:6] += cdiff * gdb[:2]
        r[7] = -adv(dp) - GAMMA * p0 * div(du)
        return r

    q = q0.copy()
    for _ in range(nsteps):
        rp = np.zeros_like(q)
        for s in range(3):
            r = L(q)
            if variant == "fx_first":
                q = np.einsum("ik,vkj->vij", Fx, q)
            q = q + RK_C[s] * dt * r + RK_D[s] * dt * rp
            if variant is None:
                q = np.einsum("ik,vkj->vij", Fx, q)
            q = Fy(q)
            rp = r
    return q


@pytest.mark.parametrize("case", ["ideal", "npe3", "dissipative", "diffusive_divb"])
def test_stage_loop_is_the_linearised_map(case):
    nx, ny, nsteps, eps = 32, 16, 10, 1e-5
    bg = [1.2, 0.15, -0.1, 0.05, 1.0, 0.4, -0.3, 0.9]   # rho, u, B, p
    kw = dict(npe=1, visc=0.0, resist=0.0, divbclean=0)
    if case == "npe3":
        kw["npe"] = 3
    elif case == "dissipative":
        kw.update(visc=2e-3, resist=3e-3)
    elif case == "diffusive_divb":
        kw["divbclean"] = 1
    rng = np.random.default_rng(11)
    v = rng.standard_normal((8, nx, ny))
    u_bg = np.broadcast_to(np.array(bg)[:, None, None], (8, nx, ny)).copy()
    dt = 0.3 * (2 * np.pi / nx) / 2.5
    runs = []
    for sgn in (1.0, -1.0):
        s = orc.Sim(nx, ny, gamma=GAMMA, resistivity=kw["resist"], viscosity=kw["visc"],
                    divbclean=kw["divbclean"], npe=kw["npe"])
        s.set_state(u_bg + sgn * eps * v)
        s.set_dt(dt)
        s.run_fixed_dt(np.full(nsteps, dt))
        runs.append(s.state())
    got = (runs[0] - runs[1]) / (2 * eps)
    want = _linear_stage_map(nx, ny, bg, dt, nsteps, v, **kw)
    scale = np.abs(want).max()
    err = np.abs(got - want).max() / scale
    assert err < 1e-8, (case, err)
    # the test has teeth: the map without the x filter of each stage, or
    # with the x filter before the update, is far from the oracle's run
    for variant in ("no_fx", "fx_first"):
        other = _linear_stage_map(nx, ny, bg, dt, nsteps, v, variant=variant, **kw)
        assert np.abs(got - other).max() > 1e-4 * scale, variant
