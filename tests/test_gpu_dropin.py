"""Parity of the drop-in entry points (reference names, reference globals)
with the FP64 oracle, one function at a time, then the reference's own stage
loop (mhd.c:154-206) driven through them.

Gate: per-field relative L2 <= 1e-12 (FP64), random and smooth inputs,
smallest and largest supported line lengths.
"""
import sys

import numpy as np
import pytest

import oracle_lib as orc
from util import FP64_TOL, assert_fields_close, random_state

pytestmark = pytest.mark.gpu

SIZES = [(64, 32), (32, 64), (16, 16), (128, 128), (2048, 16), (4096, 16), (16, 4096),
         (256, 512)]


@pytest.fixture(scope="module")
def m():
    import mhd2dcfdfftw_open_amd as m
    m.lib()
    return m


def _globals(m, nx, ny, **kw):
    print(f"[dropin] nx={nx} ny={ny} {kw}", flush=True)
    return m.RefGlobals(nx, ny, **kw)


def _smooth(nx, ny, seed):
    rng = np.random.default_rng(seed)
    x = np.arange(nx)[:, None] * 2 * np.pi / nx
    y = np.arange(ny)[None, :] * 2 * np.pi / ny
    uu = np.zeros((8, nx, ny))
    for v in range(8):
        a, b, c, d = rng.integers(1, 5, 4)
        uu[v] = np.sin(a * x + c * y + rng.uniform()) + 0.3 * np.cos(b * x - d * y)
    uu[0] += 2.0
    uu[7] += 2.0
    return uu


def _state(nx, ny, kind, seed):
    return random_state(nx, ny, seed) if kind == "random" else _smooth(nx, ny, seed)


@pytest.mark.parametrize("nx,ny", SIZES)
@pytest.mark.parametrize("kind", ["random", "smooth"])
def test_derivX_derivY(m, nx, ny, kind):
    G = _globals(m, nx, ny, Lx=3.0, Ly=5.0)
    L = m.lib()
    try:
        uu = G.array("uu")
        uu[:] = _state(nx, ny, kind, nx + ny)
        L.derivX(8, G.ptr("uu"), G.ptr("uuHaloX0"), G.ptr("uuHaloX1"), G.ptr("dudx"), 0, 0)
        L.derivY(8, G.ptr("uu"), G.ptr("dudy"), 0)
        L.mhd2d_dropin_sync_host()
        dx = 3.0 / nx
        assert_fields_close(G.array("dudx"), orc.derivX(uu, dx), what="derivX")
        assert_fields_close(G.array("dudy"), orc.derivY(uu, 5.0), what="derivY")
    finally:
        G.close()


@pytest.mark.parametrize("nx,ny", [(64, 32), (32, 64), (2048, 16), (4096, 16), (16, 4096)])
def test_derivXX_derivYY_field_offsets(m, nx, ny):
    # mhd.c:165-172 calls these on fields 1..3 and 4..6 of uu
    G = _globals(m, nx, ny, Lx=2.0, Ly=4.0, viscosity=0.01)
    L = m.lib()
    try:
        uu = G.array("uu")
        uu[:] = random_state(nx, ny, 3)
        for start, xname, yname in ((1, "d2udx2", "d2udy2"), (4, "d2Bdx2", "d2Bdy2")):
            L.derivXX(3, G.ptr("uu"), G.ptr("uuHaloX0"), G.ptr("uuHaloX1"), G.ptr(xname),
                      start, start)
            L.derivYY(3, G.ptr("uu"), G.ptr(yname), start)
            L.mhd2d_dropin_sync_host()
            sub = uu[start:start + 3]
            assert_fields_close(G.array(xname), orc.derivXX(sub, 2.0 / nx), what="derivXX")
            assert_fields_close(G.array(yname), orc.derivYY(sub, 4.0), what="derivYY")
    finally:
        G.close()


@pytest.mark.parametrize("nx,ny", SIZES)
def test_filtering(m, nx, ny):
    G = _globals(m, nx, ny)
    L = m.lib()
    try:
        uu = G.array("uu")
        u0 = random_state(nx, ny, 9)
        uu[:] = u0
        L.sendrecvHaloPointsX()
        L.filtering()
        L.mhd2d_dropin_sync_host()
        assert_fields_close(uu, orc.filterY(orc.filterX(u0)), what="filtering")
    finally:
        G.close()


def test_efield_divb(m):
    nx, ny = 64, 64
    G = _globals(m, nx, ny)
    L = m.lib()
    try:
        uu, dudx, dudy = G.array("uu"), G.array("dudx"), G.array("dudy")
        uu[:] = random_state(nx, ny, 1)
        dudx[:] = random_state(nx, ny, 2)
        dudy[:] = random_state(nx, ny, 3)
        L.calcEfield()
        L.calcDivB()
        L.mhd2d_dropin_sync_host()
        assert_fields_close(G.array("Efield"), orc.efield(uu), what="calcEfield")
        assert_fields_close(G.array("divB")[None], orc.divb(dudx, dudy)[None], what="calcDivB")
    finally:
        G.close()


@pytest.mark.parametrize("variant", ["ideal", "viscous", "resistive", "divbclean"])
def test_calcRHS(m, variant):
    nx, ny = 32, 64
    kw = dict(ideal={}, viscous=dict(viscosity=0.03), resistive=dict(resistivity=0.02),
              divbclean=dict(divbclean=1))[variant]
    G = _globals(m, nx, ny, gamma=1.4, **kw)
    L = m.lib()
    try:
        rng = np.random.default_rng(5)
        arrs = {}
        for name in ("uu", "dudx", "dudy", "dEdx", "dEdy", "d2udx2", "d2udy2", "d2Bdx2",
                     "d2Bdy2", "ddivBdx", "ddivBdy"):
            a = G.array(name)
            a[:] = rng.uniform(-1, 1, a.shape)
            arrs[name] = a
        arrs["uu"][0] = 1.0 + rng.uniform(0, 1, (nx, ny))
        L.calcRHS()
        L.mhd2d_dropin_sync_host()
        cdd = G.scalar("constDiffDivB").value
        want = orc.rhs(arrs["uu"], arrs["dudx"], arrs["dudy"], arrs["dEdx"], arrs["dEdy"],
                       gamma=1.4, resistivity=kw.get("resistivity", 0.0),
                       viscosity=kw.get("viscosity", 0.0), divbclean=kw.get("divbclean", 0),
                       constDiffDivB=cdd, d2udx2=arrs["d2udx2"], d2udy2=arrs["d2udy2"],
                       d2Bdx2=arrs["d2Bdx2"], d2Bdy2=arrs["d2Bdy2"],
                       ddivBdx=arrs["ddivBdx"][0], ddivBdy=arrs["ddivBdy"][0])
        assert_fields_close(G.array("dudt"), want, what=f"calcRHS[{variant}]")
    finally:
        G.close()


@pytest.mark.parametrize("kw", [dict(divbclean=2), dict(divbclean=2, viscosity=0.03,
                                                         resistivity=0.02)])
def test_calcRHS_hyperbolic_cleaning(m, kw):
    nx, ny = 32, 64
    G = _globals(m, nx, ny, gamma=1.4, **kw)
    L = m.lib()
    try:
        assert G.nvar == 9
        rng = np.random.default_rng(8)
        arrs = {}
        for name in ("uu", "dudx", "dudy", "dEdx", "dEdy", "d2udx2", "d2udy2", "d2Bdx2",
                     "d2Bdy2"):
            a = G.array(name)
            a[:] = rng.uniform(-1, 1, a.shape)
            arrs[name] = a
        arrs["uu"][0] = 1.0 + rng.uniform(0, 1, (nx, ny))
        G.scalar("Ch_divBcleaning").value = 1.7
        L.calc_Cp_fix_Cr()
        cp = G.scalar("Cp_divBcleaning").value
        assert cp == np.sqrt(0.18 * 1.7)
        L.calcRHS()
        L.mhd2d_dropin_sync_host()
        want = orc.rhs_glm(arrs["uu"], arrs["dudx"], arrs["dudy"], arrs["dEdx"], arrs["dEdy"],
                           1.7, cp, gamma=1.4, resistivity=kw.get("resistivity", 0.0),
                           viscosity=kw.get("viscosity", 0.0), d2udx2=arrs["d2udx2"],
                           d2udy2=arrs["d2udy2"], d2Bdx2=arrs["d2Bdx2"], d2Bdy2=arrs["d2Bdy2"])
        assert_fields_close(G.array("dudt"), want, what=f"calcRHS[glm {kw}]")
    finally:
        G.close()


@pytest.mark.parametrize("hall", [0, 1])
def test_calcDt_sets_glm_speeds_from_the_last_cell(m, hall):
    nx, ny = 64, 32
    G = _globals(m, nx, ny, Lx=2.0, Ly=3.0, cfl=0.4, hall=hall, di=0.3, divbclean=2)
    L = m.lib()
    try:
        uu = G.array("uu")
        uu[:8] = random_state(nx, ny, 12)
        uu[8] = 0.1
        want_dt = orc.calc_dt(uu[:8], 2.0 / nx, 3.0 / ny, cfl=0.4, hall=hall, di=0.3, dt0=0.0)
        got = L.calcDt(0.0)
        assert abs(got - want_dt) <= 1e-14 * want_dt
        ch = orc.glm_ch(uu[:8], 2.0 / nx, 3.0 / ny, nx - 1, ny - 1, hall=hall, di=0.3)
        assert abs(G.scalar("Ch_divBcleaning").value - ch) <= 1e-14 * ch
        assert G.scalar("Cp_divBcleaning").value == np.sqrt(0.18 * G.scalar("Ch_divBcleaning").value)
    finally:
        G.close()


def test_rkinit_rkexecute(m):
    nx, ny = 32, 32
    G = _globals(m, nx, ny)
    L = m.lib()
    try:
        rng = np.random.default_rng(6)
        uu, dudt = G.array("uu"), G.array("dudt")
        uu[:] = rng.standard_normal(uu.shape)
        u = uu.copy()
        rk = np.zeros_like(u)
        dt = 0.0123
        cc = [8 / 15 * dt, 5 / 12 * dt, 0.75 * dt]
        dd = [0.0, -17 / 60 * dt, -5 / 12 * dt]
        L.rkInit(dt)
        for s in range(3):
            r = rng.standard_normal(uu.shape)
            dudt[:] = r
            L.mhd2d_dropin_invalidate_device()   # host wrote dudt
            # the host copies of uu/dudtRK are current: re-upload is harmless
            uu[:] = u
            G.array("dudtRK")[:] = rk
            L.rkExecute(s)
            L.mhd2d_dropin_sync_host()
            u, rk = orc.rk_execute(u, r, rk, cc[s], dd[s])
            assert_fields_close(uu, u, what=f"rkExecute[{s}]")
            assert_fields_close(G.array("dudtRK"), rk, what=f"dudtRK[{s}]")
    finally:
        G.close()


@pytest.mark.parametrize("hall,eta,nu", [(0, 0.0, 0.0), (1, 0.0, 0.0), (0, 0.5, 0.0),
                                         (0, 0.0, 0.7)])
def test_calcDt(m, hall, eta, nu):
    nx, ny = 64, 32
    G = _globals(m, nx, ny, Lx=2.0, Ly=3.0, cfl=0.4, hall=hall, di=0.3, resistivity=eta,
                 viscosity=nu)
    L = m.lib()
    try:
        uu = G.array("uu")
        uu[:] = random_state(nx, ny, 11)
        want = orc.calc_dt(uu, 2.0 / nx, 3.0 / ny, cfl=0.4, hall=hall, di=0.3,
                           resistivity=eta, viscosity=nu, dt0=0.0)
        got = L.calcDt(0.0)
        assert abs(got - want) <= 1e-14 * want
        # hysteresis keeps dt0 inside [0.98, 1.02] * dtmin
        assert L.calcDt(1.01 * want) == 1.01 * want
        assert abs(L.calcDt(1.03 * want) - want) <= 1e-14 * want
    finally:
        G.close()


def _mhd_c_loop(L, G, nsteps):
    """The reference's main loop body, mhd.c:154-206, through the drop-in."""
    P = G.ptr
    nvar = G.nvar
    visc = G.scalar("viscosity").value
    eta = G.scalar("resistivity").value
    hall = G.scalar("Hall", __import__("ctypes").c_int).value
    dbc = G.scalar("divBcleaning", __import__("ctypes").c_int).value
    dt = L.calcDt(0.0)
    L.rkInit(dt)
    t = 0.0
    for _ in range(nsteps):
        for irk in range(3):
            L.sendrecvHaloPointsX()
            L.derivX(nvar, P("uu"), P("uuHaloX0"), P("uuHaloX1"), P("dudx"), 0, 0)
            L.derivY(nvar, P("uu"), P("dudy"), 0)
            if visc > 0:
                L.derivXX(3, P("uu"), P("uuHaloX0"), P("uuHaloX1"), P("d2udx2"), 1, 1)
                L.derivYY(3, P("uu"), P("d2udy2"), 1)
            if hall == 1 or eta > 0:
                L.derivXX(3, P("uu"), P("uuHaloX0"), P("uuHaloX1"), P("d2Bdx2"), 4, 4)
                L.derivYY(3, P("uu"), P("d2Bdy2"), 4)
            if dbc == 1:
                L.calcDivB()
                L.sendrecvHaloPointsdivBX()
                L.derivX(1, P("divB"), P("divBHaloX0"), P("divBHaloX1"), P("ddivBdx"), 0, 0)
                L.derivY(1, P("divB"), P("ddivBdy"), 0)
            L.calcEfield()
            L.sendrecvHaloPointsEfieldX()
            L.derivX(3, P("Efield"), P("EfieldHaloX0"), P("EfieldHaloX1"), P("dEdx"), 0, 0)
            L.derivY(3, P("Efield"), P("dEdy"), 0)
            L.calcRHS()
            L.rkExecute(irk)
            L.sendrecvHaloPointsX()
            L.filtering()
        t += dt
        dt = L.calcDt(dt)
        L.rkInit(dt)
    return t, dt


@pytest.mark.parametrize("case", ["ot_hall", "ot_visc_eta", "turb_divbclean", "turb_glm",
                                  "ot_glm_hall_visc"])
def test_reference_stage_loop_matches_oracle(m, case):
    import mhd2dcfdfftw_open_amd as pkg
    nx, ny, steps = 64, 32, 8
    kw = dict(ot_hall=dict(hall=1, di=0.05),
              ot_visc_eta=dict(viscosity=2e-3, resistivity=1e-3),
              turb_divbclean=dict(divbclean=1),
              turb_glm=dict(divbclean=2),
              ot_glm_hall_visc=dict(divbclean=2, hall=1, di=0.05, viscosity=2e-3))[case]
    u0 = pkg.orszag_tang(nx, ny) if case.startswith("ot") else pkg.turbulence(nx, ny, seed=3,
                                                                              kmax=6)
    G = _globals(m, nx, ny, **kw)
    L = m.lib()
    try:
        G.array("uu")[:8] = u0             # psi (field 8) starts at 0
        t, dt = _mhd_c_loop(L, G, steps)   # strict mode: calcDt refreshed the host arrays
        S = orc.Sim(nx, ny, **kw)
        S.set_state(u0)
        S.run(steps)
        assert abs(dt - S.dt) <= 1e-12 * S.dt and abs(t - S.time) <= 1e-12 * S.time
        assert_fields_close(G.array("uu"), S.state(), what=f"stage loop [{case}]")
        if G.nvar == 9:   # psi itself, at 10x its conditioning floor (tests/util.py)
            psi_err = np.linalg.norm(G.array("uu")[8] - S.state()[8]) / np.linalg.norm(S.state()[8])
            assert psi_err <= 1e-8, psi_err
    finally:
        G.close()
