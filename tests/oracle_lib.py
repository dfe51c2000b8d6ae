"""ctypes view of the FP64 C restatement in oracle/ (test infrastructure only).

The oracle is the checker for the HIP build: only tests, the smoke check and
the CPU-baseline leg of bench.py load it.
"""
import ctypes
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
ORACLE_SO = os.path.join(ORACLE_DIR, "libmhd2d_oracle.so")
# the same code built with OpenMP over independent lines and cells (results
# equal bit for bit); used where a large grid would make the checker slow
ORACLE_OMP_SO = os.path.join(ORACLE_DIR, "libmhd2d_oracle_omp.so")

_dp = ctypes.POINTER(ctypes.c_double)
_lib = None


def _ptr(a):
    return a.ctypes.data_as(_dp)


def build():
    subprocess.run(["make", "-s", "-C", ORACLE_DIR], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            build()
        L = ctypes.CDLL(ORACLE_SO)
        i, d, p, sz = ctypes.c_int, ctypes.c_double, _dp, ctypes.c_size_t
        sig = {
            "orc_thomas": (None, [i, p, p, p, p, p]),
            "orc_line_deriv1": (None, [i, d, p, p, p, p]),
            "orc_line_deriv2": (None, [i, d, p, p, p, p]),
            "orc_line_filter": (None, [i, p, p, p, p]),
            "orc_rfft": (None, [i, p, p, p]),
            "orc_irfft": (None, [i, p, p, p]),
            "orc_filtY_table": (None, [i, d, p]),
            "orc_derivX": (None, [i, i, d, i, p, p]),
            "orc_derivXX": (None, [i, i, d, i, p, p]),
            "orc_derivY": (None, [i, i, d, i, p, p]),
            "orc_derivYY": (None, [i, i, d, i, p, p]),
            "orc_filterX": (None, [i, i, i, p]),
            "orc_filterY": (None, [i, i, i, p]),
            "orc_derivX_npe": (None, [i, i, d, i, i, p, p]),
            "orc_derivXX_npe": (None, [i, i, d, i, i, p, p]),
            "orc_filterX_npe": (None, [i, i, i, i, p]),
            "orc_sim_set_npe": (None, [ctypes.c_void_p, i]),
            "orc_sim_set_open": (None, [ctypes.c_void_p, i]),
            "orc_derivX_open": (None, [i, i, d, i, i, p, p]),
            "orc_derivXX_open": (None, [i, i, d, i, i, p, p]),
            "orc_filterX_open": (None, [i, i, i, i, p]),
            "orc_characteristics_x": (None, [i, i, i, i, i, d, p, p, p]),
            "orc_sim_nvar": (i, [ctypes.c_void_p]),
            "orc_sim_glm_ch": (d, [ctypes.c_void_p, i]),
            "orc_sim_glm_cp": (d, [ctypes.c_void_p, i]),
            "orc_rhs_glm": (None, [i, i, d, d, d, d, d] + [p] * 10),
            "orc_glm_ch": (d, [i, i, d, d, d, i, d, p, i, i]),
            "orc_efield": (None, [i, i, p, p]),
            "orc_divb": (None, [i, i, p, p, p]),
            "orc_rhs": (None, [i, i, d, d, d, i, d] + [p] * 12),
            "orc_rk_execute": (None, [sz, d, d, p, p, p]),
            "orc_calc_dt": (d, [i, i, d, d, d, d, i, d, d, d, p, d]),
            "orc_energy": (d, [i, i, d, d, d, p]),
            "orc_sim_create": (ctypes.c_void_p, [i, i, d, d, d, d, i, d, d, d, i]),
            "orc_sim_destroy": (None, [ctypes.c_void_p]),
            "orc_sim_set_state": (None, [ctypes.c_void_p, p]),
            "orc_sim_get_state": (None, [ctypes.c_void_p, p]),
            "orc_sim_get_derivs": (None, [ctypes.c_void_p, p, p]),
            "orc_sim_run": (None, [ctypes.c_void_p, i]),
            "orc_sim_run_fixed_dt": (None, [ctypes.c_void_p, i, p]),
            "orc_sim_stage": (None, [ctypes.c_void_p, i]),
            "orc_sim_set_dt": (None, [ctypes.c_void_p, d]),
            "orc_set_threads": (None, [i]),
            "orc_threads": (i, []),
            "orc_sim_dt": (d, [ctypes.c_void_p]),
            "orc_sim_time": (d, [ctypes.c_void_p]),
            "orc_sim_steps": (ctypes.c_long, [ctypes.c_void_p]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


class threaded:
    """Context manager: calls made inside it go to the OpenMP build of the
    checker (same results bit for bit as the single-threaded build)."""

    def __enter__(self):
        global _lib, ORACLE_SO
        if not os.path.exists(ORACLE_OMP_SO):
            build()
        self._saved = (_lib, ORACLE_SO)
        _lib, ORACLE_SO = None, ORACLE_OMP_SO
        return self

    def __exit__(self, *exc):
        global _lib, ORACLE_SO
        _lib, ORACLE_SO = self._saved
        return False


def _c(a):
    return np.ascontiguousarray(a, dtype=np.float64)


# ---------------------------------------------------------------- lines
def thomas(lo, diag, up, rhs):
    lo, diag, up, rhs = map(_c, (lo, diag, up, rhs))
    x = np.zeros_like(rhs)
    lib().orc_thomas(len(rhs), _ptr(lo), _ptr(diag), _ptr(up), _ptr(rhs), _ptr(x))
    return x


def line_op(kind, A, H0, H1, DX=1.0):
    A, H0, H1 = _c(A), _c(H0), _c(H1)
    X = np.zeros_like(A)
    L = lib()
    if kind == "d1":
        L.orc_line_deriv1(len(A), DX, _ptr(A), _ptr(H0), _ptr(H1), _ptr(X))
    elif kind == "d2":
        L.orc_line_deriv2(len(A), DX, _ptr(A), _ptr(H0), _ptr(H1), _ptr(X))
    else:
        L.orc_line_filter(len(A), _ptr(A), _ptr(H0), _ptr(H1), _ptr(X))
    return X


def rfft(x):
    x = _c(x)
    n = len(x)
    re = np.zeros(n // 2 + 1)
    im = np.zeros(n // 2 + 1)
    lib().orc_rfft(n, _ptr(x), _ptr(re), _ptr(im))
    return re + 1j * im


def irfft(X, n):
    re, im = _c(np.real(X)), _c(np.imag(X))
    x = np.zeros(n)
    lib().orc_irfft(n, _ptr(re), _ptr(im), _ptr(x))
    return x


def filtY_table(ny, afy=0.45):
    t = np.zeros(ny // 2 + 1)
    lib().orc_filtY_table(ny, afy, _ptr(t))
    return t


# ---------------------------------------------------------------- fields
def derivX(arr, dx, npe=1):
    arr = _c(arr)
    nv, nx, ny = arr.shape
    out = np.zeros_like(arr)
    if npe == 1:
        lib().orc_derivX(nx, ny, dx, nv, _ptr(arr), _ptr(out))
    else:
        lib().orc_derivX_npe(nx, ny, dx, nv, npe, _ptr(arr), _ptr(out))
    return out


def derivXX(arr, dx, npe=1):
    arr = _c(arr)
    nv, nx, ny = arr.shape
    out = np.zeros_like(arr)
    if npe == 1:
        lib().orc_derivXX(nx, ny, dx, nv, _ptr(arr), _ptr(out))
    else:
        lib().orc_derivXX_npe(nx, ny, dx, nv, npe, _ptr(arr), _ptr(out))
    return out


def derivX_open(arr, dx, npe):
    arr = _c(arr)
    nv, nx, ny = arr.shape
    out = np.zeros_like(arr)
    lib().orc_derivX_open(nx, ny, dx, nv, npe, _ptr(arr), _ptr(out))
    return out


def derivXX_open(arr, dx, npe):
    arr = _c(arr)
    nv, nx, ny = arr.shape
    out = np.zeros_like(arr)
    lib().orc_derivXX_open(nx, ny, dx, nv, npe, _ptr(arr), _ptr(out))
    return out


def filterX_open(arr, npe):
    out = _c(arr).copy()
    nv, nx, ny = out.shape
    lib().orc_filterX_open(nx, ny, nv, npe, _ptr(out))
    return out


def characteristics_x(uu, dudx, i, j, boundary, gamma=5.0 / 3.0):
    uu, dudx = _c(uu), _c(dudx)
    _, nx, ny = uu.shape
    out = np.zeros(7)
    lib().orc_characteristics_x(nx, ny, i, j, boundary, gamma, _ptr(uu), _ptr(dudx), _ptr(out))
    return out


def derivY(arr, Ly):
    arr = _c(arr)
    nv, nx, ny = arr.shape
    out = np.zeros_like(arr)
    lib().orc_derivY(nx, ny, Ly, nv, _ptr(arr), _ptr(out))
    return out


def derivYY(arr, Ly):
    arr = _c(arr)
    nv, nx, ny = arr.shape
    out = np.zeros_like(arr)
    lib().orc_derivYY(nx, ny, Ly, nv, _ptr(arr), _ptr(out))
    return out


def filterX(arr, npe=1):
    out = _c(arr).copy()
    nv, nx, ny = out.shape
    if npe == 1:
        lib().orc_filterX(nx, ny, nv, _ptr(out))
    else:
        lib().orc_filterX_npe(nx, ny, nv, npe, _ptr(out))
    return out


def filterY(arr):
    out = _c(arr).copy()
    nv, nx, ny = out.shape
    lib().orc_filterY(nx, ny, nv, _ptr(out))
    return out


def efield(uu):
    uu = _c(uu)
    _, nx, ny = uu.shape
    E = np.zeros((3, nx, ny))
    lib().orc_efield(nx, ny, _ptr(uu), _ptr(E))
    return E


def divb(dudx, dudy):
    dudx, dudy = _c(dudx), _c(dudy)
    _, nx, ny = dudx.shape
    out = np.zeros((nx, ny))
    lib().orc_divb(nx, ny, _ptr(dudx), _ptr(dudy), _ptr(out))
    return out


def rhs(uu, dudx, dudy, dEdx, dEdy, gamma=5.0 / 3.0, resistivity=0.0,
        viscosity=0.0, divbclean=0, constDiffDivB=0.0, d2udx2=None,
        d2udy2=None, d2Bdx2=None, d2Bdy2=None, ddivBdx=None, ddivBdy=None):
    uu = _c(uu)
    _, nx, ny = uu.shape
    z3 = np.zeros((3, nx, ny))
    z1 = np.zeros((nx, ny))
    args = [uu, dudx, dudy, dEdx, dEdy,
            z3 if d2udx2 is None else d2udx2, z3 if d2udy2 is None else d2udy2,
            z3 if d2Bdx2 is None else d2Bdx2, z3 if d2Bdy2 is None else d2Bdy2,
            z1 if ddivBdx is None else ddivBdx, z1 if ddivBdy is None else ddivBdy]
    args = [_c(a) for a in args]
    out = np.zeros((8, nx, ny))
    lib().orc_rhs(nx, ny, gamma, resistivity, viscosity, divbclean,
                  constDiffDivB, *[_ptr(a) for a in args], _ptr(out))
    return out


def rhs_glm(uu, dudx, dudy, dEdx, dEdy, ch, cp, gamma=5.0 / 3.0, resistivity=0.0,
            viscosity=0.0, d2udx2=None, d2udy2=None, d2Bdx2=None, d2Bdy2=None):
    """calcRHS with divBcleaning == 2: 9 fields in uu, dudx, dudy and out."""
    uu = _c(uu)
    _, nx, ny = uu.shape
    z3 = np.zeros((3, nx, ny))
    args = [uu, dudx, dudy, dEdx, dEdy,
            z3 if d2udx2 is None else d2udx2, z3 if d2udy2 is None else d2udy2,
            z3 if d2Bdx2 is None else d2Bdx2, z3 if d2Bdy2 is None else d2Bdy2]
    args = [_c(a) for a in args]
    out = np.zeros((9, nx, ny))
    lib().orc_rhs_glm(nx, ny, gamma, resistivity, viscosity, ch, cp,
                      *[_ptr(a) for a in args], _ptr(out))
    return out


def glm_ch(uu, dx, dy, i, j, gamma=5.0 / 3.0, hall=0, di=0.0):
    uu = _c(uu)
    _, nx, ny = uu.shape
    return lib().orc_glm_ch(nx, ny, dx, dy, gamma, hall, di, _ptr(uu), i, j)


def rk_execute(uu, dudt, dudtRK, cc, dd):
    uu, dudt, dudtRK = _c(uu).copy(), _c(dudt), _c(dudtRK).copy()
    lib().orc_rk_execute(uu.size, cc, dd, _ptr(uu), _ptr(dudt), _ptr(dudtRK))
    return uu, dudtRK


def calc_dt(uu, dx, dy, gamma=5.0 / 3.0, cfl=0.3, hall=0, di=0.0,
            resistivity=0.0, viscosity=0.0, dt0=0.0):
    uu = _c(uu)
    _, nx, ny = uu.shape
    return lib().orc_calc_dt(nx, ny, dx, dy, gamma, cfl, hall, di, resistivity,
                             viscosity, _ptr(uu), dt0)


def energy(uu, dx, dy, gamma=5.0 / 3.0):
    uu = _c(uu)
    _, nx, ny = uu.shape
    return lib().orc_energy(nx, ny, dx, dy, gamma, _ptr(uu))


class Sim:
    """Whole-run driver replaying the reference stage loop (mhd.c:154-206)."""

    def __init__(self, nx, ny, Lx=2 * np.pi, Ly=2 * np.pi, gamma=5.0 / 3.0,
                 cfl=0.3, hall=0, di=0.0, resistivity=0.0, viscosity=0.0,
                 divbclean=0, npe=1, open_x=False):
        self.nx, self.ny = nx, ny
        self.h = lib().orc_sim_create(nx, ny, Lx, Ly, gamma, cfl, hall, di,
                                      resistivity, viscosity, divbclean)
        lib().orc_sim_set_npe(self.h, npe)
        lib().orc_sim_set_open(self.h, int(open_x))
        self.nvar = lib().orc_sim_nvar(self.h)

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_sim_destroy(self.h)
            self.h = None

    def set_state(self, uu):
        uu = _c(uu)
        if uu.shape[0] == 8 and self.nvar == 9:   # psi starts at 0 (initialize.c:221)
            uu = np.concatenate([uu, np.zeros((1, self.nx, self.ny))])
        assert uu.shape == (self.nvar, self.nx, self.ny)
        lib().orc_sim_set_state(self.h, _ptr(uu))

    def state(self):
        uu = np.zeros((self.nvar, self.nx, self.ny))
        lib().orc_sim_get_state(self.h, _ptr(uu))
        return uu

    def derivs(self):
        a = np.zeros((self.nvar, self.nx, self.ny))
        b = np.zeros((self.nvar, self.nx, self.ny))
        lib().orc_sim_get_derivs(self.h, _ptr(a), _ptr(b))
        return a, b

    def run(self, nsteps):
        lib().orc_sim_run(self.h, nsteps)

    def run_fixed_dt(self, dts):
        dts = _c(dts)
        lib().orc_sim_run_fixed_dt(self.h, len(dts), _ptr(dts))

    def stage(self, irk):
        lib().orc_sim_stage(self.h, irk)

    def set_dt(self, dt):
        """dt of the next step (instead of the calcDt value)"""
        lib().orc_sim_set_dt(self.h, dt)

    def glm(self, i=0):
        """(Ch, Cp) in force for x row i (hyperbolic cleaning)."""
        return lib().orc_sim_glm_ch(self.h, i), lib().orc_sim_glm_cp(self.h, i)

    @property
    def dt(self):
        return lib().orc_sim_dt(self.h)

    @property
    def time(self):
        return lib().orc_sim_time(self.h)

    @property
    def steps(self):
        return lib().orc_sim_steps(self.h)
