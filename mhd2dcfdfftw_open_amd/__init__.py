"""MI355X (gfx950) build of the RHS + RK3 time-advance path of
chenshihelio/MHD2DCFDFFTW_Open, behind the C ABI declared in include/mhd2d.h.

This module is a thin ctypes view of that ABI for Python callers (tests,
bench.py).  Everything it runs is the HIP library libmhd2d.so built in-tree
by ``python -m mhd2dcfdfftw_open_amd.build``; there is no fallback path, and a
missing library raises.

Layout of every state array: float64 [8, nx, ny], fields
(rho, ux, uy, uz, Bx, By, Bz, p), x along axis 1, y along axis 2.
"""
import ctypes
import os

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
# MHD2D_LIB selects another build of the same library (tuning variants)
LIB_PATH = os.environ.get("MHD2D_LIB", os.path.join(PKG, "libmhd2d.so"))

FP64, FP32, MIXED = 0, 1, 2
RHO, UX, UY, UZ, BX, BY, BZ, PR = range(8)
FIELDS = ("rho", "ux", "uy", "uz", "Bx", "By", "Bz", "p")

_lib = None
_dp = ctypes.POINTER(ctypes.c_double)


class MHD2DError(RuntimeError):
    pass


class Input(ctypes.Structure):
    """the 15 values of the reference's `input` file (initialize.c:456-525)"""
    _fields_ = [("boundaryTypeX", ctypes.c_int), ("cfl", ctypes.c_double),
                ("Tmax", ctypes.c_double), ("dtOut", ctypes.c_double),
                ("nx", ctypes.c_int), ("ny", ctypes.c_int),
                ("Lx", ctypes.c_double), ("Ly", ctypes.c_double), ("gamma", ctypes.c_double),
                ("hall", ctypes.c_int), ("di", ctypes.c_double),
                ("resistivity", ctypes.c_double), ("viscosity", ctypes.c_double),
                ("divbclean", ctypes.c_int), ("outdivb", ctypes.c_int)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class DriveOpts(ctypes.Structure):
    _fields_ = [("outdir", ctypes.c_char_p), ("mode", ctypes.c_int),
                ("precision", ctypes.c_int), ("verbose", ctypes.c_int),
                ("max_steps", ctypes.c_longlong), ("xblocks", ctypes.c_int)]


class DriveReport(ctypes.Structure):
    _fields_ = [("time", ctypes.c_double), ("dt", ctypes.c_double),
                ("steps", ctypes.c_longlong), ("outputs", ctypes.c_int),
                ("stopped", ctypes.c_int), ("seconds", ctypes.c_double),
                ("energy", ctypes.c_double), ("max_divb", ctypes.c_double)]


DRIVE_AUTO, DRIVE_FUSED, DRIVE_DROPIN = 0, 1, 2
# how an uneven ensemble launches its shape groups (mhd2d_ens_set_launch)
ENS_PER_GROUP, ENS_ONE_LAUNCH = 0, 1


class Params(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_int), ("ny", ctypes.c_int),
                ("Lx", ctypes.c_double), ("Ly", ctypes.c_double),
                ("gamma", ctypes.c_double), ("cfl", ctypes.c_double),
                ("hall", ctypes.c_int), ("di", ctypes.c_double),
                ("resistivity", ctypes.c_double), ("viscosity", ctypes.c_double),
                ("divbclean", ctypes.c_int), ("precision", ctypes.c_int),
                ("xblocks", ctypes.c_int), ("batch", ctypes.c_int), ("xgemm", ctypes.c_int),
                ("open_x", ctypes.c_int)]


def lib():
    """Load libmhd2d.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise MHD2DError(
            f"{LIB_PATH} is missing: build it with `python -m mhd2dcfdfftw_open_amd.build` "
            "(this package has no CPU fallback)")
    L = ctypes.CDLL(LIB_PATH, mode=ctypes.RTLD_GLOBAL)
    i, d, p, vp = ctypes.c_int, ctypes.c_double, _dp, ctypes.c_void_p
    ll = ctypes.c_longlong
    sigs = {
        "mhd2d_create": (i, [ctypes.POINTER(Params), ctypes.POINTER(vp)]),
        "mhd2d_destroy": (None, [vp]),
        "mhd2d_set_state": (i, [vp, p]),
        "mhd2d_get_state": (i, [vp, p]),
        "mhd2d_run": (i, [vp, i]),
        "mhd2d_run_graph": (i, [vp, i]),
        "mhd2d_sync": (i, [vp]),
        "mhd2d_get_scalars": (i, [vp, p, p, ctypes.POINTER(ll)]),
        "mhd2d_diagnostics": (i, [vp, p, p]),
        "mhd2d_time_run": (i, [vp, i, i, p]),
        "mhd2d_profile_run": (i, [vp, i, p, ctypes.POINTER(i)]),
        "mhd2d_device_bytes": (ctypes.c_size_t, [vp]),
        "mhd2d_copy_probe": (i, [ctypes.c_size_t, i, p]),
        "mhd2d_last_error": (ctypes.c_char_p, []),
        "mhd2d_ic_orszag_tang": (None, [i, i, d, d, p]),
        "mhd2d_ic_turbulence": (None, [i, i, d, d, ctypes.c_ulonglong, d, i, i, p]),
        "mhd2d_ic_linear_wave": (None, [i, i, d, d, ctypes.c_ulonglong, d, p]),
        "mhd2d_host_initialize": (i, [i, i, d, d, d, d, i, d, d, d, i]),
        "mhd2d_host_finalize": (None, []),
        "mhd2d_dropin_set_sync": (None, [i]),
        "mhd2d_dropin_sync_host": (i, []),
        "mhd2d_dropin_invalidate_device": (None, []),
        "mhd2d_dropin_release": (None, []),
        "derivX": (None, [i, p, p, p, p, i, i]),
        "derivXX": (None, [i, p, p, p, p, i, i]),
        "derivY": (None, [i, p, p, i]),
        "derivYY": (None, [i, p, p, i]),
        "filtering": (None, []),
        "calcRHS": (None, []),
        "calcDivB": (None, []),
        "calcEfield": (None, []),
        "rkInit": (None, [d]),
        "rkExecute": (None, [i]),
        "calcDt": (d, [d]),
        "sendrecvHaloPointsX": (None, []),
        "sendrecvHaloPointsdivBX": (None, []),
        "sendrecvHaloPointsEfieldX": (None, []),
        "calc_Cp_fix_Cr": (None, []),
        "mhd2d_get_divb": (i, [vp, p]),
        "mhd2d_stopped": (i, [vp, p]),
        "mhd2d_set_state_at": (i, [vp, i, p]),
        "mhd2d_get_state_at": (i, [vp, i, p]),
        "mhd2d_set_di": (i, [vp, i, d]),
        "mhd2d_set_subbatch": (i, [vp, i]),
        "mhd2d_get_subbatch": (i, [vp]),
        "mhd2d_set_xfuse": (i, [vp, i]),
        "mhd2d_get_xfuse": (i, [vp]),
        "mhd2d_get_scalars_at": (i, [vp, i, p, p, ctypes.POINTER(ll)]),
        "mhd2d_diagnostics_at": (i, [vp, i, p, p]),
        "mhd2d_stopped_at": (i, [vp, i, p]),
        "mhd2d_xderiv": (i, [vp, i, p, p]),
        "mhd2d_xgemm_panels": (ctypes.c_longlong, [vp]),
        "mhd2d_autotune": (i, [vp, i, ctypes.POINTER(i)]),
        "mhd2d_ens_autotune": (i, [vp, i, ctypes.POINTER(i)]),
        "mhd2d_ens_create": (i, [ctypes.POINTER(Params), i, ctypes.POINTER(vp)]),
        "mhd2d_ens_destroy": (None, [vp]),
        "mhd2d_ens_groups": (i, [vp]),
        "mhd2d_ens_set_state": (i, [vp, i, p]),
        "mhd2d_ens_get_state": (i, [vp, i, p]),
        "mhd2d_ens_run": (i, [vp, i, i]),
        "mhd2d_ens_sync": (i, [vp]),
        "mhd2d_ens_get_scalars": (i, [vp, i, p, p, ctypes.POINTER(ll)]),
        "mhd2d_ens_diagnostics": (i, [vp, i, p, p]),
        "mhd2d_ens_time_run": (i, [vp, i, i, p]),
        "mhd2d_ens_set_launch": (i, [vp, i]),
        "mhd2d_ens_get_launch": (i, [vp]),
        "mhd2d_ens_launches_per_step": (i, [vp]),
        "mhd2d_run_noalloc": (i, [ctypes.POINTER(Params), vp, i, vp, p, p]),
        "mhd2d_read_input": (i, [ctypes.c_char_p, ctypes.POINTER(Input)]),
        "mhd2d_write_input": (i, [ctypes.c_char_p, ctypes.POINTER(Input)]),
        "mhd2d_write_grid": (i, [ctypes.c_char_p, i, i, d, d]),
        "mhd2d_write_grid_bc": (i, [ctypes.c_char_p, i, i, d, d, i]),
        "mhd2d_write_array": (i, [ctypes.c_char_p, i, d, p, i, i, i]),
        "mhd2d_write_divb": (i, [ctypes.c_char_p, i, d, p, i, i]),
        "mhd2d_read_array": (i, [ctypes.c_char_p, p, p, ctypes.c_size_t]),
        "mhd2d_write_log": (i, [ctypes.c_char_p, d, d, d, ll]),
        "mhd2d_drive": (i, [ctypes.POINTER(Input), p, ctypes.POINTER(DriveOpts),
                            ctypes.POINTER(DriveReport)]),
    }
    for name, (res, args) in sigs.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def _ptr(a):
    return a.ctypes.data_as(_dp)


def _check(rc):
    if rc != 0:
        raise MHD2DError(f"mhd2d error {rc}: {lib().mhd2d_last_error().decode()}")


def copy_probe(nbytes=1 << 30, iters=20):
    """Achievable streaming-copy rate in GB/s (read + write bytes)."""
    g = ctypes.c_double()
    _check(lib().mhd2d_copy_probe(nbytes, iters, ctypes.byref(g)))
    return g.value


# ------------------------------------------------------------------ initial states
def orszag_tang(nx, ny, Lx=2 * np.pi, Ly=2 * np.pi):
    uu = np.zeros((8, nx, ny))
    lib().mhd2d_ic_orszag_tang(nx, ny, Lx, Ly, _ptr(uu))
    return uu


def turbulence(nx, ny, seed=42, rms=0.1, kmin=1, kmax=16, Lx=2 * np.pi, Ly=2 * np.pi):
    uu = np.zeros((8, nx, ny))
    lib().mhd2d_ic_turbulence(nx, ny, Lx, Ly, seed, rms, kmin, kmax, _ptr(uu))
    return uu


def linear_wave(nx, ny, seed=42, amp=1e-4, Lx=2 * np.pi, Ly=2 * np.pi):
    uu = np.zeros((8, nx, ny))
    lib().mhd2d_ic_linear_wave(nx, ny, Lx, Ly, seed, amp, _ptr(uu))
    return uu


# ------------------------------------------------------------------ run configuration / files
def read_input(path):
    inp = Input()
    _check(lib().mhd2d_read_input(os.fsencode(path), ctypes.byref(inp)))
    return inp


def write_input(path, **kw):
    inp = Input(**kw)
    _check(lib().mhd2d_write_input(os.fsencode(path), ctypes.byref(inp)))
    return inp


def drive(inp, uu0, outdir="", mode=DRIVE_AUTO, precision=FP64, verbose=False, max_steps=0,
          npe=1):
    """The time-driven loop of mhd.c on the GPU; returns the report dict.
    npe > 1 reproduces a run on npe MPI ranks (x split into npe blocks)."""
    uu0 = np.ascontiguousarray(uu0, dtype=np.float64)
    assert uu0.shape == (8, inp.nx, inp.ny)
    o = DriveOpts(os.fsencode(outdir), mode, precision, int(verbose), max_steps, npe)
    rep = DriveReport()
    _check(lib().mhd2d_write_grid_bc(os.fsencode(outdir), inp.nx, inp.ny, inp.Lx, inp.Ly,
                                     int(inp.boundaryTypeX == 1)))
    _check(lib().mhd2d_drive(ctypes.byref(inp), _ptr(uu0), ctypes.byref(o), ctypes.byref(rep)))
    return {k: getattr(rep, k) for k, _ in rep._fields_}


# ------------------------------------------------------------------ context API
class Context:
    """Fused device stepping of one grid (mhd2d_create ... mhd2d_destroy), or
    of `batch` independent runs of that grid advanced together."""

    def __init__(self, nx, ny, Lx=2 * np.pi, Ly=2 * np.pi, gamma=5.0 / 3.0, cfl=0.3,
                 hall=0, di=0.0, precision=FP64, npe=1, batch=1, xgemm=0, open_x=False,
                 resistivity=0.0, viscosity=0.0, divbclean=0):
        """npe > 1 splits x into npe blocks as a reference run on npe MPI
        ranks does (decomposeGrid1D, initialize.c:547-569); xgemm 3 or 6
        takes x-derivatives as split-BF16 MFMA GEMMs with the dense A^-1 B;
        resistivity and viscosity add the reference's addSourceTerm
        (calcRHS.c:498-531); divbclean = 1 the diffusive div B cleaning,
        divbclean = 2 the GLM cleaning, whose states carry psi as a ninth
        field [9][nx][ny]."""
        self.nx, self.ny, self.batch = nx, ny, batch
        self.nvar = 9 if divbclean == 2 else 8
        self.params = Params(nx, ny, Lx, Ly, gamma, cfl, hall, di, resistivity, viscosity,
                             divbclean, precision, npe, batch, xgemm, int(open_x))
        h = ctypes.c_void_p()
        _check(lib().mhd2d_create(ctypes.byref(self.params), ctypes.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            lib().mhd2d_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def set_state(self, uu, inst=0):
        uu = np.ascontiguousarray(uu, dtype=np.float64)
        assert uu.shape == (self.nvar, self.nx, self.ny)
        _check(lib().mhd2d_set_state_at(self.h, inst, _ptr(uu)))

    def state(self, inst=0):
        uu = np.zeros((self.nvar, self.nx, self.ny))
        _check(lib().mhd2d_get_state_at(self.h, inst, _ptr(uu)))
        return uu

    def xderiv(self, fields):
        """d/dx of host fields [nv][nx][ny] with this context's x-derivative"""
        f = np.ascontiguousarray(fields, dtype=np.float64)
        assert f.shape[1:] == (self.nx, self.ny)
        out = np.zeros_like(f)
        _check(lib().mhd2d_xderiv(self.h, f.shape[0], _ptr(f), _ptr(out)))
        return out

    def xgemm_panels(self):
        return lib().mhd2d_xgemm_panels(self.h)

    def autotune(self, reps=5):
        """time every x-solve launch shape as whole steps of this run (after
        set_state; the run is restored) and keep the fastest; returns
        {'deriv': (width, xcd_order), 'filter': (width, xcd_order)}"""
        c = (ctypes.c_int * 4)()
        _check(lib().mhd2d_autotune(self.h, reps, c))
        return {"deriv": (c[0], c[1]), "filter": (c[2], c[3])}

    def set_di(self, inst, di):
        _check(lib().mhd2d_set_di(self.h, inst, di))

    def set_subbatch(self, k):
        """run the batch as groups of k instances per launch (0: all at once)"""
        _check(lib().mhd2d_set_subbatch(self.h, k))

    @property
    def subbatch(self):
        """instances per launch in use"""
        r = lib().mhd2d_get_subbatch(self.h)
        if r < 0:
            _check(r)
        return r

    def set_xfuse(self, on):
        """x-derivative + RHS as the carry solve and the fused chunk kernel
        (ideal MHD on one periodic x block; bitwise the same results)"""
        _check(lib().mhd2d_set_xfuse(self.h, int(on)))

    @property
    def xfuse(self):
        """1 when the fused x-derivative + RHS path is in use"""
        r = lib().mhd2d_get_xfuse(self.h)
        if r < 0:
            _check(r)
        return r

    def run(self, nsteps, graph=False):
        f = lib().mhd2d_run_graph if graph else lib().mhd2d_run
        _check(f(self.h, nsteps))

    def sync(self):
        _check(lib().mhd2d_sync(self.h))

    def scalars(self, inst=0):
        dt, t, n = ctypes.c_double(), ctypes.c_double(), ctypes.c_longlong()
        _check(lib().mhd2d_get_scalars_at(self.h, inst, _ptr1(dt), _ptr1(t), ctypes.byref(n)))
        return dt.value, t.value, n.value

    def diagnostics(self, inst=0):
        e, m = ctypes.c_double(), ctypes.c_double()
        _check(lib().mhd2d_diagnostics_at(self.h, inst, _ptr1(e), _ptr1(m)))
        return e.value, m.value

    def time_run(self, nsteps, graph=False):
        ms = ctypes.c_double()
        _check(lib().mhd2d_time_run(self.h, nsteps, int(graph), ctypes.byref(ms)))
        return ms.value

    def profile_run(self, nsteps):
        out = (ctypes.c_double * 5)()
        cnt = (ctypes.c_int * 5)()
        _check(lib().mhd2d_profile_run(self.h, nsteps, out, cnt))
        # with the fused x-derivative + RHS the second and third classes are
        # the carry solve and the fused chunk kernel
        names = ("ysweep", "xcarry" if self.xfuse else "xderiv",
                 "xrhs_rk" if self.xfuse else "rhs_rk", "xfilter", "dt_finalize")
        return {k: (out[q], cnt[q]) for q, k in enumerate(names)}

    def device_bytes(self):
        return lib().mhd2d_device_bytes(self.h)

    def divb(self):
        out = np.zeros((self.nx, self.ny))
        _check(lib().mhd2d_get_divb(self.h, _ptr(out)))
        return out

    def stopped(self, inst=0):
        d = ctypes.c_double()
        r = lib().mhd2d_stopped_at(self.h, inst, _ptr1(d))
        if r > 1:
            _check(r)
        return bool(r), d.value


class Ensemble:
    """Runs on grids of different sizes advanced together
    (mhd2d_ens_create ... mhd2d_ens_destroy).  `runs` is a list of dicts with
    nx, ny and optionally Lx, Ly, gamma, cfl, hall, di, npe, resistivity,
    viscosity, divbclean."""

    def __init__(self, runs, precision=FP64):
        arr = (Params * len(runs))()
        for k, r in enumerate(runs):
            arr[k] = Params(r["nx"], r["ny"], r.get("Lx", 2 * np.pi), r.get("Ly", 2 * np.pi),
                            r.get("gamma", 5.0 / 3.0), r.get("cfl", 0.3), r.get("hall", 0),
                            r.get("di", 0.0), r.get("resistivity", 0.0),
                            r.get("viscosity", 0.0), r.get("divbclean", 0), precision,
                            r.get("npe", 1), 0,
                            r.get("xgemm", 0), int(r.get("open_x", False)))
        self.shapes = [(r["nx"], r["ny"]) for r in runs]
        self.nvars = [9 if r.get("divbclean", 0) == 2 else 8 for r in runs]
        h = ctypes.c_void_p()
        _check(lib().mhd2d_ens_create(arr, len(runs), ctypes.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            lib().mhd2d_ens_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    @property
    def groups(self):
        return lib().mhd2d_ens_groups(self.h)

    def set_state(self, inst, uu):
        uu = np.ascontiguousarray(uu, dtype=np.float64)
        assert uu.shape == (self.nvars[inst],) + self.shapes[inst]
        _check(lib().mhd2d_ens_set_state(self.h, inst, _ptr(uu)))

    def state(self, inst):
        uu = np.zeros((self.nvars[inst],) + self.shapes[inst])
        _check(lib().mhd2d_ens_get_state(self.h, inst, _ptr(uu)))
        return uu

    def run(self, nsteps, graph=False):
        _check(lib().mhd2d_ens_run(self.h, nsteps, int(graph)))

    def autotune(self, reps=5):
        """Context.autotune for every group; one dict per group"""
        g = lib().mhd2d_ens_groups(self.h)
        c = (ctypes.c_int * (4 * g))()
        _check(lib().mhd2d_ens_autotune(self.h, reps, c))
        return [{"deriv": (c[4 * k], c[4 * k + 1]), "filter": (c[4 * k + 2], c[4 * k + 3])}
                for k in range(g)]

    def sync(self):
        _check(lib().mhd2d_ens_sync(self.h))

    def scalars(self, inst):
        dt, t, n = ctypes.c_double(), ctypes.c_double(), ctypes.c_longlong()
        _check(lib().mhd2d_ens_get_scalars(self.h, inst, _ptr1(dt), _ptr1(t), ctypes.byref(n)))
        return dt.value, t.value, n.value

    def diagnostics(self, inst):
        e, m = ctypes.c_double(), ctypes.c_double()
        _check(lib().mhd2d_ens_diagnostics(self.h, inst, _ptr1(e), _ptr1(m)))
        return e.value, m.value

    def time_run(self, nsteps, graph=False):
        ms = ctypes.c_double()
        _check(lib().mhd2d_ens_time_run(self.h, nsteps, int(graph), _ptr1(ms)))
        return ms.value

    def set_launch(self, mode):
        """ENS_ONE_LAUNCH (one launch per kernel class and stage over every
        group) or ENS_PER_GROUP (each group's own launches)"""
        _check(lib().mhd2d_ens_set_launch(self.h, mode))

    @property
    def launch_mode(self):
        return lib().mhd2d_ens_get_launch(self.h)

    @property
    def launches_per_step(self):
        return lib().mhd2d_ens_launches_per_step(self.h)


def _ptr1(c_double):
    return ctypes.cast(ctypes.byref(c_double), _dp)


# ------------------------------------------------------------------ drop-in layer
_GLOBAL_ARRAYS = {  # name -> number of fields (initialize.c:216-294)
    "uu": "NVAR", "dudx": "NVAR", "dudy": "NVAR", "dudt": "NVAR", "dudtRK": "NVAR",
    "Efield": 3, "dEdx": 3, "dEdy": 3, "d2udx2": 3, "d2udy2": 3, "d2Bdx2": 3,
    "d2Bdy2": 3, "divB": 1, "ddivBdx": 1, "ddivBdy": 1,
}


class RefGlobals:
    """The reference's host globals as defined in libmhd2d.so's own storage
    (mhd2d_host_initialize), with numpy views of its arrays, so the drop-in
    entry points can be called exactly as mhd.c calls them."""

    def __init__(self, nx, ny, Lx=2 * np.pi, Ly=2 * np.pi, gamma=5.0 / 3.0, cfl=0.3,
                 hall=0, di=0.0, resistivity=0.0, viscosity=0.0, divbclean=0):
        L = lib()
        _check(L.mhd2d_host_initialize(nx, ny, Lx, Ly, gamma, cfl, hall, di, resistivity,
                                       viscosity, divbclean))
        self.nx, self.ny = nx, ny
        self.nvar = ctypes.c_int.in_dll(L, "NVAR").value

    def ptr(self, name):
        return ctypes.cast(ctypes.c_void_p.in_dll(lib(), name).value, _dp)

    def array(self, name):
        nf = _GLOBAL_ARRAYS[name]
        nf = self.nvar if nf == "NVAR" else nf
        n = nf * self.nx * self.ny
        buf = np.ctypeslib.as_array(self.ptr(name), shape=(n,))
        return buf.reshape(nf, self.nx, self.ny)

    def scalar(self, name, ctype=ctypes.c_double):
        return ctype.in_dll(lib(), name)

    def close(self):
        lib().mhd2d_host_finalize()
