"""Shared helpers for the parity tests."""
import numpy as np

# per-field relative L2 gate of the FP64 device path against the FP64 oracle
FP64_TOL = 1e-12


def rel_l2(a, b, floor=None):
    """||a-b|| / ||b||; for an identically-zero reference field the absolute
    L2 is divided by `floor` (the rms scale of the state) instead."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    nb = np.linalg.norm(b)
    d = np.linalg.norm(a - b)
    if nb == 0.0:
        return d / (floor if floor else 1.0) / np.sqrt(b.size)
    return d / nb


# A field whose norm is below this fraction of the state's rms norm (zero, or
# zero up to second-order terms, e.g. u_x of a small-amplitude wave) is
# measured in absolute L2 relative to that rms: its relative error would
# only measure how round-off of the O(1) fields leaks into it.
SMALL_FIELD = 1e-6


# The GLM potential psi (field 8 with hyperbolic cleaning) is driven by
# div B, a residual of O(1) derivatives that cancel to ~1e-6 of the state, so
# its own relative error is set by that cancellation: perturbing the input
# state by one ulp moves psi by ~1e-9 relative in the FP64 oracle itself
# (64x32 turbulence, 8 steps).  It is measured like a small field, against
# the rms of the physical fields, and separately (tests/test_gpu_dropin.py)
# relatively at 1e-8.
ABS_FIELDS = (8,)


def field_errors(a, b):
    """per-field relative L2 of two [nv, nx, ny] arrays"""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    scale = np.sqrt(np.mean(b[:8] ** 2)) or 1.0
    out = []
    for v in range(b.shape[0]):
        rms_v = np.sqrt(np.mean(b[v] ** 2))
        if rms_v < SMALL_FIELD * scale or v in ABS_FIELDS:
            out.append(np.sqrt(np.mean((a[v] - b[v]) ** 2)) / scale)
        else:
            out.append(rel_l2(a[v], b[v]))
    return np.array(out)


def assert_fields_close(a, b, tol=FP64_TOL, what=""):
    e = field_errors(a, b)
    assert np.all(e <= tol), f"{what} per-field rel-L2 {e} > {tol}"
    return e


def random_state(nx, ny, seed, positive=True):
    rng = np.random.default_rng(seed)
    uu = rng.uniform(-1.0, 1.0, (8, nx, ny))
    if positive:
        uu[0] = 1.0 + 0.5 * rng.uniform(0, 1, (nx, ny))
        uu[7] = 1.0 + 0.5 * rng.uniform(0, 1, (nx, ny))
    return uu
