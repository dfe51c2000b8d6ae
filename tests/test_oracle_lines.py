"""Pins for the oracle's x-direction line operators (mathlib.c restated).

Each check is chosen so that a plausible slip (a dropped or mis-scaled stencil
term, a wrong sign, a swapped halo index, a wrong matrix entry) fails it:

* Thomas solve vs LAPACK dense solve (brute force);
* polynomial exactness of every row (Taylor matching conditions);
* interior modified wavenumbers vs the published closed forms of Lele (1992)
  and Gaitonde & Visbal (2000), with the coefficients written as the
  published fractions, not via the oracle's formulas;
* the whole periodic line operator vs a dense circulant assembly solved by
  numpy (covers the wrapped halos of rows 0, 1, 2, N-3, N-2, N-1).
"""
import numpy as np
import pytest

import oracle_lib as orc


def test_thomas_matches_dense_solve():
    rng = np.random.default_rng(1)
    for n in (2, 3, 7, 64, 257):
        lo = rng.uniform(-0.4, 0.4, n - 1 if n > 1 else 1)
        up = rng.uniform(-0.4, 0.4, n - 1 if n > 1 else 1)
        dg = 1.0 + rng.uniform(0, 0.2, n)
        r = rng.standard_normal(n)
        M = np.diag(dg) + np.diag(lo[: n - 1], -1) + np.diag(up[: n - 1], 1)
        x = orc.thomas(np.r_[lo, 0], dg, np.r_[up, 0], r)
        np.testing.assert_allclose(x, np.linalg.solve(M, r), rtol=1e-13, atol=1e-13)


def _poly_line(coeffs, N, h):
    """Samples of a polynomial at i*h for i in [-2, N+1] (non-periodic)."""
    idx = np.arange(-2, N + 2) * h
    p = np.polynomial.Polynomial(coeffs)
    v = p(idx)
    return v[2:N + 2], v[:2], v[N + 2:], p, idx[2:N + 2]


@pytest.mark.parametrize("deg", [0, 1, 2, 3, 4])
def test_deriv1_exact_on_quartics(deg):
    # interior rows: 8th-order compact scheme, exact to degree 8;
    # rows 0 / N-1: explicit 4th-order central difference, exact to degree 4.
    N, h = 40, 0.05
    rng = np.random.default_rng(deg)
    A, H0, H1, p, x = _poly_line(rng.standard_normal(deg + 1), N, h)
    X = orc.line_op("d1", A, H0, H1, DX=h)
    np.testing.assert_allclose(X, p.deriv()(x), rtol=0, atol=1e-10)


def test_deriv1_interior_exact_to_degree8_far_from_ends():
    # the degree-5..8 error enters only through rows 0 and N-1 and decays
    # by |m| ~ 0.45 per row inside the tridiagonal solve.
    N, h = 160, 1.0 / 160
    rng = np.random.default_rng(7)
    c = rng.standard_normal(9)
    A, H0, H1, p, x = _poly_line(c, N, h)
    X = orc.line_op("d1", A, H0, H1, DX=h)
    d = p.deriv()(x)
    inner = slice(60, N - 60)
    np.testing.assert_allclose(X[inner], d[inner], rtol=0, atol=1e-9)
    # and the end rows are NOT exact for degree 8 (sanity of the test)
    assert abs(X[0] - d[0]) > 1e-9


@pytest.mark.parametrize("deg", [0, 1, 2, 3, 4, 5])
def test_deriv2_exact_on_quintics(deg):
    N, h = 40, 0.05
    rng = np.random.default_rng(10 + deg)
    A, H0, H1, p, x = _poly_line(rng.standard_normal(deg + 1), N, h)
    X = orc.line_op("d2", A, H0, H1, DX=h)
    np.testing.assert_allclose(X, p.deriv(2)(x), rtol=0, atol=1e-7)


@pytest.mark.parametrize("deg", [0, 1, 2, 3])
def test_filter_preserves_cubics(deg):
    # interior Gaitonde-Visbal 6th-order filter preserves degree <= 5; the
    # alpha=0.95 rows 0 / N-1 preserve degree <= 3.
    N, h = 48, 0.1
    rng = np.random.default_rng(20 + deg)
    A, H0, H1, p, x = _poly_line(rng.standard_normal(deg + 1), N, h)
    X = orc.line_op("f", A, H0, H1)
    np.testing.assert_allclose(X, A, rtol=0, atol=1e-11)


def test_filter_end_rows_not_exact_on_quartic():
    N, h = 48, 0.1
    A, H0, H1, p, x = _poly_line([0, 0, 0, 0, 1.0], N, h)
    X = orc.line_op("f", A, H0, H1)
    assert abs(X[0] - A[0]) > 1e-6


def test_filter_kills_odd_even_mode():
    N = 64
    A = (-1.0) ** np.arange(N)
    H0 = np.array([1.0, -1.0])       # f[-2], f[-1]
    H1 = np.array([1.0, -1.0])       # f[N],  f[N+1]
    X = orc.line_op("f", A, H0, H1)
    np.testing.assert_allclose(X, 0.0, atol=1e-13)


def _periodic_line(f):
    return f, f[-2:].copy(), f[:2].copy()


# published coefficients (fractions), independent of the oracle's formulas
LELE_D1 = dict(alpha=3 / 8, a=25 / 16, b=1 / 5, c=-1 / 80)
LELE_D2 = dict(alpha=3 / 8, a=21 / 32, b=6 / 5, c=-17 / 160)
GV_F6 = dict(alpha=0.49)
GV_F6.update(a0=11 / 16 + 5 * 0.49 / 8, a1=15 / 32 + 17 * 0.49 / 16,
             a2=-3 / 16 + 3 * 0.49 / 8, a3=1 / 32 - 0.49 / 16)


@pytest.mark.parametrize("m", [1, 5, 17, 40, 90])
def test_deriv1_modified_wavenumber(m):
    N = 256
    h = 2 * np.pi / N
    x = np.arange(N) * h
    phi = 0.3
    f = np.sin(m * x + phi)
    X = orc.line_op("d1", *_periodic_line(f), DX=h)
    w = m * h
    L = LELE_D1
    kmod = (L["a"] * np.sin(w) + L["b"] / 2 * np.sin(2 * w) + L["c"] / 3 * np.sin(3 * w)) / (
        1 + 2 * L["alpha"] * np.cos(w)) / h
    inner = slice(80, N - 80)
    np.testing.assert_allclose(X[inner], kmod * np.cos(m * x + phi)[inner], rtol=0, atol=2e-12 * m)
    # rows 0 and N-1 are the explicit 4th-order central difference
    k4 = (8 * np.sin(w) - np.sin(2 * w)) / 6 / h
    np.testing.assert_allclose(X[[0, N - 1]], k4 * np.cos(m * x + phi)[[0, N - 1]], rtol=0,
                               atol=1e-12 * m)


@pytest.mark.parametrize("m", [1, 5, 17, 40, 90])
def test_deriv2_modified_wavenumber(m):
    N = 256
    h = 2 * np.pi / N
    x = np.arange(N) * h
    f = np.cos(m * x - 0.7)
    X = orc.line_op("d2", *_periodic_line(f), DX=h)
    w = m * h
    L = LELE_D2
    k2 = (2 * L["a"] * (1 - np.cos(w)) + L["b"] / 2 * (1 - np.cos(2 * w)) +
          2 * L["c"] / 9 * (1 - np.cos(3 * w))) / (1 + 2 * L["alpha"] * np.cos(w)) / h ** 2
    inner = slice(80, N - 80)
    np.testing.assert_allclose(X[inner], -k2 * f[inner], rtol=0, atol=1e-10 * m * m)
    k24 = (-2 * np.cos(2 * w) / 12 + 2 * 4 / 3 * np.cos(w) - 5 / 2) / h ** 2
    np.testing.assert_allclose(X[[0, N - 1]], k24 * f[[0, N - 1]], rtol=0, atol=1e-9 * m * m)


@pytest.mark.parametrize("m", [0, 3, 30, 64, 100, 127])
def test_filter_transfer_function(m):
    N = 256
    h = 2 * np.pi / N
    x = np.arange(N) * h
    f = np.cos(m * x + 1.1)
    X = orc.line_op("f", *_periodic_line(f))
    w = m * h
    G = GV_F6
    T = (G["a0"] + G["a1"] * np.cos(w) + G["a2"] * np.cos(2 * w) + G["a3"] * np.cos(3 * w)) / (
        1 + 2 * G["alpha"] * np.cos(w))
    # the filter's decay rate per row is |m|=0.817, so stay 120 rows inside
    inner = slice(120, N - 120)
    np.testing.assert_allclose(X[inner], T * f[inner], rtol=0, atol=1e-11)


def _circulant_rows(N, stencil):
    """Dense matrix S with S[i, (i+k) % N] = stencil[k] for all rows."""
    S = np.zeros((N, N))
    for k, c in stencil.items():
        for i in range(N):
            S[i, (i + k) % N] += c
    return S


def _dense_operator(kind, N, h, sizes=None):
    """Assemble the periodic-x operator of the reference as M^-1 S.

    `sizes` splits the line into consecutive blocks, one per MPI rank of the
    reference (decomposeGrid1D, initialize.c:547-569): M becomes block
    diagonal with the end-row couplings inside each block, and the end-row
    stencils sit at every block edge while still reading the global periodic
    neighbours (the halo exchange of parallel.c)."""
    sizes = [N] if sizes is None else list(sizes)
    assert sum(sizes) == N
    if kind == "d1":
        L = LELE_D1
        inner = {1: L["a"] / 2 / h, -1: -L["a"] / 2 / h, 2: L["b"] / 4 / h,
                 -2: -L["b"] / 4 / h, 3: L["c"] / 6 / h, -3: -L["c"] / 6 / h}
        end = {-2: 1 / 12 / h, -1: -8 / 12 / h, 1: 8 / 12 / h, 2: -1 / 12 / h}
        a_in, a_end_lo, a_end_hi = L["alpha"], 0.0, 0.0
    elif kind == "d2":
        L = LELE_D2
        inner = {0: -2 * (L["a"] + L["b"] / 4 + L["c"] / 9) / h ** 2,
                 1: L["a"] / h ** 2, -1: L["a"] / h ** 2, 2: L["b"] / 4 / h ** 2,
                 -2: L["b"] / 4 / h ** 2, 3: L["c"] / 9 / h ** 2, -3: L["c"] / 9 / h ** 2}
        end = {-2: -1 / 12 / h ** 2, -1: 16 / 12 / h ** 2, 0: -30 / 12 / h ** 2,
               1: 16 / 12 / h ** 2, 2: -1 / 12 / h ** 2}
        a_in, a_end_lo, a_end_hi = L["alpha"], 0.0, 0.0
    else:
        G = GV_F6
        inner = {0: G["a0"], 1: G["a1"] / 2, -1: G["a1"] / 2, 2: G["a2"] / 2,
                 -2: G["a2"] / 2, 3: G["a3"] / 2, -3: G["a3"] / 2}
        # 5-point boundary row of the alpha_f=0.95 closure, written from its
        # defining properties: x0 + 0.95 x1 equals a stencil over f[-2..2] that
        # is exact for cubics and annihilates the odd-even mode.
        V = np.array([[1, 1, 1, 1, 1], [-2, -1, 0, 1, 2], [4, 1, 0, 1, 4],
                      [-8, -1, 0, 1, 8], [1, -1, 1, -1, 1]], dtype=float)
        rhs = np.array([1.95, 0.95, 0.95, 0.95, 0.0])
        w = np.linalg.solve(V, rhs)
        end = {k: w[k + 2] for k in range(-2, 3)}
        a_in, a_end_lo, a_end_hi = G["alpha"], 0.95, 0.95
    S = _circulant_rows(N, inner)
    Send = _circulant_rows(N, end)
    # the last row mirrors the first (x_{N-1} + 0.95 x_{N-2} for the filter);
    # derivative end rows are centred stencils, the same at both ends
    Slast = _circulant_rows(N, {-k: c for k, c in end.items()}) if kind == "f" else Send
    M = np.eye(N)
    off = 0
    for n in sizes:
        a, b = off, off + n - 1
        S[a] = Send[a]
        S[b] = Slast[b]
        for i in range(a, b):
            M[i, i + 1] = M[i + 1, i] = a_in
        M[a, a + 1] = a_end_lo
        M[b, b - 1] = a_end_hi
        off += n
    return np.linalg.solve(M, S)


@pytest.mark.parametrize("kind", ["d1", "d2", "f"])
@pytest.mark.parametrize("N", [16, 33, 64])
def test_periodic_line_vs_dense_assembly(kind, N):
    h = 2 * np.pi / N
    rng = np.random.default_rng(N)
    f = rng.standard_normal(N)
    X = orc.line_op(kind, *_periodic_line(f), DX=h)
    D = _dense_operator(kind, N, h)
    scale = np.abs(D @ f).max()
    np.testing.assert_allclose(X, D @ f, rtol=0, atol=1e-13 * max(scale, 1.0))


@pytest.mark.parametrize("kind", ["d1", "d2", "f"])
@pytest.mark.parametrize("nx,npe", [(48, 2), (64, 4), (30, 4), (64, 8)])
def test_block_closure_vs_dense_assembly(kind, nx, npe):
    """x operators split over npe MPI ranks: one block per rank, the first
    nx % npe ranks one row larger (initialize.c:547-569)."""
    h = 2 * np.pi / nx
    sizes = [nx // npe + (1 if r < nx % npe else 0) for r in range(npe)]
    D = _dense_operator(kind, nx, h, sizes)
    rng = np.random.default_rng(nx * 10 + npe)
    arr = rng.standard_normal((2, nx, 3))
    if kind == "d1":
        out = orc.derivX(arr, h, npe=npe)
    elif kind == "d2":
        out = orc.derivXX(arr, h, npe=npe)
    else:
        out = orc.filterX(arr, npe=npe)
    want = np.einsum("ik,vkj->vij", D, arr)
    scale = np.abs(want).max()
    np.testing.assert_allclose(out, want, rtol=0, atol=1e-13 * max(scale, 1.0))
    # the closure is visible: the single-block operator differs at block edges
    one = {"d1": orc.derivX, "d2": orc.derivXX}.get(kind)
    single = one(arr, h) if one else orc.filterX(arr)
    assert np.abs(single - out).max() > 1e-6 * scale
