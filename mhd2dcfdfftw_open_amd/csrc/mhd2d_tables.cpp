// One-time FP64 construction of the operator tables (host), uploaded once.
#include "mhd2d_tables.h"

#include <cmath>
#include <cstring>
#include <vector>

namespace mhd2d {

GridTables::~GridTables() { release(); }

void GridTables::release()
{
    if (blob_) (void)hipFree(blob_);
    blob_ = nullptr;
}

// Thomas factorisation (no pivoting) of the unit-diagonal tridiagonal matrix
// of the periodic branch: rows 1..n-2 carry alpha on both sides, rows 0 and
// n-1 carry alphaSub on their single coupling (initialize.c:633-654 and its
// 2nd-derivative / filter twins).  Same elimination order as the reference's
// solver (mathlib.c:17-22), done once instead of per line.
void GridTables::factor(int n, double alpha, double alphaSub, double* l, double* invd,
                        double* q)
{
    std::vector<double> lo(n), up(n);
    for (int i = 0; i < n - 2; i++) lo[i] = alpha;  // coefficient of x[i] in row i+1
    lo[n - 2] = alphaSub;
    up[0] = alphaSub;                               // coefficient of x[i+1] in row i
    for (int i = 1; i < n - 1; i++) up[i] = alpha;
    factor_rows(n, lo.data(), up.data(), l, invd, q);
}

void GridTables::factor_rows(int n, const double* lo, const double* up, double* l,
                             double* invd, double* q)
{
    std::vector<double> b(n);
    b[0] = 1.0;
    l[0] = 0.0;
    for (int i = 1; i < n; i++) {
        const double m = lo[i - 1] / b[i - 1];
        b[i] = 1.0 - m * up[i - 1];
        l[i] = m;
    }
    for (int i = 0; i < n; i++) invd[i] = 1.0 / b[i];
    for (int i = 0; i < n - 1; i++) q[i] = -up[i] / b[i];
    q[n - 1] = 0.0;
}

void GridTables::block_range(int n, int nb, int b, int* off, int* size)
{
    const int base = n / nb, rem = n % nb;
    *off = b * base + (b < rem ? b : rem);
    *size = base + (b < rem ? 1 : 0);
}

namespace {

// open-boundary scheme constants: boundary alphas (parameter.h:25-37), the
// near-boundary P1 / P2 schemes (parameter.h:39-56)
constexpr double ALPHA_B1 = 3.0, ALPHA_B2 = 5.0;
constexpr double ALPHA_F0 = 1.05, ALPHA_F1 = 0.49, ALPHA_F2 = 0.49;
constexpr double ALPHA1P1 = 1. / 4., A1P1 = 1.5, ALPHA1P2 = 1. / 3., A1P2 = 14. / 9., B1P2 = 1. / 9.;
constexpr double ALPHA2P1 = 0.1, A2P1 = 1.2, ALPHA2P2 = 2. / 11., A2P2 = 12. / 11., B2P2 = 3. / 11.;

// stencil over f[i-4 .. i+4]
struct Row9 {
    double c[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    double& at(int m) { return c[m + 4]; }
};

// right-hand side of row `pos` of a block whose end `side` (0 first row,
// 1 last row) is the ordinary closure (mathlib.c:50-52, :80-82 and twins)
Row9 end_row(const XStencil& st, int side)
{
    Row9 r;
    for (int k = 0; k < 5; k++) {
        if (side == 0) r.at(k - 2) = st.e[k];
        else r.at(2 - k) = st.esign * st.e[k];
    }
    return r;
}

// rows 0, 1, 2 (side 0) or N-1, N-2, N-3 (side 1, index 0, 1, 2) of an open
// end: mathlib.c:89-96, :124-..., :220-228, :357-370 and their mirrors
Row9 open_row(int kind, int side, int idx, double idx1, double idx2)
{
    Row9 r;
    const int s = side == 0 ? 1 : -1;  // mirror: offset m -> -m
    if (kind == XD1) {
        const double ab = ALPHA_B1;
        const double aB = -(11. + 2. * ab) / 6., bB = (6. - ab) / 2., cB = (2. * ab - 3.) / 2.,
                     dB = (2. - ab) / 6.;
        const double sg = side == 0 ? 1.0 : -1.0;  // d/dx changes sign under the mirror
        if (idx == 0) {
            r.at(0) = sg * aB * idx1;
            r.at(s * 1) = sg * bB * idx1;
            r.at(s * 2) = sg * cB * idx1;
            r.at(s * 3) = sg * dB * idx1;
        } else if (idx == 1) {
            r.at(1) = A1P1 / 2 * idx1;
            r.at(-1) = -A1P1 / 2 * idx1;
        } else {
            r.at(1) = A1P2 / 2 * idx1;
            r.at(-1) = -A1P2 / 2 * idx1;
            r.at(2) = B1P2 / 4 * idx1;
            r.at(-2) = -B1P2 / 4 * idx1;
        }
    } else if (kind == XD2) {
        const double ab = ALPHA_B2;
        const double e[5] = {(11. * ab + 35.) / 12., -(5. * ab + 26.) / 3., (ab + 19.) / 2.,
                             (ab - 14.) / 3., (11. - ab) / 12.};
        if (idx == 0) {
            for (int k = 0; k < 5; k++) r.at(s * k) = e[k] * idx2;
        } else if (idx == 1) {
            r.at(1) = A2P1 * idx2;
            r.at(-1) = A2P1 * idx2;
            r.at(0) = -2 * A2P1 * idx2;
        } else {
            r.at(1) = A2P2 * idx2;
            r.at(-1) = A2P2 * idx2;
            r.at(2) = B2P2 / 4 * idx2;
            r.at(-2) = B2P2 / 4 * idx2;
            r.at(0) = (-2 * A2P2 - 2 * B2P2 / 4) * idx2;
        }
    } else {
        const double a0 = ALPHA_F0, a1 = ALPHA_F1, a2 = ALPHA_F2;
        if (idx == 0) {
            const double e[5] = {(15. + a0) / 16., (1. + 3. * a0) / 4., 3. * (a0 - 1) / 8.,
                                 (1. - a0) / 4., (a0 - 1.) / 16.};
            for (int k = 0; k < 5; k++) r.at(s * k) = e[k];
        } else if (idx == 1) {
            // over f[0 .. 4] of the line, row 1 (offsets -1 .. 3)
            const double e[5] = {(1. + 14. * a1) / 16., (3. + 2. * a1) / 4., (3. + 2. * a1) / 8.,
                                 (-1. + 2. * a1) / 4., (1. - 2. * a1) / 16.};
            for (int k = 0; k < 5; k++) r.at(s * (k - 1)) = e[k];
        } else {
            r.at(0) = (5. + 6. * a2) / 8.;
            r.at(1) = r.at(-1) = (1. + 2. * a2) / 2. / 2;
            r.at(2) = r.at(-2) = -(1. - 2. * a2) / 8. / 2;
        }
    }
    return r;
}

// couplings of one block: lo[i] of x[i] in row i+1, up[i] of x[i+1] in row
// i; role 0 ordinary, 1 open first block, 2 open last block
// (initialize.c:633-861)
void block_rows(int kind, int role, int n, const SchemeAlphas& al, double* lo, double* up)
{
    double ai, as, ab, ap1, ap2;
    if (kind == XD1) { ai = al.alpha1; as = al.alphaS1; ab = ALPHA_B1; ap1 = ALPHA1P1; ap2 = ALPHA1P2; }
    else if (kind == XD2) { ai = al.alpha2; as = al.alphaS2; ab = ALPHA_B2; ap1 = ALPHA2P1; ap2 = ALPHA2P2; }
    else { ai = al.alphaF; as = al.alphaSF; ab = ALPHA_F0; ap1 = ALPHA_F1; ap2 = ALPHA_F2; }
    for (int i = 0; i < n - 2; i++) lo[i] = ai;
    lo[n - 2] = as;
    up[0] = as;
    for (int i = 1; i < n - 1; i++) up[i] = ai;
    up[n - 1] = lo[n - 1] = 0.0;
    if (role == 1) {
        up[0] = ab;
        lo[0] = up[1] = ap1;
        lo[1] = up[2] = ap2;
    } else if (role == 2) {
        lo[n - 2] = ab;
        lo[n - 3] = up[n - 2] = ap1;
        lo[n - 4] = up[n - 3] = ap2;
    }
}

}  // namespace

hipError_t GridTables::build(int nx_, int ny_, double Lx_, double Ly_, int elem_bytes_,
                             const SchemeAlphas& al, int xblocks_, int open_x_, void* place,
                             size_t place_bytes)
{
    release();
    xblocks = xblocks_ < 1 ? 1 : xblocks_;
    open_x = open_x_ ? 1 : 0;
    if (open_x && xblocks < 2) return hipErrorInvalidValue;
    nx = nx_;
    ny = ny_;
    Lx = Lx_;
    Ly = Ly_;
    elem_bytes = elem_bytes_;
    dx = open_x ? Lx / (nx - 1) : Lx / nx;  // initialize.c:125-133
    dy = Ly / ny;  // initialize.c:135
    xsolve_shape(nx, elem_bytes, &chunk, &width, &groups);

    // ---- stencil weights, formulas of initialize.c:574-630
    const double a1 = al.alpha1, s1 = al.alphaS1;
    const double aI1 = (a1 + 9.) / 6., bI1 = (32. * a1 - 9.) / 15., cI1 = (-3. * a1 + 1.) / 10.;
    const double e1[5] = {(1. - s1) / 12., (3. * s1 - 4.) / 6., -1.5 * s1, (4. + 5. * s1) / 6.,
                          (3. * s1 - 1.) / 12.};
    const double a2 = al.alpha2, s2 = al.alphaS2;
    const double aI2 = (6. - 9. * a2) / 4., bI2 = (-3. + 24. * a2) / 5., cI2 = (2. - 11. * a2) / 20.;
    const double e2[5] = {-(1. + s2) / 12., (s2 + 4.) / 3., (s2 - 5.) / 2., (4. - 5. * s2) / 3.,
                          (11. * s2 - 1.) / 12.};
    const double aF = al.alphaF, sF = al.alphaSF;
    const double aIF = (11. + 10. * aF) / 16., bIF = (15. + 34. * aF) / 32.,
                 cIF = (-3. + 6. * aF) / 16., dIF = (1. - 2. * aF) / 32.;
    const double eF[5] = {(sF - 1.) / 16., (1. - sF) / 4., (3. * sF + 5.) / 8., (3. * sF + 1.) / 4.,
                          (sF - 1.) / 16.};

    const double idx = 1.0 / dx, idx2 = 1.0 / (dx * dx);
    // first derivative: mathlib.c:50-82 (row N-1 carries a leading minus)
    st[XD1] = XStencil{0.0, aI1 / 2 * idx, bI1 / 4 * idx, cI1 / 6 * idx,
                       {e1[0] * idx, e1[1] * idx, e1[2] * idx, e1[3] * idx, e1[4] * idx}, -1.0};
    // second derivative: mathlib.c:182-214
    st[XD2] = XStencil{0.0, aI2 * idx2, bI2 / 4 * idx2, cI2 / 9 * idx2,
                       {e2[0] * idx2, e2[1] * idx2, e2[2] * idx2, e2[3] * idx2, e2[4] * idx2}, 1.0};
    // filter: mathlib.c:313-350
    st[XFILT] = XStencil{aIF, bIF / 2, cIF / 2, dIF / 2, {eF[0], eF[1], eF[2], eF[3], eF[4]}, 1.0};

    // ---- host tables
    const int G = groups, C = chunk, N = ny;
    const size_t nxs = (size_t)nx, gs = (size_t)G;
    const size_t x_doubles = NXKIND * (3 * nxs + 2 * gs);
    const size_t y_doubles = 5 * (size_t)N + 2 * (size_t)N;  // md, mdd, mf, mfd, mfdd + tw
    // rows with an explicit stencil: the edges of every block, and the three
    // rows of each open end; ascending, the same rows for every kind
    std::vector<int> orows;
    std::vector<Row9> ocoef[NXKIND];
    if (xblocks > 1) {
        const double idx1 = 1.0 / dx, idx2 = 1.0 / (dx * dx);
        for (int b = 0; b < xblocks; b++) {
            int bo, bn;
            block_range(nx, xblocks, b, &bo, &bn);
            const bool open_first = open_x && b == 0, open_last = open_x && b == xblocks - 1;
            if (open_first) {
                for (int r = 0; r < 3; r++) {
                    orows.push_back(bo + r);
                    for (int k = 0; k < NXKIND; k++) ocoef[k].push_back(open_row(k, 0, r, idx1, idx2));
                }
            } else {
                orows.push_back(bo);
                for (int k = 0; k < NXKIND; k++) ocoef[k].push_back(end_row(st[k], 0));
            }
            if (open_last) {
                for (int r = 2; r >= 0; r--) {
                    orows.push_back(bo + bn - 1 - r);
                    for (int k = 0; k < NXKIND; k++) ocoef[k].push_back(open_row(k, 1, r, idx1, idx2));
                }
            } else {
                orows.push_back(bo + bn - 1);
                for (int k = 0; k < NXKIND; k++) ocoef[k].push_back(end_row(st[k], 1));
            }
        }
    }
    const size_t nov = orows.size();
    const size_t m_words = xblocks > 1 ? gs + (gs + 1) / 2 + NXKIND * 9 * nov : 0;
    std::vector<double> h(x_doubles + y_doubles + m_words);
    double* p = h.data();
    const double alph[NXKIND][2] = {{al.alpha1, al.alphaS1}, {al.alpha2, al.alphaS2},
                                    {al.alphaF, al.alphaSF}};
    size_t off_l[NXKIND], off_id[NXKIND], off_q[NXKIND], off_pf[NXKIND], off_qb[NXKIND];
    size_t o = 0;
    for (int k = 0; k < NXKIND; k++) {
        off_l[k] = o;  o += nxs;
        off_id[k] = o; o += nxs;
        off_q[k] = o;  o += nxs;
        off_pf[k] = o; o += gs;
        off_qb[k] = o; o += gs;
        // one factorisation per block: l = 0 on a block's first row and q = 0
        // on its last, so both recurrences (and the chunk products) restart
        // at every block edge by themselves
        (void)alph;
        for (int b = 0; b < xblocks; b++) {
            int bo, bn;
            block_range(nx, xblocks, b, &bo, &bn);
            const int role = !open_x ? 0 : (b == 0 ? 1 : (b == xblocks - 1 ? 2 : 0));
            std::vector<double> lo(bn), up(bn);
            block_rows(k, role, bn, al, lo.data(), up.data());
            factor_rows(bn, lo.data(), up.data(), p + off_l[k] + bo, p + off_id[k] + bo,
                        p + off_q[k] + bo);
        }
        for (int g = 0; g < G; g++) {
            double pf = 1.0, qb = 1.0;
            for (int i = g * C; i < (g + 1) * C; i++) {
                pf *= -p[off_l[k] + i];
                qb *= p[off_q[k] + i];
            }
            p[off_pf[k] + g] = pf;
            p[off_qb[k] + g] = qb;
        }
    }
    // ---- y multipliers and twiddles (long double trig, rounded once)
    const size_t off_md = o, off_mdd = o + N, off_mf = o + 2 * N, off_mfd = o + 3 * N;
    const size_t off_mfdd = o + 4 * N;
    const size_t off_tw = o + 5 * N;
    const long double PI = 3.141592653589793238462643383279502884L;
    const long double afy = al.afy;
    const long double aj = (5.0L + 6.0L * afy) / 8.0L, bj = (1.0L + 2.0L * afy) / 2.0L,
                      cj = -(1.0L - 2.0L * afy) / 8.0L;  // initialize.c:166-168
    for (int k = 0; k < N; k++) {
        const int ka = (k <= N / 2) ? k : N - k;          // |k|
        const int ks = (k < N / 2) ? k : (k == N / 2 ? 0 : k - N);  // signed, Nyquist 0
        const long double kap = 2.0L * PI / (long double)Ly * ks;
        const long double kapa = 2.0L * PI / (long double)Ly * ka;
        const long double w = 2.0L * PI * ka / N;
        const long double fy = (aj + bj * cosl(w) + cj * cosl(2 * w)) / (1 + 2 * afy * cosl(w));
        p[off_md + k] = (double)(kap / N);
        p[off_mdd + k] = (double)(-kapa * kapa / N);
        p[off_mf + k] = (double)(fy / N);
        p[off_mfd + k] = (double)(fy * kap / N);
        p[off_mfdd + k] = (double)(-fy * kapa * kapa / N);
        const long double ang = 2.0L * PI * k / N;
        p[off_tw + 2 * k] = (double)cosl(ang);
        p[off_tw + 2 * k + 1] = (double)(-sinl(ang));
    }
    const size_t off_om = off_tw + 2 * (size_t)N, off_of = off_om + gs,
                 off_oc = off_of + (gs + 1) / 2;
    if (xblocks > 1) {
        std::vector<unsigned long long> om(G, 0ull);
        std::vector<int> of(G + (G & 1), 0);
        for (int g = G - 1; g >= 0; g--) of[g] = (int)nov;
        for (size_t r = 0; r < nov; r++) {
            const int row = orows[r], g = row / C;
            om[g] |= 1ull << (row % C);
            if (of[g] > (int)r) of[g] = (int)r;
        }
        std::memcpy(p + off_om, om.data(), gs * sizeof(double));
        std::memcpy(p + off_of, of.data(), of.size() * sizeof(int));
        for (int k = 0; k < NXKIND; k++)
            for (size_t r = 0; r < nov; r++)
                for (int m = 0; m < 9; m++) p[off_oc + (k * nov + r) * 9 + m] = ocoef[k][r].c[m];
    }
    // the tables go to caller-given device memory when `place` is set (no
    // allocation; the caller keeps it alive), else to one allocation
    void* dst = place;
    if (place) {
        if (h.size() * sizeof(double) > place_bytes) return hipErrorInvalidValue;
    } else {
        hipError_t e = hipMalloc(&blob_, h.size() * sizeof(double));
        if (e != hipSuccess) return e;
        dst = blob_;
    }
    hipError_t e = hipMemcpy(dst, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice);
    if (e != hipSuccess) return e;
    const double* d = (const double*)dst;
    for (int k = 0; k < NXKIND; k++) {
        // first row from which l, invd and q stay bitwise equal to their
        // values at row nx-3 up to that row (the Thomas factors' FP64 fixed
        // point: row 25 for the derivatives, 87 for the filter); the kernels
        // take the fixed point only in chunks other than the last, whose rows
        // end before nx-3
        const double* L = p + off_l[k];
        const double* ID = p + off_id[k];
        const double* Q = p + off_q[k];
        int kc = nx - 2;
        while (kc > 0 && L[kc - 1] == L[nx - 3] && ID[kc - 1] == ID[nx - 3] && Q[kc - 1] == Q[nx - 3])
            kc--;
        if (xblocks > 1) kc = nx;  // several blocks: no single fixed-point range
        const unsigned long long* omd = nullptr;
        const int* ofd = nullptr;
        const double* ocd = nullptr;
        if (xblocks > 1) {
            omd = (const unsigned long long*)(d + off_om);
            ofd = (const int*)(d + off_of);
            ocd = d + off_oc + k * nov * 9;
        }
        xt[k] = XTables{d + off_l[k], d + off_id[k], d + off_q[k], d + off_pf[k], d + off_qb[k],
                        kc, L[nx - 3], ID[nx - 3], Q[nx - 3], omd, ofd, ocd};
        kconv[k] = kc;
    }
    yt.md = d + off_md;
    yt.mdd = d + off_mdd;
    yt.mf = d + off_mf;
    yt.mfd = d + off_mfd;
    yt.mfdd = d + off_mfdd;
    yt.tw = (const double2*)(d + off_tw);
    return hipSuccess;
}

}  // namespace mhd2d
