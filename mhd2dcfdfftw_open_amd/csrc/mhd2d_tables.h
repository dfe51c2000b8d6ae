// Device-resident operator tables of one grid: stencils and Thomas factors of
// the three x operators, spectral multipliers and FFT twiddles along y.
#pragma once

#include "mhd2d_internal.h"

namespace mhd2d {

struct SchemeAlphas {
    // parameter.h:24-34 (boundaryType 0 uses the inner and "sub" alphas)
    double alpha1 = 0.375, alphaS1 = 0.0;
    double alpha2 = 0.375, alphaS2 = 0.0;
    double alphaF = 0.49, alphaSF = 0.95;
    double afy = 0.45;  // initialize.c:44
};

class GridTables {
public:
    GridTables() = default;
    ~GridTables();
    GridTables(const GridTables&) = delete;
    GridTables& operator=(const GridTables&) = delete;

    // nx, ny powers of two; dx = Lx/nx; elem_bytes selects the x-solve
    // chunking.  xblocks > 1 splits x into that many blocks sized as the
    // reference's decomposeGrid1D (initialize.c:547-569), each with its own
    // end rows and Thomas factors, as a run on xblocks MPI ranks has them.
    // open_x: the first and last blocks end in the reference's open-boundary
    // scheme (boundaryTypeX == 1, mathlib.c:87-159, :219-294, :355-440), and
    // dx = Lx / (nx - 1); needs xblocks >= 2.  place (optional): device
    // memory of place_bytes that receives the tables instead of an
    // allocation of their own (hipErrorInvalidValue when they do not fit)
    hipError_t build(int nx, int ny, double Lx, double Ly, int elem_bytes,
                     const SchemeAlphas& al = SchemeAlphas(), int xblocks = 1,
                     int open_x = 0, void* place = nullptr, size_t place_bytes = 0);
    void release();

    int nx = 0, ny = 0, elem_bytes = 0;
    double Lx = 0, Ly = 0, dx = 0, dy = 0;
    int chunk = 0, width = 0, groups = 0, xblocks = 1, open_x = 0;
    XStencil st[NXKIND];
    XTables xt[NXKIND];
    int kconv[NXKIND] = {0, 0, 0};
    YTables yt{};

    // host copies (FP64) for inspection and tests
    static void factor(int n, double alpha, double alphaSub, double* l, double* invd,
                       double* q);
    // the same for given couplings: lo[i] of x[i] in row i+1, up[i] of x[i+1] in row i
    static void factor_rows(int n, const double* lo, const double* up, double* l, double* invd,
                            double* q);
    // offset and size of block b of n rows split over nb (initialize.c:547-569)
    static void block_range(int n, int nb, int b, int* off, int* size);

private:
    void* blob_ = nullptr;  // one allocation holding every table
};

// The first-derivative operator M = A^-1 B of one grid (one x block), formed
// in FP64 and stored as BF16 planes with the non-zero panel lists, for the
// MFMA form of the x-derivative.  nsplit 3 or 6 products.
class XGemmOp {
public:
    XGemmOp() = default;
    ~XGemmOp();
    XGemmOp(const XGemmOp&) = delete;
    XGemmOp& operator=(const XGemmOp&) = delete;
    hipError_t build(const GridTables& tb, int nsplit);
    void release();

    int nx = 0, ny = 0, nsplit = 0, kmax = 0, kmax2 = 0, kmax4 = 0;
    long long panels_total = 0;  // non-zero 64 x 32 panels over all row tiles
    size_t bytes = 0;
    const unsigned short *mh = nullptr, *mm = nullptr, *ml = nullptr;
    const int *klist_d = nullptr, *kcount_d = nullptr;
    // chunk lists of 128- and 256-row tiles, entries (chunk << 4) | mask of
    // the 64-row sub-tiles that multiply the chunk
    const int *klist2_d = nullptr, *kcount2_d = nullptr;
    const int *klist4_d = nullptr, *kcount4_d = nullptr;

private:
    void* blob_ = nullptr;
};

}  // namespace mhd2d
