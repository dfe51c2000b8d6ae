// Storage of the reference's host globals for builds that do not link its
// initialize.o (the standalone library, tests and the C driver), with a
// setup routine following initialize.c:87-395 for one rank and periodic x.
// A build that links the reference's own objects leaves this file out.
#include <cstdlib>
#include <cstring>

#include "../../include/mhd2d.h"
#include "mhd2d_refglobals.h"

extern "C" {
int NVAR = 8;
int boundaryTypeX = 0;
int divBcleaning = 0;
real constDiffDivB = 2.0;
real cfl = 0.3;
size_t nx = 0, ny = 0, iOffset = 0, iSize = 0, jOffset = 0, jSize = 0, arrSize = 0, nky = 0;
real Lx = 0, dx = 0, Ly = 0, dy = 0;
real adiabaticIdx = 5.0 / 3.0, resistivity = 0, viscosity = 0;
int Hall = 0;
real di = 0;
real Ch_divBcleaning = 0, Cp_divBcleaning = 0;
real *uu = nullptr, *dudx = nullptr, *dudy = nullptr, *d2udx2 = nullptr, *d2udy2 = nullptr,
     *d2Bdx2 = nullptr, *d2Bdy2 = nullptr, *dudt = nullptr, *dudtRK = nullptr;
real *divB = nullptr, *ddivBdx = nullptr, *ddivBdy = nullptr;
real *Efield = nullptr, *dEdx = nullptr, *dEdy = nullptr;
real *uuHaloX0 = nullptr, *uuHaloX1 = nullptr;
real *divBHaloX0 = nullptr, *divBHaloX1 = nullptr;
real *EfieldHaloX0 = nullptr, *EfieldHaloX1 = nullptr;
int npe = 1;
}

namespace {
real* zalloc(size_t n)
{
    real* p = (real*)std::calloc(n ? n : 1, sizeof(real));
    return p;
}
real** all_arrays[] = {&uu,     &dudx,    &dudy,       &d2udx2,       &d2udy2,   &d2Bdx2,
                       &d2Bdy2, &dudt,    &dudtRK,     &divB,         &ddivBdx,  &ddivBdy,
                       &Efield, &dEdx,    &dEdy,       &uuHaloX0,     &uuHaloX1, &divBHaloX0,
                       &divBHaloX1, &EfieldHaloX0, &EfieldHaloX1};
}  // namespace

extern "C" {

void mhd2d_host_finalize(void)
{
    mhd2d_dropin_release();
    for (real** a : all_arrays) {
        std::free(*a);
        *a = nullptr;
    }
}

int mhd2d_host_initialize(int nx_, int ny_, double Lx_, double Ly_, double gamma, double cfl_,
                          int hall, double di_, double resistivity_, double viscosity_,
                          int divbclean)
{
    mhd2d_host_finalize();
    boundaryTypeX = 0;
    npe = 1;
    nx = (size_t)nx_;
    ny = (size_t)ny_;
    Lx = Lx_;
    Ly = Ly_;
    adiabaticIdx = gamma;
    cfl = cfl_;
    Hall = hall;
    di = di_;
    resistivity = resistivity_;
    viscosity = viscosity_;
    divBcleaning = divbclean;
    nky = ny / 2 + 1;              // initialize.c:123
    dx = Lx / nx;                  // initialize.c:127
    dy = Ly / ny;                  // initialize.c:135
    iOffset = 0;
    iSize = nx;                    // one rank: decomposeGrid1D gives the whole x range
    jOffset = 0;
    jSize = ny;
    const double m = dx > dy ? dx : dy;
    constDiffDivB = 2.0 * m * m;   // initialize.c:17, :154
    NVAR = divbclean == 2 ? 9 : 8; // initialize.c:207-214
    arrSize = (size_t)NVAR * iSize * jSize;
    const size_t f = iSize * jSize;
    uu = zalloc(arrSize);
    dudx = zalloc(arrSize);
    dudy = zalloc(arrSize);
    dudt = zalloc(arrSize);
    dudtRK = zalloc(arrSize);
    d2udx2 = zalloc(3 * f);
    d2udy2 = zalloc(3 * f);
    d2Bdx2 = zalloc(3 * f);
    d2Bdy2 = zalloc(3 * f);
    divB = zalloc(f);
    ddivBdx = zalloc(f);
    ddivBdy = zalloc(f);
    Efield = zalloc(3 * f);
    dEdx = zalloc(3 * f);
    dEdy = zalloc(3 * f);
    uuHaloX0 = zalloc((size_t)NVAR * 2 * jSize);
    uuHaloX1 = zalloc((size_t)NVAR * 2 * jSize);
    divBHaloX0 = zalloc(2 * jSize);
    divBHaloX1 = zalloc(2 * jSize);
    EfieldHaloX0 = zalloc(3 * 2 * jSize);
    EfieldHaloX1 = zalloc(3 * 2 * jSize);
    for (real** a : all_arrays)
        if (!*a) return MHD2D_ENOMEM;
    return MHD2D_OK;
}

}  // extern "C"
