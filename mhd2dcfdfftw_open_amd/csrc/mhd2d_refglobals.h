// Declarations of the reference's host globals that the drop-in entry points
// read and write.  Names and types follow the reference's definitions
// (initialize.c:15-79, declared in initialize.h:11-69) so that these entry
// points link against its initialize.o; mhd2d_globals.cpp provides the same
// storage for builds without it.
#pragma once

#include <cstddef>

extern "C" {
typedef double real;
extern int NVAR;
extern int boundaryTypeX;
extern int divBcleaning;
extern real constDiffDivB;
extern real cfl;
extern size_t nx, ny, iOffset, iSize, jOffset, jSize, arrSize, nky;
extern real Lx, dx, Ly, dy;
extern real adiabaticIdx, resistivity, viscosity;
extern int Hall;
extern real di;
extern real Ch_divBcleaning, Cp_divBcleaning;
extern real *uu, *dudx, *dudy, *d2udx2, *d2udy2, *d2Bdx2, *d2Bdy2, *dudt, *dudtRK;
extern real *divB, *ddivBdx, *ddivBdy;
extern real *Efield, *dEdx, *dEdy;
extern real *uuHaloX0, *uuHaloX1;
extern real *divBHaloX0, *divBHaloX1;
extern real *EfieldHaloX0, *EfieldHaloX1;
extern int npe;
}
