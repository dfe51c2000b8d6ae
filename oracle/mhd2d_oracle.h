/*
 * mhd2d_oracle.h -- TEST INFRASTRUCTURE ONLY.
 *
 * Plain-C FP64 restatement of the reference's right-hand-side and RK3
 * time-advance path (chenshihelio/MHD2DCFDFFTW_Open).  It is the checker for
 * the HIP implementation: only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it.  The product library never links it.
 *
 * Every function cites the reference file:line it restates.  Layout is the
 * reference's field-major [v][i][j] (macros.h:4-24): i along x (stride ny),
 * j along y (contiguous).  Periodic x with one rank, so the x-halo of field v
 * is rows (nx-2, nx-1) on the left and (0, 1) on the right (parallel.c:13-24).
 */
#ifndef MHD2D_ORACLE_H
#define MHD2D_ORACLE_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* threads of the OpenMP timing build (libmhd2d_oracle_omp.so); the plain
 * build ignores the setting and reports 1 */
void orc_set_threads(int n);
int orc_threads(void);

/* ---- scheme coefficients (initialize.c:572-631, parameter.h:24-34) ---- */
typedef struct {
    double alpha1, a1, b1, c1;          /* 1st derivative, interior */
    double aS1, bS1, cS1, dS1, eS1;     /* rows 0 / N-1 (alphaSub1)  */
    double alphaS1;
    double alpha2, a2, b2, c2;          /* 2nd derivative, interior */
    double aS2, bS2, cS2, dS2, eS2;
    double alphaS2;
    double alphaF, aF, bF, cF, dF;      /* filter, interior */
    double alphaSF, aSF, bSF, cSF, dSF, eSF;  /* filter rows 0 / N-1 */
} orc_coefs;

void orc_coefs_init(orc_coefs *c, double alpha1, double alpha2,
                    double alphaF, double alphaSF);
void orc_coefs_default(orc_coefs *c);

/* ---- line level (mathlib.c) ---- */
void orc_thomas(int n, const double *lo, const double *diag, const double *up,
                const double *rhs, double *x);
void orc_line_deriv1(int N, double DX, const double *A, const double *H0,
                     const double *H1, double *X);
void orc_line_deriv2(int N, double DX, const double *A, const double *H0,
                     const double *H1, double *X);
void orc_line_filter(int N, const double *A, const double *H0,
                     const double *H1, double *X);
/* same, with caller-chosen alphas (for convergence / option studies) */
void orc_line_deriv1_c(const orc_coefs *c, int N, double DX, const double *A,
                       const double *H0, const double *H1, double *X);
void orc_line_filter_c(const orc_coefs *c, int N, const double *A,
                       const double *H0, const double *H1, double *X);

/* ---- FFTW-convention real transforms (unnormalised) ---- */
void orc_rfft(int n, const double *x, double *re, double *im);   /* r2c */
void orc_irfft(int n, const double *re, const double *im, double *x); /* c2r */
void orc_filtY_table(int ny, double afy, double *filtY);          /* nky entries */

/* ---- field level (calcDerivs.c, filtering.c) ---- */
void orc_derivX(int nx, int ny, double dx, int nv, const double *in, double *out);
void orc_derivXX(int nx, int ny, double dx, int nv, const double *in, double *out);
void orc_derivY(int nx, int ny, double Ly, int nv, const double *in, double *out);
void orc_derivYY(int nx, int ny, double Ly, int nv, const double *in, double *out);
void orc_filterX(int nx, int ny, int nv, double *inout);
void orc_filterY(int nx, int ny, int nv, double *inout);
/* the same x operators with the x range split into npe blocks like the
 * reference's MPI ranks (decomposeGrid1D, initialize.c:547-569) */
void orc_derivX_npe(int nx, int ny, double dx, int nv, int npe, const double *in, double *out);
void orc_derivXX_npe(int nx, int ny, double dx, int nv, int npe, const double *in, double *out);
void orc_filterX_npe(int nx, int ny, int nv, int npe, double *inout);
/* the same with an open x boundary (boundaryTypeX == 1): the first and last
 * blocks close with the one-sided schemes of mathlib.c:87-159, :219-294,
 * :355-440 (npe >= 2) */
void orc_derivX_open(int nx, int ny, double dx, int nv, int npe, const double *in, double *out);
void orc_derivXX_open(int nx, int ny, double dx, int nv, int npe, const double *in, double *out);
void orc_filterX_open(int nx, int ny, int nv, int npe, double *inout);
/* characteristic x-derivative terms of cell (i, j) at an open boundary
 * (0 left, 1 right), calcRHS.c:253-495: out[7] = d/dt of rho, ux, uy, uz,
 * By, Bz, p */
void orc_characteristics_x(int nx, int ny, int i, int j, int boundary, double gamma,
                           const double *uu, const double *dudx, double *out);

/* ---- pointwise physics (calcRHS.c, timeAdvance.c) ---- */
void orc_efield(int nx, int ny, const double *uu, double *E);
void orc_divb(int nx, int ny, const double *dudx, const double *dudy, double *divB);
void orc_rhs(int nx, int ny, double gamma, double resistivity, double viscosity,
             int divbclean, double constDiffDivB,
             const double *uu, const double *dudx, const double *dudy,
             const double *dEdx, const double *dEdy,
             const double *d2udx2, const double *d2udy2,
             const double *d2Bdx2, const double *d2Bdy2,
             const double *ddivBdx, const double *ddivBdy, double *dudt);
/* the same with divBcleaning == 2 (hyperbolic GLM cleaning): uu, dudx, dudy,
 * dudt hold 9 fields, PSI last; ch, cp as calcDt / calc_Cp_fix_Cr leave them */
void orc_rhs_glm(int nx, int ny, double gamma, double resistivity, double viscosity,
                 double ch, double cp, const double *uu, const double *dudx,
                 const double *dudy, const double *dEdx, const double *dEdy,
                 const double *d2udx2, const double *d2udy2,
                 const double *d2Bdx2, const double *d2Bdy2, double *dudt);
void orc_rk_execute(size_t n, double cc, double dd, double *uu,
                    const double *dudt, double *dudtRK);
double orc_calc_dt(int nx, int ny, double dx, double dy, double gamma,
                   double cfl, int hall, double di, double resistivity,
                   double viscosity, const double *uu, double dt0);
double orc_energy(int nx, int ny, double dx, double dy, double gamma,
                  const double *uu);
/* max(cmaxx, cmaxy) of cell (i, j): the Ch calcDt leaves (timeAdvance.c:107-112) */
double orc_glm_ch(int nx, int ny, double dx, double dy, double gamma, int hall,
                  double di, const double *uu, int i, int j);

/* ---- whole-run driver replaying mhd.c:154-206 ---- */
typedef struct orc_sim orc_sim;
orc_sim *orc_sim_create(int nx, int ny, double Lx, double Ly, double gamma,
                        double cfl, int hall, double di, double resistivity,
                        double viscosity, int divbclean);
void orc_sim_destroy(orc_sim *s);
void orc_sim_set_state(orc_sim *s, const double *uu);   /* nvar fields; dt=calcDt(0) */
void orc_sim_get_state(const orc_sim *s, double *uu);
void orc_sim_get_derivs(const orc_sim *s, double *dudx, double *dudy);
void orc_sim_run(orc_sim *s, int nsteps);
void orc_sim_run_fixed_dt(orc_sim *s, int nsteps, const double *dts);
void orc_sim_stage(orc_sim *s, int iRK);                /* one RK stage */
void orc_sim_set_dt(orc_sim *s, double dt);            /* dt of the next step */
void orc_sim_set_npe(orc_sim *s, int npe);             /* x blocks (MPI ranks) */
int orc_sim_nvar(const orc_sim *s);                    /* 9 with divbclean == 2 */
void orc_sim_set_open(orc_sim *s, int open);           /* open x boundary */
double orc_sim_glm_ch(const orc_sim *s, int i);        /* Ch, Cp of x row i */
double orc_sim_glm_cp(const orc_sim *s, int i);
double orc_sim_dt(const orc_sim *s);
double orc_sim_time(const orc_sim *s);
long orc_sim_steps(const orc_sim *s);

#ifdef __cplusplus
}
#endif
#endif
