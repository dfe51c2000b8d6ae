/*
 * mhd2d_oracle.c -- TEST INFRASTRUCTURE ONLY (see mhd2d_oracle.h).
 *
 * Plain C99, FP64, single thread.  A restatement, written for this project,
 * of the reference's hot path so that the HIP build can be checked against
 * it.  The reference itself cannot be built in this image (it needs FFTW3,
 * which is absent), so this file is pinned by the tests in
 * tests/test_oracle_*.py against closed forms, brute-force dense solves,
 * numpy.fft and invariants instead.
 *
 * Citations are file:line in chenshihelio/MHD2DCFDFFTW_Open.
 */
#include "mhd2d_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define ORC_PI 3.14159265358979323846264338327950288

#ifdef _OPENMP
#include <omp.h>
#endif

/* threads of the OpenMP timing build (no effect in the plain build) */
void orc_set_threads(int n)
{
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

int orc_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

static inline double dmax(double a, double b) { return a > b ? a : b; }
static inline double dmin(double a, double b) { return a < b ? a : b; }

/* field indices, macros.h:4-12 (PSI only with divBcleaning == 2) */
enum { RHO = 0, UX, UY, UZ, BX, BY, BZ, PR, PSI };

/* ------------------------------------------------------------------------ */
/* coefficients: initialize.c:572-631 with the alphas of parameter.h:24-34   */
/* ------------------------------------------------------------------------ */
void orc_coefs_init(orc_coefs *c, double alpha1, double alpha2,
                    double alphaF, double alphaSF)
{
    double s1 = 0.0, s2 = 0.0;  /* ALPHASUB1 (parameter.h:26); initialize.c:33 */
    c->alpha1 = alpha1;
    c->a1 = (alpha1 + 9.) / 6.;
    c->b1 = (32. * alpha1 - 9.) / 15.;
    c->c1 = (-3. * alpha1 + 1.) / 10.;
    c->alphaS1 = s1;
    c->aS1 = (1. - s1) / 12.;
    c->bS1 = (3. * s1 - 4.) / 6.;
    c->cS1 = -1.5 * s1;
    c->dS1 = (4. + 5. * s1) / 6.;
    c->eS1 = (3. * s1 - 1.) / 12.;

    c->alpha2 = alpha2;
    c->a2 = (6. - 9. * alpha2) / 4.;
    c->b2 = (-3. + 24. * alpha2) / 5.;
    c->c2 = (2. - 11. * alpha2) / 20.;
    c->alphaS2 = s2;
    c->aS2 = -(1. + s2) / 12.;
    c->bS2 = (s2 + 4.) / 3.;
    c->cS2 = (s2 - 5.) / 2.;
    c->dS2 = (4. - 5. * s2) / 3.;
    c->eS2 = (11. * s2 - 1.) / 12.;

    c->alphaF = alphaF;
    c->aF = (11. + 10. * alphaF) / 16.;
    c->bF = (15. + 34. * alphaF) / 32.;
    c->cF = (-3. + 6. * alphaF) / 16.;
    c->dF = (1. - 2. * alphaF) / 32.;
    c->alphaSF = alphaSF;
    c->aSF = (alphaSF - 1.) / 16.;
    c->bSF = (1. - alphaSF) / 4.;
    c->cSF = (3. * alphaSF + 5.) / 8.;
    c->dSF = (3. * alphaSF + 1.) / 4.;
    c->eSF = (alphaSF - 1.) / 16.;
}

void orc_coefs_default(orc_coefs *c)
{
    orc_coefs_init(c, 0.375, 0.375, 0.49, 0.95);
}

/* ------------------------------------------------------------------------ */
/* Thomas algorithm without pivoting, mathlib.c:8-30                          */
/* lo[i] multiplies x[i] in row i+1; up[i] multiplies x[i+1] in row i.       */
/* ------------------------------------------------------------------------ */
void orc_thomas(int n, const double *lo, const double *diag, const double *up,
                const double *rhs, double *x)
{
    double *b = (double *)malloc(sizeof(double) * n);
    double *r = (double *)malloc(sizeof(double) * n);
    memcpy(b, diag, sizeof(double) * n);
    memcpy(r, rhs, sizeof(double) * n);
    for (int i = 1; i < n; i++) {
        double m = lo[i - 1] / b[i - 1];
        b[i] = b[i] - m * up[i - 1];
        r[i] = r[i] - m * r[i - 1];
    }
    x[n - 1] = r[n - 1] / b[n - 1];
    for (int i = n - 2; i >= 0; i--)
        x[i] = (r[i] - up[i] * x[i + 1]) / b[i];
    free(b);
    free(r);
}

/* periodic-branch matrices: initialize.c:633-654 (1st), 710-731 (2nd),
 * 788-808 (filter): unit diagonal, alpha inside, alphaSub on the two
 * couplings of rows 0 and N-1. */
static void tri_periodic(int n, double alpha, double alphaSub,
                         double *lo, double *diag, double *up)
{
    for (int i = 0; i < n; i++) diag[i] = 1.0;
    for (int i = 0; i < n - 2; i++) lo[i] = alpha;
    lo[n - 2] = alphaSub;
    up[0] = alphaSub;
    for (int i = 1; i < n - 1; i++) up[i] = alpha;
}

/* value of the line at index i in [-2, N+1], halos outside [0, N) */
static inline double at(const double *A, const double *H0, const double *H1,
                        int N, int i)
{
    if (i < 0) return H0[i + 2];         /* H0[0] = f[-2], H0[1] = f[-1] */
    if (i >= N) return H1[i - N];        /* H1[0] = f[N],  H1[1] = f[N+1] */
    return A[i];
}

/* derivCFD1st, periodic branch, mathlib.c:46-86 */
void orc_line_deriv1_c(const orc_coefs *c, int N, double DX, const double *A,
                       const double *H0, const double *H1, double *X)
{
    double *B = (double *)malloc(sizeof(double) * N);
    double *lo = (double *)malloc(sizeof(double) * N);
    double *dg = (double *)malloc(sizeof(double) * N);
    double *up = (double *)malloc(sizeof(double) * N);
    /* row 0: explicit stencil over f[-2..2], mathlib.c:50-52 */
    B[0] = 1. / DX * (c->aS1 * H0[0] + c->bS1 * H0[1] + c->cS1 * A[0] +
                      c->dS1 * A[1] + c->eS1 * A[2]);
    /* rows 1..N-2: 8th-order compact right-hand side, mathlib.c:55-77 */
    for (int i = 1; i <= N - 2; i++) {
        double f1 = at(A, H0, H1, N, i + 1) - at(A, H0, H1, N, i - 1);
        double f2 = at(A, H0, H1, N, i + 2) - at(A, H0, H1, N, i - 2);
        double f3 = at(A, H0, H1, N, i + 3) - at(A, H0, H1, N, i - 3);
        B[i] = 1. / DX * (c->a1 / 2 * f1 + c->b1 / 4 * f2 + c->c1 / 6 * f3);
    }
    /* row N-1: mirror of row 0, mathlib.c:80-82 */
    B[N - 1] = -1. / DX * (c->aS1 * H1[1] + c->bS1 * H1[0] + c->cS1 * A[N - 1] +
                           c->dS1 * A[N - 2] + c->eS1 * A[N - 3]);
    tri_periodic(N, c->alpha1, c->alphaS1, lo, dg, up);
    orc_thomas(N, lo, dg, up, B, X);
    free(B); free(lo); free(dg); free(up);
}

/* derivCFD2nd, periodic branch, mathlib.c:178-218 */
static void line_deriv2_c(const orc_coefs *c, int N, double DX, const double *A,
                          const double *H0, const double *H1, double *X)
{
    double DX2 = DX * DX;
    double *B = (double *)malloc(sizeof(double) * N);
    double *lo = (double *)malloc(sizeof(double) * N);
    double *dg = (double *)malloc(sizeof(double) * N);
    double *up = (double *)malloc(sizeof(double) * N);
    B[0] = 1. / DX2 * (c->aS2 * H0[0] + c->bS2 * H0[1] + c->cS2 * A[0] +
                       c->dS2 * A[1] + c->eS2 * A[2]);
    for (int i = 1; i <= N - 2; i++) {
        double fi = at(A, H0, H1, N, i);
        double s1 = at(A, H0, H1, N, i + 1) - 2 * fi + at(A, H0, H1, N, i - 1);
        double s2 = at(A, H0, H1, N, i + 2) - 2 * fi + at(A, H0, H1, N, i - 2);
        double s3 = at(A, H0, H1, N, i + 3) - 2 * fi + at(A, H0, H1, N, i - 3);
        B[i] = 1. / DX2 * (c->a2 * s1 + c->b2 / 4 * s2 + c->c2 / 9 * s3);
    }
    B[N - 1] = 1. / DX2 * (c->aS2 * H1[1] + c->bS2 * H1[0] + c->cS2 * A[N - 1] +
                           c->dS2 * A[N - 2] + c->eS2 * A[N - 3]);
    tri_periodic(N, c->alpha2, c->alphaS2, lo, dg, up);
    orc_thomas(N, lo, dg, up, B, X);
    free(B); free(lo); free(dg); free(up);
}

/* filterCFD, periodic branch, mathlib.c:309-354 */
void orc_line_filter_c(const orc_coefs *c, int N, const double *A,
                       const double *H0, const double *H1, double *X)
{
    double *B = (double *)malloc(sizeof(double) * N);
    double *lo = (double *)malloc(sizeof(double) * N);
    double *dg = (double *)malloc(sizeof(double) * N);
    double *up = (double *)malloc(sizeof(double) * N);
    B[0] = c->aSF * H0[0] + c->bSF * H0[1] + c->cSF * A[0] + c->dSF * A[1] +
           c->eSF * A[2];
    for (int i = 1; i <= N - 2; i++) {
        B[i] = c->aF * A[i] +
               c->bF / 2 * (at(A, H0, H1, N, i + 1) + at(A, H0, H1, N, i - 1)) +
               c->cF / 2 * (at(A, H0, H1, N, i + 2) + at(A, H0, H1, N, i - 2)) +
               c->dF / 2 * (at(A, H0, H1, N, i + 3) + at(A, H0, H1, N, i - 3));
    }
    B[N - 1] = c->aSF * H1[1] + c->bSF * H1[0] + c->cSF * A[N - 1] +
               c->dSF * A[N - 2] + c->eSF * A[N - 3];
    tri_periodic(N, c->alphaF, c->alphaSF, lo, dg, up);
    orc_thomas(N, lo, dg, up, B, X);
    free(B); free(lo); free(dg); free(up);
}

void orc_line_deriv1(int N, double DX, const double *A, const double *H0,
                     const double *H1, double *X)
{
    orc_coefs c;
    orc_coefs_default(&c);
    orc_line_deriv1_c(&c, N, DX, A, H0, H1, X);
}

void orc_line_deriv2(int N, double DX, const double *A, const double *H0,
                     const double *H1, double *X)
{
    orc_coefs c;
    orc_coefs_default(&c);
    line_deriv2_c(&c, N, DX, A, H0, H1, X);
}

void orc_line_filter(int N, const double *A, const double *H0,
                     const double *H1, double *X)
{
    orc_coefs c;
    orc_coefs_default(&c);
    orc_line_filter_c(&c, N, A, H0, H1, X);
}

/* ------------------------------------------------------------------------ */
/* x-direction field operators: calcDerivs.c:9-72, filtering.c:9-37.  The    */
/* halo of each line is the periodic wrap of the same field (parallel.c:13-24 */
/* with one rank: X0 <- rows nx-2, nx-1; X1 <- rows 0, 1).                   */
/* ------------------------------------------------------------------------ */
typedef enum { OP_D1, OP_D2, OP_F } line_op;

/* decomposeGrid1D, initialize.c:547-569: block sizes n/npe, the first
 * n%npe blocks one larger */
static void block_range(int n, int npe, int b, int *off, int *size)
{
    int base = n / npe, rem = n % npe, o = 0;
    for (int k = 0; k < b; k++) o += base + (k < rem ? 1 : 0);
    *off = o;
    *size = base + (b < rem ? 1 : 0);
}

/* ---- open x-boundary (boundaryTypeX == 1) on the first / last rank ---- */
/* boundary alphas, parameter.h:25-37 and initialize.c:27-40 */
#define OB_ALPHA_B1 3.0
#define OB_ALPHA_B2 5.0
#define OB_ALPHA_F0 1.05
#define OB_ALPHA_F1 0.49
#define OB_ALPHA_F2 0.49
/* near-boundary points, parameter.h:39-56 */
#define OB_ALPHA1P2 (1. / 3.)
#define OB_A1P2 (14. / 9.)
#define OB_B1P2 (1. / 9.)
#define OB_ALPHA2P2 (2. / 11.)
#define OB_A2P2 (12. / 11.)
#define OB_B2P2 (3. / 11.)
#define OB_ALPHA1P1 (1. / 4.)
#define OB_A1P1 1.5
#define OB_ALPHA2P1 0.1
#define OB_A2P1 1.2

/* One line of the first (side 0) or last (side 1) rank with an open x
 * boundary: derivCFD1st mathlib.c:87-159, derivCFD2nd mathlib.c:219-294,
 * filterCFD mathlib.c:355-440, matrices initialize.c:655-708, :732-786,
 * :809-861, coefficients initialize.c:581-629.  The open end reads no halo;
 * the other end is the ordinary closure with the neighbour's halo. */
static void line_open(line_op op, int side, const orc_coefs *c, int N, double DX,
                      const double *A, const double *H0, const double *H1, double *X)
{
    double *B = (double *)malloc(sizeof(double) * N);
    double *lo = (double *)malloc(sizeof(double) * N);
    double *dg = (double *)malloc(sizeof(double) * N);
    double *up = (double *)malloc(sizeof(double) * N);
    double ai, as;             /* interior and sub alphas of the operator */
    double ab, ap1, ap2;       /* boundary, P1 and P2 alphas (filter: F0, F1, F2) */
    if (op == OP_D1) { ai = c->alpha1; as = c->alphaS1; ab = OB_ALPHA_B1; ap1 = OB_ALPHA1P1; ap2 = OB_ALPHA1P2; }
    else if (op == OP_D2) { ai = c->alpha2; as = c->alphaS2; ab = OB_ALPHA_B2; ap1 = OB_ALPHA2P1; ap2 = OB_ALPHA2P2; }
    else { ai = c->alphaF; as = c->alphaSF; ab = OB_ALPHA_F0; ap1 = OB_ALPHA_F1; ap2 = OB_ALPHA_F2; }
    /* the ordinary rows (interior stencil, halo end), as in the periodic
     * branch; the open end's rows are replaced below */
    {
        for (int i = 0; i < N; i++) {
            double v;
            if (op == OP_D1) {
                if (i == 0) v = 1. / DX * (c->aS1 * H0[0] + c->bS1 * H0[1] + c->cS1 * A[0] + c->dS1 * A[1] + c->eS1 * A[2]);
                else if (i == N - 1) v = -1. / DX * (c->aS1 * H1[1] + c->bS1 * H1[0] + c->cS1 * A[N - 1] + c->dS1 * A[N - 2] + c->eS1 * A[N - 3]);
                else v = 1. / DX * (c->a1 / 2 * (at(A, H0, H1, N, i + 1) - at(A, H0, H1, N, i - 1)) +
                                    c->b1 / 4 * (at(A, H0, H1, N, i + 2) - at(A, H0, H1, N, i - 2)) +
                                    c->c1 / 6 * (at(A, H0, H1, N, i + 3) - at(A, H0, H1, N, i - 3)));
            } else if (op == OP_D2) {
                double DX2 = DX * DX;
                if (i == 0) v = 1. / DX2 * (c->aS2 * H0[0] + c->bS2 * H0[1] + c->cS2 * A[0] + c->dS2 * A[1] + c->eS2 * A[2]);
                else if (i == N - 1) v = 1. / DX2 * (c->aS2 * H1[1] + c->bS2 * H1[0] + c->cS2 * A[N - 1] + c->dS2 * A[N - 2] + c->eS2 * A[N - 3]);
                else {
                    double fi = A[i];
                    v = 1. / DX2 * (c->a2 * (at(A, H0, H1, N, i + 1) - 2 * fi + at(A, H0, H1, N, i - 1)) +
                                    c->b2 / 4 * (at(A, H0, H1, N, i + 2) - 2 * fi + at(A, H0, H1, N, i - 2)) +
                                    c->c2 / 9 * (at(A, H0, H1, N, i + 3) - 2 * fi + at(A, H0, H1, N, i - 3)));
                }
            } else {
                if (i == 0) v = c->aSF * H0[0] + c->bSF * H0[1] + c->cSF * A[0] + c->dSF * A[1] + c->eSF * A[2];
                else if (i == N - 1) v = c->aSF * H1[1] + c->bSF * H1[0] + c->cSF * A[N - 1] + c->dSF * A[N - 2] + c->eSF * A[N - 3];
                else v = c->aF * A[i] + c->bF / 2 * (at(A, H0, H1, N, i + 1) + at(A, H0, H1, N, i - 1)) +
                         c->cF / 2 * (at(A, H0, H1, N, i + 2) + at(A, H0, H1, N, i - 2)) +
                         c->dF / 2 * (at(A, H0, H1, N, i + 3) + at(A, H0, H1, N, i - 3));
            }
            B[i] = v;
        }
    }
    /* ordinary matrix */
    for (int i = 0; i < N; i++) dg[i] = 1.0;
    for (int i = 0; i < N - 2; i++) lo[i] = ai;
    lo[N - 2] = as;
    up[0] = as;
    for (int i = 1; i < N - 1; i++) up[i] = ai;
    /* the open end: three rows of the boundary scheme */
    double DX2 = DX * DX;
    if (side == 0) {
        if (op == OP_D1) {
            double aB = -(11. + 2. * ab) / 6., bB = (6. - ab) / 2., cB = (2. * ab - 3.) / 2., dB = (2. - ab) / 6.;
            B[0] = 1 / DX * (aB * A[0] + bB * A[1] + cB * A[2] + dB * A[3]);
            B[1] = 1 / DX * (OB_A1P1 / 2 * (A[2] - A[0]));
            B[2] = 1 / DX * (OB_A1P2 / 2 * (A[3] - A[1]) + OB_B1P2 / 4 * (A[4] - A[0]));
        } else if (op == OP_D2) {
            double aB = (11. * ab + 35.) / 12., bB = -(5. * ab + 26.) / 3., cB = (ab + 19.) / 2.,
                   dB = (ab - 14.) / 3., eB = (11. - ab) / 12.;
            B[0] = 1 / DX2 * (aB * A[0] + bB * A[1] + cB * A[2] + dB * A[3] + eB * A[4]);
            B[1] = 1 / DX2 * (OB_A2P1 * (A[2] - 2 * A[1] + A[0]));
            B[2] = 1 / DX2 * (OB_A2P2 * (A[3] - 2 * A[2] + A[1]) + OB_B2P2 / 4 * (A[4] - 2 * A[2] + A[0]));
        } else {
            double a0 = ab, a1 = ap1, a2 = ap2;
            B[0] = (15. + a0) / 16. * A[0] + (1. + 3. * a0) / 4. * A[1] + 3. * (a0 - 1) / 8. * A[2] +
                   (1. - a0) / 4. * A[3] + (a0 - 1.) / 16. * A[4];
            B[1] = (1. + 14. * a1) / 16. * A[0] + (3. + 2. * a1) / 4. * A[1] + (3. + 2. * a1) / 8. * A[2] +
                   (-1. + 2. * a1) / 4. * A[3] + (1. - 2. * a1) / 16. * A[4];
            B[2] = (5. + 6. * a2) / 8. * A[2] + (1. + 2. * a2) / 2. / 2 * (A[3] + A[1]) +
                   -(1. - 2. * a2) / 8. / 2 * (A[4] + A[0]);
        }
        up[0] = ab;
        lo[0] = ap1;
        up[1] = ap1;
        lo[1] = ap2;
        up[2] = ap2;
    } else {
        if (op == OP_D1) {
            double aB = -(11. + 2. * ab) / 6., bB = (6. - ab) / 2., cB = (2. * ab - 3.) / 2., dB = (2. - ab) / 6.;
            B[N - 3] = 1 / DX * (OB_A1P2 / 2 * (A[N - 2] - A[N - 4]) + OB_B1P2 / 4 * (A[N - 1] - A[N - 5]));
            B[N - 2] = 1 / DX * (OB_A1P1 / 2 * (A[N - 1] - A[N - 3]));
            B[N - 1] = -1 / DX * (aB * A[N - 1] + bB * A[N - 2] + cB * A[N - 3] + dB * A[N - 4]);
        } else if (op == OP_D2) {
            double aB = (11. * ab + 35.) / 12., bB = -(5. * ab + 26.) / 3., cB = (ab + 19.) / 2.,
                   dB = (ab - 14.) / 3., eB = (11. - ab) / 12.;
            B[N - 3] = 1 / DX2 * (OB_A2P2 * (A[N - 2] - 2 * A[N - 3] + A[N - 4]) +
                                  OB_B2P2 / 4 * (A[N - 1] - 2 * A[N - 3] + A[N - 5]));
            B[N - 2] = 1 / DX2 * (OB_A2P1 * (A[N - 1] - 2 * A[N - 2] + A[N - 3]));
            B[N - 1] = 1 / DX2 * (aB * A[N - 1] + bB * A[N - 2] + cB * A[N - 3] + dB * A[N - 4] + eB * A[N - 5]);
        } else {
            double a0 = ab, a1 = ap1, a2 = ap2;
            B[N - 3] = (5. + 6. * a2) / 8. * A[N - 3] + (1. + 2. * a2) / 2. / 2 * (A[N - 4] + A[N - 2]) +
                       -(1. - 2. * a2) / 8. / 2 * (A[N - 5] + A[N - 1]);
            B[N - 2] = (1. + 14. * a1) / 16. * A[N - 1] + (3. + 2. * a1) / 4. * A[N - 2] +
                       (3. + 2. * a1) / 8. * A[N - 3] + (-1. + 2. * a1) / 4. * A[N - 4] +
                       (1. - 2. * a1) / 16. * A[N - 5];
            B[N - 1] = (15. + a0) / 16. * A[N - 1] + (1. + 3. * a0) / 4. * A[N - 2] +
                       3. * (a0 - 1) / 8. * A[N - 3] + (1. - a0) / 4. * A[N - 4] + (a0 - 1.) / 16. * A[N - 5];
        }
        lo[N - 2] = ab;
        up[N - 2] = ap1;
        lo[N - 3] = ap1;
        up[N - 3] = ap2;
        lo[N - 4] = ap2;
    }
    orc_thomas(N, lo, dg, up, B, X);
    free(B); free(lo); free(dg); free(up);
}

/* x operator with the x range split into npe blocks as the reference's MPI
 * ranks do: each block is solved on its own with a 2-point halo from its
 * periodic neighbours (parallel.c:13-24), so rows at block edges are the
 * explicit closure rows.  open != 0: the first and last blocks close the
 * line with the open-boundary scheme instead (boundaryTypeX == 1, npe >= 2) */
static void x_apply_npe_bc(line_op op, int nx, int ny, double dx, int nv, int npe, int open,
                           const double *in, double *out);

static void x_apply_npe(line_op op, int nx, int ny, double dx, int nv, int npe,
                        const double *in, double *out)
{
    x_apply_npe_bc(op, nx, ny, dx, nv, npe, 0, in, out);
}

static void x_apply_npe_bc(line_op op, int nx, int ny, double dx, int nv, int npe, int open,
                           const double *in, double *out)
{
    orc_coefs c;
    orc_coefs_default(&c);
    /* whole column first: halos are the neighbours' unfiltered values, and
     * `in` may alias `out`.  Columns are independent: with OpenMP (the
     * CPU-baseline timing build) they are shared over threads, each with its
     * own buffers, with the same result bit for bit as one thread. */
#ifdef _OPENMP
#pragma omp parallel
#endif
    {
    double *col = (double *)malloc(sizeof(double) * nx);
    double *sol = (double *)malloc(sizeof(double) * nx);
    double H0[2], H1[2];
    long nl = (long)nv * ny;
#ifdef _OPENMP
#pragma omp for schedule(static)
#endif
    for (long l = 0; l < nl; l++) {
        int v = (int)(l / ny), j = (int)(l % ny);
        const double *f = in + (size_t)v * nx * ny;
        double *g = out + (size_t)v * nx * ny;
        {
            for (int i = 0; i < nx; i++) col[i] = f[(size_t)i * ny + j];
            for (int b = 0; b < npe; b++) {
                int off, N;
                block_range(nx, npe, b, &off, &N);
                for (int h = 0; h < 2; h++) {
                    H0[h] = col[(off - 2 + h + nx) % nx];
                    H1[h] = col[(off + N + h) % nx];
                }
                if (open && npe >= 2 && (b == 0 || b == npe - 1))
                    line_open(op, b == 0 ? 0 : 1, &c, N, dx, col + off, H0, H1, sol + off);
                else if (op == OP_D1) orc_line_deriv1_c(&c, N, dx, col + off, H0, H1, sol + off);
                else if (op == OP_D2) line_deriv2_c(&c, N, dx, col + off, H0, H1, sol + off);
                else orc_line_filter_c(&c, N, col + off, H0, H1, sol + off);
            }
            for (int i = 0; i < nx; i++) g[(size_t)i * ny + j] = sol[i];
        }
    }
    free(col);
    free(sol);
    }
}

static void x_apply(line_op op, int nx, int ny, double dx, int nv,
                    const double *in, double *out)
{
    x_apply_npe(op, nx, ny, dx, nv, 1, in, out);
}

void orc_derivX_npe(int nx, int ny, double dx, int nv, int npe, const double *in, double *out)
{
    x_apply_npe(OP_D1, nx, ny, dx, nv, npe, in, out);
}

void orc_derivXX_npe(int nx, int ny, double dx, int nv, int npe, const double *in, double *out)
{
    x_apply_npe(OP_D2, nx, ny, dx, nv, npe, in, out);
}

void orc_filterX_npe(int nx, int ny, int nv, int npe, double *inout)
{
    x_apply_npe(OP_F, nx, ny, 0.0, nv, npe, inout, inout);
}

void orc_derivX_open(int nx, int ny, double dx, int nv, int npe, const double *in, double *out)
{
    x_apply_npe_bc(OP_D1, nx, ny, dx, nv, npe, 1, in, out);
}

void orc_derivXX_open(int nx, int ny, double dx, int nv, int npe, const double *in, double *out)
{
    x_apply_npe_bc(OP_D2, nx, ny, dx, nv, npe, 1, in, out);
}

void orc_filterX_open(int nx, int ny, int nv, int npe, double *inout)
{
    x_apply_npe_bc(OP_F, nx, ny, 0.0, nv, npe, 1, inout, inout);
}

void orc_derivX(int nx, int ny, double dx, int nv, const double *in, double *out)
{
    x_apply(OP_D1, nx, ny, dx, nv, in, out);
}

void orc_derivXX(int nx, int ny, double dx, int nv, const double *in, double *out)
{
    x_apply(OP_D2, nx, ny, dx, nv, in, out);
}

void orc_filterX(int nx, int ny, int nv, double *inout)
{
    x_apply(OP_F, nx, ny, 0.0, nv, inout, inout);
}

/* ------------------------------------------------------------------------ */
/* FFT: iterative radix-2 complex transform with directly evaluated twiddles */
/* (stands in for FFTW's r2c/c2r, initialize.c:323-326, which are            */
/* unnormalised; c2r ignores the imaginary parts of bins 0 and n/2).         */
/* ------------------------------------------------------------------------ */
/* twiddles w[k] = exp(sign 2 pi i k / n), k < n/2, with the angle formed as
 * the transform's stages would form it for (k, len = n).  A stage of length
 * len = n/s takes w[k s]: (2 pi k s)/(len s) rounds exactly like
 * (2 pi k)/len (scaling by a power of two), so the table gives the same
 * twiddles, bit for bit, as evaluating cos/sin per stage. */
static void twiddles(int n, int sign, double *wr, double *wi)
{
    for (int k = 0; k < n / 2; k++) {
        double ang = sign * 2.0 * ORC_PI * (double)k / (double)n;
        wr[k] = cos(ang);
        wi[k] = sin(ang);
    }
}

static void cfft_t(int n, double *re, double *im, const double *twr, const double *twi)
{
    /* bit reversal */
    for (int i = 1, j = 0; i < n; i++) {
        int bit = n >> 1;
        for (; j & bit; bit >>= 1) j ^= bit;
        j ^= bit;
        if (i < j) {
            double t = re[i]; re[i] = re[j]; re[j] = t;
            t = im[i]; im[i] = im[j]; im[j] = t;
        }
    }
    for (int len = 2; len <= n; len <<= 1) {
        int half = len >> 1, stride = n / len;
        for (int k = 0; k < half; k++) {
            double wr = twr[k * stride], wi = twi[k * stride];
            for (int s = 0; s < n; s += len) {
                int a = s + k, b = s + k + half;
                double tr = re[b] * wr - im[b] * wi;
                double ti = re[b] * wi + im[b] * wr;
                re[b] = re[a] - tr; im[b] = im[a] - ti;
                re[a] += tr;        im[a] += ti;
            }
        }
    }
}

static void cfft(int n, double *re, double *im, int sign)
{
    double *wr = (double *)calloc(n / 2 + 1, sizeof(double));
    double *wi = (double *)calloc(n / 2 + 1, sizeof(double));
    twiddles(n, sign, wr, wi);
    cfft_t(n, re, im, wr, wi);
    free(wr);
    free(wi);
}

void orc_rfft(int n, const double *x, double *re, double *im)
{
    double *r = (double *)malloc(sizeof(double) * n);
    double *i_ = (double *)calloc(n, sizeof(double));
    memcpy(r, x, sizeof(double) * n);
    cfft(n, r, i_, -1);
    for (int k = 0; k <= n / 2; k++) { re[k] = r[k]; im[k] = i_[k]; }
    free(r);
    free(i_);
}

void orc_irfft(int n, const double *re, const double *im, double *x)
{
    double *r = (double *)malloc(sizeof(double) * n);
    double *i_ = (double *)malloc(sizeof(double) * n);
    int h = n / 2;
    r[0] = re[0]; i_[0] = 0.0;
    r[h] = re[h]; i_[h] = 0.0;
    for (int k = 1; k < h; k++) {
        r[k] = re[k];      i_[k] = im[k];
        r[n - k] = re[k];  i_[n - k] = -im[k];
    }
    cfft(n, r, i_, +1);
    memcpy(x, r, sizeof(double) * n);
    free(r);
    free(i_);
}

/* initialize.c:161-177 */
void orc_filtY_table(int ny, double afy, double *filtY)
{
    int nky = ny / 2 + 1;
    double aj = (5.0 + 6.0 * afy) / 8.0;
    double bj = (1.0 + 2.0 * afy) / 2.0;
    double cj = -(1.0 - 2 * afy) / 8.0;
    for (int k = 0; k < nky; k++) {
        double w = 2. * ORC_PI * (double)k / ny;
        filtY[k] = (aj + bj * cos(w) + cj * cos(2 * w)) / (1 + 2 * afy * cos(w));
    }
}

typedef enum { Y_D1, Y_D2, Y_F } y_op;

/* derivY calcDerivs.c:75-109, derivYY :114-150, filteringY filtering.c:40-94.
 * Lines are independent: with OpenMP (the CPU-baseline timing build) they
 * are shared over threads, each with its own buffers; the result is the
 * same bit for bit as one thread. */
static void y_apply(y_op op, int nx, int ny, double Ly, int nv,
                    const double *in, double *out)
{
    int nky = ny / 2 + 1, h = ny / 2;
    double *filtY = (double *)malloc(sizeof(double) * nky);
    double *twf_r = (double *)malloc(sizeof(double) * (h + 1));
    double *twf_i = (double *)malloc(sizeof(double) * (h + 1));
    double *twb_r = (double *)malloc(sizeof(double) * (h + 1));
    double *twb_i = (double *)malloc(sizeof(double) * (h + 1));
    orc_filtY_table(ny, 0.45, filtY);  /* afy, initialize.c:44 */
    twiddles(ny, -1, twf_r, twf_i);
    twiddles(ny, +1, twb_r, twb_i);
#ifdef _OPENMP
#pragma omp parallel
#endif
    {
        double *r = (double *)malloc(sizeof(double) * ny);
        double *q = (double *)malloc(sizeof(double) * ny);
        double *re = (double *)malloc(sizeof(double) * nky);
        double *im = (double *)malloc(sizeof(double) * nky);
        long nl = (long)nv * nx;
#ifdef _OPENMP
#pragma omp for schedule(static)
#endif
        for (long l = 0; l < nl; l++) {
            const double *f = in + (size_t)l * ny;
            double *g = out + (size_t)l * ny;
            /* r2c (orc_rfft) */
            memcpy(r, f, sizeof(double) * ny);
            memset(q, 0, sizeof(double) * ny);
            cfft_t(ny, r, q, twf_r, twf_i);
            for (int k = 0; k < nky; k++) { re[k] = r[k]; im[k] = q[k]; }
            for (int k = 0; k < nky; k++) {
                double a = re[k], b = im[k];
                if (op == Y_D1) {
                    double kk = 2 * ORC_PI / Ly * k;
                    re[k] = -b * kk / ny;
                    im[k] = a * kk / ny;
                } else if (op == Y_D2) {
                    double kk = 2 * ORC_PI / Ly * k;
                    re[k] = -a * kk * kk / ny;
                    im[k] = -b * kk * kk / ny;
                } else {
                    re[k] = a / ny * filtY[k];
                    im[k] = b / ny * filtY[k];
                }
            }
            /* c2r (orc_irfft): the imaginary parts of bins 0 and n/2 are
             * ignored */
            r[0] = re[0]; q[0] = 0.0;
            r[h] = re[h]; q[h] = 0.0;
            for (int k = 1; k < h; k++) {
                r[k] = re[k];       q[k] = im[k];
                r[ny - k] = re[k];  q[ny - k] = -im[k];
            }
            cfft_t(ny, r, q, twb_r, twb_i);
            memcpy(g, r, sizeof(double) * ny);
        }
        free(r); free(q); free(re); free(im);
    }
    free(filtY); free(twf_r); free(twf_i); free(twb_r); free(twb_i);
}

void orc_derivY(int nx, int ny, double Ly, int nv, const double *in, double *out)
{
    y_apply(Y_D1, nx, ny, Ly, nv, in, out);
}

void orc_derivYY(int nx, int ny, double Ly, int nv, const double *in, double *out)
{
    y_apply(Y_D2, nx, ny, Ly, nv, in, out);
}

void orc_filterY(int nx, int ny, int nv, double *inout)
{
    y_apply(Y_F, nx, ny, 1.0, nv, inout, inout);
}

/* ------------------------------------------------------------------------ */
/* pointwise physics                                                          */
/* ------------------------------------------------------------------------ */
/* calcEfield, calcRHS.c:24-43: E = -u x B */
void orc_efield(int nx, int ny, const double *uu, double *E)
{
    size_t n = (size_t)nx * ny;
#ifdef _OPENMP
#pragma omp parallel for schedule(static)
#endif
    for (size_t p = 0; p < n; p++) {
        double ux = uu[UX * n + p], uy = uu[UY * n + p], uz = uu[UZ * n + p];
        double bx = uu[BX * n + p], by = uu[BY * n + p], bz = uu[BZ * n + p];
        E[0 * n + p] = uz * by - uy * bz;
        E[1 * n + p] = ux * bz - uz * bx;
        E[2 * n + p] = uy * bx - ux * by;
    }
}

/* calcDivB, calcRHS.c:13-22 */
void orc_divb(int nx, int ny, const double *dudx, const double *dudy, double *divB)
{
    size_t n = (size_t)nx * ny;
    for (size_t p = 0; p < n; p++) divB[p] = dudx[BX * n + p] + dudy[BY * n + p];
}

/* calcRHSInnerPoint calcRHS.c:77-159 + addSourceTerm calcRHS.c:498-544.
 * ch/cp: per x row, the Ch and Cp of the rank owning that row (GLM only). */
/* calcCharacteristicsX, calcRHS.c:253-495: the x-derivative part of the
 * primitive equations at an open boundary cell, split into the 7 waves
 * (fast, Alfven, slow in both directions, entropy) with the waves entering
 * the domain removed; boundary 0 = left (x = 0), 1 = right.  out[0..6] =
 * d/dt of rho, ux, uy, uz, By, Bz, p. */
static void characteristics_x(size_t n, size_t p, int boundary, double gamma,
                              const double *uu, const double *dudx, double *out)
{
    double ux = uu[UX * n + p], uy = uu[UY * n + p], uz = uu[UZ * n + p];
    double bx = uu[BX * n + p], by = uu[BY * n + p], bz = uu[BZ * n + p];
    double rho = uu[RHO * n + p], pr = uu[PR * n + p];
    double dux = dudx[UX * n + p], duy = dudx[UY * n + p], duz = dudx[UZ * n + p];
    double dby = dudx[BY * n + p], dbz = dudx[BZ * n + p];
    double drho = dudx[RHO * n + p], dp = dudx[PR * n + p];
    (void)uy;
    (void)uz;
    double cFlow = ux;
    double cAlfx = bx / sqrt(rho), cAlfy = by / sqrt(rho), cAlfz = bz / sqrt(rho);
    double cSound2 = gamma * pr / rho, cSound = sqrt(cSound2);
    double cAlfx2 = cAlfx * cAlfx;
    double sgn = cAlfx >= 0 ? 1.0 : -1.0;
    double cAlf2 = cAlfx2 + cAlfy * cAlfy + cAlfz * cAlfz;
    double cAlfperp = sqrt(cAlfy * cAlfy + cAlfz * cAlfz);
    double angY, angZ;
    if (cAlfperp < 1e-7) {   /* calcRHS.c:307-315: degenerate direction */
        angY = 0.0;
        angZ = 0.0;
    } else {
        angY = cAlfy / cAlfperp;
        angZ = cAlfz / cAlfperp;
    }
    double cM2 = cSound2 + cAlf2;
    double cN2 = sqrt(dmax(cM2 * cM2 - 4 * cSound2 * cAlfx2, 0.));
    double cFast = sqrt(dmax(cM2 + cN2, 0.)) / sqrt(2.0);
    double cSlow = sqrt(dmax(cM2 - cN2, 0.)) / sqrt(2.0);
    double al1 = sqrt(fabs((cFast * cFast - cAlfx2) / (cFast * cFast - cSlow * cSlow)));
    double al2 = sqrt(fabs((cFast * cFast - cSound2) / (cFast * cFast - cSlow * cSlow)));
    double byzP = angZ * dbz + angY * dby, uyzP = angZ * duz + angY * duy;
    double byzM = angY * dbz - angZ * dby, uyzM = angY * duz - angZ * duy;
    double sr = sqrt(rho);
    double fP = (cFlow + cFast) * (al1 * (dp / rho / cFast + dux) + al2 * (byzP / sr - cAlfx / cFast * uyzP));
    double aP = (cFlow + sgn * cAlfx) * (uyzM - sgn / sr * byzM);
    double sP = (cFlow + cSlow) * (-al2 * (dp / rho / cSound + sgn * cAlfx / cFast * dux) +
                                   al1 * (cSound / cFast / sr * byzP - sgn * uyzP));
    double en = ux * cSound * (dp / gamma / pr - drho / rho);
    double sM = (cFlow - cSlow) * (-al2 * (dp / rho / cSound - sgn * cAlfx / cFast * dux) +
                                   al1 * (cSound / cFast / sr * byzP + sgn * uyzP));
    double aM = (cFlow - sgn * cAlfx) * (uyzM + sgn / sr * byzM);
    double fM = (cFlow - cFast) * (al1 * (dp / rho / cFast - dux) + al2 * (byzP / sr + cAlfx / cFast * uyzP));
    if (boundary == 0) {     /* left: waves moving in +x enter (calcRHS.c:383-419) */
        if (cFlow + cFast >= 0) fP = 0;
        if (cFlow + sgn * cAlfx >= 0) aP = 0;
        if (cFlow + cSlow >= 0) sP = 0;
        if (cFlow >= 0) en = 0;
        if (cFlow - cSlow >= 0) sM = 0;
        if (cFlow - sgn * cAlfx >= 0) aM = 0;
        if (cFlow - cFast >= 0) fM = 0;
    } else {                 /* right: waves moving in -x enter (calcRHS.c:424-457) */
        if (cFlow + cFast <= 0) fP = 0;
        if (cFlow + sgn * cAlfx <= 0) aP = 0;
        if (cFlow + cSlow <= 0) sP = 0;
        if (cFlow <= 0) en = 0;
        if (cFlow - cSlow <= 0) sM = 0;
        if (cFlow - sgn * cAlfx <= 0) aM = 0;
        if (cFlow - cFast <= 0) fM = 0;
    }
    /* calcRHS.c:461-494 */
    out[0] = rho / cSound * en + rho * (al2 / 2 / cSound * (sP + sM) - al1 / 2 / cFast * (fP + fM));
    out[1] = cSlow * al2 / 2 / cSound * (sP - sM) - al1 / 2 * (fP - fM);
    out[2] = angZ / 2 * (aP + aM) + angY * (sgn * al1 / 2 * (sP - sM) + cAlfx * al2 / 2 / cFast * (fP - fM));
    out[3] = -angY / 2 * (aP + aM) + angZ * (sgn * al1 / 2 * (sP - sM) + cAlfx * al2 / 2 / cFast * (fP - fM));
    out[4] = -sgn * sr * angZ / 2 * (aP - aM) - sr * angY * (cSound * al1 / 2 / cFast * (sP + sM) + al2 / 2 * (fP + fM));
    out[5] = sgn * sr * angY / 2 * (aP - aM) - sr * angZ * (cSound * al1 / 2 / cFast * (sP + sM) + al2 / 2 * (fP + fM));
    out[6] = gamma * pr * (al2 / 2 / cSound * (sP + sM) - al1 / 2 / cFast * (fP + fM));
}

void orc_characteristics_x(int nx, int ny, int i, int j, int boundary, double gamma,
                           const double *uu, const double *dudx, double *out)
{
    characteristics_x((size_t)nx * ny, (size_t)i * ny + j, boundary, gamma, uu, dudx, out);
}

/* calcRHSOpenBoundaryX, calcRHS.c:163-251: a cell of an open x boundary line
 * (x = 0 on the first rank, x = Lx on the last) */
static void open_boundary_cell(size_t n, size_t p, int boundary, double gamma, int divbclean,
                               double ch, double cp, const double *uu, const double *dudx,
                               const double *dudy, double *dudt)
{
    double rho = uu[RHO * n + p], ux = uu[UX * n + p], uy = uu[UY * n + p];
    double bx = uu[BX * n + p], by = uu[BY * n + p], bz = uu[BZ * n + p];
    double pr = uu[PR * n + p];
    /* Bx keeps its d/dx terms in the u.grad form (calcRHS.c:172-179) */
    double divU = dudx[UX * n + p] + dudy[UY * n + p];
    dudt[BX * n + p] = -ux * dudx[BX * n + p] - uy * dudy[BX * n + p] + bx * dudx[UX * n + p] +
                       by * dudy[UX * n + p] - divU * bx;
    /* everything else without d/dx terms (calcRHS.c:182-214) */
    double jx = dudy[BZ * n + p], jy = 0, jz = -dudy[BX * n + p];
    divU = dudy[UY * n + p];
    dudt[PR * n + p] = -uy * dudy[PR * n + p] - gamma * pr * divU;
    dudt[RHO * n + p] = -rho * divU - uy * dudy[RHO * n + p];
    dudt[UX * n + p] = -uy * dudy[UX * n + p] + jy * bz / rho - jz * by / rho;
    dudt[UY * n + p] = -uy * dudy[UY * n + p] - dudy[PR * n + p] / rho + jz * bx / rho - jx * bz / rho;
    dudt[UZ * n + p] = -uy * dudy[UZ * n + p] + jx * by / rho - jy * bx / rho;
    dudt[BY * n + p] = -uy * dudy[BY * n + p] + by * dudy[UY * n + p] - divU * by;
    dudt[BZ * n + p] = -uy * dudy[BZ * n + p] + by * dudy[UZ * n + p] - divU * bz;
    /* the outgoing waves' x-derivative terms (calcRHS.c:216-239) */
    double c[7];
    characteristics_x(n, p, boundary, gamma, uu, dudx, c);
    dudt[RHO * n + p] += c[0];
    dudt[UX * n + p] += c[1];
    dudt[UY * n + p] += c[2];
    dudt[UZ * n + p] += c[3];
    dudt[BY * n + p] += c[4];
    dudt[BZ * n + p] += c[5];
    dudt[PR * n + p] += c[6];
    if (divbclean == 2) {    /* calcRHS.c:241-249: psi terms only */
        double divB = dudx[BX * n + p] + dudy[BY * n + p];
        dudt[BX * n + p] = dudt[BX * n + p] - dudx[PSI * n + p];
        dudt[BY * n + p] = dudt[BY * n + p] - dudy[PSI * n + p];
        dudt[PSI * n + p] = -ch * ch * (divB + uu[PSI * n + p] / cp / cp);
    }
}

/* open != 0: rows 0 and nx-1 are open x boundaries (calcRHS.c:56-72) */
static void rhs_impl_bc(int nx, int ny, double gamma, double resistivity, double viscosity,
                        int divbclean, double constDiffDivB, const double *ch,
                        const double *cp, const double *uu, const double *dudx,
                        const double *dudy, const double *dEdx, const double *dEdy,
                        const double *d2udx2, const double *d2udy2,
                        const double *d2Bdx2, const double *d2Bdy2,
                        const double *ddivBdx, const double *ddivBdy, double *dudt, int open);

static void rhs_impl(int nx, int ny, double gamma, double resistivity, double viscosity,
                     int divbclean, double constDiffDivB, const double *ch,
                     const double *cp, const double *uu, const double *dudx,
                     const double *dudy, const double *dEdx, const double *dEdy,
                     const double *d2udx2, const double *d2udy2,
                     const double *d2Bdx2, const double *d2Bdy2,
                     const double *ddivBdx, const double *ddivBdy, double *dudt)
{
    rhs_impl_bc(nx, ny, gamma, resistivity, viscosity, divbclean, constDiffDivB, ch, cp, uu,
                dudx, dudy, dEdx, dEdy, d2udx2, d2udy2, d2Bdx2, d2Bdy2, ddivBdx, ddivBdy, dudt,
                0);
}

static void rhs_impl_bc(int nx, int ny, double gamma, double resistivity, double viscosity,
                        int divbclean, double constDiffDivB, const double *ch,
                        const double *cp, const double *uu, const double *dudx,
                        const double *dudy, const double *dEdx, const double *dEdy,
                        const double *d2udx2, const double *d2udy2,
                        const double *d2Bdx2, const double *d2Bdy2,
                        const double *ddivBdx, const double *ddivBdy, double *dudt, int open)
{
    size_t n = (size_t)nx * ny;
#ifdef _OPENMP
#pragma omp parallel for schedule(static)
#endif
    for (size_t p = 0; p < n; p++) {
        size_t row = p / (size_t)ny;
        if (open && (row == 0 || row == (size_t)nx - 1)) {
            open_boundary_cell(n, p, row == 0 ? 0 : 1, gamma, divbclean, ch ? ch[row] : 0,
                               cp ? cp[row] : 0, uu, dudx, dudy, dudt);
        } else {
        double rho = uu[RHO * n + p], ux = uu[UX * n + p], uy = uu[UY * n + p];
        double bx = uu[BX * n + p], by = uu[BY * n + p], bz = uu[BZ * n + p];
        double pr = uu[PR * n + p];
        double divU = dudx[UX * n + p] + dudy[UY * n + p];
        double jx = dudy[BZ * n + p];
        double jy = -dudx[BZ * n + p];
        double jz = dudx[BY * n + p] - dudy[BX * n + p];

        dudt[RHO * n + p] = -rho * divU - ux * dudx[RHO * n + p] - uy * dudy[RHO * n + p];
        dudt[UX * n + p] = -ux * dudx[UX * n + p] - uy * dudy[UX * n + p]
                           - dudx[PR * n + p] / rho + jy * bz / rho - jz * by / rho;
        dudt[UY * n + p] = -ux * dudx[UY * n + p] - uy * dudy[UY * n + p]
                           - dudy[PR * n + p] / rho + jz * bx / rho - jx * bz / rho;
        dudt[UZ * n + p] = -ux * dudx[UZ * n + p] - uy * dudy[UZ * n + p]
                           + jx * by / rho - jy * bx / rho;
        /* induction as -curl E, calcRHS.c:135-137 */
        dudt[BX * n + p] = -dEdy[2 * n + p];
        dudt[BY * n + p] = dEdx[2 * n + p];
        dudt[BZ * n + p] = dEdy[0 * n + p] - dEdx[1 * n + p];
        dudt[PR * n + p] = -ux * dudx[PR * n + p] - uy * dudy[PR * n + p] - gamma * pr * divU;

        if (divbclean == 2) {  /* Dedner et al. 2002 eq. 24, calcRHS.c:144-158 */
            size_t i = p / (size_t)ny;
            double divB = dudx[BX * n + p] + dudy[BY * n + p];
            double psi = uu[PSI * n + p];
            dudt[UX * n + p] = dudt[UX * n + p] - divB * bx;
            dudt[UY * n + p] = dudt[UY * n + p] - divB * by;
            dudt[UZ * n + p] = dudt[UZ * n + p] - divB * bz;
            dudt[BX * n + p] = dudt[BX * n + p] - dudx[PSI * n + p];
            dudt[BY * n + p] = dudt[BY * n + p] - dudy[PSI * n + p];
            dudt[PSI * n + p] = -ch[i] * ch[i] * (divB + psi / cp[i] / cp[i]);
        }
        }
        /* addSourceTerm, calcRHS.c:498-544, on every cell */
        const double rho = uu[RHO * n + p];
        if (viscosity > 0) {  /* calcRHS.c:501-515 */
            for (int c = 0; c < 3; c++)
                dudt[(UX + c) * n + p] += viscosity / rho *
                    (d2udx2[c * n + p] + d2udy2[c * n + p]);
        }
        if (resistivity > 0) {  /* calcRHS.c:517-531 */
            for (int c = 0; c < 3; c++)
                dudt[(BX + c) * n + p] += resistivity *
                    (d2Bdx2[c * n + p] + d2Bdy2[c * n + p]);
        }
        if (divbclean == 1) {  /* calcRHS.c:533-543 */
            dudt[BX * n + p] += constDiffDivB * ddivBdx[p];
            dudt[BY * n + p] += constDiffDivB * ddivBdy[p];
        }
    }
}

void orc_rhs(int nx, int ny, double gamma, double resistivity, double viscosity,
             int divbclean, double constDiffDivB,
             const double *uu, const double *dudx, const double *dudy,
             const double *dEdx, const double *dEdy,
             const double *d2udx2, const double *d2udy2,
             const double *d2Bdx2, const double *d2Bdy2,
             const double *ddivBdx, const double *ddivBdy, double *dudt)
{
    if (divbclean == 2) divbclean = 0;  /* GLM needs Ch, Cp: orc_rhs_glm */
    rhs_impl(nx, ny, gamma, resistivity, viscosity, divbclean, constDiffDivB, NULL, NULL,
             uu, dudx, dudy, dEdx, dEdy, d2udx2, d2udy2, d2Bdx2, d2Bdy2, ddivBdx,
             ddivBdy, dudt);
}

void orc_rhs_glm(int nx, int ny, double gamma, double resistivity, double viscosity,
                 double ch, double cp, const double *uu, const double *dudx,
                 const double *dudy, const double *dEdx, const double *dEdy,
                 const double *d2udx2, const double *d2udy2,
                 const double *d2Bdx2, const double *d2Bdy2, double *dudt)
{
    double *chr = (double *)malloc(sizeof(double) * nx);
    double *cpr = (double *)malloc(sizeof(double) * nx);
    for (int i = 0; i < nx; i++) { chr[i] = ch; cpr[i] = cp; }
    rhs_impl(nx, ny, gamma, resistivity, viscosity, 2, 0.0, chr, cpr, uu, dudx, dudy,
             dEdx, dEdy, d2udx2, d2udy2, d2Bdx2, d2Bdy2, NULL, NULL, dudt);
    free(chr);
    free(cpr);
}

/* rkExecute, timeAdvance.c:29-36 */
void orc_rk_execute(size_t n, double cc, double dd, double *uu,
                    const double *dudt, double *dudtRK)
{
#ifdef _OPENMP
#pragma omp parallel for schedule(static)
#endif
    for (size_t i = 0; i < n; i++) {
        uu[i] = uu[i] + cc * dudt[i] + dd * dudtRK[i];
        dudtRK[i] = dudt[i];
    }
}


/* the largest signal speeds along x and y of cell p, timeAdvance.c:49-105 */
static void cell_speeds(size_t n, size_t p, double dx, double dy, double gamma,
                        int hall, double di, const double *uu, double *cmx_out,
                        double *cmy_out)
{
    double rho = uu[RHO * n + p];
    double cs2 = gamma * uu[PR * n + p] / rho;
    double cax = uu[BX * n + p] / sqrt(rho);
    double cay = uu[BY * n + p] / sqrt(rho);
    double caz = uu[BZ * n + p] / sqrt(rho);
    double ca2 = cax * cax + cay * cay + caz * caz;
    double cms2 = cs2 + ca2;
    double nsx = sqrt(dmax(cms2 * cms2 - 4 * cs2 * cax * cax, 0.));
    double nsy = sqrt(dmax(cms2 * cms2 - 4 * cs2 * cay * cay, 0.));
    double cfx = sqrt(cms2 + nsx) / sqrt(2.);
    double cfy = sqrt(cms2 + nsy) / sqrt(2.);
    double csx = sqrt(dmax(cms2 - nsx, 0.)) / sqrt(2.);
    double csy = sqrt(dmax(cms2 - nsy, 0.)) / sqrt(2.);
    double ux = uu[UX * n + p], uy = uu[UY * n + p];
    double m;
    m = dmax(fabs(ux), fabs(ux - cax));
    m = dmax(fabs(ux - csx), m);
    m = dmax(fabs(ux - cfx), m);
    m = dmax(fabs(ux + cax), m);
    m = dmax(fabs(ux + csx), m);
    double cmx = dmax(fabs(ux + cfx), m);
    m = dmax(fabs(uy), fabs(uy - cay));
    m = dmax(fabs(uy - csy), m);
    m = dmax(fabs(uy - cfy), m);
    m = dmax(fabs(uy + cay), m);
    m = dmax(fabs(uy + csy), m);
    double cmy = dmax(fabs(uy + cfy), m);
    if (hall == 1) {  /* whistler limit, timeAdvance.c:97-105 */
        double ch = di / dmin(dx, dy) / rho *
            dmax(fabs(uu[BX * n + p]),
                 dmax(fabs(uu[BY * n + p]), fabs(uu[BZ * n + p])));
        cmx = dmax(cmx, ch);
        cmy = dmax(cmy, ch);
    }
    *cmx_out = cmx;
    *cmy_out = cmy;
}

/* calcDt, timeAdvance.c:38-153 (the divBcleaning==2 branch sets Ch: orc_glm_ch) */
double orc_calc_dt(int nx, int ny, double dx, double dy, double gamma,
                   double cfl, int hall, double di, double resistivity,
                   double viscosity, const double *uu, double dt0)
{
    size_t n = (size_t)nx * ny;
    double dtmin = -1;
    /* the minimum does not depend on the order of the cells, so the OpenMP
     * timing build finds the same value */
#ifdef _OPENMP
#pragma omp parallel
#endif
    {
    double local = -1;
#ifdef _OPENMP
#pragma omp for schedule(static)
#endif
    for (size_t p = 0; p < n; p++) {
        double cmx, cmy;
        cell_speeds(n, p, dx, dy, gamma, hall, di, uu, &cmx, &cmy);
        double dtx = dx / cmx, dty = dy / cmy;
        if (resistivity > 0) {
            dtx = dmin(dtx, dx * dx / resistivity);
            dty = dmin(dty, dy * dy / resistivity);
        }
        if (viscosity > 0) {
            dtx = dmin(dtx, dx * dx / viscosity);
            dty = dmin(dty, dy * dy / viscosity);
        }
        double it = dmin(dtx, dty);
        if (local < 0 || it < local) local = it;
    }
#ifdef _OPENMP
#pragma omp critical
#endif
    if (local >= 0 && (dtmin < 0 || local < dtmin)) dtmin = local;
    }
    dtmin = dtmin * cfl;
    /* hysteresis, timeAdvance.c:144-151 */
    if (dt0 < 0.98 * dtmin || dt0 > 1.02 * dtmin) return dtmin;
    return dt0;
}

/* Ch of the divBcleaning==2 branch of calcDt (timeAdvance.c:107-112): it is
 * assigned inside the cell loop, so the value left is that of the last cell
 * the rank visits, (iSize-1, jSize-1) of its block, not a maximum. */
double orc_glm_ch(int nx, int ny, double dx, double dy, double gamma, int hall,
                  double di, const double *uu, int i, int j)
{
    double cmx, cmy;
    cell_speeds((size_t)nx * ny, (size_t)i * ny + j, dx, dy, gamma, hall, di, uu, &cmx, &cmy);
    return dmax(cmx, cmy);
}

/* total energy: sum of 1/2 rho u^2 + p/(gamma-1) + 1/2 B^2 times the cell
 * area (no such diagnostic exists in the reference). */
double orc_energy(int nx, int ny, double dx, double dy, double gamma,
                  const double *uu)
{
    size_t n = (size_t)nx * ny;
    double s = 0.0;
    for (size_t p = 0; p < n; p++) {
        double rho = uu[RHO * n + p];
        double u2 = uu[UX * n + p] * uu[UX * n + p] + uu[UY * n + p] * uu[UY * n + p] +
                    uu[UZ * n + p] * uu[UZ * n + p];
        double b2 = uu[BX * n + p] * uu[BX * n + p] + uu[BY * n + p] * uu[BY * n + p] +
                    uu[BZ * n + p] * uu[BZ * n + p];
        s += 0.5 * rho * u2 + uu[PR * n + p] / (gamma - 1.0) + 0.5 * b2;
    }
    return s * dx * dy;
}

/* ------------------------------------------------------------------------ */
/* run driver: the stage order of mhd.c:154-206                               */
/* ------------------------------------------------------------------------ */
struct orc_sim {
    int nx, ny, nvar, hall, divbclean, npe, open;
    double Lx, Ly, dx, dy, gamma, cfl, di, resistivity, viscosity, constDiffDivB;
    double dt, time;
    long steps;
    double cc[3], dd[3];
    double *uu, *dudx, *dudy, *E, *dEdx, *dEdy, *dudt, *dudtRK;
    double *d2udx2, *d2udy2, *d2Bdx2, *d2Bdy2, *divB, *ddivBdx, *ddivBdy;
    double *ch, *cp;   /* per x row: Ch and Cp of the owning rank (GLM) */
};

orc_sim *orc_sim_create(int nx, int ny, double Lx, double Ly, double gamma,
                        double cfl, int hall, double di, double resistivity,
                        double viscosity, int divbclean)
{
    orc_sim *s = (orc_sim *)calloc(1, sizeof(orc_sim));
    size_t n = (size_t)nx * ny;
    s->nx = nx; s->ny = ny; s->npe = 1;
    s->nvar = divbclean == 2 ? 9 : 8;   /* initialize.c:205-214 */
    s->Lx = Lx; s->Ly = Ly;
    s->dx = Lx / nx;             /* periodic x, initialize.c:125-128 */
    s->dy = Ly / ny;             /* initialize.c:135 */
    s->gamma = gamma; s->cfl = cfl; s->hall = hall; s->di = di;
    s->resistivity = resistivity; s->viscosity = viscosity;
    s->divbclean = divbclean;
    /* initialize.c:17 and :154 */
    s->constDiffDivB = 2.0 * dmax(s->dx, s->dy) * dmax(s->dx, s->dy);
    size_t nv = (size_t)s->nvar;
    s->uu = (double *)calloc(nv * n, sizeof(double));
    s->dudx = (double *)calloc(nv * n, sizeof(double));
    s->dudy = (double *)calloc(nv * n, sizeof(double));
    s->dudt = (double *)calloc(nv * n, sizeof(double));
    s->dudtRK = (double *)calloc(nv * n, sizeof(double));
    s->ch = (double *)calloc(nx, sizeof(double));
    s->cp = (double *)calloc(nx, sizeof(double));
    s->E = (double *)calloc(3 * n, sizeof(double));
    s->dEdx = (double *)calloc(3 * n, sizeof(double));
    s->dEdy = (double *)calloc(3 * n, sizeof(double));
    s->d2udx2 = (double *)calloc(3 * n, sizeof(double));
    s->d2udy2 = (double *)calloc(3 * n, sizeof(double));
    s->d2Bdx2 = (double *)calloc(3 * n, sizeof(double));
    s->d2Bdy2 = (double *)calloc(3 * n, sizeof(double));
    s->divB = (double *)calloc(n, sizeof(double));
    s->ddivBdx = (double *)calloc(n, sizeof(double));
    s->ddivBdy = (double *)calloc(n, sizeof(double));
    return s;
}

void orc_sim_destroy(orc_sim *s)
{
    if (!s) return;
    free(s->uu); free(s->dudx); free(s->dudy); free(s->dudt); free(s->dudtRK);
    free(s->E); free(s->dEdx); free(s->dEdy);
    free(s->d2udx2); free(s->d2udy2); free(s->d2Bdx2); free(s->d2Bdy2);
    free(s->divB); free(s->ddivBdx); free(s->ddivBdy);
    free(s->ch); free(s->cp);
    free(s);
}

/* rkInit, timeAdvance.c:20-27 */
static void rk_init(orc_sim *s, double dt)
{
    s->cc[0] = 8.0 / 15.0 * dt; s->dd[0] = 0.0;
    s->cc[1] = 5.0 / 12.0 * dt; s->dd[1] = -17. / 60. * dt;
    s->cc[2] = 0.75 * dt;       s->dd[2] = -5. / 12. * dt;
    memset(s->dudtRK, 0, sizeof(double) * s->nvar * (size_t)s->nx * s->ny);
}

/* the Ch / Cp side effect of calcDt (timeAdvance.c:107-112,
 * divBcleaning.c:7-10), per rank */
static void sim_glm_update(orc_sim *s)
{
    if (s->divbclean != 2) return;
    for (int b = 0; b < s->npe; b++) {
        int off, N;
        block_range(s->nx, s->npe, b, &off, &N);
        double ch = orc_glm_ch(s->nx, s->ny, s->dx, s->dy, s->gamma, s->hall, s->di,
                               s->uu, off + N - 1, s->ny - 1);
        double cp = sqrt(0.18 * ch);   /* Cr_divBCleaning, divBcleaning.c:4 */
        for (int i = off; i < off + N; i++) { s->ch[i] = ch; s->cp[i] = cp; }
    }
}

static double sim_calc_dt(orc_sim *s, double dt0)
{
    sim_glm_update(s);
    return orc_calc_dt(s->nx, s->ny, s->dx, s->dy, s->gamma, s->cfl, s->hall,
                       s->di, s->resistivity, s->viscosity, s->uu, dt0);
}

void orc_sim_set_state(orc_sim *s, const double *uu)
{
    memcpy(s->uu, uu, sizeof(double) * s->nvar * (size_t)s->nx * s->ny);
    s->time = 0.0;
    s->steps = 0;
    s->dt = sim_calc_dt(s, 0.0);   /* mhd.c:58 */
    rk_init(s, s->dt);             /* mhd.c:63 */
}

void orc_sim_get_state(const orc_sim *s, double *uu)
{
    memcpy(uu, s->uu, sizeof(double) * s->nvar * (size_t)s->nx * s->ny);
}

void orc_sim_get_derivs(const orc_sim *s, double *dudx, double *dudy)
{
    size_t n = s->nvar * (size_t)s->nx * s->ny;
    if (dudx) memcpy(dudx, s->dudx, sizeof(double) * n);
    if (dudy) memcpy(dudy, s->dudy, sizeof(double) * n);
}

/* one pass of mhd.c:156-198 */
void orc_sim_stage(orc_sim *s, int iRK)
{
    int nx = s->nx, ny = s->ny;
    size_t n = (size_t)nx * ny;
    int nv = s->nvar;
    x_apply_npe_bc(OP_D1, nx, ny, s->dx, nv, s->npe, s->open, s->uu, s->dudx);           /* mhd.c:160 */
    orc_derivY(nx, ny, s->Ly, nv, s->uu, s->dudy);           /* mhd.c:161 */
    if (s->viscosity > 0) {                                  /* mhd.c:163-167 */
        x_apply_npe_bc(OP_D2, nx, ny, s->dx, 3, s->npe, s->open, s->uu + UX * n, s->d2udx2);
        orc_derivYY(nx, ny, s->Ly, 3, s->uu + UX * n, s->d2udy2);
    }
    /* mhd.c:169-173 also runs these when Hall==1, but their output is read
     * only when resistivity > 0 (calcRHS.c:517), so they are skipped here. */
    if (s->resistivity > 0) {
        x_apply_npe_bc(OP_D2, nx, ny, s->dx, 3, s->npe, s->open, s->uu + BX * n, s->d2Bdx2);
        orc_derivYY(nx, ny, s->Ly, 3, s->uu + BX * n, s->d2Bdy2);
    }
    if (s->divbclean == 1) {                                 /* mhd.c:175-182 */
        orc_divb(nx, ny, s->dudx, s->dudy, s->divB);
        x_apply_npe_bc(OP_D1, nx, ny, s->dx, 1, s->npe, s->open, s->divB, s->ddivBdx);
        orc_derivY(nx, ny, s->Ly, 1, s->divB, s->ddivBdy);
    }
    orc_efield(nx, ny, s->uu, s->E);                         /* mhd.c:186 */
    x_apply_npe_bc(OP_D1, nx, ny, s->dx, 3, s->npe, s->open, s->E, s->dEdx);             /* mhd.c:188 */
    orc_derivY(nx, ny, s->Ly, 3, s->E, s->dEdy);             /* mhd.c:189 */
    rhs_impl_bc(nx, ny, s->gamma, s->resistivity, s->viscosity, s->divbclean,
                s->constDiffDivB, s->ch, s->cp, s->uu, s->dudx, s->dudy, s->dEdx, s->dEdy,
                s->d2udx2, s->d2udy2, s->d2Bdx2, s->d2Bdy2, s->ddivBdx,
                s->ddivBdy, s->dudt, s->open);               /* mhd.c:192 */
    orc_rk_execute(nv * n, s->cc[iRK], s->dd[iRK], s->uu, s->dudt, s->dudtRK);
    x_apply_npe_bc(OP_F, nx, ny, 0.0, nv, s->npe, s->open, s->uu, s->uu);              /* filtering.c:100 */
    orc_filterY(nx, ny, nv, s->uu);                          /* filtering.c:101 */
}

static void sim_step(orc_sim *s, const double *forced_dt)
{
    for (int iRK = 0; iRK < 3; iRK++) orc_sim_stage(s, iRK);
    s->time += s->dt;                                        /* mhd.c:201 */
    s->steps++;
    if (forced_dt) {
        sim_glm_update(s);   /* calcDt's Ch side effect still happens */
        s->dt = *forced_dt;
    } else {
        s->dt = sim_calc_dt(s, s->dt);                       /* mhd.c:205 */
    }
    rk_init(s, s->dt);                                       /* mhd.c:206 */
}

void orc_sim_run(orc_sim *s, int nsteps)
{
    for (int k = 0; k < nsteps; k++) sim_step(s, NULL);
}

/* replay a given dt sequence: dts[k] is the dt used after step k */
void orc_sim_run_fixed_dt(orc_sim *s, int nsteps, const double *dts)
{
    for (int k = 0; k < nsteps; k++) sim_step(s, &dts[k]);
}

/* use dt for the next step (replaces the calcDt value of set_state or of
 * the last step): rkInit(dt), timeAdvance.c:20-27.  A test hook, so that
 * runs from slightly different states advance with the same dt sequence. */
void orc_sim_set_dt(orc_sim *s, double dt)
{
    s->dt = dt;
    rk_init(s, dt);
}

/* emulate the reference run on npe MPI ranks along x (block closure) */
void orc_sim_set_npe(orc_sim *s, int npe) { s->npe = npe < 1 ? 1 : npe; }

/* boundaryTypeX == 1 (open x boundary, meaningful with npe >= 2): dx =
 * Lx/(nx-1) (initialize.c:125-133) and constDiffDivB from it (:154); call
 * before orc_sim_set_state */
void orc_sim_set_open(orc_sim *s, int open)
{
    s->open = open ? 1 : 0;
    s->dx = s->open ? s->Lx / (s->nx - 1) : s->Lx / s->nx;
    s->constDiffDivB = 2.0 * dmax(s->dx, s->dy) * dmax(s->dx, s->dy);
}

int orc_sim_nvar(const orc_sim *s) { return s->nvar; }
double orc_sim_glm_ch(const orc_sim *s, int i) { return s->ch[i]; }
double orc_sim_glm_cp(const orc_sim *s, int i) { return s->cp[i]; }

double orc_sim_dt(const orc_sim *s) { return s->dt; }
double orc_sim_time(const orc_sim *s) { return s->time; }
long orc_sim_steps(const orc_sim *s) { return s->steps; }
