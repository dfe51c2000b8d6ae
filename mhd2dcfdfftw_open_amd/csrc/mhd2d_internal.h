// Internal declarations shared by the HIP kernels, the context API and the
// drop-in layer.  Not part of the public C ABI (include/mhd2d.h).
#pragma once

#include <hip/hip_runtime.h>
#include <cstddef>
#include <cstdint>

namespace mhd2d {

// field indices of the state, same order as the reference (macros.h:4-11)
enum Field { RHO = 0, UX, UY, UZ, BX, BY, BZ, PR, NFIELD = 8 };
// psi of the hyperbolic (GLM) div B cleaning in the fused engine's state
// buffer (after the 8 fields and E_y, E_z), and of its derivative buffers
constexpr int PSI_SLOT = 10;

struct RunScalars;

// ---------------------------------------------------------------- x operators
enum XKind { XD1 = 0, XD2 = 1, XFILT = 2, NXKIND = 3 };
// two solve variants beside the XKind ones (template values of the x-solve
// kernels): the x-derivative of a periodic line that stops after the two
// chunk scans and stores, per chunk, the values entering it (y at the row
// before it, x at the row after it), from which the fused x-derivative +
// RHS re-solves the chunks; and the filter whose input rows at chunk edges
// come from XSolveArgs::edge
constexpr int XD1_CARRY = 3;
constexpr int XFILT_EDGE = 4;

// Scaled stencil weights of one x operator.  Interior rows:
//   D1: r = c1 (f+1 - f-1) + c2 (f+2 - f-2) + c3 (f+3 - f-3)
//   D2: r = c1 (f+1 - 2f + f-1) + c2 (f+2 - 2f + f-2) + c3 (f+3 - 2f + f-3)
//   F : r = c0 f + c1 (f+1 + f-1) + c2 (f+2 + f-2) + c3 (f+3 + f-3)
// Row 0: r = sum_k e[k] f[k-2]; row N-1: r = esign * sum_k e[k] f[N+1-k].
struct XStencil {
    double c0, c1, c2, c3;
    double e[5];
    double esign;
};

// Thomas factors of one x operator's matrix (computed once, FP64) and the
// chunk products used to carry the two recurrences across thread chunks.
//   forward : y_i = r_i - l_i y_{i-1}           (l_0 = 0)
//   backward: x_i = y_i * invd_i + q_i x_{i+1}  (q_{N-1} = 0)
struct XTables {
    const double* l;      // [nx]
    const double* invd;   // [nx]
    const double* q;      // [nx]
    const double* pf;     // [G] product of -l over each chunk
    const double* qb;     // [G] product of q over each chunk
    int kconv;            // l, invd, q equal (bitwise) their fixed point on rows kconv..nx-3
    double linf, dinf, qinf;
    // x split into blocks solved separately (the reference's MPI ranks),
    // possibly with open x ends: rows whose right-hand side is not the
    // interior stencil (block edges, the rows of an open end) take an
    // explicit stencil over f[i-4 .. i+4].  Bit k of omask[g] marks row
    // g*C + k; the marked rows of chunk g use ocoef[9 * (ofirst[g] + m)],
    // m = 0, 1, ... in row order.  Null when x is one periodic block (rows 0
    // and nx-1 are the only end rows and use XStencil.e).
    const unsigned long long* omask;
    const int* ofirst;
    const double* ocoef;
};

// T: arithmetic and output type; TI: input type (a FP64 state feeding FP32
// derivative solves in the mixed form)
template <typename T, typename TI = T>
struct XSolveArgs {
    const TI* in[16];
    T* out[16];
    XStencil st;
    XTables tb;
    const RunScalars* rs = nullptr;  // optional: skip the launch once the run has stopped
    int kind;
    int nx, ny;
    int accumulate = 0;      // out += result (second x-derivatives onto the y part)
    // launch shape: strip width (0: the default rule of xsolve_shape, else 8
    // or 16 columns) and the XCD-contiguous workgroup order (-1: default,
    // on for 8-column strips; 0 / 1: off / on), as chosen by the autotuner
    int width = 0;
    int remap = -1;
    // instance z of a batch (blockIdx.z) reads in[v] + z * bs_in, writes
    // out[v] + z * bs_out and checks rs[z]
    size_t bs_in = 0, bs_out = 0;
    // derivative solves (XD1, one periodic block) that stop after the chunk
    // scans: with carry set, out is not written; instead carry + z * bs_carry
    // receives, per field v and chunk g of C rows, the values entering the
    // chunk, [v][2][nx / C][ny]: y at row g C - 1 of the forward recurrence
    // (0 for g = 0), then x at row g C + C of the backward one (0 for the
    // last chunk)
    T* carry = nullptr;
    size_t bs_carry = 0;
    // filter solves (XFILT, one periodic block, in == out) whose input's
    // first and last 3 rows of every chunk of C rows (the rows other
    // chunks read as halos) are not in in[v] but in edge + z * bs_edge +
    // v * edge_fs, [v][nx / C][6][ny]: rows g C .. g C + 2, then g C + C - 3
    // .. g C + C - 1 (what k_xrhs_rk leaves there)
    const T* edge = nullptr;
    size_t edge_fs = 0, bs_edge = 0;
};

// x-derivative through the dense operator M = A^-1 B split into BF16 planes
// (mh, mm, ml: row-major nx x nx), applied as an MFMA GEMM over the y-lines;
// klist[t * kmax ...] lists the XGEMM_KC-column chunks of row tile t (64
// rows) that hold non-zeros, kcount[t] how many
constexpr int XGEMM_KC = 32;
template <typename TI>
struct XGemmArgs {
    const TI* in[16];
    float* out[16];
    const unsigned short* mh = nullptr;
    const unsigned short* mm = nullptr;
    const unsigned short* ml = nullptr;
    const int* klist = nullptr;
    const int* kcount = nullptr;
    int kmax = 0;
    // the same for 128- and 256-row tiles: entries (chunk << 4) | mask of
    // the 64-row sub-tiles that multiply the chunk
    const int* klist2 = nullptr;
    const int* kcount2 = nullptr;
    int kmax2 = 0;
    const int* klist4 = nullptr;
    const int* kcount4 = nullptr;
    int kmax4 = 0;
    int nsplit = 3;
    const RunScalars* rs = nullptr;
    int nx = 0, ny = 0;
    size_t bs_in = 0, bs_out = 0;
};

// ---------------------------------------------------------------- y operators
// Spectral multipliers, precomputed in FP64 for k = 0..ny-1 (full complex
// spectrum, signed wavenumbers):
//   md [k] = kappa(k) / ny          (first derivative, Nyquist -> 0)
//   mdd[k] = -kappa(k)^2 / ny       (second derivative, Nyquist kept)
//   mf [k] = filtY(|k|) / ny        (y filter)
//   mfd[k] = filtY(|k|) kappa(k) / ny
// with kappa(k) = 2 pi / Ly * (k <= ny/2 ? k : k - ny) and the FFT twiddles
// tw[t] = exp(-2 pi i t / ny).
struct YTables {
    const double* md;
    const double* mdd;
    const double* mf;
    const double* mfd;
    const double* mfdd;      // filter times second derivative
    const double2* tw;
};

enum YMode { Y_D1 = 0, Y_D2 = 1, Y_F = 2 };

template <typename T>
struct YApplyArgs {          // generic multiplier pass (drop-in entry points)
    const T* in[16];
    T* out[16];
    YTables tb;
    int nv, mode;
    int nx, ny;
    // instance z of a batch (blockIdx.y): in[v] + z * bs_in, out[v] + z * bs_out
    size_t bs_in = 0, bs_out = 0;
};

struct DtParams {
    double dx, dy, gamma, cfl, di, resistivity, viscosity;
    int hall;
};

// device-side scalars of one run
struct RunScalars {
    unsigned long long dtmin_bits;   // running min of the local dt limits
    int nonfinite;                   // set when a dt limit is not finite
    int stop;                        // set when the new dt is < 1e-8 (mhd.c:131)
    double dt;                       // dt of the current step
    double dt_prev;                  // dt of the previous step (0 at start)
    double time;                     // simulated time
    double cc[3], dd[3];             // RK3 coefficients times dt
    long long step;
    double dt_stop;                  // the dt that triggered the stop
    double glm_ch;                   // fastest speed of the last cell (GLM Ch)
    double di;                       // ion inertial length of this instance (Hall CFL)
};

// T: state and arithmetic type; TD: type of the stored y derivatives
template <typename T, typename TD = T>
struct YSweepArgs {          // fused y-sweep of the driver
    T* u;                    // [8][nx][ny], in place (filtered)
    TD* dudy;                // [8][nx][ny]
    T* E;                    // [2][nx][ny]: Ey, Ez (for the x-solves)
    TD* dEdy;                // [2][nx][ny]: dEx/dy, dEz/dy
    YTables tb;
    DtParams dp;
    RunScalars* rs;
    int nx, ny;
    int filter;              // apply the pending y filter first
    int dt_limits;           // reduce the local dt limits into rs
    int derivs;              // compute the y derivatives and E
    // second y-derivatives (with derivs) for the source terms: of ux, uy, uz
    // (viscosity) into lap[0..2] and of Bx, By, Bz (resistivity) into
    // lap[3..5]; the x-solve adds the second x-derivatives
    TD* lap = nullptr;       // [6][nx][ny]
    int lap_u = 0, lap_b = 0;
    int glm = 0;             // psi (slot PSI_SLOT) is filtered and differentiated too
    // instance z of a batch (blockIdx.y): u, E + z * bs_u; dudy, dEdy +
    // z * bs_d; lap + z * bs_l; rs + z (whose di replaces dp.di)
    size_t bs_u = 0, bs_d = 0, bs_l = 0;
};

template <typename T>
struct RhsArgs {
    const T* u;
    const T* dudx;
    const T* dudy;
    const T* dEdx;
    const T* dEdy;
    const T* d2udx2;  // optional (viscosity)
    const T* d2udy2;
    const T* d2Bdx2;  // optional (resistivity)
    const T* d2Bdy2;
    const T* ddivBdx; // optional (diffusive div B cleaning)
    const T* ddivBdy;
    T* dudt;
    double gamma, resistivity, viscosity, constDiffDivB;
    double ch, cp;    // hyperbolic cleaning (divbclean == 2): field 8 is psi
    int divbclean;
    size_t n;
};

// T: state and arithmetic type; TD: type of the derivatives and of R_prev
template <typename T, typename TD = T>
struct RhsRkArgs {           // fused RHS + low-storage RK update (driver)
    T* u;
    const TD* dudx;
    const TD* dudy;
    const TD* dEdx;          // [2][nx][ny]: dEy/dx, dEz/dx
    const TD* dEdy;          // [2][nx][ny]: dEx/dy, dEz/dy
    TD* rprev;
    const TD* lap = nullptr; // [6][nx][ny] Laplacians of u and B (source terms)
    double viscosity = 0, resistivity = 0;
    // diffusive div B cleaning (divBcleaning == 1): d(div B)/dx, d(div B)/dy
    const TD* ddivb = nullptr;  // [2][nx][ny]
    double cdiff = 0;           // constDiffDivB
    // hyperbolic cleaning (divBcleaning == 2): psi in slot PSI_SLOT of u,
    // dudx, dudy and in field 8 of rprev; Ch, Cp of each of the npe x blocks
    // of instance z at glm_cp[2 (z npe + b)]
    int glm = 0;
    int npe = 1;
    const double* glm_cp = nullptr;
    const RunScalars* rs;
    double gamma;
    int stage;
    int write_rprev;
    size_t n;
    // instance z of a batch (blockIdx.y): u + z * bs_u, dudx/dudy/dEdx/dEdy
    // + z * bs_d, rprev + z * bs_r, lap + z * bs_l, ddivb + z * bs_b, rs + z
    size_t bs_u = 0, bs_d = 0, bs_r = 0, bs_l = 0, bs_b = 0;
    // open x boundary: rows 0 and nx-1 are left to launch_rhs_rk_open
    int open_x = 0;
    int nx = 0, ny = 0;
};

// fused x-derivative + RHS + RK (ideal MHD, one periodic x block): x holds
// the derivative solve's arguments (in[0..8]: rho, ux, uy, uz, By, Bz, p,
// Ey, Ez; st, tb of XD1), r the RHS + RK arguments (dudx, dEdx unused; the
// state is read from r.u), carry / bs_carry what the carry solve
// (XSolveArgs::carry) stored.  The updated state goes to u_out + z * bs_uo
// The updated state of a chunk's rows that other chunks read as x halos
// (its first and last 3) goes to edge + z * bs_edge + f * edge_fs in the
// layout of XSolveArgs::edge, the others back to r.u: the workgroups of
// neighbouring chunks may still be reading those rows of r.u.
template <typename T, typename TD>
struct XRhsArgs {
    XSolveArgs<TD, T> x;
    RhsRkArgs<T, TD> r;
    const TD* carry = nullptr;
    size_t bs_carry = 0;
    T* edge = nullptr;
    size_t edge_fs = 0, bs_edge = 0;
};

// ---------------------------------------------------------------- shape groups
// One launch over the shape groups of an uneven ensemble: group g owns the
// blocks [first[g], first[g + 1]) of a flattened grid, holds nb[g]
// instances, per[g] blocks per instance; aux[g] and code[g] are kernel-class
// specific (x-solves: column strips per field and the (chunk, width) code).
constexpr int MULTI_MAXG = 16;
struct MultiMap {
    int ng = 0;
    int first[MULTI_MAXG + 1] = {};
    int nb[MULTI_MAXG] = {};
    int per[MULTI_MAXG] = {};
    int aux[MULTI_MAXG] = {};
    int code[MULTI_MAXG] = {};
};

struct DtMultiArgs {
    RunScalars* rs;  // the group's instances
    double cfl;
};

// ---------------------------------------------------------------- launchers
template <typename T, typename TI = T>
hipError_t launch_xsolve(const XSolveArgs<T, TI>& a, int nv, hipStream_t s, int nb = 1);
template <typename T>
hipError_t launch_yapply(const YApplyArgs<T>& a, hipStream_t s, int nb = 1);
template <typename T, typename TD = T>
hipError_t launch_ysweep(const YSweepArgs<T, TD>& a, hipStream_t s, int nb = 1);
template <typename TI>
hipError_t launch_xgemm(const XGemmArgs<TI>& a, int nv, hipStream_t s, int nb = 1);
template <typename T>
hipError_t launch_rhs(const RhsArgs<T>& a, hipStream_t s);
template <typename T, typename TD = T>
hipError_t launch_rhs_rk(const RhsRkArgs<T, TD>& a, hipStream_t s, int nb = 1);
template <typename T, typename TD = T>
hipError_t launch_xrhs_rk(const XRhsArgs<T, TD>& a, hipStream_t s, int nb = 1);
// the fused x-derivative + RHS + RK covers this grid (elem_bytes: the state
// word): the x chunks are C = xsolve_shape's rows, a workgroup's columns
// 512 / C (FP64 state) or 1024 / C (FP32) must divide ny
int xrhs_supported(int nx, int ny, int elem_bytes);
// columns of one workgroup tile of the fused kernel (elem_bytes: the state word)
int xrhs_tile_columns(int nx, int elem_bytes);
// the two open boundary lines (x = 0, x = Lx): calcRHSOpenBoundaryX with the
// characteristic terms, then the same RK update
template <typename T, typename TD = T>
hipError_t launch_rhs_rk_open(const RhsRkArgs<T, TD>& a, hipStream_t s, int nb = 1);
template <typename T>
hipError_t launch_efield(const T* u, T* E, size_t n, hipStream_t s);
template <typename T>
hipError_t launch_divb(const T* dudx, const T* dudy, T* divB, size_t n, hipStream_t s);
// GLM Ch of each x block (calcDt's value: the fastest speed of the block's
// last cell, timeAdvance.c:107-112) and Cp = sqrt(0.18 Ch) (divBcleaning.c:
// 7-10), for every instance: out[2 (z npe + b)] = {Ch, Cp}
template <typename T>
hipError_t launch_glm_blocks(const T* u, size_t bs_u, const DtParams& dp, const RunScalars* rs,
                             double* out, int nx, int ny, int npe, int nb, hipStream_t s);
// batched: instance z reads dudx, dudy + z * bs_d and writes divB + z * bs_o
template <typename T>
hipError_t launch_divb_b(const T* dudx, const T* dudy, T* divB, size_t n, size_t bs_d,
                         size_t bs_o, int nb, hipStream_t s);
template <typename T>
hipError_t launch_rk(T* u, const T* dudt, T* dudtRK, double cc, double dd, size_t n,
                     hipStream_t s);
template <typename T>
hipError_t launch_dt_limits(const T* u, const DtParams& dp, RunScalars* rs, int nx,
                            int ny, hipStream_t s);
hipError_t launch_dt_finalize(RunScalars* rs, double cfl, int first, hipStream_t s,
                              int nb = 1);
template <typename T>
hipError_t launch_glm_ch(const T* u, const DtParams& dp, RunScalars* rs, size_t n,
                         hipStream_t s);
hipError_t launch_dt_reset(RunScalars* rs, hipStream_t s);
template <typename T>
hipError_t launch_fill(T* p, T v, size_t n, hipStream_t s);
template <typename T>
hipError_t launch_diag(const T* u, const T* dbx, const T* dby, double* partial,
                       int nblocks, size_t n, double gamma, hipStream_t s);

hipError_t launch_copy(const void* src, void* dst, size_t bytes, hipStream_t s);
template <typename T>
hipError_t launch_add(const T* a, const T* b, T* out, size_t n, hipStream_t s);

// shape-group launches (tab: device array of one argument set per group)
template <typename T, typename TD>
hipError_t launch_ysweep_multi(const YSweepArgs<T, TD>* tab, const MultiMap& m, int mode,
                               hipStream_t s);
template <typename T, typename TI>
hipError_t launch_xsolve_multi(const XSolveArgs<T, TI>* tab, const MultiMap& m, int kind,
                               size_t dyn_lds, hipStream_t s);
template <typename T, typename TD>
hipError_t launch_rhs_rk_multi(const RhsRkArgs<T, TD>* tab, const MultiMap& m, int mode,
                               hipStream_t s);
// the fused x-derivative + RHS + RK: per group per[g] blocks per instance
// (aux[g] column tiles of XRhsShape cells / code[g] columns, times the
// chunks), code[g] = the chunk rows
template <typename T, typename TD>
hipError_t launch_xrhs_rk_multi(const XRhsArgs<T, TD>* tab, const MultiMap& m, hipStream_t s);
hipError_t launch_dt_finalize_multi(const DtMultiArgs* tab, const MultiMap& m, hipStream_t s);
// rows per y-sweep workgroup of a shape-group launch (0: ny not supported
// there); elem_bytes: the state word
int ysweep_multi_rows(int ny, int elem_bytes);
// (chunk, width) code of an x-solve in a shape-group launch (1024 threads:
// width = 1024 / (nx / chunk)), or -1 when the grid does not fit it
int xsolve_multi_code(int nx, int ny, int* chunk, int* width);

// chunk length (rows per thread) and strip width used by the x solves
// (elem_bytes: size of the input word)
void xsolve_shape(int nx, int elem_bytes, int* chunk, int* width, int* groups);
int ysweep_supported(int ny);

}  // namespace mhd2d
