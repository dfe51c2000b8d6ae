// Hand-written gfx950 kernels of the RHS / RK3 hot path.
//
//  k_xsolve   compact x-derivatives (1st, 2nd) and the compact x-filter:
//             7-point stencil + constant-matrix tridiagonal solve along x,
//             batched over (field, 16-column strip).  Each thread owns a chunk
//             of rows of one column; the Thomas recurrences are carried across
//             chunks by a segmented wavefront-shuffle scan.  The periodic
//             x-halo of the reference (parallel.c:13-24) is an index wrap.
//  k_yapply   spectral y operators (derivative, 2nd derivative, filter) on
//             two real fields packed into one complex line, radix-8/4/2
//             Stockham FFT with the line held in registers + LDS.
//  k_ysweep   fused y-sweep of the driver: pending y-filter, y-derivatives of
//             the 8 state fields, E = -u x B and its y-derivative, and the
//             CFL limits of the new state in one pass over each row.
//  pointwise  E field, div B, RHS, RK update, fused RHS + RK, dt reduction.
#include "mhd2d_fft.h"
#include "mhd2d_internal.h"

#include <cmath>
#include <type_traits>

namespace mhd2d {

// ============================================================== dt limits
// per-cell CFL limit of calcDt, timeAdvance.c:50-136 (divBcleaning==2 excluded)
// fastest signal speeds along x and y of one cell (timeAdvance.c:49-105)
__device__ __forceinline__ void cell_speeds(double rho, double ux, double uy, double bx,
                                            double by, double bz, double p,
                                            const DtParams& dp, double& cmx_out,
                                            double& cmy_out)
{
    const double cs2 = dp.gamma * p / rho;
    const double sr = sqrt(rho);
    const double cax = bx / sr, cay = by / sr, caz = bz / sr;
    const double ca2 = cax * cax + cay * cay + caz * caz;
    const double cms2 = cs2 + ca2;
    const double nsx = sqrt(fmax(cms2 * cms2 - 4 * cs2 * cax * cax, 0.));
    const double nsy = sqrt(fmax(cms2 * cms2 - 4 * cs2 * cay * cay, 0.));
    const double r2 = 1.0 / sqrt(2.);
    const double cfx = sqrt(cms2 + nsx) * r2, cfy = sqrt(cms2 + nsy) * r2;
    const double csx = sqrt(fmax(cms2 - nsx, 0.)) * r2, csy = sqrt(fmax(cms2 - nsy, 0.)) * r2;
    double cmx = fmax(fmax(fmax(fabs(ux), fabs(ux - cax)), fmax(fabs(ux - csx), fabs(ux - cfx))),
                      fmax(fmax(fabs(ux + cax), fabs(ux + csx)), fabs(ux + cfx)));
    double cmy = fmax(fmax(fmax(fabs(uy), fabs(uy - cay)), fmax(fabs(uy - csy), fabs(uy - cfy))),
                      fmax(fmax(fabs(uy + cay), fabs(uy + csy)), fabs(uy + cfy)));
    if (dp.hall == 1) {
        const double ch = dp.di / fmin(dp.dx, dp.dy) / rho * fmax(fabs(bx), fmax(fabs(by), fabs(bz)));
        cmx = fmax(cmx, ch);
        cmy = fmax(cmy, ch);
    }
    cmx_out = cmx;
    cmy_out = cmy;
}

__device__ __forceinline__ double cell_dt(double rho, double ux, double uy, double bx,
                                          double by, double bz, double p,
                                          const DtParams& dp)
{
    double cmx, cmy;
    cell_speeds(rho, ux, uy, bx, by, bz, p, dp, cmx, cmy);
    double dtx = dp.dx / cmx, dty = dp.dy / cmy;
    if (dp.resistivity > 0) {
        dtx = fmin(dtx, dp.dx * dp.dx / dp.resistivity);
        dty = fmin(dty, dp.dy * dp.dy / dp.resistivity);
    }
    if (dp.viscosity > 0) {
        dtx = fmin(dtx, dp.dx * dp.dx / dp.viscosity);
        dty = fmin(dty, dp.dy * dp.dy / dp.viscosity);
    }
    return fmin(dtx, dty);
}

// block-wide min of one value per thread, then one atomic per block.
// Positive doubles order like their bit patterns, so the running minimum is
// an integer atomicMin.  A non-finite or non-positive limit of a thread that
// holds cells (valid) raises a flag; threads without cells (the spare rows
// of a y-sweep workgroup) take no part.
__device__ __forceinline__ void block_min_dt(double v, bool valid, RunScalars* rs, double* red)
{
    if (!valid) {
        v = INFINITY;
    } else if (!(v > 0.0) || !isfinite(v)) {
        atomicOr(&rs->nonfinite, 1);
        v = INFINITY;
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmin(v, __shfl_xor(v, o));
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) red[w] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        const int nw = (blockDim.x + 63) >> 6;
        double m = red[0];
        for (int k = 1; k < nw; k++) m = fmin(m, red[k]);
        atomicMin(&rs->dtmin_bits, (unsigned long long)__double_as_longlong(m));
    }
}

// ============================================================== streaming stores
// Results that a later kernel reads once go out as nontemporal stores
// (MHD2D_NT_X / _Y / _R: the x-solves, the y-sweep, the RHS + RK).  They do
// not leave dirty lines in L2 for the next kernel to write back: measured
// per step -2.8 % (2048^2 FP64), -5.6 % (mixed), -4.3 % (4096^2 FP64).
// The exceptions are the y-sweep's filtered state (MHD2D_NT_YS): the same
// kernel reads six of its fields back to form E, and ordinary stores let
// part of that come from cache (-3.3 % y-sweep time at 2048^2 FP64); and
// the x-filter's output (MHD2D_NT_XF), which the next y-sweep reads first
// (-7 % y-sweep time; the derivative solves keep nontemporal stores, whose
// reader, the RHS + RK, measured 7 % slower after ordinary ones).  Ordinary
// stores of E (read by the x-derivative next) measured +9 % y-sweep and
// x-derivative time, of the RHS + RK's state no change.  The y-sweep's
// second y-derivatives (viscosity / resistivity, MHD2D_NT_LAP) go out
// nontemporal too: -2.1 % per step (2048^2 FP64), -0.5 % (mixed).
#ifndef MHD2D_NT_X
#define MHD2D_NT_X 1
#endif
#ifndef MHD2D_NT_Y
#define MHD2D_NT_Y 1
#endif
#ifndef MHD2D_NT_R
#define MHD2D_NT_R 1
#endif
#ifndef MHD2D_NT_LAP
#define MHD2D_NT_LAP 1
#endif
#ifndef MHD2D_NT_YS
#define MHD2D_NT_YS 0
#endif
#ifndef MHD2D_NT_XF
#define MHD2D_NT_XF 0
#endif
template <bool NT, typename T>
__device__ __forceinline__ void st_stream(T* p, T v)
{
    if constexpr (NT) __builtin_nontemporal_store(v, p);
    else *p = v;
}

// ============================================================== x solves
constexpr bool xsolve_lds_tables(int C) { return C <= 32; }
// Factor rows of a line in LDS.  With one periodic block the rows kconv ..
// nx-3 all hold the factors' FP64 fixed point, so the table keeps rows 0 ..
// K-1 (K: kconv rounded up to whole chunks of C rows), one chunk of copies of
// the fixed point and the last chunk's rows, K + 2 C rows instead of nx (at
// nx = 2048, C = 32: 96 or 160 of 2048 rows; xsolve_body maps each chunk to
// its rows).  Lines with several x blocks, and lines too short for the
// compact table to be smaller, keep all nx rows.
__host__ __device__ inline int xtable_head(int kconv, int C) { return (kconv + C - 1) / C * C; }
__host__ __device__ inline int xtable_rows(int nx, int kconv, int C, bool blk)
{
    const int K = xtable_head(kconv, C);
    return (!blk && K + 2 * C < nx) ? K + 2 * C : nx;
}
#ifndef MHD2D_X_MINB
#define MHD2D_X_MINB 1
#endif
// Workgroups are dispatched round-robin over the 8 XCDs (each with its own
// L2).  With remap the linear workgroup id is mapped so that every XCD takes
// one contiguous range of (strip, field, instance) tiles and neighbouring
// strips meet in one L2.  It pays for 8-column strips (two strips per
// 128-byte FP64 line segment: -12 % x-derivative, -17 % x-filter time at
// nx = 4096) and measured slower for 16-column strips, so that is the
// default; the autotuner may choose either.
__device__ __forceinline__ void xcd_tile(int& x, int& y, int& z, bool remap)
{
    x = blockIdx.x;
    y = blockIdx.y;
    z = blockIdx.z;
    if (remap) {
        const unsigned S = gridDim.x, V = gridDim.y;
        const unsigned T = S * V * gridDim.z;
        if ((T & 7u) == 0) {
            const unsigned L = blockIdx.x + S * (blockIdx.y + V * blockIdx.z);
            const unsigned Lp = (L & 7u) * (T >> 3) + (L >> 3);
            x = (int)(Lp % S);
            y = (int)((Lp / S) % V);
            z = (int)(Lp / (S * V));
        }
    }
}

// BLK: x is split into blocks (and may have open ends): the rows marked in
// tb.omask take explicit stencils over f[i-4 .. i+4] (tb.ocoef); otherwise
// rows 0 and nx-1 are the only end rows.
// per-chunk carries, row g of W (+1 pad: the scans read them with
// consecutive lanes one row apart, which the pad spreads over the banks);
// up to 1024 threads: 32 chunks of 32 columns, 64 of 16 or 128 of 8
template <int W>
struct XShared {
    static constexpr int WP = W + 1;
    static constexpr int GMAX = 1024 / W;
    static constexpr int WORDS = WP * GMAX;
};
constexpr int xshared_max_words()
{
    return XShared<8>::WORDS > XShared<16>::WORDS
               ? (XShared<8>::WORDS > XShared<32>::WORDS ? XShared<8>::WORDS : XShared<32>::WORDS)
               : (XShared<16>::WORDS > XShared<32>::WORDS ? XShared<16>::WORDS
                                                           : XShared<32>::WORDS);
}

// Pieces of one column's compact solve over its chunk of C rows starting at
// row s, shared by the full-line solve (k_xsolve) and the chunk re-solve of
// the fused x-derivative + RHS (k_xrhs_rk), so that both produce the same
// values bit for bit.  They (and the RHS / RK helpers below) contract a * b
// + c into an FMA only within one source expression (fp contract(on)):
// under the default cross-statement contraction the backend chooses, per
// kernel, which product of an a * b + c * d becomes the FMA, and the two
// kernels then differed in the last bits.

// rows s-3 .. s+C+2 of the column at f; only the 3-row halos can wrap in x
template <typename T, typename TI, int C>
__device__ __forceinline__ void xcol_load(T (&w)[C + 6], const TI* __restrict__ f, int s, int nx,
                                          size_t ny)
{
    const TI* __restrict__ lo = f + (size_t)(s > 0 ? s - 3 : nx - 3) * ny;
    const TI* __restrict__ hi = f + (size_t)(s + C < nx ? s + C : 0) * ny;
    const TI* __restrict__ mid = f + (size_t)s * ny;
    // A narrowing load (FP64 state into an FP32 derivative solve) first
    // subtracts the column's row-0 value in TI: the operator annihilates
    // constants, and the rounding then scales with the variation along
    // the line instead of with a large uniform background.
    TI c0 = 0;
    if constexpr (!std::is_same<T, TI>::value) c0 = f[0];
#pragma unroll
    for (int k = 0; k < 3; k++) w[k] = (T)(lo[k * ny] - c0);
#pragma unroll
    for (int k = 0; k < C; k++) w[k + 3] = (T)(mid[k * ny] - c0);
#pragma unroll
    for (int k = 0; k < 3; k++) w[C + 3 + k] = (T)(hi[k * ny] - c0);
}

// the same rows for a filter input whose chunk edge rows (first and last 3
// of every chunk of C rows) are in the edge buffer e (the column's word of
// row 0 of chunk 0 there, [G][6][ny]) and the others in f
template <typename T, int C>
__device__ __forceinline__ void xcol_load_edge(T (&w)[C + 6], const T* __restrict__ f,
                                               const T* __restrict__ e, int g, int G, size_t ny)
{
    const int s = g * C;
    const T* __restrict__ lo = e + ((size_t)(g > 0 ? g - 1 : G - 1) * 6 + 3) * ny;
    const T* __restrict__ own = e + (size_t)g * 6 * ny;
    const T* __restrict__ hi = e + (size_t)(g + 1 < G ? g + 1 : 0) * 6 * ny;
    const T* __restrict__ mid = f + (size_t)s * ny;
#pragma unroll
    for (int k = 0; k < 3; k++) w[k] = lo[k * ny];
#pragma unroll
    for (int k = 0; k < 3; k++) w[k + 3] = own[k * ny];
#pragma unroll
    for (int k = 3; k < C - 3; k++) w[k + 3] = mid[k * ny];
#pragma unroll
    for (int k = 0; k < 3; k++) w[C + k] = own[(3 + k) * ny];
#pragma unroll
    for (int k = 0; k < 3; k++) w[C + 3 + k] = hi[k * ny];
}

// right-hand side of chunk g in place: w[k] = r at row s + k.  first / last:
// the chunk holds row 0 / row N-1 of a periodic line (explicit end rows).
// BLK: x is split into blocks (and may have open ends): the rows marked in
// tb.omask take explicit stencils over f[i-4 .. i+4] (tb.ocoef); otherwise
// rows 0 and nx-1 are the only end rows.
template <typename T, int C, int KIND, bool BLK>
__device__ __forceinline__ void xcol_rhs(T (&w)[C + 6], const XStencil& st, const XTables& tb,
                                         int g, bool first, bool last)
{
#pragma clang fp contract(on)
    const T c0 = (T)st.c0, c1 = (T)st.c1, c2 = (T)st.c2, c3 = (T)st.c3;
    unsigned long long om = 0;
    int oi = 0;
    if constexpr (BLK) {
        om = tb.omask[g];
        oi = tb.ofirst[g];
    }
    T saved = 0;  // f at row - 4 of the current row (w[k - 1] before it is overwritten)
    T rfirst = 0, rlast = 0;
    if (first)   // row 0 over f[-2..2]
        rfirst = (T)st.e[0] * w[1] + (T)st.e[1] * w[2] + (T)st.e[2] * w[3] +
                 (T)st.e[3] * w[4] + (T)st.e[4] * w[5];
    if (last)    // row N-1 over f[N+1..N-3]
        rlast = (T)st.esign * ((T)st.e[0] * w[C + 4] + (T)st.e[1] * w[C + 3] +
                               (T)st.e[2] * w[C + 2] + (T)st.e[3] * w[C + 1] +
                               (T)st.e[4] * w[C]);
#pragma unroll
    for (int k = 0; k < C; k++) {
        const T fm3 = w[k], fm2 = w[k + 1], fm1 = w[k + 2], f0 = w[k + 3];
        const T fp1 = w[k + 4], fp2 = w[k + 5], fp3 = w[k + 6];
        T r;
        if constexpr (KIND == XD1) {
            r = c1 * (fp1 - fm1) + c2 * (fp2 - fm2) + c3 * (fp3 - fm3);
        } else if constexpr (KIND == XD2) {
            r = c1 * (fp1 - 2 * f0 + fm1) + c2 * (fp2 - 2 * f0 + fm2) + c3 * (fp3 - 2 * f0 + fm3);
        } else {
            r = c0 * f0 + c1 * (fp1 + fm1) + c2 * (fp2 + fm2) + c3 * (fp3 + fm3);
        }
        if constexpr (BLK) {
            // block edges and open-end rows: explicit stencil over f[-4 .. 4]
            // (f[-4] exists from the second row of the chunk on, f[+4] up to
            // the one before last; the tables put no weight where it does not)
            if ((om >> k) & 1ull) {
                const double* __restrict__ cf = tb.ocoef + 9 * oi;
                T acc = (T)cf[1] * fm3 + (T)cf[2] * fm2 + (T)cf[3] * fm1 + (T)cf[4] * f0 +
                        (T)cf[5] * fp1 + (T)cf[6] * fp2 + (T)cf[7] * fp3;
                if (k >= 1) acc += (T)cf[0] * saved;
                if (k + 7 <= C + 5) acc += (T)cf[8] * w[(k + 7 <= C + 5) ? k + 7 : 0];
                r = acc;
                oi++;
            }
            saved = fm3;
        }
        w[k] = r;  // w[k] is not read again by later rows
    }
    if (first) w[0] = rfirst;
    if (last) w[C - 1] = rlast;
}

// forward recurrence y_i = r_i - l_i y_{i-1} over the chunk from the value
// y entering it, in place
template <typename T, typename FT, int C>
__device__ __forceinline__ void xcol_forward(T (&w)[C + 6], T y, bool interior, T linf,
                                             const FT* __restrict__ L)
{
#pragma clang fp contract(on)
    if (interior) {
#pragma unroll
        for (int k = 0; k < C; k++) {
            y = w[k] - linf * y;
            w[k] = y;
        }
    } else {
#pragma unroll
        for (int k = 0; k < C; k++) {
            y = w[k] - (T)L[k] * y;
            w[k] = y;
        }
    }
}

// backward recurrence x_i = y_i invd_i + q_i x_{i+1} over the chunk from
// the value x entering it from above, in place
template <typename T, typename FT, int C>
__device__ __forceinline__ void xcol_backward(T (&w)[C + 6], T x, bool interior, T dinf, T qinf,
                                              const FT* __restrict__ ID, const FT* __restrict__ Q)
{
#pragma clang fp contract(on)
    if (interior) {
#pragma unroll
        for (int k = C - 1; k >= 0; k--) {
            x = w[k] * dinf + qinf * x;
            w[k] = x;
        }
    } else {
#pragma unroll
        for (int k = C - 1; k >= 0; k--) {
            x = w[k] * (T)ID[k] + (T)Q[k] * x;
            w[k] = x;
        }
    }
}

// Rows from tb.kconv to N-3 carry the factors' exact FP64 fixed point, so a
// chunk inside that range (never the last chunk) uses the scalar constants (no loads).  (FP32 only:
// in FP64 the second code path costs more registers than the table loads it
// saves)
#ifndef MHD2D_X_CONST
#define MHD2D_X_CONST(T) (sizeof(T) == 4)
#endif


// One workgroup's compact solves: column strip bx, field v, instance z of
// the launch described by `a`.  sh_in / sh_fw / sh_bw hold at least
// XShared<W>::WORDS words, sh_half W words; x_dyn the factor tables (3 nx
// words) when xsolve_lds_tables(C).
template <typename T, typename TI, int C, int W, int KIND, bool BLK>
__device__ __forceinline__ void xsolve_body(const XSolveArgs<T, TI>& a, int bx, int v, int z,
                                            T* __restrict__ sh_in, T* __restrict__ sh_fw,
                                            T* __restrict__ sh_bw, T* __restrict__ sh_half,
                                            unsigned char* x_dyn)
{
#pragma clang fp contract(on)
    if (a.rs && a.rs[z].stop) return;  // the run stopped (dt < 1e-8): uniform per block
    constexpr bool CARRY = KIND == XD1_CARRY, EDGE = KIND == XFILT_EDGE;
    constexpr int SK = CARRY ? XD1 : (EDGE ? XFILT : KIND);  // the stencil
    const int G = blockDim.x / W;
    const int c = threadIdx.x % W;
    const int g = threadIdx.x / W;
    const int nx = a.nx;
    const size_t ny = (size_t)a.ny;
    const size_t j = (size_t)bx * W + c;
    const TI* __restrict__ f = a.in[v] + z * a.bs_in + j;
    T* __restrict__ o = CARRY ? nullptr : a.out[v] + z * a.bs_out + j;
    const int s = g * C;

    constexpr int WP = XShared<W>::WP;

    T w[C + 6];
    if constexpr (EDGE)
        xcol_load_edge<T, C>(w, f, a.edge + z * a.bs_edge + v * a.edge_fs + j, g, G, ny);
    else
        xcol_load<T, TI, C>(w, f, s, nx, ny);

    // Factor rows.  With chunks of up to 32 rows (every line length now) the
    // three tables of the line go to LDS (xtable_rows words each, dynamic)
    // and the recurrences read them as broadcasts instead of 6 C vector loads per
    // thread (measured -8 % x-solve time at nx = 2048, -11 % at 512, -8 % at
    // 4096 with 128 chunks).  The fill is
    // issued right behind the line's loads, so both wait out one memory
    // latency (-1.5 % x-solve time at nx = 2048); the barrier before the
    // recurrences publishes it.
    using FT = typename std::conditional<xsolve_lds_tables(C), T, double>::type;
    const FT* L;
    const FT* ID;
    const FT* Q;
    if constexpr (xsolve_lds_tables(C)) {
        const int TR = xtable_rows(nx, a.tb.kconv, C, BLK);
        const int K = xtable_head(a.tb.kconv, C);
        const bool compact = TR < nx;
        T* const tl = reinterpret_cast<T*>(x_dyn);
        T* const tid = tl + TR;
        T* const tqq = tid + TR;
        for (int i = threadIdx.x; i < TR; i += blockDim.x) {
            // row of entry i: the head rows, the fixed point (row nx - 3,
            // equal bit for bit to rows kconv .. nx - 3), the last chunk
            const int r = (!compact || i < K) ? i : (i < K + C ? nx - 3 : i + nx - 2 * C - K);
            tl[i] = (T)a.tb.l[r];
            tid[i] = (T)a.tb.invd[r];
            tqq[i] = (T)a.tb.q[r];
        }
        const int base = !compact ? s : (g == G - 1 ? K + C : (s >= K ? K : s));
        L = tl + base;
        ID = tid + base;
        Q = tqq + base;
    } else {
        L = a.tb.l + s;
        ID = a.tb.invd + s;
        Q = a.tb.q + s;
    }

    const bool first = !BLK && (g == 0), last = !BLK && (g == G - 1);
    xcol_rhs<T, C, SK, BLK>(w, a.st, a.tb, g, first, last);

    if constexpr (xsolve_lds_tables(C)) __syncthreads();  // the tables are in LDS
    const bool interior = MHD2D_X_CONST(T) && !BLK && (s >= a.tb.kconv) && !last;
    const T linf = (T)a.tb.linf, dinf = (T)a.tb.dinf, qinf = (T)a.tb.qinf;

    // ---- forward recurrence y_i = r_i - l_i y_{i-1}
    T y = 0;
    if (interior) {
#pragma unroll
        for (int k = 0; k < C; k++) y = w[k] - linf * y;
    } else {
#pragma unroll
        for (int k = 0; k < C; k++) y = w[k] - (T)L[k] * y;
    }
    // The chunk scans run on wavefront shuffles over up to 64 chunks; with
    // 128 chunks (8-column strips of 4096 rows) each half is scanned in its
    // own wavefront and the upper half then starts from the lower half's end.
    const int SW = G > 64 ? 64 : G;
    sh_in[g * WP + c] = y;
    __syncthreads();
    {
        const int c2i = threadIdx.x / G, g2 = threadIdx.x % G;
        const int gl = g2 % SW;
        T B = sh_in[g2 * WP + c2i];
        T A = (T)a.tb.pf[g2];
        for (int d = 1; d < SW; d <<= 1) {
            const T Ap = __shfl_up(A, d, SW);
            const T Bp = __shfl_up(B, d, SW);
            if (gl >= d) {
                B = B + A * Bp;
                A = A * Ap;
            }
        }
        if (G > 64) {
            if (g2 == 63) sh_half[c2i] = B;
            __syncthreads();
            if (g2 >= 64) B = B + A * sh_half[c2i];
        }
        sh_fw[g2 * WP + c2i] = B;  // y at the last row of chunk g2
    }
    __syncthreads();
    y = (g > 0) ? sh_fw[(g - 1) * WP + c] : (T)0;
    if (interior) {
#pragma unroll
        for (int k = 0; k < C; k++) {
            y = w[k] - linf * y;
            w[k] = y;
        }
    } else {
#pragma unroll
        for (int k = 0; k < C; k++) {
            y = w[k] - (T)L[k] * y;
            w[k] = y;
        }
    }

    // ---- backward recurrence x_i = y_i invd_i + q_i x_{i+1}
    T x = 0;
    if (interior) {
#pragma unroll
        for (int k = C - 1; k >= 0; k--) x = w[k] * dinf + qinf * x;
    } else {
#pragma unroll
        for (int k = C - 1; k >= 0; k--) x = w[k] * (T)ID[k] + (T)Q[k] * x;
    }
    sh_in[g * WP + c] = x;
    __syncthreads();
    {
        const int c2i = threadIdx.x / G, g2 = G - 1 - (threadIdx.x % G);
        const int gr = threadIdx.x % G;  // position in the reversed scan
        const int gl = gr % SW;
        T B = sh_in[g2 * WP + c2i];
        T A = (T)a.tb.qb[g2];
        for (int d = 1; d < SW; d <<= 1) {
            const T Ap = __shfl_up(A, d, SW);
            const T Bp = __shfl_up(B, d, SW);
            if (gl >= d) {
                B = B + A * Bp;
                A = A * Ap;
            }
        }
        if (G > 64) {
            if (gr == 63) sh_half[c2i] = B;
            __syncthreads();
            if (gr >= 64) B = B + A * sh_half[c2i];
        }
        sh_bw[g2 * WP + c2i] = B;  // x at the first row of chunk g2
    }
    __syncthreads();
    x = (g < G - 1) ? sh_bw[(g + 1) * WP + c] : (T)0;
    if constexpr (CARRY) {
        // the two values entering chunk g, [z][v][2][G][ny]
        T* __restrict__ cy = a.carry + z * a.bs_carry + ((size_t)v * 2 * G + g) * ny + j;
        cy[0] = (g > 0) ? sh_fw[(g - 1) * WP + c] : (T)0;
        cy[(size_t)G * ny] = x;
        return;
    }
    // (the same recurrences as xcol_forward / xcol_backward, written out
    // here: through the helpers the compiler allocates 72 instead of 64
    // VGPRs to the FP32 16-row solve, which then fits one workgroup per CU
    // instead of two)
    if (interior) {
#pragma unroll
        for (int k = C - 1; k >= 0; k--) {
            x = w[k] * dinf + qinf * x;
            w[k] = x;
        }
    } else {
#pragma unroll
        for (int k = C - 1; k >= 0; k--) {
            x = w[k] * (T)ID[k] + (T)Q[k] * x;
            w[k] = x;
        }
    }
    T* __restrict__ po = o + (size_t)s * ny;
    // row segments of at least 64 bytes; narrower ones (8 FP32 columns)
    // measured 1.5x slower as nontemporal stores
    constexpr bool NT = (SK == XFILT ? MHD2D_NT_XF : MHD2D_NT_X) && W * sizeof(T) >= 64;
    if (a.accumulate) {
        // second x-derivative onto the second y-derivative already there
        // (addSourceTerm's d2/dx2 + d2/dy2, calcRHS.c:498-531)
#pragma unroll
        for (int k = 0; k < C; k++) {
            const T yy = po[k * ny];
            st_stream<NT>(po + k * ny, w[k] + yy);
        }
    } else {
#pragma unroll
        for (int k = 0; k < C; k++) st_stream<NT>(po + k * ny, w[k]);
    }
}

template <typename T, typename TI, int C, int W, int KIND, bool BLK>
__global__ void __launch_bounds__(1024, MHD2D_X_MINB) k_xsolve(XSolveArgs<T, TI> a)
{
    int bx, v, z;  // column strip, field, instance of a batch
    xcd_tile(bx, v, z, a.remap > 0);
    __shared__ T sh_in[XShared<W>::WORDS];
    __shared__ T sh_fw[XShared<W>::WORDS];
    __shared__ T sh_bw[XShared<W>::WORDS];
    __shared__ T sh_half[W];  // a column's value at the end of its lower 64 chunks
    extern __shared__ unsigned char x_dyn[];
    xsolve_body<T, TI, C, W, KIND, BLK>(a, bx, v, z, sh_in, sh_fw, sh_bw, sh_half, x_dyn);
}

#ifndef MHD2D_X_W
#define MHD2D_X_W 16
#endif

void xsolve_shape(int nx, int elem_bytes, int* chunk, int* width, int* groups)
{
    // 32 chunks per line, of 8 to 32 rows (64-row ones need ~170 VGPRs in
    // FP64, one 512-thread workgroup per CU): longer lines take more chunks
    // of 32 rows (4096 rows: 128 chunks of 8 columns).  Against 64 chunks
    // of 8 rows, 512-row lines in 16-row chunks read 22 instead of 28 rows
    // (with the 3-row halos) per 16 and run 512-thread workgroups, two to a
    // CU: 64 x 512^2 +6 % FP64, +5 % mixed, +19 % FP32 per step; 1024-row
    // lines in 32-row chunks: uneven ensemble +2.6 % FP64, unchanged mixed
    // and FP32
    int C = nx / 32;
    if (C < 8) C = 8;
    if (C > 32) C = 32;
    const int G = nx / C;
    int W = MHD2D_X_W;
    // 128 chunks: 8-column strips keep 1024 threads per block.  Lines of
    // <= 512 rows read as FP32 (elem_bytes: the input word) measured faster
    // in 8-column strips too (64 x 512^2 FP32: -5 %); with FP64 inputs they
    // take 16 columns (128-byte rows: -3 % FP64, -6 % mixed).
    if (G > 64 || (nx <= 512 && elem_bytes == 4)) W = 8;
    *chunk = C;
    *width = W;
    *groups = G;
}

template <typename T, typename TI, int C, int W, bool BLK>
static void launch_xsolve_kind(const XSolveArgs<T, TI>& a, dim3 grid, dim3 block, hipStream_t s)
{
    const size_t shm =
        xsolve_lds_tables(C) ? 3 * (size_t)xtable_rows(a.nx, a.tb.kconv, C, BLK) * sizeof(T) : 0;
    if constexpr (!BLK) {
        if (a.carry) {
            hipLaunchKernelGGL((k_xsolve<T, TI, C, W, XD1_CARRY, false>), grid, block, shm, s, a);
            return;
        }
        if constexpr (std::is_same<T, TI>::value) {
            if (a.edge) {
                hipLaunchKernelGGL((k_xsolve<T, TI, C, W, XFILT_EDGE, false>), grid, block, shm, s,
                                   a);
                return;
            }
        }
    }
    if constexpr (std::is_same<T, TI>::value) {
        switch (a.kind) {
        case XD1: hipLaunchKernelGGL((k_xsolve<T, TI, C, W, XD1, BLK>), grid, block, shm, s, a); break;
        case XD2: hipLaunchKernelGGL((k_xsolve<T, TI, C, W, XD2, BLK>), grid, block, shm, s, a); break;
        default: hipLaunchKernelGGL((k_xsolve<T, TI, C, W, XFILT, BLK>), grid, block, shm, s, a); break;
        }
    } else {
        // converting inputs: derivatives of a FP64 state (both annihilate
        // the subtracted column constant; the filter does not)
        if (a.kind == XD2)
            hipLaunchKernelGGL((k_xsolve<T, TI, C, W, XD2, BLK>), grid, block, shm, s, a);
        else
            hipLaunchKernelGGL((k_xsolve<T, TI, C, W, XD1, BLK>), grid, block, shm, s, a);
    }
}

template <typename T, typename TI, int C, int W>
static hipError_t launch_xsolve_cw(const XSolveArgs<T, TI>& a, int nv, hipStream_t s, int nb)
{
    const int G = a.nx / C;
    dim3 grid(a.ny / W, nv, nb), block(W * G);
    if (a.tb.omask)
        launch_xsolve_kind<T, TI, C, W, true>(a, grid, block, s);
    else
        launch_xsolve_kind<T, TI, C, W, false>(a, grid, block, s);
    return hipGetLastError();
}

template <typename T, typename TI>
hipError_t launch_xsolve(const XSolveArgs<T, TI>& a, int nv, hipStream_t s, int nb)
{
    if (nb < 1 || nb > 65535) return hipErrorInvalidValue;
    // converting inputs: derivatives only (they annihilate the subtracted
    // constant; the filter does not)
    if (!std::is_same<T, TI>::value && a.kind == XFILT) return hipErrorInvalidValue;
    // carries: first derivatives of one periodic block, nothing accumulated;
    // edge rows: filters of one periodic block
    if (a.carry && (a.kind != XD1 || a.tb.omask || a.accumulate)) return hipErrorInvalidValue;
    if (a.edge && (a.kind != XFILT || a.tb.omask || a.accumulate)) return hipErrorInvalidValue;
    int C, W, G;
    xsolve_shape(a.nx, sizeof(TI), &C, &W, &G);
    if (a.width) {
        // autotuned width: 8, 16 or 32 columns, at most 1024 threads (128
        // chunks only fit 8, 64 chunks 16)
        if ((a.width != 8 && a.width != 16 && a.width != 32) || a.width * G > 1024)
            return hipErrorInvalidValue;
        W = a.width;
    }
    if (a.nx % C || G > 128 || G < 2 || W * G > 1024 || a.ny % W || nv < 1 || nv > 16)
        return hipErrorInvalidValue;
    XSolveArgs<T, TI> b = a;
    b.remap = a.remap < 0 ? (W == 8) : a.remap;
    if (W == 8) {
        switch (C) {
        case 8: return launch_xsolve_cw<T, TI, 8, 8>(b, nv, s, nb);
        case 16: return launch_xsolve_cw<T, TI, 16, 8>(b, nv, s, nb);
        case 32: return launch_xsolve_cw<T, TI, 32, 8>(b, nv, s, nb);
        default: return hipErrorInvalidValue;
        }
    }
    if (W == 32) {
        switch (C) {
        case 8: return launch_xsolve_cw<T, TI, 8, 32>(b, nv, s, nb);
        case 16: return launch_xsolve_cw<T, TI, 16, 32>(b, nv, s, nb);
        case 32: return launch_xsolve_cw<T, TI, 32, 32>(b, nv, s, nb);
        default: return hipErrorInvalidValue;
        }
    }
    switch (C) {
    case 8: return launch_xsolve_cw<T, TI, 8, 16>(b, nv, s, nb);
    case 16: return launch_xsolve_cw<T, TI, 16, 16>(b, nv, s, nb);
    case 32: return launch_xsolve_cw<T, TI, 32, 16>(b, nv, s, nb);
    default: return hipErrorInvalidValue;
    }
}

// ============================================================== y operators
// elements per thread of the row FFTs.  FP64 lines up to 2048 points: 8
// (a line plus its spectrum within ~100 VGPRs, 4 waves per SIMD; 4 elements
// and 6 waves measured +30 % y-sweep time at 2048^2, 16 elements and 2
// waves +11 %); 4096-point FP64 lines: 16 (2 LDS exchanges per transform
// instead of 3, at 2 waves per SIMD: -6.5 % y-sweep time at 4096^2).
// FP32 lines take 16.
#ifndef MHD2D_Y_E64
#define MHD2D_Y_E64 8
#endif
#ifndef MHD2D_Y_E64_4096
#define MHD2D_Y_E64_4096 16
#endif
template <typename T, int N>
struct YElems {
    static constexpr int E = sizeof(T) == 8 ? (N >= 4096 ? MHD2D_Y_E64_4096 : MHD2D_Y_E64) : 16;
};

// occupancy request for the fused y-sweep: 4 waves per SIMD (128 VGPRs, a
// few spills) measured -17 % y-sweep time at 4096^2 and neutral at 2048^2
// with 8 elements per thread; 2 waves (256 VGPRs) with 16 FP64 elements
#ifndef MHD2D_Y_WAVES
#define MHD2D_Y_WAVES 4
#endif
template <typename T, int N>
constexpr int ysweep_waves()
{
    return (sizeof(T) == 8 && YElems<T, N>::E == 16) ? 2 : MHD2D_Y_WAVES;
}

// threads of a y-sweep workgroup for lines shorter than that many threads
// (several rows per workgroup): 256 measured within 1.4 % of 128 on the
// batched and uneven ensembles and the FP32 2048^2 run
#ifndef MHD2D_Y_THREADS
#define MHD2D_Y_THREADS 128
#endif
#ifndef MHD2D_Y_SMALL_GRID
#define MHD2D_Y_SMALL_GRID 1024
#endif
template <typename T, int N, int THR = MHD2D_Y_THREADS>
struct YShape {
    static constexpr int E = YElems<T, N>::E;
    static constexpr int TPR = FFTShape<N, E>::TPR;
    static constexpr int ROWS = (TPR >= THR) ? 1 : THR / TPR;
    static constexpr int THREADS = TPR * ROWS;
    static constexpr int LDSW = FFTShape<N, E>::LDS;
};

int ysweep_supported(int ny)
{
    return ny >= 16 && ny <= 4096 && (ny & (ny - 1)) == 0;
}

// generic spectral multiplier on nv fields (drop-in derivY / derivYY / filteringY)
template <typename T, int N>
__global__ void __launch_bounds__((YShape<T, N>::THREADS)) k_yapply(YApplyArgs<T> a)
{
    using Y = YShape<T, N>;
    constexpr int TPR = Y::TPR;
    constexpr int E = Y::E;
    __shared__ T s_re[Y::ROWS * Y::LDSW];
    __shared__ T s_im[Y::ROWS * Y::LDSW];
    const int t = threadIdx.x % TPR, rl = threadIdx.x / TPR;
    const int i_raw = blockIdx.x * Y::ROWS + rl;
    const bool active = i_raw < a.nx;
    const size_t row = (size_t)(active ? i_raw : a.nx - 1) * N;
    const size_t zi = blockIdx.y * a.bs_in, zo = blockIdx.y * a.bs_out;  // instance
    T* lre = s_re + rl * Y::LDSW;
    T* lim = s_im + rl * Y::LDSW;
    const double* __restrict__ mul =
        a.mode == Y_D1 ? a.tb.md : (a.mode == Y_D2 ? a.tb.mdd : a.tb.mf);
    for (int p = 0; p < a.nv; p += 2) {
        const bool two = (p + 1) < a.nv;
        const T* fa = a.in[p] + zi + row;
        const T* fb = two ? a.in[p + 1] + zi + row : fa;
        cx<T> v[E];
#pragma unroll
        for (int m = 0; m < E; m++) {
            const int k = t + m * TPR;
            v[m] = cx<T>{fa[k], two ? fb[k] : (T)0};
        }
        fft_fwd<T, N, E>(v, t, lre, lim, a.tb.tw);
#pragma unroll
        for (int m = 0; m < E; m++) {
            const T s = (T)mul[t + m * TPR];
            v[m] = (a.mode == Y_D1) ? mul_i_scal(v[m], s) : mul_scal(v[m], s);
        }
        fft_inv<T, N, E>(v, t, lre, lim, a.tb.tw);
        if (active) {
            T* oa = a.out[p] + zo + row;
#pragma unroll
            for (int m = 0; m < E; m++) oa[t + m * TPR] = v[m].re;
            if (two) {
                T* ob = a.out[p + 1] + zo + row;
#pragma unroll
                for (int m = 0; m < E; m++) ob[t + m * TPR] = v[m].im;
            }
        }
    }
}

// Fused y-sweep of one RK stage (driver).  Each thread owns the same 16
// positions of its row in physical and spectral space, so the filtered state
// it writes is re-read by the same thread only (E and the CFL limits).
// MODE 0: ideal; 1: with the second derivatives of the source terms; 2: also
// the GLM psi line (separate instances, so each keeps its own registers)
// One workgroup of the y-sweep: rows bx * ROWS .. of instance z; s_re and
// s_im hold ROWS * LDSW words each, red 16 doubles.
template <typename T, typename TD, int N, int MODE, int THR = MHD2D_Y_THREADS>
__device__ __forceinline__ void ysweep_body(const YSweepArgs<T, TD>& a, int bx, int z,
                                            T* __restrict__ s_re, T* __restrict__ s_im,
                                            double* red)
{
    constexpr bool LAP = MODE >= 1, GLM = MODE == 2;
    using Y = YShape<T, N, THR>;
    constexpr int TPR = Y::TPR;
    constexpr int E = Y::E;
    RunScalars* const rs = a.rs + z;
    if (rs->stop) return;  // the run stopped (dt < 1e-8): uniform per block
    const int t = threadIdx.x % TPR, rl = threadIdx.x / TPR;
    const int i_raw = bx * Y::ROWS + rl;
    const bool active = i_raw < a.nx;
    const size_t row = (size_t)(active ? i_raw : a.nx - 1) * N;
    const size_t n = (size_t)a.nx * N;
    T* lre = s_re + rl * Y::LDSW;
    T* lim = s_im + rl * Y::LDSW;
    const double2* __restrict__ tw = a.tb.tw;
    T* const u = a.u + z * a.bs_u;
    T* const Ef = a.E + z * a.bs_u;
    TD* const dudy = a.dudy + z * a.bs_d;
    TD* const dEdy = a.dEdy + z * a.bs_d;

    // field pairs packed as re/im of one complex line; the velocity and field
    // components share pairs so that the second derivatives of the source
    // terms take at most three more inverse transforms, and they come last so
    // that E finds them as recent as possible in cache; the GLM psi
    // (slot PSI_SLOT) is a fifth line of its own, in the GLM instance only,
    // so the others keep four full pairs and their register budgets
    constexpr int PA[5] = {RHO, UX, UZ, BX, PSI_SLOT};
    constexpr int PB[5] = {PR, UY, BZ, BY, -1};
    constexpr int npairs = GLM ? 5 : 4;
    // Ideal MHD (MODE 0): the last pair, (Bx, By), runs outside the loop and
    // leaves its state values (filtered when the sweep filters, as loaded
    // otherwise) in nxt, whose registers the next-pair load no longer needs;
    // E takes them from there and reads four fields back instead of six.
    // Its spectrum is saved in Z and differentiated there, so no more values
    // are live than in the loop.  (The CFL loop, unrolled by 4 only, reads
    // them back: a runtime index would move nxt to scratch memory.)
    constexpr bool KEEPB = MODE == 0;
    static_assert(!KEEPB || (PA[3] == BX && PB[3] == BY && npairs == 4),
                  "the kept pair is (Bx, By), the last one");
    cx<T> nxt[E];
    if (a.filter || a.derivs) {
        // The next pair is loaded while the last transform of the current
        // pair runs (its registers replace the saved spectrum Z), so the HBM
        // latency of every pair but the first is hidden behind FFT work.
#pragma unroll
        for (int m = 0; m < E; m++) {
            const int k = t + m * TPR;
            nxt[m] = cx<T>{u[PA[0] * n + row + k], u[PB[0] * n + row + k]};
        }
        TD* const lapz = a.lap ? a.lap + z * a.bs_l : nullptr;
        // one field pair; NEXT: load the next pair during the last transform
        // (a compile-time flag, so the loop body has no conditional load)
        auto pair = [&](int p, auto next_tag) {
            constexpr bool NEXT = decltype(next_tag)::value;
            const int fa = PA[p], fb = PB[p];
            const bool two = !GLM || fb >= 0;
            T* ua = u + fa * n + row;
            T* ub = u + (two ? fb : fa) * n + row;
            // second y-derivatives wanted from this pair (re: fa, im: fb)
            const bool la = LAP && a.derivs &&
                            ((a.lap_u && (fa == UX || fa == UZ)) || (a.lap_b && fa == BX));
            const bool lb = LAP && a.derivs &&
                            ((a.lap_u && fb == UY) || (a.lap_b && (fb == BZ || fb == BY)));
            // opaque copy of t: the multiplier loads stay inside the loop
            // instead of being hoisted into 24 live registers
            int tq = t;
            asm volatile("" : "+v"(tq));
            cx<T> v[E];
#pragma unroll
            for (int m = 0; m < E; m++) v[m] = nxt[m];
            fft_fwd<T, N, E>(v, t, lre, lim, tw);
            cx<T> Z[E];
            if (LAP || a.filter) {
#pragma unroll
                for (int m = 0; m < E; m++) Z[m] = v[m];
            }
            if (a.filter) {
#pragma unroll
                for (int m = 0; m < E; m++) v[m] = mul_scal(Z[m], (T)a.tb.mf[tq + m * TPR]);
                fft_inv<T, N, E>(v, t, lre, lim, tw);
                if (active) {
#pragma unroll
                    for (int m = 0; m < E; m++) {
                        const int k = t + m * TPR;
                        st_stream<MHD2D_NT_YS>(ua + k, v[m].re);
                        if (two) st_stream<MHD2D_NT_YS>(ub + k, v[m].im);
                    }
                }
                if (!LAP && a.derivs) {
#pragma unroll
                    for (int m = 0; m < E; m++)
                        v[m] = mul_i_scal(Z[m], (T)a.tb.mfd[tq + m * TPR]);
                }
            } else if (!LAP) {
#pragma unroll
                for (int m = 0; m < E; m++) v[m] = mul_i_scal(v[m], (T)a.tb.md[tq + m * TPR]);
            }
            if (LAP && (la || lb)) {
                // d2/dy2 of the filtered fields (derivYY, calcDerivs.c:132-140)
                const double* __restrict__ m2 = a.filter ? a.tb.mfdd : a.tb.mdd;
#pragma unroll
                for (int m = 0; m < E; m++) v[m] = mul_scal(Z[m], (T)m2[tq + m * TPR]);
                fft_inv<T, N, E>(v, t, lre, lim, tw);
                if (active) {
                    TD* pa = lapz + (fa - UX) * n + row;
                    TD* pb = lapz + (fb - UX) * n + row;
#pragma unroll
                    for (int m = 0; m < E; m++) {
                        const int k = t + m * TPR;
                        if (la) st_stream<MHD2D_NT_LAP>(pa + k, (TD)v[m].re);
                        if (lb) st_stream<MHD2D_NT_LAP>(pb + k, (TD)v[m].im);
                    }
                }
            }
            if (LAP && a.derivs) {
                const double* __restrict__ m1 = a.filter ? a.tb.mfd : a.tb.md;
#pragma unroll
                for (int m = 0; m < E; m++) v[m] = mul_i_scal(Z[m], (T)m1[tq + m * TPR]);
            }
            if (NEXT) {
                const T* na = u + PA[p + 1] * n + row;
                const bool two_n = !GLM || PB[p + 1] >= 0;
                const T* nb = u + (two_n ? PB[p + 1] : PA[p + 1]) * n + row;
#pragma unroll
                for (int m = 0; m < E; m++) {
                    const int k = t + m * TPR;
                    nxt[m] = cx<T>{na[k], two_n ? nb[k] : (T)0};
                }
            }
            if (a.derivs) {
                fft_inv<T, N, E>(v, t, lre, lim, tw);
                if (active) {
                    TD* da = dudy + fa * n + row;
                    TD* db = dudy + (two ? fb : fa) * n + row;
#pragma unroll
                    for (int m = 0; m < E; m++) {
                        const int k = t + m * TPR;
                        st_stream<MHD2D_NT_Y>(da + k, (TD)v[m].re);
                        if (two) st_stream<MHD2D_NT_Y>(db + k, (TD)v[m].im);
                    }
                }
            }
        };
#pragma unroll 1
        for (int p = 0; p < npairs - 1; p++) pair(p, std::true_type{});
        // the last pair: with the Laplacian terms or the GLM line, the loop
        // body without the next-pair load (ideal MHD keeps its (Bx, By)
        // values, below)
        if constexpr (!KEEPB) pair(npairs - 1, std::false_type{});
        if constexpr (KEEPB) {
            // the last pair (Bx, By): the same operations as the loop body
            // without the Laplacian terms and with no next pair to load
            T* const ua = u + BX * n + row;
            T* const ub = u + BY * n + row;
            int tq = t;
            asm volatile("" : "+v"(tq));
            cx<T> Z[E];
#pragma unroll
            for (int m = 0; m < E; m++) Z[m] = nxt[m];
            fft_fwd<T, N, E>(Z, t, lre, lim, tw);
            if (a.filter) {
                cx<T> F[E];
#pragma unroll
                for (int m = 0; m < E; m++) F[m] = mul_scal(Z[m], (T)a.tb.mf[tq + m * TPR]);
                fft_inv<T, N, E>(F, t, lre, lim, tw);
#pragma unroll
                for (int m = 0; m < E; m++) nxt[m] = F[m];
                if (active) {
#pragma unroll
                    for (int m = 0; m < E; m++) {
                        const int k = t + m * TPR;
                        st_stream<MHD2D_NT_YS>(ua + k, F[m].re);
                        st_stream<MHD2D_NT_YS>(ub + k, F[m].im);
                    }
                }
            }
            if (a.derivs) {
                const double* __restrict__ m1 = a.filter ? a.tb.mfd : a.tb.md;
#pragma unroll
                for (int m = 0; m < E; m++) Z[m] = mul_i_scal(Z[m], (T)m1[tq + m * TPR]);
                fft_inv<T, N, E>(Z, t, lre, lim, tw);
                if (active) {
                    TD* const da = dudy + BX * n + row;
                    TD* const db = dudy + BY * n + row;
#pragma unroll
                    for (int m = 0; m < E; m++) {
                        const int k = t + m * TPR;
                        st_stream<MHD2D_NT_Y>(da + k, (TD)Z[m].re);
                        st_stream<MHD2D_NT_Y>(db + k, (TD)Z[m].im);
                    }
                }
            }
        }
    }

    if (a.dt_limits) {
        // CFL limits of the filtered state (stage 1; timeAdvance.c:50-136),
        // ahead of E: E's other inputs are then read back from cache
        DtParams dp = a.dp;
        dp.di = rs->di;
        double mn = INFINITY;
        if (active) {
#pragma unroll 4
            for (int m = 0; m < E; m++) {
                const size_t k = row + t + m * TPR;
                mn = fmin(mn, cell_dt(u[RHO * n + k], u[UX * n + k], u[UY * n + k],
                                      u[BX * n + k], u[BY * n + k], u[BZ * n + k],
                                      u[PR * n + k], dp));
            }
        }
        block_min_dt(mn, active, rs, red);
    }

    if (a.derivs) {
        // E = -u x B of the filtered state (calcRHS.c:24-43).  The induction
        // terms take dEz/dy, dEz/dx, dEx/dy and dEy/dx (calcRHS.c:135-137):
        // Ey and Ez are stored for the x-solves, and (Ex, Ez) share one
        // complex line for d/dy; Ex itself and dEy/dy are never needed.
        cx<T> v[E];
#pragma unroll
        for (int m = 0; m < E; m++) {
            const size_t k = row + t + m * TPR;
            const T ux = u[UX * n + k], uy = u[UY * n + k], uz = u[UZ * n + k];
            T bx, by;
            if constexpr (KEEPB) {
                bx = nxt[m].re;
                by = nxt[m].im;
            } else {
                bx = u[BX * n + k];
                by = u[BY * n + k];
            }
            const T bz = u[BZ * n + k];
            const T ex = uz * by - uy * bz;
            const T ey = ux * bz - uz * bx;
            const T ez = uy * bx - ux * by;
            v[m] = cx<T>{ex, ez};
            if (active) {
                st_stream<MHD2D_NT_Y>(Ef + 0 * n + k, ey);
                st_stream<MHD2D_NT_Y>(Ef + 1 * n + k, ez);
            }
        }
        fft_fwd<T, N, E>(v, t, lre, lim, tw);
#pragma unroll
        for (int m = 0; m < E; m++) v[m] = mul_i_scal(v[m], (T)a.tb.md[t + m * TPR]);
        fft_inv<T, N, E>(v, t, lre, lim, tw);
        if (active) {
#pragma unroll
            for (int m = 0; m < E; m++) {
                const size_t k = row + t + m * TPR;
                st_stream<MHD2D_NT_Y>(dEdy + 0 * n + k, (TD)v[m].re);
                st_stream<MHD2D_NT_Y>(dEdy + 1 * n + k, (TD)v[m].im);
            }
        }
    }
}

template <typename T, typename TD, int N, int MODE, int THR>
__global__ void __launch_bounds__((YShape<T, N, THR>::THREADS))
    __attribute__((amdgpu_waves_per_eu(ysweep_waves<T, N>()))) k_ysweep(YSweepArgs<T, TD> a)
{
    using Y = YShape<T, N, THR>;
    __shared__ T s_re[Y::ROWS * Y::LDSW];
    __shared__ T s_im[Y::ROWS * Y::LDSW];
    __shared__ double red[16];
    // instance of a batch, taken last to first: the x-filter before this
    // sweep ends on the last instance, whose data the 256 MB Infinity Cache
    // still holds (with the same order in k_rhs_rk, -4 % step time for
    // 2 x 2048^2 and 8 x 1024^2 batches; the results do not depend on it)
    const int z = (int)(gridDim.y - 1 - blockIdx.y);
    ysweep_body<T, TD, N, MODE, THR>(a, blockIdx.x, z, s_re, s_im, red);
}

template <typename T, typename TD, int N>
static hipError_t launch_y_n(const YApplyArgs<T>* ya, const YSweepArgs<T, TD>* ys, hipStream_t s,
                             int nb)
{
    const int nx = ya ? ya->nx : ys->nx;
    using Y = YShape<T, N>;
    dim3 grid((nx + Y::ROWS - 1) / Y::ROWS, nb), block(Y::THREADS);
    if constexpr (std::is_same<T, TD>::value) {
        if (ya) {
            hipLaunchKernelGGL((k_yapply<T, N>), grid, block, 0, s, *ya);
            return hipGetLastError();
        }
    }
    // Short lines on a small grid (fewer than MHD2D_Y_SMALL_GRID workgroups
    // of MHD2D_Y_THREADS threads): 64-thread workgroups, half the rows each,
    // twice as many to spread over the CUs (256^2: -9 % step time, the
    // y-sweep 0.031 -> 0.027 ms; larger grids keep the wider workgroups:
    // the 64 x 512^2 batch took +3 % with 64 threads).  Each row's
    // arithmetic is the same, so are the results.
    using Ys = YShape<T, N, 64>;
    if constexpr (N >= 128 && Ys::ROWS != Y::ROWS) {
        if ((size_t)grid.x * nb < MHD2D_Y_SMALL_GRID) {
            const dim3 gs((nx + Ys::ROWS - 1) / Ys::ROWS, nb), bs(Ys::THREADS);
            if (ys->glm)
                hipLaunchKernelGGL((k_ysweep<T, TD, N, 2, 64>), gs, bs, 0, s, *ys);
            else if (ys->lap)
                hipLaunchKernelGGL((k_ysweep<T, TD, N, 1, 64>), gs, bs, 0, s, *ys);
            else
                hipLaunchKernelGGL((k_ysweep<T, TD, N, 0, 64>), gs, bs, 0, s, *ys);
            return hipGetLastError();
        }
    }
    if (ys->glm)
        hipLaunchKernelGGL((k_ysweep<T, TD, N, 2, MHD2D_Y_THREADS>), grid, block, 0, s, *ys);
    else if (ys->lap)
        hipLaunchKernelGGL((k_ysweep<T, TD, N, 1, MHD2D_Y_THREADS>), grid, block, 0, s, *ys);
    else
        hipLaunchKernelGGL((k_ysweep<T, TD, N, 0, MHD2D_Y_THREADS>), grid, block, 0, s, *ys);
    return hipGetLastError();
}

template <typename T, typename TD = T>
static hipError_t launch_y(int ny, const YApplyArgs<T>* ya, const YSweepArgs<T, TD>* ys,
                           hipStream_t s, int nb = 1)
{
    if (nb < 1 || nb > 65535) return hipErrorInvalidValue;
    switch (ny) {
    case 16: return launch_y_n<T, TD, 16>(ya, ys, s, nb);
    case 32: return launch_y_n<T, TD, 32>(ya, ys, s, nb);
    case 64: return launch_y_n<T, TD, 64>(ya, ys, s, nb);
    case 128: return launch_y_n<T, TD, 128>(ya, ys, s, nb);
    case 256: return launch_y_n<T, TD, 256>(ya, ys, s, nb);
    case 512: return launch_y_n<T, TD, 512>(ya, ys, s, nb);
    case 1024: return launch_y_n<T, TD, 1024>(ya, ys, s, nb);
    case 2048: return launch_y_n<T, TD, 2048>(ya, ys, s, nb);
    case 4096: return launch_y_n<T, TD, 4096>(ya, ys, s, nb);
    default: return hipErrorInvalidValue;
    }
}

template <typename T>
hipError_t launch_yapply(const YApplyArgs<T>& a, hipStream_t s, int nb)
{
    if (a.nv < 1 || a.nv > 16) return hipErrorInvalidValue;
    return launch_y<T, T>(a.ny, &a, nullptr, s, nb);
}

template <typename T, typename TD>
hipError_t launch_ysweep(const YSweepArgs<T, TD>& a, hipStream_t s, int nb)
{
    return launch_y<T, TD>(a.ny, nullptr, &a, s, nb);
}

// ============================================================== pointwise
static inline dim3 pw_grid(size_t n, int block)
{
    size_t b = (n + block - 1) / block;
    if (b > 65535u * 16u) b = 65535u * 16u;
    return dim3((unsigned)b);
}

template <typename T>
__global__ void k_efield(const T* __restrict__ u, T* __restrict__ E, size_t n)
{
    for (size_t p = blockIdx.x * (size_t)blockDim.x + threadIdx.x; p < n;
         p += (size_t)gridDim.x * blockDim.x) {
        const T ux = u[UX * n + p], uy = u[UY * n + p], uz = u[UZ * n + p];
        const T bx = u[BX * n + p], by = u[BY * n + p], bz = u[BZ * n + p];
        E[0 * n + p] = uz * by - uy * bz;
        E[1 * n + p] = ux * bz - uz * bx;
        E[2 * n + p] = uy * bx - ux * by;
    }
}

template <typename T>
__global__ void k_divb(const T* __restrict__ dudx, const T* __restrict__ dudy,
                       T* __restrict__ divB, size_t n)
{
    for (size_t p = blockIdx.x * (size_t)blockDim.x + threadIdx.x; p < n;
         p += (size_t)gridDim.x * blockDim.x)
        divB[p] = dudx[BX * n + p] + dudy[BY * n + p];
}

// calcDivB (calcRHS.c:11-21) for instance blockIdx.y of a batch
template <typename T>
__global__ void k_divb_b(const T* __restrict__ dudx, const T* __restrict__ dudy,
                         T* __restrict__ divB, size_t n, size_t bs_d, size_t bs_o)
{
    const T* const dx = dudx + blockIdx.y * bs_d;
    const T* const dy = dudy + blockIdx.y * bs_d;
    T* const o = divB + blockIdx.y * bs_o;
    for (size_t p = blockIdx.x * (size_t)blockDim.x + threadIdx.x; p < n;
         p += (size_t)gridDim.x * blockDim.x)
        o[p] = dx[BX * n + p] + dy[BY * n + p];
}

// ideal-MHD right-hand side at one cell, calcRHS.c:85-142
template <typename T>
__device__ __forceinline__ void rhs_cell(const T* __restrict__ u, const T* __restrict__ dx,
                                         const T* __restrict__ dy, const T* __restrict__ dEx,
                                         const T* __restrict__ dEy, size_t n, size_t p,
                                         T gamma, T (&R)[8], T& rho_out)
{
    const T rho = u[RHO * n + p], ux = u[UX * n + p], uy = u[UY * n + p];
    const T bx = u[BX * n + p], by = u[BY * n + p], bz = u[BZ * n + p], pr = u[PR * n + p];
    const T divU = dx[UX * n + p] + dy[UY * n + p];
    const T jx = dy[BZ * n + p];
    const T jy = -dx[BZ * n + p];
    const T jz = dx[BY * n + p] - dy[BX * n + p];
    const T ir = (T)1 / rho;
    R[RHO] = -rho * divU - ux * dx[RHO * n + p] - uy * dy[RHO * n + p];
    R[UX] = -ux * dx[UX * n + p] - uy * dy[UX * n + p] - dx[PR * n + p] * ir + jy * bz * ir -
            jz * by * ir;
    R[UY] = -ux * dx[UY * n + p] - uy * dy[UY * n + p] - dy[PR * n + p] * ir + jz * bx * ir -
            jx * bz * ir;
    R[UZ] = -ux * dx[UZ * n + p] - uy * dy[UZ * n + p] + jx * by * ir - jy * bx * ir;
    R[BX] = -dEy[2 * n + p];
    R[BY] = dEx[2 * n + p];
    R[BZ] = dEy[0 * n + p] - dEx[1 * n + p];
    R[PR] = -ux * dx[PR * n + p] - uy * dy[PR * n + p] - gamma * pr * divU;
    rho_out = rho;
}

// 256-thread launches: without the bound the compiler assumed 1024 threads
// and capped the kernel at 128 VGPRs (6 spilled)
template <typename T>
__global__ void __launch_bounds__(256) k_rhs(RhsArgs<T> a)
{
    const size_t n = a.n;
    for (size_t p = blockIdx.x * (size_t)blockDim.x + threadIdx.x; p < n;
         p += (size_t)gridDim.x * blockDim.x) {
        T R[8], rho, rpsi = 0;
        rhs_cell<T>(a.u, a.dudx, a.dudy, a.dEdx, a.dEdy, n, p, (T)a.gamma, R, rho);
        if (a.divbclean == 2) {
            // hyperbolic cleaning, Dedner et al. 2002 eq. 24 (calcRHS.c:144-158)
            const T divB = a.dudx[BX * n + p] + a.dudy[BY * n + p];
            const T ch = (T)a.ch, cp = (T)a.cp;
            R[UX] = R[UX] - divB * a.u[BX * n + p];
            R[UY] = R[UY] - divB * a.u[BY * n + p];
            R[UZ] = R[UZ] - divB * a.u[BZ * n + p];
            R[BX] = R[BX] - a.dudx[8 * n + p];
            R[BY] = R[BY] - a.dudy[8 * n + p];
            rpsi = -ch * ch * (divB + a.u[8 * n + p] / cp / cp);
        }
        // addSourceTerm, calcRHS.c:498-544
        if (a.viscosity > 0) {
            for (int c = 0; c < 3; c++)
                R[UX + c] += (T)a.viscosity / rho * (a.d2udx2[c * n + p] + a.d2udy2[c * n + p]);
        }
        if (a.resistivity > 0) {
            for (int c = 0; c < 3; c++)
                R[BX + c] += (T)a.resistivity * (a.d2Bdx2[c * n + p] + a.d2Bdy2[c * n + p]);
        }
        if (a.divbclean == 1) {
            R[BX] += (T)a.constDiffDivB * a.ddivBdx[p];
            R[BY] += (T)a.constDiffDivB * a.ddivBdy[p];
        }
#pragma unroll
        for (int f = 0; f < 8; f++) a.dudt[f * n + p] = R[f];
        if (a.divbclean == 2) a.dudt[8 * n + p] = rpsi;
    }
}

// x block (MPI rank of decomposeGrid1D, initialize.c:547-569) of row i
__device__ __forceinline__ int x_block_of(int i, int nx, int npe)
{
    const int base = nx / npe, rem = nx % npe;
    const int big = rem * (base + 1);
    return i < big ? i / (base + 1) : rem + (i - big) / base;
}

// Ch of every x block of every instance: calcDt on each rank leaves the
// fastest speed of its own last cell (timeAdvance.c:107-112), and
// calc_Cp_fix_Cr sets Cp = sqrt(Cr Ch) with Cr = 0.18 (divBcleaning.c:4-10)
template <typename T>
__global__ void k_glm_blocks(const T* __restrict__ u, size_t bs_u, DtParams dp,
                             const RunScalars* __restrict__ rs, double* __restrict__ out,
                             int nx, int ny, int npe)
{
    const int z = blockIdx.x;
    const T* const uz = u + z * bs_u;
    const size_t n = (size_t)nx * ny;
    dp.di = rs[z].di;
    for (int b = threadIdx.x; b < npe; b += blockDim.x) {
        const int base = nx / npe, rem = nx % npe;
        const int off = b * base + (b < rem ? b : rem), N = base + (b < rem ? 1 : 0);
        const size_t p = (size_t)(off + N - 1) * ny + (ny - 1);
        double cmx, cmy;
        cell_speeds(uz[RHO * n + p], uz[UX * n + p], uz[UY * n + p], uz[BX * n + p],
                    uz[BY * n + p], uz[BZ * n + p], uz[PR * n + p], dp, cmx, cmy);
        const double ch = fmax(cmx, cmy);
        out[2 * ((size_t)z * npe + b)] = ch;
        out[2 * ((size_t)z * npe + b) + 1] = sqrt(0.18 * ch);
    }
}

// Ch of calcDt's divBcleaning == 2 branch (timeAdvance.c:107-112): the value
// the cell loop leaves, i.e. max(cmaxx, cmaxy) of the last cell (n - 1)
template <typename T>
__global__ void k_glm_ch(const T* __restrict__ u, DtParams dp, RunScalars* rs, size_t n)
{
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    const size_t p = n - 1;
    double cmx, cmy;
    cell_speeds(u[RHO * n + p], u[UX * n + p], u[UY * n + p], u[BX * n + p], u[BY * n + p],
                u[BZ * n + p], u[PR * n + p], dp, cmx, cmy);
    rs->glm_ch = fmax(cmx, cmy);
}

// W consecutive cells of one field, loaded and stored as one vector
template <typename T, int W>
struct alignas(sizeof(T) * W) Pack {
    T x[W];
};

template <int W, typename T>
__device__ __forceinline__ Pack<T, W> ldp(const T* __restrict__ p)
{
    return *reinterpret_cast<const Pack<T, W>*>(p);
}

// streaming inputs of the RHS + RK (derivatives and R_prev, read once per
// stage) as nontemporal loads (MHD2D_NT_LOAD): measured -15 % RHS + RK time
// at 2048^2 FP64 (it then streams at the copy kernel's rate), -13 % mixed,
// -7 % at 4096^2.  The same for the x-solves' rows and the y-sweep's field
// pairs measured no gain, and for the state itself a loss (the x-filter that
// follows reads it back 12 % slower).
#ifndef MHD2D_NT_LOAD
#define MHD2D_NT_LOAD 1
#endif

template <int W, typename T>
__device__ __forceinline__ Pack<T, W> ldn(const T* __restrict__ p)
{
    if constexpr (MHD2D_NT_LOAD) {
        typedef T V __attribute__((ext_vector_type(W)));
        const V x = __builtin_nontemporal_load(reinterpret_cast<const V*>(p));
        Pack<T, W> r;
#pragma unroll
        for (int w = 0; w < W; w++) r.x[w] = x[w];
        return r;
    } else {
        return ldp<W>(p);
    }
}

template <bool NT, int W, typename T>
__device__ __forceinline__ void stp(T* p, const Pack<T, W>& v)
{
    if constexpr (NT) {
        typedef T V __attribute__((ext_vector_type(W)));
        V x;
#pragma unroll
        for (int w = 0; w < W; w++) x[w] = v.x[w];
        __builtin_nontemporal_store(x, reinterpret_cast<V*>(p));
    } else {
        *reinterpret_cast<Pack<T, W>*>(p) = v;
    }
}

// calcRHSInnerPoint (calcRHS.c:85-142) at one cell of ideal MHD: R of the 8
// fields from the state uc, its x- and y-derivatives d_x, d_y (state slots;
// d_x[BX] and d_y[BY] are not read), dEy/dx, dEz/dx (ex0, ex1) and dEx/dy,
// dEz/dy (ey0, ey1)
template <typename T>
__device__ __forceinline__ void rhs_point(const T (&uc)[8], const T (&d_x)[8], const T (&d_y)[8],
                                          T ex0, T ex1, T ey0, T ey1, T g, T (&R)[8])
{
#pragma clang fp contract(on)
    const T rho = uc[RHO], ux = uc[UX], uy = uc[UY];
    const T bx = uc[BX], by = uc[BY], bz = uc[BZ], pr = uc[PR];
    const T divU = d_x[UX] + d_y[UY];
    const T jx = d_y[BZ];
    const T jy = -d_x[BZ];
    const T jz = d_x[BY] - d_y[BX];
    const T ir = (T)1 / rho;
    R[RHO] = -rho * divU - ux * d_x[RHO] - uy * d_y[RHO];
    R[UX] = -ux * d_x[UX] - uy * d_y[UX] - d_x[PR] * ir + jy * bz * ir - jz * by * ir;
    R[UY] = -ux * d_x[UY] - uy * d_y[UY] - d_y[PR] * ir + jz * bx * ir - jx * bz * ir;
    R[UZ] = -ux * d_x[UZ] - uy * d_y[UZ] + jx * by * ir - jy * bx * ir;
    // calcRHS.c:135-137
    R[BX] = -ey1;
    R[BY] = ex1;
    R[BZ] = ey0 - ex0;
    R[PR] = -ux * d_x[PR] - uy * d_y[PR] - g * pr * divU;
}

// rkExecute (timeAdvance.c:31-35): u + cc R + dd R_prev (R_prev is zero in
// stage 0)
template <typename T>
__device__ __forceinline__ T rk_update(T u, T R, T rp, T cc, T dd)
{
#pragma clang fp contract(on)
    return u + cc * R + dd * rp;
}

// fused RHS + RK update; each thread handles W = 16 / sizeof(T) consecutive
// cells (one 16-byte vector of each state field) per field (n is a multiple
// of W: nx, ny are powers of two).  Derivatives and R_prev are read and
// written as TD, the arithmetic and the state update run in T.  MODE 1 adds
// the viscous and resistive source terms and the diffusive cleaning, MODE 2
// also the GLM terms (separate instances, so each keeps its own register
// allocation).  The FP64 instance is held at 2
// waves per SIMD (measured -2.5 % against the 3 its registers allow; the
// mixed instance measured +7 % at 2, so the others keep the compiler's
// choice).
template <typename T, typename TD>
struct RhsRkWaves {
    static constexpr int MAX = (sizeof(T) == 8 && sizeof(TD) == 8) ? 2 : 10;
};
// RHS + RK over the cells of instance z with blocks bid = 0 .. nblk-1
template <typename T, typename TD, int MODE>
__device__ __forceinline__ void rhs_rk_body(const RhsRkArgs<T, TD>& a, int z, int bid, int nblk)
{
    constexpr bool LAP = MODE >= 1, GLM = MODE == 2;
    constexpr int W = 16 / sizeof(T);
    const RunScalars* const rs = a.rs + z;
    if (rs->stop) return;  // the run stopped (dt < 1e-8)
    const size_t n = a.n;
    const T cc = (T)rs->cc[a.stage];
    const T dd = (T)rs->dd[a.stage];
    T* const uz = a.u + z * a.bs_u;
    TD* const rpz = a.rprev + z * a.bs_r;
    const TD* const dxz = a.dudx + z * a.bs_d;
    const TD* const dyz = a.dudy + z * a.bs_d;
    const TD* const exz = a.dEdx + z * a.bs_d;
    const TD* const eyz = a.dEdy + z * a.bs_d;
    const T g = (T)a.gamma;
    const bool use_prev = a.stage > 0;
    // with open x ends the first and last rows belong to the boundary kernel
    const size_t q0 = a.open_x ? (size_t)a.ny / W : 0;
    const size_t q1 = a.open_x ? (n - (size_t)a.ny) / W : n / W;
    for (size_t q = q0 + (size_t)bid * blockDim.x + threadIdx.x; q < q1;
         q += (size_t)nblk * blockDim.x) {
        const size_t p = q * W;
        Pack<T, W> u[8];
        Pack<TD, W> dx[8], dy[8], ex[2], ey[2];
#pragma unroll
        for (int f = 0; f < 8; f++) {
            u[f] = ldp<W>(uz + f * n + p);
            dx[f] = ldn<W>(dxz + f * n + p);
            dy[f] = ldn<W>(dyz + f * n + p);
        }
#pragma unroll
        for (int f = 0; f < 2; f++) {
            ex[f] = ldn<W>(exz + f * n + p);
            ey[f] = ldn<W>(eyz + f * n + p);
        }
        Pack<T, W> R[8];
#pragma unroll
        for (int w = 0; w < W; w++) {
            T uc[8], d_x[8], d_y[8], Rc[8];
#pragma unroll
            for (int f = 0; f < 8; f++) {
                uc[f] = u[f].x[w];
                d_x[f] = (T)dx[f].x[w];
                d_y[f] = (T)dy[f].x[w];
            }
            rhs_point<T>(uc, d_x, d_y, (T)ex[0].x[w], (T)ex[1].x[w], (T)ey[0].x[w],
                         (T)ey[1].x[w], g, Rc);
#pragma unroll
            for (int f = 0; f < 8; f++) R[f].x[w] = Rc[f];
        }
        Pack<T, W> rpsi;
        if (GLM) {
            // hyperbolic cleaning, Dedner et al. 2002 eq. 24 (calcRHS.c:144-158),
            // with the Ch, Cp of this row's x block
            const int blk = x_block_of((int)(p / (size_t)a.ny), a.nx, a.npe);
            const double* const gc = a.glm_cp + 2 * ((size_t)z * a.npe + blk);
            const T ch = (T)gc[0], cp = (T)gc[1];
            const Pack<TD, W> px = ldp<W>(dxz + PSI_SLOT * n + p);
            const Pack<TD, W> py = ldp<W>(dyz + PSI_SLOT * n + p);
            const Pack<T, W> ps = ldp<W>(uz + PSI_SLOT * n + p);
#pragma unroll
            for (int w = 0; w < W; w++) {
                const T divB = (T)dx[BX].x[w] + (T)dy[BY].x[w];
                R[UX].x[w] = R[UX].x[w] - divB * u[BX].x[w];
                R[UY].x[w] = R[UY].x[w] - divB * u[BY].x[w];
                R[UZ].x[w] = R[UZ].x[w] - divB * u[BZ].x[w];
                R[BX].x[w] = R[BX].x[w] - (T)px.x[w];
                R[BY].x[w] = R[BY].x[w] - (T)py.x[w];
                rpsi.x[w] = -ch * ch * (divB + ps.x[w] / cp / cp);
            }
        }
        if (LAP && a.lap) {
            // addSourceTerm (calcRHS.c:498-531): lap holds d2/dx2 + d2/dy2
            const TD* const lz = a.lap + z * a.bs_l;
            if (a.viscosity > 0) {
#pragma unroll
                for (int c = 0; c < 3; c++) {
                    const Pack<TD, W> l = ldp<W>(lz + c * n + p);
#pragma unroll
                    for (int w = 0; w < W; w++)
                        R[UX + c].x[w] = R[UX + c].x[w] +
                                         (T)a.viscosity / u[RHO].x[w] * (T)l.x[w];
                }
            }
            if (a.resistivity > 0) {
#pragma unroll
                for (int c = 0; c < 3; c++) {
                    const Pack<TD, W> l = ldp<W>(lz + (3 + c) * n + p);
#pragma unroll
                    for (int w = 0; w < W; w++)
                        R[BX + c].x[w] = R[BX + c].x[w] + (T)a.resistivity * (T)l.x[w];
                }
            }
        }
        if (LAP && a.ddivb) {
            // diffusive div B cleaning (calcRHS.c:533-543)
            const TD* const gz = a.ddivb + z * a.bs_b;
#pragma unroll
            for (int c = 0; c < 2; c++) {
                const Pack<TD, W> gd = ldp<W>(gz + c * n + p);
#pragma unroll
                for (int w = 0; w < W; w++)
                    R[BX + c].x[w] = R[BX + c].x[w] + (T)a.cdiff * (T)gd.x[w];
            }
        }
#pragma unroll
        for (int f = 0; f < 8; f++) {
            // rkExecute, timeAdvance.c:31-35 (R_prev is zero in stage 0)
            Pack<TD, W> rp;
            if (use_prev) rp = ldn<W>(rpz + f * n + p);
            Pack<T, W> o;
            Pack<TD, W> rs_out;
#pragma unroll
            for (int w = 0; w < W; w++) {
                o.x[w] = rk_update<T>(u[f].x[w], R[f].x[w], use_prev ? (T)rp.x[w] : (T)0, cc, dd);
                rs_out.x[w] = (TD)R[f].x[w];
            }
            stp<MHD2D_NT_R, W>(uz + f * n + p, o);
            if (a.write_rprev) stp<MHD2D_NT_R, W>(rpz + f * n + p, rs_out);
        }
        if (GLM) {
            // psi: field 8 of R_prev, slot PSI_SLOT of the state
            const Pack<T, W> ps = ldp<W>(uz + PSI_SLOT * n + p);
            Pack<TD, W> rp;
            if (use_prev) rp = ldp<W>(rpz + 8 * n + p);
            Pack<T, W> o;
            Pack<TD, W> rs_out;
#pragma unroll
            for (int w = 0; w < W; w++) {
                o.x[w] = ps.x[w] + cc * rpsi.x[w] + dd * (use_prev ? (T)rp.x[w] : (T)0);
                rs_out.x[w] = (TD)rpsi.x[w];
            }
            stp<MHD2D_NT_R, W>(uz + PSI_SLOT * n + p, o);
            if (a.write_rprev) stp<MHD2D_NT_R, W>(rpz + 8 * n + p, rs_out);
        }
    }
}

template <typename T, typename TD, int MODE>
__global__ void __launch_bounds__(256)
    __attribute__((amdgpu_waves_per_eu(1, RhsRkWaves<T, TD>::MAX))) k_rhs_rk(RhsRkArgs<T, TD> a)
{
    // instance of a batch, last to first like k_ysweep (the x solves between
    // them take the instances first to last), so that the next kernel
    // starts on the instance this one touched last
    const int z = (int)(gridDim.y - 1 - blockIdx.y);
    rhs_rk_body<T, TD, MODE>(a, z, blockIdx.x, gridDim.x);
}

// ------------------------------------------ fused x-derivative + RHS + RK
// Ideal MHD on one periodic x block.  A workgroup takes one chunk of C rows
// (the chunks of the full x-solve) by WB columns of one instance.  It
// re-solves the x-derivatives of the nine fields the RHS reads (in[0..8] of
// the derivative arguments: rho, ux, uy, uz, By, Bz, p, Ey, Ez) on its chunk
// from the carries k_xsolve<..., XD1_CARRY> stored, with the recurrences of
// the full solve (so the values are the same bit for bit), into LDS; then
// it applies the RHS + RK update to the chunk's cells.  The new state of
// the chunk's first and last 3 rows, which the chunks above and below read
// as halos meanwhile, goes to a small edge buffer (6 rows per chunk), the
// rest back in place; the x-filter reads the edge rows from there.  The
// derivatives are neither written to nor read back from HBM (18 words per
// cell and stage in FP64).
template <typename T>
struct XRhsShape {
    static constexpr int CELLS = sizeof(T) == 8 ? 512 : 1024;  // cells per workgroup
    static constexpr int PW = 16 / (int)sizeof(T);             // cells per thread: 256 threads
    static constexpr int NV = 9;                               // derivative fields
};

template <typename T>
constexpr int xrhs_fs(int C)
{
    // LDS words per field: the pad puts the fields that one wavefront writes
    // together (same column, same row) on different banks
    return XRhsShape<T>::CELLS + XRhsShape<T>::CELLS / C;
}

// One workgroup of the fused kernel: column tile bx, chunk g, instance z;
// sdx holds at least XRhsShape<TD>::NV * xrhs_fs<T>(C) words (16-byte
// aligned), stab 3 C.  EARLY: the RHS phase's streaming inputs are requested
// before the chunk solves.
template <typename T, typename TD, int C, bool EARLY>
__device__ __forceinline__ void xrhs_body(const XRhsArgs<T, TD>& a, int bx, int g, int z,
                                          TD* __restrict__ sdx, TD* __restrict__ stab)
{
    using S = XRhsShape<T>;
    constexpr int PW = S::PW, NV = S::NV;
    constexpr int WB = S::CELLS / C;          // columns of the chunk
    constexpr int NI = NV * WB;               // column solves
    constexpr int IPT = (NI + 255) / 256;     // column solves per thread
    constexpr int FS = xrhs_fs<T>(C);
    const XSolveArgs<TD, T>& xa = a.x;
    const RhsRkArgs<T, TD>& ra = a.r;
    const RunScalars* const rs = ra.rs + z;
    if (rs->stop) return;  // the run stopped (dt < 1e-8): uniform per block
    const int nx = xa.nx;
    const size_t ny = (size_t)xa.ny;
    const int G = nx / C;
    const int s = g * C;
    const bool first = g == 0, last = g == G - 1;

    // this thread's cells in the RHS + RK phase: PW consecutive cells of
    // row kr of the chunk
    const int kr = threadIdx.x / (WB / PW);
    const int col = (threadIdx.x % (WB / PW)) * PW;
    const size_t n = ra.n;
    const size_t p = (size_t)(s + kr) * ny + (size_t)bx * WB + col;
    T* const uz = ra.u + z * ra.bs_u;
    TD* const rpz = ra.rprev + z * ra.bs_r;
    const TD* const dyz = ra.dudy + z * ra.bs_d;
    const TD* const eyz = ra.dEdy + z * ra.bs_d;
    const bool use_prev = ra.stage > 0;
    Pack<T, PW> u[8];
    Pack<TD, PW> dy[8], ey[2], rp[8];
    // EARLY: the RHS phase's streaming inputs (y-derivatives, R_prev) are
    // requested before the chunk solves, so that their latency overlaps the
    // solves' own loads and arithmetic (MHD2D_XRHS_PREFETCH 2: the state as
    // well)
#ifndef MHD2D_XRHS_PREFETCH
#define MHD2D_XRHS_PREFETCH 0
#endif
    auto load_streams = [&]() {
#pragma unroll
        for (int f = 0; f < 8; f++) dy[f] = ldn<PW>(dyz + f * n + p);
#pragma unroll
        for (int f = 0; f < 2; f++) ey[f] = ldn<PW>(eyz + f * n + p);
        if (use_prev) {
#pragma unroll
            for (int f = 0; f < 8; f++) rp[f] = ldn<PW>(rpz + f * n + p);
        }
    };
    auto load_state = [&]() {
#pragma unroll
        for (int f = 0; f < 8; f++) u[f] = ldp<PW>(uz + f * n + p);
    };
    constexpr bool PF1 = EARLY || MHD2D_XRHS_PREFETCH >= 1;
    if constexpr (PF1) load_streams();
    if constexpr (MHD2D_XRHS_PREFETCH >= 2) load_state();

    // the chunk's rows (with halos) and its two carries, for every column
    // solve of this thread; then the factor rows of the chunk
    TD w[IPT][C + 6];
    TD yin[IPT], xin[IPT];
#pragma unroll
    for (int i = 0; i < IPT; i++) {
        const int it = threadIdx.x + 256 * i;
        if (it < NI) {
            const int v = it / WB, c = it % WB;
            const size_t j = (size_t)bx * WB + c;
            xcol_load<TD, T, C>(w[i], xa.in[v] + z * xa.bs_in + j, s, nx, ny);
            const TD* __restrict__ cy =
                a.carry + z * a.bs_carry + ((size_t)v * 2 * G + g) * ny + j;
            yin[i] = cy[0];
            xin[i] = cy[(size_t)G * ny];
        }
    }
    if (threadIdx.x < C) {
        stab[threadIdx.x] = (TD)xa.tb.l[s + threadIdx.x];
        stab[C + threadIdx.x] = (TD)xa.tb.invd[s + threadIdx.x];
        stab[2 * C + threadIdx.x] = (TD)xa.tb.q[s + threadIdx.x];
    }
    __syncthreads();
    const bool interior = MHD2D_X_CONST(TD) && (s >= xa.tb.kconv) && !last;
    const TD linf = (TD)xa.tb.linf, dinf = (TD)xa.tb.dinf, qinf = (TD)xa.tb.qinf;
#pragma unroll
    for (int i = 0; i < IPT; i++) {
        const int it = threadIdx.x + 256 * i;
        if (it < NI) {
            const int v = it / WB, c = it % WB;
            xcol_rhs<TD, C, XD1, false>(w[i], xa.st, xa.tb, g, first, last);
            xcol_forward<TD, TD, C>(w[i], yin[i], interior, linf, stab);
            xcol_backward<TD, TD, C>(w[i], xin[i], interior, dinf, qinf, stab + C, stab + 2 * C);
            TD* const d = sdx + v * FS + c;
#pragma unroll
            for (int k = 0; k < C; k++) d[k * WB] = w[i][k];
        }
    }
    __syncthreads();

    // RHS + RK
    if constexpr (!PF1) load_streams();
    if constexpr (MHD2D_XRHS_PREFETCH < 2) load_state();
    const T cc = (T)rs->cc[ra.stage];
    const T dd = (T)rs->dd[ra.stage];
    // this row's destination: the edge buffer for the first and last 3 rows
    // of the chunk (other chunks read them as halos), else the state in place
    T* uo;
    size_t po, fs;
    if (kr < 3 || kr >= C - 3) {
        uo = a.edge + z * a.bs_edge;
        po = ((size_t)g * 6 + (kr < 3 ? kr : kr - C + 6)) * ny + (size_t)bx * WB + col;
        fs = a.edge_fs;
    } else {
        uo = uz;
        po = p;
        fs = n;
    }
    const T g_ = (T)ra.gamma;
    Pack<TD, PW> dxv[NV];
#pragma unroll
    for (int v = 0; v < NV; v++)
        dxv[v] = *reinterpret_cast<const Pack<TD, PW>*>(sdx + v * FS + kr * WB + col);
    // derivative field v -> state slot (Bx is not differentiated)
    constexpr int SLOT[7] = {RHO, UX, UY, UZ, BY, BZ, PR};
    Pack<T, PW> R[8];
#pragma unroll
    for (int w = 0; w < PW; w++) {
        T uc[8], d_x[8], d_y[8], Rc[8];
#pragma unroll
        for (int f = 0; f < 8; f++) {
            uc[f] = u[f].x[w];
            d_x[f] = (T)0;
            d_y[f] = (T)dy[f].x[w];
        }
#pragma unroll
        for (int v = 0; v < 7; v++) d_x[SLOT[v]] = (T)dxv[v].x[w];
        rhs_point<T>(uc, d_x, d_y, (T)dxv[7].x[w], (T)dxv[8].x[w], (T)ey[0].x[w], (T)ey[1].x[w],
                     g_, Rc);
#pragma unroll
        for (int f = 0; f < 8; f++) R[f].x[w] = Rc[f];
    }
#pragma unroll
    for (int f = 0; f < 8; f++) {
        Pack<T, PW> o;
        Pack<TD, PW> rs_out;
#pragma unroll
        for (int w = 0; w < PW; w++) {
            o.x[w] = rk_update<T>(u[f].x[w], R[f].x[w], use_prev ? (T)rp[f].x[w] : (T)0, cc, dd);
            rs_out.x[w] = (TD)R[f].x[w];
        }
        stp<MHD2D_NT_R, PW>(uo + f * fs + po, o);
        if (ra.write_rprev) stp<MHD2D_NT_R, PW>(rpz + f * n + p, rs_out);
    }
}

template <typename T, typename TD, int C>
__global__ void __launch_bounds__(256) k_xrhs_rk(XRhsArgs<T, TD> a)
{
    // read back as PW-word vectors (16 or 8 bytes)
    __shared__ __align__(16) TD sdx[XRhsShape<T>::NV * xrhs_fs<T>(C)];
    __shared__ TD stab[3 * C];
    // instances last to first, as k_rhs_rk (the carry solve before it runs
    // them first to last)
    // FP64 with 32-row chunks (lines of >= 1024 rows): the streaming inputs
    // early (2048^2 -2.7 % kernel time, the uneven ensemble's per-group step
    // -2 %); the mixed and FP32 instances, shorter chunks (the 64 x 512^2
    // batch: +6.6 % per step) and the multi-shape kernel measured slower
    // with them early (more VGPRs, fewer waves)
    constexpr bool EARLY = sizeof(T) == 8 && sizeof(TD) == 8 && C == 32;
    xrhs_body<T, TD, C, EARLY>(a, blockIdx.x, blockIdx.y, (int)(gridDim.z - 1 - blockIdx.z), sdx,
                               stab);
}

// calcCharacteristicsX (calcRHS.c:253-495) at one open boundary cell:
// c[0..6] = outgoing-wave parts of d/dt of rho, ux, uy, uz, By, Bz, p
template <typename T>
__device__ __forceinline__ void characteristics_x(const T (&w)[8], const T (&d)[8], T gamma,
                                                  int boundary, T (&c)[7])
{
    const T rho = w[RHO], ux = w[UX], bx = w[BX], by = w[BY], bz = w[BZ], pr = w[PR];
    const T dux = d[UX], duy = d[UY], duz = d[UZ], dby = d[BY], dbz = d[BZ];
    const T drho = d[RHO], dp = d[PR];
    const T sr = sqrt(rho);
    const T cFlow = ux;
    const T cAlfx = bx / sr, cAlfy = by / sr, cAlfz = bz / sr;
    const T cs2 = gamma * pr / rho, cs = sqrt(cs2);
    const T cAlfx2 = cAlfx * cAlfx;
    const T sgn = cAlfx >= 0 ? (T)1 : (T)-1;
    const T cAlf2 = cAlfx2 + cAlfy * cAlfy + cAlfz * cAlfz;
    const T cperp = sqrt(cAlfy * cAlfy + cAlfz * cAlfz);
    T angY = 0, angZ = 0;
    if (!(cperp < (T)1e-7)) {
        angY = cAlfy / cperp;
        angZ = cAlfz / cperp;
    }
    const T cM2 = cs2 + cAlf2;
    const T cN2 = sqrt(fmax(cM2 * cM2 - 4 * cs2 * cAlfx2, (T)0));
    const T r2 = sqrt((T)2);
    const T cf = sqrt(fmax(cM2 + cN2, (T)0)) / r2, csl = sqrt(fmax(cM2 - cN2, (T)0)) / r2;
    const T al1 = sqrt(fabs((cf * cf - cAlfx2) / (cf * cf - csl * csl)));
    const T al2 = sqrt(fabs((cf * cf - cs2) / (cf * cf - csl * csl)));
    const T byzP = angZ * dbz + angY * dby, uyzP = angZ * duz + angY * duy;
    const T byzM = angY * dbz - angZ * dby, uyzM = angY * duz - angZ * duy;
    T fP = (cFlow + cf) * (al1 * (dp / rho / cf + dux) + al2 * (byzP / sr - cAlfx / cf * uyzP));
    T aP = (cFlow + sgn * cAlfx) * (uyzM - sgn / sr * byzM);
    T sP = (cFlow + csl) * (-al2 * (dp / rho / cs + sgn * cAlfx / cf * dux) +
                            al1 * (cs / cf / sr * byzP - sgn * uyzP));
    T en = ux * cs * (dp / gamma / pr - drho / rho);
    T sM = (cFlow - csl) * (-al2 * (dp / rho / cs - sgn * cAlfx / cf * dux) +
                            al1 * (cs / cf / sr * byzP + sgn * uyzP));
    T aM = (cFlow - sgn * cAlfx) * (uyzM + sgn / sr * byzM);
    T fM = (cFlow - cf) * (al1 * (dp / rho / cf - dux) + al2 * (byzP / sr + cAlfx / cf * uyzP));
    // drop the waves that enter the domain
    if (boundary == 0) {
        if (cFlow + cf >= 0) fP = 0;
        if (cFlow + sgn * cAlfx >= 0) aP = 0;
        if (cFlow + csl >= 0) sP = 0;
        if (cFlow >= 0) en = 0;
        if (cFlow - csl >= 0) sM = 0;
        if (cFlow - sgn * cAlfx >= 0) aM = 0;
        if (cFlow - cf >= 0) fM = 0;
    } else {
        if (cFlow + cf <= 0) fP = 0;
        if (cFlow + sgn * cAlfx <= 0) aP = 0;
        if (cFlow + csl <= 0) sP = 0;
        if (cFlow <= 0) en = 0;
        if (cFlow - csl <= 0) sM = 0;
        if (cFlow - sgn * cAlfx <= 0) aM = 0;
        if (cFlow - cf <= 0) fM = 0;
    }
    c[0] = rho / cs * en + rho * (al2 / 2 / cs * (sP + sM) - al1 / 2 / cf * (fP + fM));
    c[1] = csl * al2 / 2 / cs * (sP - sM) - al1 / 2 * (fP - fM);
    c[2] = angZ / 2 * (aP + aM) + angY * (sgn * al1 / 2 * (sP - sM) + cAlfx * al2 / 2 / cf * (fP - fM));
    c[3] = -angY / 2 * (aP + aM) + angZ * (sgn * al1 / 2 * (sP - sM) + cAlfx * al2 / 2 / cf * (fP - fM));
    c[4] = -sgn * sr * angZ / 2 * (aP - aM) - sr * angY * (cs * al1 / 2 / cf * (sP + sM) + al2 / 2 * (fP + fM));
    c[5] = sgn * sr * angY / 2 * (aP - aM) - sr * angZ * (cs * al1 / 2 / cf * (sP + sM) + al2 / 2 * (fP + fM));
    c[6] = gamma * pr * (al2 / 2 / cs * (sP + sM) - al1 / 2 / cf * (fP + fM));
}

// the two open boundary lines: calcRHSOpenBoundaryX (calcRHS.c:163-239) and
// the RK update of k_rhs_rk; blockIdx.y = 0 the line x = 0, 1 the line x = Lx
template <typename T, typename TD>
__global__ void __launch_bounds__(256) k_rhs_rk_open(RhsRkArgs<T, TD> a)
{
    const int z = blockIdx.z;
    const RunScalars* const rs = a.rs + z;
    if (rs->stop) return;
    const size_t n = a.n, ny = (size_t)a.ny;
    const int boundary = blockIdx.y;
    const size_t j = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (j >= ny) return;
    const size_t p = (boundary == 0 ? 0 : (size_t)(a.nx - 1) * ny) + j;
    T* const uz = a.u + z * a.bs_u;
    TD* const rpz = a.rprev + z * a.bs_r;
    const TD* const dxz = a.dudx + z * a.bs_d;
    const TD* const dyz = a.dudy + z * a.bs_d;
    const T cc = (T)rs->cc[a.stage], dd = (T)rs->dd[a.stage];
    const T g = (T)a.gamma;
    T w[8], d[8], e[8];
#pragma unroll
    for (int f = 0; f < 8; f++) {
        w[f] = uz[f * n + p];
        d[f] = (T)dxz[f * n + p];
        e[f] = (T)dyz[f * n + p];
    }
    const T rho = w[RHO], ux = w[UX], uy = w[UY], bx = w[BX], by = w[BY], bz = w[BZ], pr = w[PR];
    T R[8];
    // Bx keeps its d/dx terms in the u.grad form
    R[BX] = -ux * d[BX] - uy * e[BX] + bx * d[UX] + by * e[UX] - (d[UX] + e[UY]) * bx;
    const T jx = e[BZ], jy = 0, jz = -e[BX];
    const T divU = e[UY];
    R[PR] = -uy * e[PR] - g * pr * divU;
    R[RHO] = -rho * divU - uy * e[RHO];
    R[UX] = -uy * e[UX] + jy * bz / rho - jz * by / rho;
    R[UY] = -uy * e[UY] - e[PR] / rho + jz * bx / rho - jx * bz / rho;
    R[UZ] = -uy * e[UZ] + jx * by / rho - jy * bx / rho;
    R[BY] = -uy * e[BY] + by * e[UY] - divU * by;
    R[BZ] = -uy * e[BZ] + by * e[UZ] - divU * bz;
    T c[7];
    characteristics_x<T>(w, d, g, boundary, c);
    R[RHO] += c[0];
    R[UX] += c[1];
    R[UY] += c[2];
    R[UZ] += c[3];
    R[BY] += c[4];
    R[BZ] += c[5];
    R[PR] += c[6];
    T rpsi = 0;
    if (a.glm) {
        // hyperbolic cleaning at the open line: psi terms only (calcRHS.c:
        // 241-249), with the Ch, Cp of the first / last x block
        const double* const gc = a.glm_cp + 2 * ((size_t)z * a.npe + (boundary == 0 ? 0 : a.npe - 1));
        const T ch = (T)gc[0], cp = (T)gc[1];
        const T divB = d[BX] + e[BY];
        R[BX] = R[BX] - (T)dxz[PSI_SLOT * n + p];
        R[BY] = R[BY] - (T)dyz[PSI_SLOT * n + p];
        rpsi = -ch * ch * (divB + uz[PSI_SLOT * n + p] / cp / cp);
    }
    if (a.lap) {
        // addSourceTerm covers the boundary lines too (calcRHS.c:498-531)
        const TD* const lz = a.lap + z * a.bs_l;
        if (a.viscosity > 0) {
#pragma unroll
            for (int k = 0; k < 3; k++) R[UX + k] = R[UX + k] + (T)a.viscosity / rho * (T)lz[k * n + p];
        }
        if (a.resistivity > 0) {
#pragma unroll
            for (int k = 0; k < 3; k++) R[BX + k] = R[BX + k] + (T)a.resistivity * (T)lz[(3 + k) * n + p];
        }
    }
    if (a.ddivb) {
        const TD* const gz = a.ddivb + z * a.bs_b;
        R[BX] = R[BX] + (T)a.cdiff * (T)gz[p];
        R[BY] = R[BY] + (T)a.cdiff * (T)gz[n + p];
    }
    const bool use_prev = a.stage > 0;
#pragma unroll
    for (int f = 0; f < 8; f++) {
        const T rp = use_prev ? (T)rpz[f * n + p] : (T)0;
        uz[f * n + p] = w[f] + cc * R[f] + dd * rp;
        if (a.write_rprev) rpz[f * n + p] = (TD)R[f];
    }
    if (a.glm) {
        const T rp = use_prev ? (T)rpz[8 * n + p] : (T)0;
        uz[PSI_SLOT * n + p] = uz[PSI_SLOT * n + p] + cc * rpsi + dd * rp;
        if (a.write_rprev) rpz[8 * n + p] = (TD)rpsi;
    }
}

template <typename T>
__global__ void k_rk(T* __restrict__ u, const T* __restrict__ dudt, T* __restrict__ dudtRK,
                     T cc, T dd, size_t n)
{
    for (size_t p = blockIdx.x * (size_t)blockDim.x + threadIdx.x; p < n;
         p += (size_t)gridDim.x * blockDim.x) {
        u[p] = u[p] + cc * dudt[p] + dd * dudtRK[p];
        dudtRK[p] = dudt[p];
    }
}

template <typename T>
__global__ void __launch_bounds__(256) k_dt_limits(const T* __restrict__ u, DtParams dp,
                                                   RunScalars* rs, size_t n)
{
    __shared__ double red[16];
    double mn = INFINITY;
    const size_t p0 = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    for (size_t p = p0; p < n; p += (size_t)gridDim.x * blockDim.x)
        mn = fmin(mn, cell_dt(u[RHO * n + p], u[UX * n + p], u[UY * n + p], u[BX * n + p],
                              u[BY * n + p], u[BZ * n + p], u[PR * n + p], dp));
    block_min_dt(mn, p0 < n, rs, red);
}

// dt for the coming step: cfl * global min, with the +-2 % hysteresis of
// timeAdvance.c:144-151 against the previous step's dt (0 before the first
// step, mhd.c:58), then rkInit's coefficients (timeAdvance.c:20-27).
__device__ __forceinline__ void dt_finalize_body(RunScalars* rs, double cfl)
{
    if (rs->stop) return;
    const double m = __longlong_as_double((long long)rs->dtmin_bits) * cfl;
    const bool first = rs->step == 0;
    const double dt0 = first ? 0.0 : rs->dt;
    const double dt = (dt0 < 0.98 * m || dt0 > 1.02 * m) ? m : dt0;
    if (!first) rs->time += rs->dt;
    rs->dt_prev = first ? 0.0 : rs->dt;
    if (dt < 1e-8) {
        // mhd.c:131-151: the step is not taken; the time stays at the
        // previous step's end (time + dt with dt = 0)
        rs->stop = 1;
        rs->dt_stop = dt;
        rs->dt = 0.0;
        rs->dtmin_bits = 0x7ff0000000000000ull;
        return;
    }
    rs->dt = dt;
    rs->cc[0] = 8.0 / 15.0 * dt;
    rs->dd[0] = 0.0;
    rs->cc[1] = 5.0 / 12.0 * dt;
    rs->dd[1] = -17. / 60. * dt;
    rs->cc[2] = 0.75 * dt;
    rs->dd[2] = -5. / 12. * dt;
    rs->step += 1;
    rs->dtmin_bits = 0x7ff0000000000000ull;  // +inf for the next reduction
}

__global__ void k_dt_finalize(RunScalars* rs0, double cfl)
{
    dt_finalize_body(rs0 + blockIdx.x, cfl);  // one block per instance
}

__global__ void k_dt_reset(RunScalars* rs)
{
    rs->dtmin_bits = 0x7ff0000000000000ull;
    rs->nonfinite = 0;
}

template <typename T>
__global__ void k_add(const T* __restrict__ a, const T* __restrict__ b, T* __restrict__ out,
                      size_t n)
{
    for (size_t p = blockIdx.x * (size_t)blockDim.x + threadIdx.x; p < n;
         p += (size_t)gridDim.x * blockDim.x)
        out[p] = a[p] + b[p];
}

template <typename T>
__global__ void k_fill(T* p, T v, size_t n)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        p[i] = v;
}

// per-block partial sums (deterministic order): [0] energy density,
// [1] max |dBx/dx + dBy/dy| (from the two given derivative fields; skipped if null)
template <typename T>
__global__ void __launch_bounds__(256) k_diag(const T* __restrict__ u, const T* __restrict__ dbx,
                                              const T* __restrict__ dby, double* partial,
                                              size_t n, double gamma)
{
    __shared__ double se[256], sm[256];
    double e = 0.0, m = 0.0;
    for (size_t p = blockIdx.x * (size_t)blockDim.x + threadIdx.x; p < n;
         p += (size_t)gridDim.x * blockDim.x) {
        const double rho = u[RHO * n + p];
        const double ux = u[UX * n + p], uy = u[UY * n + p], uz = u[UZ * n + p];
        const double bx = u[BX * n + p], by = u[BY * n + p], bz = u[BZ * n + p];
        e += 0.5 * rho * (ux * ux + uy * uy + uz * uz) + u[PR * n + p] / (gamma - 1.0) +
             0.5 * (bx * bx + by * by + bz * bz);
        if (dbx) m = fmax(m, fabs((double)dbx[p] + (double)dby[p]));
    }
    se[threadIdx.x] = e;
    sm[threadIdx.x] = m;
    __syncthreads();
    for (int o = blockDim.x / 2; o > 0; o >>= 1) {
        if (threadIdx.x < o) {
            se[threadIdx.x] += se[threadIdx.x + o];
            sm[threadIdx.x] = fmax(sm[threadIdx.x], sm[threadIdx.x + o]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        partial[2 * blockIdx.x] = se[0];
        partial[2 * blockIdx.x + 1] = sm[0];
    }
}

// plain streaming copy of known size (bandwidth calibration of the profiler
// counters and of the achievable HBM rate)
__global__ void __launch_bounds__(256) k_copy(const double2* __restrict__ a,
                                              double2* __restrict__ b, size_t n)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        b[i] = a[i];
}

hipError_t launch_copy(const void* src, void* dst, size_t bytes, hipStream_t s)
{
    const size_t n = bytes / 16;
    hipLaunchKernelGGL(k_copy, pw_grid(n, 256), dim3(256), 0, s, (const double2*)src,
                       (double2*)dst, n);
    return hipGetLastError();
}

// ============================================================== launchers
template <typename T>
hipError_t launch_efield(const T* u, T* E, size_t n, hipStream_t s)
{
    hipLaunchKernelGGL((k_efield<T>), pw_grid(n, 256), dim3(256), 0, s, u, E, n);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_divb(const T* dudx, const T* dudy, T* divB, size_t n, hipStream_t s)
{
    hipLaunchKernelGGL((k_divb<T>), pw_grid(n, 256), dim3(256), 0, s, dudx, dudy, divB, n);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_glm_blocks(const T* u, size_t bs_u, const DtParams& dp, const RunScalars* rs,
                             double* out, int nx, int ny, int npe, int nb, hipStream_t s)
{
    if (nb < 1 || npe < 1) return hipErrorInvalidValue;
    hipLaunchKernelGGL((k_glm_blocks<T>), dim3(nb), dim3(64), 0, s, u, bs_u, dp, rs, out, nx, ny,
                       npe);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_divb_b(const T* dudx, const T* dudy, T* divB, size_t n, size_t bs_d,
                         size_t bs_o, int nb, hipStream_t s)
{
    if (nb < 1 || nb > 65535) return hipErrorInvalidValue;
    dim3 grid = pw_grid(n, 256);
    grid.y = nb;
    hipLaunchKernelGGL((k_divb_b<T>), grid, dim3(256), 0, s, dudx, dudy, divB, n, bs_d, bs_o);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_rhs(const RhsArgs<T>& a, hipStream_t s)
{
    hipLaunchKernelGGL((k_rhs<T>), pw_grid(a.n, 256), dim3(256), 0, s, a);
    return hipGetLastError();
}

template <typename T, typename TD>
hipError_t launch_rhs_rk_open(const RhsRkArgs<T, TD>& a, hipStream_t s, int nb)
{
    if (nb < 1 || nb > 65535 || a.nx < 2 || a.ny < 1) return hipErrorInvalidValue;
    dim3 grid((a.ny + 255) / 256, 2, nb);
    hipLaunchKernelGGL((k_rhs_rk_open<T, TD>), grid, dim3(256), 0, s, a);
    return hipGetLastError();
}

template <typename T, typename TD>
hipError_t launch_rhs_rk(const RhsRkArgs<T, TD>& a, hipStream_t s, int nb)
{
    if (nb < 1 || nb > 65535) return hipErrorInvalidValue;
    dim3 grid = pw_grid(a.n / (16 / sizeof(T)), 256);
    grid.y = nb;
    if (a.glm)
        hipLaunchKernelGGL((k_rhs_rk<T, TD, 2>), grid, dim3(256), 0, s, a);
    else if (a.lap || a.ddivb)
        hipLaunchKernelGGL((k_rhs_rk<T, TD, 1>), grid, dim3(256), 0, s, a);
    else
        hipLaunchKernelGGL((k_rhs_rk<T, TD, 0>), grid, dim3(256), 0, s, a);
    return hipGetLastError();
}

int xrhs_tile_columns(int nx, int elem_bytes)
{
    int C, W, G;
    xsolve_shape(nx, elem_bytes, &C, &W, &G);
    return (elem_bytes == 8 ? XRhsShape<double>::CELLS : XRhsShape<float>::CELLS) / C;
}

int xrhs_supported(int nx, int ny, int elem_bytes)
{
    int C, W, G;
    xsolve_shape(nx, elem_bytes, &C, &W, &G);  // C depends on nx only
    const int WB = (elem_bytes == 8 ? 512 : 1024) / C;
    return nx % C == 0 && G >= 2 && G <= 128 && ny % WB == 0;
}

template <typename T, typename TD>
hipError_t launch_xrhs_rk(const XRhsArgs<T, TD>& a, hipStream_t s, int nb)
{
    if (nb < 1 || nb > 65535 || !a.carry || !a.edge || a.x.tb.omask || a.r.open_x || a.r.lap ||
        a.r.ddivb || a.r.glm)
        return hipErrorInvalidValue;
    if (!xrhs_supported(a.x.nx, a.x.ny, sizeof(T))) return hipErrorInvalidValue;
    int C, W, G;
    xsolve_shape(a.x.nx, sizeof(T), &C, &W, &G);
    const int WB = XRhsShape<T>::CELLS / C;
    dim3 grid(a.x.ny / WB, G, nb), block(256);
    switch (C) {
    case 8: hipLaunchKernelGGL((k_xrhs_rk<T, TD, 8>), grid, block, 0, s, a); break;
    case 16: hipLaunchKernelGGL((k_xrhs_rk<T, TD, 16>), grid, block, 0, s, a); break;
    case 32: hipLaunchKernelGGL((k_xrhs_rk<T, TD, 32>), grid, block, 0, s, a); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <typename T>
hipError_t launch_rk(T* u, const T* dudt, T* dudtRK, double cc, double dd, size_t n,
                     hipStream_t s)
{
    hipLaunchKernelGGL((k_rk<T>), pw_grid(n, 256), dim3(256), 0, s, u, dudt, dudtRK, (T)cc,
                       (T)dd, n);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_dt_limits(const T* u, const DtParams& dp, RunScalars* rs, int nx, int ny,
                            hipStream_t s)
{
    const size_t n = (size_t)nx * ny;
    size_t b = (n + 255) / 256;
    if (b > 2048) b = 2048;
    hipLaunchKernelGGL((k_dt_limits<T>), dim3((unsigned)b), dim3(256), 0, s, u, dp, rs, n);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_glm_ch(const T* u, const DtParams& dp, RunScalars* rs, size_t n,
                         hipStream_t s)
{
    hipLaunchKernelGGL((k_glm_ch<T>), dim3(1), dim3(64), 0, s, u, dp, rs, n);
    return hipGetLastError();
}
template hipError_t launch_glm_ch<double>(const double*, const DtParams&, RunScalars*, size_t,
                                          hipStream_t);

hipError_t launch_dt_finalize(RunScalars* rs, double cfl, int first, hipStream_t s, int nb)
{
    (void)first;
    if (nb < 1) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_dt_finalize, dim3(nb), dim3(1), 0, s, rs, cfl);
    return hipGetLastError();
}

hipError_t launch_dt_reset(RunScalars* rs, hipStream_t s)
{
    hipLaunchKernelGGL(k_dt_reset, dim3(1), dim3(1), 0, s, rs);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_add(const T* a, const T* b, T* out, size_t n, hipStream_t s)
{
    hipLaunchKernelGGL((k_add<T>), pw_grid(n, 256), dim3(256), 0, s, a, b, out, n);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_fill(T* p, T v, size_t n, hipStream_t s)
{
    hipLaunchKernelGGL((k_fill<T>), pw_grid(n, 256), dim3(256), 0, s, p, v, n);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_diag(const T* u, const T* dbx, const T* dby, double* partial, int nblocks,
                       size_t n, double gamma, hipStream_t s)
{
    hipLaunchKernelGGL((k_diag<T>), dim3(nblocks), dim3(256), 0, s, u, dbx, dby, partial, n,
                       gamma);
    return hipGetLastError();
}

// ============================================================== shape groups
// One launch covers several grid shapes (an uneven ensemble): each block
// finds its group from the flattened block index and runs that group's
// code path, specialised on its line length; the shared memory is the
// largest any of them needs.  Per cell and line the arithmetic is that of
// the single-shape kernels, so results are the same bit for bit.
__device__ __forceinline__ int multi_group(const MultiMap& m, int b)
{
    int g = 0;
    while (g + 1 < m.ng && b >= m.first[g + 1]) g++;
    return g;
}

// line lengths a y-sweep workgroup of MHD2D_Y_THREADS threads covers
template <typename T, int N>
constexpr bool ymulti_ok()
{
    return YShape<T, N>::THREADS == MHD2D_Y_THREADS;
}
template <typename T>
constexpr int ymulti_lds()
{
    int m = 0;
    if (ymulti_ok<T, 128>() && YShape<T, 128>::ROWS * YShape<T, 128>::LDSW > m)
        m = YShape<T, 128>::ROWS * YShape<T, 128>::LDSW;
    if (ymulti_ok<T, 256>() && YShape<T, 256>::ROWS * YShape<T, 256>::LDSW > m)
        m = YShape<T, 256>::ROWS * YShape<T, 256>::LDSW;
    if (ymulti_ok<T, 512>() && YShape<T, 512>::ROWS * YShape<T, 512>::LDSW > m)
        m = YShape<T, 512>::ROWS * YShape<T, 512>::LDSW;
    if (ymulti_ok<T, 1024>() && YShape<T, 1024>::ROWS * YShape<T, 1024>::LDSW > m)
        m = YShape<T, 1024>::ROWS * YShape<T, 1024>::LDSW;
    if (ymulti_ok<T, 2048>() && YShape<T, 2048>::ROWS * YShape<T, 2048>::LDSW > m)
        m = YShape<T, 2048>::ROWS * YShape<T, 2048>::LDSW;
    return m;
}

// rows per workgroup of the multi-shape y-sweep, 0 where the line length
// does not fit its MHD2D_Y_THREADS threads
template <int N>
static int ymulti_rows(int elem_bytes)
{
    if (elem_bytes == 8) return ymulti_ok<double, N>() ? YShape<double, N>::ROWS : 0;
    return ymulti_ok<float, N>() ? YShape<float, N>::ROWS : 0;
}

int ysweep_multi_rows(int ny, int elem_bytes)
{
    switch (ny) {
    case 128: return ymulti_rows<128>(elem_bytes);
    case 256: return ymulti_rows<256>(elem_bytes);
    case 512: return ymulti_rows<512>(elem_bytes);
    case 1024: return ymulti_rows<1024>(elem_bytes);
    case 2048: return ymulti_rows<2048>(elem_bytes);
    default: return 0;
    }
}

template <typename T, typename TD, int MODE>
__global__ void __launch_bounds__(MHD2D_Y_THREADS) __attribute__((amdgpu_waves_per_eu(MHD2D_Y_WAVES)))
    k_ysweep_multi(const YSweepArgs<T, TD>* __restrict__ tab, MultiMap m)
{
    __shared__ T s_re[ymulti_lds<T>()];
    __shared__ T s_im[ymulti_lds<T>()];
    __shared__ double red[16];
    const int b = blockIdx.x;
    const int g = multi_group(m, b);
    const YSweepArgs<T, TD>& a = tab[g];
    const int local = b - m.first[g];
    const int bx = local % m.per[g];
    // instances last to first, as in k_ysweep
    const int z = m.nb[g] - 1 - local / m.per[g];
    // only the line lengths ysweep_multi_rows admits (a workgroup of
    // MHD2D_Y_THREADS threads, shared memory sized for them)
    switch (a.ny) {
    case 128:
        if constexpr (ymulti_ok<T, 128>()) ysweep_body<T, TD, 128, MODE>(a, bx, z, s_re, s_im, red);
        break;
    case 256:
        if constexpr (ymulti_ok<T, 256>()) ysweep_body<T, TD, 256, MODE>(a, bx, z, s_re, s_im, red);
        break;
    case 512:
        if constexpr (ymulti_ok<T, 512>()) ysweep_body<T, TD, 512, MODE>(a, bx, z, s_re, s_im, red);
        break;
    case 1024:
        if constexpr (ymulti_ok<T, 1024>()) ysweep_body<T, TD, 1024, MODE>(a, bx, z, s_re, s_im, red);
        break;
    case 2048:
        if constexpr (ymulti_ok<T, 2048>()) ysweep_body<T, TD, 2048, MODE>(a, bx, z, s_re, s_im, red);
        break;
    default: break;
    }
}

int xsolve_multi_code(int nx, int ny, int* chunk, int* width)
{
    int C, W, G;
    xsolve_shape(nx, 8, &C, &W, &G);
    const int w = G > 0 ? 1024 / G : 0;
    int code = -1;
    // the shapes xsolve_shape gives lines of 256 ... 4096 rows
    if (C == 8 && w == 32) code = 0;
    else if (C == 16 && w == 32) code = 1;
    else if (C == 32 && w == 32) code = 2;
    else if (C == 32 && w == 16) code = 3;
    else if (C == 32 && w == 8) code = 4;
    if (code < 0 || w * G != 1024 || ny % w) return -1;
    if (chunk) *chunk = C;
    if (width) *width = w;
    return code;
}

template <typename T, typename TI, int KIND>
__global__ void __launch_bounds__(1024, MHD2D_X_MINB)
    k_xsolve_multi(const XSolveArgs<T, TI>* __restrict__ tab, MultiMap m)
{
    __shared__ T sh_in[xshared_max_words()];
    __shared__ T sh_fw[xshared_max_words()];
    __shared__ T sh_bw[xshared_max_words()];
    __shared__ T sh_half[32];
    extern __shared__ unsigned char x_dyn[];
    // The groups in reverse order (each group's blocks in their own order):
    // the y-sweep and the RHS + RK end on the last group, the x-solves
    // start there and end on the first group, where the next kernel starts,
    // so every launch begins on the lines the one before it touched last
    // (one-launch step -0.6 % in FP64, FP32 and mixed on the 96-run
    // ensemble against one order everywhere).
    const int b2 = m.first[m.ng] - 1 - (int)blockIdx.x;
    const int g = multi_group(m, b2);
    const int local = m.first[g + 1] - 1 - b2;
    const XSolveArgs<T, TI>& a = tab[g];
    const int strips = m.aux[g];
    const int nv = m.per[g] / strips;
    const int bx = local % strips;
    const int v = (local / strips) % nv;
    const int z = local / m.per[g];
    switch (m.code[g]) {
    case 0: xsolve_body<T, TI, 8, 32, KIND, false>(a, bx, v, z, sh_in, sh_fw, sh_bw, sh_half, x_dyn); break;
    case 1: xsolve_body<T, TI, 16, 32, KIND, false>(a, bx, v, z, sh_in, sh_fw, sh_bw, sh_half, x_dyn); break;
    case 2: xsolve_body<T, TI, 32, 32, KIND, false>(a, bx, v, z, sh_in, sh_fw, sh_bw, sh_half, x_dyn); break;
    case 3: xsolve_body<T, TI, 32, 16, KIND, false>(a, bx, v, z, sh_in, sh_fw, sh_bw, sh_half, x_dyn); break;
    case 4: xsolve_body<T, TI, 32, 8, KIND, false>(a, bx, v, z, sh_in, sh_fw, sh_bw, sh_half, x_dyn); break;
    default: break;
    }
}

template <typename T, typename TD, int MODE>
__global__ void __launch_bounds__(256)
    __attribute__((amdgpu_waves_per_eu(1, RhsRkWaves<T, TD>::MAX)))
        k_rhs_rk_multi(const RhsRkArgs<T, TD>* __restrict__ tab, MultiMap m)
{
    const int b = blockIdx.x;
    const int g = multi_group(m, b);
    const int local = b - m.first[g];
    const int per = m.per[g];
    // instances last to first, as in k_rhs_rk
    rhs_rk_body<T, TD, MODE>(tab[g], m.nb[g] - 1 - local / per, local % per, per);
}

// the fused x-derivative + RHS + RK over every group: per[g] blocks per
// instance (aux[g] column tiles x the chunks), code[g] the chunk rows
template <typename T, typename TD>
__global__ void __launch_bounds__(256) k_xrhs_rk_multi(const XRhsArgs<T, TD>* __restrict__ tab,
                                                       MultiMap m)
{
    // sized for the largest tile table: 8-row chunks
    __shared__ __align__(16) TD sdx[XRhsShape<T>::NV * xrhs_fs<T>(8)];
    __shared__ TD stab[3 * 32];
    const int b = blockIdx.x;
    const int g = multi_group(m, b);
    const int local = b - m.first[g];
    const int per = m.per[g], tiles = m.aux[g];
    // instances last to first, as k_xrhs_rk
    const int z = m.nb[g] - 1 - local / per;
    const int bx = (local % per) % tiles, ch = (local % per) / tiles;
    switch (m.code[g]) {
    case 8: xrhs_body<T, TD, 8, false>(tab[g], bx, ch, z, sdx, stab); break;
    case 16: xrhs_body<T, TD, 16, false>(tab[g], bx, ch, z, sdx, stab); break;
    case 32: xrhs_body<T, TD, 32, false>(tab[g], bx, ch, z, sdx, stab); break;
    default: break;
    }
}

__global__ void k_dt_finalize_multi(const DtMultiArgs* __restrict__ tab, MultiMap m)
{
    const int b = blockIdx.x;
    const int g = multi_group(m, b);
    dt_finalize_body(tab[g].rs + (b - m.first[g]), tab[g].cfl);
}

static int multi_blocks(const MultiMap& m)
{
    return m.ng > 0 ? m.first[m.ng] : 0;
}

template <typename T, typename TD>
hipError_t launch_ysweep_multi(const YSweepArgs<T, TD>* tab, const MultiMap& m, int mode,
                               hipStream_t s)
{
    const int nb = multi_blocks(m);
    if (m.ng < 1 || m.ng > MULTI_MAXG || nb < 1 || mode < 0 || mode > 1) return hipErrorInvalidValue;
    if (mode == 1)
        hipLaunchKernelGGL((k_ysweep_multi<T, TD, 1>), dim3(nb), dim3(MHD2D_Y_THREADS), 0, s, tab, m);
    else
        hipLaunchKernelGGL((k_ysweep_multi<T, TD, 0>), dim3(nb), dim3(MHD2D_Y_THREADS), 0, s, tab, m);
    return hipGetLastError();
}

template <typename T, typename TI>
hipError_t launch_xsolve_multi(const XSolveArgs<T, TI>* tab, const MultiMap& m, int kind,
                               size_t dyn_lds, hipStream_t s)
{
    const int nb = multi_blocks(m);
    if (m.ng < 1 || m.ng > MULTI_MAXG || nb < 1) return hipErrorInvalidValue;
    if constexpr (std::is_same<T, TI>::value) {
        switch (kind) {
        case XD1: hipLaunchKernelGGL((k_xsolve_multi<T, TI, XD1>), dim3(nb), dim3(1024), dyn_lds, s, tab, m); break;
        case XD2: hipLaunchKernelGGL((k_xsolve_multi<T, TI, XD2>), dim3(nb), dim3(1024), dyn_lds, s, tab, m); break;
        case XD1_CARRY: hipLaunchKernelGGL((k_xsolve_multi<T, TI, XD1_CARRY>), dim3(nb), dim3(1024), dyn_lds, s, tab, m); break;
        case XFILT_EDGE: hipLaunchKernelGGL((k_xsolve_multi<T, TI, XFILT_EDGE>), dim3(nb), dim3(1024), dyn_lds, s, tab, m); break;
        default: hipLaunchKernelGGL((k_xsolve_multi<T, TI, XFILT>), dim3(nb), dim3(1024), dyn_lds, s, tab, m); break;
        }
    } else {
        // converting inputs: derivatives only (they annihilate the subtracted
        // column constant; the filter does not)
        if (kind == XD1)
            hipLaunchKernelGGL((k_xsolve_multi<T, TI, XD1>), dim3(nb), dim3(1024), dyn_lds, s, tab, m);
        else if (kind == XD2)
            hipLaunchKernelGGL((k_xsolve_multi<T, TI, XD2>), dim3(nb), dim3(1024), dyn_lds, s, tab, m);
        else if (kind == XD1_CARRY)
            hipLaunchKernelGGL((k_xsolve_multi<T, TI, XD1_CARRY>), dim3(nb), dim3(1024), dyn_lds, s, tab, m);
        else
            return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <typename T, typename TD>
hipError_t launch_rhs_rk_multi(const RhsRkArgs<T, TD>* tab, const MultiMap& m, int mode,
                               hipStream_t s)
{
    const int nb = multi_blocks(m);
    if (m.ng < 1 || m.ng > MULTI_MAXG || nb < 1 || mode < 0 || mode > 1) return hipErrorInvalidValue;
    if (mode == 1)
        hipLaunchKernelGGL((k_rhs_rk_multi<T, TD, 1>), dim3(nb), dim3(256), 0, s, tab, m);
    else
        hipLaunchKernelGGL((k_rhs_rk_multi<T, TD, 0>), dim3(nb), dim3(256), 0, s, tab, m);
    return hipGetLastError();
}

template <typename T, typename TD>
hipError_t launch_xrhs_rk_multi(const XRhsArgs<T, TD>* tab, const MultiMap& m, hipStream_t s)
{
    const int nb = multi_blocks(m);
    if (m.ng < 1 || m.ng > MULTI_MAXG || nb < 1) return hipErrorInvalidValue;
    hipLaunchKernelGGL((k_xrhs_rk_multi<T, TD>), dim3(nb), dim3(256), 0, s, tab, m);
    return hipGetLastError();
}

hipError_t launch_dt_finalize_multi(const DtMultiArgs* tab, const MultiMap& m, hipStream_t s)
{
    const int nb = multi_blocks(m);
    if (m.ng < 1 || m.ng > MULTI_MAXG || nb < 1) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_dt_finalize_multi, dim3(nb), dim3(1), 0, s, tab, m);
    return hipGetLastError();
}

// explicit instantiations
#define MHD2D_INST(T)                                                                        \
    template hipError_t launch_xsolve<T, T>(const XSolveArgs<T, T>&, int, hipStream_t, int); \
    template hipError_t launch_yapply<T>(const YApplyArgs<T>&, hipStream_t, int);            \
    template hipError_t launch_ysweep<T, T>(const YSweepArgs<T, T>&, hipStream_t, int);     \
    template hipError_t launch_rhs<T>(const RhsArgs<T>&, hipStream_t);                       \
    template hipError_t launch_rhs_rk<T, T>(const RhsRkArgs<T, T>&, hipStream_t, int);      \
    template hipError_t launch_rhs_rk_open<T, T>(const RhsRkArgs<T, T>&, hipStream_t, int); \
    template hipError_t launch_efield<T>(const T*, T*, size_t, hipStream_t);                 \
    template hipError_t launch_divb<T>(const T*, const T*, T*, size_t, hipStream_t);         \
    template hipError_t launch_divb_b<T>(const T*, const T*, T*, size_t, size_t, size_t, int, \
                                         hipStream_t);                                       \
    template hipError_t launch_glm_blocks<T>(const T*, size_t, const DtParams&,               \
                                             const RunScalars*, double*, int, int, int, int,  \
                                             hipStream_t);                                    \
    template hipError_t launch_rk<T>(T*, const T*, T*, double, double, size_t, hipStream_t); \
    template hipError_t launch_dt_limits<T>(const T*, const DtParams&, RunScalars*, int, int, \
                                            hipStream_t);                                    \
    template hipError_t launch_fill<T>(T*, T, size_t, hipStream_t);                          \
    template hipError_t launch_add<T>(const T*, const T*, T*, size_t, hipStream_t);          \
    template hipError_t launch_diag<T>(const T*, const T*, const T*, double*, int, size_t,   \
                                       double, hipStream_t);

MHD2D_INST(double)
MHD2D_INST(float)

#define MHD2D_INST_MULTI(T, TD)                                                              \
    template hipError_t launch_ysweep_multi<T, TD>(const YSweepArgs<T, TD>*, const MultiMap&, \
                                                   int, hipStream_t);                        \
    template hipError_t launch_xrhs_rk_multi<T, TD>(const XRhsArgs<T, TD>*, const MultiMap&,  \
                                                    hipStream_t);                            \
    template hipError_t launch_rhs_rk_multi<T, TD>(const RhsRkArgs<T, TD>*, const MultiMap&, \
                                                   int, hipStream_t);
MHD2D_INST_MULTI(double, double)
MHD2D_INST_MULTI(float, float)
MHD2D_INST_MULTI(double, float)
template hipError_t launch_xsolve_multi<double, double>(const XSolveArgs<double, double>*,
                                                        const MultiMap&, int, size_t,
                                                        hipStream_t);
template hipError_t launch_xsolve_multi<float, float>(const XSolveArgs<float, float>*,
                                                      const MultiMap&, int, size_t, hipStream_t);
template hipError_t launch_xsolve_multi<float, double>(const XSolveArgs<float, double>*,
                                                       const MultiMap&, int, size_t,
                                                       hipStream_t);

// mixed form: FP64 state, FP32 derivatives and R_prev
template hipError_t launch_xsolve<float, double>(const XSolveArgs<float, double>&, int,
                                                 hipStream_t, int);
template hipError_t launch_ysweep<double, float>(const YSweepArgs<double, float>&, hipStream_t,
                                                 int);
template hipError_t launch_rhs_rk<double, float>(const RhsRkArgs<double, float>&, hipStream_t,
                                                 int);
template hipError_t launch_rhs_rk_open<double, float>(const RhsRkArgs<double, float>&,
                                                      hipStream_t, int);

template hipError_t launch_xrhs_rk<double, double>(const XRhsArgs<double, double>&, hipStream_t,
                                                   int);
template hipError_t launch_xrhs_rk<float, float>(const XRhsArgs<float, float>&, hipStream_t, int);
template hipError_t launch_xrhs_rk<double, float>(const XRhsArgs<double, float>&, hipStream_t,
                                                  int);

}  // namespace mhd2d
