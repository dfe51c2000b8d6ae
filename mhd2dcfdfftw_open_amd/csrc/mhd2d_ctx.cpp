// Context API: whole RK3 steps on the device with the fused schedule
//   stage s:  y-sweep  (pending y-filter of stage s-1, d/dy of u, E, dEx/dy, dEz/dy,
//                       CFL limits when s == 0)
//             [s == 0: dt finalize, the reference's calcDt + rkInit]
//             x-derivative of u, Ey and Ez   (10 compact line solves)
//             RHS + low-storage RK update    (pointwise)
//             x-filter of u                  (8 compact line solves)
// which is the reference order of mhd.c:154-206 with the y-filter of each
// stage carried into the next stage's y-sweep.  A context holds `batch`
// independent runs of one grid; every launch covers all of them (the
// instance index is a grid dimension) and each has its own RunScalars.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/mhd2d.h"
#include "mhd2d_internal.h"
#include "mhd2d_tables.h"

using namespace mhd2d;

namespace {
thread_local std::string g_err;

int fail(int code, const char* what, hipError_t e = hipSuccess)
{
    g_err = what;
    if (e != hipSuccess) {
        g_err += ": ";
        g_err += hipGetErrorString(e);
    }
    return code;
}

// a device temporary released when it goes out of scope (early error
// returns included)
template <typename T>
struct DevTemp {
    T* p = nullptr;
    ~DevTemp()
    {
        if (p) (void)hipFree(p);
    }
};
}  // namespace

#define HIPCHK(expr)                                                        \
    do {                                                                    \
        hipError_t e_ = (expr);                                             \
        if (e_ != hipSuccess) return fail(MHD2D_EHIP, #expr, e_);           \
    } while (0)

struct mhd2d_ctx {
    mhd2d_params p{};
    int elem = 8;           // bytes per state value
    int delem = 8;          // bytes per derivative / R_prev value
    int B = 1;              // instances
    size_t n = 0;
    GridTables tb;
    XGemmOp xg;             // built when params.xgemm != 0
    hipStream_t stream = nullptr;
    // u: [B][11][nx][ny]: 8 state fields, Ey, Ez, one scratch field (divB)
    // dudx: [B][11][nx][ny]: d/dx of the 8 state fields, dEy/dx, dEz/dx
    // dudy: [B][11][nx][ny]: d/dy of the 8 state fields, dEx/dy, dEz/dy
    void* u = nullptr;
    void* dudx = nullptr;
    void* dudy = nullptr;
    void* rprev = nullptr;  // [B][8][nx][ny]
    // [B][6][nx][ny] (derivative type), with viscosity or resistivity: the
    // Laplacians of ux, uy, uz and Bx, By, Bz (d2/dy2 from the y-sweep, then
    // d2/dx2 added by the x-solve)
    void* lap = nullptr;
    // [B][3][nx][ny] (derivative type), with diffusive div B cleaning: div B,
    // d(div B)/dx, d(div B)/dy
    void* divb = nullptr;
    // [B][xblocks][2] FP64, with GLM cleaning: Ch, Cp of every x block
    double* glm_cp = nullptr;
    RunScalars* rs = nullptr;  // [B]
    double* partial = nullptr;
    int npartial = 0;
    bool pending_filter = false;
    size_t bytes = 0;
    hipGraphExec_t graph = nullptr;
    // launch shapes of the derivative and filter x-solves (strip width,
    // XCD-contiguous order); defaults until mhd2d_autotune measures them
    int xw_d = 0, xr_d = -1, xw_f = 0, xr_f = -1;
    // instances per launch: a step runs the batch as groups of this many
    // instances, one whole step per group (0: the whole batch at once)
    int subbatch = 0;
    // x-derivative and RHS + RK as the carry solve and the fused chunk
    // kernel (k_xrhs_rk) instead of the full solve and k_rhs_rk; only
    // where xfuse_ok holds
    int xfuse = 0;
    // [B][8][nx / C][6][ny] (state type), allocated with xfuse: the fused
    // kernel's updated state of the first and last 3 rows of every x chunk
    // of C rows, which the neighbouring chunks' workgroups still read from
    // u as halos; the x-filter reads them from here
    void* xedge = nullptr;
    size_t edge_fs() const
    {
        int C, W, G;
        xsolve_shape(p.nx, elem, &C, &W, &G);
        return (size_t)G * 6 * p.ny;
    }
    double cdiff = 0;  // constDiffDivB = 2 max(dx, dy)^2 (initialize.c:17, 154)
    std::vector<char> have_state;
    std::vector<double> di;
    // per-instance strides (elements); with GLM cleaning psi takes slot
    // PSI_SLOT of u and the diagnostics scratch moves to slot 11
    bool glm() const { return p.divbclean == 2; }
    // dBx/dx of a stage is read only by the cleaning terms and the
    // open-boundary lines; the ideal periodic RHS does without it (dBy/dy
    // comes out of the (Bx, By) transform pair and is stored regardless:
    // skipping the store cost the y-sweep registers and time)
    bool divb_derivs() const { return p.divbclean != 0 || p.open_x; }
    size_t su() const { return (glm() ? 12 : 11) * n; }
    int nv_state() const { return glm() ? 9 : 8; }
    size_t sd() const { return 11 * n; }
    size_t sr() const { return (glm() ? 9 : 8) * n; }
    size_t sl() const { return 6 * n; }
    size_t sb() const { return 3 * n; }
    bool lap_u() const { return p.viscosity > 0; }
    bool lap_b() const { return p.resistivity > 0; }
};

namespace {

DtParams dt_params(const mhd2d_ctx* c)
{
    DtParams d;
    d.dx = c->tb.dx;
    d.dy = c->tb.dy;
    d.gamma = c->p.gamma;
    d.cfl = c->p.cfl;
    d.di = c->p.di;
    d.resistivity = c->p.resistivity;
    d.viscosity = c->p.viscosity;
    d.hall = c->p.hall;
    return d;
}

template <typename T, typename TD>
YSweepArgs<T, TD> ysweep_args(mhd2d_ctx* c, int filter, int derivs, int dt_limits)
{
    YSweepArgs<T, TD> a;
    TD* dudy = (TD*)c->dudy;
    a.u = (T*)c->u;
    a.dudy = dudy;
    a.E = (T*)c->u + 8 * c->n;
    a.dEdy = dudy + 8 * c->n;
    a.tb = c->tb.yt;
    a.dp = dt_params(c);
    a.rs = c->rs;
    a.nx = c->p.nx;
    a.ny = c->p.ny;
    a.filter = filter;
    a.derivs = derivs;
    a.dt_limits = dt_limits;
    a.bs_u = c->su();
    a.bs_d = c->sd();
    a.glm = c->glm();
    if (c->lap && derivs) {
        a.lap = (TD*)c->lap;
        a.lap_u = c->lap_u();
        a.lap_b = c->lap_b();
        a.bs_l = c->sl();
    }
    return a;
}

template <typename T, typename TD>
hipError_t ysweep(mhd2d_ctx* c, int filter, int derivs, int dt_limits)
{
    return launch_ysweep<T, TD>(ysweep_args<T, TD>(c, filter, derivs, dt_limits), c->stream,
                                c->B);
}

// x-derivative of nv fields: in (state type, stride bs_in per instance) ->
// out (derivative type); the tridiagonal solve or the MFMA GEMM form.  Field
// v is slot fl[v] of in and out (slot v without a list).
template <typename T, typename TD>
hipError_t xderiv_fields(mhd2d_ctx* c, const T* in, TD* out, int nv, size_t bs_in,
                         size_t bs_out, int nb, const RunScalars* rs, const int* fl = nullptr)
{
    if constexpr (std::is_same<TD, float>::value) {
        if (c->p.xgemm) {
            XGemmArgs<T> g;
            for (int v = 0; v < nv; v++) {
                g.in[v] = in + (fl ? fl[v] : v) * c->n;
                g.out[v] = out + (fl ? fl[v] : v) * c->n;
            }
            g.mh = c->xg.mh;
            g.mm = c->xg.mm;
            g.ml = c->xg.ml;
            g.klist = c->xg.klist_d;
            g.kcount = c->xg.kcount_d;
            g.kmax = c->xg.kmax;
            g.klist2 = c->xg.klist2_d;
            g.kcount2 = c->xg.kcount2_d;
            g.kmax2 = c->xg.kmax2;
            g.klist4 = c->xg.klist4_d;
            g.kcount4 = c->xg.kcount4_d;
            g.kmax4 = c->xg.kmax4;
            g.nsplit = c->xg.nsplit;
            g.rs = rs;
            g.nx = c->p.nx;
            g.ny = c->p.ny;
            g.bs_in = bs_in;
            g.bs_out = bs_out;
            return launch_xgemm<T>(g, nv, c->stream, nb);
        }
    }
    XSolveArgs<TD, T> a;
    for (int v = 0; v < nv; v++) {
        a.in[v] = in + (fl ? fl[v] : v) * c->n;
        a.out[v] = out + (fl ? fl[v] : v) * c->n;
    }
    a.rs = rs;
    a.st = c->tb.st[XD1];
    a.tb = c->tb.xt[XD1];
    a.kind = XD1;
    a.nx = c->p.nx;
    a.ny = c->p.ny;
    a.bs_in = bs_in;
    a.bs_out = bs_out;
    a.width = c->xw_d;
    a.remap = c->xr_d;
    return launch_xsolve<TD, T>(a, nv, c->stream, nb);
}

template <typename T, typename TD>
hipError_t xderiv(mhd2d_ctx* c)
{
    // 8 state fields, Ey and Ez, solved in the derivative type (and psi in
    // slot PSI_SLOT = 10 with GLM cleaning); Bx only where dBx/dx is read
    static const int ideal[9] = {RHO, UX, UY, UZ, BY, BZ, PR, 8, 9};
    hipError_t e =
        c->divb_derivs()
            ? xderiv_fields<T, TD>(c, (const T*)c->u, (TD*)c->dudx, c->glm() ? 11 : 10, c->su(),
                                   c->sd(), c->B, c->rs)
            : xderiv_fields<T, TD>(c, (const T*)c->u, (TD*)c->dudx, 9, c->su(), c->sd(), c->B,
                                   c->rs, ideal);
    if (e != hipSuccess || !c->lap) return e;
    // second x-derivatives of ux..uz (viscosity) and / or Bx..Bz
    // (resistivity), added onto the second y-derivatives in lap
    const int f0 = c->lap_u() ? UX : BX, f1 = c->lap_b() ? BZ : UZ;
    XSolveArgs<TD, T> a;
    for (int v = 0; v <= f1 - f0; v++) {
        a.in[v] = (const T*)c->u + (f0 + v) * c->n;
        a.out[v] = (TD*)c->lap + (f0 + v - UX) * c->n;
    }
    a.rs = c->rs;
    a.st = c->tb.st[XD2];
    a.tb = c->tb.xt[XD2];
    a.kind = XD2;
    a.nx = c->p.nx;
    a.ny = c->p.ny;
    a.accumulate = 1;
    a.width = c->xw_d;
    a.remap = c->xr_d;
    a.bs_in = c->su();
    a.bs_out = c->sl();
    return launch_xsolve<TD, T>(a, f1 - f0 + 1, c->stream, c->B);
}

// u in place; with edge, the chunk edge rows of the input come from the
// fused path's edge buffer
template <typename T>
XSolveArgs<T> xfilter_args(mhd2d_ctx* c, bool edge)
{
    XSolveArgs<T> a;
    T* u = (T*)c->u;
    for (int v = 0; v < 8; v++) a.in[v] = a.out[v] = u + v * c->n;
    if (edge) {
        a.edge = (const T*)c->xedge;
        a.edge_fs = c->edge_fs();
        a.bs_edge = 8 * a.edge_fs;
    }
    if (c->glm()) a.in[8] = a.out[8] = u + PSI_SLOT * c->n;
    a.rs = c->rs;
    a.st = c->tb.st[XFILT];
    a.tb = c->tb.xt[XFILT];
    a.kind = XFILT;
    a.nx = c->p.nx;
    a.ny = c->p.ny;
    a.bs_in = a.bs_out = c->su();
    a.width = c->xw_f;
    a.remap = c->xr_f;
    return a;
}

template <typename T>
hipError_t xfilter(mhd2d_ctx* c)
{
    return launch_xsolve<T>(xfilter_args<T>(c, c->xfuse != 0), c->nv_state(), c->stream, c->B);
}

template <typename T, typename TD>
RhsRkArgs<T, TD> rhs_rk_args(mhd2d_ctx* c, int stage)
{
    RhsRkArgs<T, TD> a;
    TD* dx = (TD*)c->dudx;
    TD* dy = (TD*)c->dudy;
    a.u = (T*)c->u;
    a.dudx = dx;
    a.dudy = dy;
    a.dEdx = dx + 8 * c->n;
    a.dEdy = dy + 8 * c->n;
    a.rprev = (TD*)c->rprev;
    a.rs = c->rs;
    a.gamma = c->p.gamma;
    a.stage = stage;
    a.write_rprev = stage < 2;  // rkInit zeroes R_prev before it is read again
    a.n = c->n;
    a.bs_u = c->su();
    a.bs_d = c->sd();
    a.bs_r = c->sr();
    a.open_x = c->p.open_x;
    a.nx = c->p.nx;
    a.ny = c->p.ny;
    if (c->lap) {
        a.lap = (const TD*)c->lap;
        a.bs_l = c->sl();
        a.viscosity = c->p.viscosity;
        a.resistivity = c->p.resistivity;
    }
    if (c->glm()) {
        a.glm = 1;
        a.npe = c->p.xblocks > 1 ? c->p.xblocks : 1;
        a.glm_cp = c->glm_cp;
    }
    return a;
}

template <typename T, typename TD>
hipError_t rhs_rk(mhd2d_ctx* c, int stage)
{
    RhsRkArgs<T, TD> a = rhs_rk_args<T, TD>(c, stage);
    TD* dx = (TD*)c->dudx;
    TD* dy = (TD*)c->dudy;
    hipError_t e = hipSuccess;
    if (c->divb) {
        // diffusive cleaning (mhd.c:175-182): div B of this stage's
        // derivatives, then its x- and y-derivatives
        TD* const db = (TD*)c->divb;
        e = launch_divb_b<TD>(dx, dy, db, c->n, c->sd(), c->sb(), c->B, c->stream);
        if (e == hipSuccess) {
            YApplyArgs<TD> ya;
            ya.in[0] = db;
            ya.out[0] = db + 2 * c->n;
            ya.tb = c->tb.yt;
            ya.nv = 1;
            ya.mode = Y_D1;
            ya.nx = c->p.nx;
            ya.ny = c->p.ny;
            ya.bs_in = ya.bs_out = c->sb();
            e = launch_yapply<TD>(ya, c->stream, c->B);
        }
        if (e == hipSuccess) {
            XSolveArgs<TD> xa;
            xa.in[0] = db;
            xa.out[0] = db + c->n;
            xa.rs = c->rs;
            xa.st = c->tb.st[XD1];
            xa.tb = c->tb.xt[XD1];
            xa.kind = XD1;
            xa.nx = c->p.nx;
            xa.ny = c->p.ny;
            xa.bs_in = xa.bs_out = c->sb();
            xa.width = c->xw_d;
            xa.remap = c->xr_d;
            e = launch_xsolve<TD>(xa, 1, c->stream, c->B);
        }
        a.ddivb = db + c->n;
        a.bs_b = c->sb();
        a.cdiff = c->cdiff;
    }
    if (e == hipSuccess) e = launch_rhs_rk<T, TD>(a, c->stream, c->B);
    if (e == hipSuccess && c->p.open_x) e = launch_rhs_rk_open<T, TD>(a, c->stream, c->B);
    return e;
}

// the fused x-derivative + RHS + RK applies: ideal MHD on one periodic x
// block, the tridiagonal x-derivative, a grid its workgroups tile
bool xfuse_ok(const mhd2d_ctx* c)
{
    return !c->lap && c->p.divbclean == 0 && !c->p.open_x && c->p.xblocks <= 1 &&
           !c->p.xgemm && xrhs_supported(c->p.nx, c->p.ny, c->elem);
}

// the x-derivative solve of the fields the ideal RHS reads, stopping at the
// chunk carries (into the dudx buffer, which the fused path leaves unused:
// 18 nx ny / C <= 2.25 nx ny words per instance)
template <typename T, typename TD>
XSolveArgs<TD, T> xcarry_args(mhd2d_ctx* c)
{
    static const int ideal[9] = {RHO, UX, UY, UZ, BY, BZ, PR, 8, 9};
    XSolveArgs<TD, T> a;
    for (int v = 0; v < 9; v++) {
        a.in[v] = (const T*)c->u + ideal[v] * c->n;
        a.out[v] = nullptr;
    }
    a.rs = c->rs;
    a.st = c->tb.st[XD1];
    a.tb = c->tb.xt[XD1];
    a.kind = XD1;
    a.nx = c->p.nx;
    a.ny = c->p.ny;
    a.bs_in = c->su();
    a.width = c->xw_d;
    a.remap = c->xr_d;
    a.carry = (TD*)c->dudx;
    a.bs_carry = c->sd();
    return a;
}

// the fused path's edge-row buffer, allocated on first use
hipError_t ensure_xedge(mhd2d_ctx* c)
{
    if (c->xedge) return hipSuccess;
    const size_t b = 8 * c->edge_fs() * c->elem * c->B;
    const hipError_t e = hipMalloc(&c->xedge, b);
    if (e == hipSuccess) c->bytes += b;
    return e;
}

template <typename T, typename TD>
XRhsArgs<T, TD> xrhs_args(mhd2d_ctx* c, int stage)
{
    XRhsArgs<T, TD> a;
    a.x = xcarry_args<T, TD>(c);
    a.r = rhs_rk_args<T, TD>(c, stage);
    a.carry = (const TD*)c->dudx;
    a.bs_carry = c->sd();
    a.edge = (T*)c->xedge;
    a.edge_fs = c->edge_fs();
    a.bs_edge = 8 * a.edge_fs;
    return a;
}

int nsub(const mhd2d_ctx* c)
{
    const int k = c->subbatch > 0 && c->subbatch < c->B ? c->subbatch : c->B;
    return (c->B + k - 1) / k;
}

// one RK3 step of instances [z0, z0 + c->B) of the buffers (the caller
// shifts them); when ev is given, an event is recorded after every launch
template <typename T, typename TD>
hipError_t step_group(mhd2d_ctx* c, hipEvent_t* ev, int* nev)
{
    hipError_t e;
    auto mark = [&](int) {
        if (ev) (void)hipEventRecord(ev[(*nev)++], c->stream);
    };
    for (int s = 0; s < 3; s++) {
        e = ysweep<T, TD>(c, (s > 0 || c->pending_filter) ? 1 : 0, 1, s == 0 ? 1 : 0);
        if (e != hipSuccess) return e;
        mark(0);
        if (s == 0 && c->glm()) {
            // calcDt's Ch / Cp of the state it sees (timeAdvance.c:107-112)
            DtParams dp = dt_params(c);
            e = launch_glm_blocks<T>((const T*)c->u, c->su(), dp, c->rs, c->glm_cp, c->p.nx,
                                     c->p.ny, c->p.xblocks > 1 ? c->p.xblocks : 1, c->B,
                                     c->stream);
            if (e != hipSuccess) return e;
        }
        if (s == 0) {
            e = launch_dt_finalize(c->rs, c->p.cfl, 0, c->stream, c->B);
            if (e != hipSuccess) return e;
            mark(0);
        }
        if (c->xfuse) {
            e = launch_xsolve<TD, T>(xcarry_args<T, TD>(c), 9, c->stream, c->B);
            if (e != hipSuccess) return e;
            mark(0);
            e = launch_xrhs_rk<T, TD>(xrhs_args<T, TD>(c, s), c->stream, c->B);
            if (e != hipSuccess) return e;
            mark(0);
        } else {
            e = xderiv<T, TD>(c);
            if (e != hipSuccess) return e;
            mark(0);
            e = rhs_rk<T, TD>(c, s);
            if (e != hipSuccess) return e;
            mark(0);
        }
        e = xfilter<T>(c);
        if (e != hipSuccess) return e;
        mark(0);
    }
    c->pending_filter = true;
    return hipSuccess;
}

// One RK3 step of the batch.  With subbatch k < B the instances go through
// the step k at a time, each group's data still in the 256 MB Infinity Cache
// from kernel to kernel; the instances are independent, so the results are
// those of one launch over the batch (bit for bit).
template <typename T, typename TD>
hipError_t step(mhd2d_ctx* c, hipEvent_t* ev, int* nev)
{
    const int ng = nsub(c);
    if (ng == 1) return step_group<T, TD>(c, ev, nev);
    const int B = c->B, k = (B + ng - 1) / ng;
    const int npe = c->p.xblocks > 1 ? c->p.xblocks : 1;
    T* const u = (T*)c->u;
    TD* const dx = (TD*)c->dudx;
    TD* const dy = (TD*)c->dudy;
    TD* const rp = (TD*)c->rprev;
    TD* const lp = (TD*)c->lap;
    TD* const db = (TD*)c->divb;
    T* const xe = (T*)c->xedge;
    double* const cp = c->glm_cp;
    RunScalars* const rs = c->rs;
    const bool pf = c->pending_filter;
    hipError_t e = hipSuccess;
    for (int z0 = 0; z0 < B && e == hipSuccess; z0 += k) {
        const size_t z = (size_t)z0;
        c->B = std::min(k, B - z0);
        c->u = u + z * c->su();
        c->dudx = dx + z * c->sd();
        c->dudy = dy + z * c->sd();
        c->rprev = rp + z * c->sr();
        c->lap = lp ? lp + z * c->sl() : nullptr;
        c->divb = db ? db + z * c->sb() : nullptr;
        c->xedge = xe ? xe + z * 8 * c->edge_fs() : nullptr;
        c->glm_cp = cp ? cp + 2 * z * npe : nullptr;
        c->rs = rs + z;
        c->pending_filter = pf;
        e = step_group<T, TD>(c, ev, nev);
    }
    c->B = B;
    c->u = u;
    c->dudx = dx;
    c->dudy = dy;
    c->rprev = rp;
    c->lap = lp;
    c->divb = db;
    c->xedge = xe;
    c->glm_cp = cp;
    c->rs = rs;
    c->pending_filter = e == hipSuccess ? true : pf;
    return e;
}

template <typename T, typename TD>
hipError_t flush(mhd2d_ctx* c)
{
    if (!c->pending_filter) return hipSuccess;
    hipError_t e = ysweep<T, TD>(c, 1, 0, 1);
    c->pending_filter = false;
    return e;
}

hipError_t read_scalars(mhd2d_ctx* c, std::vector<RunScalars>& r)
{
    r.resize(c->B);
    hipError_t e = hipMemcpyAsync(r.data(), c->rs, sizeof(RunScalars) * c->B,
                                  hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    return e;
}

// After a stop (dt < 1e-8, mhd.c:131) the stopping step's y-sweep has
// already applied the pending filter and every later launch was a no-op;
// the flush y-sweep skips stopped instances by itself.
hipError_t check_stop(mhd2d_ctx* c, int inst, int* stopped, double* dt_stop)
{
    std::vector<RunScalars> r;
    hipError_t e = read_scalars(c, r);
    if (e != hipSuccess) return e;
    bool all = true;
    for (auto& q : r) all = all && q.stop;
    if (all) c->pending_filter = false;
    if (stopped) *stopped = r[inst].stop;
    if (dt_stop) *dt_stop = r[inst].dt_stop;
    return hipSuccess;
}

hipError_t do_flush(mhd2d_ctx* c)
{
    if (c->pending_filter) {
        hipError_t e = check_stop(c, 0, nullptr, nullptr);
        if (e != hipSuccess) return e;
    }
    if (c->elem == 4) return flush<float, float>(c);
    return c->delem == 8 ? flush<double, double>(c) : flush<double, float>(c);
}

hipError_t do_step(mhd2d_ctx* c, hipEvent_t* ev, int* nev)
{
    if (c->elem == 4) return step<float, float>(c, ev, nev);
    return c->delem == 8 ? step<double, double>(c, ev, nev) : step<double, float>(c, ev, nev);
}

__global__ void k_to_float(const double* __restrict__ a, float* __restrict__ b, size_t n)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        b[i] = (float)a[i];
}
__global__ void k_to_double(const float* __restrict__ a, double* __restrict__ b, size_t n)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        b[i] = (double)a[i];
}

int check_params(const mhd2d_params* p)
{
    auto pow2 = [](int v) { return v > 0 && (v & (v - 1)) == 0; };
    if (!pow2(p->nx) || !pow2(p->ny) || p->nx < 16 || p->nx > 4096 || p->ny < 16 ||
        p->ny > 4096)
        return fail(MHD2D_EINVAL, "nx and ny must be powers of two in [16, 4096]");
    if (!(p->Lx > 0) || !(p->Ly > 0) || !(p->cfl > 0) || !(p->gamma > 1))
        return fail(MHD2D_EINVAL, "Lx, Ly, cfl must be > 0 and gamma > 1");
    if (!(p->resistivity >= 0) || !(p->viscosity >= 0))
        return fail(MHD2D_EINVAL, "resistivity and viscosity must be >= 0");
    if (p->divbclean < 0 || p->divbclean > 2)
        return fail(MHD2D_EINVAL, "divbclean must be 0, 1 or 2");
    if (p->precision != MHD2D_FP64 && p->precision != MHD2D_FP32 &&
        p->precision != MHD2D_MIXED)
        return fail(MHD2D_EINVAL, "precision must be MHD2D_FP64, MHD2D_FP32 or MHD2D_MIXED");
    if (p->xblocks < 0 || p->xblocks > p->nx / 4)
        return fail(MHD2D_EINVAL, "xblocks must be in [0, nx/4] (blocks of at least 4 rows)");
    if (p->batch < 0 || p->batch > 65535)
        return fail(MHD2D_EINVAL, "batch must be in [0, 65535]");
    if (p->open_x != 0 && p->open_x != 1) return fail(MHD2D_EINVAL, "open_x must be 0 or 1");
    if (p->open_x && (p->xblocks < 2 || p->nx / p->xblocks < 8 || p->xgemm))
        return fail(MHD2D_EINVAL,
                    "an open x boundary needs xblocks >= 2 (the reference's npe >= 2), blocks "
                    "of at least 8 rows and the tridiagonal x-derivative");
    if (p->xgemm != 0 && p->xgemm != 3 && p->xgemm != 6)
        return fail(MHD2D_EINVAL, "xgemm must be 0, 3 or 6");
    if (p->xgemm && (p->precision == MHD2D_FP64 || p->xblocks > 1 || p->nx < 64 || p->ny < 64))
        return fail(MHD2D_EINVAL,
                    "the GEMM x-derivative needs precision FP32 or MIXED, one x block and "
                    "nx, ny >= 64");
    return MHD2D_OK;
}

}  // namespace

extern "C" {

const char* mhd2d_last_error(void) { return g_err.c_str(); }

int mhd2d_create(const mhd2d_params* p, mhd2d_ctx** out)
{
    if (!p || !out) return fail(MHD2D_EINVAL, "null argument");
    *out = nullptr;
    int rc = check_params(p);
    if (rc) return rc;
    mhd2d_ctx* c = new mhd2d_ctx();
    c->p = *p;
    c->elem = p->precision == MHD2D_FP32 ? 4 : 8;
    c->delem = p->precision == MHD2D_FP64 ? 8 : 4;
    c->B = p->batch > 1 ? p->batch : 1;
    c->n = (size_t)p->nx * p->ny;
    c->have_state.assign(c->B, 0);
    c->di.assign(c->B, p->di);
    hipError_t e = c->tb.build(p->nx, p->ny, p->Lx, p->Ly, c->delem, SchemeAlphas(), p->xblocks,
                               p->open_x);
    if (e == hipSuccess && p->xgemm) e = c->xg.build(c->tb, p->xgemm);
    if (e != hipSuccess) {
        delete c;
        return fail(MHD2D_EHIP, "table build", e);
    }
    c->bytes += c->xg.bytes;
    {
        const double h = std::max(c->tb.dx, c->tb.dy);
        c->cdiff = 2.0 * h * h;
    }
    const size_t fb = c->n * c->elem * c->B, db = c->n * c->delem * c->B;
    c->npartial = 1024;
    size_t sizes[] = {(p->divbclean == 2 ? 12 : 11) * fb, 11 * db, 11 * db,
                      (p->divbclean == 2 ? 9 : 8) * db};
    void** ptrs[] = {&c->u, &c->dudx, &c->dudy, &c->rprev};
    for (int k = 0; k < 4; k++) {
        e = hipMalloc(ptrs[k], sizes[k]);
        if (e != hipSuccess) {
            mhd2d_destroy(c);
            return fail(MHD2D_ENOMEM, "device allocation", e);
        }
        c->bytes += sizes[k];
    }
    if (p->divbclean == 2) {
        const size_t gb = 2 * sizeof(double) * c->B * (p->xblocks > 1 ? p->xblocks : 1);
        e = hipMalloc(&c->glm_cp, gb);
        if (e != hipSuccess) {
            mhd2d_destroy(c);
            return fail(MHD2D_ENOMEM, "device allocation", e);
        }
        c->bytes += gb;
    }
    if (p->divbclean == 1) {
        e = hipMalloc(&c->divb, 3 * db);
        if (e != hipSuccess) {
            mhd2d_destroy(c);
            return fail(MHD2D_ENOMEM, "device allocation", e);
        }
        c->bytes += 3 * db;
    }
    if (c->lap_u() || c->lap_b()) {
        e = hipMalloc(&c->lap, 6 * db);
        if (e != hipSuccess) {
            mhd2d_destroy(c);
            return fail(MHD2D_ENOMEM, "device allocation", e);
        }
        c->bytes += 6 * db;
    }
    e = hipMalloc(&c->rs, sizeof(RunScalars) * c->B);
    if (e == hipSuccess) e = hipMalloc(&c->partial, 2 * c->npartial * sizeof(double));
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    // the fused x-derivative + RHS from the start where it applies, in FP64
    // and FP32 (measured as fast or faster on every grid and batch tried;
    // the mixed form at 2048^2 runs slower fused, so it starts unfused)
    if (e == hipSuccess && xfuse_ok(c) && p->precision != MHD2D_MIXED) {
        e = ensure_xedge(c);
        c->xfuse = e == hipSuccess;
    }
    if (e != hipSuccess) {
        mhd2d_destroy(c);
        return fail(MHD2D_EHIP, "context setup", e);
    }
    *out = c;
    return MHD2D_OK;
}

void mhd2d_destroy(mhd2d_ctx* c)
{
    if (!c) return;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    if (c->graph) (void)hipGraphExecDestroy(c->graph);
    void* ptrs[] = {c->u,    c->dudx,   c->dudy, c->rprev,   c->lap,
                    c->divb, c->glm_cp, c->rs,   c->partial, c->xedge};
    for (void* q : ptrs)
        if (q) (void)hipFree(q);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    c->tb.release();
    delete c;
}

size_t mhd2d_device_bytes(const mhd2d_ctx* c) { return c ? c->bytes : 0; }

int mhd2d_copy_probe(size_t bytes, int iters, double* gbps)
{
    if (!gbps || bytes < 16 || iters < 1) return fail(MHD2D_EINVAL, "bad argument");
    bytes &= ~(size_t)15;
    // released on every return, also after a failed call
    struct Res {
        DevTemp<char> a, b;
        hipStream_t s = nullptr;
        hipEvent_t e0 = nullptr, e1 = nullptr;
        ~Res()
        {
            if (e0) (void)hipEventDestroy(e0);
            if (e1) (void)hipEventDestroy(e1);
            if (s) (void)hipStreamDestroy(s);
        }
    } r;
    HIPCHK(hipMalloc(&r.a.p, bytes));
    HIPCHK(hipMalloc(&r.b.p, bytes));
    HIPCHK(hipStreamCreateWithFlags(&r.s, hipStreamNonBlocking));
    HIPCHK(hipMemsetAsync(r.a.p, 0, bytes, r.s));
    HIPCHK(launch_copy(r.a.p, r.b.p, bytes, r.s));  // warm-up
    HIPCHK(hipEventCreate(&r.e0));
    HIPCHK(hipEventCreate(&r.e1));
    HIPCHK(hipEventRecord(r.e0, r.s));
    for (int k = 0; k < iters; k++) HIPCHK(launch_copy(r.a.p, r.b.p, bytes, r.s));
    HIPCHK(hipEventRecord(r.e1, r.s));
    HIPCHK(hipEventSynchronize(r.e1));
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, r.e0, r.e1));
    *gbps = 2.0 * (double)bytes * iters / (ms * 1e-3) / 1e9;
    return MHD2D_OK;
}

int mhd2d_set_state_at(mhd2d_ctx* c, int inst, const double* uu)
{
    if (!c || !uu) return fail(MHD2D_EINVAL, "null argument");
    if (inst < 0 || inst >= c->B) return fail(MHD2D_EINVAL, "instance out of range");
    // the other instances keep their pending y filter: complete it first
    HIPCHK(do_flush(c));
    // host fields 0..7 -> slots 0..7; psi (host field 8, GLM) -> PSI_SLOT
    const size_t cnt = 8 * c->n;
    const double* hpsi = uu + cnt;
    if (c->elem == 8) {
        double* dst = (double*)c->u + inst * c->su();
        HIPCHK(hipMemcpyAsync(dst, uu, cnt * 8, hipMemcpyHostToDevice, c->stream));
        if (c->glm())
            HIPCHK(hipMemcpyAsync(dst + PSI_SLOT * c->n, hpsi, c->n * 8, hipMemcpyHostToDevice,
                                  c->stream));
    } else {
        DevTemp<double> t;  // freed on every return
        double*& tmp = t.p;
        const size_t tot = cnt + (c->glm() ? c->n : 0);
        HIPCHK(hipMalloc(&tmp, tot * 8));
        HIPCHK(hipMemcpyAsync(tmp, uu, tot * 8, hipMemcpyHostToDevice, c->stream));
        float* dst = (float*)c->u + inst * c->su();
        hipLaunchKernelGGL(k_to_float, dim3(2048), dim3(256), 0, c->stream, tmp, dst, cnt);
        HIPCHK(hipGetLastError());
        if (c->glm()) {
            hipLaunchKernelGGL(k_to_float, dim3(2048), dim3(256), 0, c->stream, tmp + cnt,
                               dst + PSI_SLOT * c->n, c->n);
            HIPCHK(hipGetLastError());
        }
        HIPCHK(hipStreamSynchronize(c->stream));
    }
    RunScalars z;
    std::memset(&z, 0, sizeof(z));
    z.dtmin_bits = 0x7ff0000000000000ull;
    z.di = c->di[inst];
    HIPCHK(hipMemcpyAsync(c->rs + inst, &z, sizeof(z), hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    c->have_state[inst] = 1;
    return MHD2D_OK;
}

int mhd2d_set_state(mhd2d_ctx* c, const double* uu) { return mhd2d_set_state_at(c, 0, uu); }

int mhd2d_set_di(mhd2d_ctx* c, int inst, double di)
{
    if (!c) return fail(MHD2D_EINVAL, "null argument");
    if (inst < 0 || inst >= c->B) return fail(MHD2D_EINVAL, "instance out of range");
    if (!(di >= 0)) return fail(MHD2D_EINVAL, "di must be >= 0");
    c->di[inst] = di;
    HIPCHK(hipMemcpyAsync(&c->rs[inst].di, &c->di[inst], sizeof(double), hipMemcpyHostToDevice,
                          c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return MHD2D_OK;
}

int mhd2d_get_state_at(mhd2d_ctx* c, int inst, double* uu)
{
    if (!c || !uu) return fail(MHD2D_EINVAL, "null argument");
    if (inst < 0 || inst >= c->B) return fail(MHD2D_EINVAL, "instance out of range");
    HIPCHK(do_flush(c));
    const size_t cnt = 8 * c->n;
    if (c->elem == 8) {
        const double* src = (const double*)c->u + inst * c->su();
        HIPCHK(hipMemcpyAsync(uu, src, cnt * 8, hipMemcpyDeviceToHost, c->stream));
        if (c->glm())
            HIPCHK(hipMemcpyAsync(uu + cnt, src + PSI_SLOT * c->n, c->n * 8,
                                  hipMemcpyDeviceToHost, c->stream));
    } else {
        DevTemp<double> t;  // freed on every return
        double*& tmp = t.p;
        const size_t tot = cnt + (c->glm() ? c->n : 0);
        HIPCHK(hipMalloc(&tmp, tot * 8));
        const float* src = (const float*)c->u + inst * c->su();
        hipLaunchKernelGGL(k_to_double, dim3(2048), dim3(256), 0, c->stream, src, tmp, cnt);
        HIPCHK(hipGetLastError());
        if (c->glm()) {
            hipLaunchKernelGGL(k_to_double, dim3(2048), dim3(256), 0, c->stream,
                               src + PSI_SLOT * c->n, tmp + cnt, c->n);
            HIPCHK(hipGetLastError());
        }
        HIPCHK(hipMemcpyAsync(uu, tmp, tot * 8, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
    }
    HIPCHK(hipStreamSynchronize(c->stream));
    return MHD2D_OK;
}

int mhd2d_get_state(mhd2d_ctx* c, double* uu) { return mhd2d_get_state_at(c, 0, uu); }

namespace {
int check_ready(const mhd2d_ctx* c)
{
    for (char h : c->have_state)
        if (!h) return fail(MHD2D_EINVAL, "every instance needs mhd2d_set_state first");
    return MHD2D_OK;
}
}  // namespace

int mhd2d_run(mhd2d_ctx* c, int nsteps)
{
    if (!c || nsteps < 0) return fail(MHD2D_EINVAL, "bad argument");
    if (int rc = check_ready(c)) return rc;
    for (int k = 0; k < nsteps; k++) HIPCHK(do_step(c, nullptr, nullptr));
    return MHD2D_OK;
}

int mhd2d_run_graph(mhd2d_ctx* c, int nsteps)
{
    if (!c || nsteps < 0) return fail(MHD2D_EINVAL, "bad argument");
    if (int rc = check_ready(c)) return rc;
    if (nsteps == 0) return MHD2D_OK;
    int k = 0;
    if (!c->pending_filter) {  // the first step differs (no pending y filter)
        HIPCHK(do_step(c, nullptr, nullptr));
        k = 1;
    }
    if (k == nsteps) return MHD2D_OK;
    if (!c->graph) {
        hipGraph_t g;
        HIPCHK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        hipError_t e = do_step(c, nullptr, nullptr);
        hipError_t e2 = hipStreamEndCapture(c->stream, &g);
        if (e != hipSuccess) return fail(MHD2D_EHIP, "capture", e);
        HIPCHK(e2);
        HIPCHK(hipGraphInstantiate(&c->graph, g, nullptr, nullptr, 0));
        HIPCHK(hipGraphDestroy(g));
    }
    for (; k < nsteps; k++) HIPCHK(hipGraphLaunch(c->graph, c->stream));
    return MHD2D_OK;
}

int mhd2d_sync(mhd2d_ctx* c)
{
    if (!c) return fail(MHD2D_EINVAL, "null argument");
    std::vector<RunScalars> r;
    HIPCHK(read_scalars(c, r));
    for (int i = 0; i < c->B; i++)
        if (r[i].nonfinite)
            return fail(MHD2D_ENONFINITE, "a CFL limit was not finite (state blew up)");
    return MHD2D_OK;
}

int mhd2d_get_scalars(mhd2d_ctx* c, double* dt, double* time, long long* steps)
{
    return mhd2d_get_scalars_at(c, 0, dt, time, steps);
}

int mhd2d_get_scalars_at(mhd2d_ctx* c, int inst, double* dt, double* time, long long* steps)
{
    if (!c) return fail(MHD2D_EINVAL, "null argument");
    if (inst < 0 || inst >= c->B) return fail(MHD2D_EINVAL, "instance out of range");
    RunScalars r;
    HIPCHK(hipStreamSynchronize(c->stream));
    HIPCHK(hipMemcpy(&r, c->rs + inst, sizeof(r), hipMemcpyDeviceToHost));
    // rs->time holds the sum of the dts of all steps but the last
    if (dt) *dt = r.dt;
    if (time) *time = r.step > 0 ? r.time + r.dt : 0.0;
    if (steps) *steps = r.step;
    return MHD2D_OK;
}

int mhd2d_stopped(mhd2d_ctx* c, double* dt_stop) { return mhd2d_stopped_at(c, 0, dt_stop); }

int mhd2d_stopped_at(mhd2d_ctx* c, int inst, double* dt_stop)
{
    if (!c) return fail(MHD2D_EINVAL, "null argument");
    if (inst < 0 || inst >= c->B) return fail(MHD2D_EINVAL, "instance out of range");
    int st = 0;
    HIPCHK(check_stop(c, inst, &st, dt_stop));
    return st;
}

int mhd2d_get_divb(mhd2d_ctx* c, double* divb)
{
    if (!c || !divb) return fail(MHD2D_EINVAL, "null argument");
    HIPCHK(do_flush(c));
    const size_t n = c->n;
    // dBx/dx + dBy/dy of the current state of instance 0, through its E
    // fields (scratch: the next y-sweep recomputes E)
    if (c->elem == 8) {
        double* u = (double*)c->u;
        double* sc = u + 8 * n;
        XSolveArgs<double> xa;
        xa.in[0] = u + BX * n;
        xa.out[0] = sc;
        xa.st = c->tb.st[XD1];
        xa.tb = c->tb.xt[XD1];
        xa.kind = XD1;
        xa.nx = c->p.nx;
        xa.ny = c->p.ny;
        HIPCHK(launch_xsolve<double>(xa, 1, c->stream));
        YApplyArgs<double> ya;
        ya.in[0] = u + BY * n;
        ya.out[0] = sc + n;
        ya.tb = c->tb.yt;
        ya.nv = 1;
        ya.mode = Y_D1;
        ya.nx = c->p.nx;
        ya.ny = c->p.ny;
        HIPCHK(launch_yapply<double>(ya, c->stream));
        double* s3 = u + (c->glm() ? 11 : 10) * n;  // not psi
        HIPCHK(launch_add<double>(sc, sc + n, s3, n, c->stream));
        HIPCHK(hipMemcpyAsync(divb, s3, n * 8, hipMemcpyDeviceToHost, c->stream));
    } else {
        float* u = (float*)c->u;
        float* sc = u + 8 * n;
        XSolveArgs<float> xa;
        xa.in[0] = u + BX * n;
        xa.out[0] = sc;
        xa.st = c->tb.st[XD1];
        xa.tb = c->tb.xt[XD1];
        xa.kind = XD1;
        xa.nx = c->p.nx;
        xa.ny = c->p.ny;
        HIPCHK(launch_xsolve<float>(xa, 1, c->stream));
        YApplyArgs<float> ya;
        ya.in[0] = u + BY * n;
        ya.out[0] = sc + n;
        ya.tb = c->tb.yt;
        ya.nv = 1;
        ya.mode = Y_D1;
        ya.nx = c->p.nx;
        ya.ny = c->p.ny;
        HIPCHK(launch_yapply<float>(ya, c->stream));
        float* s3 = u + (c->glm() ? 11 : 10) * n;  // not psi
        HIPCHK(launch_add<float>(sc, sc + n, s3, n, c->stream));
        std::vector<float> tmp(n);
        HIPCHK(hipMemcpyAsync(tmp.data(), s3, n * 4, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        for (size_t i = 0; i < n; i++) divb[i] = tmp[i];
    }
    HIPCHK(hipStreamSynchronize(c->stream));
    return MHD2D_OK;
}

int mhd2d_diagnostics(mhd2d_ctx* c, double* energy, double* max_divb)
{
    return mhd2d_diagnostics_at(c, 0, energy, max_divb);
}

int mhd2d_diagnostics_at(mhd2d_ctx* c, int inst, double* energy, double* max_divb)
{
    if (!c) return fail(MHD2D_EINVAL, "null argument");
    if (inst < 0 || inst >= c->B) return fail(MHD2D_EINVAL, "instance out of range");
    HIPCHK(do_flush(c));
    // d(Bx)/dx and d(By)/dy of the instance's current state into its E
    // fields (scratch: the next y-sweep recomputes E)
    const size_t n = c->n;
    int nb = c->npartial;
    if (c->elem == 8) {
        double* u = (double*)c->u + inst * c->su();
        double* scratch = u + 8 * n;
        XSolveArgs<double> xa;
        xa.in[0] = u + BX * n;
        xa.out[0] = scratch;
        xa.st = c->tb.st[XD1];
        xa.tb = c->tb.xt[XD1];
        xa.kind = XD1;
        xa.nx = c->p.nx;
        xa.ny = c->p.ny;
        HIPCHK(launch_xsolve<double>(xa, 1, c->stream));
        YApplyArgs<double> ya;
        ya.in[0] = u + BY * n;
        ya.out[0] = scratch + n;
        ya.tb = c->tb.yt;
        ya.nv = 1;
        ya.mode = Y_D1;
        ya.nx = c->p.nx;
        ya.ny = c->p.ny;
        HIPCHK(launch_yapply<double>(ya, c->stream));
        HIPCHK(launch_diag<double>(u, scratch, scratch + n, c->partial, nb, n, c->p.gamma,
                                   c->stream));
    } else {
        float* u = (float*)c->u + inst * c->su();
        float* scratch = u + 8 * n;
        XSolveArgs<float> xa;
        xa.in[0] = u + BX * n;
        xa.out[0] = scratch;
        xa.st = c->tb.st[XD1];
        xa.tb = c->tb.xt[XD1];
        xa.kind = XD1;
        xa.nx = c->p.nx;
        xa.ny = c->p.ny;
        HIPCHK(launch_xsolve<float>(xa, 1, c->stream));
        YApplyArgs<float> ya;
        ya.in[0] = u + BY * n;
        ya.out[0] = scratch + n;
        ya.tb = c->tb.yt;
        ya.nv = 1;
        ya.mode = Y_D1;
        ya.nx = c->p.nx;
        ya.ny = c->p.ny;
        HIPCHK(launch_yapply<float>(ya, c->stream));
        HIPCHK(launch_diag<float>(u, scratch, scratch + n, c->partial, nb, n, c->p.gamma,
                                  c->stream));
    }
    std::vector<double> h(2 * nb);
    HIPCHK(hipMemcpyAsync(h.data(), c->partial, h.size() * sizeof(double),
                          hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    double e = 0.0, m = 0.0;
    for (int b = 0; b < nb; b++) {
        e += h[2 * b];
        m = std::fmax(m, h[2 * b + 1]);
    }
    if (energy) *energy = e * c->tb.dx * c->tb.dy;
    if (max_divb) *max_divb = m;
    return MHD2D_OK;
}

}  // extern "C"

namespace {
template <typename T, typename TD>
int xderiv_host(mhd2d_ctx* c, int nv, const double* in, double* out)
{
    const size_t cnt = (size_t)nv * c->n;
    double* dd = nullptr;
    T* tin = nullptr;
    TD* tout = nullptr;
    hipError_t e = hipMalloc(&dd, cnt * sizeof(double));
    if (e == hipSuccess) e = hipMalloc(&tin, cnt * sizeof(T));
    if (e == hipSuccess) e = hipMalloc(&tout, cnt * sizeof(TD));
    if (e == hipSuccess) e = hipMemcpyAsync(dd, in, cnt * 8, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) {
        if constexpr (std::is_same<T, double>::value)
            e = hipMemcpyAsync(tin, dd, cnt * 8, hipMemcpyDeviceToDevice, c->stream);
        else {
            hipLaunchKernelGGL(k_to_float, dim3(2048), dim3(256), 0, c->stream, dd, tin, cnt);
            e = hipGetLastError();
        }
    }
    if (e == hipSuccess) e = xderiv_fields<T, TD>(c, tin, tout, nv, 0, 0, 1, nullptr);
    if (e == hipSuccess) {
        if constexpr (std::is_same<TD, double>::value)
            e = hipMemcpyAsync(dd, tout, cnt * 8, hipMemcpyDeviceToDevice, c->stream);
        else {
            hipLaunchKernelGGL(k_to_double, dim3(2048), dim3(256), 0, c->stream, tout, dd, cnt);
            e = hipGetLastError();
        }
    }
    if (e == hipSuccess) e = hipMemcpyAsync(out, dd, cnt * 8, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    (void)hipFree(dd);
    (void)hipFree(tin);
    (void)hipFree(tout);
    if (e != hipSuccess) return fail(MHD2D_EHIP, "x-derivative", e);
    return MHD2D_OK;
}
// mean time of `reps` launches of f() after one untimed launch
template <typename F>
hipError_t time_launches(hipStream_t s, int reps, F f, float* ms)
{
    hipEvent_t a, b;
    hipError_t e = hipEventCreate(&a);
    if (e != hipSuccess) return e;
    e = hipEventCreate(&b);
    if (e == hipSuccess) e = f();
    if (e == hipSuccess) e = hipEventRecord(a, s);
    for (int r = 0; r < reps && e == hipSuccess; r++) e = f();
    if (e == hipSuccess) e = hipEventRecord(b, s);
    if (e == hipSuccess) e = hipEventSynchronize(b);
    float t = 0;
    if (e == hipSuccess) e = hipEventElapsedTime(&t, a, b);
    *ms = t / reps;
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    return e;
}

// Autotuner of the x-solve launch shapes: every valid (strip width, XCD
// order) of the derivative solves, then of the filter, is timed as whole RK3
// steps of this context's own run (grid, precision, batch, settings), each
// from the same saved state (state, R_prev, run scalars, pending filter),
// which is restored afterwards; where the fused x-derivative + RHS applies,
// once for each x path, and the faster path keeps its shapes.  A shape is
// judged where it runs: alone, the kernels' own timings ranked the XCD
// orders differently from the step.  Two interleaved rounds, each candidate
// keeps its faster mean.
template <typename T, typename TD>
hipError_t autotune_impl(mhd2d_ctx* c, int reps, int* out)
{
    struct Cand {
        int w, r;
    };
    int C0, W0, G0;
    xsolve_shape(c->p.nx, 8, &C0, &W0, &G0);
    std::vector<Cand> cands;
    for (int w : {16, 8, 32})
        for (int r : {0, 1}) {
            if (w * G0 > 1024 || c->p.ny % w) continue;  // at most 1024 threads
            cands.push_back({w, r});
        }
    hipError_t e = hipStreamSynchronize(c->stream);
    const size_t ub = c->su() * sizeof(T) * c->B, rb = c->sr() * sizeof(TD) * c->B;
    const size_t sb = sizeof(RunScalars) * c->B;
    const bool pf = c->pending_filter;
    char* saved = nullptr;
    if (e == hipSuccess) e = hipMalloc(&saved, ub + rb + sb);
    if (e == hipSuccess) e = hipMemcpyAsync(saved, c->u, ub, hipMemcpyDeviceToDevice, c->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(saved + ub, c->rprev, rb, hipMemcpyDeviceToDevice, c->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(saved + ub + rb, c->rs, sb, hipMemcpyDeviceToDevice, c->stream);
    auto restore = [&]() {
        hipError_t r = hipMemcpyAsync(c->u, saved, ub, hipMemcpyDeviceToDevice, c->stream);
        if (r == hipSuccess)
            r = hipMemcpyAsync(c->rprev, saved + ub, rb, hipMemcpyDeviceToDevice, c->stream);
        if (r == hipSuccess)
            r = hipMemcpyAsync(c->rs, saved + ub + rb, sb, hipMemcpyDeviceToDevice, c->stream);
        c->pending_filter = pf;
        return r;
    };
    // ms per step of `reps` steps after one untimed step, from the saved state
    auto time_steps = [&](float* ms) {
        hipError_t r = restore();
        if (r == hipSuccess)
            r = time_launches(c->stream, reps, [&]() { return step<T, TD>(c, nullptr, nullptr); },
                              ms);
        return r;
    };
    const int xw_f0 = c->xw_f, xr_f0 = c->xr_f;
    // shape search for one x path: the derivative shapes (the full solve,
    // or the carry solve of the fused path), then the filter shapes with
    // the chosen derivative shape; returns the best step time
    auto search = [&](int (&pk)[2][2]) {
        float best = 1e30f;
        for (int phase = 0; phase < 2 && e == hipSuccess; phase++) {
            std::vector<float> tm(cands.size(), 1e30f);
            for (size_t it = 0; it < 2 * cands.size() && e == hipSuccess; it++) {
                const size_t ci = it % cands.size();
                if (phase == 0) {
                    c->xw_d = cands[ci].w;
                    c->xr_d = cands[ci].r;
                    c->xw_f = xw_f0;
                    c->xr_f = xr_f0;
                } else {
                    c->xw_d = pk[0][0];
                    c->xr_d = pk[0][1];
                    c->xw_f = cands[ci].w;
                    c->xr_f = cands[ci].r;
                }
                float t = 0;
                e = time_steps(&t);
                if (e == hipSuccess) tm[ci] = std::min(tm[ci], t);
            }
            best = 1e30f;
            for (size_t ci = 0; ci < cands.size(); ci++)
                if (tm[ci] < best) {
                    best = tm[ci];
                    pk[phase][0] = cands[ci].w;
                    pk[phase][1] = cands[ci].r;
                }
        }
        return best;
    };
    // the fused x-derivative + RHS, where it applies, against the unfused
    // path, each with its own shapes (the carry solve and the full
    // derivative solve prefer different strips): the faster pair wins.  The
    // fused path moves ~13 words per cell and stage less, but its chunk
    // kernel streams less evenly than the RHS + RK alone; which wins depends
    // on the grid, the precision and the batch.
    const int fuse0 = c->xfuse;
    int pick_fuse = fuse0;
    int pick[2][2] = {{0, -1}, {0, -1}};  // [deriv, filter][width, order]
    if (e == hipSuccess && xfuse_ok(c) && ensure_xedge(c) == hipSuccess) {
        int pk[2][2][2] = {{{0, -1}, {0, -1}}, {{0, -1}, {0, -1}}};
        float tf[2] = {1e30f, 1e30f};
        for (int f = 0; f < 2 && e == hipSuccess; f++) {
            c->xfuse = f;
            tf[f] = search(pk[f]);
        }
        pick_fuse = tf[1] < tf[0] ? 1 : 0;
        std::memcpy(pick, pk[pick_fuse], sizeof(pick));
        c->xfuse = pick_fuse;
    } else {
        (void)search(pick);
    }
    // instances per launch (batches only): the whole batch, then groups of
    // half as many while a group keeps at least 2^21 cells (fewer leave the
    // chip short of work); a group size must beat the whole batch by 1 %
    int pick_sb = c->subbatch;
    if (e == hipSuccess && c->B > 1) {
        c->xw_d = pick[0][0];
        c->xr_d = pick[0][1];
        c->xw_f = pick[1][0];
        c->xr_f = pick[1][1];
        std::vector<int> sbc = {0};
        for (int k = (c->B + 1) / 2; (size_t)k * c->n >= ((size_t)1 << 21); k = (k + 1) / 2) {
            sbc.push_back(k);
            if (k == 1) break;
        }
        std::vector<float> tm(sbc.size(), 1e30f);
        for (size_t it = 0; sbc.size() > 1 && it < 2 * sbc.size() && e == hipSuccess; it++) {
            const size_t ci = it % sbc.size();
            c->subbatch = sbc[ci];
            float t = 0;
            e = time_steps(&t);
            if (e == hipSuccess) tm[ci] = std::min(tm[ci], t);
        }
        pick_sb = 0;
        float best = tm[0] * 0.99f;
        for (size_t ci = 1; ci < sbc.size(); ci++)
            if (tm[ci] < best) {
                best = tm[ci];
                pick_sb = sbc[ci];
            }
    }
    if (saved) {
        const hipError_t e2 = restore();
        const hipError_t e3 = hipStreamSynchronize(c->stream);
        (void)hipFree(saved);
        if (e == hipSuccess) e = e2 != hipSuccess ? e2 : e3;
    }
    c->pending_filter = pf;
    if (e != hipSuccess) {
        c->xw_d = c->xw_f = 0;
        c->xr_d = c->xr_f = -1;
        c->subbatch = 0;
        c->xfuse = fuse0;
        return e;
    }
    c->subbatch = pick_sb;
    c->xfuse = pick_fuse;
    c->xw_d = pick[0][0];
    c->xr_d = pick[0][1];
    c->xw_f = pick[1][0];
    c->xr_f = pick[1][1];
    if (c->graph) {  // the captured step holds the old launch shapes
        (void)hipGraphExecDestroy(c->graph);
        c->graph = nullptr;
    }
    if (out) {
        out[0] = c->xw_d;
        out[1] = c->xr_d;
        out[2] = c->xw_f;
        out[3] = c->xr_f;
    }
    return hipSuccess;
}
}  // namespace

extern "C" {

int mhd2d_autotune(mhd2d_ctx* c, int reps, int* choice)
{
    if (!c || reps < 1) return fail(MHD2D_EINVAL, "bad argument");
    if (c->p.xgemm) return fail(MHD2D_EINVAL, "the GEMM x-derivative has no shapes to tune");
    if (int rc = check_ready(c)) return rc;  // the candidates are timed as steps of the run
    hipError_t e;
    if (c->elem == 4)
        e = autotune_impl<float, float>(c, reps, choice);
    else if (c->delem == 8)
        e = autotune_impl<double, double>(c, reps, choice);
    else
        e = autotune_impl<double, float>(c, reps, choice);
    if (e != hipSuccess) return fail(MHD2D_EHIP, "autotune", e);
    return MHD2D_OK;
}

int mhd2d_set_subbatch(mhd2d_ctx* c, int k)
{
    if (!c || k < 0) return fail(MHD2D_EINVAL, "bad argument");
    c->subbatch = k >= c->B ? 0 : k;
    if (c->graph) {  // the captured step holds the old launches
        (void)hipGraphExecDestroy(c->graph);
        c->graph = nullptr;
    }
    return MHD2D_OK;
}

int mhd2d_set_xfuse(mhd2d_ctx* c, int on)
{
    if (!c || on < 0 || on > 1) return fail(MHD2D_EINVAL, "bad argument");
    if (on && !xfuse_ok(c))
        return fail(MHD2D_EINVAL, "fused x-derivative + RHS: needs ideal MHD on one periodic "
                                  "x block, the tridiagonal x-derivative and a grid it tiles");
    if (on) HIPCHK(ensure_xedge(c));
    c->xfuse = on;
    if (c->graph) {  // the captured step holds the old launches
        (void)hipGraphExecDestroy(c->graph);
        c->graph = nullptr;
    }
    return MHD2D_OK;
}

int mhd2d_get_xfuse(const mhd2d_ctx* c)
{
    if (!c) return fail(MHD2D_EINVAL, "bad argument");
    return c->xfuse;
}

int mhd2d_get_subbatch(const mhd2d_ctx* c)
{
    if (!c) return fail(MHD2D_EINVAL, "bad argument");
    return c->subbatch > 0 && c->subbatch < c->B ? c->subbatch : c->B;
}

long long mhd2d_xgemm_panels(const mhd2d_ctx* c) { return c && c->p.xgemm ? c->xg.panels_total : 0; }

int mhd2d_xderiv(mhd2d_ctx* c, int nv, const double* in, double* out)
{
    if (!c || !in || !out || nv < 1 || nv > 11) return fail(MHD2D_EINVAL, "bad argument");
    if (c->elem == 4) return xderiv_host<float, float>(c, nv, in, out);
    if (c->delem == 8) return xderiv_host<double, double>(c, nv, in, out);
    return xderiv_host<double, float>(c, nv, in, out);
}

int mhd2d_time_run(mhd2d_ctx* c, int nsteps, int use_graph, double* ms)
{
    if (!c || !ms) return fail(MHD2D_EINVAL, "null argument");
    hipEvent_t a, b;
    HIPCHK(hipEventCreate(&a));
    HIPCHK(hipEventCreate(&b));
    HIPCHK(hipStreamSynchronize(c->stream));
    HIPCHK(hipEventRecord(a, c->stream));
    int rc = use_graph ? mhd2d_run_graph(c, nsteps) : mhd2d_run(c, nsteps);
    if (rc) return rc;
    HIPCHK(hipEventRecord(b, c->stream));
    HIPCHK(hipEventSynchronize(b));
    float f = 0;
    HIPCHK(hipEventElapsedTime(&f, a, b));
    *ms = f;
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    return MHD2D_OK;
}

int mhd2d_profile_run(mhd2d_ctx* c, int nsteps, double* out_ms, int* out_launches)
{
    if (!c || !out_ms || nsteps < 1) return fail(MHD2D_EINVAL, "bad argument");
    // per step and instance group: 3 stages x (ysweep, xderiv, rhs_rk,
    // xfilter) + one dt finalize; event k+1 closes the launch opened by event
    // k.  A class's time sums its launches over the groups of a step, so it
    // is the time of one launch over the whole batch.
    const int per_step = 3 * 4 + 1, ng = nsub(c);
    std::vector<hipEvent_t> ev(per_step * ng + 1);
    for (auto& e : ev) HIPCHK(hipEventCreate(&e));
    double acc[5] = {0, 0, 0, 0, 0};
    int cnt[5] = {0, 0, 0, 0, 0};
    // classes in launch order within a step
    const int cls0[] = {0, 4, 1, 2, 3};
    const int cls[] = {0, 1, 2, 3};
    for (int k = 0; k < nsteps; k++) {
        HIPCHK(hipEventRecord(ev[0], c->stream));
        int nev = 1;
        HIPCHK(do_step(c, ev.data(), &nev));
        HIPCHK(hipStreamSynchronize(c->stream));
        int idx = 0;
        for (int g = 0; g < ng; g++)
            for (int s = 0; s < 3; s++) {
                const int* cl = s == 0 ? cls0 : cls;
                const int m = s == 0 ? 5 : 4;
                for (int q = 0; q < m; q++, idx++) {
                    float f = 0;
                    HIPCHK(hipEventElapsedTime(&f, ev[idx], ev[idx + 1]));
                    acc[cl[q]] += f;
                    if (g == 0) cnt[cl[q]] += 1;
                }
            }
    }
    for (int q = 0; q < 5; q++) {
        out_ms[q] = cnt[q] ? acc[q] / cnt[q] : 0.0;
        if (out_launches) out_launches[q] = cnt[q];
    }
    for (auto& e : ev) (void)hipEventDestroy(e);
    return MHD2D_OK;
}

}  // extern "C"

// ------------------------------------------------------------------ ensembles
// The shape groups can advance per group (each group's batched context on
// its own stream, forked from and joined back into the master stream) or
// together: one launch per kernel class and stage over every group
// (k_*_multi), with the groups' argument sets in a device table.
struct EnsMulti {
    void* tab = nullptr;       // device: the argument sets below
    size_t off_y[3] = {0, 0, 0};  // y-sweeps: stage 0 without / with the pending filter, stages 1-2
    size_t off_xd = 0, off_x2 = 0, off_xf = 0, off_dt = 0;
    size_t off_r[3] = {0, 0, 0};  // RHS + RK of each stage
    // the fused x-derivative + RHS (every group ideal MHD on one periodic
    // block): carry solve, fused kernel of each stage, x-filter with the
    // chunk edge rows
    size_t off_xc = 0, off_xfe = 0;
    size_t off_xr[3] = {0, 0, 0};
    MultiMap my, mxd, mx2, mxf, mr, mdt, mxr;
    size_t xdyn = 0;           // dynamic LDS of the x-solves: the largest group's tables
    int lap = 0;               // viscosity / resistivity in every group
    int fused = 0;             // the fused x path
};

struct mhd2d_ens {
    std::vector<mhd2d_ctx*> groups;
    std::vector<std::pair<int, int>> where;  // instance -> (group, index in group)
    hipStream_t master = nullptr;
    std::vector<hipEvent_t> ev;             // [0] fork, [1 + g] join of group g
    hipGraphExec_t graph = nullptr;
    EnsMulti multi;
    bool multi_ok = false;     // every group fits the one-launch path
    int mode = MHD2D_ENS_PER_GROUP;
};

namespace {

// Argument tables and block maps of the one-launch path; false when some
// group does not fit it (cleaning, open x, x blocks, GEMM x-derivative, a
// line length outside the multi-shape kernels, or groups that differ in
// having the Laplacian terms).  Groups go largest first, so the longest
// blocks start first (the x-solves take the groups in reverse order, to
// start on the lines the kernel before them touched last).
template <typename T, typename TD>
hipError_t build_multi_t(mhd2d_ens* e, bool* ok)
{
    *ok = false;
    const int ng = (int)e->groups.size();
    if (ng < 1 || ng > MULTI_MAXG) return hipSuccess;
    std::vector<int> ord(ng);
    for (int g = 0; g < ng; g++) ord[g] = g;
    std::sort(ord.begin(), ord.end(), [&](int a, int b) {
        const mhd2d_ctx *x = e->groups[a], *y = e->groups[b];
        return x->n * x->B > y->n * y->B;
    });
    const int lap = (e->groups[0]->lap_u() || e->groups[0]->lap_b()) ? 1 : 0;
    for (auto* c : e->groups) {
        const int cl = (c->lap_u() || c->lap_b()) ? 1 : 0;
        if (c->p.divbclean || c->p.open_x || c->p.xblocks > 1 || c->p.xgemm || cl != lap)
            return hipSuccess;
        if (!ysweep_multi_rows(c->p.ny, c->elem) || xsolve_multi_code(c->p.nx, c->p.ny, 0, 0) < 0)
            return hipSuccess;
    }
    EnsMulti& M = e->multi;
    M.lap = lap;
    // the fused x path where every group takes it; not in the mixed form,
    // whose contexts keep the full x-derivative solve by default (the fused
    // one-launch step measured 2.5 % slower there on the uneven ensemble)
    bool fused = true;
    for (auto* c : e->groups) fused = fused && xfuse_ok(c) && c->p.precision != MHD2D_MIXED;
    if (fused) {
        for (auto* c : e->groups) {
            const hipError_t r = ensure_xedge(c);
            if (r != hipSuccess) return r;
        }
    }
    std::vector<XSolveArgs<TD, T>> xc;
    std::vector<XRhsArgs<T, TD>> xr[3];
    std::vector<XSolveArgs<T>> xfe;
    MultiMap mxr;
    mxr.ng = ng;
    std::vector<YSweepArgs<T, TD>> ya[3];
    std::vector<XSolveArgs<TD, T>> xd, x2;
    std::vector<XSolveArgs<T>> xf;
    std::vector<RhsRkArgs<T, TD>> ra[3];
    std::vector<DtMultiArgs> dta;
    MultiMap my, mxd, mx2, mxf, mr, mdt;
    for (MultiMap* m : {&my, &mxd, &mx2, &mxf, &mr, &mdt}) m->ng = ng;
    size_t xdyn = 0;
    static const int ideal[9] = {RHO, UX, UY, UZ, BY, BZ, PR, 8, 9};
    for (int k = 0; k < ng; k++) {
        mhd2d_ctx* c = e->groups[ord[k]];
        ya[0].push_back(ysweep_args<T, TD>(c, 0, 1, 1));
        ya[1].push_back(ysweep_args<T, TD>(c, 1, 1, 1));
        ya[2].push_back(ysweep_args<T, TD>(c, 1, 1, 0));
        for (int st = 0; st < 3; st++) ra[st].push_back(rhs_rk_args<T, TD>(c, st));
        dta.push_back(DtMultiArgs{c->rs, c->p.cfl});
        int C, W;
        const int code = xsolve_multi_code(c->p.nx, c->p.ny, &C, &W);
        const int strips = c->p.ny / W;
        // x-derivative: the ideal periodic fields (no dBx/dx)
        XSolveArgs<TD, T> a;
        for (int v = 0; v < 9; v++) {
            a.in[v] = (const T*)c->u + ideal[v] * c->n;
            a.out[v] = (TD*)c->dudx + ideal[v] * c->n;
        }
        a.rs = c->rs;
        a.st = c->tb.st[XD1];
        a.tb = c->tb.xt[XD1];
        a.kind = XD1;
        a.nx = c->p.nx;
        a.ny = c->p.ny;
        a.bs_in = c->su();
        a.bs_out = c->sd();
        xd.push_back(a);
        int nv2 = 0;
        if (lap) {
            const int f0 = c->lap_u() ? UX : BX, f1 = c->lap_b() ? BZ : UZ;
            XSolveArgs<TD, T> b;
            for (int v = 0; v <= f1 - f0; v++) {
                b.in[v] = (const T*)c->u + (f0 + v) * c->n;
                b.out[v] = (TD*)c->lap + (f0 + v - UX) * c->n;
            }
            b.rs = c->rs;
            b.st = c->tb.st[XD2];
            b.tb = c->tb.xt[XD2];
            b.kind = XD2;
            b.nx = c->p.nx;
            b.ny = c->p.ny;
            b.accumulate = 1;
            b.bs_in = c->su();
            b.bs_out = c->sl();
            x2.push_back(b);
            nv2 = f1 - f0 + 1;
        }
        xf.push_back(xfilter_args<T>(c, false));
        if (fused) {
            xc.push_back(xcarry_args<T, TD>(c));
            for (int st = 0; st < 3; st++) xr[st].push_back(xrhs_args<T, TD>(c, st));
            xfe.push_back(xfilter_args<T>(c, true));
        }
        xdyn = std::max(xdyn, 3 * (size_t)c->p.nx * std::max(sizeof(T), sizeof(TD)));
        const int rows = ysweep_multi_rows(c->p.ny, c->elem);
        const int B = c->B;
        auto set = [&](MultiMap& m, int per, int aux, int cd) {
            m.nb[k] = B;
            m.per[k] = per;
            m.aux[k] = aux;
            m.code[k] = cd;
            m.first[k + 1] = m.first[k] + per * B;
        };
        set(my, (c->p.nx + rows - 1) / rows, 0, 0);
        set(mxd, strips * 9, strips, code);
        set(mx2, strips * (nv2 > 0 ? nv2 : 1), strips, code);
        set(mxf, strips * 8, strips, code);
        {
            // blocks per instance of the RHS + RK: as launch_rhs_rk sizes
            // its grid (W cells per thread, 256 threads)
            const size_t q = c->n / (16 / sizeof(T));
            size_t b = (q + 255) / 256;
            if (b > 65535u * 16u) b = 65535u * 16u;
            set(mr, (int)b, 0, 0);
        }
        set(mdt, 1, 0, 0);
        if (fused) {
            // the fused kernel's tiles: column tiles x chunks per instance
            const int tiles = c->p.ny / xrhs_tile_columns(c->p.nx, c->elem);
            set(mxr, tiles * (c->p.nx / C), tiles, C);
        }
    }
    // one device table: each argument set 256-byte aligned
    size_t off = 0;
    auto take = [&](size_t bytes) {
        const size_t o = off;
        off += (bytes + 255) & ~(size_t)255;
        return o;
    };
    for (int v = 0; v < 3; v++) M.off_y[v] = take(ya[v].size() * sizeof(YSweepArgs<T, TD>));
    M.off_xd = take(xd.size() * sizeof(XSolveArgs<TD, T>));
    M.off_x2 = take(std::max<size_t>(1, x2.size()) * sizeof(XSolveArgs<TD, T>));
    M.off_xf = take(xf.size() * sizeof(XSolveArgs<T>));
    for (int st = 0; st < 3; st++) M.off_r[st] = take(ra[st].size() * sizeof(RhsRkArgs<T, TD>));
    M.off_dt = take(dta.size() * sizeof(DtMultiArgs));
    if (fused) {
        M.off_xc = take(xc.size() * sizeof(XSolveArgs<TD, T>));
        for (int st = 0; st < 3; st++) M.off_xr[st] = take(xr[st].size() * sizeof(XRhsArgs<T, TD>));
        M.off_xfe = take(xfe.size() * sizeof(XSolveArgs<T>));
    }
    std::vector<char> h(off, 0);
    for (int v = 0; v < 3; v++)
        std::memcpy(h.data() + M.off_y[v], ya[v].data(), ya[v].size() * sizeof(ya[v][0]));
    std::memcpy(h.data() + M.off_xd, xd.data(), xd.size() * sizeof(xd[0]));
    if (!x2.empty()) std::memcpy(h.data() + M.off_x2, x2.data(), x2.size() * sizeof(x2[0]));
    std::memcpy(h.data() + M.off_xf, xf.data(), xf.size() * sizeof(xf[0]));
    for (int st = 0; st < 3; st++)
        std::memcpy(h.data() + M.off_r[st], ra[st].data(), ra[st].size() * sizeof(ra[st][0]));
    std::memcpy(h.data() + M.off_dt, dta.data(), dta.size() * sizeof(dta[0]));
    if (fused) {
        std::memcpy(h.data() + M.off_xc, xc.data(), xc.size() * sizeof(xc[0]));
        for (int st = 0; st < 3; st++)
            std::memcpy(h.data() + M.off_xr[st], xr[st].data(), xr[st].size() * sizeof(xr[st][0]));
        std::memcpy(h.data() + M.off_xfe, xfe.data(), xfe.size() * sizeof(xfe[0]));
    }
    hipError_t r = hipMalloc(&M.tab, off);
    if (r != hipSuccess) return r;
    r = hipMemcpy(M.tab, h.data(), off, hipMemcpyHostToDevice);
    if (r != hipSuccess) return r;
    M.my = my;
    M.mxd = mxd;
    M.mx2 = mx2;
    M.mxf = mxf;
    M.mr = mr;
    M.mdt = mdt;
    M.mxr = mxr;
    M.fused = fused ? 1 : 0;
    M.xdyn = xdyn;
    *ok = true;
    return hipSuccess;
}

hipError_t build_multi(mhd2d_ens* e, bool* ok)
{
    const mhd2d_ctx* c = e->groups[0];
    if (c->elem == 4) return build_multi_t<float, float>(e, ok);
    return c->delem == 8 ? build_multi_t<double, double>(e, ok) : build_multi_t<double, float>(e, ok);
}

// one RK3 step of every group with one launch per kernel class and stage,
// on the master stream; every group must have the same pending filter
template <typename T, typename TD>
hipError_t ens_step_multi_t(mhd2d_ens* e)
{
    EnsMulti& M = e->multi;
    char* tb = (char*)M.tab;
    const bool pf = e->groups[0]->pending_filter;
    const int mode = M.lap;
    hipStream_t s = e->master;
    hipError_t r = hipSuccess;
    for (int st = 0; st < 3 && r == hipSuccess; st++) {
        const size_t oy = st > 0 ? M.off_y[2] : (pf ? M.off_y[1] : M.off_y[0]);
        r = launch_ysweep_multi<T, TD>((const YSweepArgs<T, TD>*)(tb + oy), M.my, mode, s);
        if (r == hipSuccess && st == 0)
            r = launch_dt_finalize_multi((const DtMultiArgs*)(tb + M.off_dt), M.mdt, s);
        if (M.fused) {
            if (r == hipSuccess)
                r = launch_xsolve_multi<TD, T>((const XSolveArgs<TD, T>*)(tb + M.off_xc), M.mxd,
                                               XD1_CARRY, M.xdyn, s);
            if (r == hipSuccess)
                r = launch_xrhs_rk_multi<T, TD>((const XRhsArgs<T, TD>*)(tb + M.off_xr[st]), M.mxr,
                                                s);
            if (r == hipSuccess)
                r = launch_xsolve_multi<T, T>((const XSolveArgs<T>*)(tb + M.off_xfe), M.mxf,
                                              XFILT_EDGE, M.xdyn, s);
            continue;
        }
        if (r == hipSuccess)
            r = launch_xsolve_multi<TD, T>((const XSolveArgs<TD, T>*)(tb + M.off_xd), M.mxd, XD1,
                                           M.xdyn, s);
        if (r == hipSuccess && M.lap)
            r = launch_xsolve_multi<TD, T>((const XSolveArgs<TD, T>*)(tb + M.off_x2), M.mx2, XD2,
                                           M.xdyn, s);
        if (r == hipSuccess)
            r = launch_rhs_rk_multi<T, TD>((const RhsRkArgs<T, TD>*)(tb + M.off_r[st]), M.mr,
                                           mode, s);
        if (r == hipSuccess)
            r = launch_xsolve_multi<T, T>((const XSolveArgs<T>*)(tb + M.off_xf), M.mxf, XFILT,
                                          M.xdyn, s);
    }
    if (r == hipSuccess)
        for (auto* c : e->groups) c->pending_filter = true;
    return r;
}

hipError_t ens_step_multi(mhd2d_ens* e)
{
    const mhd2d_ctx* c = e->groups[0];
    if (c->elem == 4) return ens_step_multi_t<float, float>(e);
    return c->delem == 8 ? ens_step_multi_t<double, double>(e)
                         : ens_step_multi_t<double, float>(e);
}

bool same_group(const mhd2d_params& a, const mhd2d_params& b)
{
    return a.nx == b.nx && a.ny == b.ny && a.Lx == b.Lx && a.Ly == b.Ly && a.gamma == b.gamma &&
           a.cfl == b.cfl && a.hall == b.hall && a.resistivity == b.resistivity &&
           a.viscosity == b.viscosity && a.divbclean == b.divbclean &&
           a.precision == b.precision && a.xblocks == b.xblocks;
}

// the master stream waits for every group's queued work
hipError_t join_into_master(mhd2d_ens* e)
{
    for (size_t g = 0; g < e->groups.size(); g++) {
        hipError_t r = hipEventRecord(e->ev[1 + g], e->groups[g]->stream);
        if (r == hipSuccess) r = hipStreamWaitEvent(e->master, e->ev[1 + g], 0);
        if (r != hipSuccess) return r;
    }
    return hipSuccess;
}

// every group stream waits for the master stream's queued work
hipError_t fork_from_master(mhd2d_ens* e)
{
    hipError_t r = hipEventRecord(e->ev[0], e->master);
    for (auto* c : e->groups)
        if (r == hipSuccess) r = hipStreamWaitEvent(c->stream, e->ev[0], 0);
    return r;
}

// Before a one-launch step every group needs the same pending filter: a
// group whose y filter is pending while another's is not (a state was set
// on one group only) applies it now, on the master stream.  Not during
// graph capture: the caller arranges the first (eager) step.
hipError_t ens_align_filters(mhd2d_ens* e)
{
    bool any_fresh = false;
    for (auto* c : e->groups) any_fresh = any_fresh || !c->pending_filter;
    if (!any_fresh) return hipSuccess;
    for (auto* c : e->groups) {
        if (!c->pending_filter) continue;
        hipStream_t own = c->stream;
        c->stream = e->master;
        hipError_t r = do_flush(c);
        c->stream = own;
        if (r != hipSuccess) return r;
    }
    return hipSuccess;
}

hipError_t ens_step_all(mhd2d_ens* e)
{
    if (e->mode == MHD2D_ENS_ONE_LAUNCH) return ens_step_multi(e);
    for (auto* c : e->groups) {
        hipError_t r = do_step(c, nullptr, nullptr);
        if (r != hipSuccess) return r;
    }
    return hipSuccess;
}

// kernel launches of one step (after the first) in the current mode
int ens_launches(const mhd2d_ens* e)
{
    if (e->mode == MHD2D_ENS_ONE_LAUNCH) return 13 + (e->multi.lap ? 3 : 0);
    int n = 0;
    for (const auto* c : e->groups) {
        int per = 13;
        if (c->lap) per += 3;
        if (c->divb) per += 9;
        if (c->glm()) per += 1;
        if (c->p.open_x) per += 3;
        n += per * nsub(c);
    }
    return n;
}

}  // namespace

extern "C" {

int mhd2d_ens_create(const mhd2d_params* p, int count, mhd2d_ens** out)
{
    if (!p || !out || count < 1) return fail(MHD2D_EINVAL, "bad argument");
    *out = nullptr;
    for (int i = 0; i < count; i++) {
        if (int rc = check_params(&p[i])) return rc;
        if (p[i].precision != p[0].precision)
            return fail(MHD2D_EINVAL, "all runs of an ensemble share one precision");
    }
    // group the runs: first member's parameters, di per run
    std::vector<std::vector<int>> members;
    for (int i = 0; i < count; i++) {
        size_t g = 0;
        for (; g < members.size(); g++)
            if (same_group(p[members[g][0]], p[i])) break;
        if (g == members.size()) members.push_back({});
        members[g].push_back(i);
    }
    mhd2d_ens* e = new mhd2d_ens();
    e->where.resize(count);
    for (size_t g = 0; g < members.size(); g++) {
        mhd2d_params q = p[members[g][0]];
        q.batch = (int)members[g].size();
        mhd2d_ctx* c = nullptr;
        int rc = mhd2d_create(&q, &c);
        if (rc) {
            mhd2d_ens_destroy(e);
            return rc;
        }
        e->groups.push_back(c);
        for (size_t k = 0; k < members[g].size(); k++) {
            const int i = members[g][k];
            e->where[i] = {(int)g, (int)k};
            rc = mhd2d_set_di(c, (int)k, p[i].di);
            if (rc) {
                mhd2d_ens_destroy(e);
                return rc;
            }
        }
    }
    hipError_t r = hipStreamCreateWithFlags(&e->master, hipStreamNonBlocking);
    e->ev.resize(1 + e->groups.size(), nullptr);
    for (auto& v : e->ev)
        if (r == hipSuccess) r = hipEventCreateWithFlags(&v, hipEventDisableTiming);
    if (r == hipSuccess) r = build_multi(e, &e->multi_ok);
    if (r != hipSuccess) {
        mhd2d_ens_destroy(e);
        return fail(MHD2D_EHIP, "ensemble setup", r);
    }
    e->mode = e->multi_ok ? MHD2D_ENS_ONE_LAUNCH : MHD2D_ENS_PER_GROUP;
    *out = e;
    return MHD2D_OK;
}

int mhd2d_ens_set_launch(mhd2d_ens* e, int mode)
{
    if (!e) return fail(MHD2D_EINVAL, "null argument");
    if (mode != MHD2D_ENS_PER_GROUP && mode != MHD2D_ENS_ONE_LAUNCH)
        return fail(MHD2D_EINVAL, "mode must be MHD2D_ENS_PER_GROUP or MHD2D_ENS_ONE_LAUNCH");
    if (mode == MHD2D_ENS_ONE_LAUNCH && !e->multi_ok)
        return fail(MHD2D_EINVAL,
                    "one launch per kernel class needs every group periodic, without div B "
                    "cleaning, x blocks or the GEMM x-derivative, with the same source terms, "
                    "nx in [256, 4096] and ny in [128, 1024] (FP32: [128, 2048])");
    if (mode != e->mode && e->graph) {
        (void)hipGraphExecDestroy(e->graph);
        e->graph = nullptr;
    }
    e->mode = mode;
    return MHD2D_OK;
}

int mhd2d_ens_get_launch(const mhd2d_ens* e) { return e ? e->mode : -1; }

int mhd2d_ens_launches_per_step(const mhd2d_ens* e) { return e ? ens_launches(e) : -1; }

void mhd2d_ens_destroy(mhd2d_ens* e)
{
    if (!e) return;
    if (e->master) (void)hipStreamSynchronize(e->master);
    if (e->graph) (void)hipGraphExecDestroy(e->graph);
    for (auto v : e->ev)
        if (v) (void)hipEventDestroy(v);
    for (auto* c : e->groups) mhd2d_destroy(c);
    if (e->master) (void)hipStreamDestroy(e->master);
    if (e->multi.tab) (void)hipFree(e->multi.tab);
    delete e;
}

int mhd2d_ens_groups(const mhd2d_ens* e) { return e ? (int)e->groups.size() : 0; }

int mhd2d_ens_set_state(mhd2d_ens* e, int inst, const double* uu)
{
    if (!e || inst < 0 || inst >= (int)e->where.size()) return fail(MHD2D_EINVAL, "bad argument");
    HIPCHK(hipStreamSynchronize(e->master));
    auto w = e->where[inst];
    return mhd2d_set_state_at(e->groups[w.first], w.second, uu);
}

int mhd2d_ens_get_state(mhd2d_ens* e, int inst, double* uu)
{
    if (!e || inst < 0 || inst >= (int)e->where.size()) return fail(MHD2D_EINVAL, "bad argument");
    HIPCHK(hipStreamSynchronize(e->master));
    auto w = e->where[inst];
    return mhd2d_get_state_at(e->groups[w.first], w.second, uu);
}

int mhd2d_ens_get_scalars(mhd2d_ens* e, int inst, double* dt, double* time, long long* steps)
{
    if (!e || inst < 0 || inst >= (int)e->where.size()) return fail(MHD2D_EINVAL, "bad argument");
    HIPCHK(hipStreamSynchronize(e->master));
    auto w = e->where[inst];
    return mhd2d_get_scalars_at(e->groups[w.first], w.second, dt, time, steps);
}

int mhd2d_ens_diagnostics(mhd2d_ens* e, int inst, double* energy, double* max_divb)
{
    if (!e || inst < 0 || inst >= (int)e->where.size()) return fail(MHD2D_EINVAL, "bad argument");
    HIPCHK(hipStreamSynchronize(e->master));
    auto w = e->where[inst];
    return mhd2d_diagnostics_at(e->groups[w.first], w.second, energy, max_divb);
}

int mhd2d_ens_sync(mhd2d_ens* e)
{
    if (!e) return fail(MHD2D_EINVAL, "null argument");
    HIPCHK(hipStreamSynchronize(e->master));
    for (auto* c : e->groups)
        if (int rc = mhd2d_sync(c)) return rc;
    return MHD2D_OK;
}

// every group tunes its own x-solve shapes (its grid, precision and batch);
// choice[4 g ..] receives group g's.  Where the one-launch path fits, the
// two launch modes are then timed as whole steps of the ensemble from the
// same saved state (restored afterwards); one launch per kernel class is
// kept unless the per-group launches are faster by more than 1 %.
int mhd2d_ens_autotune(mhd2d_ens* e, int reps, int* choice)
{
    if (!e || reps < 1) return fail(MHD2D_EINVAL, "bad argument");
    HIPCHK(join_into_master(e));
    HIPCHK(hipStreamSynchronize(e->master));
    for (size_t g = 0; g < e->groups.size(); g++)
        if (int rc = mhd2d_autotune(e->groups[g], reps, choice ? choice + 4 * g : nullptr))
            return rc;
    if (e->graph) {
        (void)hipGraphExecDestroy(e->graph);
        e->graph = nullptr;
    }
    if (!e->multi_ok) return MHD2D_OK;
    // snapshot of every group: state, R_prev, run scalars, pending filter
    struct Snap {
        char* p = nullptr;
        size_t ub = 0, rb = 0, sb = 0;
        bool pf = false;
    };
    std::vector<Snap> sn(e->groups.size());
    auto release = [&]() {
        for (auto& q : sn)
            if (q.p) (void)hipFree(q.p);
    };
    hipError_t r = hipSuccess;
    for (size_t g = 0; g < e->groups.size() && r == hipSuccess; g++) {
        mhd2d_ctx* c = e->groups[g];
        Snap& q = sn[g];
        q.ub = c->su() * c->elem * c->B;
        q.rb = c->sr() * c->delem * c->B;
        q.sb = sizeof(RunScalars) * c->B;
        q.pf = c->pending_filter;
        r = hipMalloc(&q.p, q.ub + q.rb + q.sb);
        if (r == hipSuccess) r = hipMemcpyAsync(q.p, c->u, q.ub, hipMemcpyDeviceToDevice, e->master);
        if (r == hipSuccess)
            r = hipMemcpyAsync(q.p + q.ub, c->rprev, q.rb, hipMemcpyDeviceToDevice, e->master);
        if (r == hipSuccess)
            r = hipMemcpyAsync(q.p + q.ub + q.rb, c->rs, q.sb, hipMemcpyDeviceToDevice, e->master);
    }
    auto restore = [&]() {
        hipError_t x = hipSuccess;
        for (size_t g = 0; g < e->groups.size() && x == hipSuccess; g++) {
            mhd2d_ctx* c = e->groups[g];
            const Snap& q = sn[g];
            x = hipMemcpyAsync(c->u, q.p, q.ub, hipMemcpyDeviceToDevice, e->master);
            if (x == hipSuccess)
                x = hipMemcpyAsync(c->rprev, q.p + q.ub, q.rb, hipMemcpyDeviceToDevice, e->master);
            if (x == hipSuccess)
                x = hipMemcpyAsync(c->rs, q.p + q.ub + q.rb, q.sb, hipMemcpyDeviceToDevice,
                                   e->master);
            c->pending_filter = q.pf;
        }
        if (x == hipSuccess) x = fork_from_master(e);
        if (x == hipSuccess) x = hipStreamSynchronize(e->master);
        return x;
    };
    float best[2] = {1e30f, 1e30f};
    const int saved_mode = e->mode;
    for (int it = 0; it < 4 && r == hipSuccess; it++) {
        const int mode = it % 2 ? MHD2D_ENS_PER_GROUP : MHD2D_ENS_ONE_LAUNCH;
        e->mode = mode;
        r = restore();
        double ms = 0;
        if (r == hipSuccess) {
            // one untimed step (the first has no pending filter), then reps
            if (mhd2d_ens_run(e, 1, 0) != MHD2D_OK) r = hipErrorUnknown;
            if (r == hipSuccess && mhd2d_ens_time_run(e, reps, 0, &ms) != MHD2D_OK)
                r = hipErrorUnknown;
        }
        if (r == hipSuccess) best[mode] = std::min(best[mode], (float)ms);
    }
    e->mode = saved_mode;
    if (r == hipSuccess) r = restore();
    release();
    if (r != hipSuccess) return fail(MHD2D_EHIP, "ensemble autotune", r);
    e->mode = best[MHD2D_ENS_PER_GROUP] < 0.99f * best[MHD2D_ENS_ONE_LAUNCH] ? MHD2D_ENS_PER_GROUP
                                                                             : MHD2D_ENS_ONE_LAUNCH;
    return MHD2D_OK;
}

// Steps are queued on the master stream's timeline: the groups fork from it
// and join back, so events on the master stream bracket all of them.
int mhd2d_ens_run(mhd2d_ens* e, int nsteps, int use_graph)
{
    if (!e || nsteps < 0) return fail(MHD2D_EINVAL, "bad argument");
    for (auto* c : e->groups)
        if (int rc = check_ready(c)) return rc;
    if (nsteps == 0) return MHD2D_OK;
    HIPCHK(join_into_master(e));  // earlier work queued on the group streams
    int k = 0;
    bool fresh = false;
    for (auto* c : e->groups) fresh = fresh || !c->pending_filter;
    if (e->mode == MHD2D_ENS_ONE_LAUNCH) HIPCHK(ens_align_filters(e));
    if (!use_graph || fresh) {
        // eager steps (and the first step, which has no pending y filter)
        const int ne = use_graph ? 1 : nsteps;
        HIPCHK(fork_from_master(e));
        for (; k < ne; k++) HIPCHK(ens_step_all(e));
        HIPCHK(join_into_master(e));
    }
    if (k < nsteps) {
        if (!e->graph) {
            hipGraph_t g;
            HIPCHK(hipStreamBeginCapture(e->master, hipStreamCaptureModeThreadLocal));
            hipError_t r = fork_from_master(e);
            if (r == hipSuccess) r = ens_step_all(e);
            if (r == hipSuccess) r = join_into_master(e);
            hipError_t r2 = hipStreamEndCapture(e->master, &g);
            if (r != hipSuccess) return fail(MHD2D_EHIP, "ensemble capture", r);
            HIPCHK(r2);
            HIPCHK(hipGraphInstantiate(&e->graph, g, nullptr, nullptr, 0));
            HIPCHK(hipGraphDestroy(g));
        }
        for (; k < nsteps; k++) HIPCHK(hipGraphLaunch(e->graph, e->master));
    }
    // later work on the group streams follows the steps
    HIPCHK(fork_from_master(e));
    return MHD2D_OK;
}

int mhd2d_ens_time_run(mhd2d_ens* e, int nsteps, int use_graph, double* ms)
{
    if (!e || !ms) return fail(MHD2D_EINVAL, "null argument");
    hipEvent_t a, b;
    HIPCHK(hipEventCreate(&a));
    HIPCHK(hipEventCreate(&b));
    HIPCHK(join_into_master(e));
    HIPCHK(hipStreamSynchronize(e->master));
    HIPCHK(hipEventRecord(a, e->master));
    int rc = mhd2d_ens_run(e, nsteps, use_graph);
    if (rc) return rc;
    HIPCHK(hipEventRecord(b, e->master));
    HIPCHK(hipEventSynchronize(b));
    float f = 0;
    HIPCHK(hipEventElapsedTime(&f, a, b));
    *ms = f;
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    return MHD2D_OK;
}

}  // extern "C"

// =====================================================================
// mhd2d_run_noalloc: the context API's set_state + run + get_state for one
// FP64 run held in caller device memory, allocating nothing.  Every working
// array lives in `scratch`, one fixed array per GPU for the life of the
// process.
// =====================================================================
namespace {

constexpr size_t NA_MISC = 64u << 10;     // run scalars, GLM Ch / Cp per x block
constexpr size_t NA_TABLES = 1u << 20;    // operator tables of the grid
// per cell at most 50 words: state + E + scratch (11, 12 with GLM), dudx and
// dudy (11 each), R_prev (8, 9), Laplacians (6), div B and its derivatives (3)
constexpr size_t NA_WORDS = 50;
static_assert(NA_MISC + NA_TABLES + NA_WORDS * 8 * (size_t)MHD2D_NOALLOC_MAX_CELLS <=
                  (size_t)MHD2D_NOALLOC_BYTES,
              "scratch too small for the largest grid");

// Host code never uses this name as an address: each device's address comes
// from hipGetSymbolAddress (noalloc_base).
static __device__ __align__(256) double scratch[MHD2D_NOALLOC_BYTES / sizeof(double)];

constexpr int NA_MAXDEV = 64;
std::mutex g_na_mu;
void* g_na_base[NA_MAXDEV] = {};

// the address of the current device's copy of scratch, looked up once per
// device with that device current
hipError_t noalloc_base(void** out)
{
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= NA_MAXDEV) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(g_na_mu);
    if (!g_na_base[dev]) {
        void* q = nullptr;
        e = hipGetSymbolAddress(&q, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        if (!q) return hipErrorInvalidSymbol;
        g_na_base[dev] = q;
    }
    *out = g_na_base[dev];
    return hipSuccess;
}

}  // namespace

extern "C" int mhd2d_run_noalloc(const mhd2d_params* p, double* d_uu, int nsteps, void* stream,
                                 double* dt_out, double* time_out)
{
    if (!p || !d_uu || nsteps < 0) return fail(MHD2D_EINVAL, "bad argument");
    if (int rc = check_params(p)) return rc;
    if (p->precision != MHD2D_FP64 || p->batch > 1 || p->xgemm)
        return fail(MHD2D_EINVAL,
                    "mhd2d_run_noalloc runs one FP64 run (precision MHD2D_FP64, batch <= 1) "
                    "with the tridiagonal x-derivative (xgemm 0)");
    const size_t n = (size_t)p->nx * p->ny;
    if (n > (size_t)MHD2D_NOALLOC_MAX_CELLS)
        return fail(MHD2D_EINVAL,
                    "mhd2d_run_noalloc: nx * ny exceeds MHD2D_NOALLOC_MAX_CELLS (65536 cells; "
                    "its working arrays are a fixed 28 MiB per GPU)");
    void* base = nullptr;
    HIPCHK(noalloc_base(&base));
    hipStream_t st = (hipStream_t)stream;

    mhd2d_ctx c;  // views into scratch only: nothing to release
    c.p = *p;
    c.elem = c.delem = 8;
    c.B = 1;
    c.n = n;
    c.stream = st;
    char* const q = (char*)base;
    c.rs = (RunScalars*)q;
    if (c.glm()) c.glm_cp = (double*)(q + 4096);  // 2 npe <= nx / 2 doubles: 16 KiB at most
    HIPCHK(c.tb.build(p->nx, p->ny, p->Lx, p->Ly, 8, SchemeAlphas(), p->xblocks, p->open_x,
                      q + NA_MISC, NA_TABLES));
    {
        // constDiffDivB = 2 max(dx, dy)^2 (initialize.c:17, 154)
        const double h = std::max(c.tb.dx, c.tb.dy);
        c.cdiff = 2.0 * h * h;
    }
    double* f = (double*)(q + NA_MISC + NA_TABLES);
    c.u = f;
    f += c.su();
    c.dudx = f;
    f += c.sd();
    c.dudy = f;
    f += c.sd();
    c.rprev = f;
    f += c.sr();
    if (c.lap_u() || c.lap_b()) {
        c.lap = f;
        f += c.sl();
    }
    if (p->divbclean == 1) {
        c.divb = f;
        f += c.sb();
    }
    if ((size_t)((char*)f - q) > (size_t)MHD2D_NOALLOC_BYTES)
        return fail(MHD2D_EINVAL, "mhd2d_run_noalloc: working arrays exceed the scratch size");

    // the state: host layout fields 0..7 (psi 8 with GLM) -> slots 0..7 (PSI_SLOT)
    double* u = (double*)c.u;
    HIPCHK(hipMemcpyAsync(u, d_uu, 8 * n * sizeof(double), hipMemcpyDeviceToDevice, st));
    if (c.glm())
        HIPCHK(hipMemcpyAsync(u + PSI_SLOT * n, d_uu + 8 * n, n * sizeof(double),
                              hipMemcpyDeviceToDevice, st));
    // a fresh run: time 0, and the first dt is calcDt(0) of the state (mhd.c:58)
    RunScalars z;
    std::memset(&z, 0, sizeof(z));
    z.dtmin_bits = 0x7ff0000000000000ull;
    z.di = p->di;
    HIPCHK(hipMemcpyAsync(c.rs, &z, sizeof(z), hipMemcpyHostToDevice, st));
    HIPCHK(hipStreamSynchronize(st));
    c.pending_filter = false;
    for (int k = 0; k < nsteps; k++) HIPCHK((step<double, double>(&c, nullptr, nullptr)));
    // the last stage's y filter (skipped when the run stopped: its stopping
    // step's y-sweep applied it already)
    RunScalars r;
    HIPCHK(hipMemcpyAsync(&r, c.rs, sizeof(r), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    if (r.stop) c.pending_filter = false;
    HIPCHK((flush<double, double>(&c)));
    HIPCHK(hipMemcpyAsync(d_uu, u, 8 * n * sizeof(double), hipMemcpyDeviceToDevice, st));
    if (c.glm())
        HIPCHK(hipMemcpyAsync(d_uu + 8 * n, u + PSI_SLOT * n, n * sizeof(double),
                              hipMemcpyDeviceToDevice, st));
    HIPCHK(hipMemcpyAsync(&r, c.rs, sizeof(r), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    if (dt_out) *dt_out = r.dt;
    if (time_out) *time_out = r.step > 0 ? r.time + r.dt : 0.0;
    if (r.nonfinite) return fail(MHD2D_ENONFINITE, "a CFL limit was not finite (state blew up)");
    return MHD2D_OK;
}
