// Drop-in entry points with the reference's names and signatures
// (calcDerivs.h:6-15, filtering.h:6, calcRHS.h:4-6, timeAdvance.h:6-8,
// parallel.h:4-6, divBcleaning.h:10-11).  They act on the reference's host
// globals through device mirrors: each host array gets one device copy,
// uploaded when the device copy is stale and downloaded at the sync points
// (calcDt in strict mode, or mhd2d_dropin_sync_host).  Like the reference,
// these functions return nothing and are not re-entrant; a HIP failure or an
// unsupported configuration prints file:line and aborts.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <vector>

#include "../../include/mhd2d.h"
#include "mhd2d_internal.h"
#include "mhd2d_refglobals.h"
#include "mhd2d_tables.h"

using namespace mhd2d;

extern "C" {
real Cr_divBCleaning = 0.18;  // divBcleaning.c:4
}

namespace {

struct Mirror {
    const real* host;
    double* dev;
    size_t count;
    bool host_newer;
    bool dev_newer;
};

std::vector<Mirror> g_mirrors;
int g_sync = MHD2D_SYNC_STRICT;
hipStream_t g_stream = nullptr;
std::unique_ptr<GridTables> g_tb;
RunScalars* g_rs = nullptr;
double g_cc[3] = {0, 0, 0}, g_dd[3] = {0, 0, 0};

[[noreturn]] void die(const char* file, int line, const char* what)
{
    std::fprintf(stderr, "mhd2d drop-in: %s:%d: %s\n", file, line, what);
    std::fflush(stderr);
    std::abort();
}

#define DIE(msg) die(__FILE__, __LINE__, msg)
#define CHK(expr)                                                                   \
    do {                                                                            \
        hipError_t e_ = (expr);                                                     \
        if (e_ != hipSuccess) die(__FILE__, __LINE__, hipGetErrorString(e_));      \
    } while (0)

size_t plane() { return iSize * jSize; }

// number of fields of a known global array (initialize.c:216-294)
size_t fields_of(const real* p, size_t fallback)
{
    if (!p) DIE("null array");
    if (p == uu || p == dudx || p == dudy || p == dudt || p == dudtRK) return (size_t)NVAR;
    if (p == Efield || p == dEdx || p == dEdy || p == d2udx2 || p == d2udy2 || p == d2Bdx2 ||
        p == d2Bdy2)
        return 3;
    if (p == divB || p == ddivBdx || p == ddivBdy) return 1;
    return fallback;
}

void setup()
{
    if (boundaryTypeX != 0) DIE("only the periodic x boundary (boundaryTypeX == 0) is supported");
    if (npe != 1) DIE("the whole grid runs on one GPU: npe must be 1");
    if (iSize != nx || jSize != ny) DIE("iSize/jSize must equal nx/ny with one rank");
    if (!g_stream) CHK(hipStreamCreateWithFlags(&g_stream, hipStreamNonBlocking));
    if (!g_rs) CHK(hipMalloc(&g_rs, sizeof(RunScalars)));
    if (!g_tb || g_tb->nx != (int)nx || g_tb->ny != (int)ny || g_tb->Lx != Lx ||
        g_tb->Ly != Ly) {
        auto pow2 = [](size_t v) { return v > 0 && (v & (v - 1)) == 0; };
        if (!pow2(nx) || !pow2(ny) || nx < 16 || nx > 4096 || ny < 16 || ny > 4096)
            DIE("nx and ny must be powers of two in [16, 4096]");
        g_tb.reset(new GridTables());
        CHK(g_tb->build((int)nx, (int)ny, Lx, Ly, 8));
    }
}

Mirror& mirror(const real* host, size_t fields)
{
    const size_t cnt = fields * plane();
    for (auto& m : g_mirrors)
        if (m.host == host) {
            if (m.count < cnt) DIE("array used with more fields than it was registered with");
            return m;
        }
    Mirror m{host, nullptr, cnt, true, false};
    CHK(hipMalloc(&m.dev, cnt * sizeof(double)));
    g_mirrors.push_back(m);
    return g_mirrors.back();
}

// device copy of a host array, current for reading
double* dev_in(const real* host, size_t fields)
{
    Mirror& m = mirror(host, fields);
    if (m.host_newer) {
        CHK(hipMemcpyAsync(m.dev, m.host, m.count * sizeof(double), hipMemcpyHostToDevice,
                           g_stream));
        m.host_newer = false;
    }
    return m.dev;
}

// device copy that the caller is about to (partly) overwrite
double* dev_out(const real* host, size_t fields)
{
    double* d = dev_in(host, fields);
    for (auto& m : g_mirrors)
        if (m.host == host) m.dev_newer = true;
    return d;
}

void sync_all()
{
    for (auto& m : g_mirrors)
        if (m.dev_newer) {
            CHK(hipMemcpyAsync(const_cast<real*>(m.host), m.dev, m.count * sizeof(double),
                               hipMemcpyDeviceToHost, g_stream));
            m.dev_newer = false;
        }
    CHK(hipStreamSynchronize(g_stream));
}

void x_op(int kind, int nv, real* arrMain, real* arrDeriv, int ivm)
{
    setup();
    if (nv < 1) return;
    const double* din = dev_in(arrMain, fields_of(arrMain, (size_t)(ivm + nv)));
    double* dout = dev_out(arrDeriv, fields_of(arrDeriv, (size_t)nv));
    const size_t n = plane();
    for (int v0 = 0; v0 < nv; v0 += 16) {
        const int cnt = nv - v0 < 16 ? nv - v0 : 16;
        XSolveArgs<double> a;
        for (int v = 0; v < cnt; v++) {
            a.in[v] = din + (size_t)(ivm + v0 + v) * n;
            a.out[v] = dout + (size_t)(v0 + v) * n;
        }
        a.st = g_tb->st[kind];
        a.tb = g_tb->xt[kind];
        a.kind = kind;
        a.nx = (int)nx;
        a.ny = (int)ny;
        CHK(launch_xsolve<double>(a, cnt, g_stream));
    }
}

void y_op(int mode, int nv, const real* arrMain, real* arrDeriv, int ivm)
{
    setup();
    if (nv < 1) return;
    const double* din = dev_in(arrMain, fields_of(arrMain, (size_t)(ivm + nv)));
    double* dout = dev_out(arrDeriv, fields_of(arrDeriv, (size_t)nv));
    const size_t n = plane();
    for (int v0 = 0; v0 < nv; v0 += 16) {
        const int cnt = nv - v0 < 16 ? nv - v0 : 16;
        YApplyArgs<double> a;
        for (int v = 0; v < cnt; v++) {
            a.in[v] = din + (size_t)(ivm + v0 + v) * n;
            a.out[v] = dout + (size_t)(v0 + v) * n;
        }
        a.tb = g_tb->yt;
        a.nv = cnt;
        a.mode = mode;
        a.nx = (int)nx;
        a.ny = (int)ny;
        CHK(launch_yapply<double>(a, g_stream));
    }
}

}  // namespace

extern "C" {

void derivX(int nv, real* arrMain, real* arrHalo0, real* arrHalo1, real* arrDeriv,
            int iVarStartMain, int iVarStartHalo)
{
    (void)arrHalo0;
    (void)arrHalo1;
    (void)iVarStartHalo;
    x_op(XD1, nv, arrMain, arrDeriv, iVarStartMain);
}

void derivXX(int nv, real* arrMain, real* arrHalo0, real* arrHalo1, real* arrDeriv,
             int iVarStartMain, int iVarStartHalo)
{
    (void)arrHalo0;
    (void)arrHalo1;
    (void)iVarStartHalo;
    x_op(XD2, nv, arrMain, arrDeriv, iVarStartMain);
}

void derivY(int nv, real* arrMain, real* arrDeriv, int iVarStartMain)
{
    y_op(Y_D1, nv, arrMain, arrDeriv, iVarStartMain);
}

void derivYY(int nv, real* arrMain, real* arrDeriv, int iVarStartMain)
{
    y_op(Y_D2, nv, arrMain, arrDeriv, iVarStartMain);
}

// filtering.c:98-102: x filter then y filter, in place on all NVAR fields
void filtering(void)
{
    x_op(XFILT, NVAR, uu, uu, 0);
    y_op(Y_F, NVAR, uu, uu, 0);
}

void calcEfield(void)
{
    setup();
    const double* du = dev_in(uu, (size_t)NVAR);
    double* dE = dev_out(Efield, 3);
    CHK(launch_efield<double>(du, dE, plane(), g_stream));
}

void calcDivB(void)
{
    setup();
    const double* dx_ = dev_in(dudx, (size_t)NVAR);
    const double* dy_ = dev_in(dudy, (size_t)NVAR);
    double* d = dev_out(divB, 1);
    CHK(launch_divb<double>(dx_, dy_, d, plane(), g_stream));
}

// calcRHS.c:46-75 for the periodic branch, with addSourceTerm (:498-544) and,
// for divBcleaning == 2, the hyperbolic cleaning terms with the Ch / Cp the
// last calcDt left in the globals (calcRHS.c:144-158)
void calcRHS(void)
{
    setup();
    if (divBcleaning == 2 && NVAR != 9) DIE("divBcleaning == 2 needs NVAR == 9 (psi)");
    RhsArgs<double> a;
    a.u = dev_in(uu, (size_t)NVAR);
    a.dudx = dev_in(dudx, (size_t)NVAR);
    a.dudy = dev_in(dudy, (size_t)NVAR);
    a.dEdx = dev_in(dEdx, 3);
    a.dEdy = dev_in(dEdy, 3);
    a.d2udx2 = viscosity > 0 ? dev_in(d2udx2, 3) : nullptr;
    a.d2udy2 = viscosity > 0 ? dev_in(d2udy2, 3) : nullptr;
    a.d2Bdx2 = resistivity > 0 ? dev_in(d2Bdx2, 3) : nullptr;
    a.d2Bdy2 = resistivity > 0 ? dev_in(d2Bdy2, 3) : nullptr;
    a.ddivBdx = divBcleaning == 1 ? dev_in(ddivBdx, 1) : nullptr;
    a.ddivBdy = divBcleaning == 1 ? dev_in(ddivBdy, 1) : nullptr;
    a.dudt = dev_out(dudt, (size_t)NVAR);
    a.gamma = adiabaticIdx;
    a.resistivity = resistivity;
    a.viscosity = viscosity;
    a.constDiffDivB = constDiffDivB;
    a.ch = Ch_divBcleaning;
    a.cp = Cp_divBcleaning;
    a.divbclean = divBcleaning;
    a.n = plane();
    CHK(launch_rhs<double>(a, g_stream));
}

// timeAdvance.c:20-27
void rkInit(real dt)
{
    setup();
    g_cc[0] = 8.0 / 15.0 * dt;
    g_dd[0] = 0.0;
    g_cc[1] = 5.0 / 12.0 * dt;
    g_dd[1] = -17. / 60. * dt;
    g_cc[2] = 0.75 * dt;
    g_dd[2] = -5. / 12. * dt;
    double* d = dev_out(dudtRK, (size_t)NVAR);
    CHK(launch_fill<double>(d, 0.0, arrSize, g_stream));
}

// timeAdvance.c:29-36
void rkExecute(int iRK)
{
    setup();
    if (iRK < 0 || iRK > 2) DIE("iRK must be 0, 1 or 2");
    double* du = dev_out(uu, (size_t)NVAR);
    const double* dt_ = dev_in(dudt, (size_t)NVAR);
    double* drk = dev_out(dudtRK, (size_t)NVAR);
    CHK(launch_rk<double>(du, dt_, drk, g_cc[iRK], g_dd[iRK], arrSize, g_stream));
}

// timeAdvance.c:38-153; with divBcleaning == 2 it also sets Ch / Cp
// (timeAdvance.c:107-112) from the last cell of the loop
real calcDt(real dt0)
{
    setup();
    const double* du = dev_in(uu, (size_t)NVAR);
    DtParams dp;
    dp.dx = dx;
    dp.dy = dy;
    dp.gamma = adiabaticIdx;
    dp.cfl = cfl;
    dp.di = di;
    dp.resistivity = resistivity;
    dp.viscosity = viscosity;
    dp.hall = Hall;
    CHK(launch_dt_reset(g_rs, g_stream));
    CHK(launch_dt_limits<double>(du, dp, g_rs, (int)nx, (int)ny, g_stream));
    if (divBcleaning == 2) CHK(launch_glm_ch<double>(du, dp, g_rs, plane(), g_stream));
    RunScalars r;
    CHK(hipMemcpyAsync(&r, g_rs, sizeof(r), hipMemcpyDeviceToHost, g_stream));
    CHK(hipStreamSynchronize(g_stream));
    double dtmin;
    std::memcpy(&dtmin, &r.dtmin_bits, sizeof(double));
    dtmin *= cfl;
    const double dt = (dt0 < 0.98 * dtmin || dt0 > 1.02 * dtmin) ? dtmin : dt0;
    if (divBcleaning == 2) {
        Ch_divBcleaning = r.glm_ch;
        calc_Cp_fix_Cr();
    }
    if (g_sync == MHD2D_SYNC_STRICT) sync_all();
    return dt;
}

void sendrecvHaloPointsX(void) {}
void sendrecvHaloPointsdivBX(void) {}
void sendrecvHaloPointsEfieldX(void) {}

// divBcleaning.c:7-10
void calc_Cp_fix_Cr(void) { Cp_divBcleaning = std::sqrt(Cr_divBCleaning * Ch_divBcleaning); }

void mhd2d_dropin_set_sync(int mode) { g_sync = mode; }

int mhd2d_dropin_sync_host(void)
{
    if (!g_stream) return MHD2D_OK;
    sync_all();
    return MHD2D_OK;
}

void mhd2d_dropin_invalidate_device(void)
{
    for (auto& m : g_mirrors) {
        m.host_newer = true;
        m.dev_newer = false;
    }
}

void mhd2d_dropin_release(void)
{
    if (g_stream) (void)hipStreamSynchronize(g_stream);
    for (auto& m : g_mirrors) (void)hipFree(m.dev);
    g_mirrors.clear();
    g_tb.reset();
    if (g_rs) (void)hipFree(g_rs);
    g_rs = nullptr;
    if (g_stream) (void)hipStreamDestroy(g_stream);
    g_stream = nullptr;
}

}  // extern "C"
