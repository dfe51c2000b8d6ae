// Time-driven run loop of the reference's main program (mhd.c:58-230) on
// the GPU: output cadence dtOut, log cadence dtOut/10, the dt < 1e-8 stop,
// and the final output at Tmax.  Two engines:
//   MHD2D_DRIVE_FUSED  - the context API (fused kernels, dt on the device);
//                        Hall MHD with viscosity, resistivity and diffusive
//                        or GLM div B cleaning.
//   MHD2D_DRIVE_DROPIN - the reference's own call sequence through the
//                        drop-in entry points on the host globals; any
//                        supported setting (viscosity, resistivity,
//                        diffusive div B cleaning), with the reference's
//                        stale checkDivB (mhd.c:342-367).
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "../../include/mhd2d.h"
#include "mhd2d_refglobals.h"

namespace {

double now_s()
{
    using namespace std::chrono;
    return duration<double>(steady_clock::now().time_since_epoch()).count();
}

struct Out {
    const mhd2d_input* in;
    const mhd2d_drive_opts* o;
    int iout = 0;
    int nvar = 8;   // fields per output file (NVAR: 9 with hyperbolic cleaning)
};

int emit(Out& w, double t, const double* uu, const double* divb, double maxdivb, double dt)
{
    if (w.o->verbose)
        std::printf("  Output at time = %10.4f, max(div B) = %12.4e, dt = %12.4e\n", t, maxdivb,
                    dt);
    int rc = mhd2d_write_array(w.o->outdir, w.iout, t, uu, w.nvar, w.in->nx, w.in->ny);
    if (!rc && w.in->outdivb == 1 && divb)
        rc = mhd2d_write_divb(w.o->outdir, w.iout, t, divb, w.in->nx, w.in->ny);
    w.iout++;
    return rc;
}

int drive_fused(const mhd2d_input* in, const double* uu0, const mhd2d_drive_opts* o,
                mhd2d_drive_report* rep)
{
    mhd2d_params p{};
    p.nx = in->nx;
    p.ny = in->ny;
    p.Lx = in->Lx;
    p.Ly = in->Ly;
    p.gamma = in->gamma;
    p.cfl = in->cfl;
    p.hall = in->hall;
    p.di = in->di;
    p.resistivity = in->resistivity;
    p.viscosity = in->viscosity;
    p.divbclean = in->divbclean;
    p.precision = o->precision;
    p.xblocks = o->xblocks;
    p.open_x = in->boundaryTypeX == 1;
    mhd2d_ctx* c = nullptr;
    int rc = mhd2d_create(&p, &c);
    if (rc) return rc;
    const size_t n = (size_t)in->nx * in->ny;
    const int nvar = in->divbclean == 2 ? 9 : 8;      // psi starts at 0
    std::vector<double> uu(nvar * n, 0.0), divb(n);
    std::memcpy(uu.data(), uu0, 8 * n * sizeof(double));
    rc = mhd2d_set_state(c, uu.data());
    Out w{in, o, 0, nvar};
    const double t0 = now_s();
    double t = 0.0, dt = 0.0, tOut = 0.0, tLog = 0.0, maxdivb = 0.0, energy = 0.0;
    const double dtLog = in->dtOut / 10.0;  // mhd.c:66
    long long steps = 0;
    int stopped = 0;
    auto snapshot = [&]() -> int {
        int r = mhd2d_get_state(c, uu.data());
        if (!r) r = mhd2d_get_divb(c, divb.data());
        if (!r) r = mhd2d_diagnostics(c, &energy, &maxdivb);
        return r;
    };
    // output and log at t = 0 (mhd.c:74-95)
    if (!rc) rc = snapshot();
    if (!rc) rc = emit(w, t, uu.data(), divb.data(), maxdivb, dt);
    tOut += in->dtOut;
    if (!rc) rc = mhd2d_write_log(o->outdir, t, dt, now_s() - t0, steps);
    tLog += dtLog;
    while (!rc && t <= in->Tmax) {                     // mhd.c:99
        if (t >= tOut) {                               // mhd.c:102-116
            rc = snapshot();
            if (!rc) rc = emit(w, t, uu.data(), divb.data(), maxdivb, dt);
            tOut += in->dtOut;
        }
        if (!rc && t >= tLog) {                        // mhd.c:119-128
            rc = mhd2d_write_log(o->outdir, t, dt, now_s() - t0, steps);
            tLog += dtLog;
        }
        if (rc) break;
        if (o->max_steps > 0 && steps >= o->max_steps) break;
        rc = mhd2d_run(c, 1);                          // mhd.c:154-206
        if (rc) break;
        double dt_stop = 0.0;
        const int st = mhd2d_stopped(c, &dt_stop);
        if (st > 1) {
            rc = st;
            break;
        }
        if (st == 1) {                                 // mhd.c:131-151
            stopped = 1;
            dt = dt_stop;
            if (o->verbose)
                std::printf("  STOP: dt < 1e-8\n  time = %10.4f, dt = %12.4e\n", t, dt);
            rc = snapshot();
            if (!rc) rc = emit(w, t, uu.data(), divb.data(), maxdivb, dt);
            break;
        }
        rc = mhd2d_get_scalars(c, &dt, &t, &steps);
    }
    if (!rc && !stopped) {                             // mhd.c:211-230
        rc = snapshot();
        if (!rc) rc = emit(w, t, uu.data(), divb.data(), maxdivb, dt);
        if (!rc) rc = mhd2d_write_log(o->outdir, t, dt, now_s() - t0, steps);
    }
    if (!rc) rc = mhd2d_sync(c);
    if (rep) {
        rep->time = t;
        rep->dt = dt;
        rep->steps = steps;
        rep->outputs = w.iout;
        rep->stopped = stopped;
        rep->seconds = now_s() - t0;
        rep->energy = energy;
        rep->max_divb = maxdivb;
    }
    mhd2d_destroy(c);
    return rc;
}

// mhd.c:342-367 on the host copies (derivatives of the last stage's state)
double check_divb()
{
    double m = 0.0;
    const size_t n = iSize * jSize;
    for (size_t p = 0; p < n; p++) {
        divB[p] = dudx[4 * n + p] + dudy[5 * n + p];
        const double a = std::fabs(divB[p]);
        if (a > m) m = a;
    }
    return m;
}

int drive_dropin(const mhd2d_input* in, const double* uu0, const mhd2d_drive_opts* o,
                 mhd2d_drive_report* rep)
{
    int rc = mhd2d_host_initialize(in->nx, in->ny, in->Lx, in->Ly, in->gamma, in->cfl, in->hall,
                                   in->di, in->resistivity, in->viscosity, in->divbclean);
    if (rc) return rc;
    mhd2d_dropin_set_sync(MHD2D_SYNC_STRICT);
    const size_t n = (size_t)in->nx * in->ny;
    std::memcpy(uu, uu0, 8 * n * sizeof(double));   // usrInitialize(); psi starts at 0
    Out w{in, o, 0, NVAR};
    const double t0 = now_s();
    double dt = calcDt(0.0);                          // mhd.c:58
    rkInit(dt);                                       // mhd.c:63
    const double dtLog = in->dtOut / 10.0;
    double t = 0.0, tOut = 0.0, tLog = 0.0;
    long long steps = 0;
    int stopped = 0;
    double maxdivb = check_divb();                    // mhd.c:75
    rc = emit(w, t, uu, divB, maxdivb, dt);
    tOut += in->dtOut;
    if (!rc) rc = mhd2d_write_log(o->outdir, t, dt, now_s() - t0, steps);
    tLog += dtLog;
    while (!rc && t <= in->Tmax) {
        if (t >= tOut) {
            maxdivb = check_divb();
            rc = emit(w, t, uu, divB, maxdivb, dt);
            tOut += in->dtOut;
        }
        if (!rc && t >= tLog) {
            rc = mhd2d_write_log(o->outdir, t, dt, now_s() - t0, steps);
            tLog += dtLog;
        }
        if (rc) break;
        if (dt < 1e-8) {
            stopped = 1;
            if (o->verbose) std::printf("  STOP: dt < 1e-8\n  time = %10.4f, dt = %12.4e\n", t, dt);
            rc = emit(w, t, uu, divB, maxdivb, dt);
            break;
        }
        if (o->max_steps > 0 && steps >= o->max_steps) break;
        for (int iRK = 0; iRK < 3; iRK++) {           // mhd.c:154-199
            sendrecvHaloPointsX();
            derivX(NVAR, uu, uuHaloX0, uuHaloX1, dudx, 0, 0);
            derivY(NVAR, uu, dudy, 0);
            if (viscosity > 0) {
                derivXX(3, uu, uuHaloX0, uuHaloX1, d2udx2, 1, 1);
                derivYY(3, uu, d2udy2, 1);
            }
            if (Hall == 1 || resistivity > 0) {
                derivXX(3, uu, uuHaloX0, uuHaloX1, d2Bdx2, 4, 4);
                derivYY(3, uu, d2Bdy2, 4);
            }
            if (divBcleaning == 1) {
                calcDivB();
                sendrecvHaloPointsdivBX();
                derivX(1, divB, divBHaloX0, divBHaloX1, ddivBdx, 0, 0);
                derivY(1, divB, ddivBdy, 0);
            }
            calcEfield();
            sendrecvHaloPointsEfieldX();
            derivX(3, Efield, EfieldHaloX0, EfieldHaloX1, dEdx, 0, 0);
            derivY(3, Efield, dEdy, 0);
            calcRHS();
            rkExecute(iRK);
            sendrecvHaloPointsX();
            filtering();
        }
        t += dt;                                      // mhd.c:201-206
        steps++;
        dt = calcDt(dt);
        rkInit(dt);
    }
    if (!rc && !stopped) {
        maxdivb = check_divb();
        rc = emit(w, t, uu, divB, maxdivb, dt);
        if (!rc) rc = mhd2d_write_log(o->outdir, t, dt, now_s() - t0, steps);
    }
    if (rep) {
        rep->time = t;
        rep->dt = dt;
        rep->steps = steps;
        rep->outputs = w.iout;
        rep->stopped = stopped;
        rep->seconds = now_s() - t0;
        rep->energy = 0.0;
        rep->max_divb = maxdivb;
    }
    mhd2d_host_finalize();
    return rc;
}

}  // namespace

extern "C" int mhd2d_drive(const mhd2d_input* in, const double* uu0, const mhd2d_drive_opts* o,
                           mhd2d_drive_report* rep)
{
    if (!in || !uu0 || !o) return MHD2D_EINVAL;
    if (in->boundaryTypeX != 0 && in->boundaryTypeX != 1) {
        std::fprintf(stderr, "mhd2d_drive: boundaryTypeX must be 0 (periodic) or 1 (open)\n");
        return MHD2D_EINVAL;
    }
    if (in->boundaryTypeX == 1 && (o->xblocks < 2 || o->mode == MHD2D_DRIVE_DROPIN)) {
        // the reference's open boundary needs npe >= 2 (parallel.c:25-61)
        std::fprintf(stderr, "mhd2d_drive: the open x boundary needs the fused engine and "
                             "npe >= 2 x blocks\n");
        return MHD2D_EINVAL;
    }
    int mode = o->mode;
    if (mode == MHD2D_DRIVE_AUTO) mode = MHD2D_DRIVE_FUSED;
    if (mode == MHD2D_DRIVE_DROPIN && o->xblocks > 1) {
        std::fprintf(stderr, "mhd2d_drive: x blocks (npe > 1) need the fused engine\n");
        return MHD2D_EINVAL;
    }
    if (mode == MHD2D_DRIVE_FUSED) return drive_fused(in, uu0, o, rep);
    return drive_dropin(in, uu0, o, rep);
}
