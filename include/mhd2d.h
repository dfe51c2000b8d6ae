/*
 * mhd2d.h -- C ABI of the MI355X (gfx950) build of the right-hand-side and
 * RK3 time-advance path of chenshihelio/MHD2DCFDFFTW_Open.
 *
 * Two layers:
 *
 *  1. Drop-in entry points with the reference's own names and signatures.
 *     They operate on the reference's host globals (uu, dudx, dudy, Efield,
 *     nx, ny, ... as defined by its initialize.c) exactly like the object
 *     files they replace, so its mhd.c stage loop can drive them unchanged.
 *     Arrays stay resident on the GPU between calls; host copies are
 *     refreshed at calcDt (the one per-step call mhd.c makes), see
 *     mhd2d_dropin_set_sync().
 *
 *  2. A context API for running whole RK3 steps on the device with fused
 *     kernels (the timed path).  Plain pointers, sizes and int status codes.
 *
 * Layout everywhere: field-major [v][i][j], i along x (stride ny), j along y
 * (contiguous), fields (rho, ux, uy, uz, Bx, By, Bz, p) (macros.h:4-11).
 * Grid: periodic in x and y, nx and ny powers of two in [16, 4096].
 */
#ifndef MHD2D_H
#define MHD2D_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef double real; /* parameter.h:21 */

/* ===================================================================== */
/* 1. drop-in entry points                                               */
/* ===================================================================== */

/* replaces calcDerivs.c:9-39 (declared calcDerivs.h:6-8): compact 8th-order
 * d/dx of fields iVarStartMain.. of arrMain into fields 0..nv-1 of arrDeriv.
 * The halo arrays are accepted and not read: the periodic x-halo is an index
 * wrap on the device (parallel.c:13-24 with one rank). */
void derivX(int nv, real *arrMain, real *arrHalo0, real *arrHalo1,
            real *arrDeriv, int iVarStartMain, int iVarStartHalo);
/* replaces calcDerivs.c:41-72 (calcDerivs.h:10-12): compact d2/dx2 */
void derivXX(int nv, real *arrMain, real *arrHalo0, real *arrHalo1,
             real *arrDeriv, int iVarStartMain, int iVarStartHalo);
/* replaces calcDerivs.c:75-109 (calcDerivs.h:14): spectral d/dy */
void derivY(int nv, real *arrMain, real *arrDeriv, int iVarStartMain);
/* replaces calcDerivs.c:114-150 (calcDerivs.h:15): spectral d2/dy2 */
void derivYY(int nv, real *arrMain, real *arrDeriv, int iVarStartMain);
/* replaces filtering.c:98-102 (filtering.h:6): compact x filter then
 * spectral y filter of all NVAR fields of uu, in place */
void filtering(void);
/* replace calcRHS.c:46-75, :13-22, :24-43 (calcRHS.h:4-6).  calcRHS covers
 * the periodic branch with viscosity, resistivity and both div B cleanings:
 * diffusive (divBcleaning 1) and hyperbolic GLM (2, NVAR = 9 with psi last,
 * calcRHS.c:144-158, using the Ch / Cp the last calcDt set) */
void calcRHS(void);
void calcDivB(void);
void calcEfield(void);
/* replace timeAdvance.c:20-27, :29-36, :38-153 (timeAdvance.h:6-8).  With
 * divBcleaning == 2, calcDt also sets Ch_divBcleaning to the fastest speed
 * of the loop's last cell (nx-1, ny-1) and Cp via calc_Cp_fix_Cr, as the
 * reference's loop leaves them (timeAdvance.c:107-112) */
void rkInit(real dt);
void rkExecute(int iRK);
real calcDt(real dt0);
/* replace parallel.c:4, :66, :130 (parallel.h:4-6): with the whole grid on
 * one GPU the periodic halo is read by index wrap, so these do nothing */
void sendrecvHaloPointsX(void);
void sendrecvHaloPointsdivBX(void);
void sendrecvHaloPointsEfieldX(void);
/* replace divBcleaning.c:4-10 (divBcleaning.h:10-11) */
extern real Cr_divBCleaning;
void calc_Cp_fix_Cr(void);

/* Host-copy policy of the drop-in layer.
 *  MHD2D_SYNC_STRICT (default): calcDt copies uu, dudx and dudy back to the
 *    host every step, so mhd.c's checkDivB (mhd.c:342-367) and writeArray
 *    (writeOutput.c:47-76), which read host memory directly, see current data.
 *  MHD2D_SYNC_EXPLICIT: host copies are refreshed only by
 *    mhd2d_dropin_sync_host(); call it before any host-side read. */
enum { MHD2D_SYNC_STRICT = 0, MHD2D_SYNC_EXPLICIT = 1 };
void mhd2d_dropin_set_sync(int mode);
/* copy every device-newer array back to its host array */
int mhd2d_dropin_sync_host(void);
/* declare that the host has written the arrays (e.g. a restart): the next
 * call uploads them again */
void mhd2d_dropin_invalidate_device(void);
/* free all device mirrors (end of run) */
void mhd2d_dropin_release(void);

/* Host storage of the reference globals for builds that do not link the
 * reference's initialize.c (tests, the standalone driver): allocates uu,
 * dudx, ... like initialize.c:87-395 for one rank and periodic x, with the
 * parameters given.  Not part of the reference ABI. */
int mhd2d_host_initialize(int nx, int ny, double Lx, double Ly, double gamma,
                          double cfl, int hall, double di, double resistivity,
                          double viscosity, int divbclean);
void mhd2d_host_finalize(void);

/* ===================================================================== */
/* 2. context API (fused device stepping)                                */
/* ===================================================================== */
typedef struct mhd2d_ctx mhd2d_ctx;

/* precision: FP64 everywhere; FP32 everywhere; MIXED = FP64 state, filters
 * and RK accumulation with FP32 derivative solves, derivative arrays and
 * R_prev (the form that meets 1e-6 against FP64 over 200 steps) */
enum { MHD2D_FP64 = 0, MHD2D_FP32 = 1, MHD2D_MIXED = 2 };

typedef struct mhd2d_params {
    int nx, ny;             /* grid, powers of two in [16, 4096] */
    double Lx, Ly;          /* periodic box; dx = Lx/nx, dy = Ly/ny */
    double gamma;           /* adiabatic index */
    double cfl;
    int hall;               /* 1: whistler CFL limit with di (timeAdvance.c:97) */
    double di;
    double resistivity;     /* >= 0: eta Lap B (addSourceTerm, calcRHS.c:498-531) */
    double viscosity;       /* >= 0: nu/rho Lap u (same) */
    int divbclean;          /* 0, 1 (diffusive) or 2 (GLM: the state gains
                               psi as a ninth field, host arrays [9][nx][ny]) */
    int precision;          /* MHD2D_FP64, MHD2D_FP32 or MHD2D_MIXED */
    int xblocks;            /* 0 or 1: x is one block (the reference on one MPI
                               rank); npe > 1: x split into npe blocks sized as
                               decomposeGrid1D (initialize.c:547-569), each
                               solved with its own end rows, so results match a
                               reference run on npe ranks; nx / npe >= 4 */
    int batch;              /* 0 or 1: one run; B > 1: B independent runs of
                               this grid advanced together, one launch per
                               kernel and stage for all of them; each has its
                               own state, d_i (mhd2d_set_di), dt, time and
                               stop.  Instance 0 is the one the calls without
                               an _at suffix address */
    int xgemm;              /* 0: x-derivatives by the compact tridiagonal
                               solve; 3 or 6: by the dense operator A^-1 B
                               split into BF16 planes and applied as MFMA GEMMs
                               with FP32 accumulation, 3 or 6 split products
                               (precision FP32 or MIXED, xblocks <= 1, nx and
                               ny >= 64) */
    int open_x;             /* 0: periodic x (boundaryTypeX 0); 1: open x
                               boundary (boundaryTypeX 1): one-sided closures
                               on the first and last x block and the
                               characteristic outflow condition on the lines
                               x = 0 and x = Lx (calcRHS.c:163-495), dx =
                               Lx/(nx-1); needs xblocks >= 2 (the reference
                               needs npe >= 2) and nx/xblocks >= 8 */
} mhd2d_params;

/* status codes */
enum {
    MHD2D_OK = 0,
    MHD2D_EINVAL = 1,      /* unsupported size or parameter */
    MHD2D_EHIP = 2,        /* HIP runtime error */
    MHD2D_ENONFINITE = 3,  /* a CFL limit was NaN/inf/<=0 */
    MHD2D_ENOMEM = 4
};

int mhd2d_create(const mhd2d_params *p, mhd2d_ctx **out);
void mhd2d_destroy(mhd2d_ctx *c);
/* upload a state [8][nx][ny] (FP64 host; [9][nx][ny] with psi last when
 * divbclean == 2); resets time, step and dt, and the first step's dt is
 * calcDt(0) of this state (mhd.c:58) */
int mhd2d_set_state(mhd2d_ctx *c, const double *uu);
/* download the state after the last step (applies the pending y filter) */
int mhd2d_get_state(mhd2d_ctx *c, double *uu);
/* advance nsteps full RK3 steps (mhd.c:154-206), asynchronously */
int mhd2d_run(mhd2d_ctx *c, int nsteps);
/* same, with the step captured once into a hipGraph and replayed */
int mhd2d_run_graph(mhd2d_ctx *c, int nsteps);
int mhd2d_sync(mhd2d_ctx *c);
/* dt of the last step, simulated time, number of steps done */
int mhd2d_get_scalars(mhd2d_ctx *c, double *dt, double *time, long long *steps);
/* total energy sum(rho u^2/2 + p/(gamma-1) + B^2/2) dx dy and max |div B|
 * of the current state, FP64 reductions */
int mhd2d_diagnostics(mhd2d_ctx *c, double *energy, double *max_divb);
/* dBx/dx + dBy/dy of the current state, [nx][ny] FP64 host array */
int mhd2d_get_divb(mhd2d_ctx *c, double *divb);
/* 1 when the run has stopped because the new dt fell below 1e-8 (mhd.c:131):
 * that step and all later ones are not taken; *dt_stop gets the small dt.
 * 0 while running, or an error code > 1 */
int mhd2d_stopped(mhd2d_ctx *c, double *dt_stop);
/* per-instance forms for batch > 1 (inst in [0, batch)); every instance
 * needs a state before mhd2d_run.  Setting a state completes the pending
 * filter of the others and restarts only that instance's time and dt. */
int mhd2d_set_state_at(mhd2d_ctx *c, int inst, const double *uu);
int mhd2d_get_state_at(mhd2d_ctx *c, int inst, double *uu);
/* ion inertial length of one instance (default: params.di), used by its
 * Hall CFL limit; takes effect from the instance's next dt */
int mhd2d_set_di(mhd2d_ctx *c, int inst, double di);
int mhd2d_get_scalars_at(mhd2d_ctx *c, int inst, double *dt, double *time,
                         long long *steps);
int mhd2d_diagnostics_at(mhd2d_ctx *c, int inst, double *energy, double *max_divb);
int mhd2d_stopped_at(mhd2d_ctx *c, int inst, double *dt_stop);
/* d/dx of nv host fields [nv][nx][ny] (nv <= 11) with the context's own
 * x-derivative (tridiagonal or GEMM, in its precision), into FP64 host out;
 * the run's state is untouched (temporary device buffers) */
int mhd2d_xderiv(mhd2d_ctx *c, int nv, const double *in, double *out);
/* GEMM form: number of non-zero 64 x 32 panels of the split operator over
 * all 64-row tiles (the MFMA work is panels * 64 * 32 * 2 * ny flops per
 * field and split product); 0 for the tridiagonal form */
long long mhd2d_xgemm_panels(const mhd2d_ctx *c);
/* Autotuner (no counterpart in the reference): times every valid x-solve
 * launch shape - strip width of 8, 16 or 32 y-lines per workgroup and the
 * XCD-contiguous workgroup order - for the derivative solves, then for the
 * filter, as reps whole RK3 steps of this context's own run, each from the
 * same state, and keeps the fastest.  Needs a state for every instance; the
 * state, R_prev, time, dt and step count are restored afterwards.  choice
 * (may be NULL) receives {width_d, remap_d, width_f, remap_f}.  Results do
 * not depend on the shape (each y-line is solved the same way).  For a
 * batch it then times instance groups (mhd2d_set_subbatch) of half, a
 * quarter, ... of the batch, down to 2^21 cells a group, and keeps a group
 * size only where it beats the whole batch by 1 %.  Where the fused
 * x-derivative + RHS applies (mhd2d_set_xfuse) it times the step with and
 * without it (before the group sizes) and keeps the faster; mhd2d_get_xfuse
 * tells which.  Tridiagonal form only. */
int mhd2d_autotune(mhd2d_ctx *c, int reps, int *choice);
/* Instances per launch of a batch (no counterpart in the reference, whose
 * runs are one grid each): a step runs the batch as groups of k instances,
 * one whole RK3 step per group, so that a group's fields stay in the GPU's
 * 256 MB Infinity Cache from kernel to kernel.  k = 0 or k >= batch: the
 * whole batch in every launch (the default).  Results are bitwise the same
 * for every k.  mhd2d_get_subbatch returns the group size in use. */
int mhd2d_set_subbatch(mhd2d_ctx *c, int k);
int mhd2d_get_subbatch(const mhd2d_ctx *c);
/* Fused x-derivative + RHS (no counterpart in the reference, whose calcRHS
 * reads the derivative arrays derivX wrote, mhd.c:186-198): on = 1 runs each
 * stage's x-derivative as a solve that keeps only the values entering each
 * chunk of rows, then one kernel that re-solves every chunk from them in
 * on-chip memory and applies calcRHS + rkExecute there, so the x-derivative
 * arrays never reach device memory.  Results are bitwise those of on = 0.
 * Ideal MHD (no viscosity, resistivity or div B cleaning) on one periodic x
 * block with the tridiagonal x-derivative only: MHD2D_EINVAL otherwise.
 * A context starts with it on where it applies in MHD2D_FP64 and MHD2D_FP32
 * and off in MHD2D_MIXED; mhd2d_autotune keeps whichever runs faster.
 * mhd2d_get_xfuse returns the setting. */
int mhd2d_set_xfuse(mhd2d_ctx *c, int on);
int mhd2d_get_xfuse(const mhd2d_ctx *c);
/* wall time of nsteps on the context's stream (hipEvents), in ms */
int mhd2d_time_run(mhd2d_ctx *c, int nsteps, int use_graph, double *ms);
/* average device time per launch of each kernel class over nsteps, in ms:
 * out[0] y-sweep, [1] x-derivative, [2] RHS+RK, [3] x-filter, [4] dt */
int mhd2d_profile_run(mhd2d_ctx *c, int nsteps, double *out_ms, int *out_launches);
/* bytes of device memory held by the context */
size_t mhd2d_device_bytes(const mhd2d_ctx *c);
/* measurement helper: a plain device-to-device streaming copy of `bytes`,
 * run `iters` times; *gbps = 2 * bytes / average kernel time (read + write).
 * Used to calibrate the profiler's HBM byte counters and to state the
 * achievable copy rate beside the 8 TB/s spec peak. */
int mhd2d_copy_probe(size_t bytes, int iters, double *gbps);
const char *mhd2d_last_error(void);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): the same as mhd2d_create + mhd2d_set_state + mhd2d_run
 * (nsteps) + mhd2d_get_state, bit for bit, for one FP64 run whose state
 * [8][nx][ny] ([9][nx][ny] with psi last when divbclean == 2) is in caller
 * device memory d_uu and is advanced in place, from time 0 (the first dt is
 * calcDt(0) of the state).  The working arrays (derivatives, E, R_prev,
 * source-term buffers, run scalars, operator tables) live in one fixed
 * library array of MHD2D_NOALLOC_BYTES on each GPU the library runs on, held
 * until the process ends; nothing is allocated or freed per call.  Limits:
 * precision MHD2D_FP64, batch <= 1, xgemm 0, nx * ny <= MHD2D_NOALLOC_MAX_CELLS
 * (larger grids: MHD2D_EINVAL with a message).  Every call on one GPU uses
 * that GPU's array, so calls that overlap on one GPU (other host threads,
 * other streams) race.  stream: a hipStream_t, or NULL for the null stream;
 * the call returns when the work is done.  dt_out, time_out (may be NULL):
 * the dt of the last step and the simulated time reached. */
#define MHD2D_NOALLOC_MAX_CELLS 65536
#define MHD2D_NOALLOC_BYTES (28u << 20)
int mhd2d_run_noalloc(const mhd2d_params *p, double *d_uu, int nsteps, void *stream,
                      double *dt_out, double *time_out);

/* Ensembles of runs on grids of different sizes.  p[0..count) gives each
 * run's parameters (the batch field is ignored); runs whose parameters agree
 * except for di are grouped into one batched context per grid shape, and a
 * step advances every group: eagerly on one stream per group, or as one
 * hipGraph that forks into the groups' streams and joins them again.  All
 * entries must share one precision.  inst indexes p. */
typedef struct mhd2d_ens mhd2d_ens;
int mhd2d_ens_create(const mhd2d_params *p, int count, mhd2d_ens **out);
void mhd2d_ens_destroy(mhd2d_ens *e);
int mhd2d_ens_groups(const mhd2d_ens *e);
int mhd2d_ens_set_state(mhd2d_ens *e, int inst, const double *uu);
int mhd2d_ens_get_state(mhd2d_ens *e, int inst, double *uu);
int mhd2d_ens_run(mhd2d_ens *e, int nsteps, int use_graph);
/* mhd2d_autotune for every group; choice (may be NULL) receives 4 ints per
 * group in group order */
int mhd2d_ens_autotune(mhd2d_ens *e, int reps, int *choice);
int mhd2d_ens_sync(mhd2d_ens *e);
int mhd2d_ens_get_scalars(mhd2d_ens *e, int inst, double *dt, double *time,
                          long long *steps);
int mhd2d_ens_diagnostics(mhd2d_ens *e, int inst, double *energy, double *max_divb);
/* wall time of nsteps of every group (hipEvents around the fork / join) */
int mhd2d_ens_time_run(mhd2d_ens *e, int nsteps, int use_graph, double *ms);
/* How a step launches the groups (no counterpart in the reference):
 * MHD2D_ENS_ONE_LAUNCH runs every group with one launch per kernel class
 * and stage (13 launches a step, 16 with viscosity or resistivity), each
 * block finding its group and line length; MHD2D_ENS_PER_GROUP runs each
 * group's own batched launches on its own stream, forked from and joined
 * into one master stream.  Results are the same bit for bit.  ONE_LAUNCH is
 * the default where every group fits it: periodic x, no div B cleaning, one
 * x block, the tridiagonal x-derivative, the same source terms in every
 * group, nx in [256, 4096], ny in [128, 1024] (FP32 [128, 2048]); elsewhere
 * it is refused with MHD2D_EINVAL.  mhd2d_ens_autotune keeps the faster. */
enum { MHD2D_ENS_PER_GROUP = 0, MHD2D_ENS_ONE_LAUNCH = 1 };
int mhd2d_ens_set_launch(mhd2d_ens *e, int mode);
int mhd2d_ens_get_launch(const mhd2d_ens *e);
/* kernel launches of one step in the current mode */
int mhd2d_ens_launches_per_step(const mhd2d_ens *e);

/* ===================================================================== */
/* 3. the reference's run configuration, output files and run loop       */
/* ===================================================================== */
/* the 15 values of the reference's `input` file, in its order
 * (initialize.c:456-525) */
typedef struct mhd2d_input {
    int boundaryTypeX;
    double cfl, Tmax, dtOut;
    int nx, ny;
    double Lx, Ly, gamma;
    int hall;
    double di, resistivity, viscosity;
    int divbclean, outdivb;
} mhd2d_input;

/* parse an `input` file: each value from the first 16 characters of its
 * line with atoi/atof, as readLine/readInput do (initialize.c:442-525) */
int mhd2d_read_input(const char *path, mhd2d_input *in);
/* write an `input` file the reference's reader accepts */
int mhd2d_write_input(const char *path, const mhd2d_input *in);
/* xgrid.dat / ygrid.dat: [size_t n][n doubles] (writeOutput.c:5-45) */
int mhd2d_write_grid(const char *dir, int nx, int ny, double Lx, double Ly);
/* the same with the open-x grid spacing Lx/(nx-1) when open_x != 0 */
int mhd2d_write_grid_bc(const char *dir, int nx, int ny, double Lx, double Ly, int open_x);
/* out%04d.dat: [double t][nvar*nx*ny doubles] (writeOutput.c:47-76) */
int mhd2d_write_array(const char *dir, int iout, double t, const double *uu,
                      int nvar, int nx, int ny);
/* divB%04d.dat: [double t][nx*ny doubles] (writeOutput.c:78-107) */
int mhd2d_write_divb(const char *dir, int iout, double t, const double *divb,
                     int nx, int ny);
/* read an out%04d.dat back (restart; the reference has no reader) */
int mhd2d_read_array(const char *path, double *t, double *uu, size_t count);
/* the `log` file of mhd.c:319-338 */
int mhd2d_write_log(const char *dir, double t, double dt, double seconds,
                    long long step);

enum { MHD2D_DRIVE_AUTO = 0, MHD2D_DRIVE_FUSED = 1, MHD2D_DRIVE_DROPIN = 2 };

typedef struct mhd2d_drive_opts {
    const char *outdir;     /* directory of the output files ("" = cwd) */
    int mode;               /* MHD2D_DRIVE_*: AUTO = FUSED */
    int precision;          /* MHD2D_FP64 / MHD2D_FP32 (FUSED only) */
    int verbose;            /* print the reference's progress lines */
    long long max_steps;    /* stop after this many steps (0: no cap) */
    int xblocks;            /* emulate a run on this many MPI ranks (FUSED
                               only; 0 or 1: one rank) */
} mhd2d_drive_opts;

typedef struct mhd2d_drive_report {
    double time, dt;        /* simulated time reached, last dt */
    long long steps;        /* RK3 steps taken */
    int outputs;            /* out%04d.dat files written */
    int stopped;            /* 1: ended by the dt < 1e-8 stop (mhd.c:131) */
    double seconds;         /* wall time of the loop */
    double energy;          /* total energy at the end (FUSED only) */
    double max_divb;        /* max |div B| at the last output */
} mhd2d_drive_report;

/* The time-driven loop of mhd.c:58-230 on the GPU: outputs every dtOut,
 * log every dtOut/10, the dt < 1e-8 stop, final output after Tmax.  Writes
 * xgrid/ygrid are the caller's (mhd2d_write_grid).  Single thread; the
 * DROPIN engine uses the library's own storage of the reference globals
 * (libmhd2d.so only).  uu0 holds the 8 physical fields; with divbclean == 2
 * psi starts at 0 and the output files hold NVAR = 9 fields. */
int mhd2d_drive(const mhd2d_input *in, const double *uu0,
                const mhd2d_drive_opts *o, mhd2d_drive_report *rep);

/* ===================================================================== */
/* initial states (host, FP64) shared by the device runs and the checker  */
/* ===================================================================== */
/* Orszag-Tang vortex on [0,Lx)x[0,Ly) */
void mhd2d_ic_orszag_tang(int nx, int ny, double Lx, double Ly, double *uu);
/* rho=p=1, B0=(1,0,0); du and dB summed over integer wavevectors
 * kmin <= |k| <= kmax with |k|^(-5/3) amplitudes and SplitMix64 phases;
 * in-plane parts from stream functions (dB from A_z, so div B = 0); each of
 * du, dB scaled to the given rms */
void mhd2d_ic_turbulence(int nx, int ny, double Lx, double Ly,
                         unsigned long long seed, double rms, int kmin,
                         int kmax, double *uu);
/* rho=p=1, B0=(1,0,0) plus one linear wave with integer (kx, ky) in
 * [1,8]^2 drawn from the seed, amplitude amp, random phase */
void mhd2d_ic_linear_wave(int nx, int ny, double Lx, double Ly,
                          unsigned long long seed, double amp, double *uu);

#ifdef __cplusplus
}
#endif
#endif /* MHD2D_H */
