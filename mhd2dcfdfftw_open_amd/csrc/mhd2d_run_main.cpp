// mhd2d_run: the reference's executable (mhd.exe) for periodic runs on one
// GPU.  Reads the reference's `input` file, builds a seeded initial state
// (or restarts from an out%04d.dat file), and runs the time-driven loop of
// mhd.c, writing xgrid.dat, ygrid.dat, out%04d.dat, divB%04d.dat and log in
// the reference's formats.
//
//   mhd2d_run [--input FILE] [--out DIR] [--ic ot|turbulence|wave|FILE]
//             [--seed N] [--rms R] [--kmax K] [--amp A]
//             [--engine auto|fused|dropin] [--fp32] [--npe N] [--max-steps N] [--quiet]
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/mhd2d.h"

static void usage()
{
    std::fprintf(stderr,
                 "usage: mhd2d_run [--input FILE] [--out DIR] [--ic ot|turbulence|wave|FILE]\n"
                 "                 [--seed N] [--rms R] [--kmax K] [--amp A]\n"
                 "                 [--engine auto|fused|dropin] [--fp32] [--npe N] [--max-steps N] [--quiet]\n");
}

int main(int argc, char** argv)
{
    std::string input = "./input", out = "", ic = "turbulence", engine = "auto";
    unsigned long long seed = 42;
    double rms = 0.1, amp = 1e-4;
    int kmax = 16, fp32 = 0, quiet = 0, npe = 1;
    long long max_steps = 0;
    for (int k = 1; k < argc; k++) {
        std::string a = argv[k];
        auto next = [&]() -> const char* {
            if (k + 1 >= argc) {
                usage();
                std::exit(2);
            }
            return argv[++k];
        };
        if (a == "--input") input = next();
        else if (a == "--out") out = next();
        else if (a == "--ic") ic = next();
        else if (a == "--seed") seed = std::strtoull(next(), nullptr, 10);
        else if (a == "--rms") rms = std::atof(next());
        else if (a == "--kmax") kmax = std::atoi(next());
        else if (a == "--amp") amp = std::atof(next());
        else if (a == "--engine") engine = next();
        else if (a == "--fp32") fp32 = 1;
        else if (a == "--npe") npe = std::atoi(next());  // results of an npe-rank run
        else if (a == "--max-steps") max_steps = std::atoll(next());
        else if (a == "--quiet") quiet = 1;
        else {
            usage();
            return 2;
        }
    }
    mhd2d_input in;
    if (mhd2d_read_input(input.c_str(), &in)) {
        std::fprintf(stderr, "mhd2d_run: cannot read %s\n", input.c_str());
        return 1;
    }
    const size_t n = (size_t)in.nx * in.ny;
    std::vector<double> uu(8 * n);
    if (ic == "ot") mhd2d_ic_orszag_tang(in.nx, in.ny, in.Lx, in.Ly, uu.data());
    else if (ic == "turbulence")
        mhd2d_ic_turbulence(in.nx, in.ny, in.Lx, in.Ly, seed, rms, 1, kmax, uu.data());
    else if (ic == "wave") mhd2d_ic_linear_wave(in.nx, in.ny, in.Lx, in.Ly, seed, amp, uu.data());
    else {
        double t0;
        if (mhd2d_read_array(ic.c_str(), &t0, uu.data(), 8 * n)) {
            std::fprintf(stderr, "mhd2d_run: cannot read initial state %s\n", ic.c_str());
            return 1;
        }
    }
    if (!quiet) {
        std::printf("****************************************************\n");
        std::printf("  (nx, ny) = (%d, %d), Lx = %f, Ly = %f\n", in.nx, in.ny, in.Lx, in.Ly);
        std::printf("  cfl = %f, Tmax = %f, dtOut = %f, gamma = %f\n", in.cfl, in.Tmax, in.dtOut,
                    in.gamma);
        std::printf("  Hall = %d, di = %f, resistivity = %e, viscosity = %e, divBcleaning = %d\n",
                    in.hall, in.di, in.resistivity, in.viscosity, in.divbclean);
        std::printf("****************************************************\n");
    }
    if (mhd2d_write_grid_bc(out.c_str(), in.nx, in.ny, in.Lx, in.Ly, in.boundaryTypeX == 1)) {
        std::fprintf(stderr, "mhd2d_run: cannot write the grid files\n");
        return 1;
    }
    mhd2d_drive_opts o;
    o.outdir = out.c_str();
    o.mode = engine == "fused" ? MHD2D_DRIVE_FUSED
                               : (engine == "dropin" ? MHD2D_DRIVE_DROPIN : MHD2D_DRIVE_AUTO);
    o.precision = fp32 ? MHD2D_FP32 : MHD2D_FP64;
    o.xblocks = npe;
    o.verbose = !quiet;
    o.max_steps = max_steps;
    mhd2d_drive_report rep;
    std::memset(&rep, 0, sizeof rep);
    const int rc = mhd2d_drive(&in, uu.data(), &o, &rep);
    if (rc) {
        std::fprintf(stderr, "mhd2d_run: error %d: %s\n", rc, mhd2d_last_error());
        return 1;
    }
    std::printf("Simulation end... time = %.6f, steps = %lld, outputs = %d, %.2f s%s\n",
                rep.time, rep.steps, rep.outputs, rep.seconds, rep.stopped ? " (dt < 1e-8)" : "");
    return 0;
}
