// The reference's run-time configuration and output formats (host side):
//  - `input`: 15 lines, each value parsed from the line's first 16
//    characters with atoi/atof (initialize.c:442-525);
//  - xgrid.dat / ygrid.dat: [size_t n][n doubles] (writeOutput.c:5-45);
//  - out%04d.dat: [double t][NVAR*nx*ny doubles, C order v,i,j]
//    (writeOutput.c:47-76), divB%04d.dat: [double t][nx*ny doubles]
//    (writeOutput.c:78-107) -- the layout data_proc_scripts/read_output.py reads;
//  - log: the text of mhd.c:319-338.
#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/mhd2d.h"

namespace {

std::string join(const char* dir, const std::string& name)
{
    if (!dir || !*dir) return name;
    std::string d(dir);
    if (d.back() != '/') d += '/';
    return d + name;
}

// one line of `input`: the value sits in the first 16 characters
bool read_field(FILE* f, char out[17])
{
    char line[256];
    if (!std::fgets(line, sizeof line, f)) return false;
    std::memcpy(out, line, 16);
    out[16] = '\0';
    return true;
}

int write_block(const std::string& path, double t, const double* a, size_t count)
{
    FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) return MHD2D_EINVAL;
    bool ok = std::fwrite(&t, sizeof(double), 1, f) == 1 &&
              std::fwrite(a, sizeof(double), count, f) == count;
    ok = (std::fclose(f) == 0) && ok;
    return ok ? MHD2D_OK : MHD2D_EINVAL;
}

}  // namespace

extern "C" {

int mhd2d_read_input(const char* path, mhd2d_input* in)
{
    if (!path || !in) return MHD2D_EINVAL;
    FILE* f = std::fopen(path, "r");
    if (!f) return MHD2D_EINVAL;
    char v[15][17];
    for (int k = 0; k < 15; k++)
        if (!read_field(f, v[k])) {
            std::fclose(f);
            return MHD2D_EINVAL;
        }
    std::fclose(f);
    // order of initialize.c:465-522
    in->boundaryTypeX = std::atoi(v[0]);
    in->cfl = std::atof(v[1]);
    in->Tmax = std::atof(v[2]);
    in->dtOut = std::atof(v[3]);
    in->nx = std::atoi(v[4]);
    in->ny = std::atoi(v[5]);
    in->Lx = std::atof(v[6]);
    in->Ly = std::atof(v[7]);
    in->gamma = std::atof(v[8]);
    in->hall = std::atoi(v[9]);
    in->di = std::atof(v[10]);
    in->resistivity = std::atof(v[11]);
    in->viscosity = std::atof(v[12]);
    in->divbclean = std::atoi(v[13]);
    in->outdivb = std::atoi(v[14]);
    return MHD2D_OK;
}

int mhd2d_write_input(const char* path, const mhd2d_input* in)
{
    if (!path || !in) return MHD2D_EINVAL;
    FILE* f = std::fopen(path, "w");
    if (!f) return MHD2D_EINVAL;
    std::fprintf(f, "%-20d; boundary condition along x. 0: periodic, 1: open\n", in->boundaryTypeX);
    std::fprintf(f, "%-20.12g; cfl\n", in->cfl);
    std::fprintf(f, "%-20.12g; Tmax\n", in->Tmax);
    std::fprintf(f, "%-20.12g; dtOut\n", in->dtOut);
    std::fprintf(f, "%-20d; nx\n", in->nx);
    std::fprintf(f, "%-20d; ny\n", in->ny);
    std::fprintf(f, "%-20.12g; Lx\n", in->Lx);
    std::fprintf(f, "%-20.12g; Ly\n", in->Ly);
    std::fprintf(f, "%-20.12g; adiabatic_index\n", in->gamma);
    std::fprintf(f, "%-20d; Hall effect 0: off; 1: on\n", in->hall);
    std::fprintf(f, "%-20.12g; normalized ion inertial length di\n", in->di);
    std::fprintf(f, "%-20.12g; resistivity\n", in->resistivity);
    std::fprintf(f, "%-20.12g; viscosity\n", in->viscosity);
    std::fprintf(f, "%-20d; divB cleaning method: 0: none; 1: diffusion\n", in->divbclean);
    std::fprintf(f, "%-20d; whether output divB: 0: no; 1: yes\n", in->outdivb);
    return std::fclose(f) == 0 ? MHD2D_OK : MHD2D_EINVAL;
}

int mhd2d_write_grid(const char* dir, int nx, int ny, double Lx, double Ly)
{
    return mhd2d_write_grid_bc(dir, nx, ny, Lx, Ly, 0);
}

int mhd2d_write_grid_bc(const char* dir, int nx, int ny, double Lx, double Ly, int open_x)
{
    // xgrid[i] = i * dx with dx = Lx / nx (periodic) or Lx / (nx - 1) (open),
    // initialize.c:125-133, :145-148; ygrid[j] = j * Ly / ny
    const double dx = open_x ? Lx / (nx - 1) : Lx / nx, dy = Ly / ny;
    struct {
        const char* name;
        size_t n;
        double d;
    } g[2] = {{"xgrid.dat", (size_t)nx, dx}, {"ygrid.dat", (size_t)ny, dy}};
    for (auto& gi : g) {
        std::vector<double> v(gi.n);
        for (size_t i = 0; i < gi.n; i++) v[i] = (double)i * gi.d;
        FILE* f = std::fopen(join(dir, gi.name).c_str(), "wb");
        if (!f) return MHD2D_EINVAL;
        bool ok = std::fwrite(&gi.n, sizeof(size_t), 1, f) == 1 &&
                  std::fwrite(v.data(), sizeof(double), gi.n, f) == gi.n;
        ok = (std::fclose(f) == 0) && ok;
        if (!ok) return MHD2D_EINVAL;
    }
    return MHD2D_OK;
}

int mhd2d_write_array(const char* dir, int iout, double t, const double* uu, int nvar, int nx,
                      int ny)
{
    char name[32];
    std::snprintf(name, sizeof name, "out%04d.dat", iout);
    return write_block(join(dir, name), t, uu, (size_t)nvar * nx * ny);
}

int mhd2d_write_divb(const char* dir, int iout, double t, const double* divb, int nx, int ny)
{
    char name[32];
    std::snprintf(name, sizeof name, "divB%04d.dat", iout);
    return write_block(join(dir, name), t, divb, (size_t)nx * ny);
}

int mhd2d_read_array(const char* path, double* t, double* uu, size_t count)
{
    FILE* f = std::fopen(path, "rb");
    if (!f) return MHD2D_EINVAL;
    bool ok = std::fread(t, sizeof(double), 1, f) == 1 &&
              std::fread(uu, sizeof(double), count, f) == count;
    std::fclose(f);
    return ok ? MHD2D_OK : MHD2D_EINVAL;
}

// mhd.c:319-338
int mhd2d_write_log(const char* dir, double t, double dt, double seconds, long long step)
{
    const long s = (long)seconds;
    const int hour = (int)(s / 3600), minute = (int)((s % 3600) / 60), second = (int)(s % 60);
    FILE* f = std::fopen(join(dir, "log").c_str(), "w");
    if (!f) return MHD2D_EINVAL;
    std::fprintf(f, "  Simulation time = %f\n", t);
    std::fprintf(f, "  dt = %12.4e\n", dt);
    std::fprintf(f, "  Real time used (sec) = %12ld\n", s);
    std::fprintf(f, "  Real time used (hh:mm:ss) = (%03d:%02d:%02d)\n", hour, minute, second);
    std::fprintf(f, "  Iteration = %14lld\n", step);
    return std::fclose(f) == 0 ? MHD2D_OK : MHD2D_EINVAL;
}

}  // extern "C"
