// Seeded initial states (host, FP64).  The reference's own user hook
// (usrInitialize.c) draws phases from unseeded rand(); these generators are
// deterministic so the device runs and the checker start from bit-identical
// inputs.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstdint>
#include <vector>

#include "../../include/mhd2d.h"

namespace {

const double kPi = 3.14159265358979323846264338327950288;

// SplitMix64 (Steele, Lea, Flood 2014)
struct SplitMix64 {
    uint64_t s;
    uint64_t next()
    {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    double uniform() { return (next() >> 11) * (1.0 / 9007199254740992.0); }  // [0,1)
};

struct Mode {
    int kx, ky;
    double amp;
    double phase;
};

// F[i][j] += sum_m amp_m * trig(2 pi (kx i / nx + ky j / ny) + phase_m),
// trig = sin (use_cos == false) or cos, grouped by kx so the work is
// O(#kx * nx * ny) instead of O(#modes * nx * ny).
void add_modes(int nx, int ny, const std::vector<Mode>& modes, bool use_cos, double* F)
{
    int kmax = 0;
    for (const auto& m : modes) kmax = std::max(kmax, std::abs(m.kx));
    std::vector<double> gs(ny), gc(ny), sx(nx), cx(nx);
    for (int kx = -kmax; kx <= kmax; kx++) {
        bool any = false;
        std::fill(gs.begin(), gs.end(), 0.0);
        std::fill(gc.begin(), gc.end(), 0.0);
        for (const auto& m : modes) {
            if (m.kx != kx) continue;
            any = true;
            for (int j = 0; j < ny; j++) {
                const double Y = 2 * kPi * (double)m.ky * j / ny + m.phase;
                gs[j] += m.amp * std::cos(Y);
                gc[j] += m.amp * std::sin(Y);
            }
        }
        if (!any) continue;
        for (int i = 0; i < nx; i++) {
            const double X = 2 * kPi * (double)kx * i / nx;
            sx[i] = std::sin(X);
            cx[i] = std::cos(X);
        }
        for (int i = 0; i < nx; i++) {
            double* row = F + (size_t)i * ny;
            // sin(X+Y) = sin X cos Y + cos X sin Y ; cos(X+Y) = cos X cos Y - sin X sin Y
            if (!use_cos)
                for (int j = 0; j < ny; j++) row[j] += sx[i] * gs[j] + cx[i] * gc[j];
            else
                for (int j = 0; j < ny; j++) row[j] += cx[i] * gs[j] - sx[i] * gc[j];
        }
    }
}

}  // namespace

extern "C" {

void mhd2d_ic_orszag_tang(int nx, int ny, double Lx, double Ly, double* uu)
{
    const size_t n = (size_t)nx * ny;
    const double s = 1.0 / std::sqrt(4 * kPi);
    for (int i = 0; i < nx; i++)
        for (int j = 0; j < ny; j++) {
            const size_t p = (size_t)i * ny + j;
            const double x = (double)i * Lx / nx * (2 * kPi / Lx);
            const double y = (double)j * Ly / ny * (2 * kPi / Ly);
            uu[0 * n + p] = 25.0 / (36.0 * kPi);
            uu[1 * n + p] = -std::sin(y);
            uu[2 * n + p] = std::sin(x);
            uu[3 * n + p] = 0.0;
            uu[4 * n + p] = -std::sin(y) * s;
            uu[5 * n + p] = std::sin(2 * x) * s;
            uu[6 * n + p] = 0.0;
            uu[7 * n + p] = 5.0 / (12.0 * kPi);
        }
}

void mhd2d_ic_turbulence(int nx, int ny, double Lx, double Ly, unsigned long long seed,
                         double rms, int kmin, int kmax, double* uu)
{
    const size_t n = (size_t)nx * ny;
    SplitMix64 rng{seed};
    // one representative per +-k pair: ky > 0, or ky == 0 and kx > 0
    std::vector<Mode> psi, az, wz, bz;
    for (int kx = -kmax; kx <= kmax; kx++)
        for (int ky = 0; ky <= kmax; ky++) {
            if (ky == 0 && kx <= 0) continue;
            const int k2 = kx * kx + ky * ky;
            if (k2 < kmin * kmin || k2 > kmax * kmax) continue;
            const double k = std::sqrt((double)k2);
            const double a = std::pow(k, -5.0 / 3.0);
            psi.push_back({kx, ky, a / k, 2 * kPi * rng.uniform()});
            az.push_back({kx, ky, a / k, 2 * kPi * rng.uniform()});
            wz.push_back({kx, ky, a, 2 * kPi * rng.uniform()});
            bz.push_back({kx, ky, a, 2 * kPi * rng.uniform()});
        }
    const double kxs = 2 * kPi / Lx, kys = 2 * kPi / Ly;
    // in-plane parts from stream functions: (d/dy, -d/dx) of sum a sin(k.x + t)
    auto grad = [&](const std::vector<Mode>& src, bool dy, double sign) {
        std::vector<Mode> out = src;
        for (auto& m : out) m.amp *= sign * (dy ? kys * m.ky : kxs * m.kx);
        return out;
    };
    std::vector<double> f(6 * n, 0.0);
    add_modes(nx, ny, grad(psi, true, 1.0), true, &f[0 * n]);   // du_x
    add_modes(nx, ny, grad(psi, false, -1.0), true, &f[1 * n]); // du_y
    add_modes(nx, ny, wz, false, &f[2 * n]);                    // du_z
    add_modes(nx, ny, grad(az, true, 1.0), true, &f[3 * n]);    // dB_x
    add_modes(nx, ny, grad(az, false, -1.0), true, &f[4 * n]);  // dB_y
    add_modes(nx, ny, bz, false, &f[5 * n]);                    // dB_z
    double su = 0, sb = 0;
    for (size_t p = 0; p < n; p++) {
        for (int c = 0; c < 3; c++) {
            su += f[c * n + p] * f[c * n + p];
            sb += f[(3 + c) * n + p] * f[(3 + c) * n + p];
        }
    }
    const double scu = su > 0 ? rms / std::sqrt(su / n) : 0.0;
    const double scb = sb > 0 ? rms / std::sqrt(sb / n) : 0.0;
    for (size_t p = 0; p < n; p++) {
        uu[0 * n + p] = 1.0;
        uu[1 * n + p] = scu * f[0 * n + p];
        uu[2 * n + p] = scu * f[1 * n + p];
        uu[3 * n + p] = scu * f[2 * n + p];
        uu[4 * n + p] = 1.0 + scb * f[3 * n + p];
        uu[5 * n + p] = scb * f[4 * n + p];
        uu[6 * n + p] = scb * f[5 * n + p];
        uu[7 * n + p] = 1.0;
    }
}

void mhd2d_ic_linear_wave(int nx, int ny, double Lx, double Ly, unsigned long long seed,
                          double amp, double* uu)
{
    (void)Lx;
    (void)Ly;
    const size_t n = (size_t)nx * ny;
    SplitMix64 rng{seed};
    const int kx = 1 + (int)(rng.next() % 8);
    const int ky = 1 + (int)(rng.next() % 8);
    const double ph = 2 * kPi * rng.uniform();
    // shear Alfven eigenmode of B0 = x, rho = 1: dB_z = a cos(k.x + t), du_z = -dB_z
    for (int i = 0; i < nx; i++)
        for (int j = 0; j < ny; j++) {
            const size_t p = (size_t)i * ny + j;
            const double c = amp * std::cos(2 * kPi * ((double)kx * i / nx + (double)ky * j / ny) + ph);
            uu[0 * n + p] = 1.0;
            uu[1 * n + p] = 0.0;
            uu[2 * n + p] = 0.0;
            uu[3 * n + p] = -c;
            uu[4 * n + p] = 1.0;
            uu[5 * n + p] = 0.0;
            uu[6 * n + p] = c;
            uu[7 * n + p] = 1.0;
        }
}

}  // extern "C"
