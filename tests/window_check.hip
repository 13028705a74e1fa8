// CPU check of the voxelizer's windowed t evaluation (voxel_window.h): for
// random and adversarial (column, facet) pairs, the grid points flipped by
// window_toggle equal those of the literal expression evaluated at every k.
// Built and run by test_voxel_window.py; prints the counts and exits non-zero
// on a mismatch.  Host code only: no kernel is launched.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>

#include "voxel_window.h"

struct HostBits {
    uint32_t w[64];
    void flip(int k) { w[k >> 5] ^= 1u << (k & 31); }
    void flip_range(int lo, int hi) {
        for (int k = lo; k <= hi; ++k) flip(k);
    }
};

int main(int argc, char** argv) {
    const long iters = argc > 1 ? atol(argv[1]) : 1000000;
    std::mt19937_64 rng(12345);
    std::uniform_real_distribution<double> U(0.0, 1.0);
    long bad = 0, nontrivial = 0;
    for (long it = 0; it < iters; ++it) {
        lbm_vox::WinRay g{};
        const int mode = (int)(it % 8);
        const double scale = mode == 5 ? 1e4 : (mode == 6 ? 1e-3 : 1.0);
        g.nz = 1 + (int)(U(rng) * (mode == 7 ? 2047 : 300));
        g.sp = (mode == 6 ? 1e-3 : 1.0) * (0.05 + U(rng) * 2.0);
        if (mode == 3) g.sp = 1.0;  // integer lattice: exact ties
        g.oz = (U(rng) - 0.5) * 200 * scale;
        if (mode == 3) g.oz = std::floor(g.oz);
        const double zmin = g.oz + (U(rng) - 0.3) * 50 * scale;
        const double zmax = mode == 4 ? zmin : zmin + U(rng) * g.nz * g.sp * 1.2;  // mode 4: flat, L = 0
        g.L = zmax - zmin;
        g.pad = mode == 2 ? U(rng) * 5 : 0.1;
        g.teps = mode == 1 ? U(rng) * 0.5 : 1e-7;
        double T[12];
        for (int c = 0; c < 9; ++c) T[c] = (U(rng) - 0.5) * 400 * scale;
        for (int v = 0; v < 3; ++v) T[3 * v + 2] = g.oz + U(rng) * g.nz * g.sp * 1.3 - 0.15 * g.nz * g.sp;
        if (mode == 3)
            for (int c = 0; c < 9; ++c) T[c] = std::round(T[c]);
        // file normals: random, tiny z, tiny z near 1e-200, huge x, powers of two, axial
        double nx = U(rng) - 0.5, ny = U(rng) - 0.5, nz = U(rng) - 0.5;
        const int nm = (int)((it / 8) % 6);
        if (nm == 1) nz *= 1e-9;
        if (nm == 2) nz *= 1e-200;
        if (nm == 3) nx *= 1e200;
        if (nm == 4) nz = (U(rng) < 0.5 ? -1 : 1) * std::ldexp(1.0, -(int)(U(rng) * 60));
        if (nm == 5) {
            nx = 0;
            ny = 0;
            nz = U(rng) < 0.5 ? -1.0 : 1.0;
        }
        T[9] = nx;
        T[10] = ny;
        T[11] = nz;
        double sx = (U(rng) - 0.5) * 400 * scale, sy = (U(rng) - 0.5) * 400 * scale;
        if (mode == 3) {
            sx = std::round(sx);
            sy = std::round(sy);
        }
        double ex = sx, ey = sy;
        if (it % 13 == 0) {  // an offset column
            sx += 0.01;
            ex += 0.01;
        }
        // put an end of the window on (or within rounding of) a grid point
        if (it % 3 == 0 && nz != 0) {
            const int k = (int)(U(rng) * g.nz);
            const double A = (T[0] - sx) * nx + (T[1] - sy) * ny;
            double target = g.oz + k * g.sp + (U(rng) - 0.5) * 1e-12 * scale;
            if (it % 6 == 0) target -= g.teps * (g.L + g.pad);
            else target += (1 + g.teps) * (g.L + g.pad);
            T[2] = target - A / nz;
        }
        HostBits lit, win;
        std::memset(&lit, 0, sizeof lit);
        std::memset(&win, 0, sizeof win);
        lbm_vox::toggle_exact(g, T, sx, sy, ex, ey, 0, g.nz - 1, lit);
        lbm_vox::window_toggle(g, T, sx, sy, ex, ey, win);
        int cnt = 0;
        for (int k = 0; k < g.nz; ++k) cnt += (lit.w[k >> 5] >> (k & 31)) & 1;
        if (cnt > 0 && cnt < g.nz) nontrivial++;
        if (std::memcmp(lit.w, win.w, sizeof lit.w) != 0) {
            if (bad < 10) std::printf("mismatch at pair %ld (mode %d, normal %d)\n", it, mode, nm);
            bad++;
        }
    }
    std::printf("pairs %ld mismatches %ld nontrivial %ld\n", iters, bad, nontrivial);
    return bad != 0;
}
