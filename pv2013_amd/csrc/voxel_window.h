// The reference's ray-parameter window for one column and one facet, and the
// set of grid points of the column whose t lies in it (ComplexGeometry.cpp:
// 441-458 for the ray, :315-316 for t, :320 for the window).  Shared by the
// voxelizer kernel and the CPU check of the argument below, so both run the
// same code.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

namespace lbm_vox {

// The z part of a column's rays: grid point k starts at oz + k sp and the ray
// ends L + pad above it; t is in the window when -teps <= t <= 1 + teps.
struct WinRay {
    double oz, sp, L, pad, teps;
    int32_t nz;
};

// The literal expression for grid point k (T: v1, v2, v3, file normal).
__host__ __device__ inline bool t_window(const WinRay& g, const double* T, double sx, double sy, double ex,
                                         double ey, int k) {
    const double nx = T[9], ny = T[10], nz = T[11];
    const double sz = g.oz + k * g.sp;
    const double ez = sz + g.L + g.pad;
    const double tt = ((T[0] - sx) * nx + (T[1] - sy) * ny + (T[2] - sz) * nz) /
                      ((ex - sx) * nx + (ey - sy) * ny + (ez - sz) * nz);
    return tt >= -g.teps && tt <= 1 + g.teps;
}

template <class Bits>
__host__ __device__ inline void toggle_exact(const WinRay& g, const double* T, double sx, double sy, double ex,
                                             double ey, int lo, int hi, Bits& bits) {
    for (int k = lo; k <= hi; ++k)
        if (t_window(g, T, sx, sy, ex, ey, k)) bits.flip(k);
}

// Flips the bit of every grid point k of the column whose t is in the window
// (Bits provides flip(k) and flip_range(lo, hi), both ends included).
//
// With ex == sx and ey == sy the reference's t at height z is, in real
// arithmetic, (zp - z) / H with H = L + pad and zp = T2 + A / nz, where
// A = (T0 - sx) nx + (T1 - sy) ny is the reference's own partial sum.  The
// window is then the interval zp - (1 + teps) H <= z <= zp + teps H.  Every
// rounding step (z_k, the numerator, the denominator ez - sz, the constant
// 1 + teps, and this estimate) errs by at most a few units in the last place
// of the magnitudes involved, so M below, 64 ulp of their sum, bounds how far
// the rounded decision can move an end of the interval.  Grid points within M
// of an end (plus one on each side) are decided by the literal expression;
// those strictly between the two ends are in the window and the rest are not.
// Pairs where that argument does not hold (non-zero transverse ray component,
// non-finite or extreme values, an end region wider than 16 points) evaluate
// every k literally.
template <class Bits>
__host__ __device__ inline void window_toggle(const WinRay& g, const double* T, double sx, double sy, double ex,
                                              double ey, Bits& bits) {
    const double nx = T[9], ny = T[10], nz = T[11];
    const double A = (T[0] - sx) * nx + (T[1] - sy) * ny;
    const double D0 = (ex - sx) * nx + (ey - sy) * ny;
    const double H = g.L + g.pad;
    const double q = A / nz;
    const double S = fabs(g.oz) + (double)g.nz * g.sp;
    const double mag = fabs(T[2]) + fabs(q) + 2.0 * (S + fabs(g.L) + fabs(g.pad)) + (2.0 + g.teps) * H;
    const double M = 64.0 * 0x1p-53 * mag;
    const bool ok = D0 == 0.0 && fabs(nz) >= 1e-150 && fabs(nz) <= 1e150 && mag <= 1e100 && H > 1e-9 * mag &&
                    g.teps >= 0.0 && g.teps <= 1e6 && g.sp > 1e-9 * mag && fabs(A) <= 1e250;
    const int last = g.nz - 1;
    if (!ok) {
        toggle_exact(g, T, sx, sy, ex, ey, 0, last, bits);
        return;
    }
    const double zp = T[2] + q;
    const double zlo = zp - (1.0 + g.teps) * H, zhi = zp + g.teps * H;
    // grid index range [floor((z - M - oz) / sp) - 1, floor((z + M - oz) / sp) + 1], clamped
    auto kidx = [&](double z, double off) {
        const double v = floor((z - g.oz) / g.sp) + off;
        return (int)fmin(fmax(v, -2.0), (double)g.nz + 1.0);
    };
    const int a0 = kidx(zlo - M, -1.0), a1 = kidx(zlo + M, 1.0);
    const int b0 = kidx(zhi - M, -1.0), b1 = kidx(zhi + M, 1.0);
    auto lo0 = [](int v) { return v < 0 ? 0 : v; };
    auto hiN = [&](int v) { return v > last ? last : v; };
    if (a1 - a0 > 16 || b1 - b0 > 16) {
        toggle_exact(g, T, sx, sy, ex, ey, 0, last, bits);
        return;
    }
    if (b0 <= a1 + 1) {  // the two end regions touch: decide all of them literally
        toggle_exact(g, T, sx, sy, ex, ey, lo0(a0), hiN(b1), bits);
        return;
    }
    toggle_exact(g, T, sx, sy, ex, ey, lo0(a0), hiN(a1), bits);
    const int m0 = lo0(a1 + 1), m1 = hiN(b0 - 1);
    if (m0 <= m1) bits.flip_range(m0, m1);
    toggle_exact(g, T, sx, sy, ex, ey, lo0(b0), hiN(b1), bits);
}

}  // namespace lbm_vox
