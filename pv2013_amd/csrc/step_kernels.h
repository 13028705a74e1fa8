// Fused stream-collide kernels for gfx950 (templated on lattice, precision and
// mode).  Included by kernels_f64.hip (built with -ffp-contract=off so the
// FP64 path rounds exactly like the reference) and kernels_lp.hip (FP32 /
// FP16 shifted storage, FP32 collision).
//
// One thread per fluid node, node ids packed in an order whose fastest axis is
// the vessel direction, populations SoA [d][n].  Per step and node:
//   pull   f[d] = g_in[d][node at x - c_d]      (Simulation.cpp:184-198 as a pull)
//   links  missing source: bounce-back g_in[opp d][x] or the stale value
//          (Simulation.cpp:1353-1370; the reference leaves unlinked
//          populations untouched)
//   BC     Zou-He velocity / pressure on inlet / outlet nodes
//          (Boundary.cpp:127-148, Simulation.cpp:952-1348)
//   BGK    rho, u, f_eq, g_out = omega f_eq + (1 - omega) f (Simulation.cpp:21-179)
// The stored g is the post-collision state; the reference's state f after the
// last step is recovered by the same pull + BC without the collision (mode PRE,
// two lattices) or written by the last step itself (mode SNAP, AA, whose single
// lattice no longer holds the previous state).
#pragma once

#include <type_traits>
#include "lattice.h"
#include "lbm_internal.h"

namespace lbm {

struct PrecF64 {
    using store_t = double;
    using calc_t = double;
};
struct PrecF32S {
    using store_t = float;
    using calc_t = float;
};
struct PrecF16S {
    using store_t = _Float16;
    using calc_t = float;
};

// Blocks are dispatched round-robin over the 8 XCDs (block b on XCD b & 7).
// tile 0: each XCD takes one contiguous range of logical blocks, so
// neighbouring rows share that XCD's L2.  tile T > 0: the XCDs take tiles of T
// consecutive logical blocks in turn, so the eight of them advance together
// as one front through the lattice (the blocks past the last whole round of
// tiles split into one range per XCD).  The grid is a multiple of 8 blocks,
// and either mapping is a bijection onto [0, gridDim.x).  rev: each XCD takes
// its blocks from the top, so a launch starts on the lines the previous
// launch touched last.
__device__ __forceinline__ int64_t logical_block(int rev = 0, int tile = 0) {
    const int64_t per = gridDim.x >> 3;  // blocks per XCD
    const int x = (int)(blockIdx.x & 7);
    int64_t k = blockIdx.x >> 3;         // this block's turn on its XCD
    if (rev) k = per - 1 - k;
    const int64_t T = tile > 0 && tile < per ? tile : per;
    const int64_t full = per / T * T;
    if (k < full) return ((k / T) * 8 + x) * T + k % T;
    const int64_t r = per - full;
    return 8 * full + x * r + (k - full);
}

// ------------------------------------------------------------- waveforms (FP64)

// pulsatile.h:12-28
__device__ inline double pv_eq1(double x) {
    return (-8.579e+08 * pow(x, 8) + 4.059e+08 * pow(x, 7) - 7.817e+07 * pow(x, 6) +
            7.994e+06 * pow(x, 5) - 4.835e+05 * pow(x, 4) + 1.782e+04 * pow(x, 3) -
            321.6 * pow(x, 2) + 3.215 * x + 0.1423);
}
__device__ inline double pv_eq2(double x) {
    return (-4.32e+06 * pow(x, 8) + 7.94e+06 * pow(x, 7) - 6.298e+06 * pow(x, 6) +
            2.82e+06 * pow(x, 5) - 7.809e+05 * pow(x, 4) + 1.37e+05 * pow(x, 3) -
            1.484e+04 * pow(x, 2) + 901.9 * x - 22.87);
}
__device__ inline double pv_eq3(double x) {
    return (-5.123 * pow(x, 4) + 14.15 * pow(x, 3) - 14.21 * pow(x, 2) + 5.865 * x - 0.5696);
}

// Waveform factor at iteration `it` (Simulation.cpp:875-900 for the PV2013
// polynomial: the first evaluation, at t = 0, skips the period reduction).
__device__ inline double wave_value(const WaveTab& w, int64_t it) {
    if (w.kind == LBM_WAVE_FOURIER) {
        double acc = 0.0;
        for (int k = 0; k < w.nk; ++k) {
            const double arg = 6.283185307179586 * (double)(k + 1) * (double)it / w.period + w.phi[k];
            acc = acc + w.a[k] * sin(arg);
        }
        return 1.0 + acc;
    }
    if (w.kind == LBM_WAVE_PV2013) {
        double x = (double)it * w.dt;
        const double period = w.period;
        if (it != 0) x -= period * floor(x / period);
        const double nx = fmod(x, period);
        if (nx >= 0 && nx < (0.1 * period)) return pv_eq1(nx);
        if (nx >= (0.1 * period) && nx < (0.3 * period)) return pv_eq2(nx);
        if (nx >= (0.3 * period) && nx < (0.9 * period)) return pv_eq3(nx);
        return 0.0;
    }
    return 1.0;
}

// ------------------------------------------------------------ FP64 (reference)

// Simulation.cpp:32-78 / :82-152: rho summed in index order; u from the +/-
// groups in ascending index order, divided by rho.
template <int Q>
__device__ __forceinline__ void moments_f64(const double* f, double& rho, double* u) {
    using L = Lat<Q>;
    double r = f[0];
#pragma unroll
    for (int d = 1; d < Q; ++d) r = r + f[d];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        double p = 0.0, m = 0.0;
        bool fp = true, fm = true;
#pragma unroll
        for (int d = 0; d < Q; ++d) {
            if (L::c(d, a) == 1) { p = fp ? f[d] : p + f[d]; fp = false; }
            if (L::c(d, a) == -1) { m = fm ? f[d] : m + f[d]; fm = false; }
        }
        u[a] = (p - m) / r;
    }
    rho = r;
}

// Simulation.cpp:154-176, written as the reference writes it.
template <int Q>
__device__ __forceinline__ void collide_f64(const double* f, double omega, double* g) {
    using L = Lat<Q>;
    double rho, u[3];
    moments_f64<Q>(f, rho, u);
#pragma unroll
    for (int d = 0; d < Q; ++d) {
        const double du = u[0] * (double)L::c(d, 0) + u[1] * (double)L::c(d, 1) + u[2] * (double)L::c(d, 2);
        const double usq = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
        const double feq = L::w(d) * rho * (1.0 + 3.0 * du + 9.0 / 2.0 * du * du - 3.0 / 2.0 * usq);
        g[d] = omega * feq + (1.0 - omega) * f[d];
    }
}

// Zou-He on the face whose unknown populations have c[AX] == SG, in the
// rounding order of Boundary.cpp:131-147 (velocity) and Simulation.cpp:966-1346
// (pressure); D3Q27 follows Boundary.cpp:101-121.  All index sets are resolved
// at compile time.
template <int Q, int AX, int SG>
__device__ void zouhe_f64(const FaceTab& F, const double* uv, bool typo, double* f) {
    using L = Lat<Q>;
    constexpr int T0 = AX == 0 ? 1 : 0;
    constexpr int T1 = AX == 2 ? 1 : 2;
    double S0 = 0.0, S1 = 0.0;
    bool f0 = true, f1 = true;
#pragma unroll
    for (int d = 0; d < Q; ++d) {
        if (L::c(d, AX) == 0) { S0 = f0 ? f[d] : S0 + f[d]; f0 = false; }
        if (L::c(d, AX) == -SG) { S1 = f1 ? f[d] : S1 + f[d]; f1 = false; }
    }
    double rho, u[3];
    if (F.kind == LBM_BC_VELOCITY) {
        u[0] = uv[0]; u[1] = uv[1]; u[2] = uv[2];
        const double un = SG > 0 ? u[AX] : -u[AX];
        rho = (S0 + 2.0 * S1) / (1 - un);
    } else {
        rho = F.rho;
        const double v = 1 - (1.0 / rho) * (S0 + 2.0 * S1);
        u[0] = 0.0; u[1] = 0.0; u[2] = 0.0;
        u[AX] = SG > 0 ? v : -v;
    }
    double N[3];
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        const int t = k == 0 ? T0 : T1;
        double P = 0.0, M = 0.0;
        bool fp = true, fm = true;
#pragma unroll
        for (int d = 0; d < Q; ++d) {
            if (L::c(d, AX) == 0 && L::c(d, t) == 1) { P = fp ? f[d] : P + f[d]; fp = false; }
            if (L::c(d, AX) == 0 && L::c(d, t) == -1) {
                if (Q == 19) { M = fm ? f[d] : M + f[d]; fm = false; }
            }
        }
        if (Q == 19) {
            if (AX == 0 && SG == 1 && t == 2 && typo && F.kind == LBM_BC_PRESSURE)
                P = f[4] + f[10] + f[14];  // Simulation.cpp:1124-1129
            N[t] = 0.5 * (P - M) - (1.0 / 3.0) * rho * u[t];
        } else {
            double s = P;  // Boundary.cpp:101-105: ((p1 + p2 + p3) - m1 - m2 - m3)
#pragma unroll
            for (int d = 0; d < Q; ++d)
                if (L::c(d, AX) == 0 && L::c(d, t) == -1) s = s - f[d];
            N[t] = (1.0 / 6) * s - (1.0 / 9) * (rho * u[t]);
        }
    }
    double g[Q];
#pragma unroll
    for (int d = 0; d < Q; ++d) g[d] = f[d];
#pragma unroll
    for (int d = 0; d < Q; ++d) {
        if (L::c(d, AX) != SG) continue;
        const int e = L::opp(d);
        double cu = SG > 0 ? u[AX] : -u[AX];
        int nt = 0;
        if (L::c(d, T0) == 1) { cu = cu + u[T0]; ++nt; }
        if (L::c(d, T0) == -1) { cu = cu - u[T0]; ++nt; }
        if (L::c(d, T1) == 1) { cu = cu + u[T1]; ++nt; }
        if (L::c(d, T1) == -1) { cu = cu - u[T1]; ++nt; }
        double coef;
        if (Q == 19) coef = nt == 0 ? 1.0 / 3.0 : 1.0 / 6.0;
        else coef = nt == 0 ? 4 / 9.0 : (nt == 1 ? 1 / 9.0 : 1 / 36.0);
        double v = g[e] + coef * rho * cu;
        if (L::c(d, T0) == 1) v = v - N[T0];
        if (L::c(d, T0) == -1) v = v + N[T0];
        if (L::c(d, T1) == 1) v = v - N[T1];
        if (L::c(d, T1) == -1) v = v + N[T1];
        f[d] = v;
    }
}

// ----------------------------------------------------- FP32 on shifted values
//
// Constants such as the lattice weights and omega are applied as a float pair
// (hi, lo), hi = (float)c, lo = (float)(c - hi), through one fused
// multiply-add: x*c = fma(x, hi, x*lo).  A plain float constant carries a
// relative error of up to 6e-8 that is the same at every node and every step:
// (1/36)f and (1/18)f are both 7.45e-9 too large, so the equilibrium would hold
// (1 + 7.45e-9) times the node's momentum and every collision would add
// omega * 7.45e-9 of it, a bias that compounds over the run (about 1e-5 of
// the velocity after a few thousand steps).  With the pair the constant is
// exact to about 2^-48 and only unbiased per-node rounding remains.
struct CPair {
    float hi, lo;
};
__host__ __device__ constexpr CPair cpair(double c) { return {(float)c, (float)(c - (double)(float)c)}; }
__device__ __forceinline__ float mulc(float x, CPair c) { return __builtin_fmaf(x, c.hi, x * c.lo); }

// BGK on deviations: g - w = (f - w) + omega ((f_eq - w) - (f - w)),
// f_eq - w = w (drho + rho (3 cu + 4.5 cu^2 - 1.5 u^2)).  Populations are taken
// in opposite pairs (c_e = -c_d): the pair sum feeds rho, the pair difference
// feeds j, and the two equilibria share S = drho + rho (4.5 cu^2 - 1.5 u^2) and
// differ only in the sign of A = 3 rho cu.
template <int Q>
__device__ __forceinline__ void collide_dev(float* fd, CPair om) {
    using L = Lat<Q>;
    float r = fd[L::rest], j0 = 0.f, j1 = 0.f, j2 = 0.f;
#pragma unroll
    for (int d = 0; d < Q; ++d) {
        if (d == L::rest || d > L::opp(d)) continue;
        const int e = L::opp(d);
        const float sum = fd[d] + fd[e], dif = fd[d] - fd[e];
        r += sum;
        if (L::c(d, 0)) j0 = L::c(d, 0) > 0 ? j0 + dif : j0 - dif;
        if (L::c(d, 1)) j1 = L::c(d, 1) > 0 ? j1 + dif : j1 - dif;
        if (L::c(d, 2)) j2 = L::c(d, 2) > 0 ? j2 + dif : j2 - dif;
    }
    const float inv = 1.f / (1.f + r);
    const float u[3] = {j0 * inv, j1 * inv, j2 * inv};
    const float rho = 1.f + r;
    const float q[3] = {u[0] * u[0], u[1] * u[1], u[2] * u[2]};
    const float usq15 = 1.5f * (q[0] + q[1] + q[2]);
    const float r3 = 3.f * rho;
    const float a[3] = {r3 * u[0], r3 * u[1], r3 * u[2]};
    {
        const int d = L::rest;
        const float feqd = mulc(__builtin_fmaf(-rho, usq15, r), cpair(L::w(d)));
        const float x = feqd - fd[d];
        fd[d] = __builtin_fmaf(x, om.hi, __builtin_fmaf(x, om.lo, fd[d]));
    }
#pragma unroll
    for (int d = 0; d < Q; ++d) {
        if (d == L::rest || d > L::opp(d)) continue;
        const int e = L::opp(d);
        float A = 0.f, cu = 0.f, cu2 = 0.f;
        int nz = 0, k1 = 0;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            if (!L::c(d, k)) continue;
            A = nz == 0 ? (L::c(d, k) > 0 ? a[k] : -a[k]) : (L::c(d, k) > 0 ? A + a[k] : A - a[k]);
            cu = nz == 0 ? (L::c(d, k) > 0 ? u[k] : -u[k]) : (L::c(d, k) > 0 ? cu + u[k] : cu - u[k]);
            k1 = nz == 0 ? k : k1;
            ++nz;
        }
        cu2 = nz == 1 ? q[k1] : cu * cu;
        const float S = __builtin_fmaf(rho, __builtin_fmaf(4.5f, cu2, -usq15), r);
        const CPair w = cpair(L::w(d));
        const float xd = mulc(S + A, w) - fd[d];
        const float xe = mulc(S - A, w) - fd[e];
        fd[d] = __builtin_fmaf(xd, om.hi, __builtin_fmaf(xd, om.lo, fd[d]));
        fd[e] = __builtin_fmaf(xe, om.hi, __builtin_fmaf(xe, om.lo, fd[e]));
    }
}

// Zou-He on deviations.  The weight parts cancel: the c_axis = 0 set plus twice
// the outgoing set weighs 1, the +/- tangential sets weigh the same, and w_e =
// w_opp(e), so the closure is the FP64 one with sum f -> 1 + sum dev.
template <int Q, int AX, int SG>
__device__ void zouhe_dev(const FaceTab& F, const float* uv, bool typo, float* fd) {
    using L = Lat<Q>;
    constexpr int T0 = AX == 0 ? 1 : 0;
    constexpr int T1 = AX == 2 ? 1 : 2;
    float D = 0.f;
#pragma unroll
    for (int d = 0; d < Q; ++d) {
        if (L::c(d, AX) == 0) D += fd[d];
        if (L::c(d, AX) == -SG) D += 2.f * fd[d];
    }
    float rho, u[3];
    if (F.kind == LBM_BC_VELOCITY) {
        u[0] = uv[0]; u[1] = uv[1]; u[2] = uv[2];
        const float un = SG > 0 ? u[AX] : -u[AX];
        rho = (1.f + D) / (1.f - un);
    } else {
        rho = (float)F.rho;
        const float v = ((float)(F.rho - 1.0) - D) / rho;
        u[0] = 0.f; u[1] = 0.f; u[2] = 0.f;
        u[AX] = SG > 0 ? v : -v;
    }
    float N[3];
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        const int t = k == 0 ? T0 : T1;
        float PM = 0.f;
#pragma unroll
        for (int d = 0; d < Q; ++d) {
            if (L::c(d, AX) == 0 && L::c(d, t) == 1) PM += fd[d];
            if (L::c(d, AX) == 0 && L::c(d, t) == -1) PM -= fd[d];
        }
        if (Q == 19) {
            if (AX == 0 && SG == 1 && t == 2 && typo && F.kind == LBM_BC_PRESSURE)
                PM += fd[10] - fd[16];  // Simulation.cpp:1124-1129
            N[t] = 0.5f * PM - mulc(rho * u[t], cpair(1.0 / 3.0));
        } else {
            N[t] = mulc(PM, cpair(1.0 / 6.0)) - mulc(rho * u[t], cpair(1.0 / 9.0));
        }
    }
    float g[Q];
#pragma unroll
    for (int d = 0; d < Q; ++d) g[d] = fd[d];
#pragma unroll
    for (int d = 0; d < Q; ++d) {
        if (L::c(d, AX) != SG) continue;
        float cu = SG > 0 ? u[AX] : -u[AX];
        int nt = 0;
        if (L::c(d, T0)) { cu += L::c(d, T0) > 0 ? u[T0] : -u[T0]; ++nt; }
        if (L::c(d, T1)) { cu += L::c(d, T1) > 0 ? u[T1] : -u[T1]; ++nt; }
        double coef;
        if (Q == 19) coef = nt == 0 ? 1.0 / 3.0 : 1.0 / 6.0;
        else coef = nt == 0 ? 4.0 / 9.0 : (nt == 1 ? 1.0 / 9.0 : 1.0 / 36.0);
        float v = g[L::opp(d)] + mulc(rho * cu, cpair(coef));
        if (L::c(d, T0)) v += L::c(d, T0) > 0 ? -N[T0] : N[T0];
        if (L::c(d, T1)) v += L::c(d, T1) > 0 ? -N[T1] : N[T1];
        fd[d] = v;
    }
}

// ---------------------------------------------------------------- dispatch

template <int Q, class T>
struct ZouHe;

template <int Q>
struct ZouHe<Q, double> {
    __device__ static void apply(const FaceTab& F, const double* uv, double* f) {
        const bool typo = F.typo != 0;
        switch (F.axis * 2 + (F.sigma > 0)) {
            case 0: zouhe_f64<Q, 0, -1>(F, uv, typo, f); break;
            case 1: zouhe_f64<Q, 0, 1>(F, uv, typo, f); break;
            case 2: zouhe_f64<Q, 1, -1>(F, uv, typo, f); break;
            case 3: zouhe_f64<Q, 1, 1>(F, uv, typo, f); break;
            case 4: zouhe_f64<Q, 2, -1>(F, uv, typo, f); break;
            default: zouhe_f64<Q, 2, 1>(F, uv, typo, f); break;
        }
    }
};

template <int Q>
struct ZouHe<Q, float> {
    __device__ static void apply(const FaceTab& F, const double* uvd, float* f) {
        const bool typo = F.typo != 0;
        const float uv[3] = {(float)uvd[0], (float)uvd[1], (float)uvd[2]};
        switch (F.axis * 2 + (F.sigma > 0)) {
            case 0: zouhe_dev<Q, 0, -1>(F, uv, typo, f); break;
            case 1: zouhe_dev<Q, 0, 1>(F, uv, typo, f); break;
            case 2: zouhe_dev<Q, 1, -1>(F, uv, typo, f); break;
            case 3: zouhe_dev<Q, 1, 1>(F, uv, typo, f); break;
            case 4: zouhe_dev<Q, 2, -1>(F, uv, typo, f); break;
            default: zouhe_dev<Q, 2, 1>(F, uv, typo, f); break;
        }
    }
};

// Element i of population array d (g + d*N).  For 2- and 4-byte storage the
// access is a buffer load/store: the array's base goes into a resource
// descriptor built with scalar instructions (the array is uniform), the lane
// supplies a 32-bit byte offset (node counts stay below 2^30, so offsets fit),
// and the descriptor's size (N elements) makes the hardware drop any access
// outside the array.  No per-lane 64-bit address arithmetic is left.  FP64
// arrays may exceed 4 GiB and use plain global addressing.
template <class S>
__device__ __forceinline__ __amdgpu_buffer_rsrc_t pop_rsrc(const S* g, int d, int64_t N) {
    return __builtin_amdgcn_make_buffer_rsrc((void*)(g + (int64_t)d * N), (short)0,
                                             (int)(uint32_t)((uint64_t)N * sizeof(S)), 0x00020000);
}
// BUF = false: plain global addressing, for the record path whose scattered,
// conditional accesses would each need their own scalar descriptor set-up.
template <class P, bool BUF = true>
__device__ __forceinline__ typename P::calc_t ldp(const typename P::store_t* g, int d, int64_t N, int64_t i) {
    using S = typename P::store_t;
    if constexpr (BUF && sizeof(S) == 4) {
        const uint32_t v = __builtin_amdgcn_raw_buffer_load_b32(pop_rsrc(g, d, N), (int)((uint32_t)i * 4u), 0, 0);
        return (typename P::calc_t)__builtin_bit_cast(float, v);
    } else if constexpr (BUF && sizeof(S) == 2) {
        const unsigned short v = __builtin_amdgcn_raw_buffer_load_b16(pop_rsrc(g, d, N), (int)((uint32_t)i * 2u), 0, 0);
        return (typename P::calc_t)__builtin_bit_cast(_Float16, v);
    } else {
        return (typename P::calc_t)g[(int64_t)d * N + i];
    }
}
template <class S, bool BUF = true>
__device__ __forceinline__ void stp(S* g, int d, int64_t N, int64_t i, S v) {
    if constexpr (BUF && sizeof(S) == 4) {
        __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, v), pop_rsrc(g, d, N), (int)((uint32_t)i * 4u),
                                              0, 0);
    } else if constexpr (BUF && sizeof(S) == 2) {
        __builtin_amdgcn_raw_buffer_store_b16(__builtin_bit_cast(unsigned short, v), pop_rsrc(g, d, N),
                                              (int)((uint32_t)i * 2u), 0, 0);
    } else {
        g[(int64_t)d * N + i] = v;
    }
}

// One buffer descriptor over a whole lattice, all Q population arrays of
// stride N; element e of the lattice is array e / N, node e % N.  Only for
// lattices below 4 GiB (the descriptor's size and the byte offsets are 32-bit).
template <class S>
__device__ __forceinline__ __amdgpu_buffer_rsrc_t lat_rsrc(const void* g, int q, int64_t N) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(g), (short)0,
                                             (int)(uint32_t)((uint64_t)q * (uint64_t)N * sizeof(S)), 0x00020000);
}
template <class S>
__device__ __forceinline__ uint32_t lat_ld(__amdgpu_buffer_rsrc_t r, uint32_t e) {
    if constexpr (sizeof(S) == 4) return __builtin_amdgcn_raw_buffer_load_b32(r, (int)(e * 4u), 0, 0);
    else return __builtin_amdgcn_raw_buffer_load_b16(r, (int)(e * 2u), 0, 0);
}
template <class S>
__device__ __forceinline__ void lat_st(__amdgpu_buffer_rsrc_t r, uint32_t e, S v) {
    if constexpr (sizeof(S) == 4)
        __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, v), r, (int)(e * 4u), 0, 0);
    else
        __builtin_amdgcn_raw_buffer_store_b16(__builtin_bit_cast(unsigned short, v), r, (int)(e * 2u), 0, 0);
}
template <class S, class T>
__device__ __forceinline__ T from_bits(uint32_t b) {
    if constexpr (sizeof(S) == 4) return (T)__builtin_bit_cast(float, b);
    else return (T)__builtin_bit_cast(_Float16, (unsigned short)b);
}

// Conversion of a computed population to the storage type.  FP16 storage
// rounds the FP32 result (the collision's last operation is an FMA rounded to
// FP32) to FP16.  The compiler would otherwise fuse that FMA with the
// conversion into one mixed-precision FMA that rounds once, straight to FP16,
// in some kernels and not in others, and kernels that must agree bit for bit
// would differ in the last place: the empty register constraint keeps the FP32
// value.
template <class S, class T>
__device__ __forceinline__ S to_store(T v) {
    if constexpr (sizeof(S) == 2) {
        float x = (float)v;
        asm volatile("" : "+v"(x));
        return (S)x;
    } else {
        return (S)v;
    }
}

// Streaming schemes.
//   SCH_TWO:       pull from g_in, write g_out (two lattices).
//   SCH_AA_GATHER: one lattice A; read A[opp d][x - c_d], write A[d][x + c_d]
//                  (a missing destination keeps the value at A[opp d][x], where
//                  the next local step reads it as the bounced-back population).
//   SCH_AA_LOCAL:  read A[d][x], write A[opp d][x].
// Every slot of A is read and written by one node only, so both AA steps run
// in place without races.
enum { SCH_TWO = 0, SCH_AA_GATHER = 1, SCH_AA_LOCAL = 2 };

template <class P, int Q>
__device__ __forceinline__ void write_pre(const StepArgs& a, int64_t n, const typename P::calc_t* f) {
    using L = Lat<Q>;
    using T = typename P::calc_t;
    constexpr bool F64 = sizeof(T) == 8;
    const int64_t N = a.n;
    double fd[Q];
#pragma unroll
    for (int d = 0; d < Q; ++d) fd[d] = F64 ? (double)f[d] : L::w(d) + (double)f[d];
#pragma unroll
    for (int d = 0; d < Q; ++d) a.fpre[(int64_t)d * N + n] = fd[d];
    double rho, u[3];
    if constexpr (F64) {
        moments_f64<Q>(fd, rho, u);
    } else {
        double dr = 0.0, j[3] = {0.0, 0.0, 0.0};
#pragma unroll
        for (int d = 0; d < Q; ++d) {
            const double v = (double)f[d];
            dr += v;
            j[0] += L::c(d, 0) * v; j[1] += L::c(d, 1) * v; j[2] += L::c(d, 2) * v;
        }
        rho = 1.0 + dr;
        u[0] = j[0] / rho; u[1] = j[1] / rho; u[2] = j[2] / rho;
    }
    a.rho[n] = rho;
    a.u[n] = u[0]; a.u[N + n] = u[1]; a.u[2 * N + n] = u[2];
}

// ------------------------------------------------------------------- kernel

template <int Q, class P>
__device__ __forceinline__ void apply_bc(const StepArgs& a, int type, int64_t r, typename P::calc_t* f);

// Bounce-back, stale populations and Zou-He of boundary record r on the
// gathered populations f (miss: links whose source node does not exist).
template <int Q, class P, int SCH>
__device__ __forceinline__ void apply_record(const StepArgs& a, int64_t r, int64_t n, uint32_t miss,
                                             typename P::calc_t* f) {
    using L = Lat<Q>;
    using T = typename P::calc_t;
    const typename P::store_t* gin = (const typename P::store_t*)a.gin;
    const uint32_t bb = a.bbb[r];
    const int type = a.btype[r];
    if (SCH == SCH_AA_LOCAL) miss = a.bmiss[r];
    // bounce-back: this node's own post-collision g[opp d] (SCH_AA_LOCAL:
    // A[d][x] already holds it, kept there by the gather step)
    if (SCH != SCH_AA_LOCAL && (miss & bb)) {
#pragma unroll
        for (int d = 0; d < Q; ++d) {
            if (!((miss & bb) >> d & 1u)) continue;
            if (SCH == SCH_TWO) f[d] = ldp<P>(gin, L::opp(d), a.ns, n);
            if (SCH == SCH_AA_GATHER) f[d] = ldp<P>(gin, d, a.ns, n);
        }
    }
    // unlinked populations the reference never rewrites keep their value
    if (miss & ~bb) {
        const typename P::store_t* stale = (const typename P::store_t*)a.bstale;
#pragma unroll
        for (int d = 0; d < Q; ++d)
            if ((miss & ~bb) >> d & 1u) f[d] = (T)stale[(int64_t)d * a.nrec + r];
    }
    apply_bc<Q, P>(a, type, r, f);
}

// Zou-He on the lanes whose record r carries a boundary condition (type >= 3,
// condition type - 3).  The callers load type next to their other
// per-record words, ahead of the population loads.  The condition is made
// wave-uniform one condition at a time (a wavefront's boundary lanes nearly
// always share one): its table entry and waveform value are then uniform
// loads and the closure runs for one face orientation.  The inlet profile
// factor is read here, by the boundary lanes only.
template <int Q, class P>
__device__ __forceinline__ void apply_bc(const StepArgs& a, int type, int64_t r, typename P::calc_t* f) {
    using T = typename P::calc_t;
    uint64_t todo = __builtin_amdgcn_ballot_w64(type >= 3);
    while (todo) {
        const int t0 = __builtin_amdgcn_readlane(type, (int)__builtin_ctzll(todo));
        const bool mine = type == t0;
        todo &= ~__builtin_amdgcn_ballot_w64(mine);
        if (mine) {
            const FaceTab& F = a.faces[t0 - 3];
            double uv[3] = {0.0, 0.0, 0.0};
            if (F.kind == LBM_BC_VELOCITY) {
                const double w = F.waveform >= 0 ? a.wvals[F.waveform * LBM_WCHUNK + a.wk] : 1.0;
                const double fac = a.bprof[r] * w;
                uv[0] = F.u[0] * fac; uv[1] = F.u[1] * fac; uv[2] = F.u[2] * fac;
            }
            ZouHe<Q, T>::apply(F, uv, f);
        }
    }
}

// Pre-state output (modes PRE / SNAP), collision and store of node n.
// dst(d) (AA gather only): destination node of population d, -1 if none;
// evaluated at the store so no per-direction index stays live in registers.
// ALLDST: every destination exists (segments, record-free chunks), so the
// stores need no per-lane test.  LAT: each store picks its element of the
// whole lattice (32-bit, one descriptor), no per-lane branch.
template <int Q, class P, int MODE, int SCH, bool BUF = true, bool ALLDST = false, bool LAT = false, class Dst>
__device__ __forceinline__ void emit(const StepArgs& a, int64_t n, typename P::calc_t* f, Dst dst) {
    using L = Lat<Q>;
    using S = typename P::store_t;
    using T = typename P::calc_t;
    if (MODE == LBM_MODE_PRE || MODE == LBM_MODE_SNAP) write_pre<P, Q>(a, n, f);
    if (MODE == LBM_MODE_PRE) return;
    S* gout = (S*)a.gout;
    T g[Q];
    if constexpr (sizeof(T) == 8) {
        collide_f64<Q>((const double*)f, a.omega, (double*)g);
    } else {
#pragma unroll
        for (int d = 0; d < Q; ++d) g[d] = f[d];
        collide_dev<Q>((float*)g, cpair(a.omega));
    }
#pragma unroll
    for (int d = 0; d < Q; ++d) {
        if (SCH == SCH_TWO) {
            stp<S, BUF>(gout, d, a.ns, n, to_store<S>(g[d]));
        } else if (SCH == SCH_AA_LOCAL) {
            stp<S, BUF>(gout, L::opp(d), a.ns, n, to_store<S>(g[d]));
        } else {
            const int32_t t = dst(d);
            if constexpr (LAT) {
                const uint32_t nS = (uint32_t)a.ns;
                const uint32_t e = t >= 0 ? (uint32_t)d * nS + (uint32_t)t : (uint32_t)L::opp(d) * nS + (uint32_t)n;
                lat_st<S>(lat_rsrc<S>(a.gout, Q, a.ns), e, to_store<S>(g[d]));
            } else if constexpr (ALLDST) {
                stp<S, BUF>(gout, d, a.ns, t, to_store<S>(g[d]));
            } else if constexpr (BUF) {
                if (t >= 0) stp<S, BUF>(gout, d, a.ns, t, to_store<S>(g[d]));
                else stp<S, BUF>(gout, L::opp(d), a.ns, n, to_store<S>(g[d]));
            } else {
                // one store to a per-lane selected slot
                S* q = t >= 0 ? gout + (int64_t)d * a.ns + t : gout + (int64_t)L::opp(d) * a.ns + n;
                *q = to_store<S>(g[d]);
            }
        }
    }
}

// Generic kernel: one thread per node, sources found through the padded site
// map, records through the boundary bitmap.  Wavefronts marked in `wreg` (64
// interior nodes of one contiguous row) take scalar per-direction bases.  This
// path shares no addressing with the record/segment kernel below, and the
// parity tests run both.
template <int Q, class P, int MODE, int SCH>
__global__ __launch_bounds__(512) void k_step(StepArgs a) {
    using L = Lat<Q>;
    using S = typename P::store_t;
    using T = typename P::calc_t;
    const int64_t lb = logical_block(a.rev, a.tile);
    const int64_t n = lb * blockDim.x + threadIdx.x;
    if (lb >= a.nblocks || n >= a.n) return;
    // gin and gout are the same buffer in the AA scheme, where a thread
    // reads and then writes the same slots: no restrict qualifiers.
    const S* gin = (const S*)a.gin;
    const int64_t N = a.n;
    T f[Q];
    int32_t dst[Q];
    // Wavefront index, made wave-uniform so the loads below that depend only
    // on it are scalar loads.
    const int wv = __builtin_amdgcn_readfirstlane((int)(n >> 6));
    bool reg = a.wreg && ((a.wreg[wv >> 5] >> (wv & 31)) & 1u);
    int32_t b_src[Q], b_dst[Q];
    if (reg) {
        // All 64 lanes are interior nodes of one contiguous row, so for every
        // direction their sources (and destinations) are one contiguous run
        // of node ids.  The wave-uniform bases are checked to lie in
        // [0, n - 64] before any lane uses them.
        const int32_t p0 = a.pnode[(int64_t)wv << 6];
        bool ok = ((int64_t)wv << 6) + 63 < N;
#pragma unroll
        for (int d = 0; d < Q; ++d) {
            b_src[d] = (d == L::rest || SCH == SCH_AA_LOCAL) ? (wv << 6) : a.map[p0 - a.off[d]];
            b_dst[d] = d == L::rest ? (wv << 6) : (SCH == SCH_AA_GATHER ? a.map[p0 + a.off[d]] : 0);
            ok = ok && b_src[d] >= 0 && (int64_t)b_src[d] + 63 < N && b_dst[d] >= 0 && (int64_t)b_dst[d] + 63 < N;
        }
        reg = ok;
    }
    if (reg) {
        const int lane = (int)(n & 63);
#pragma unroll
        for (int d = 0; d < Q; ++d) {
            if (SCH == SCH_AA_LOCAL || d == L::rest) {
                f[d] = ldp<P>(gin, d, a.ns, n);
                dst[d] = (int32_t)n;
            } else {
                const int64_t src = (int64_t)b_src[d] + lane;
                f[d] = ldp<P>(gin, (SCH == SCH_AA_GATHER ? L::opp(d) : d), a.ns, src);
                if (SCH == SCH_AA_GATHER) dst[d] = b_dst[d] + lane;
            }
        }
        if (SCH == SCH_AA_GATHER) dst[L::rest] = (int32_t)n;
    } else {
        const uint64_t word = a.bbits[n >> 6];
        const int bit = (int)(n & 63);
        uint32_t miss = 0;
        const int32_t p = SCH != SCH_AA_LOCAL ? a.pnode[n] : 0;
#pragma unroll
        for (int d = 0; d < Q; ++d) {
            if (SCH == SCH_AA_LOCAL || d == L::rest) {
                f[d] = ldp<P>(gin, d, a.ns, n);
            } else {
                const int32_t src = a.map[p - a.off[d]];
                if (src >= 0) f[d] = ldp<P>(gin, (SCH == SCH_AA_GATHER ? L::opp(d) : d), a.ns, src);
                else { f[d] = T(0); miss |= 1u << d; }
                if (SCH == SCH_AA_GATHER) dst[d] = a.map[p + a.off[d]];
            }
        }
        if (SCH == SCH_AA_GATHER) dst[L::rest] = (int32_t)n;
        if ((word >> bit) & 1ull) {
            const int64_t r = a.bpre[n >> 6] + __popcll(word & ((1ull << bit) - 1ull));
            apply_record<Q, P, SCH>(a, r, n, miss, f);
        }
    }
    emit<Q, P, MODE, SCH>(a, n, f, [&](int d) { return dst[d]; });
}

// Record/segment kernel.  Records (nodes with a missing link or a boundary
// condition) take one thread each, their sources from the per-record table
// rsrc[d][r] (coalesced; no map lookups).  Every other node belongs to a row
// segment: at most 64 consecutive node ids along the fast axis whose sources
// in each direction are again consecutive ids, so one wavefront reads each
// direction as one contiguous run from a base in the segment table.
//
// One launch (PART_FUSED): each 64-lane wavefront takes one row chunk, records
// and interior nodes alike, in node order; blocks map to XCDs in contiguous
// ranges (logical_block).  Two launches (PART_RECORDS, then PART_SEGMENTS):
// per XCD (block id mod 8) a contiguous share of record blocks or of segment
// blocks (one segment per wavefront).
enum { PART_FUSED = 0, PART_RECORDS = 1, PART_SEGMENTS = 2 };

// One boundary record per lane.  Every population has one source, chosen per
// lane: the neighbour x - c_d when it exists, else this node's own slot
// (bounce-back) or the stale value (unlinked populations the reference never
// rewrites); one load follows the choice.  Zou-He runs on boundary-condition
// records only.
template <int Q, class P, int MODE, int SCH>
__device__ __forceinline__ void record_node(const StepArgs& a, int64_t r) {
    using L = Lat<Q>;
    using S = typename P::store_t;
    using T = typename P::calc_t;
    const S* gin = (const S*)a.gin;
    const S* stale = (const S*)a.bstale;
    const int64_t NS = a.ns, R = a.nrec;
    const int64_t n = a.rec_node[r];
    const uint32_t bb = a.bbb[r];
    const uint32_t lmiss = SCH == SCH_AA_LOCAL ? a.bmiss[r] : 0u;
    const int type = a.btype[r];
    T f[Q];
#pragma unroll
    for (int d = 0; d < Q; ++d) {
        const S* st = stale + (int64_t)d * R + r;
        const S* p;
        if (SCH == SCH_AA_LOCAL || d == L::rest) {
            // AA local: a bounced-back population already sits in A[d][x]
            p = ((lmiss & ~bb) >> d & 1u) ? st : gin + (int64_t)d * NS + n;
        } else {
            const int32_t sd = a.rsrc[(int64_t)d * R + r];
            const S* nb = gin + (int64_t)(SCH == SCH_AA_GATHER ? L::opp(d) : d) * NS + sd;
            const S* own = gin + (int64_t)(SCH == SCH_TWO ? L::opp(d) : d) * NS + n;
            p = sd >= 0 ? nb : ((bb >> d & 1u) ? own : st);
        }
        f[d] = (T)*p;
    }
    apply_bc<Q, P>(a, type, r, f);
    // AA gather: the node at x + c_d is the source of direction opp d
    emit<Q, P, MODE, SCH, false>(a, n, f, [&](int d) {
        return d == L::rest ? (int32_t)n : a.rsrc[(int64_t)L::opp(d) * R + r];
    });
}

template <int Q, class P, int MODE, int SCH>
__device__ __forceinline__ void segment_node(const StepArgs& a, int64_t s, int lane) {
    using L = Lat<Q>;
    using S = typename P::store_t;
    using T = typename P::calc_t;
    const int32_t* row = a.seg + s * LBM_SEG_W;  // wave-uniform
    const int32_t start = row[0], len = row[1];
    if (lane >= len) return;
    const S* gin = (const S*)a.gin;
    const int64_t n = (int64_t)start + lane;
    T f[Q];
#pragma unroll
    for (int d = 0; d < Q; ++d) {
        if (SCH == SCH_AA_LOCAL || d == L::rest) {
            f[d] = ldp<P>(gin, d, a.ns, n);
        } else {
            f[d] = ldp<P>(gin, (SCH == SCH_AA_GATHER ? L::opp(d) : d), a.ns, row[2 + d] + lane);
        }
    }
    emit<Q, P, MODE, SCH, true, true>(a, n, f,
                                      [&](int d) { return d == L::rest ? (int32_t)n : row[2 + L::opp(d)] + lane; });
}

// One node of a row chunk (lane l = node start + l).  A chunk without records
// (the common case, a wave-uniform test) reads every direction as one
// contiguous run base_d + l, exactly like a segment.  A chunk with records
// picks, per lane and direction, the neighbour base_d + l on the lanes of
// [lo_d, hi_d), else the node's own slot (bounce-back) or its stale value, and
// issues one load; record lanes find their record index from the chunk's
// record mask (records are numbered in node order, so a chunk may span two
// 64-node words).
//
// LAT (16-bit storage, lattice below 4 GiB): in a chunk with records each lane
// picks, per direction, one 32-bit element index into the whole lattice (its
// neighbour or its own slot) for one buffer load; stale values take a second
// load only in chunks that have such lanes.  Otherwise the choice is among
// 64-bit addresses, which the compiler turns into per-lane branches.
//
// TWO: the wavefront takes two consecutive chunks of one aligned 64-node window
// (piece A on lanes [0, lenA), piece B on the lanes after it), so a window cut
// by a row end still fills its lanes.  The two chunk rows are copied into the
// wavefront's slot of LDS (srow, 2 CW ints) first, and each lane reads its
// piece's base and interval per direction from there where it uses them,
// indexed by its lane within the piece; the two record masks join into one over
// the wavefront (piece B's records follow piece A's in node order).  The rows
// stay out of scalar and vector registers, so the pair path does not raise the
// register count of the single-chunk path it shares a kernel with.  Pairs serve
// the two-lattice pull only: its stores go to the wavefront's own nodes and
// need no per-piece base.
template <int Q, class P, int MODE, int SCH, bool LAT = false, bool TWO = false>
__device__ __forceinline__ void chunk_node(const StepArgs& a, int64_t c, int lane, int32_t* srow = nullptr) {
    using L = Lat<Q>;
    using S = typename P::store_t;
    using T = typename P::calc_t;
    constexpr int CW = lbm_chunk_w(Q);
    static_assert(2 * CW > 64 && 2 * CW <= 128, "two chunk rows take two LDS words per lane at most");
    static_assert(!TWO || SCH == SCH_TWO, "chunk pairs serve the two-lattice pull, whose stores need no bases");
    const int32_t* h = a.chunk + c * CW;  // wave-uniform; piece B's row follows at h + CW
    if constexpr (TWO) {
        // every lane is still active here
        srow[lane] = h[lane];
        if (lane + 64 < 2 * CW) srow[lane + 64] = h[lane + 64];
        __builtin_amdgcn_wave_barrier();
    }
    const int32_t start = h[0], lenA = h[1];
    const int32_t len = TWO ? lenA + h[CW + 1] : lenA;
    if (lane >= len) return;
    const S* gin = (const S*)a.gin;
    const int64_t NS = a.ns;
    const int64_t n = (int64_t)start + lane;
    uint32_t mlo = (uint32_t)h[4], mhi = (uint32_t)h[5];
    const bool pb = TWO && lane >= lenA;     // lane of piece B
    const int li = pb ? lane - lenA : lane;  // lane within its piece
    const int32_t* lr = TWO ? srow + (pb ? CW : 0) : nullptr;  // this lane's piece row (LDS)
    if constexpr (TWO) {
        // lenA <= 63: piece B holds at least one node of the window
        const uint64_t mb = (uint64_t)(uint32_t)h[CW + 5] << 32 | (uint32_t)h[CW + 4];
        const uint64_t m = ((uint64_t)mhi << 32 | mlo) | (mb << lenA);
        mlo = (uint32_t)m;
        mhi = (uint32_t)(m >> 32);
    }
    // One chunk: the whole row, read once into scalar registers before any
    // store (a table read after a store could not be proved unclobbered and
    // would become a per-lane vector load with a wait of its own).
    int32_t hb[Q], hl[Q];
    if constexpr (!TWO) {
#pragma unroll
        for (int d = 0; d < Q; ++d) {
            hb[d] = h[6 + d];
            hl[d] = (int32_t)(((uint32_t)h[6 + Q + (d >> 1)] >> (16 * (d & 1))) & 0xffffu);
        }
    }
    // base and lo | hi << 8 of direction d for this lane's piece
    auto base = [&](int d) -> int32_t {
        if constexpr (TWO) return lr[6 + d];
        else return hb[d];
    };
    auto ivl = [&](int d) -> uint32_t {
        if constexpr (TWO) return ((uint32_t)lr[6 + Q + (d >> 1)] >> (16 * (d & 1))) & 0xffffu;
        else return (uint32_t)hl[d];
    };
    T f[Q];
    if (!TWO && (mlo | mhi) == 0u) {
#pragma unroll
        for (int d = 0; d < Q; ++d) {
            if (SCH == SCH_AA_LOCAL || d == L::rest) f[d] = ldp<P>(gin, d, NS, n);
            else f[d] = ldp<P>(gin, SCH == SCH_AA_GATHER ? L::opp(d) : d, NS, (int64_t)hb[d] + lane);
        }
        emit<Q, P, MODE, SCH, true, true>(a, n, f,
                                          [&](int d) { return d == L::rest ? (int32_t)n : hb[L::opp(d)] + lane; });
        return;
    }
    const bool isrec = lane < 32 ? (mlo >> lane) & 1u : (mhi >> (lane - 32)) & 1u;
    const int64_t r = (int64_t)h[2] + __builtin_amdgcn_mbcnt_hi(mhi, __builtin_amdgcn_mbcnt_lo(mlo, 0u));
    const uint32_t bb = isrec ? a.bbb[r] : 0u;
    const uint32_t lmiss = (SCH == SCH_AA_LOCAL && isrec) ? a.bmiss[r] : 0u;
    const int type = isrec ? (int)a.btype[r] : 0;
    const S* stale = (const S*)a.bstale;
    uint32_t pm = 0u;  // directions whose source node exists for this lane
    if constexpr (LAT) {
        static_assert(sizeof(S) != 8 && SCH != SCH_AA_LOCAL, "lattice descriptor path: 16/32-bit, pull or gather");
        const uint32_t nS = (uint32_t)NS;
        const __amdgpu_buffer_rsrc_t lat = lat_rsrc<S>(gin, Q, NS);
#pragma unroll
        for (int d = 0; d < Q; ++d)
            if (d != L::rest) {
                // lo <= li < hi as one unsigned compare
                const uint32_t w = ivl(d), lo = w & 0xffu, hi = (w >> 8) & 0xffu;
                pm |= (uint32_t)((uint32_t)li - lo < hi - lo) << d;
            }
        constexpr uint32_t nonrest = (uint32_t)((1ull << Q) - 1ull) & ~(1u << L::rest);
        // Whether any lane keeps a stale population comes from the chunk row,
        // so the loads of the common chunks do not wait for the record words.
        const bool anyst = (h[3] | (TWO ? h[CW + 3] : 0)) != 0;  // wave-uniform
        auto gather = [&](auto with_stale) {
            const uint32_t stn = isrec ? (~pm & ~bb & nonrest) : 0u;  // links that take the stale value
#pragma unroll
            for (int d = 0; d < Q; ++d) {
                uint32_t e;
                if (d == L::rest) {
                    e = (uint32_t)d * nS + (uint32_t)n;
                } else {
                    const uint32_t enb =
                        (uint32_t)(SCH == SCH_AA_GATHER ? L::opp(d) : d) * nS + (uint32_t)(base(d) + li);
                    const uint32_t eown = (uint32_t)(SCH == SCH_TWO ? L::opp(d) : d) * nS + (uint32_t)n;
                    e = (pm >> d & 1u) ? enb : eown;
                }
                uint32_t bits = lat_ld<S>(lat, e);
                if constexpr (decltype(with_stale)::value) {
                    if (d != L::rest) {
                        const bool us = stn >> d & 1u;
                        const uint32_t sv = lat_ld<S>(pop_rsrc(stale, d, a.nrec), us ? (uint32_t)r : 0u);
                        bits = us ? sv : bits;
                    }
                }
                f[d] = from_bits<S, T>(bits);
            }
        };
        if (anyst) gather(std::true_type{});
        else gather(std::false_type{});
        apply_bc<Q, P>(a, type, r, f);
        emit<Q, P, MODE, SCH, true, false, true>(a, n, f, [&](int d) {
            if (d == L::rest) return (int32_t)n;
            return (pm >> L::opp(d)) & 1u ? base(L::opp(d)) + li : -1;
        });
        return;
    }
    // as above: in the pull a chunk whose lanes keep no stale population (the
    // chunk row says so) picks between neighbour and own slot without the
    // record words.  The AA gather keeps the one three-way choice: split in two
    // it ran 0.3% slower on the benchmark case (FP32).
    const bool anyst = SCH != SCH_TWO || (h[3] | (TWO ? h[CW + 3] : 0)) != 0;  // wave-uniform
    auto gather = [&](auto with_stale) {
#pragma unroll
        for (int d = 0; d < Q; ++d) {
            const S* st = stale + (int64_t)d * a.nrec + r;
            const S* p;
            if (SCH == SCH_AA_LOCAL || d == L::rest) {
                p = ((lmiss & ~bb) >> d & 1u) ? st : gin + (int64_t)d * NS + n;
            } else {
                const uint32_t w = ivl(d);
                const bool pres = (uint32_t)li >= (w & 0xffu) && (uint32_t)li < ((w >> 8) & 0xffu);
                pm |= (uint32_t)pres << d;
                const S* nb = gin + (int64_t)(SCH == SCH_AA_GATHER ? L::opp(d) : d) * NS + ((int64_t)base(d) + li);
                const S* own = gin + (int64_t)(SCH == SCH_TWO ? L::opp(d) : d) * NS + n;
                if constexpr (decltype(with_stale)::value) p = pres ? nb : ((bb >> d & 1u) ? own : st);
                else p = pres ? nb : own;
            }
            f[d] = (T)*p;
        }
    };
    if (anyst) gather(std::true_type{});
    else gather(std::false_type{});
    apply_bc<Q, P>(a, type, r, f);
    // AA gather: the node at x + c_d is the source of direction opp d
    emit<Q, P, MODE, SCH, false>(a, n, f, [&](int d) {
        if (d == L::rest) return (int32_t)n;
        return (pm >> L::opp(d)) & 1u ? base(L::opp(d)) + li : -1;
    });
}

// D3Q27 with 16-bit storage: at most 96 VGPRs, five wavefronts per SIMD instead
// of four (the compiler's own choice, 102 VGPRs; the fit needs no scratch).
// FP32 and FP64 keep the default, whose extra registers they use.
template <int Q, class P, int MODE, int SCH, int PART, bool LAT = false>
__global__ __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(
    Q == 27 && sizeof(typename P::store_t) == 2 ? 5 : 1))) void k_step_rs(StepArgs a) {
    if constexpr (PART == PART_FUSED) {
        const int64_t lb = logical_block(a.rev, a.tile);
        if (lb >= a.nblocks) return;
        const int64_t j = __builtin_amdgcn_readfirstlane((int)(lb * (blockDim.x >> 6) + (threadIdx.x >> 6)));
        const int lane = (int)(threadIdx.x & 63);
        if constexpr (SCH == SCH_TWO) {
            // Two-lattice pull (stores to the wavefront's own aligned nodes):
            // wavefronts through the wave table, chunk pairs or single chunks.
            // An AA gather's pair lanes would have to shift their stores as
            // well as their loads, which measured slower than one chunk each.
            if (a.wave) {
                // LDS slot per wavefront for the rows of a chunk pair
                __shared__ int32_t srow[8][2 * lbm_chunk_w(Q)];
                if (j < a.nwave) {
                    const int32_t e = a.wave[j];
                    const int64_t c = e & (LBM_WAVE_PAIR - 1);
                    if (e & LBM_WAVE_PAIR) chunk_node<Q, P, MODE, SCH, LAT, true>(a, c, lane, srow[threadIdx.x >> 6]);
                    else chunk_node<Q, P, MODE, SCH, LAT, false>(a, c, lane);
                }
                return;
            }
        }
        if (j < a.nchunk) {
            chunk_node<Q, P, MODE, SCH, LAT>(a, j, lane);
        }
        return;
    }
    const int xcd = blockIdx.x & 7, j = blockIdx.x >> 3;
    const int rb = PART == PART_SEGMENTS ? 0 : a.rb8;
    if (PART != PART_SEGMENTS && j < rb) {
        const int64_t r = ((int64_t)xcd * rb + j) * blockDim.x + threadIdx.x;
        if (r < a.nrec) record_node<Q, P, MODE, SCH>(a, r);
    } else if (PART != PART_RECORDS) {
        const int64_t lb = (int64_t)xcd * a.sb8 + (j - rb);
        const int64_t s = __builtin_amdgcn_readfirstlane((int)(lb * (blockDim.x >> 6) + (threadIdx.x >> 6)));
        if (s < a.nseg) segment_node<Q, P, MODE, SCH>(a, s, (int)(threadIdx.x & 63));
    }
}

// Initial lattice from a given pre-collision state (or f = w when fpre is
// null): collide once and store in the layout the first step reads (swapped
// for AA, whose first step is a gather).
template <int Q, class P, bool SWAP>
__global__ __launch_bounds__(256) void k_init(InitArgs a) {
    using L = Lat<Q>;
    using S = typename P::store_t;
    using T = typename P::calc_t;
    const int64_t lb = logical_block();
    const int64_t n = lb * blockDim.x + threadIdx.x;
    if (lb >= a.nblocks || n >= a.n) return;
    S* gout = (S*)a.gout;
    T g[Q];
    if constexpr (sizeof(T) == 8) {
        double f[Q];
#pragma unroll
        for (int d = 0; d < Q; ++d) f[d] = a.fpre ? a.fpre[(int64_t)d * a.n + n] : L::w(d);
        collide_f64<Q>(f, a.omega, (double*)g);
    } else {
#pragma unroll
        for (int d = 0; d < Q; ++d) g[d] = a.fpre ? (float)(a.fpre[(int64_t)d * a.n + n] - L::w(d)) : 0.f;
        collide_dev<Q>((float*)g, cpair(a.omega));
    }
#pragma unroll
    for (int d = 0; d < Q; ++d) gout[(int64_t)(SWAP ? L::opp(d) : d) * a.ns + n] = to_store<S>(g[d]);
}

// Every pointer a launch dereferences must be set; a missing one is an error
// returned to the caller, never a launch.
inline bool step_args_ok(int mode, int sch, int layout, const StepArgs& a) {
    if (!a.gin || !a.faces || !a.wvals || a.wk < 0 || a.wk >= LBM_WCHUNK || a.ns < a.n || (a.ns & 63)) return false;
    if (a.block != 64 && a.block != 128 && a.block != 256 && a.block != 512) return false;
    if (layout == LBM_LAYOUT_GENERIC && (!a.pnode || !a.map || !a.bbits || !a.bpre)) return false;
    if (layout == LBM_LAYOUT_SPLIT && ((a.nseg > 0 && !a.seg) || (a.nrec > 0 && (!a.rsrc || !a.rec_node))))
        return false;
    if (layout == LBM_LAYOUT_SPLIT && a.nseg + a.nrec <= 0) return false;
    if (layout == LBM_LAYOUT_FUSED && (!a.chunk || a.nchunk <= 0 || (a.wave && a.nwave <= 0))) return false;
    // the wave table (chunk pairs) serves the two-lattice pull only
    if (a.wave && sch != SCH_TWO) return false;
    // the fused layout runs the AA local step on the generic kernel
    if (layout == LBM_LAYOUT_FUSED && sch == SCH_AA_LOCAL && (!a.bbits || !a.bpre)) return false;
    if (a.nrec > 0 && (!a.btype || !a.bbb || !a.bprof || !a.bstale || !a.bmiss)) return false;
    if (mode != LBM_MODE_PRE && !a.gout) return false;
    if (mode != LBM_MODE_STEP && (!a.fpre || !a.rho || !a.u)) return false;
    if (mode == LBM_MODE_PRE && sch != SCH_TWO) return false;
    return sch >= SCH_TWO && sch <= SCH_AA_LOCAL && layout >= LBM_LAYOUT_GENERIC && layout <= LBM_LAYOUT_SPLIT;
}

// Wavefronts of a fused-layout launch: one per wave-table entry (two-lattice
// pull with chunk pairs) or one per chunk.
inline int64_t fused_waves(const StepArgs& a, int sch) { return a.wave && sch == SCH_TWO ? a.nwave : a.nchunk; }

template <int Q, class P, int MODE, int SCH>
void launch_layout(int layout, const StepArgs& b, hipStream_t s) {
    if (layout == LBM_LAYOUT_GENERIC) {
        hipLaunchKernelGGL((k_step<Q, P, MODE, SCH>), dim3(lbm_grid_blocks(b.n, b.block)), dim3(b.block), 0, s, b);
    } else if (layout == LBM_LAYOUT_FUSED) {
        if constexpr (SCH == SCH_AA_LOCAL) {
            // The AA local step touches only each node's own slots: one dense
            // launch in node order, fully coalesced, record lanes fix up in place.
            StepArgs c = b;
            c.wreg = nullptr;
            c.nblocks = (int)((b.n + b.block - 1) / b.block);
            hipLaunchKernelGGL((k_step<Q, P, MODE, SCH>), dim3(lbm_grid_blocks(b.n, b.block)), dim3(b.block), 0, s, c);
        } else if constexpr (sizeof(typename P::store_t) == 2 ||
                             (sizeof(typename P::store_t) == 4 && SCH == SCH_TWO)) {
            // 16-bit storage, and the FP32 pull: chunks with records through
            // one lattice-wide descriptor while the lattice is below 4 GiB
            if (b.wide)
                hipLaunchKernelGGL((k_step_rs<Q, P, MODE, SCH, PART_FUSED, false>),
                                   dim3(lbm_grid_blocks(fused_waves(b, SCH), b.block / 64)), dim3(b.block), 0, s, b);
            else
                hipLaunchKernelGGL((k_step_rs<Q, P, MODE, SCH, PART_FUSED, true>),
                                   dim3(lbm_grid_blocks(fused_waves(b, SCH), b.block / 64)), dim3(b.block), 0, s, b);
        } else {
            hipLaunchKernelGGL((k_step_rs<Q, P, MODE, SCH, PART_FUSED>), dim3(lbm_grid_blocks(fused_waves(b, SCH), b.block / 64)),
                               dim3(b.block), 0, s, b);
        }
    } else {
        // records first, then segments: the two launches write disjoint nodes
        if (b.rb8 > 0)
            hipLaunchKernelGGL((k_step_rs<Q, P, MODE, SCH, PART_RECORDS>), dim3(8 * b.rb8), dim3(b.block), 0, s, b);
        if (b.sb8 > 0)
            hipLaunchKernelGGL((k_step_rs<Q, P, MODE, SCH, PART_SEGMENTS>), dim3(8 * b.sb8), dim3(b.block), 0, s, b);
    }
}

template <int Q, class P, int MODE>
void launch_sch(int sch, int layout, const StepArgs& b, hipStream_t s) {
    if (sch == SCH_TWO) launch_layout<Q, P, MODE, SCH_TWO>(layout, b, s);
    else if (sch == SCH_AA_GATHER) launch_layout<Q, P, MODE, SCH_AA_GATHER>(layout, b, s);
    else launch_layout<Q, P, MODE, SCH_AA_LOCAL>(layout, b, s);
}

// layout: LBM_LAYOUT_GENERIC (one per-node launch through the site map),
// LBM_LAYOUT_FUSED (records and segments in one launch) or LBM_LAYOUT_SPLIT
// (the same work in two launches).
template <int Q, class P>
hipError_t launch_step_t(int mode, int sch, int layout, const StepArgs& a, hipStream_t s) {
    if (!step_args_ok(mode, sch, layout, a)) return hipErrorInvalidValue;
    StepArgs b = a;
    const int wpb = a.block / 64;  // wavefronts per block
    const int64_t fw = fused_waves(a, sch);  // wavefronts of the fused launch
    b.nblocks = (int)(layout == LBM_LAYOUT_FUSED ? (fw + wpb - 1) / wpb : (a.n + a.block - 1) / a.block);
    b.rb8 = (int)(((a.nrec + a.block - 1) / a.block + 7) / 8);
    b.sb8 = (int)(((a.nseg + wpb - 1) / wpb + 7) / 8);
    if (mode == LBM_MODE_STEP) launch_sch<Q, P, LBM_MODE_STEP>(sch, layout, b, s);
    else if (mode == LBM_MODE_SNAP) launch_sch<Q, P, LBM_MODE_SNAP>(sch, layout, b, s);
    else launch_layout<Q, P, LBM_MODE_PRE, SCH_TWO>(layout, b, s);
    return hipGetLastError();
}

template <int Q, class P>
hipError_t launch_init_t(const InitArgs& a, bool swap, hipStream_t s) {
    if (!a.gout || a.ns < a.n) return hipErrorInvalidValue;
    const int grid = lbm_grid_blocks(a.n, 256);
    InitArgs b = a;
    b.nblocks = (int)((a.n + 255) / 256);
    if (swap) hipLaunchKernelGGL((k_init<Q, P, true>), dim3(grid), dim3(256), 0, s, b);
    else hipLaunchKernelGGL((k_init<Q, P, false>), dim3(grid), dim3(256), 0, s, b);
    return hipGetLastError();
}

}  // namespace lbm
