// FP64 kernels: the reference-exact step, wall shear stress and the FP64
// conservation reductions.  Built with -ffp-contract=off: every expression
// rounds as written, which is what makes the FP64 path bit-identical to the
// reference arithmetic.
#include <algorithm>

#include "step_kernels.h"

using namespace lbm;

// One-thread kernels that move the device iteration base (graph chunks read
// it; kernels keep it in device memory so replays need no argument updates).
__global__ void k_advance_iter(int64_t* iter, int64_t by) { *iter = *iter + by; }
__global__ void k_set_iter(int64_t* iter, int64_t value) { *iter = value; }

hipError_t launch_advance_iter(int64_t* iter, int64_t by, hipStream_t s) {
    if (!iter) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_advance_iter, dim3(1), dim3(1), 0, s, iter, by);
    return hipGetLastError();
}

hipError_t launch_set_iter(int64_t* iter, int64_t value, hipStream_t s) {
    if (!iter) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_set_iter, dim3(1), dim3(1), 0, s, iter, value);
    return hipGetLastError();
}

// Probe contents: a 32-bit integer hash of the element index, read as floats
// in [1, 2), so every word is non-zero and the low mantissa bits vary.  The
// memory system streams all-zero data measurably faster than data of a real
// lattice, so a probe over zeros would overstate the rate.
__global__ __launch_bounds__(256) void k_stream_fill(float4* a, int64_t n) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        uint32_t h = (uint32_t)i * 0x9E3779B1u;
        float v[4];
        for (int k = 0; k < 4; ++k) {
            h ^= h >> 15; h *= 0x2C1B3C6Du; h ^= h >> 12;
            v[k] = __builtin_bit_cast(float, 0x3F800000u | (h >> 9));
        }
        a[i] = make_float4(v[0], v[1], v[2], v[3]);
    }
}

hipError_t launch_stream_fill(float4* a, int64_t n, hipStream_t s) {
    if (!a || n <= 0) return hipErrorInvalidValue;
    const int64_t blocks = std::min<int64_t>((n + 255) / 256, 256 * 64);
    hipLaunchKernelGGL(k_stream_fill, dim3((unsigned)blocks), dim3(256), 0, s, a, n);
    return hipGetLastError();
}

// Streaming probe for the measured-bandwidth figure: one float4 per thread,
// a grid that covers the whole buffer, in place (b == nullptr: a[i] += 1, so no
// store can be dropped) or a copy.  A grid-stride loop over a capped grid
// moves the same bytes markedly slower on this GPU (about 4.7 against 6.2 TB/s
// for a 1 GiB copy), so it would understate what the memory delivers.
__global__ __launch_bounds__(256) void k_stream_probe(float4* a, float4* b, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    float4 v = a[i];
    if (b) {
        b[i] = v;
    } else {
        v.x += 1.f; v.y += 1.f; v.z += 1.f; v.w += 1.f;
        a[i] = v;
    }
}

hipError_t launch_stream_probe(float4* a, float4* b, int64_t n, hipStream_t s) {
    if (!a || n <= 0 || (n + 255) / 256 > INT32_MAX) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_stream_probe, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, a, b, n);
    return hipGetLastError();
}

// Lattice-shaped probe for the practical ceiling of a step's traffic: Q arrays
// of stride ns (SoA, as the step stores them), one thread per node reads its Q
// elements and then writes Q elements, in place into the arrays in reverse
// order (the AA local step's pattern) or into a second lattice (the two-lattice
// pull's pattern).  SHIFT = 1 moves the reads (and, in place, the writes) one
// node along, so every 64-lane run starts off its cache line as a neighbour
// run does (the pull's reads; the AA gather's reads and writes).  Integer adds
// keep every store live.  Blocks map to nodes as in the step kernels (tile:
// one range per XCD, or tiles in turn for lattices beyond the Infinity
// Cache), walked downward when rev is set.
template <int Q, class E, int SHIFT>
__global__ __launch_bounds__(256) void k_lattice_probe(E* a, E* b, int64_t n, int64_t ns, int rev, int tile) {
    const int64_t i = logical_block(rev, tile) * 256 + threadIdx.x;
    if (i >= n) return;
    const int64_t j = i + SHIFT;  // < ns: the stride leaves room for the shift
    E v[Q];
#pragma unroll
    for (int d = 0; d < Q; ++d) v[d] = a[d * ns + j];
    E* o = b ? b : a;
    const int64_t k = b ? i : j;
#pragma unroll
    for (int d = 0; d < Q; ++d) o[(Q - 1 - d) * ns + k] = (E)(v[d] + 1);
}

template <class E, int SHIFT>
static hipError_t launch_lattice_probe_t(int q, int rev, void* a, void* b, int64_t n, int64_t ns, hipStream_t s) {
    const dim3 grid((unsigned)lbm_grid_blocks(n, 256));
    const int tile = lbm_launch_tile((uint64_t)q * (uint64_t)ns * sizeof(E) * (b ? 2u : 1u), 256);
    if (q == 19)
        hipLaunchKernelGGL((k_lattice_probe<19, E, SHIFT>), grid, dim3(256), 0, s, (E*)a, (E*)b, n, ns, rev, tile);
    else
        hipLaunchKernelGGL((k_lattice_probe<27, E, SHIFT>), grid, dim3(256), 0, s, (E*)a, (E*)b, n, ns, rev, tile);
    return hipGetLastError();
}

hipError_t launch_lattice_probe(int q, int elem_bytes, int shift, int rev, void* a, void* b, int64_t n, int64_t ns,
                                hipStream_t s) {
    if (!a || n <= 0 || ns < n + shift || (shift != 0 && shift != 1) || (q != 19 && q != 27) ||
        (n + 255) / 256 > INT32_MAX - 8)
        return hipErrorInvalidValue;
    if (elem_bytes == 2) return shift ? launch_lattice_probe_t<uint16_t, 1>(q, rev, a, b, n, ns, s)
                                      : launch_lattice_probe_t<uint16_t, 0>(q, rev, a, b, n, ns, s);
    if (elem_bytes == 4) return shift ? launch_lattice_probe_t<uint32_t, 1>(q, rev, a, b, n, ns, s)
                                      : launch_lattice_probe_t<uint32_t, 0>(q, rev, a, b, n, ns, s);
    if (elem_bytes == 8) return shift ? launch_lattice_probe_t<uint64_t, 1>(q, rev, a, b, n, ns, s)
                                      : launch_lattice_probe_t<uint64_t, 0>(q, rev, a, b, n, ns, s);
    return hipErrorInvalidValue;
}

// One block per used waveform slot, one thread per iteration of the chunk.
__global__ void k_wave_chunk(const WaveTab* waves, const int32_t* used, const int64_t* iter, int64_t it0, int it_abs,
                             double* wvals) {
    const int slot = used[blockIdx.x];
    const int k = threadIdx.x;
    const int64_t it = (it_abs ? it0 : *iter) + k;
    wvals[slot * LBM_WCHUNK + k] = wave_value(waves[slot], it);
}

hipError_t launch_wave_chunk(const WaveTab* waves, const int32_t* used, int nused, const int64_t* iter,
                             int64_t it0, int it_abs, double* wvals, hipStream_t s) {
    if (nused <= 0) return hipSuccess;
    if (!waves || !used || !wvals || (!it_abs && !iter)) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_wave_chunk, dim3(nused), dim3(LBM_WCHUNK), 0, s, waves, used, iter, it0, it_abs, wvals);
    return hipGetLastError();
}

hipError_t launch_step_f64(int mode, int sch, int layout, const StepArgs& a, hipStream_t s) {
    if (a.q == 19) return launch_step_t<19, PrecF64>(mode, sch, layout, a, s);
    return launch_step_t<27, PrecF64>(mode, sch, layout, a, s);
}

hipError_t launch_init_f64(int q, bool swap, const InitArgs& a, hipStream_t s) {
    if (q == 19) return launch_init_t<19, PrecF64>(a, swap, s);
    return launch_init_t<27, PrecF64>(a, swap, s);
}

// Simulation.cpp:1374-1421: mean over present neighbours x + c_d (the rest
// link makes the node its own neighbour) of u_n - u_x, times mu, then |.|.
template <int Q>
__global__ __launch_bounds__(256) void k_wss(WssArgs a) {
    using L = Lat<Q>;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.nwall) return;
    const int32_t x = a.wall[i];
    const int32_t p = a.pnode[x];
    const int64_t N = a.n;
    const double ux = a.u[x], uy = a.u[N + x], uz = a.u[2 * N + x];
    double g0 = 0.0, g1 = 0.0, g2 = 0.0;
    int cnt = 0;
#pragma unroll
    for (int d = 0; d < Q; ++d) {
        const int32_t nb = d == L::rest ? x : a.map[p + a.off[d]];
        if (nb < 0) continue;
        g0 += (a.u[nb] - ux);
        g1 += (a.u[N + nb] - uy);
        g2 += (a.u[2 * N + nb] - uz);
        cnt++;
    }
    if (cnt > 0) {
        g0 /= cnt;
        g1 /= cnt;
        g2 /= cnt;
    }
    const double s0 = a.mu * g0, s1 = a.mu * g1, s2 = a.mu * g2;
    a.wss[i] = sqrt(s0 * s0 + s1 * s1 + s2 * s2);
}

hipError_t launch_wss(int q, const WssArgs& a, hipStream_t s) {
    if (a.nwall == 0) return hipSuccess;
    if (!a.wall || !a.pnode || !a.map || !a.u || !a.wss) return hipErrorInvalidValue;
    const int grid = (int)((a.nwall + 255) / 256);
    if (q == 19) hipLaunchKernelGGL(k_wss<19>, dim3(grid), dim3(256), 0, s, a);
    else hipLaunchKernelGGL(k_wss<27>, dim3(grid), dim3(256), 0, s, a);
    return hipGetLastError();
}

// FP64 health check, summed in a fixed order so that every run and every
// device gives the same bits:
//   out[0]       = sum of the finite rho over the nodes,
//   out[1 + b]   = sum of rho u_axis over the nodes of boundary condition b
//                  (axis = its face normal),
//   out[1 + nbc] = number of nodes whose rho is not finite.
// Stage 1 (k_reduce_part).  Block g < LBM_RED_BLOCKS takes nodes
// [g C, (g + 1) C) with C = 256 ceil(n / (256 LBM_RED_BLOCKS)); thread t adds
// nodes g C + t, g C + t + 256, ... to 0.0 in turn, then the 256 thread sums
// are folded as a tree (sum[t] += sum[t + h] for h = 128, 64, ..., 1).  Each
// later block writes the flux term rho u_axis of 256 boundary nodes, which are
// listed grouped by condition and in node order within one.
// Stage 2 (k_reduce_final).  Block o sums its inputs (the stage-1 partials,
// or the flux terms of one condition) in the same thread-then-tree order.
__device__ __forceinline__ double tree_fold(double v, double* part) {
    part[threadIdx.x] = v;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) part[threadIdx.x] = part[threadIdx.x] + part[threadIdx.x + h];
        __syncthreads();
    }
    const double r = part[0];
    __syncthreads();
    return r;
}

__global__ __launch_bounds__(256) void k_reduce_part(const double* rho, const double* u, int64_t n, int64_t chunk,
                                                     const int32_t* bnode, const int32_t* bface, int64_t nbnd,
                                                     const FaceTab* faces, double* scratch) {
    __shared__ double part[256];
    const int64_t g = blockIdx.x;
    if (g < LBM_RED_BLOCKS) {
        const int64_t lo = g * chunk, hi = lo + chunk < n ? lo + chunk : n;
        double m = 0.0, bad = 0.0;
        for (int64_t i = lo + threadIdx.x; i < hi; i += 256) {
            const double r = rho[i];
            if (isfinite(r)) m = m + r;
            else bad = bad + 1.0;
        }
        m = tree_fold(m, part);
        bad = tree_fold(bad, part);
        if (threadIdx.x == 0) {
            scratch[g] = m;
            scratch[LBM_RED_BLOCKS + g] = bad;
        }
    } else {
        const int64_t i = (g - LBM_RED_BLOCKS) * 256 + threadIdx.x;
        if (i < nbnd) {
            const int32_t x = bnode[i];
            scratch[2 * LBM_RED_BLOCKS + i] = rho[x] * u[(int64_t)faces[bface[i]].axis * n + x];
        }
    }
}

__global__ __launch_bounds__(256) void k_reduce_final(const double* scratch, const int32_t* bstart, int nbc,
                                                      double* out) {
    __shared__ double part[256];
    const int o = blockIdx.x;
    int64_t lo, hi;
    if (o == 0) { lo = 0; hi = LBM_RED_BLOCKS; }
    else if (o <= nbc) { lo = 2 * LBM_RED_BLOCKS + bstart[o - 1]; hi = 2 * LBM_RED_BLOCKS + bstart[o]; }
    else { lo = LBM_RED_BLOCKS; hi = 2 * LBM_RED_BLOCKS; }
    double v = 0.0;
    for (int64_t i = lo + threadIdx.x; i < hi; i += 256) v = v + scratch[i];
    v = tree_fold(v, part);
    if (threadIdx.x == 0) out[o] = v;
}

hipError_t launch_reduce(const double* rho, const double* u, int64_t n, const int32_t* bnode,
                         const int32_t* bface, const int32_t* bstart, int64_t nbnd, const FaceTab* faces,
                         int nbc, double* scratch, double* out, hipStream_t s) {
    if (!rho || !u || !out || !scratch || !bstart || (nbnd > 0 && (!bnode || !bface || !faces)) || nbc < 0)
        return hipErrorInvalidValue;
    const int64_t per = (n + 256LL * LBM_RED_BLOCKS - 1) / (256LL * LBM_RED_BLOCKS);
    const int64_t chunk = 256 * (per > 0 ? per : 1);
    const int64_t blocks = LBM_RED_BLOCKS + (nbnd + 255) / 256;
    if (blocks > INT32_MAX) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_reduce_part, dim3((unsigned)blocks), dim3(256), 0, s, rho, u, n, chunk, bnode, bface, nbnd,
                       faces, scratch);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_reduce_final, dim3((unsigned)(nbc + 2)), dim3(256), 0, s, scratch, bstart, nbc, out);
    return hipGetLastError();
}
