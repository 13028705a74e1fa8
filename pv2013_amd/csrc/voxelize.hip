// STL voxelizer and boundary classification for gfx950.
//
// Geometry::labelPoints (ComplexGeometry.cpp:431-506) casts, for every grid
// point, a +z ray of length (max_z - min_z) + 0.1 and counts the facets whose
// plane (through v1, with the normal read from the file) it crosses with
// t in [-1e-7, 1 + 1e-7] at a point inside the facet's XY projection
// (inclusive sign test, ComplexGeometry.cpp:264-281).  The point is fluid iff
// the count is odd.
//
// The XY test does not depend on z, so the work is organised per column:
//   host   facets with nz == 0 are dropped (their t is +-inf or NaN, never in
//          the window); facets whose XY projection is thinner than ~1e-6 rad
//          go to a list every column tests; the rest are binned into 8 x 8
//          column tiles by their XY bbox widened by a margin that bounds the
//          rounding error of the sign test (so binning never loses a facet the
//          exact test would accept);
//   device one workgroup of four wavefronts per tile (tiles with the longest
//          lists launch first), one lane per column in each wavefront; the
//          wavefronts take the tile's facets in turns of 64, staged through
//          LDS; for each facet that passes the exact XY test the lane toggles
//          the parity bits (held in LDS) of the grid points whose t is in the
//          window: the reference's t expression is evaluated at the few points
//          near the two ends of that interval and the points between them are
//          toggled as a range (voxel_window.h); bits are written out as flags.
// Built with -ffp-contract=off so t and the sign test round as the reference.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <cstring>
#include <thread>
#include <vector>

#include "lbm_internal.h"
#include "voxel_window.h"

int lbm_ctx_fail(lbm_ctx* ctx, int code, const std::string& msg);
hipStream_t lbm_ctx_stream(lbm_ctx* ctx);
int lbm_ctx_device(lbm_ctx* ctx);

namespace {

constexpr int TX = 8, TY = 8;     // a tile of TX x TY columns per workgroup, one per lane
static_assert(TX * TY == 64, "one column per lane of a wavefront");
constexpr int VWAVES = 4;         // wavefronts per workgroup sharing a tile's facet list
constexpr int MAXNZ = 2048;       // parity bits per column held in LDS

struct VoxArgs {
    const double* tri;        // 12 doubles per facet
    const int32_t* tile_off;  // ntiles + 1
    const int32_t* order;     // tiles by decreasing list length, one per workgroup
    const int32_t* tile_tri;
    const int32_t* global_tri;
    int32_t nglobal;
    int32_t nx, ny, nz;
    int32_t ntx;
    double ox, oy, oz, sp;
    double bminx, bmaxx, bminy, bmaxy;
    double L, pad, off, teps;
    uint8_t* flags;
};

// ComplexGeometry.cpp:264-281
__device__ __forceinline__ bool inside_xy(double x, double y, const double* T) {
    const double x1 = T[0], y1 = T[1], x2 = T[3], y2 = T[4], x3 = T[6], y3 = T[7];
    const double d1 = ((x - x2) * (y1 - y2) - (y - y2) * (x1 - x2));
    const double d2 = ((x - x3) * (y2 - y3) - (y - y3) * (x2 - x3));
    const double d3 = ((x - x1) * (y3 - y1) - (y - y1) * (x3 - x1));
    const bool neg = (d1 < 0) || (d2 < 0) || (d3 < 0);
    const bool pos = (d1 > 0) || (d2 > 0) || (d3 > 0);
    return !(neg && pos);
}

// Parity bits of the workgroup's 64 columns in LDS: word w of column c at
// bits[w * 64 + c].  The workgroup's wavefronts flip them concurrently, so
// every flip is an LDS atomic xor.
struct LdsColumnBits {
    uint32_t* bits;
    int col;
    __device__ void flip(int k) { atomicXor(&bits[(k >> 5) * 64 + col], 1u << (k & 31)); }
    __device__ void flip_range(int lo, int hi) {
        for (int w = lo >> 5; w <= (hi >> 5); ++w) {
            uint32_t m = 0xffffffffu;
            if (w == (lo >> 5)) m &= 0xffffffffu << (lo & 31);
            if (w == (hi >> 5)) m &= 0xffffffffu >> (31 - (hi & 31));
            atomicXor(&bits[w * 64 + col], m);
        }
    }
};

// The workgroup's wavefronts take the list's batches of 64 facets in turn
// (batch r * VWAVES + wave in round r), each staging its batch through its own
// LDS slice; all threads run the same number of rounds, so the barriers are
// uniform.
__device__ void process_list(const VoxArgs& a, const int32_t* list, int count, double* sT,
                             uint32_t* bits, bool valid, double sx, double sy, double ex, double ey) {
    const int tid = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int rounds = (count + 64 * VWAVES - 1) / (64 * VWAVES);
    double* mine = sT + wave * 12 * 64;
    for (int r = 0; r < rounds; ++r) {
        const int base = (r * VWAVES + wave) * 64;
        const int m = max(0, min(64, count - base));
        __syncthreads();
        if (tid < m) {
            const double* T = a.tri + 12 * (int64_t)list[base + tid];
#pragma unroll
            for (int c = 0; c < 12; ++c) mine[c * 64 + tid] = T[c];
        }
        __syncthreads();
        if (!valid) continue;
        const lbm_vox::WinRay g{a.oz, a.sp, a.L, a.pad, a.teps, a.nz};
        LdsColumnBits col{bits, tid};
        for (int t = 0; t < m; ++t) {
            double T[12];
#pragma unroll
            for (int c = 0; c < 12; ++c) T[c] = mine[c * 64 + t];
            if (inside_xy(sx, sy, T)) lbm_vox::window_toggle(g, T, sx, sy, ex, ey, col);
        }
    }
}

__global__ __launch_bounds__(64 * VWAVES) void k_voxelize(VoxArgs a) {
    __shared__ double sT[VWAVES * 12 * 64];
    extern __shared__ uint32_t bits[];  // ceil(nz / 32) * 64 words
    const int tid = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int tile = a.order[blockIdx.x];
    const int tx = tile % a.ntx, ty = tile / a.ntx;
    const int i = tx * TX + (tid % TX), j = ty * TY + (tid / TX);
    const bool valid = i < a.nx && j < a.ny;
    const int nw = (a.nz + 31) >> 5;
    for (int w = wave; w < nw; w += VWAVES) bits[w * 64 + tid] = 0u;
    // the ray of ComplexGeometry.cpp:441-458 with adjustLineCoordinates case 0
    double sx = a.ox + i * a.sp, sy = a.oy + j * a.sp;
    double ex = sx, ey = sy;
    if (sx == a.bminx) { sx += a.off; ex += a.off; }
    if (sx == a.bmaxx) { sx -= a.off; ex -= a.off; }
    if (sy == a.bminy) { sy += a.off; ey += a.off; }
    if (sy == a.bmaxy) { sy -= a.off; ey -= a.off; }
    const int beg = a.tile_off[tile], end = a.tile_off[tile + 1];
    process_list(a, a.tile_tri + beg, end - beg, sT, bits, valid, sx, sy, ex, ey);
    process_list(a, a.global_tri, a.nglobal, sT, bits, valid, sx, sy, ex, ey);
    __syncthreads();
    if (!valid) return;
    const int64_t nxy = (int64_t)a.nx * a.ny;
    const int64_t col = i + (int64_t)a.nx * j;
    for (int k = wave; k < a.nz; k += VWAVES)
        a.flags[col + nxy * k] = (uint8_t)((bits[(k >> 5) * 64 + tid] >> (k & 31)) & 1u);
}

struct ClsArgs {
    const uint8_t* in;
    uint8_t* out;
    int32_t nx, ny, nz;
    int32_t nplanes;
    const int32_t* box;  // per plane: lo[3], hi[3] index intervals
    int32_t bx;          // workgroups per x-row
};

// ComplexGeometry.cpp:509-556 (boundary: fluid with an outside 26-neighbour or
// on a grid face), :579-638 (plane membership, here as index intervals worked
// out on the host in long double) and Subdomain.cpp:413-473 (priority).
// Workgroup b covers 256 sites of x-row b / bx (row = j + ny k), so the
// indices come from one 32-bit division instead of 64-bit ones per site.
__global__ __launch_bounds__(256) void k_classify(ClsArgs a) {
    const uint32_t b = blockIdx.x;
    const uint32_t row = b / (uint32_t)a.bx;
    const int i = (int)(b - row * (uint32_t)a.bx) * 256 + (int)threadIdx.x;
    if (i >= a.nx) return;
    const int j = (int)(row % (uint32_t)a.ny), k = (int)(row / (uint32_t)a.ny);
    const int64_t s = (int64_t)row * a.nx + i;
    const uint8_t v = a.in[s];
    if (v != 1) { a.out[s] = 0; return; }
    bool bnd = (i == 0 || i == a.nx - 1 || j == 0 || j == a.ny - 1 || k == 0 || k == a.nz - 1);
    if (!bnd) {
        // off the grid faces all 26 neighbours exist: issue their loads together
        const int64_t sy = a.nx, sz = (int64_t)a.nx * a.ny;
        uint32_t outside = 0u;
#pragma unroll
        for (int dz = -1; dz <= 1; ++dz)
#pragma unroll
            for (int dy = -1; dy <= 1; ++dy)
#pragma unroll
                for (int dx = -1; dx <= 1; ++dx)
                    if (dx != 0 || dy != 0 || dz != 0) outside |= a.in[s + dx + sy * dy + sz * dz] == 0;
        bnd = outside != 0u;
    }
    if (!bnd) { a.out[s] = 1; return; }
    uint8_t f = 2;
    for (int p = 0; p < a.nplanes; ++p) {
        const int32_t* b = a.box + 6 * p;
        if (i >= b[0] && i <= b[3] && j >= b[1] && j <= b[4] && k >= b[2] && k <= b[5]) {
            f = (uint8_t)(3 + p);
            break;
        }
    }
    a.out[s] = f;
}

// A pair of timing events, destroyed on every return path.
struct EventPair {
    hipEvent_t e0 = nullptr, e1 = nullptr;
    ~EventPair() {
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
    }
};

struct DevMem {
    std::vector<void*> ptrs;
    ~DevMem() {
        for (void* p : ptrs) (void)hipFree(p);
    }
    template <class T>
    hipError_t alloc(T** p, size_t count) {
        void* v = nullptr;
        hipError_t e = hipMalloc(&v, std::max<size_t>(count, 1) * sizeof(T));
        if (e == hipSuccess) ptrs.push_back(v);
        *p = (T*)v;
        return e;
    }
};

// Sliver test: a facet whose XY projection has its smallest angle below about
// 1e-6 rad is tested at every column instead of being binned.
bool thin_xy(const double* T) {
    const double ax = T[3] - T[0], ay = T[4] - T[1], bx = T[6] - T[0], by = T[7] - T[1];
    const double cx = T[6] - T[3], cy = T[7] - T[4];
    const double area2 = std::fabs(ax * by - ay * bx);
    const double l = std::max(ax * ax + ay * ay, std::max(bx * bx + by * by, cx * cx + cy * cy));
    return !(area2 >= 1e-6 * l) || l == 0.0;
}

}  // namespace

#define VCHECK(call)                                                                        \
    do {                                                                                    \
        hipError_t e_ = (call);                                                             \
        if (e_ != hipSuccess)                                                               \
            return lbm_ctx_fail(ctx, LBM_EIO, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

extern "C" int lbm_voxelize(lbm_ctx* ctx, const double* tri, int64_t ntri, const lbm_grid_desc* g,
                            uint8_t* flags_out, double* kernel_ms) {
    if (!ctx) return LBM_EINVAL;
    if (!g || !flags_out || (ntri > 0 && !tri) || ntri < 0)
        return lbm_ctx_fail(ctx, LBM_EINVAL, "lbm_voxelize: null argument");
    const int nx = g->dims[0], ny = g->dims[1], nz = g->dims[2];
    if (nx <= 0 || ny <= 0 || nz <= 0 || !(g->spacing > 0))
        return lbm_ctx_fail(ctx, LBM_EINVAL, "lbm_voxelize: bad grid");
    if (nz > MAXNZ) return lbm_ctx_fail(ctx, LBM_ERANGE, "lbm_voxelize: nz above 2048");
    if (ntri > INT32_MAX) return lbm_ctx_fail(ctx, LBM_ERANGE, "lbm_voxelize: too many facets");
    VCHECK(hipSetDevice(lbm_ctx_device(ctx)));
    hipStream_t st = lbm_ctx_stream(ctx);
    const double ox = g->origin[0], oy = g->origin[1], sp = g->spacing;
    double ext = 0.0;
    for (int a = 0; a < 3; ++a)
        ext = std::max(ext, std::fabs(g->raybox[2 * a + 1] - g->raybox[2 * a]) + std::fabs(g->origin[a]) +
                                g->dims[a] * g->spacing);
    const double margin = 1e-7 * std::max(ext, 1.0);
    const int ntx = (nx + TX - 1) / TX, nty = (ny + TY - 1) / TY;
    const int64_t ntiles = (int64_t)ntx * nty;
    if (ntiles > INT32_MAX) return lbm_ctx_fail(ctx, LBM_ERANGE, "lbm_voxelize: more than 2^31 column tiles");
    // Bin facets into column tiles.  Up to 16 host threads take contiguous facet
    // ranges with per-thread tile counts; the counts then become per-thread
    // write offsets, so every tile's list holds its facets in facet order, as a
    // single pass would (the flags do not depend on the order either way).
    struct Part {
        std::vector<int32_t> keep, global, cnt;
        std::vector<std::array<int32_t, 4>> rng;
    };
    const int64_t hw = std::max(1u, std::thread::hardware_concurrency());
    int nthr = (int)std::min<int64_t>(std::min<int64_t>(16, hw), ntri / 16384 + 1);
    nthr = (int)std::max<int64_t>(1, std::min<int64_t>(nthr, ((int64_t)1 << 26) / ntiles));
    std::vector<Part> parts(nthr);
    bool oom = false;
    auto run = [&](auto&& work) {
        std::vector<std::thread> th;
        std::vector<char> failed(nthr, 0);
        auto guarded = [&](int p) {
            try {
                work(p);
            } catch (...) {
                failed[p] = 1;
            }
        };
        for (int p = 1; p < nthr; ++p) th.emplace_back(guarded, p);
        guarded(0);
        for (auto& t : th) t.join();
        for (char f : failed) oom = oom || f;
    };
    const double slack = margin + g->ray_offset;
    run([&](int p) {
        Part& P = parts[p];
        P.cnt.assign((size_t)ntiles, 0);
        for (int64_t t = ntri * p / nthr; t < ntri * (p + 1) / nthr; ++t) {
            const double* T = tri + 12 * t;
            if (T[11] == 0.0) continue;  // t = +-inf or NaN: never counted
            if (thin_xy(T)) { P.global.push_back((int32_t)t); continue; }
            const double xlo = std::min({T[0], T[3], T[6]}) - slack, xhi = std::max({T[0], T[3], T[6]}) + slack;
            const double ylo = std::min({T[1], T[4], T[7]}) - slack, yhi = std::max({T[1], T[4], T[7]}) + slack;
            long i0 = (long)std::floor((xlo - ox) / sp) - 1, i1 = (long)std::ceil((xhi - ox) / sp) + 1;
            long j0 = (long)std::floor((ylo - oy) / sp) - 1, j1 = (long)std::ceil((yhi - oy) / sp) + 1;
            i0 = std::max(i0, 0L); j0 = std::max(j0, 0L);
            i1 = std::min(i1, (long)nx - 1); j1 = std::min(j1, (long)ny - 1);
            if (i0 > i1 || j0 > j1) continue;
            std::array<int32_t, 4> r = {(int32_t)(i0 / TX), (int32_t)(i1 / TX), (int32_t)(j0 / TY),
                                        (int32_t)(j1 / TY)};
            P.keep.push_back((int32_t)t);
            P.rng.push_back(r);
            for (int ty = r[2]; ty <= r[3]; ++ty)
                for (int tx = r[0]; tx <= r[1]; ++tx) P.cnt[(int64_t)ty * ntx + tx]++;
        }
    });
    if (oom) return lbm_ctx_fail(ctx, LBM_ENOMEM, "lbm_voxelize: out of host memory while binning");
    // tile offsets, then each thread's write offset inside every tile
    std::vector<int32_t> cnt(ntiles + 1, 0);
    int64_t total = 0;
    for (int64_t i = 0; i < ntiles; ++i) {
        cnt[i] = (int32_t)std::min<int64_t>(total, INT32_MAX);
        for (int p = 0; p < nthr; ++p) {
            const int32_t c = parts[p].cnt[i];
            parts[p].cnt[i] = (int32_t)std::min<int64_t>(total, INT32_MAX);
            total += c;
        }
    }
    if (total > INT32_MAX / 2) return lbm_ctx_fail(ctx, LBM_ERANGE, "lbm_voxelize: bin lists too long");
    cnt[ntiles] = (int32_t)total;
    std::vector<int32_t> lists(std::max<int64_t>(total, 1)), global;
    for (const Part& P : parts) global.insert(global.end(), P.global.begin(), P.global.end());
    // longest lists first, so the slowest workgroups start first
    std::vector<int32_t> order(ntiles);
    for (int64_t i = 0; i < ntiles; ++i) order[i] = (int32_t)i;
    std::stable_sort(order.begin(), order.end(),
                     [&](int32_t x, int32_t y) { return cnt[x + 1] - cnt[x] > cnt[y + 1] - cnt[y]; });
    run([&](int p) {
        Part& P = parts[p];
        for (size_t u = 0; u < P.keep.size(); ++u) {
            const auto& r = P.rng[u];
            for (int ty = r[2]; ty <= r[3]; ++ty)
                for (int tx = r[0]; tx <= r[1]; ++tx) lists[P.cnt[(int64_t)ty * ntx + tx]++] = P.keep[u];
        }
    });
    DevMem m;
    double* d_tri;
    int32_t *d_off, *d_list, *d_glob, *d_order;
    uint8_t* d_flags;
    const int64_t nsites = (int64_t)nx * ny * nz;
    VCHECK(m.alloc(&d_tri, (size_t)std::max<int64_t>(ntri, 1) * 12));
    VCHECK(m.alloc(&d_off, (size_t)ntiles + 1));
    VCHECK(m.alloc(&d_order, (size_t)ntiles));
    VCHECK(m.alloc(&d_list, lists.size()));
    VCHECK(m.alloc(&d_glob, std::max<size_t>(global.size(), 1)));
    VCHECK(m.alloc(&d_flags, (size_t)nsites));
    if (ntri > 0) VCHECK(hipMemcpyAsync(d_tri, tri, sizeof(double) * 12 * ntri, hipMemcpyHostToDevice, st));
    VCHECK(hipMemcpyAsync(d_off, cnt.data(), sizeof(int32_t) * (ntiles + 1), hipMemcpyHostToDevice, st));
    VCHECK(hipMemcpyAsync(d_order, order.data(), sizeof(int32_t) * ntiles, hipMemcpyHostToDevice, st));
    VCHECK(hipMemcpyAsync(d_list, lists.data(), sizeof(int32_t) * lists.size(), hipMemcpyHostToDevice, st));
    if (!global.empty())
        VCHECK(hipMemcpyAsync(d_glob, global.data(), sizeof(int32_t) * global.size(), hipMemcpyHostToDevice, st));
    VoxArgs a;
    a.tri = d_tri; a.tile_off = d_off; a.order = d_order; a.tile_tri = d_list; a.global_tri = d_glob;
    a.nglobal = (int32_t)global.size();
    a.nx = nx; a.ny = ny; a.nz = nz; a.ntx = ntx;
    a.ox = ox; a.oy = oy; a.oz = g->origin[2]; a.sp = sp;
    a.bminx = g->raybox[0]; a.bmaxx = g->raybox[1]; a.bminy = g->raybox[2]; a.bmaxy = g->raybox[3];
    a.L = g->raybox[5] - g->raybox[4];
    a.pad = g->ray_pad; a.off = g->ray_offset; a.teps = g->t_eps;
    a.flags = d_flags;
    EventPair ev;
    VCHECK(hipEventCreate(&ev.e0));
    VCHECK(hipEventCreate(&ev.e1));
    hipEvent_t e0 = ev.e0, e1 = ev.e1;
    VCHECK(hipEventRecord(e0, st));
    const size_t lds_bits = sizeof(uint32_t) * 64 * (size_t)((nz + 31) / 32);
    hipLaunchKernelGGL(k_voxelize, dim3((unsigned)ntiles), dim3(64 * VWAVES), lds_bits, st, a);
    hipError_t le = hipGetLastError();
    VCHECK(hipEventRecord(e1, st));
    VCHECK(le);
    VCHECK(hipMemcpyAsync(flags_out, d_flags, (size_t)nsites, hipMemcpyDeviceToHost, st));
    VCHECK(hipStreamSynchronize(st));
    float ms = 0.f;
    VCHECK(hipEventElapsedTime(&ms, e0, e1));
    if (kernel_ms) *kernel_ms = ms;
    return LBM_OK;
}

// Index interval of grid points along one axis that satisfy a plane's
// conditions on that axis, evaluated in long double as
// ComplexGeometry.cpp:579-633 does (coordinates are monotone in the index).
static void axis_interval(const lbm_grid_desc* g, const lbm_plane_desc* p, int axis, int32_t* lo, int32_t* hi) {
    int32_t first = -1, last = -2;
    const long double v = (long double)p->value_hi + (long double)p->value_lo;
    for (int i = 0; i < g->dims[axis]; ++i) {
        const double c = g->origin[axis] + i * g->spacing;
        bool ok = true;
        if (p->axis == axis) ok = std::fabs((long double)c - v) <= (long double)p->tol;
        for (int r = 0; r < p->nranges && ok; ++r)
            if (p->range_axis[r] == axis)
                ok = ((long double)p->range_lo[r] < (long double)c) && ((long double)c < (long double)p->range_hi[r]);
        if (ok) {
            if (first < 0) first = i;
            last = i;
        }
    }
    *lo = first < 0 ? 1 : first;
    *hi = first < 0 ? 0 : last;
}

extern "C" int lbm_classify(lbm_ctx* ctx, const lbm_grid_desc* g, const lbm_plane_desc* planes, int nplanes,
                            uint8_t* flags, double* kernel_ms) {
    if (!ctx) return LBM_EINVAL;
    if (!g || !flags || nplanes < 0 || (nplanes > 0 && !planes))
        return lbm_ctx_fail(ctx, LBM_EINVAL, "lbm_classify: null argument");
    if (nplanes > 252) return lbm_ctx_fail(ctx, LBM_ERANGE, "lbm_classify: more than 252 planes");
    for (int p = 0; p < nplanes; ++p) {
        if (planes[p].axis < 0 || planes[p].axis > 2 || planes[p].nranges < 0 || planes[p].nranges > 3)
            return lbm_ctx_fail(ctx, LBM_EINVAL, "lbm_classify: bad plane");
        for (int r = 0; r < planes[p].nranges; ++r)
            if (planes[p].range_axis[r] < 0 || planes[p].range_axis[r] > 2)
                return lbm_ctx_fail(ctx, LBM_EINVAL, "lbm_classify: bad plane range axis");
    }
    const int nx = g->dims[0], ny = g->dims[1], nz = g->dims[2];
    if (nx <= 0 || ny <= 0 || nz <= 0) return lbm_ctx_fail(ctx, LBM_EINVAL, "lbm_classify: bad grid");
    VCHECK(hipSetDevice(lbm_ctx_device(ctx)));
    hipStream_t st = lbm_ctx_stream(ctx);
    std::vector<int32_t> box(6 * std::max(nplanes, 1));
    for (int p = 0; p < nplanes; ++p)
        for (int a = 0; a < 3; ++a) axis_interval(g, &planes[p], a, &box[6 * p + a], &box[6 * p + 3 + a]);
    const int64_t ns = (int64_t)nx * ny * nz;
    DevMem m;
    uint8_t *d_in, *d_out;
    int32_t* d_box;
    VCHECK(m.alloc(&d_in, (size_t)ns));
    VCHECK(m.alloc(&d_out, (size_t)ns));
    VCHECK(m.alloc(&d_box, box.size()));
    VCHECK(hipMemcpyAsync(d_in, flags, (size_t)ns, hipMemcpyHostToDevice, st));
    VCHECK(hipMemcpyAsync(d_box, box.data(), sizeof(int32_t) * box.size(), hipMemcpyHostToDevice, st));
    const int64_t bx = (nx + 255) / 256, blocks = bx * ny * (int64_t)nz;
    if (blocks > INT32_MAX) return lbm_ctx_fail(ctx, LBM_ERANGE, "lbm_classify: grid too large");
    ClsArgs a{d_in, d_out, nx, ny, nz, nplanes, d_box, (int32_t)bx};
    EventPair ev;
    VCHECK(hipEventCreate(&ev.e0));
    VCHECK(hipEventCreate(&ev.e1));
    hipEvent_t e0 = ev.e0, e1 = ev.e1;
    VCHECK(hipEventRecord(e0, st));
    hipLaunchKernelGGL(k_classify, dim3((unsigned)blocks), dim3(256), 0, st, a);
    hipError_t le = hipGetLastError();
    VCHECK(hipEventRecord(e1, st));
    VCHECK(le);
    VCHECK(hipMemcpyAsync(flags, d_out, (size_t)ns, hipMemcpyDeviceToHost, st));
    VCHECK(hipStreamSynchronize(st));
    float ms = 0.f;
    VCHECK(hipEventElapsedTime(&ms, e0, e1));
    if (kernel_ms) *kernel_ms = ms;
    return LBM_OK;
}
