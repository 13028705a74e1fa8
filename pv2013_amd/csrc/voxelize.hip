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
//   bins   facets with nz == 0 are dropped (their t is +-inf or NaN, never in
//          the window); facets whose XY projection is thinner than ~1e-6 rad
//          go to a list every column tests; the rest are binned into 8 x 8
//          column tiles by their XY bbox widened by a margin that bounds the
//          rounding error of the sign test (so binning never loses a facet the
//          exact test would accept) -- three device kernels, see k_bin_count;
//   voxels one workgroup of four wavefronts per tile (tiles with the longest
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
#include <mutex>
#include <vector>

#include "lbm_internal.h"
#include "voxel_window.h"

int lbm_ctx_fail(lbm_ctx* ctx, int code, const std::string& msg);
hipStream_t lbm_ctx_stream(lbm_ctx* ctx);
int lbm_ctx_device(lbm_ctx* ctx);

#define VCHECK(call)                                                                        \
    do {                                                                                    \
        hipError_t e_ = (call);                                                             \
        if (e_ != hipSuccess)                                                               \
            return lbm_ctx_fail(ctx, LBM_EIO, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

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
__device__ bool thin_xy(const double* T) {
    const double ax = T[3] - T[0], ay = T[4] - T[1], bx = T[6] - T[0], by = T[7] - T[1];
    const double cx = T[6] - T[3], cy = T[7] - T[4];
    const double area2 = fabs(ax * by - ay * bx);
    const double l = fmax(ax * ax + ay * ay, fmax(bx * bx + by * by, cx * cx + cy * cy));
    return !(area2 >= 1e-6 * l) || l == 0.0;
}

// ---------------------------------------------------------- facet binning
//
// On the device, in three launches:
//   k_bin_count  one thread per facet: facets with nz == 0 are dropped (their
//                t is +-inf or NaN, never in the window); slivers (thin_xy) are
//                appended to the list every column tests; the rest get the
//                range of 8 x 8 column tiles their XY bbox touches, widened by
//                a margin that bounds the rounding error of the sign test plus
//                the ray's 0.01 shift at the ray box faces (binning never
//                loses a facet the exact test would accept), and count into
//                each of those tiles;
//   k_bin_scan   one workgroup: tile list offsets (exclusive prefix sum), the
//                fill cursors, and the tile launch order (tiles bucketed by
//                floor(log2(list length)), longest buckets first, so the
//                slowest workgroups start first);
//   k_bin_fill   one thread per facet: appends the facet to each tile's list.
// List order within a tile varies from run to run; the flags do not, since a
// column's parity is the xor of independent toggles.

struct BinBufs {
    unsigned long long* hdr;  // [0] list entries, [1] facets every column tests, [2] entries above 2^31 - 1
    int32_t* cnt;             // per tile: count, then the fill cursor
    int32_t* off;             // ntiles + 1 list offsets
    int32_t* order;           // tiles in launch order
    int4* rng;                // per facet: tile range x0, x1, y0, y1 (x0 > x1: none)
    int32_t* glob;            // facets every column tests
    int32_t* list;            // the tile lists
    int64_t list_cap;
};

struct BinArgs {
    const double* tri;
    int64_t ntri, ntiles;
    int32_t nx, ny, ntx;
    double ox, oy, sp, slack;
    BinBufs b;
};

constexpr int HDR_WORDS = 4;
constexpr int SCAN_THREADS = 1024;
constexpr int NBUCKETS = 33;  // list length 0, then [2^(b-1), 2^b)

__global__ __launch_bounds__(256) void k_bin_zero(uint32_t* p, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0u;
}

// Tile range of facet t (x0 > x1: none); slivers go to the every-column list.
__device__ __forceinline__ int4 facet_tiles(const BinArgs& a, int64_t t, unsigned long long* nglob_slot) {
    const double* T = a.tri + 12 * t;
    int4 r = make_int4(1, 0, 1, 0);
    if (T[11] == 0.0) return r;
    if (thin_xy(T)) {
        const unsigned long long pos = atomicAdd(nglob_slot, 1ull);
        a.b.glob[pos] = (int32_t)t;
        return r;
    }
    const double xlo = fmin(T[0], fmin(T[3], T[6])) - a.slack, xhi = fmax(T[0], fmax(T[3], T[6])) + a.slack;
    const double ylo = fmin(T[1], fmin(T[4], T[7])) - a.slack, yhi = fmax(T[1], fmax(T[4], T[7])) + a.slack;
    const double i0 = fmax(floor((xlo - a.ox) / a.sp) - 1.0, 0.0);
    const double i1 = fmin(ceil((xhi - a.ox) / a.sp) + 1.0, (double)(a.nx - 1));
    const double j0 = fmax(floor((ylo - a.oy) / a.sp) - 1.0, 0.0);
    const double j1 = fmin(ceil((yhi - a.oy) / a.sp) + 1.0, (double)(a.ny - 1));
    if (i0 <= i1 && j0 <= j1) r = make_int4((int)i0 / TX, (int)i1 / TX, (int)j0 / TY, (int)j1 / TY);
    return r;
}

__global__ __launch_bounds__(256) void k_bin_count(BinArgs a) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= a.ntri) return;
    const int4 r = facet_tiles(a, t, &a.b.hdr[1]);
    a.b.rng[t] = r;
    if (r.x > r.y) return;
    for (int ty = r.z; ty <= r.w; ++ty)
        for (int tx = r.x; tx <= r.y; ++tx) atomicAdd(&a.b.cnt[(int64_t)ty * a.ntx + tx], 1);
    atomicAdd(&a.b.hdr[0], (unsigned long long)(r.y - r.x + 1) * (unsigned long long)(r.w - r.z + 1));
}

// The same two passes with the tile histogram privatised in LDS, for grids of
// at most LDS_TILES tiles: a workgroup takes FPW consecutive facets (meshes
// list neighbouring facets together, so a workgroup touches few tiles), counts
// into LDS, then adds each non-zero count to the global count with one atomic;
// the fill pass reserves one block of positions per tile per workgroup and
// hands out positions inside it with LDS atomics.
constexpr int LDS_TILES = 16384;  // 64 KiB of LDS counters
constexpr int FPW = 2048;         // facets per workgroup

__global__ __launch_bounds__(256) void k_bin_count_lds(BinArgs a) {
    extern __shared__ int32_t hist[];  // ntiles counters
    const int n = (int)a.ntiles;
    for (int i = threadIdx.x; i < n; i += 256) hist[i] = 0;
    __syncthreads();
    const int64_t t0 = (int64_t)blockIdx.x * FPW;
    const int64_t t1 = t0 + FPW < a.ntri ? t0 + FPW : a.ntri;
    unsigned long long entries = 0;
    for (int64_t t = t0 + threadIdx.x; t < t1; t += 256) {
        const int4 r = facet_tiles(a, t, &a.b.hdr[1]);
        a.b.rng[t] = r;
        for (int ty = r.z; ty <= r.w; ++ty)
            for (int tx = r.x; tx <= r.y; ++tx) atomicAdd(&hist[ty * a.ntx + tx], 1);
        if (r.x <= r.y) entries += (unsigned long long)(r.y - r.x + 1) * (unsigned long long)(r.w - r.z + 1);
    }
    if (entries) atomicAdd(&a.b.hdr[0], entries);
    __syncthreads();
    for (int i = threadIdx.x; i < n; i += 256)
        if (hist[i]) atomicAdd(&a.b.cnt[i], hist[i]);
}

__global__ __launch_bounds__(256) void k_bin_fill_lds(BinArgs a) {
    extern __shared__ int32_t pos[];  // per tile: this workgroup's count, then its next position
    const int n = (int)a.ntiles;
    for (int i = threadIdx.x; i < n; i += 256) pos[i] = 0;
    __syncthreads();
    const int64_t t0 = (int64_t)blockIdx.x * FPW;
    const int64_t t1 = t0 + FPW < a.ntri ? t0 + FPW : a.ntri;
    for (int64_t t = t0 + threadIdx.x; t < t1; t += 256) {
        const int4 r = a.b.rng[t];
        for (int ty = r.z; ty <= r.w; ++ty)
            for (int tx = r.x; tx <= r.y; ++tx) atomicAdd(&pos[ty * a.ntx + tx], 1);
    }
    __syncthreads();
    for (int i = threadIdx.x; i < n; i += 256)
        if (pos[i]) pos[i] = atomicAdd(&a.b.cnt[i], pos[i]);  // the cursor: reserve a block
    __syncthreads();
    for (int64_t t = t0 + threadIdx.x; t < t1; t += 256) {
        const int4 r = a.b.rng[t];
        for (int ty = r.z; ty <= r.w; ++ty)
            for (int tx = r.x; tx <= r.y; ++tx) {
                const int32_t p = atomicAdd(&pos[ty * a.ntx + tx], 1);
                if (p < a.b.list_cap) a.b.list[p] = (int32_t)t;
            }
    }
}

__device__ __forceinline__ int length_bucket(int len) { return len <= 0 ? 0 : 32 - __clz(len); }

__global__ __launch_bounds__(SCAN_THREADS) void k_bin_scan(BinArgs a) {
    __shared__ long long part[SCAN_THREADS];
    __shared__ int bcnt[NBUCKETS], bcur[NBUCKETS];
    const int tid = threadIdx.x;
    const int64_t n = a.ntiles;
    const int64_t c0 = n * tid / SCAN_THREADS, c1 = n * (tid + 1) / SCAN_THREADS;
    if (tid < NBUCKETS) bcnt[tid] = 0;
    long long sum = 0;
    for (int64_t i = c0; i < c1; ++i) sum += a.b.cnt[i];
    part[tid] = sum;
    __syncthreads();
    for (int h = 1; h < SCAN_THREADS; h <<= 1) {
        const long long v = tid >= h ? part[tid - h] : 0;
        __syncthreads();
        part[tid] += v;
        __syncthreads();
    }
    long long run = part[tid] - sum;
    for (int64_t i = c0; i < c1; ++i) {
        const int len = a.b.cnt[i];
        const int32_t o = (int32_t)(run < INT32_MAX ? run : INT32_MAX);
        a.b.off[i] = o;
        a.b.cnt[i] = o;  // the fill cursor
        atomicAdd(&bcnt[length_bucket(len)], 1);
        run += len;
    }
    if (tid == SCAN_THREADS - 1) {
        a.b.off[n] = (int32_t)(run < INT32_MAX ? run : INT32_MAX);
        a.b.hdr[2] = run > INT32_MAX ? 1ull : 0ull;
    }
    __syncthreads();
    if (tid == 0) {
        int acc = 0;
        for (int b = NBUCKETS - 1; b >= 0; --b) {
            bcur[b] = acc;
            acc += bcnt[b];
        }
    }
    __syncthreads();
    for (int64_t i = c0; i < c1; ++i) {
        // offsets saturated at 2^31 - 1 (then the call is rejected) could change
        // a length's bucket: never write past the order array
        const int pos = atomicAdd(&bcur[length_bucket(a.b.off[i + 1] - a.b.off[i])], 1);
        if (pos < n) a.b.order[pos] = (int32_t)i;
    }
}

__global__ __launch_bounds__(256) void k_bin_fill(BinArgs a) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= a.ntri) return;
    const int4 r = a.b.rng[t];
    for (int ty = r.z; ty <= r.w; ++ty)
        for (int tx = r.x; tx <= r.y; ++tx) {
            const int32_t pos = atomicAdd(&a.b.cnt[(int64_t)ty * a.ntx + tx], 1);
            if (pos < a.b.list_cap) a.b.list[pos] = (int32_t)t;
        }
}

size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

// Bytes of the binning buffers other than the tile lists, laid out from base
// (every array 256-byte aligned); fills b when base is not null.
size_t bin_layout(int64_t ntri, int64_t ntiles, char* base, BinBufs* b) {
    size_t o = 0;
    auto take = [&](size_t bytes) {
        const size_t at = o;
        o = align256(o + bytes);
        return base ? base + at : nullptr;
    };
    char* hdr = take(sizeof(unsigned long long) * HDR_WORDS);
    char* cnt = take(sizeof(int32_t) * (size_t)ntiles);
    char* off = take(sizeof(int32_t) * (size_t)(ntiles + 1));
    char* order = take(sizeof(int32_t) * (size_t)ntiles);
    char* rng = take(sizeof(int4) * (size_t)std::max<int64_t>(ntri, 1));
    char* glob = take(sizeof(int32_t) * (size_t)std::max<int64_t>(ntri, 1));
    if (b) {
        b->hdr = (unsigned long long*)hdr;
        b->cnt = (int32_t*)cnt;
        b->off = (int32_t*)off;
        b->order = (int32_t*)order;
        b->rng = (int4*)rng;
        b->glob = (int32_t*)glob;
        b->list = nullptr;
        b->list_cap = 0;
    }
    return o;
}

// Grid checks shared by the voxelizer entry points.
int check_grid(lbm_ctx* ctx, const char* fn, int64_t ntri, const lbm_grid_desc* g) {
    const int nx = g->dims[0], ny = g->dims[1], nz = g->dims[2];
    if (nx <= 0 || ny <= 0 || nz <= 0 || !(g->spacing > 0))
        return lbm_ctx_fail(ctx, LBM_EINVAL, std::string(fn) + ": bad grid");
    if (nz > MAXNZ) return lbm_ctx_fail(ctx, LBM_ERANGE, std::string(fn) + ": nz above 2048");
    if (ntri > INT32_MAX) return lbm_ctx_fail(ctx, LBM_ERANGE, std::string(fn) + ": too many facets");
    const int64_t ntiles = (int64_t)((nx + TX - 1) / TX) * ((ny + TY - 1) / TY);
    if (ntiles > INT32_MAX) return lbm_ctx_fail(ctx, LBM_ERANGE, std::string(fn) + ": more than 2^31 column tiles");
    return LBM_OK;
}

// Bins the device-resident facets d_tri and voxelizes into the device flags
// d_flags on the context stream.  The binning buffers are `b` (bin_layout);
// the tile lists come from list_buf(entries), which returns a device buffer of
// at least that many int32 or null (an error already recorded).  *ms gets the
// device time from the first binning kernel to the end of the voxelizer.
template <class ListBuf>
int bin_and_voxelize(lbm_ctx* ctx, const char* fn, const double* d_tri, int64_t ntri, const lbm_grid_desc* g,
                     uint8_t* d_flags, BinBufs b, ListBuf&& list_buf, double* ms) {
    hipStream_t st = lbm_ctx_stream(ctx);
    const int nx = g->dims[0], ny = g->dims[1], nz = g->dims[2];
    const int ntx = (nx + TX - 1) / TX, nty = (ny + TY - 1) / TY;
    const int64_t ntiles = (int64_t)ntx * nty;
    double ext = 0.0;
    for (int a = 0; a < 3; ++a)
        ext = std::max(ext, std::fabs(g->raybox[2 * a + 1] - g->raybox[2 * a]) + std::fabs(g->origin[a]) +
                                g->dims[a] * g->spacing);
    BinArgs ba;
    ba.tri = d_tri;
    ba.ntri = ntri;
    ba.ntiles = ntiles;
    ba.nx = nx;
    ba.ny = ny;
    ba.ntx = ntx;
    ba.ox = g->origin[0];
    ba.oy = g->origin[1];
    ba.sp = g->spacing;
    ba.slack = 1e-7 * std::max(ext, 1.0) + g->ray_offset;
    ba.b = b;
    EventPair ev;
    VCHECK(hipEventCreate(&ev.e0));
    VCHECK(hipEventCreate(&ev.e1));
    VCHECK(hipEventRecord(ev.e0, st));
    // counters left by the previous call start at zero (a fill kernel, no memset node)
    const int64_t zw = (int64_t)((char*)b.off - (char*)b.hdr) / 4;  // hdr and cnt
    hipLaunchKernelGGL(k_bin_zero, dim3((unsigned)((zw + 255) / 256)), dim3(256), 0, st, (uint32_t*)b.hdr, zw);
    VCHECK(hipGetLastError());
    const bool lds = ntiles <= LDS_TILES;
    const unsigned fgrid = (unsigned)std::max<int64_t>(1, (ntri + 255) / 256);
    const unsigned lgrid = (unsigned)std::max<int64_t>(1, (ntri + FPW - 1) / FPW);
    const size_t lds_hist = sizeof(int32_t) * (size_t)ntiles;
    if (ntri > 0) {
        if (lds) hipLaunchKernelGGL(k_bin_count_lds, dim3(lgrid), dim3(256), lds_hist, st, ba);
        else hipLaunchKernelGGL(k_bin_count, dim3(fgrid), dim3(256), 0, st, ba);
        VCHECK(hipGetLastError());
    }
    hipLaunchKernelGGL(k_bin_scan, dim3(1), dim3(SCAN_THREADS), 0, st, ba);
    VCHECK(hipGetLastError());
    unsigned long long hdr[HDR_WORDS];
    VCHECK(hipMemcpyAsync(hdr, b.hdr, sizeof hdr, hipMemcpyDeviceToHost, st));
    VCHECK(hipStreamSynchronize(st));
    if (hdr[2] || hdr[0] > (unsigned long long)INT32_MAX)
        return lbm_ctx_fail(ctx, LBM_ERANGE, std::string(fn) + ": bin lists too long");
    ba.b.list = list_buf((int64_t)hdr[0]);
    if (!ba.b.list) return LBM_ERANGE;
    ba.b.list_cap = (int64_t)hdr[0];
    if (ntri > 0) {
        if (lds) hipLaunchKernelGGL(k_bin_fill_lds, dim3(lgrid), dim3(256), lds_hist, st, ba);
        else hipLaunchKernelGGL(k_bin_fill, dim3(fgrid), dim3(256), 0, st, ba);
        VCHECK(hipGetLastError());
    }
    VoxArgs a;
    a.tri = d_tri; a.tile_off = b.off; a.order = b.order; a.tile_tri = ba.b.list; a.global_tri = b.glob;
    a.nglobal = (int32_t)hdr[1];
    a.nx = nx; a.ny = ny; a.nz = nz; a.ntx = ntx;
    a.ox = g->origin[0]; a.oy = g->origin[1]; a.oz = g->origin[2]; a.sp = g->spacing;
    a.bminx = g->raybox[0]; a.bmaxx = g->raybox[1]; a.bminy = g->raybox[2]; a.bmaxy = g->raybox[3];
    a.L = g->raybox[5] - g->raybox[4];
    a.pad = g->ray_pad; a.off = g->ray_offset; a.teps = g->t_eps;
    a.flags = d_flags;
    const size_t lds_bits = sizeof(uint32_t) * 64 * (size_t)((nz + 31) / 32);
    hipLaunchKernelGGL(k_voxelize, dim3((unsigned)ntiles), dim3(64 * VWAVES), lds_bits, st, a);
    VCHECK(hipGetLastError());
    VCHECK(hipEventRecord(ev.e1, st));
    VCHECK(hipEventSynchronize(ev.e1));
    float t = 0.f;
    VCHECK(hipEventElapsedTime(&t, ev.e0, ev.e1));
    if (ms) *ms = t;
    return LBM_OK;
}

// The no-allocation entry point's binning buffers: one array per device the
// library runs on, held until the process ends.  Host code uses only the
// device address hipGetSymbolAddress gives for it on each device.
constexpr size_t VOX_SCRATCH_BYTES = (size_t)64 << 20;
constexpr int VOX_MAX_DEVICES = 64;
static __device__ __align__(256) unsigned char scratch[VOX_SCRATCH_BYTES];
void* g_scratch_ptr[VOX_MAX_DEVICES];
std::mutex g_scratch_mu;

}  // namespace

// Called by lbm_ctx_create with `device` current: the address of this
// device's scratch array, once per device.
int lbm_vox_scratch_init(int device) {
    if (device < 0 || device >= VOX_MAX_DEVICES) return LBM_ERANGE;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_ptr[device]) return LBM_OK;
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return LBM_EIO;
    g_scratch_ptr[device] = p;
    return LBM_OK;
}

extern "C" int lbm_voxelize(lbm_ctx* ctx, const double* tri, int64_t ntri, const lbm_grid_desc* g,
                            uint8_t* flags_out, double* kernel_ms) {
    if (!ctx) return LBM_EINVAL;
    if (!g || !flags_out || (ntri > 0 && !tri) || ntri < 0)
        return lbm_ctx_fail(ctx, LBM_EINVAL, "lbm_voxelize: null argument");
    if (int rc = check_grid(ctx, "lbm_voxelize", ntri, g)) return rc;
    VCHECK(hipSetDevice(lbm_ctx_device(ctx)));
    hipStream_t st = lbm_ctx_stream(ctx);
    const int nx = g->dims[0], ny = g->dims[1], nz = g->dims[2];
    const int64_t ntiles = (int64_t)((nx + TX - 1) / TX) * ((ny + TY - 1) / TY);
    const int64_t nsites = (int64_t)nx * ny * nz;
    DevMem m;
    double* d_tri;
    uint8_t* d_flags;
    char* d_bins;
    VCHECK(m.alloc(&d_tri, (size_t)std::max<int64_t>(ntri, 1) * 12));
    VCHECK(m.alloc(&d_flags, (size_t)nsites));
    VCHECK(m.alloc(&d_bins, bin_layout(ntri, ntiles, nullptr, nullptr)));
    BinBufs b;
    bin_layout(ntri, ntiles, d_bins, &b);
    if (ntri > 0) VCHECK(hipMemcpyAsync(d_tri, tri, sizeof(double) * 12 * ntri, hipMemcpyHostToDevice, st));
    int32_t* d_list = nullptr;
    auto list_buf = [&](int64_t n) -> int32_t* {
        if (m.alloc(&d_list, (size_t)std::max<int64_t>(n, 1)) != hipSuccess) {
            lbm_ctx_fail(ctx, LBM_ENOMEM, "lbm_voxelize: tile lists");
            return nullptr;
        }
        return d_list;
    };
    int rc = bin_and_voxelize(ctx, "lbm_voxelize", d_tri, ntri, g, d_flags, b, list_buf, kernel_ms);
    if (rc) return rc;
    VCHECK(hipMemcpyAsync(flags_out, d_flags, (size_t)nsites, hipMemcpyDeviceToHost, st));
    VCHECK(hipStreamSynchronize(st));
    return LBM_OK;
}

extern "C" int lbm_voxelize_dev(lbm_ctx* ctx, const double* tri_dev, int64_t ntri, const lbm_grid_desc* g,
                                uint8_t* flags_dev, double* kernel_ms) {
    if (!ctx) return LBM_EINVAL;
    if (!g || !flags_dev || (ntri > 0 && !tri_dev) || ntri < 0)
        return lbm_ctx_fail(ctx, LBM_EINVAL, "lbm_voxelize_dev: null argument");
    if (int rc = check_grid(ctx, "lbm_voxelize_dev", ntri, g)) return rc;
    const int dev = lbm_ctx_device(ctx);
    VCHECK(hipSetDevice(dev));
    void* base = nullptr;
    {
        std::lock_guard<std::mutex> lock(g_scratch_mu);
        base = dev >= 0 && dev < VOX_MAX_DEVICES ? g_scratch_ptr[dev] : nullptr;
    }
    if (!base) return lbm_ctx_fail(ctx, LBM_EIO, "lbm_voxelize_dev: no scratch address for this device");
    const int nx = g->dims[0], ny = g->dims[1];
    const int64_t ntiles = (int64_t)((nx + TX - 1) / TX) * ((ny + TY - 1) / TY);
    const size_t fixed = bin_layout(ntri, ntiles, nullptr, nullptr);
    if (fixed > VOX_SCRATCH_BYTES)
        return lbm_ctx_fail(ctx, LBM_ERANGE, "lbm_voxelize_dev: " + std::to_string(fixed) +
                                                 " bytes of bins exceed the 64 MiB scratch");
    BinBufs b;
    bin_layout(ntri, ntiles, (char*)base, &b);
    const int64_t cap = (int64_t)((VOX_SCRATCH_BYTES - fixed) / sizeof(int32_t));
    auto list_buf = [&](int64_t n) -> int32_t* {
        if (n > cap) {
            lbm_ctx_fail(ctx, LBM_ERANGE, "lbm_voxelize_dev: " + std::to_string(n) +
                                              " tile-list entries exceed the 64 MiB scratch");
            return nullptr;
        }
        return (int32_t*)((char*)base + fixed);
    };
    return bin_and_voxelize(ctx, "lbm_voxelize_dev", tri_dev, ntri, g, flags_dev, b, list_buf, kernel_ms);
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
