// Internal interfaces between the host domain builder (domain.cpp) and the
// kernel translation units (kernels_f64.hip, kernels_lp.hip, voxelize.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

#include "lbm_capi.h"

constexpr int LBM_MAXBC = LBM_MAX_BCS;
constexpr int LBM_MAXWAVE = LBM_MAX_WAVEFORMS;
constexpr int LBM_MAXK = 16;

// One Zou-He boundary condition as the kernels see it.  The population index
// sets of each face are resolved at compile time from the lattice tables.
struct FaceTab {
    int32_t kind, axis, sigma, waveform;
    int32_t typo;  // x-face pressure outlet N_xz typo (Simulation.cpp:1124-1129)
    int32_t pad_;
    double rho;
    double u[3];
};

struct WaveTab {
    int32_t kind, nk;
    double period, dt;
    double a[LBM_MAXK], phi[LBM_MAXK];
};

// Everything a step / pre-state kernel reads.  Node arrays are indexed by the
// packed node id n; boundary arrays by the boundary rank r of n.
struct StepArgs {
    int64_t n;              // fluid nodes
    int64_t ns;             // stride of the population arrays (n rounded up to 64)
    int64_t nrec;           // boundary records
    int32_t nblocks;        // logical blocks (grid is rounded up to a multiple of 8)
    int32_t q;
    int32_t block;          // threads per block of the step kernels: 64, 128, 256 or 512
    int32_t wide;           // 16-bit row chunks with records: 64-bit addresses (lattice >= 4 GiB)
    const int32_t* pnode;   // padded map index of node n
    const int32_t* map;     // padded site -> node id or -1
    const uint64_t* bbits;  // bit n: node n has a boundary record
    const uint32_t* wreg;   // bit w: the 64 nodes of wavefront w are interior nodes of one
                            // contiguous row (null: no fast path)
    const int32_t* bpre;    // records before word n/64
    const uint8_t* btype;   // per record: flag value (1 fluid, 2 wall, 3+b)
    const uint32_t* bbb;    // per record: populations taken by bounce-back
    const uint32_t* bmiss;  // per record: populations whose source node is missing
    const int32_t* rec_node;  // per record: its node id
    const int32_t* rsrc;    // [d][r] source node of record r in direction d (x - c_d), -1: none
    const int32_t* seg;     // [s][LBM_SEG_W] row segments: start, length, source base per direction
    int64_t nseg;
    const int32_t* chunk;   // [c][lbm_chunk_w(q)] row chunks (fused layout)
    int64_t nchunk;
    // Wavefront table of the fused layout (null: wavefront j takes chunk j).
    // Entry j: chunk index c in bits 0-29; bit 30 set: the wavefront also takes
    // chunk c + 1, the next piece of the same aligned 64-node window.
    const int32_t* wave;
    int64_t nwave;
    int32_t rb8, sb8;       // record / segment blocks per XCD (two-launch layout, set by the launcher)
    const double* bprof;    // per record: velocity profile factor
    const void* bstale;     // [e][r] stale populations (double, or float deviations)
    const FaceTab* faces;
    const double* wvals;    // [slot][LBM_WCHUNK] waveform values of the current chunk
    int32_t wk;             // this launch's row in the chunk
    int32_t rev;            // 1: each XCD takes its logical blocks in descending order
    int32_t tile;           // 0: one range of logical blocks per XCD; T: tiles of T blocks in turn
    const void* gin;        // [d][ns] post-collision populations
    void* gout;
    double omega;
    int32_t off[27];        // padded map offset of c_d
    // pre-state outputs (mode PRE): f' [d][n] FP64, rho [n], u [c][n]
    double* fpre;
    double* rho;
    double* u;
};

struct InitArgs {
    int64_t n;
    int64_t ns;          // population array stride
    int32_t nblocks;
    const double* fpre;  // [d][n]
    void* gout;
    double omega;
};

struct WssArgs {
    int64_t nwall;
    int64_t n;
    const int32_t* wall;   // wall node ids
    const int32_t* pnode;
    const int32_t* map;
    const double* u;       // [c][n]
    double mu;
    int32_t off[27];
    double* wss;           // per wall node
};

enum { LBM_MODE_STEP = 0, LBM_MODE_PRE = 1, LBM_MODE_SNAP = 2 };

// How a step's nodes map onto launches: the generic per-node kernel, or the
// record/segment kernel in one launch or in two.
enum { LBM_LAYOUT_GENERIC = 0, LBM_LAYOUT_FUSED = 1, LBM_LAYOUT_SPLIT = 2 };

// Row segment table entry: {start node, length (1..64), base[Q]} padded to 32
// ints; base[d] is the node id at x_start - c_d, so lane l of the segment reads
// direction d from node base[d] + l.
constexpr int LBM_SEG_W = 32;

// Row chunk table entry (fused layout): every node belongs to one chunk, up to
// 64 consecutive node ids along one row.
//   [0] start node, [1] length, [2] records before start, [3] 1 if a record
//   lane of the chunk keeps a stale population (an unlinked, unbounced link), else 0,
//   [4], [5] record mask (bit l: node start + l has a boundary record),
//   [6 + d] base_d,
//   [6 + Q + d / 2] bits 16 (d & 1) .. 16 (d & 1) + 15: lo_d | hi_d << 8:
//   lanes l in [lo_d, hi_d) read direction d from node base_d + l; the other
//   lanes' source x - c_d is missing (bounce-back or stale value).
// 36 ints (144 B) for D3Q19, 48 ints (192 B) for D3Q27.
constexpr int lbm_chunk_w(int q) { return q == 19 ? 36 : 48; }
constexpr int32_t LBM_WAVE_PAIR = 1 << 30;  // wave table: the entry's chunk and the next one
static_assert(6 + 19 + 10 <= 36 && 6 + 27 + 14 <= 48, "chunk row fits");

// Waveform values are tabulated once per chunk of LBM_WCHUNK consecutive
// iterations: wvals[slot][k] = W_slot(base + k), base = it0 (it_abs) or *iter.
constexpr int LBM_WCHUNK = 64;

// kernels_f64.hip
hipError_t launch_advance_iter(int64_t* iter, int64_t by, hipStream_t s);
hipError_t launch_set_iter(int64_t* iter, int64_t value, hipStream_t s);
// bandwidth probe: a[i] += 1 over n float4 (in place) or b[i] = a[i]
hipError_t launch_stream_probe(float4* a, float4* b, int64_t n, hipStream_t s);
hipError_t launch_stream_fill(float4* a, int64_t n, hipStream_t s);
// lattice-shaped probe: q arrays of elem_bytes elements, stride ns; b == nullptr:
// in place (reversed array order), else into lattice b; shift 1: reads (in place:
// reads and writes) one node along, off their cache lines (needs ns >= n + 1);
// rev 1: the blocks walk the lattice downward; blocks map to nodes as in the step
// kernels (lbm_launch_tile of the bytes the pass touches)
hipError_t launch_lattice_probe(int q, int elem_bytes, int shift, int rev, void* a, void* b, int64_t n, int64_t ns,
                                hipStream_t s);
hipError_t launch_wave_chunk(const WaveTab* waves, const int32_t* used, int nused, const int64_t* iter,
                             int64_t it0, int it_abs, double* wvals, hipStream_t s);
// sch: 0 two-lattice pull, 1 AA gather step, 2 AA local step
// layout: LBM_LAYOUT_*
hipError_t launch_step_f64(int mode, int sch, int layout, const StepArgs& a, hipStream_t s);
hipError_t launch_init_f64(int q, bool swap, const InitArgs& a, hipStream_t s);
hipError_t launch_wss(int q, const WssArgs& a, hipStream_t s);
// FP64 reductions in a fixed order (kernels_f64.hip): bnode/bface list the
// boundary-condition nodes grouped by condition (bstart[b] .. bstart[b + 1]),
// scratch holds lbm_reduce_scratch(nbnd) doubles.
constexpr int LBM_RED_BLOCKS = 1024;
inline int64_t lbm_reduce_scratch(int64_t nbnd) { return 2 * LBM_RED_BLOCKS + (nbnd > 0 ? nbnd : 1); }
hipError_t launch_reduce(const double* rho, const double* u, int64_t n, const int32_t* bnode,
                         const int32_t* bface, const int32_t* bstart, int64_t nbnd, const FaceTab* faces,
                         int nbc, double* scratch, double* out /* 1 + nbc + 1 */, hipStream_t s);
// kernels_lp.hip (storage LBM_STORE_F32 or LBM_STORE_F16S)
hipError_t launch_step_lp(int storage, int mode, int sch, int layout, const StepArgs& a, hipStream_t s);
hipError_t launch_init_lp(int storage, int q, bool swap, const InitArgs& a, hipStream_t s);

// Grid rounded up to whole groups of 8 blocks so the XCD-aware block remap is a
// bijection (blocks are dispatched round-robin over the 8 XCDs).
// Tile of the block remap, in blocks, for a launch of `block`-thread blocks
// that touches `bytes` of lattice: one range per XCD while the lattice fits the
// 256 MB Infinity Cache (neighbour lines stay in one L2), else tiles of 64
// wavefronts so the XCDs stream one front.
constexpr uint64_t LBM_INFINITY_CACHE = (uint64_t)256 << 20;
inline int lbm_launch_tile(uint64_t bytes, int block) {
    return bytes > LBM_INFINITY_CACHE ? 64 * 64 / block : 0;
}
inline int lbm_grid_blocks(int64_t items, int block) {
    int64_t b = (items + block - 1) / block;
    return (int)(((b + 7) / 8) * 8);
}
