// Internal declarations shared by the HIP translation units of libfsr.
//
// Data layout in HBM (one frame, N particles, C spatial cells, V scalar
// vertices, B band vertices):
//   particles      float3[N] caller buffer, sorted in place by cell id
//   packed         float4[N] (x, y, z, density) copy taken by the hash pass
//   sorted         float4[N] (x, y, z, density) in cell order, used by every
//                  neighbour gather (16-byte aligned loads)
//   ranges         fsr_index_range[C]
//   band bits      u32[Wx * resy * resz], Wx = ceil(resx / 32): one bit per
//                  scalar vertex instead of the reference's dense u32[V] flags
//   band index     u32[B] ascending linear vertex ids (the reference's
//                  surfaceVerticesIndexArray)
//   cell-major     u32[B] the same ids grouped by the hash cell each vertex
//                  falls in, cut into 64-vertex chunks for the field kernels
//   phi            float[V] caller buffer
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string>
#include <vector>

#include "../../include/fsr.h"

namespace fsr {

// ---------------------------------------------------------------------------
// Error handling
// ---------------------------------------------------------------------------
void set_error(const std::string& msg);
fsr_status hip_fail(hipError_t e, const char* what, const char* file, int line);

#define FSR_HIP(expr)                                                      \
    do {                                                                   \
        hipError_t _e = (expr);                                            \
        if (_e != hipSuccess) return ::fsr::hip_fail(_e, #expr, __FILE__, __LINE__); \
    } while (0)

#define FSR_LAUNCH_CHECK(name)                                             \
    do {                                                                   \
        hipError_t _e = hipGetLastError();                                 \
        if (_e != hipSuccess) return ::fsr::hip_fail(_e, name, __FILE__, __LINE__); \
    } while (0)

#define FSR_TRY(expr)                                                      \
    do {                                                                   \
        fsr_status _s = (expr);                                            \
        if (_s != FSR_OK) return _s;                                       \
    } while (0)

// ---------------------------------------------------------------------------
// Grid description used inside kernels (passed by value)
// ---------------------------------------------------------------------------
struct GridDev {
    float minx, miny, minz;
    float cell;
    uint32_t rx, ry, rz;
};

inline GridDev to_dev(const fsr_grid_info& g) {
    GridDev d;
    d.minx = g.min_pos.x; d.miny = g.min_pos.y; d.minz = g.min_pos.z;
    d.cell = g.cell_size;
    d.rx = g.resolution.x; d.ry = g.resolution.y; d.rz = g.resolution.z;
    return d;
}

inline uint64_t grid_count(const fsr_grid_info& g) {
    return (uint64_t)g.resolution.x * g.resolution.y * g.resolution.z;
}

// Words per x-row of a band bit mask.
__host__ __device__ inline uint32_t row_words(uint32_t rx) { return (rx + 31u) / 32u; }

// ---------------------------------------------------------------------------
// Device helpers
// ---------------------------------------------------------------------------
#ifdef __HIPCC__
// The reference's getIndex3D (ReconstructionCUDA.cuh:354-358): IEEE FP32
// subtract and divide, then truncation toward zero (int3(float3)).  Returns the
// FP32 quotient so callers can range-check it before truncating.
__device__ __forceinline__ float cell_coord(float p, float mn, float cell) {
    return (p - mn) / cell;
}

// The reference's getVertexPos (ReconstructionCUDA.cuh:366-370), evaluated as
// one fused multiply-add per axis (the contraction nvcc applies to it).
__device__ __forceinline__ float vertex_coord(uint32_t i, float mn, float cell) {
    return __builtin_fmaf((float)i, cell, mn);
}

__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }

// Number of active lanes below this one whose bit is set in `mask`.
__device__ __forceinline__ uint32_t mask_rank(uint64_t mask) {
    uint32_t lo = __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u);
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32), lo);
}

// Wave-wide inclusive sum of a u32 (64 lanes).
__device__ __forceinline__ uint32_t wave_inclusive_sum(uint32_t v) {
    const uint32_t lane = __lane_id();
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        uint32_t o = __shfl_up(v, d, 64);
        if (lane >= (uint32_t)d) v += o;
    }
    return v;
}

__device__ __forceinline__ uint32_t wave_sum(uint32_t v) {
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
    return v;
}

__device__ __forceinline__ uint32_t wave_or(uint32_t v) {
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) v |= __shfl_xor(v, d, 64);
    return v;
}

// Band bits of one mask word (vertices 32 w .. 32 w + 31 of a row) restricted
// to the vertex range [xlo, xhi).
__device__ __forceinline__ uint32_t clip_word(uint32_t word, uint32_t w, uint32_t xlo,
                                              uint32_t xhi) {
    const uint32_t lo = w * 32u, hi = lo + 32u;
    if (xlo > lo) word &= ~((1u << (xlo - lo)) - 1u);
    if (xhi < hi) word &= (1u << (xhi - lo)) - 1u;
    return word;
}

// XCD-aware dynamic work fetch for persistent kernels.  The work list [0, total)
// is cut into 64 contiguous sub-segments, 8 per XCD in list order (XCD x owns
// the x-th eighth of the list); workgroups are dispatched round-robin over the
// 8 XCDs, so blockIdx % 8 names the XCD.  A wave takes items from its XCD's
// sub-segments first (neighbouring items share particle data, which then stays
// in that XCD's L2), starting at the sub-segment its workgroup index picks, then
// from the other XCDs' eighths.  `step` (0 at start) is the wave's position in
// that order.  Same-address atomics serialise (about 30 ns each here), so the
// list has 64 counters, next64[32 k] on separate 128-byte lines, all starting at
// 0, and a sub-segment seen exhausted by a plain (possibly stale, never early)
// read is passed without an atomic.  Returns total when every item is taken.
// Wave-uniform.
constexpr uint32_t kFetchSegments = 64;

__device__ __forceinline__ uint32_t next_item_xcd(uint32_t* next64, uint32_t total,
                                                  uint32_t& step) {
    const uint32_t xcd = blockIdx.x & 7u;
    const uint32_t sub0 = (blockIdx.x >> 3) & 7u;
    const uint32_t seg = (total + kFetchSegments - 1u) / kFetchSegments;
    while (step < kFetchSegments) {
        const uint32_t k = ((xcd + (step >> 3)) & 7u) * 8u + ((sub0 + step) & 7u);
        const uint32_t lo = k * seg;
        const uint32_t hi = min(lo + seg, total);
        if (lo < hi) {
            uint32_t* ctr = &next64[32u * k];
            const uint32_t seen = __hip_atomic_load(ctr, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (seen < hi - lo) {
                uint32_t idx = 0;
                if (__lane_id() == 0) idx = atomicAdd(ctr, 1u);
                idx = __builtin_amdgcn_readfirstlane(idx);
                if (lo + idx < hi) return lo + idx;
            }
        }
        ++step;
    }
    return total;
}

// Adds a per-lane count into a global counter with one atomic per wave.
__device__ __forceinline__ void wave_count_add(uint32_t* counter, uint32_t v) {
    uint32_t s = wave_sum(v);
    if (__lane_id() == 0 && s) atomicAdd(counter, s);
}

// Adds two per-thread counts into two global counters with one atomic each per
// block (thousands of waves adding to one address serialise in L2).  Every
// thread of the block must call it: it synchronises the block.
__device__ __forceinline__ void block_count_add2(uint32_t* c0, uint32_t v0, uint32_t* c1,
                                                 uint32_t v1) {
    __shared__ uint32_t part[2][16];
    const uint32_t s0 = wave_sum(v0), s1 = wave_sum(v1);
    const uint32_t w = threadIdx.x >> 6;
    if (__lane_id() == 0) {
        part[0][w] = s0;
        part[1][w] = s1;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t t0 = 0, t1 = 0;
        for (uint32_t k = 0; k < (blockDim.x + 63u) / 64u; ++k) {
            t0 += part[0][k];
            t1 += part[1][k];
        }
        if (t0) atomicAdd(c0, t0);
        if (t1) atomicAdd(c1, t1);
    }
}
#endif

// ---------------------------------------------------------------------------
// Scratch arena: named device buffers that only grow.
// ---------------------------------------------------------------------------
enum ArenaSlot {
    A_PACKED = 0, A_SORTED, A_KEYS0, A_KEYS1, A_VALS0, A_VALS1, A_RADIX_HIST,
    A_RANGES, A_OCCUPANCY, A_SURF_CELL, A_SURF_FLAGS, A_SBITS, A_BAND_TMP,
    A_BAND_BITS, A_EXCEPTIONS, A_WORD_OFF, A_BAND_INDEX, A_AXIS_MAPS,
    A_CELL_CNT, A_CELL_OFF, A_CHUNK_OFF, A_CELL_BAND,
    A_CHUNK_CELL, A_NVERTS, A_VALID, A_VALID_OFF, A_VERT_OFF, A_VALID_IDX,
    A_COUNTERS, A_SCAN_PARTIALS, A_STATS, A_INVOLVE, A_INVOLVE_OFF,
    A_INVOLVED_IDX, A_MEAN, A_SMOOTH, A_GMAT, A_SIGMA, A_AUX0, A_AUX1, A_REC,
    A_ADJ_VERT, A_ADJ_COEF, A_ADJ_AUX, A_ADJ_TOUCH, A_EDGE_MASK, A_EDGE_OFF, A_SURF_CHUNKS, A_SLAB, A_BAND_PARTIALS,
    A_FIELD_WORK, A_COLOR_WORK, A_MOM_CELL, A_MOM_T1, A_MOM_T2,
    A_MOM_BLOCK, A_CELL_BOX, A_ANISO_WORK, A_CELL_BOX_RAW, A_ADJ_CORNER, A_INVOLVE_TMP, A_NUM_SLOTS
};

// Band vertices grouped by hash cell (band.hip).
struct BandCompactOut {
    uint32_t* band_index;   // u32[B]
    uint32_t* cell_off;     // per virtual cell offset into cell_band
    uint32_t* cell_cnt;
    uint32_t* chunk_off;
    uint32_t* cell_band;    // u32[B] cell-major, or null when not materialised:
                            // then a chunk's vertices are read from the mask
    uint32_t* chunk_cell;   // chunk -> virtual cell
    uint32_t vcx, vcy, vcz; // virtual cell grid (hash grid padded by `radius`)
    int radius;             // 1 (3^3 blocks) or 2 (5^3 blocks)
    uint64_t max_chunks;
    // what a kernel needs to list a chunk's vertices from the band mask: the
    // mask, its words per row, each virtual cell's vertex interval per axis
    // (axis: lo_x[vcx] hi_x[vcx] lo_y[vcy] hi_y[vcy] lo_z[vcz] hi_z[vcz]) and
    // the per-slab band counts (slab = one scalar layer of a cell; a cell's
    // slabs are consecutive, layer by layer)
    const uint32_t* bits;
    const uint32_t* axis;
    const uint32_t* slab;
    uint32_t wx;
};

struct Arena {
    void* ptr[A_NUM_SLOTS] = {};
    size_t bytes[A_NUM_SLOTS] = {};
    uint64_t count[A_NUM_SLOTS] = {};
};

// Internal device counters (reset by a fill kernel at frame start, so a frame
// can be captured in a graph without memset nodes).
struct Counters {
    uint32_t field_next_chunk;
    uint32_t num_chunks;
    uint32_t num_exceptions;
    uint32_t mc_next;
    uint32_t aux[12];
    uint32_t pad[12];
    // fetch counters of the persistent kernels (next_item_xcd), 128 B apart
    uint32_t xcd_next[kFetchSegments * 32];
    // the same for the ZB05 field kernel, used once per frame, so the frame's
    // initial clear is its reset
    uint32_t xcd_field[kFetchSegments * 32];
};

}  // namespace fsr

namespace fsr {
// Launch-shape classes of the persistent kernels (workgroups per CU), set by
// the autotuner.
enum TuneClass { TUNE_FIELD = 0, TUNE_MC = 1, TUNE_ANISO = 2, TUNE_BAND = 3, TUNE_NUM = 4 };

// What the last frame left for its adjoint (valid while no other frame runs on
// the context).
struct AdjointState {
    bool valid = false;
    int cpu_flavor = 0;                      // marching cubes flavour of the frame
    uint32_t n = 0;
    GridDev sp{}, sc{};
    float radius_sq = 0.f, iso = 0.f;
    const float4* sorted = nullptr;          // hash-sorted particles
    const fsr_index_range* ranges = nullptr;
    const uint32_t* perm = nullptr;          // sorted -> caller index
    BandCompactOut bc{};
    const float4* saved = nullptr;           // cell-major (x - xbar, sum w)
    const float* phi = nullptr;
    const uint32_t* bits = nullptr;          // band mask
    const uint32_t* word_off = nullptr;      // band rank of each mask word
    const uint32_t* nverts = nullptr;        // soup vertices of the cube at band slot b
    const uint32_t* vert_off = nullptr;      // their first soup index
    const uint32_t* valid_idx = nullptr;     // band slots of the valid cubes
    uint32_t tri_capacity = 0;
    const fsr_frame_stats* stats = nullptr;
    Counters* counters = nullptr;
};
}  // namespace fsr

struct fsr_context {
    int device = 0;
    int num_cus = 256;
    int last_sort_passes = 0;
    int tune[fsr::TUNE_NUM] = {0, 0, 0, 0};  // 0: kernel default
    bool capturing = false;                  // inside a graph capture: no arena growth
    bool save_adjoint = false;               // frames keep what the adjoint needs
    int surface_mode = 0;                    // fsr_surface_mode of ZB05 / Akinci12 frames
    fsr::AdjointState adj;
    fsr::Arena arena;
    // stage profiling: ring of (FSR_NUM_STAGES + 1) events per frame
    std::vector<hipEvent_t> prof_events;
    uint32_t prof_ring = 0;
    uint32_t prof_frames = 0;  // frames recorded since the last read
};

namespace fsr {
// Records stage boundary `k` (0..FSR_NUM_STAGES) of the current frame.
void prof_mark(fsr_context* ctx, int k, hipStream_t s);
}  // namespace fsr

namespace fsr {

fsr_status arena_get(fsr_context* ctx, ArenaSlot s, size_t bytes, void** out,
                     uint64_t count_for_report);

template <typename T>
fsr_status arena_typed(fsr_context* ctx, ArenaSlot s, uint64_t n, T** out) {
    void* p = nullptr;
    size_t b = (size_t)(n ? n : 1) * sizeof(T);
    FSR_TRY(arena_get(ctx, s, b, &p, n));
    *out = (T*)p;
    return FSR_OK;
}

// ---------------------------------------------------------------------------
// Launchers (defined in the .hip files)
// ---------------------------------------------------------------------------

// scan.hip ------------------------------------------------------------------
enum ScanInput { SCAN_U32 = 0, SCAN_POPC = 1, SCAN_CHUNKS64 = 2, SCAN_NONZERO = 3 };
// Exclusive scan of f(in[i]) for i < n, where n = *d_count when d_count is not
// null (bounded by n_max) and n_max otherwise.  Writes the total to *d_total
// when d_total is not null.
fsr_status scan_exclusive(fsr_context* ctx, ScanInput f, const uint32_t* in,
                          uint32_t* out, uint64_t n_max, const uint32_t* d_count,
                          uint32_t* d_total, hipStream_t s);
fsr_status fill_u32(uint32_t* p, uint32_t value, uint64_t n, hipStream_t s);
// Sets up the per-device scratch of the allocation-free scan (device address of
// the scratch array, looked up once per device).
fsr_status noalloc_scratch_init(int device);
// Exclusive scan of in[0, n) with its workgroup partials in that scratch.
fsr_status scan_exclusive_noalloc(int device, const uint32_t* in, uint32_t* out, uint64_t n,
                                  uint32_t* d_total, hipStream_t s);

// Cell-major work list offsets from the per-cell band counts cnt[n]: cell_off
// (exclusive scan), chunk_off (exclusive scan of the 64-vertex chunk counts),
// chunk_cell (chunk -> cell, at most max_chunks entries) and the chunk total.
fsr_status scan_cells(fsr_context* ctx, const uint32_t* cnt, uint64_t n, uint32_t* cell_off,
                      uint32_t* chunk_off, uint32_t* chunk_cell, uint64_t max_chunks,
                      uint32_t* d_num_chunks, hipStream_t s);

// Marching-cubes offsets: vert_off = exclusive scan of the per-cube vertex
// counts nv (live length *d_count, at most n_max) and valid_idx = the cubes
// with nv > 0 in ascending order; the totals go to the two device counters.
fsr_status scan_mc(fsr_context* ctx, const uint32_t* nv, uint64_t n_max, const uint32_t* d_count,
                   uint32_t* vert_off, uint32_t* valid_idx, uint32_t* d_total_verts,
                   uint32_t* d_total_valid, hipStream_t s);

// Several independent u32 fills in one launch.
constexpr int kMaxFillJobs = 8;
struct FillList {
    uint32_t* p[kMaxFillJobs];
    uint32_t value[kMaxFillJobs];
    uint64_t n[kMaxFillJobs];
    int count = 0;
    void add(void* ptr, uint32_t v, uint64_t words) {
        p[count] = static_cast<uint32_t*>(ptr);
        value[count] = v;
        n[count] = words;
        ++count;
    }
};
fsr_status fill_multi(const FillList& l, hipStream_t s);
// Zeroes the per-XCD fetch counters before a persistent kernel that uses them.
inline fsr_status reset_xcd(Counters* c, hipStream_t s) {
    return fill_u32(c->xcd_next, 0u, sizeof(c->xcd_next) / 4, s);
}

// sort.hip ------------------------------------------------------------------
struct HashOut {
    float4* sorted;        // A_SORTED
    uint32_t* keys;        // sorted cell ids
    uint32_t* perm;        // sorted -> original index
    fsr_index_range* ranges;
    float* occupancy;
};
// `pre` (may be null): fills of the caller's buffers, issued in the same launch
// as the hash grid's own initial fills.
fsr_status spatial_hash(fsr_context* ctx, float* pos, float* dens, uint32_t n,
                        const fsr_grid_info& sp, fsr_frame_stats* st,
                        HashOut* out, hipStream_t s, const FillList* pre = nullptr);

// surface.hip ---------------------------------------------------------------
struct SurfaceArgs {
    const float4* sorted;
    const uint32_t* keys;
    const fsr_index_range* ranges;
    const float* occupancy;
    uint32_t n;
    GridDev sp;
    GridDev sc;
    float h;
    uint32_t* flags;        // u32[N] surface flags (sorted order)
    uint32_t* sbits;        // seed mask: one byte per scalar cell of the surface
                            // particles, rows padded to whole band-mask words
                            // (32 bytes per word, 8 words of storage each)
    uint32_t* exceptions;   // particles routed to the exact band path
    Counters* counters;
    fsr_frame_stats* st;
    int all_exceptions;     // 1: every surface particle takes the exact path
    // 1: surface particles are those with fewer than 25 other particles at
    // squared distance (FP32) below h^2, instead of the method's own test
    int radius_count;
};
fsr_status surface_occupancy(fsr_context* ctx, const SurfaceArgs& a, uint8_t* surf_cell,
                             uint32_t* involve, hipStream_t s);
fsr_status surface_colorfield(fsr_context* ctx, const SurfaceArgs& a,
                              float mass, hipStream_t s);

// band.hip ------------------------------------------------------------------
struct BandArgs {
    const float4* sorted;
    const uint32_t* exceptions;
    uint32_t max_exceptions;
    const uint32_t* sbits;
    uint32_t* tmp;
    uint32_t* bits;
    GridDev sc;
    float h;
    Counters* counters;
    fsr_frame_stats* st;
    // When every surface particle takes the exact path, the exact pass walks
    // the surface flags in particle order (sorted: neighbours share a wave)
    // instead of the exception list, whose order is that of the appends.
    const uint32_t* flags;
    uint32_t num_particles;
    int all_exceptions;
};
fsr_status band_build(fsr_context* ctx, const BandArgs& a, hipStream_t s);
// Returns 1 when every vertex of the reference's band box lies within 1.5 h of
// the particle, so the band equals a box dilation of the surface cells.
int band_dilation_exact(const fsr_grid_info& sc, float h);

// materialize: also write the cell-major list (cell_band); otherwise the
// field kernel lists each chunk's vertices from the mask.
fsr_status band_compact(fsr_context* ctx, const uint32_t* bits, const GridDev& sc,
                        const GridDev& sp, int radius, Counters* counters,
                        fsr_frame_stats* st, BandCompactOut* out, hipStream_t s,
                        bool materialize = true);

// field_zb05.hip ------------------------------------------------------------
struct FieldArgs {
    const float4* sorted;
    const fsr_index_range* ranges;
    const BandCompactOut* bc;
    GridDev sp;
    GridDev sc;
    float radius_sq;       // R^2
    float particle_radius;
    float* phi;
    Counters* counters;
    float4* saved;         // optional, cell-major [B]: (x - xbar, sum w) for the adjoint
    uint32_t* work;        // set by field_zb05: per chunk, candidates staged, candidates
                           // kept by the box test, 16 x 16 tile steps evaluated
};
fsr_status field_zb05(fsr_context* ctx, const FieldArgs& a, hipStream_t s);
fsr_status fill_f32(float* p, float value, uint64_t n, hipStream_t s);

// mc.hip --------------------------------------------------------------------
struct McArgs {
    const uint32_t* band_index;
    const uint32_t* bits;
    const float* phi;
    GridDev sc;
    float iso;
    int cpu_flavor;        // 1: the reference's host MC semantics (Akinci12)
    uint32_t* nverts;
    uint32_t* valid;
    uint32_t* valid_off;
    uint32_t* vert_off;
    uint32_t* valid_idx;
    float* tri_pos;
    float* tri_nrm;
    uint32_t capacity;
    uint64_t max_band;
    fsr_frame_stats* st;
    // indexed (deduplicated) mode: capacity bounds the index count
    const uint32_t* word_off;  // band rank of each mask word
    uint32_t* edge_mask;       // [B] bit a: the edge (v, v + e_a) carries a mesh vertex
    uint32_t* edge_off;        // [B] first mesh vertex of band vertex b
    float* vtx_pos;            // float3[vtx_capacity]
    float* vtx_nrm;            // float3[vtx_capacity] (may be null)
    uint32_t vtx_capacity;
    uint32_t* indices;         // u32[capacity]
};
fsr_status mc_run(fsr_context* ctx, const McArgs& a, hipStream_t s);
// classify, scans, compaction, then the deduplicated mesh
fsr_status mc_run_indexed(fsr_context* ctx, const McArgs& a, hipStream_t s);
// single steps of mc_run (classify reads a.st->num_band_vertices, emit reads
// a.st->num_valid_cubes)
fsr_status mc_classify(fsr_context* ctx, const McArgs& a, hipStream_t s);
fsr_status mc_emit(fsr_context* ctx, const McArgs& a, hipStream_t s);

// aniso.hip -----------------------------------------------------------------
struct AnisoArgs {
    const float4* sorted;
    const uint32_t* keys;
    const fsr_index_range* ranges;
    const uint32_t* involve;      // u32[C] involve counts (0: not involved)
    uint32_t* involve_off;        // u32[C] exclusive scan of involve
    uint32_t n;
    GridDev sp, sc;
    float aniso_radius;           // covariance / smoothing radius r
    float lambda;                 // smoothing blend
    float h2, inv_h;              // smoothing radius squared, inverse
    float kr, kn;
    float mass;
    uint32_t min_neighbors;
    float4* mean;                 // [N] xyz
    float4* smooth;               // [N] xyz; w = squared culling radius once the field starts
    uint32_t* involved_idx;       // [I]
    fsr_matrix* gmat;             // [I]
    float4* sigma;                // [I] covariance eigenvalues, descending
    float4* rec;                  // [2N] field records (M = G^T G, coefficient)
    Counters* counters;
    fsr_frame_stats* st;
    int noninv_moments;           // 1: the field sums non-involved particles by moments
    // per hash cell, the box of its sorted positions for the cell test of the
    // smoothing and covariance (float4 lo[C], hi[C]); computed by the first
    // of those kernels when null
    const float4* cell_box = nullptr;
};
fsr_status aniso_run(fsr_context* ctx, const AnisoArgs& a, hipStream_t s);
// steps of aniso_run: smoothing (K10), involved list + non-involved field
// records (K8), covariance -> G (K11)
fsr_status aniso_smooth(fsr_context* ctx, const AnisoArgs& a, hipStream_t s);
fsr_status aniso_involved(fsr_context* ctx, const AnisoArgs& a, hipStream_t s);
fsr_status aniso_cov(fsr_context* ctx, const AnisoArgs& a, hipStream_t s);
fsr_status aniso_field(fsr_context* ctx, const AnisoArgs& a, const BandCompactOut& bc,
                       float* phi, hipStream_t s);
// Chunks of up to 64 particles per hash cell (cells with filter[c] == 0 are
// skipped when filter is not null): counts, exclusive chunk offsets, chunk ->
// cell map, total chunk count into *d_total.
fsr_status cell_chunk_list(fsr_context* ctx, const fsr_index_range* ranges,
                           const uint32_t* filter, uint32_t C, uint32_t* cnt,
                           uint32_t* chunk_off, uint32_t* chunk_cell, uint32_t* d_total,
                           hipStream_t s);

// adjoint.hip: backward passes over the state a frame left in ctx->adj
fsr_status mc_vertex_backward(fsr_context* ctx, const float* g_tri_pos, float* g_phi,
                              hipStream_t s);
fsr_status field_zb05_backward(fsr_context* ctx, const float* g_phi, float* g_pos,
                               hipStream_t s);

// Grid of a persistent kernel: CUs x workgroups per CU (the tuned value of
// `cls` when set, else `per_cu`).
int persistent_blocks(fsr_context* ctx, int threads, int per_cu, TuneClass cls);

// util.hip: writes the 100 MHz device clock (s_memrealtime) to *out
fsr_status device_timestamp(uint64_t* out, hipStream_t s);

}  // namespace fsr
