// ZB05 (Zhu & Bridson 2005) field on the surface band.
//
// Reference: updateScalarGridValuesCompacted -> updateScalarFieldValueZB05
// (ReconstructionCUDA.cu:308-324, :230-289) with kernelZB05
// (ReconstructionCUDA.cuh:262-268).  For a band vertex x in hash cell c, over
// the particles p of the clamped 3^3 cell block around c:
//   w(p) = (1 - |x - p|^2 / R^2)^3 for |x - p| < R, else 0
//   xbar = sum w p / sum w
//   phi(x) = |x - xbar| - r_p      if sum w > 0, otherwise phi is left as is
// (the caller's fill, 0).  The same field serves the Akinci12 method.
//
// MI355X mapping.  A wave owns a chunk of up to 64 band vertices of one hash
// cell (the cell-major list built in band.hip), so all lanes share one
// candidate block: the 9 x-runs of the 3^3 block (x-adjacent cells are
// contiguous in the sorted particle array).  The wave stages candidates in LDS
// (one coalesced 16-byte load per lane, one candidate per lane) and evaluates
// them against the chunk's vertices in 16 x 16 tiles on the matrix cores:
//   t = 1 - |x - p|^2/R^2 = c_x + a_x . p' - q_p
// with p' = p - o and x' = x - o relative to the cell corner o, a = 2 x'/R^2,
// c_x = 1 - |x'|^2/R^2 and q_p = |p'|^2/R^2, is a K = 4 product (one
// v_mfma_f32_16x16x4_f32 per tile), w = max(t, 0)^3 takes 2 VALU ops per
// pair, and the sums of w and w p' are rank-1 updates of 4 x 4 blocks
// (v_mfma_f32_4x4x1_16b_f32), so a pair costs 2 VALU ops where the scalar
// form needs 11.  Coordinates relative to the cell keep the expansion well
// conditioned and make |x - xbar| free of the cancellation the
// absolute-coordinate sum has.  Candidates are staged in the reference's
// order (z, y, x cells, then range order); each vertex's sums are formed as
// four interleaved partial sums that are added at the end.  Work is
// distributed dynamically, one chunk per atomic fetch, each XCD starting on
// its own contiguous eighth of the chunk list.
//
// Culling: a chunk's vertices are a compact slab of its cell (the cell-major
// list is in ascending vertex id), so many block candidates lie farther than R
// from the chunk's bounding box.  At staging every candidate is tested against
// that box with R^2 (1 + 1e-4); survivors are compacted in order (ballot rank)
// into a 128-entry LDS queue that is drained 64 at a time.  A culled candidate
// is at least 1.00005 R from every vertex of the chunk, so its computed t is
// negative (the rounding of t is below 1e-5 there) and its weight exactly 0:
// culling leaves every sum unchanged.
#include <type_traits>

#include "common.hpp"

namespace fsr {

namespace {

constexpr int kFieldThreads = 256;

__device__ __forceinline__ int clampi(int v, int lo, int hi) {
    return v < lo ? lo : (v > hi ? hi : v);
}

__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Lane exchange i <-> i ^ m within each half-wave (m < 32), with no address
// register: ds_swizzle in bit mode (and-mask 31, xor-mask m).
template <int M>
__device__ __forceinline__ float swz_xor(float v) {
    return __int_as_float(__builtin_amdgcn_ds_swizzle(__float_as_int(v), 0x1F | (M << 10)));
}

// Minimum / maximum over the wave, returned in a scalar register.
__device__ __forceinline__ float wave_min_u(float v) {
    v = fminf(v, swz_xor<1>(v));
    v = fminf(v, swz_xor<2>(v));
    v = fminf(v, swz_xor<4>(v));
    v = fminf(v, swz_xor<8>(v));
    v = fminf(v, swz_xor<16>(v));
    return fminf(__int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 0)),
                 __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 32)));
}

__device__ __forceinline__ float wave_max_u(float v) {
    v = fmaxf(v, swz_xor<1>(v));
    v = fmaxf(v, swz_xor<2>(v));
    v = fmaxf(v, swz_xor<4>(v));
    v = fmaxf(v, swz_xor<8>(v));
    v = fmaxf(v, swz_xor<16>(v));
    return fmaxf(__int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 0)),
                 __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 32)));
}

// Inclusive sum within each row of 16 lanes (DPP row shifts; lanes whose
// source lies before the row start add 0).
__device__ __forceinline__ uint32_t row16_inclusive_sum(uint32_t v) {
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x111, 0xF, 0xF, true);
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x112, 0xF, 0xF, true);
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x114, 0xF, 0xF, true);
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x118, 0xF, 0xF, true);
    return v;
}

// x, y, z of a sorted particle (one 12-byte load)
__device__ __forceinline__ float3 load_xyz(const float4* p) {
    return *reinterpret_cast<const float3*>(p);
}

// Division by a launch-constant divisor d with a multiply-high: the
// round-up method of Granlund and Montgomery (1994), exact for every 32-bit
// dividend.  m = floor(2^32 (2^l - d) / d) + 1 with l = ceil(log2 d).
struct FastDiv {
    uint32_t m, s1, s2;
};

FastDiv make_fast_div(uint32_t d) {
    const uint32_t l = d <= 1u ? 0u : 32u - (uint32_t)__builtin_clz(d - 1u);
    FastDiv f;
    f.m = (uint32_t)((((uint64_t)1 << l) - d) * ((uint64_t)1 << 32) / d + 1u);
    f.s1 = l < 1u ? l : 1u;
    f.s2 = l > 0u ? l - 1u : 0u;
    return f;
}

__device__ __forceinline__ uint32_t fast_div(uint32_t n, const FastDiv& f) {
    const uint32_t t = __umulhi(n, f.m);
    return (t + ((n - t) >> f.s1)) >> f.s2;
}

// Divisors of the vertex id: x extent and xy extent of the scalar grid.
struct VertexDiv {
    FastDiv rx, rxy;
};

// The vertex ids of chunk k of virtual cell v, listed from the band mask in
// the cell-major order (layer z, row y, then x ascending), when the cell-major
// list is not materialised: lane l < cnt receives the (64 k + l)-th band vertex
// of the cell.  The cell's layers (slabs) are consecutive in the slab counts;
// 64 layers at a time, those that overlap [64 k, 64 k + 64) are expanded word
// by word, each lane taking one (layer, row, word) of them, and the wanted
// ranks go to `scratch` (64 LDS words of the wave).
__device__ uint32_t chunk_vertex(const BandCompactOut& bc, uint32_t rx, uint32_t ry, uint32_t v,
                                 uint32_t k, uint32_t cnt, uint32_t* scratch) {
    const uint32_t lane = __lane_id();
    const uint32_t cxv = v % bc.vcx, cyv = (v / bc.vcx) % bc.vcy, czv = v / (bc.vcx * bc.vcy);
    const uint32_t* lo_x = bc.axis;
    const uint32_t* lo_y = lo_x + 2u * bc.vcx;
    const uint32_t* lo_z = lo_y + 2u * bc.vcy;
    const uint32_t zlo = lo_z[czv], nz = lo_z[bc.vcz + czv] - zlo;
    const uint64_t sbase = (uint64_t)bc.vcx * bc.vcy * zlo + ((uint64_t)cyv * bc.vcx + cxv) * nz;
    const uint32_t ylo = lo_y[cyv], ny = lo_y[bc.vcy + cyv] - ylo;
    const uint32_t xlo = lo_x[cxv], xhi = lo_x[bc.vcx + cxv];
    const uint32_t w0 = xlo >> 5, nw = ((xhi + 31u) >> 5) - w0;
    const uint32_t per = ny * nw;
    const uint32_t lo = 64u * k, hi = lo + 64u;
    // every slot starts empty, so a slot the expansion does not fill reads as
    // "no vertex" rather than as a stale id
    scratch[lane] = 0xFFFFFFFFu;
    wave_lds_sync();
    uint32_t before = 0;  // band vertices of the cell's layers before the group
    for (uint32_t g0 = 0; g0 < nz && before < hi; g0 += 64u) {
        const uint32_t sn = g0 + lane < nz ? bc.slab[sbase + g0 + lane] : 0u;
        const uint32_t sincl = before + wave_inclusive_sum(sn);
        const uint32_t sexcl = sincl - sn;
        const uint64_t hit = __ballot(g0 + lane < nz && sn != 0u && sexcl < hi && sincl > lo);
        if (hit) {
            // the overlapping layers of the group are consecutive: j0 .. j1
            const uint32_t j0 = (uint32_t)__builtin_ctzll(hit);
            const uint32_t j1 = 63u - (uint32_t)__builtin_clzll(hit);
            const uint32_t items = (j1 - j0 + 1u) * per;
            uint32_t rank0 = (uint32_t)__builtin_amdgcn_readlane((int)sexcl, (int)j0);
            for (uint32_t i0 = 0; i0 < items; i0 += 64u) {
                const uint32_t i = i0 + lane;
                uint32_t word = 0, base = 0;
                if (i < items) {
                    const uint32_t js = g0 + j0 + i / per, rem = i % per;
                    const uint32_t y = ylo + rem / nw, w = w0 + rem % nw;
                    const uint64_t row = (uint64_t)(zlo + js) * ry + y;
                    word = clip_word(bc.bits[row * bc.wx + w], w, xlo, xhi);
                    base = (uint32_t)(row * rx) + w * 32u;
                }
                const uint32_t c = (uint32_t)__popc(word);
                const uint32_t incl = wave_inclusive_sum(c);
                uint32_t r = rank0 + incl - c;
                while (word) {
                    if (r >= lo && r < hi) scratch[r - lo] = base + (uint32_t)__builtin_ctz(word);
                    ++r;
                    word &= word - 1u;
                }
                rank0 += (uint32_t)__builtin_amdgcn_readlane((int)incl, 63);
            }
        }
        before = (uint32_t)__builtin_amdgcn_readlane((int)sincl, 63);
    }
    wave_lds_sync();
    const uint32_t id = lane < cnt ? scratch[lane] : 0xFFFFFFFFu;
    wave_lds_sync();
    return id;
}

typedef float v4f __attribute__((ext_vector_type(4)));

// LDS queue of one wave, structure of arrays: rows x', y', z', a row of ones
// and a row of -q, kQueue entries each.  The first four rows start 144 floats
// apart, so the 16-entry slices that one matrix operand reads from them fall
// into distinct LDS banks.
constexpr int kQueue = 128;
constexpr int kRowStride = 144;
constexpr int kRowX = 0, kRowY = 144, kRowZ = 288, kRowOne = 432, kRowQ = 576;
constexpr int kQueueFloats = 704;

// Sums of w and w p' for the NT 16-vertex tiles of one chunk over n staged
// candidates (n a multiple of 16), on the matrix cores.  Per 16 candidates and
// vertex tile:
//   1. one v_mfma_f32_16x16x4_f32 gives t = c_x + a_x . p' - q_p for the
//      16 x 16 pairs: A = (x', y', z', 1) of the candidates (rows),
//      B = (a_x, a_y, a_z, c_x) of the vertices (columns), C = -q_p;
//      lane l holds t of candidates 4(l/16)+r (r = 0..3) for vertex l%16;
//   2. w = max(t, 0)^3 (2 VALU ops per value: t^3 with a clamp to [0, 1]);
//   3. four v_mfma_f32_4x4x1_16b_f32 (one per r) accumulate, in each 4 x 4
//      block b = l/4, the outer product of (x', y', z', 1) of candidate
//      4(b/4)+r with the weights of vertices 4(b%4)..+3, so lane l collects
//      sum w p' (x, y, z) and sum w for vertex l%16 over the candidates
//      congruent to l/16 in each group of 16.
// The caller adds the four partial sums of a vertex (lanes l%16 + 16m).
template <int NT>
__device__ __forceinline__ void accumulate(const float* __restrict__ rows, uint32_t n,
                                           const float (&b1)[NT], v4f (&acc)[NT]) {
    const uint32_t lane = __lane_id();
    const float* a1p = rows + (lane >> 4) * kRowStride + (lane & 15u);
    const float* a2p = rows + (lane & 3u) * kRowStride + 4u * (lane >> 4);
    const float* cqp = rows + kRowQ + 4u * (lane >> 4);
    for (uint32_t tb = 0; tb < n; tb += 16u) {
        const float a1 = a1p[tb];
        const float4 a2 = *reinterpret_cast<const float4*>(a2p + tb);
        const float4 cq = *reinterpret_cast<const float4*>(cqp + tb);
        const v4f c1 = v4f{cq.x, cq.y, cq.z, cq.w};
        // the tiles' t products back to back, then the weights, then the
        // rank-1 updates (independent across tiles, so no chain stalls)
        v4f w[NT];
#pragma unroll
        for (int T = 0; T < NT; ++T) w[T] = __builtin_amdgcn_mfma_f32_16x16x4f32(a1, b1[T], c1, 0, 0, 0);
#pragma unroll
        for (int T = 0; T < NT; ++T) {
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                // max(t, 0)^3 as t^3 clamped to [0, 1] (the clamp bit of the
                // last multiply): t <= 1 up to rounding, and t^3 < 0 for t < 0
                const float t = w[T][r];
                w[T][r] = __builtin_amdgcn_fmed3f(t * t * t, 0.0f, 1.0f);
            }
        }
#pragma unroll
        for (int T = 0; T < NT; ++T) acc[T] = __builtin_amdgcn_mfma_f32_4x4x1f32(a2.x, w[T][0], acc[T], 0, 0, 0);
#pragma unroll
        for (int T = 0; T < NT; ++T) acc[T] = __builtin_amdgcn_mfma_f32_4x4x1f32(a2.y, w[T][1], acc[T], 0, 0, 0);
#pragma unroll
        for (int T = 0; T < NT; ++T) acc[T] = __builtin_amdgcn_mfma_f32_4x4x1f32(a2.z, w[T][2], acc[T], 0, 0, 0);
#pragma unroll
        for (int T = 0; T < NT; ++T) acc[T] = __builtin_amdgcn_mfma_f32_4x4x1f32(a2.w, w[T][3], acc[T], 0, 0, 0);
        // keep that order in the schedule: NT matrix ops, 8 NT VALU ops, 4 NT
        // matrix ops (VALU and matrix ops of one SIMD do not overlap, so
        // fewer switches between them pay)
        __builtin_amdgcn_sched_group_barrier(0x008, NT, 0);
        __builtin_amdgcn_sched_group_barrier(0x002, 8 * NT, 0);
        __builtin_amdgcn_sched_group_barrier(0x008, 4 * NT, 0);
    }
}

__global__ __launch_bounds__(kFieldThreads, 6) void k_field_zb05(FieldArgs a, BandCompactOut bc,
                                                              VertexDiv vd) {
    // one LDS block per wave, so every per-lane LDS address is one base
    // register plus an immediate offset
    struct WaveLds {
        float queue[kQueueFloats];
        // per lane: x', y', z' of its vertex (for the end) and (a_x, a_y, a_z, c_x)
        float vtx[7][64];
    };
    __shared__ __align__(16) WaveLds lds[kFieldThreads / 64];
    const uint32_t lane = __lane_id();
    const uint32_t wave = threadIdx.x >> 6;
    const uint32_t num_chunks = a.counters->num_chunks;
    const float inv_r2 = __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(1.0f / a.radius_sq)));
    const float cull_r2 = a.radius_sq * 1.0001f;
    const int r = bc.radius;
    float* q = lds[wave].queue;
    q[kRowOne + lane] = 1.0f;
    q[kRowOne + 64 + lane] = 1.0f;
    // the work record ends with a marker after the last chunk
    if (blockIdx.x == 0 && threadIdx.x == 0) a.work[3u * num_chunks] = 0xFFFFFFFFu;

    uint32_t step = 0;
    for (;;) {
        const uint32_t ch = next_item_xcd(a.counters->xcd_field, num_chunks, step);
        if (ch >= num_chunks) break;

        const uint32_t v = bc.chunk_cell[ch];
        const int cx = (int)(v % bc.vcx) - r;
        const int cy = (int)((v / bc.vcx) % bc.vcy) - r;
        const int cz = (int)(v / (bc.vcx * bc.vcy)) - r;
        const uint32_t k = ch - bc.chunk_off[v];
        const uint32_t ccount = bc.cell_cnt[v];
        const uint32_t first = bc.cell_off[v] + 64u * k;
        const uint32_t cnt = min(64u, ccount - 64u * k);
        // the chunk's vertex ids (before the candidate rows, whose registers
        // live through the staging loop)
        uint32_t vid = 0;
        if (bc.cell_band) {
            if (lane < cnt) vid = bc.cell_band[first + lane];
        } else {
            vid = chunk_vertex(bc, a.sc.rx, a.sc.ry, v, k, cnt, reinterpret_cast<uint32_t*>(q));
        }

        // candidate rows of the clamped 3^3 block
        const int x0 = clampi(cx - 1, 0, a.sp.rx - 1), x1 = clampi(cx + 1, 0, a.sp.rx - 1);
        const int y0 = clampi(cy - 1, 0, a.sp.ry - 1), y1 = clampi(cy + 1, 0, a.sp.ry - 1);
        const int z0 = clampi(cz - 1, 0, a.sp.rz - 1), z1 = clampi(cz + 1, 0, a.sp.rz - 1);
        const int ny = y1 - y0 + 1;
        const int nrows = ny * (z1 - z0 + 1);
        uint32_t len = 0, rs = 0;
        if ((int)lane < nrows) {
            const int yy = y0 + (int)lane % ny, zz = z0 + (int)lane / ny;
            const uint32_t rowc = ((uint32_t)zz * a.sp.ry + yy) * a.sp.rx;
            // the row's 3 cells as clamped offsets (a repeated edge cell leaves
            // the run unchanged), loaded together
            fsr_index_range rg[3];
#pragma unroll
            for (int d = 0; d < 3; ++d) rg[d] = a.ranges[rowc + clampi(cx - 1 + d, x0, x1)];
            uint32_t s = 0xFFFFFFFFu, e = 0;
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                if (rg[d].start != 0xFFFFFFFFu) {
                    if (s == 0xFFFFFFFFu) s = rg[d].start;
                    e = rg[d].end;
                }
            }
            if (s != 0xFFFFFFFFu) { rs = s; len = e - s; }
        }
        // nrows <= 9: the prefix over the first 16 lanes is the whole prefix.
        // Lane j < nrows keeps the first candidate index of row j (rp) and the
        // offset from candidate index to sorted index in that row (ro).
        const uint32_t incl = row16_inclusive_sum(len);
        const uint32_t total = (uint32_t)__builtin_amdgcn_readlane((int)incl, 15);
        const uint32_t rp = incl - len;
        const uint32_t ro = rs - rp;

        // cell corner as the local origin
        const float ox = __builtin_fmaf((float)cx, a.sp.cell, a.sp.minx);
        const float oy = __builtin_fmaf((float)cy, a.sp.cell, a.sp.miny);
        const float oz = __builtin_fmaf((float)cz, a.sp.cell, a.sp.minz);

        float vx = 0.f, vy = 0.f, vz = 0.f;
        // lanes with a vertex (an id past the grid would mean an empty slot)
        const bool has_vertex = lane < cnt && vid < a.sc.rx * a.sc.ry * a.sc.rz;
        if (has_vertex) {
            const uint32_t iz = fast_div(vid, vd.rxy);
            const uint32_t rem = vid - iz * (a.sc.rx * a.sc.ry);
            const uint32_t iy = fast_div(rem, vd.rx);
            const uint32_t ix = rem - iy * a.sc.rx;
            vx = vertex_coord(ix, a.sc.minx, a.sc.cell) - ox;
            vy = vertex_coord(iy, a.sc.miny, a.sc.cell) - oy;
            vz = vertex_coord(iz, a.sc.minz, a.sc.cell) - oz;
        }
        float* vt = &lds[wave].vtx[0][0];
        vt[lane] = vx;
        vt[64 + lane] = vy;
        vt[128 + lane] = vz;
        vt[192 + lane] = 2.0f * vx * inv_r2;
        vt[256 + lane] = 2.0f * vy * inv_r2;
        vt[320 + lane] = 2.0f * vz * inv_r2;
        vt[384 + lane] = 1.0f - __builtin_fmaf(vz, vz, __builtin_fmaf(vy, vy, vx * vx)) * inv_r2;
        // bounding box of the chunk's vertices
        const bool live = has_vertex;
        const float bx0 = wave_min_u(live ? vx : 3.0e38f), bx1 = wave_max_u(live ? vx : -3.0e38f);
        const float by0 = wave_min_u(live ? vy : 3.0e38f), by1 = wave_max_u(live ? vy : -3.0e38f);
        const float bz0 = wave_min_u(live ? vz : 3.0e38f), bz1 = wave_max_u(live ? vz : -3.0e38f);
        wave_lds_sync();
        const int nt = (int)((cnt + 15u) >> 4);
        // The rest of the chunk, compiled once per tile count NT, so that the
        // accumulators keep fixed registers across the staging loop.
        auto run = [&](auto ntc) {
            constexpr int NT = decltype(ntc)::value;
            // matrix operand of the vertex tiles: lane l carries component l/16
            // of (a_x, a_y, a_z, c_x) of vertex 16T + l%16
            float b1[NT];
            v4f acc[NT];
            const float* bsrc = vt + 192 + 64 * (lane >> 4) + (lane & 15u);
#pragma unroll
            for (int T = 0; T < NT; ++T) {
                b1[T] = bsrc[16 * T];
                acc[T] = v4f{0.f, 0.f, 0.f, 0.f};
            }
            // the queue is a ring of kQueue entries; entries [head, head + fill)
            // wait for evaluation, and head is 0 or 64
            uint32_t head = 0, fill = 0;
            // work of the chunk: candidates kept by the box test, tile steps
            uint32_t n_kept = 0, n_steps = 0;
            // this lane's candidate of the next round is loaded one round ahead,
            // so the load overlaps the culling and the accumulation of the
            // current one
            //
            // Row lookup: jc is a row whose first candidate index is at most
            // the first index nb of the round being located (rows only grow
            // with the index).  The rows that start inside the round are
            // visited with wave-uniform reads of rp and ro; a lane takes the
            // last one that starts at or before its index gg.  Empty rows
            // share their first index with the next row, which is visited
            // after them.
            uint32_t jc = 0;
            auto row_offset = [&](uint32_t nb, uint32_t gg) {
                uint32_t off = (uint32_t)__builtin_amdgcn_readlane((int)ro, (int)jc);
                for (uint32_t jj = jc + 1u; jj < (uint32_t)nrows; ++jj) {
                    const uint32_t pj = (uint32_t)__builtin_amdgcn_readlane((int)rp, (int)jj);
                    if (pj >= nb + 64u) break;
                    if (gg >= pj) off = (uint32_t)__builtin_amdgcn_readlane((int)ro, (int)jj);
                    jc = jj;
                }
                return off;
            };
            float3 next = make_float3(0.f, 0.f, 0.f);
            {
                const uint32_t off = row_offset(0u, lane);
                if (lane < total) next = load_xyz(a.sorted + (off + lane));
            }
            for (uint32_t base = 0; base < total; base += 64u) {
                const uint32_t g = base + lane;
                const float3 p = next;
                if (base + 64u < total) {
                    const uint32_t gn = g + 64u;
                    const uint32_t off = row_offset(base + 64u, gn);
                    if (gn < total) next = load_xyz(a.sorted + (off + gn));
                }
                bool keep = false;
                float px = 0.f, py = 0.f, pz = 0.f;
                if (g < total) {
                    px = p.x - ox;
                    py = p.y - oy;
                    pz = p.z - oz;
                    // distance to the box per axis: p - clamp(p, lo, hi), whose
                    // square equals that of max(lo - p, p - hi, 0) exactly
                    const float dx = px - __builtin_amdgcn_fmed3f(px, bx0, bx1);
                    const float dy = py - __builtin_amdgcn_fmed3f(py, by0, by1);
                    const float dz = pz - __builtin_amdgcn_fmed3f(pz, bz0, bz1);
                    keep = __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, dx * dx)) < cull_r2;
                }
                const uint64_t m = __ballot(keep);
                if (keep) {
                    const uint32_t slot = (head + fill + mask_rank(m)) & (kQueue - 1);
                    q[kRowX + slot] = px;
                    q[kRowY + slot] = py;
                    q[kRowZ + slot] = pz;
                    q[kRowQ + slot] =
                        -(__builtin_fmaf(pz, pz, __builtin_fmaf(py, py, px * px)) * inv_r2);
                }
                fill += (uint32_t)__popcll(m);
                n_kept += (uint32_t)__popcll(m);
                if (fill >= 64u) {
                    n_steps += 4u * NT;
                    wave_lds_sync();
                    accumulate<NT>(q + head, 64u, b1, acc);
                    wave_lds_sync();
                    head ^= 64u;
                    fill -= 64u;
                }
            }
            if (fill) {
                // pad to a multiple of 16 with entries of weight 0 (t < 0 for
                // every vertex); head + 64 <= kQueue, so the pad does not wrap
                const uint32_t n = (fill + 15u) & ~15u;
                if (lane >= fill && lane < n) {
                    q[kRowX + head + lane] = 0.f;
                    q[kRowY + head + lane] = 0.f;
                    q[kRowZ + head + lane] = 0.f;
                    q[kRowQ + head + lane] = -1.0e30f;
                }
                n_steps += (n >> 4) * NT;
                wave_lds_sync();
                accumulate<NT>(q + head, n, b1, acc);
                wave_lds_sync();
            }
            // add the four partial sums of each vertex; lane l then holds the
            // sums of vertex 16T + l%16 in acc[T]
            v4f s4 = acc[0];
            const uint32_t my = lane >> 4;
#pragma unroll
            for (int T = 0; T < NT; ++T) {
#pragma unroll
                for (int c = 0; c < 4; ++c) {
                    float x = acc[T][c];
                    x += swz_xor<16>(x);
                    x += __shfl_xor(x, 32, 64);
                    acc[T][c] = x;
                }
                if (my == (uint32_t)T) s4 = acc[T];
            }
            if (has_vertex) {
                float4 keep_out = make_float4(0.f, 0.f, 0.f, 0.f);
                const float sw = s4[3];
                if (sw > 0.0f) {
                    const float inv = 1.0f / sw;
                    const float dx = vt[lane] - s4[0] * inv;
                    const float dy = vt[64 + lane] - s4[1] * inv;
                    const float dz = vt[128 + lane] - s4[2] * inv;
                    a.phi[vid] =
                        __builtin_sqrtf(__builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, dx * dx))) -
                        a.particle_radius;
                    keep_out = make_float4(dx, dy, dz, sw);
                }
                // for the adjoint: x - xbar and sum w, at the vertex's cell-major slot
                if (a.saved) a.saved[first + lane] = keep_out;
            }
            if (lane == 0) {
                uint32_t* w = a.work + 3u * ch;
                w[0] = total;
                w[1] = n_kept;
                w[2] = n_steps;
            }
        };
        switch (nt) {
            case 1: run(std::integral_constant<int, 1>{}); break;
            case 2: run(std::integral_constant<int, 2>{}); break;
            case 3: run(std::integral_constant<int, 3>{}); break;
            default: run(std::integral_constant<int, 4>{}); break;
        }
    }
}

}  // namespace

fsr_status field_zb05(fsr_context* ctx, const FieldArgs& a, hipStream_t s) {
    // a.counters was cleared when the frame (or the stage call) began, and no
    // other kernel fetches through xcd_field
    // one workgroup per resident slot: 6 per CU (6 waves per SIMD); workgroups
    // beyond the resident ones start late and upset the per-XCD fetch order
    const int blocks = persistent_blocks(ctx, kFieldThreads, 6, TUNE_FIELD);
    FieldArgs fa = a;
    FSR_TRY(arena_typed(ctx, A_FIELD_WORK, 3ull * a.bc->max_chunks + 3, &fa.work));
    VertexDiv vd;
    vd.rx = make_fast_div(a.sc.rx);
    vd.rxy = make_fast_div(a.sc.rx * a.sc.ry);
    hipLaunchKernelGGL(k_field_zb05, dim3(blocks), dim3(kFieldThreads), 0, s, fa, *a.bc, vd);
    FSR_LAUNCH_CHECK("k_field_zb05");
    return FSR_OK;
}

}  // namespace fsr
