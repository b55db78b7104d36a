// Surface band of scalar vertices, its compaction, and the cell-major work
// list of the field kernels.
//
// The reference (estimationOfSurfaceVertices, ReconstructionCUDA.cu:554-595)
// loops, for every surface particle in scalar cell (i, j, k), over the vertices
//   x in [clamp(i-3), clamp(i+3))   (upper bound exclusive)
//   y in [clamp(j-3), clamp(j+3)],  z in [clamp(k-3), clamp(k+3)]
// (clamp to [0, res-1]) and marks those with FP32 |v - p| < 1.5 h (compared in
// double) in a dense u32[V] grid, which is then scanned (thrust) and compacted
// (compactSurfaceVertex, ReconstructionCUDA.cu:293-304).
//
// Here the band is a bit mask (V/8 bytes instead of 4V).  When every vertex of
// such a box lies within 1.5 h of any point of the particle's scalar cell
// (4*sqrt(3) scalar cells < 1.5 h, checked on the host with a rounding margin),
// the distance test never rejects and the band of in-grid particles is exactly
// the dilation of the set S of their scalar cells by the box
//   x - s_x in [-3, 2] (and x <= resx-2),  |y - s_y| <= 3,  |z - s_z| <= 3,
// done as three separable passes over the mask.  Particles outside the scalar
// grid, and every particle when the condition fails, go through the exact
// per-particle loop of the reference (k_band_exceptions).
//
// Compaction: popcount per word -> exclusive scan -> each word writes its set
// bits in ascending order (the reference's surfaceVerticesIndexArray).
//
// Cell-major list: the field kernels process the band grouped by the hash cell
// each vertex falls in (getIndex3D of the vertex position, exactly as the
// reference's field kernels compute it), so that all 64 lanes of a wave share
// one neighbour-cell block.  Cells outside the hash grid are folded into a ring
// of `radius` virtual cells on each side (their clamped neighbour blocks
// coincide).  The order of vertices inside a cell's list does not affect any
// value: every vertex sums its candidates in the cell's fixed order.
#include <algorithm>
#include <cmath>

#include "common.hpp"

namespace fsr {

namespace {

// Bits of four 0/1 bytes (byte k -> bit k).
__device__ __forceinline__ uint32_t pack4(uint32_t x) {
    return (x & 1u) | ((x >> 7) & 2u) | ((x >> 14) & 4u) | ((x >> 21) & 8u);
}

// The x pass reads the seed mask as bytes (one per scalar cell, rows padded to
// whole words: word w covers bytes 32w .. 32w + 31) and writes bits.
__global__ __launch_bounds__(256) void k_dilate_x(const uint8_t* __restrict__ in,
                                                  uint32_t* __restrict__ out, uint32_t wx,
                                                  uint64_t nwords, uint32_t rx) {
    const uint64_t w = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= nwords) return;
    const uint32_t xw = (uint32_t)(w % wx);
    const uint4* chunk = reinterpret_cast<const uint4*>(in + (w << 5));
    const uint4 lo = chunk[0], hi = chunk[1];
    const uint32_t cur = pack4(lo.x) | (pack4(lo.y) << 4) | (pack4(lo.z) << 8) |
                         (pack4(lo.w) << 12) | (pack4(hi.x) << 16) | (pack4(hi.y) << 20) |
                         (pack4(hi.z) << 24) | (pack4(hi.w) << 28);
    // the two top bits of the previous word and the three low bits of the next
    const uint32_t prev =
        xw > 0 ? pack4(reinterpret_cast<const uint32_t*>(in + (w << 5))[-1]) << 28 : 0u;
    const uint32_t next = xw + 1 < wx ? pack4(reinterpret_cast<const uint32_t*>(in + (w << 5))[8]) : 0u;
    uint32_t o = cur;
    o |= (cur >> 1) | (next << 31);
    o |= (cur >> 2) | (next << 30);
    o |= (cur >> 3) | (next << 29);
    o |= (cur << 1) | (prev >> 31);
    o |= (cur << 2) | (prev >> 30);
    // the reference's exclusive x bound after clamping: x <= rx - 2
    const uint32_t lim = rx >= 1 ? rx - 1 : 0;
    const uint32_t x0 = xw * 32u;
    if (x0 >= lim) o = 0;
    else if (lim - x0 < 32u) o &= (1u << (lim - x0)) - 1u;
    out[w] = o;
}

// out(w, row) = OR of in over rows row + d*stride, |d| <= 3, inside [0, extent)
__global__ __launch_bounds__(256) void k_dilate_rows(const uint32_t* __restrict__ in,
                                                     uint32_t* __restrict__ out,
                                                     uint64_t nwords, uint64_t stride,
                                                     uint32_t extent, uint64_t period) {
    const uint64_t w = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= nwords) return;
    const uint32_t c = (uint32_t)((w / stride) % period);
    uint32_t o = 0;
#pragma unroll
    for (int d = -3; d <= 3; ++d) {
        const int cc = (int)c + d;
        if (cc >= 0 && cc < (int)extent) o |= in[(uint64_t)((int64_t)w + (int64_t)d * (int64_t)stride)];
    }
    out[w] = o;
}

// x dilation of one band-mask word from the byte seed mask (the body of
// k_dilate_x): w = word index in the row, rowbase = first byte of the row.
__device__ __forceinline__ uint32_t dilate_x_word(const uint8_t* __restrict__ in, uint64_t rowbase,
                                                  uint32_t xw, uint32_t wx, uint32_t rx) {
    const uint8_t* src = in + rowbase + ((uint64_t)xw << 5);
    const uint4* chunk = reinterpret_cast<const uint4*>(src);
    const uint4 lo = chunk[0], hi = chunk[1];
    const uint32_t cur = pack4(lo.x) | (pack4(lo.y) << 4) | (pack4(lo.z) << 8) |
                         (pack4(lo.w) << 12) | (pack4(hi.x) << 16) | (pack4(hi.y) << 20) |
                         (pack4(hi.z) << 24) | (pack4(hi.w) << 28);
    const uint32_t prev = xw > 0 ? pack4(reinterpret_cast<const uint32_t*>(src)[-1]) << 28 : 0u;
    const uint32_t next = xw + 1 < wx ? pack4(reinterpret_cast<const uint32_t*>(src)[8]) : 0u;
    uint32_t o = cur;
    o |= (cur >> 1) | (next << 31);
    o |= (cur >> 2) | (next << 30);
    o |= (cur >> 3) | (next << 29);
    o |= (cur << 1) | (prev >> 31);
    o |= (cur << 2) | (prev >> 30);
    const uint32_t lim = rx >= 1 ? rx - 1 : 0;
    const uint32_t x0 = xw * 32u;
    if (x0 >= lim) o = 0;
    else if (lim - x0 < 32u) o &= (1u << (lim - x0)) - 1u;
    return o;
}

// The three dilation passes in one launch: a workgroup owns a tile of kTW
// words x kTY rows x kTZ layers of the output; it x-dilates the tile's input
// with a halo of 3 rows and 3 layers on each side into LDS (from the byte
// mask), ORs the 7 rows (y), then the 7 layers (z), and writes the tile.
// Rows and layers outside the grid contribute nothing, as in the separate
// passes.  1024 threads per tile: a thread loads the tile's 3,872 input
// words in 4 rounds rather than 16, each round waiting for its loads.
constexpr int kTW = 8, kTY = 16, kTZ = 16, kH = 3;
constexpr int kDilateThreads = 1024;
__global__ __launch_bounds__(kDilateThreads) void k_dilate_box(const uint8_t* __restrict__ in,
                                                    uint32_t* __restrict__ out, uint32_t wx,
                                                    GridDev sc) {
    __shared__ uint32_t xs[kTZ + 2 * kH][kTY + 2 * kH][kTW];
    __shared__ uint32_t ys[kTZ + 2 * kH][kTY][kTW];
    const uint32_t w0 = blockIdx.x * kTW;
    const int y0 = (int)(blockIdx.y * kTY), z0 = (int)(blockIdx.z * kTZ);
    constexpr int nxs = (kTZ + 2 * kH) * (kTY + 2 * kH) * kTW;
    for (int t = threadIdx.x; t < nxs; t += blockDim.x) {
        const int w = t % kTW, r = t / kTW;
        const int yy = y0 - kH + r % (kTY + 2 * kH), zz = z0 - kH + r / (kTY + 2 * kH);
        uint32_t v = 0;
        if (w0 + w < wx && yy >= 0 && yy < (int)sc.ry && zz >= 0 && zz < (int)sc.rz)
            v = dilate_x_word(in, (((uint64_t)zz * sc.ry + yy) * wx) << 5, w0 + w, wx, sc.rx);
        xs[r / (kTY + 2 * kH)][r % (kTY + 2 * kH)][w] = v;
    }
    __syncthreads();
    constexpr int nys = (kTZ + 2 * kH) * kTY * kTW;
    for (int t = threadIdx.x; t < nys; t += blockDim.x) {
        const int w = t % kTW, y = (t / kTW) % kTY, zr = t / (kTW * kTY);
        uint32_t v = 0;
#pragma unroll
        for (int d = 0; d <= 2 * kH; ++d) v |= xs[zr][y + d][w];
        ys[zr][y][w] = v;
    }
    __syncthreads();
    constexpr int nout = kTZ * kTY * kTW;
    for (int t = threadIdx.x; t < nout; t += blockDim.x) {
        const int w = t % kTW, y = (t / kTW) % kTY, z = t / (kTW * kTY);
        const int yy = y0 + y, zz = z0 + z;
        if (w0 + w >= wx || yy >= (int)sc.ry || zz >= (int)sc.rz) continue;
        uint32_t v = 0;
#pragma unroll
        for (int d = 0; d <= 2 * kH; ++d) v |= ys[z + d][y][w];
        out[((uint64_t)zz * sc.ry + yy) * wx + w0 + w] = v;
    }
}

__device__ __forceinline__ int clampi(int v, int lo, int hi) {
    return v < lo ? lo : (v > hi ? hi : v);
}

__device__ __forceinline__ int trunc_clamped(float q, int res) {
    if (!(q == q)) return 0;
    const float lo = -4.0f, hi = (float)res + 4.0f;
    q = q < lo ? lo : (q > hi ? hi : q);
    return (int)q;
}

// Rows of one exception particle's box (the reference's per-particle loop):
// calls emit(z, y, word, bits) for every row word with vertices in the band.
// The x-run of a row has at most 6 vertices (x in [clamp(i-3), clamp(i+3)))
// and so touches at most two words; its 6 tests are unrolled and branch-free,
// so they run as independent chains.
template <typename Emit>
__device__ __forceinline__ void exception_rows(const BandArgs& a, const float4& p, int x0, int x1,
                                               int y0, int y1, int z0, int z1, float limit,
                                               Emit emit) {
    const uint32_t w0 = (uint32_t)x0 >> 5;
    for (int z = z0; z <= z1; ++z) {
        const float vz = vertex_coord((uint32_t)z, a.sc.minz, a.sc.cell);
        const float dz = vz - p.z;
        for (int y = y0; y <= y1; ++y) {
            const float vy = vertex_coord((uint32_t)y, a.sc.miny, a.sc.cell);
            const float dy = vy - p.y;
            uint32_t m0 = 0, m1 = 0;
#pragma unroll
            for (int k = 0; k < 6; ++k) {
                const int x = x0 + k;
                const float vx = vertex_coord((uint32_t)x, a.sc.minx, a.sc.cell);
                const float dx = vx - p.x;
                // the reference's |v - p| (same operation order as fma(dz, dz, fma(dy, dy, dx dx)))
                const float dist =
                    __builtin_sqrtf(__builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, dx * dx)));
                const uint32_t bit = (x < x1 && dist < limit) ? 1u << ((uint32_t)x & 31u) : 0u;
                if (((uint32_t)x >> 5) == w0) m0 |= bit;
                else m1 |= bit;
            }
            if (m0) emit(z, y, w0, m0);
            if (m1) emit(z, y, w0 + 1u, m1);
        }
    }
}

// The same tests with wave-uniform loops (the box has at most 7 x 7 x 6
// vertices), for the windowed path: every lane runs all 294 slots, invalid
// slots test nothing, and each row ORs its nonzero bits into the LDS window.
__device__ __forceinline__ void exception_rows_window(const BandArgs& a, const float4& p,
                                                      bool live, int x0, int x1, int y0, int y1,
                                                      int z0, int z1, float limit, uint32_t* w,
                                                      int W0, int Y0, int Z0, int nw, int ny) {
    const uint32_t w0 = (uint32_t)x0 >> 5;
    for (int kz = 0; kz < 7; ++kz) {
        const int z = z0 + kz;
        const bool zok = live && z <= z1;
        const float dz = vertex_coord((uint32_t)z, a.sc.minz, a.sc.cell) - p.z;
#pragma unroll
        for (int ky = 0; ky < 7; ++ky) {
            const int y = y0 + ky;
            const bool ok = zok && y <= y1;
            const float dy = vertex_coord((uint32_t)y, a.sc.miny, a.sc.cell) - p.y;
            uint32_t m0 = 0, m1 = 0;
#pragma unroll
            for (int kx = 0; kx < 6; ++kx) {
                const int x = x0 + kx;
                const float dx = vertex_coord((uint32_t)x, a.sc.minx, a.sc.cell) - p.x;
                const float dist =
                    __builtin_sqrtf(__builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, dx * dx)));
                const uint32_t bit = (x < x1 && dist < limit) ? 1u << ((uint32_t)x & 31u) : 0u;
                if (((uint32_t)x >> 5) == w0) m0 |= bit;
                else m1 |= bit;
            }
            if (ok && (m0 | m1)) {
                const int base = ((z - Z0) * ny + (y - Y0)) * nw + ((int)w0 - W0);
                if (m0) atomicOr(&w[base], m0);
                if (m1) atomicOr(&w[base + 1], m1);  // m1 != 0: word w0 + 1 is in the window
            }
        }
    }
}

__device__ __forceinline__ int wave_min_i(int v) {
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) v = min(v, __shfl_xor(v, d, 64));
    return v;
}

__device__ __forceinline__ int wave_max_i(int v) {
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) v = max(v, __shfl_xor(v, d, 64));
    return v;
}

__device__ __forceinline__ void wave_sync_lds() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// The reference's per-particle band loop for the exception list, one wave per
// 64 exceptions.  Device-scope atomics are slow here, so a wave whose boxes
// fit a window of kWinWords x kWinRows x kWinRows mask words (the common case:
// the list is filled in sorted order, so a wave's particles are neighbours)
// ORs its bits into that window in LDS and then adds each nonzero window word
// to the mask with one atomic; other waves OR straight into the mask.
constexpr int kWinWords = 3, kWinRows = 16;
constexpr int kWin = kWinWords * kWinRows * kWinRows;

__global__ __launch_bounds__(256) void k_band_exceptions(BandArgs a) {
    __shared__ uint32_t win[4][kWin];
    const bool by_flags = a.all_exceptions != 0 && a.flags != nullptr;
    const uint32_t count = by_flags ? a.num_particles : a.counters->num_exceptions;
    const uint32_t wx = row_words(a.sc.rx);
    // The reference compares the FP32 distance with 1.5 h in double.  For a
    // float d, (double)d < L holds exactly when d < the smallest float >= L.
    const double L = (double)a.h * 1.5;
    float limit = (float)L;
    if ((double)limit < L) limit = __int_as_float(__float_as_int(limit) + 1);
    const uint32_t lane = __lane_id();
    uint32_t* w = win[threadIdx.x >> 6];
    const uint32_t stride = gridDim.x * blockDim.x;
    // wave-uniform loop: every lane of the wave runs the same iterations
    for (uint32_t e0 = blockIdx.x * blockDim.x + (threadIdx.x & ~63u); e0 < count; e0 += stride) {
        const uint32_t e = e0 + lane;
        bool live = e < count;
        uint32_t pi = 0;
        if (live) {
            if (by_flags) {
                live = a.flags[e] != 0u;
                pi = e;
            } else {
                pi = a.exceptions[e];
            }
        }
        float4 p = make_float4(0.f, 0.f, 0.f, 0.f);
        int x0 = 0x7fffffff, x1 = -1, y0 = 0x7fffffff, y1 = -1, z0 = 0x7fffffff, z1 = -1;
        if (live) {
            p = a.sorted[pi];
            const int i = trunc_clamped(cell_coord(p.x, a.sc.minx, a.sc.cell), a.sc.rx);
            const int j = trunc_clamped(cell_coord(p.y, a.sc.miny, a.sc.cell), a.sc.ry);
            const int k = trunc_clamped(cell_coord(p.z, a.sc.minz, a.sc.cell), a.sc.rz);
            x0 = clampi(i - 3, 0, a.sc.rx - 1); x1 = clampi(i + 3, 0, a.sc.rx - 1);
            y0 = clampi(j - 3, 0, a.sc.ry - 1); y1 = clampi(j + 3, 0, a.sc.ry - 1);
            z0 = clampi(k - 3, 0, a.sc.rz - 1); z1 = clampi(k + 3, 0, a.sc.rz - 1);
        }
        // a wave without a surface particle has nothing to do (wave-uniform)
        if (__ballot(live) == 0) continue;
        // the wave's window: words [W0, W1] of rows [Y0, Y1] of layers [Z0, Z1]
        // (over the live lanes; at least one is live)
        const int W0 = wave_min_i(x0) >> 5, W1 = wave_max_i(x1) >> 5;
        const int Y0 = wave_min_i(y0), Y1 = wave_max_i(y1);
        const int Z0 = wave_min_i(z0), Z1 = wave_max_i(z1);
        const int nw = W1 - W0 + 1, ny = Y1 - Y0 + 1, nz = Z1 - Z0 + 1;
        if (nw <= kWinWords && ny <= kWinRows && nz <= kWinRows) {
            const int n = nw * ny * nz;
            for (int t = (int)lane; t < n; t += 64) w[t] = 0u;
            wave_sync_lds();
            // lanes without a particle test nothing (in-range dummy bounds)
            exception_rows_window(a, p, live, live ? x0 : 0, live ? x1 : 0, live ? y0 : 0,
                                  live ? y1 : 0, live ? z0 : 0, live ? z1 : 0, limit, w, W0, Y0,
                                  Z0, nw, ny);
            wave_sync_lds();
            for (int t = (int)lane; t < n; t += 64) {
                const uint32_t m = w[t];
                if (m) {
                    const int wi = t % nw, yi = (t / nw) % ny, zi = t / (nw * ny);
                    const uint64_t row = (uint64_t)(Z0 + zi) * a.sc.ry + (uint32_t)(Y0 + yi);
                    atomicOr(&a.bits[row * wx + (uint32_t)(W0 + wi)], m);
                }
            }
            wave_sync_lds();
        } else if (live) {
            exception_rows(a, p, x0, x1, y0, y1, z0, z1, limit,
                           [&](int z, int y, uint32_t word, uint32_t m) {
                               atomicOr(&a.bits[((uint64_t)z * a.sc.ry + y) * wx + word], m);
                           });
        }
    }
}

// Compaction of the band mask into the ascending band index (the reference's
// scan + compactSurfaceVertex), in two launches: k_band_partials counts the
// set bits of every 256-word block; k_band_emit gives each block its offset
// (the sum of the counts of the blocks before it, at most 8 words per thread
// from L2), scans its words' counts, writes each word's offset (word_off, the
// band rank of the word's first vertex) and the block's vertices.  A block's
// vertices own one contiguous output range: they are assembled in LDS and
// stored with consecutive lanes on consecutive entries (direct per-word stores
// when the range exceeds the staging buffer).  The last block writes the band
// size.
__global__ __launch_bounds__(256) void k_band_partials(const uint32_t* __restrict__ bits,
                                                       uint64_t nwords,
                                                       uint32_t* __restrict__ partials) {
    __shared__ uint32_t part[4];
    const uint64_t w = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t c = wave_sum(w < nwords ? (uint32_t)__popc(bits[w]) : 0u);
    if (__lane_id() == 0) part[threadIdx.x >> 6] = c;
    __syncthreads();
    if (threadIdx.x == 0) partials[blockIdx.x] = part[0] + part[1] + part[2] + part[3];
}

__global__ __launch_bounds__(256) void k_band_emit(const uint32_t* __restrict__ bits,
                                                   const uint32_t* __restrict__ partials,
                                                   uint64_t nwords, uint32_t wx, uint32_t rx,
                                                   uint32_t* __restrict__ word_off,
                                                   uint32_t* __restrict__ band_index,
                                                   uint64_t cap, uint32_t* __restrict__ total,
                                                   int scanned) {
    constexpr uint32_t kStage = 4096;
    __shared__ uint32_t stage[kStage];
    __shared__ uint32_t part[8];
    const uint64_t w = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const bool live = w < nwords;
    uint32_t b = live ? bits[w] : 0u;
    const uint32_t c = (uint32_t)__popc(b);
    // offset of this block: the counts of the blocks before it (or, when
    // `scanned`, partials already holds the exclusive offsets)
    uint32_t before = 0;
    if (scanned) {
        before = threadIdx.x == 0 ? partials[blockIdx.x] : 0u;
    } else {
        // at most 4096 blocks: up to 16 counts per thread, loaded together
        uint32_t pv[16];
#pragma unroll
        for (int k = 0; k < 16; ++k) {
            const uint32_t i = threadIdx.x + 256u * k;
            pv[k] = i < blockIdx.x ? partials[i] : 0u;
        }
#pragma unroll
        for (int k = 0; k < 16; ++k) before += pv[k];
    }
    before = wave_sum(before);
    // exclusive scan of the words' counts inside the block
    const uint32_t incl = wave_inclusive_sum(c);
    const uint32_t wave = threadIdx.x >> 6;
    if (__lane_id() == 63) {
        part[wave] = incl;
        part[4 + wave] = before;
    }
    __syncthreads();
    uint32_t lo = part[4] + part[5] + part[6] + part[7];  // the block's first output slot
    uint32_t wave_off = 0, block_total = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        if (k < (int)wave) wave_off += part[k];
        block_total += part[k];
    }
    const uint32_t o = lo + wave_off + incl - c;
    if (live) word_off[w] = o;
    if (blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) *total = lo + block_total;
    const uint64_t row = w / wx;
    const uint32_t vbase = (uint32_t)(row * rx) + (uint32_t)(w % wx) * 32u;
    if (block_total <= kStage) {
        uint32_t k = o - lo;
        while (b) {
            if (k < kStage) stage[k] = vbase + __builtin_ctz(b);
            ++k;
            b &= b - 1u;
        }
        __syncthreads();
        for (uint32_t t = threadIdx.x; t < block_total; t += blockDim.x)
            if ((uint64_t)lo + t < cap) band_index[lo + t] = stage[t];
    } else {
        uint32_t k = o;
        while (b) {
            if (k < cap) band_index[k] = vbase + __builtin_ctz(b);
            ++k;
            b &= b - 1u;
        }
    }
}

// Per axis: virtual cell of every scalar vertex, and each virtual cell's
// vertex interval [lo, hi).  Layout of `maps`:
//   vc_x[rx] vc_y[ry] vc_z[rz] lo_x[vcx] hi_x[vcx] lo_y[vcy] hi_y[vcy] lo_z[vcz] hi_z[vcz]
struct AxisMaps {
    uint32_t *vc[3], *lo[3], *hi[3];
    uint32_t* row_any;   // bit (z ry + y): scalar row (y, z) holds a band vertex
    uint64_t row_words;  // words of row_any
};

// blockIdx.y = 0, 1, 2: the maps of axis x, y, z; blockIdx.y = 3: the row
// occupancy bits, one wave per 64 rows (each wave writes its two words);
// blockIdx.y = 4: clears the per-cell band counts cnt[ncells].
__global__ __launch_bounds__(256) void k_axis_maps(GridDev sc, GridDev sp, int radius, AxisMaps m,
                                                   const uint32_t* __restrict__ bits, uint32_t wx,
                                                   uint32_t* __restrict__ cnt, uint64_t ncells) {
    const uint32_t axis = blockIdx.y;
    if (axis == 4) {
        for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < ncells;
             i += (uint64_t)gridDim.x * blockDim.x)
            cnt[i] = 0u;
        return;
    }
    if (axis == 3) {
        const uint64_t rows = (uint64_t)sc.ry * sc.rz;
        for (uint64_t r0 = (uint64_t)blockIdx.x * blockDim.x + (threadIdx.x & ~63u); r0 < rows;
             r0 += (uint64_t)gridDim.x * blockDim.x) {
            const uint64_t r = r0 + __lane_id();
            uint32_t any = 0;
            if (r < rows)
                for (uint32_t w = 0; w < wx; ++w) any |= bits[r * wx + w];
            const uint64_t b = __ballot(any != 0);
            if (__lane_id() == 0) {
                const uint64_t k = r0 >> 5;
                m.row_any[k] = (uint32_t)b;
                if (k + 1 < m.row_words) m.row_any[k + 1] = (uint32_t)(b >> 32);
            }
        }
        return;
    }
    const uint32_t res = axis == 0 ? sc.rx : (axis == 1 ? sc.ry : sc.rz);
    const float smin = axis == 0 ? sc.minx : (axis == 1 ? sc.miny : sc.minz);
    const float pmin = axis == 0 ? sp.minx : (axis == 1 ? sp.miny : sp.minz);
    const int pres = (int)(axis == 0 ? sp.rx : (axis == 1 ? sp.ry : sp.rz));
    const uint32_t nvc = (uint32_t)(pres + 2 * radius);
    auto vc_of = [&](uint32_t v) -> uint32_t {
        const float q = cell_coord(vertex_coord(v, smin, sc.cell), pmin, sp.cell);
        int c = trunc_clamped(q, pres + radius);
        c = clampi(c, -radius, pres - 1 + radius);
        return (uint32_t)(c + radius);
    };
    // first vertex whose virtual cell is >= c (the map is non-decreasing:
    // vertex positions, the cell coordinate, truncation and clamping all are)
    auto first_at_least = [&](uint32_t c) -> uint32_t {
        uint32_t lo = 0, hi = res;
        while (lo < hi) {
            const uint32_t mid = (lo + hi) >> 1;
            if (vc_of(mid) < c) lo = mid + 1;
            else hi = mid;
        }
        return lo;
    };
    // threads [0, res): the vertex -> cell map; [res, res + nvc): each virtual
    // cell's vertex interval [lo, hi), empty (lo = hi) when no vertex maps to it
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < res + nvc;
         i += gridDim.x * blockDim.x) {
        if (i < res) {
            m.vc[axis][i] = vc_of(i);
        } else {
            const uint32_t c = i - res;
            m.lo[axis][c] = first_at_least(c);
            m.hi[axis][c] = first_at_least(c + 1);
        }
    }
}

// 32 row occupancy bits starting at row `bit0` (bits past the end read 0).
__device__ __forceinline__ uint32_t row_bits32(const AxisMaps& m, uint64_t bit0) {
    const uint64_t w = bit0 >> 5;
    const uint32_t sh = (uint32_t)(bit0 & 31u);
    const uint32_t lo = w < m.row_words ? m.row_any[w] : 0u;
    if (!sh) return lo;
    const uint32_t hi = w + 1 < m.row_words ? m.row_any[w + 1] : 0u;
    return (lo >> sh) | (hi << (32u - sh));
}

// Occupied rows y in [ylo + yb, min(ylo + yb + 32, yhi)) of layer z, bit i = row ylo + yb + i.
__device__ __forceinline__ uint32_t layer_rows(const AxisMaps& m, const GridDev& sc, uint32_t z,
                                               uint32_t ylo, uint32_t yb, uint32_t ny) {
    uint32_t b = row_bits32(m, (uint64_t)z * sc.ry + ylo + yb);
    const uint32_t left = ny - yb;
    if (left < 32u) b &= (1u << left) - 1u;
    return b;
}

// Cell-major band list, built per slab: the slab (cx, cy, z) is the part of
// virtual cell v = (cz vcy + cy) vcx + cx (cz = vc_z[z]) on scalar layer z.
// Slabs are stored cell by cell and by layer inside a cell:
//   slab(cx, cy, z) = vcx vcy lo_z[cz] + (cy vcx + cx) nz(cz) + (z - lo_z[cz]),
// so the exclusive scan of the slab counts is each slab's offset in the cell
// list, and a cell's vertices come out in ascending id order (layers z, rows
// y, then x), i.e. a 64-vertex chunk is a compact slab of the cell, which
// keeps the candidate culling of the field kernels effective.  The count pass
// takes one thread per slab (consecutive threads = consecutive cx, which read
// the same row words); rows without band vertices are skipped with the row
// occupancy bits.  It writes the slab's band vertex count and adds it to its
// cell's count (cnt, cleared by k_axis_maps).  The fill pass (k_slab_fill)
// writes the slab's vertex ids at its scanned offset.  The list itself is
// deterministic.
__global__ __launch_bounds__(256) void k_slab_rows(const uint32_t* __restrict__ bits, GridDev sc,
                                                   uint32_t wx, AxisMaps m, uint32_t vcx,
                                                   uint32_t vcy, uint32_t* __restrict__ slab,
                                                   uint32_t* __restrict__ cnt) {
    const uint64_t total = (uint64_t)vcx * vcy * sc.rz;
    for (uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total;
         t += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t cx = (uint32_t)(t % vcx);
        const uint32_t cy = (uint32_t)((t / vcx) % vcy);
        const uint32_t z = (uint32_t)(t / ((uint64_t)vcx * vcy));
        const uint32_t cz = m.vc[2][z];
        const uint32_t zlo = m.lo[2][cz], nz = m.hi[2][cz] - zlo;
        const uint64_t first = (uint64_t)vcx * vcy * zlo + ((uint64_t)cy * vcx + cx) * nz;
        const uint64_t idx = first + (z - zlo);
        const uint32_t v = (cz * vcy + cy) * vcx + cx;
        uint32_t o = 0;
        const uint32_t ylo = m.lo[1][cy], yhi = m.hi[1][cy];
        const uint32_t xlo = m.lo[0][cx], xhi = m.hi[0][cx];
        const uint32_t ny = yhi > ylo ? yhi - ylo : 0u;
        const uint32_t w0 = xlo >> 5;
        const uint32_t nw = xhi > xlo ? ((xhi + 31u) >> 5) - w0 : 0u;
        for (uint32_t yb = 0; nw && yb < ny; yb += 32u) {
            uint32_t rows = layer_rows(m, sc, z, ylo, yb, ny);
            while (rows) {
                // up to 4 occupied rows per round; with at most 2 words per
                // row (the common case) their 8 words are loaded together
                uint64_t rw[4];
                bool ok[4];
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    ok[k] = rows != 0u;
                    const uint32_t dy = ok[k] ? (uint32_t)__builtin_ctz(rows) : 0u;
                    rw[k] = ((uint64_t)z * sc.ry + ylo + yb + dy) * wx;
                    rows &= rows - 1u;
                }
                if (nw <= 2u) {
                    uint32_t v0[4], v1[4];
#pragma unroll
                    for (int k = 0; k < 4; ++k) {
                        v0[k] = ok[k] ? bits[rw[k] + w0] : 0u;
                        v1[k] = ok[k] && nw == 2u ? bits[rw[k] + w0 + 1u] : 0u;
                    }
#pragma unroll
                    for (int k = 0; k < 4; ++k) {
                        o += __popc(clip_word(v0[k], w0, xlo, xhi));
                        if (nw == 2u) o += __popc(clip_word(v1[k], w0 + 1u, xlo, xhi));
                    }
                } else {
                    for (int k = 0; k < 4; ++k) {
                        if (!ok[k]) break;
                        for (uint32_t j = 0; j < nw; ++j) {
                            const uint32_t w = w0 + j;
                            o += __popc(clip_word(bits[rw[k] + w], w, xlo, xhi));
                        }
                    }
                }
            }
        }
        slab[idx] = o;
        if (o) atomicAdd(&cnt[v], o);
    }
}

// Fill pass in slab order: thread t of block b takes slab idx = 256 b + t, so
// the block's slabs own one contiguous range of the cell list.  The slab's ids
// are written to LDS (serial over its set bits) and the block then copies the
// range out with consecutive threads on consecutive entries.  A slab idx
// decodes as: layer q = idx / (vcx vcy) lies in the layer range of the slab's
// z-cell cz = vc_z[q]; r = idx - vcx vcy lo_z[cz] = (cy vcx + cx) nz + (z - lo_z).
__global__ __launch_bounds__(256) void k_slab_fill(const uint32_t* __restrict__ bits, GridDev sc,
                                                   uint32_t wx, AxisMaps m, uint32_t vcx,
                                                   uint32_t vcy, const uint32_t* __restrict__ slab,
                                                   const uint32_t* __restrict__ slab_off,
                                                   uint32_t* __restrict__ cell_band, uint64_t cap,
                                                   uint64_t nslabs) {
    constexpr uint32_t kStage = 8192;
    __shared__ uint32_t stage[kStage];
    __shared__ uint32_t range[2];
    const uint64_t first = (uint64_t)blockIdx.x * blockDim.x;
    const uint64_t idx = first + threadIdx.x;
    const uint64_t last = min(first + blockDim.x, nslabs) - 1;
    if (threadIdx.x == 0) {
        range[0] = slab_off[first];
        range[1] = slab_off[last] + slab[last];
    }
    __syncthreads();
    const uint32_t lo = range[0], total = range[1] - range[0];
    const bool staged = total <= kStage;
    uint32_t n = idx < nslabs ? slab[idx] : 0u;
    if (n) {
        const uint64_t layer = (uint64_t)vcx * vcy;
        const uint32_t q = (uint32_t)(idx / layer);
        const uint32_t cz = m.vc[2][q];
        const uint32_t zlo = m.lo[2][cz], nz = m.hi[2][cz] - zlo;
        const uint64_t r = idx - layer * zlo;
        const uint32_t c = (uint32_t)(r / nz);
        const uint32_t z = zlo + (uint32_t)(r - (uint64_t)c * nz);
        const uint32_t cx = c % vcx, cy = c / vcx;
        uint32_t o = slab_off[idx];
        const uint32_t ylo = m.lo[1][cy], yhi = m.hi[1][cy];
        const uint32_t xlo = m.lo[0][cx], xhi = m.hi[0][cx];
        const uint32_t ny = yhi > ylo ? yhi - ylo : 0u;
        const uint32_t w0 = xlo >> 5;
        const uint32_t nw = xhi > xlo ? ((xhi + 31u) >> 5) - w0 : 0u;
        // the slab's ids in row order, then word order
        auto put = [&](uint64_t row, uint32_t w, uint32_t word) {
            word = clip_word(word, w, xlo, xhi);
            const uint32_t base = (uint32_t)(row * sc.rx) + w * 32u;
            while (word) {
                const uint32_t id = base + __builtin_ctz(word);
                if (staged) stage[o - lo] = id;
                else if (o < cap) cell_band[o] = id;
                ++o;
                word &= word - 1u;
            }
        };
        for (uint32_t yb = 0; nw && yb < ny; yb += 32u) {
            uint32_t rows = layer_rows(m, sc, z, ylo, yb, ny);
            while (rows) {
                // up to 4 occupied rows per round; with at most 2 words per
                // row (the common case) their 8 words are loaded together
                uint64_t rr[4];
                bool ok[4];
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    ok[k] = rows != 0u;
                    const uint32_t dy = ok[k] ? (uint32_t)__builtin_ctz(rows) : 0u;
                    rr[k] = (uint64_t)z * sc.ry + ylo + yb + dy;
                    rows &= rows - 1u;
                }
                if (nw <= 2u) {
                    uint32_t v0[4], v1[4];
#pragma unroll
                    for (int k = 0; k < 4; ++k) {
                        v0[k] = ok[k] ? bits[rr[k] * wx + w0] : 0u;
                        v1[k] = ok[k] && nw == 2u ? bits[rr[k] * wx + w0 + 1u] : 0u;
                    }
                    for (int k = 0; k < 4; ++k) {
                        if (!ok[k]) break;
                        put(rr[k], w0, v0[k]);
                        if (nw == 2u) put(rr[k], w0 + 1u, v1[k]);
                    }
                } else {
                    for (int k = 0; k < 4; ++k) {
                        if (!ok[k]) break;
                        for (uint32_t j = 0; j < nw; ++j) put(rr[k], w0 + j, bits[rr[k] * wx + w0 + j]);
                    }
                }
            }
        }
    }
    if (staged) {
        __syncthreads();
        for (uint32_t k = threadIdx.x; k < total; k += blockDim.x)
            if ((uint64_t)lo + k < cap) cell_band[lo + k] = stage[k];
    }
}

}  // namespace

int band_dilation_exact(const fsr_grid_info& sc, float h) {
    // farthest box vertex from any point of the particle's scalar cell:
    // |dx| < 4, |dy| < 4, |dz| < 4 cells
    const double reach = 4.0 * 1.7320508075688772 * (double)sc.cell_size;
    double maxabs = 0.0;
    const double lo[3] = {sc.min_pos.x, sc.min_pos.y, sc.min_pos.z};
    const uint32_t r[3] = {sc.resolution.x, sc.resolution.y, sc.resolution.z};
    for (int a = 0; a < 3; ++a) {
        const double hi = lo[a] + (double)sc.cell_size * (double)r[a];
        maxabs = std::max(maxabs, std::max(std::fabs(lo[a]), std::fabs(hi)));
    }
    // FP32 rounding of the vertex position, the difference and the length
    const double slack = 64.0 * 1.1920928955078125e-7 * std::max(maxabs, (double)h);
    return reach + slack < 1.5 * (double)h ? 1 : 0;
}

fsr_status band_build(fsr_context* ctx, const BandArgs& a, hipStream_t s) {
    const uint32_t wx = row_words(a.sc.rx);
    const uint64_t nwords = (uint64_t)wx * a.sc.ry * a.sc.rz;
    const uint32_t blocks = (uint32_t)((nwords + 255) / 256);
    // S (bytes) -> tmp (x), tmp -> S (y, its first nwords words), S -> bits
    // (z); `sbits` is consumed
    uint32_t* sb = const_cast<uint32_t*>(a.sbits);
    const dim3 tiles((wx + kTW - 1) / kTW, (a.sc.ry + kTY - 1) / kTY, (a.sc.rz + kTZ - 1) / kTZ);
    if (tiles.y <= 65535u && tiles.z <= 65535u) {
        hipLaunchKernelGGL(k_dilate_box, tiles, dim3(kDilateThreads), 0, s, (const uint8_t*)sb, a.bits, wx,
                           a.sc);
        FSR_LAUNCH_CHECK("k_dilate_box");
    } else {
        hipLaunchKernelGGL(k_dilate_x, dim3(blocks), dim3(256), 0, s, (const uint8_t*)sb, a.tmp,
                           wx, nwords, a.sc.rx);
        FSR_LAUNCH_CHECK("k_dilate_x");
        hipLaunchKernelGGL(k_dilate_rows, dim3(blocks), dim3(256), 0, s, a.tmp, sb, nwords,
                           (uint64_t)wx, a.sc.ry, (uint64_t)a.sc.ry);
        FSR_LAUNCH_CHECK("k_dilate_rows(y)");
        hipLaunchKernelGGL(k_dilate_rows, dim3(blocks), dim3(256), 0, s, sb, a.bits, nwords,
                           (uint64_t)wx * a.sc.ry, a.sc.rz, (uint64_t)a.sc.rz);
        FSR_LAUNCH_CHECK("k_dilate_rows(z)");
    }
    const int eb = persistent_blocks(ctx, 256, 4, TUNE_BAND);
    hipLaunchKernelGGL(k_band_exceptions, dim3(eb), dim3(256), 0, s, a);
    FSR_LAUNCH_CHECK("k_band_exceptions");
    return FSR_OK;
}

fsr_status band_compact(fsr_context* ctx, const uint32_t* bits, const GridDev& sc,
                        const GridDev& sp, int radius, Counters* counters, fsr_frame_stats* st,
                        BandCompactOut* out, hipStream_t s, bool materialize) {
    const uint32_t wx = row_words(sc.rx);
    const uint64_t nwords = (uint64_t)wx * sc.ry * sc.rz;
    const uint64_t V = (uint64_t)sc.rx * sc.ry * sc.rz;
    uint32_t *word_off, *band_index;
    FSR_TRY(arena_typed(ctx, A_WORD_OFF, nwords, &word_off));
    FSR_TRY(arena_typed(ctx, A_BAND_INDEX, V, &band_index));
    const uint32_t emit_blocks = (uint32_t)((nwords + 255) / 256);
    uint32_t* partials;
    FSR_TRY(arena_typed(ctx, A_BAND_PARTIALS, emit_blocks, &partials));
    hipLaunchKernelGGL(k_band_partials, dim3(emit_blocks), dim3(256), 0, s, bits, nwords,
                       partials);
    FSR_LAUNCH_CHECK("k_band_partials");
    // up to 4096 blocks each block sums the counts before it (at most 16 words
    // per thread); beyond, the counts are scanned first
    const int scanned = emit_blocks > 4096u ? 1 : 0;
    if (scanned)
        FSR_TRY(scan_exclusive(ctx, SCAN_U32, partials, partials, emit_blocks, nullptr, nullptr,
                               s));
    hipLaunchKernelGGL(k_band_emit, dim3(emit_blocks), dim3(256), 0, s, bits, partials, nwords,
                       wx, sc.rx, word_off, band_index, V, &st->num_band_vertices, scanned);
    FSR_LAUNCH_CHECK("k_band_emit");
    out->band_index = band_index;

    // cell-major work list
    const uint32_t vcx = sp.rx + 2 * radius, vcy = sp.ry + 2 * radius, vcz = sp.rz + 2 * radius;
    const uint64_t nvc = (uint64_t)vcx * vcy * vcz;
    uint32_t *maps, *cnt, *cell_off, *chunk_off, *cell_band, *chunk_cell;
    const uint64_t rows = (uint64_t)sc.ry * sc.rz;
    const uint64_t row_words = 2ull * ((rows + 63) / 64);
    const uint64_t map_len =
        (uint64_t)sc.rx + sc.ry + sc.rz + 2ull * (vcx + vcy + vcz) + row_words;
    FSR_TRY(arena_typed(ctx, A_AXIS_MAPS, map_len, &maps));
    FSR_TRY(arena_typed(ctx, A_CELL_CNT, nvc, &cnt));
    FSR_TRY(arena_typed(ctx, A_CELL_OFF, nvc, &cell_off));
    FSR_TRY(arena_typed(ctx, A_CHUNK_OFF, nvc, &chunk_off));
    cell_band = nullptr;
    if (materialize) FSR_TRY(arena_typed(ctx, A_CELL_BAND, V, &cell_band));
    const uint64_t max_chunks = V / 64 + nvc + 1;
    FSR_TRY(arena_typed(ctx, A_CHUNK_CELL, max_chunks, &chunk_cell));
    AxisMaps m;
    m.vc[0] = maps;
    m.vc[1] = m.vc[0] + sc.rx;
    m.vc[2] = m.vc[1] + sc.ry;
    m.lo[0] = m.vc[2] + sc.rz;
    m.hi[0] = m.lo[0] + vcx;
    m.lo[1] = m.hi[0] + vcx;
    m.hi[1] = m.lo[1] + vcy;
    m.lo[2] = m.hi[1] + vcy;
    m.hi[2] = m.lo[2] + vcz;
    m.row_any = m.hi[2] + vcz;
    m.row_words = row_words;
    const uint32_t maxres = std::max(sc.rx + vcx, std::max(sc.ry + vcy, sc.rz + vcz));
    const uint32_t map_blocks = (uint32_t)std::max<uint64_t>(
        (maxres + 255) / 256, std::min<uint64_t>((rows + 255) / 256, 1024));
    hipLaunchKernelGGL(k_axis_maps, dim3(map_blocks, 5), dim3(256), 0, s, sc, sp, radius, m, bits,
                       wx, cnt, nvc);
    FSR_LAUNCH_CHECK("k_axis_maps");
    // slab counts (and cell counts) -> slab offsets; cell offsets, chunk
    // offsets and the chunk list in one scan -> slab fill
    const uint64_t nslabs = (uint64_t)vcx * vcy * sc.rz;
    uint32_t *slab, *slab_off;
    FSR_TRY(arena_typed(ctx, A_SLAB, 2 * nslabs, &slab));
    slab_off = slab + nslabs;
    const uint32_t slab_blocks = (uint32_t)std::min<uint64_t>((nslabs + 255) / 256, 16384);
    hipLaunchKernelGGL(k_slab_rows, dim3(slab_blocks), dim3(256), 0, s, bits, sc, wx, m, vcx, vcy,
                       slab, cnt);
    FSR_LAUNCH_CHECK("k_slab_rows");
    FSR_TRY(scan_cells(ctx, cnt, nvc, cell_off, chunk_off, chunk_cell, max_chunks,
                       &counters->num_chunks, s));
    if (materialize) {
        FSR_TRY(scan_exclusive(ctx, SCAN_U32, slab, slab_off, nslabs, nullptr, nullptr, s));
        hipLaunchKernelGGL(k_slab_fill, dim3((uint32_t)((nslabs + 255) / 256)), dim3(256), 0, s,
                           bits, sc, wx, m, vcx, vcy, (const uint32_t*)slab,
                           (const uint32_t*)slab_off, cell_band, V, nslabs);
        FSR_LAUNCH_CHECK("k_slab_fill");
    }
    out->cell_off = cell_off;
    out->cell_cnt = cnt;
    out->chunk_off = chunk_off;
    out->cell_band = materialize ? cell_band : nullptr;
    out->chunk_cell = chunk_cell;
    out->bits = bits;
    out->wx = wx;
    out->axis = m.lo[0];  // lo_x, hi_x, lo_y, hi_y, lo_z, hi_z follow one another
    out->slab = slab;
    out->vcx = vcx;
    out->vcy = vcy;
    out->vcz = vcz;
    out->radius = radius;
    out->max_chunks = max_chunks;
    return FSR_OK;
}

}  // namespace fsr
