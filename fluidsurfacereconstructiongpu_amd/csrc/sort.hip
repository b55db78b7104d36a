// Spatial hashing: cell id per particle, stable LSD radix sort of (cell id,
// particle id), gather of the particle data in cell order and cell ranges.
//
// Replaces launchSpatialGridBuilding (ReconstructionCUDA.cu:1437-1481):
//   calcParticlesHashKernel (ReconstructionCUDA.cu:1379-1396)  -> k_hash
//   thrust::sort_by_key     (ReconstructionCUDA.cu:1461-1466)  -> k_hash (first histogram),
//                                                                 k_radix_hist, k_radix_scatter
//   findCellRangeKernel     (ReconstructionCUDA.cu:1399-1435)  -> k_gather_ranges
// Semantics kept: cell id = z*rx*ry + y*rx + x with x = int((p.x-min.x)/cell)
// in IEEE FP32 (ReconstructionCUDA.cuh:338-358); the sort is stable, so
// particles of one cell keep their input order; positions and densities are
// permuted in place; empty cells hold start = end = 0xFFFFFFFF; the flag grid
// holds 1.0 for occupied cells.
// Difference: a particle whose cell lies outside the grid (or is NaN) gets the
// key C (one past the last cell), sorts to the end and belongs to no cell; the
// reference indexes out of bounds for it.
//
// The sort: 8-bit digits, ceil(bits(C)/8) passes (2 for C < 65536).  A block
// owns a 2048-key tile.  Pass = per-tile digit histogram in LDS (for the first
// pass, made by k_hash) -> per digit, an exclusive scan over the tiles and the
// digit total (one block per digit) -> stable scatter (which adds the totals of
// the smaller digits): each of 8 waves ranks its 256 contiguous keys with
// eight ballots per round and wave-private running digit counts, the block
// combines the waves' counts once, reorders the tile by digit in LDS and
// writes each digit's run with consecutive threads.
#include "common.hpp"

namespace fsr {

namespace {

constexpr int kSortThreads = 256;
constexpr int kSortRounds = 8;
constexpr int kSortTile = kSortThreads * kSortRounds;
// The histogram kernels take a tile with 1024 threads in 2 rounds, so that
// four times as many waves have their loads in flight.
constexpr int kHistThreads = 1024;
constexpr int kHistRounds = kSortTile / kHistThreads;

// Cell id per particle, the packed (x, y, z, density) copy, and the first
// pass's digit histogram of the block's 2048-key tile.
__global__ __launch_bounds__(kHistThreads) void k_hash(const float* __restrict__ pos,
                                                       const float* __restrict__ dens, uint32_t n,
                                                       GridDev sp, float4* __restrict__ packed,
                                                       uint32_t* __restrict__ keys,
                                                       uint32_t* __restrict__ hist, uint32_t tiles) {
    __shared__ uint32_t cnt[256];
    if (threadIdx.x < 256) cnt[threadIdx.x] = 0;
    __syncthreads();
    const uint32_t C = sp.rx * sp.ry * sp.rz;
    const uint32_t base = blockIdx.x * kSortTile;
#pragma unroll
    for (int r = 0; r < kHistRounds; ++r) {
        const uint32_t i = base + r * kHistThreads + threadIdx.x;
        if (i >= n) continue;
        const float px = pos[3ull * i + 0];
        const float py = pos[3ull * i + 1];
        const float pz = pos[3ull * i + 2];
        const float d = dens[i];
        packed[i] = make_float4(px, py, pz, d);
        const float qx = cell_coord(px, sp.minx, sp.cell);
        const float qy = cell_coord(py, sp.miny, sp.cell);
        const float qz = cell_coord(pz, sp.minz, sp.cell);
        const bool ok = qx > -1.0f && qx < (float)sp.rx && qy > -1.0f && qy < (float)sp.ry &&
                        qz > -1.0f && qz < (float)sp.rz;
        uint32_t key = C;
        if (ok) {
            // truncation toward zero, as int3(float3) does
            const uint32_t ix = (uint32_t)(int)qx, iy = (uint32_t)(int)qy, iz = (uint32_t)(int)qz;
            key = iz * sp.rx * sp.ry + iy * sp.rx + ix;
        }
        keys[i] = key;
        atomicAdd(&cnt[key & 255u], 1u);
    }
    __syncthreads();
    if (threadIdx.x < 256) hist[threadIdx.x * tiles + blockIdx.x] = cnt[threadIdx.x];
}

__global__ __launch_bounds__(kHistThreads) void k_radix_hist(const uint32_t* __restrict__ keys,
                                                             uint32_t n, uint32_t shift,
                                                             uint32_t* __restrict__ hist,
                                                             uint32_t tiles) {
    __shared__ uint32_t cnt[256];
    if (threadIdx.x < 256) cnt[threadIdx.x] = 0;
    __syncthreads();
    const uint32_t base = blockIdx.x * kSortTile;
#pragma unroll
    for (int r = 0; r < kHistRounds; ++r) {
        const uint32_t i = base + r * kHistThreads + threadIdx.x;
        if (i < n) atomicAdd(&cnt[(keys[i] >> shift) & 255u], 1u);
    }
    __syncthreads();
    if (threadIdx.x < 256) hist[threadIdx.x * tiles + blockIdx.x] = cnt[threadIdx.x];
}

// Per digit (one block each): exclusive prefix of the digit's per-tile counts
// over the tiles, in place, and the digit's total in totals[digit].
__global__ __launch_bounds__(kSortThreads) void k_digit_scan(uint32_t* __restrict__ hist,
                                                             uint32_t tiles,
                                                             uint32_t* __restrict__ totals) {
    __shared__ uint32_t wsum[4];
    __shared__ uint32_t carry;
    uint32_t* row = hist + (uint64_t)blockIdx.x * tiles;
    const uint32_t t = threadIdx.x, wave = t >> 6;
    if (t == 0) carry = 0;
    for (uint32_t b0 = 0; b0 < tiles; b0 += kSortThreads) {
        __syncthreads();
        const uint32_t b = b0 + t;
        const uint32_t c = b < tiles ? row[b] : 0u;
        const uint32_t incl = wave_inclusive_sum(c);
        if ((t & 63u) == 63u) wsum[wave] = incl;
        __syncthreads();
        uint32_t before = carry;
#pragma unroll
        for (int w = 0; w < 4; ++w) before += (w < (int)wave) ? wsum[w] : 0u;
        if (b < tiles) row[b] = before + incl - c;
        __syncthreads();
        if (t == kSortThreads - 1) carry = before + incl;
    }
    __syncthreads();
    if (t == 0) totals[blockIdx.x] = carry;
}

__device__ __forceinline__ void wave_sync_lds() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Stable scatter of one 2048-key tile by kScatterWaves waves.  Wave w owns the
// contiguous keys [256 w, 256 w + 256) of the tile, in 4 rounds of 64:
//   1. per round, the lanes holding the same digit are found with 8 ballots;
//      a key's rank in the wave's sub-tile is the wave's running count of its
//      digit plus its rank among those lanes (wave-private LDS counters);
//   2. one block step (threads 0..255, one digit each) turns the waves' digit
//      counts into the start of every (wave, digit) run in the tile's digit
//      order;
//   3. keys and values are placed in LDS in that order and written out with
//      consecutive threads on consecutive positions of each digit's run.
constexpr int kScatterWaves = 8;
constexpr int kScatterThreads = 64 * kScatterWaves;
constexpr int kScatterRounds = kSortTile / kScatterThreads;
template <bool kFirst>
__global__ __launch_bounds__(kScatterThreads) void k_radix_scatter(
    const uint32_t* __restrict__ keys_in, const uint32_t* __restrict__ vals_in,
    uint32_t* __restrict__ keys_out, uint32_t* __restrict__ vals_out, uint32_t n,
    uint32_t shift, const uint32_t* __restrict__ offsets, uint32_t tiles,
    const uint32_t* __restrict__ totals) {
    __shared__ uint32_t wcnt[kScatterWaves][256];  // counts, then run starts, of (wave, digit)
    __shared__ uint32_t dstart[256];               // first tile slot of each digit
    __shared__ uint32_t gbase[256];  // output position of the digit's first key of the tile
    __shared__ uint32_t wsum[4];
    __shared__ uint32_t skey[kSortTile];
    __shared__ uint32_t sval[kSortTile];
    const uint32_t t = threadIdx.x;
    const uint32_t wave = t >> 6, lane = t & 63u;
    const bool digit_thread = t < 256u;  // waves 0..3: thread t stands for digit t
    for (uint32_t k = t; k < kScatterWaves * 256u; k += kScatterThreads) (&wcnt[0][0])[k] = 0;
    {
        // output position of digit t's first key of this tile: the keys of
        // all smaller digits, then digit t's keys of the tiles before this one
        const uint32_t tot = digit_thread ? totals[t] : 0u;
        const uint32_t incl = wave_inclusive_sum(tot);
        if (digit_thread && lane == 63u) wsum[wave] = incl;
        __syncthreads();
        if (digit_thread) {
            uint32_t before = 0;
#pragma unroll
            for (int w = 0; w < 4; ++w) before += (w < (int)wave) ? wsum[w] : 0u;
            gbase[t] = before + incl - tot + offsets[t * tiles + blockIdx.x];
        }
    }
    __syncthreads();
    const uint32_t tile0 = blockIdx.x * kSortTile;
    const uint32_t base = tile0 + wave * (kSortTile / kScatterWaves);
    uint32_t key[kScatterRounds], val[kScatterRounds], loc[kScatterRounds];
#pragma unroll
    for (int j = 0; j < kScatterRounds; ++j) {
        const uint32_t i = base + j * 64u + lane;
        const bool valid = i < n;
        key[j] = valid ? keys_in[i] : 0u;
        val[j] = kFirst ? i : (valid ? vals_in[i] : 0u);
    }
#pragma unroll
    for (int j = 0; j < kScatterRounds; ++j) {
        const bool valid = base + j * 64u + lane < n;
        const uint32_t d = (key[j] >> shift) & 255u;
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            const bool bit = (d >> b) & 1u;
            const uint64_t m = __ballot(bit);
            peers &= bit ? m : ~m;
        }
        const uint32_t rank = mask_rank(peers);
        const uint32_t pre = wcnt[wave][d];
        wave_sync_lds();
        if (valid && rank == 0) wcnt[wave][d] = pre + (uint32_t)__popcll(peers);
        wave_sync_lds();
        loc[j] = pre + rank;
    }
    __syncthreads();
    {
        // thread t < 256: digit t.  Tile total of the digit, its start in the
        // tile (exclusive scan over digits), and the start of each wave's run.
        uint32_t c[kScatterWaves];
        uint32_t tot = 0;
#pragma unroll
        for (int w = 0; w < kScatterWaves; ++w) {
            c[w] = digit_thread ? wcnt[w][t] : 0u;
            tot += c[w];
        }
        const uint32_t incl = wave_inclusive_sum(tot);
        if (digit_thread && lane == 63u) wsum[wave] = incl;
        __syncthreads();
        if (digit_thread) {
            uint32_t before = 0;
#pragma unroll
            for (int w = 0; w < 4; ++w) before += (w < (int)wave) ? wsum[w] : 0u;
            uint32_t st = before + incl - tot;
            dstart[t] = st;
#pragma unroll
            for (int w = 0; w < kScatterWaves; ++w) {
                wcnt[w][t] = st;
                st += c[w];
            }
        }
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < kScatterRounds; ++j) {
        if (base + j * 64u + lane < n) {
            const uint32_t slot = wcnt[wave][(key[j] >> shift) & 255u] + loc[j];
            skey[slot] = key[j];
            sval[slot] = val[j];
        }
    }
    __syncthreads();
    const uint32_t count = min((uint32_t)kSortTile, n - tile0);
    for (uint32_t q = t; q < count; q += kScatterThreads) {
        const uint32_t k = skey[q];
        const uint32_t d = (k >> shift) & 255u;
        const uint32_t p = gbase[d] + (q - dstart[d]);
        keys_out[p] = k;
        vals_out[p] = sval[q];
    }
}

__global__ __launch_bounds__(256) void k_gather_ranges(
    const uint32_t* __restrict__ keys, const uint32_t* __restrict__ ids,
    const float4* __restrict__ packed, uint32_t n, uint32_t C, float* __restrict__ pos,
    float* __restrict__ dens, float4* __restrict__ sorted,
    fsr_index_range* __restrict__ ranges, float* __restrict__ occupancy,
    fsr_frame_stats* __restrict__ st) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t j = ids[i];
    const float4 p = packed[j];
    pos[3ull * i + 0] = p.x;
    pos[3ull * i + 1] = p.y;
    pos[3ull * i + 2] = p.z;
    dens[i] = p.w;
    sorted[i] = p;
    const uint32_t key = keys[i];
    const bool first = (i == 0) || keys[i - 1] != key;
    const bool last = (i == n - 1) || keys[i + 1] != key;
    if (key < C) {
        if (first) {
            ranges[key].start = i;
            occupancy[key] = 1.0f;
        }
        if (last) ranges[key].end = i + 1;
    } else if (first) {
        st->num_out_of_grid = n - i;
    }
    if (i == 0) st->num_particles = n;
}

}  // namespace

fsr_status spatial_hash(fsr_context* ctx, float* pos, float* dens, uint32_t n,
                        const fsr_grid_info& spg, fsr_frame_stats* st, HashOut* out,
                        hipStream_t s, const FillList* pre) {
    const GridDev sp = to_dev(spg);
    const uint64_t C64 = grid_count(spg);
    if (C64 == 0 || C64 >= 0xFFFFFFFFull) {
        set_error("spatial_hash: spatial grid resolution must be nonzero and hold < 2^32 cells");
        return FSR_ERR_INVALID_ARG;
    }
    const uint32_t C = (uint32_t)C64;
    float4 *packed, *sorted;
    uint32_t *k0, *k1, *v0, *v1, *hist;
    fsr_index_range* ranges;
    float* occ;
    const uint64_t nn = n ? n : 1;
    FSR_TRY(arena_typed(ctx, A_PACKED, nn, &packed));
    FSR_TRY(arena_typed(ctx, A_SORTED, nn, &sorted));
    FSR_TRY(arena_typed(ctx, A_KEYS0, nn, &k0));
    FSR_TRY(arena_typed(ctx, A_KEYS1, nn, &k1));
    FSR_TRY(arena_typed(ctx, A_VALS0, nn, &v0));
    FSR_TRY(arena_typed(ctx, A_VALS1, nn, &v1));
    FSR_TRY(arena_typed(ctx, A_RANGES, C, &ranges));
    FSR_TRY(arena_typed(ctx, A_OCCUPANCY, C, &occ));
    const uint32_t tiles = (n + kSortTile - 1) / kSortTile;
    // [256 x tiles] per-tile digit counts, then 256 digit totals
    FSR_TRY(arena_typed(ctx, A_RADIX_HIST, 256ull * (tiles ? tiles : 1) + 256ull, &hist));
    uint32_t* totals = hist + 256ull * (tiles ? tiles : 1);

    FillList fills;
    if (pre) fills = *pre;
    if (fills.count + 2 > kMaxFillJobs) {
        set_error("spatial_hash: too many fills");
        return FSR_ERR_INVALID_ARG;
    }
    fills.add(ranges, 0xFFFFFFFFu, 2ull * C);
    fills.add(occ, 0u, C);
    FSR_TRY(fill_multi(fills, s));
    out->sorted = sorted;
    out->ranges = ranges;
    out->occupancy = occ;
    if (n == 0) {
        ctx->last_sort_passes = 0;
        out->keys = k0;
        out->perm = v0;
        return FSR_OK;
    }
    const uint32_t blocks = (n + 255) / 256;
    hipLaunchKernelGGL(k_hash, dim3(tiles), dim3(kHistThreads), 0, s, pos, dens, n, sp, packed, k0,
                       hist, tiles);
    FSR_LAUNCH_CHECK("k_hash");

    // key C (out of grid) must be representable
    int bits = 0;
    while (bits < 32 && (C >> bits) != 0) ++bits;
    const int passes = (bits + 7) / 8;
    ctx->last_sort_passes = passes;
    uint32_t *kin = k0, *kout = k1, *vin = v0, *vout = v1;
    for (int p = 0; p < passes; ++p) {
        const uint32_t shift = 8u * p;
        // the first pass's histogram comes from k_hash
        if (p > 0) {
            hipLaunchKernelGGL(k_radix_hist, dim3(tiles), dim3(kHistThreads), 0, s, kin, n, shift,
                               hist, tiles);
            FSR_LAUNCH_CHECK("k_radix_hist");
        }
        hipLaunchKernelGGL(k_digit_scan, dim3(256), dim3(kSortThreads), 0, s, hist, tiles, totals);
        FSR_LAUNCH_CHECK("k_digit_scan");
        if (p == 0)
            hipLaunchKernelGGL(k_radix_scatter<true>, dim3(tiles), dim3(kScatterThreads), 0, s, kin,
                               vin, kout, vout, n, shift, hist, tiles, totals);
        else
            hipLaunchKernelGGL(k_radix_scatter<false>, dim3(tiles), dim3(kScatterThreads), 0, s, kin,
                               vin, kout, vout, n, shift, hist, tiles, totals);
        FSR_LAUNCH_CHECK("k_radix_scatter");
        uint32_t* t;
        t = kin; kin = kout; kout = t;
        t = vin; vin = vout; vout = t;
    }
    // kin/vin now hold the sorted keys and the sorted original ids (the permutation)
    hipLaunchKernelGGL(k_gather_ranges, dim3(blocks), dim3(256), 0, s, kin, vin, packed, n, C,
                       pos, dens, sorted, ranges, occ, st);
    FSR_LAUNCH_CHECK("k_gather_ranges");
    out->keys = kin;
    out->perm = vin;
    return FSR_OK;
}

}  // namespace fsr
