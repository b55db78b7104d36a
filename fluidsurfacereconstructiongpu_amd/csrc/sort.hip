// Spatial hashing: cell id per particle, stable LSD radix sort of (cell id,
// particle id), gather of the particle data in cell order and cell ranges.
//
// Replaces launchSpatialGridBuilding (ReconstructionCUDA.cu:1437-1481):
//   calcParticlesHashKernel (ReconstructionCUDA.cu:1379-1396)  -> k_hash
//   thrust::sort_by_key     (ReconstructionCUDA.cu:1461-1466)  -> k_radix_hist / k_radix_scatter
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
// owns a 2048-key tile.  Pass = per-tile digit histogram in LDS -> one exclusive
// scan over the digit-major [256 x tiles] table -> stable scatter in which each
// 256-key round is ranked with eight wave ballots (lanes that share a digit),
// per-wave digit counts exchanged through LDS and running per-digit offsets.
#include "common.hpp"

namespace fsr {

namespace {

constexpr int kSortThreads = 256;
constexpr int kSortRounds = 8;
constexpr int kSortTile = kSortThreads * kSortRounds;

__global__ __launch_bounds__(256) void k_hash(const float* __restrict__ pos,
                                              const float* __restrict__ dens, uint32_t n,
                                              GridDev sp, float4* __restrict__ packed,
                                              uint32_t* __restrict__ keys) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float px = pos[3ull * i + 0];
    const float py = pos[3ull * i + 1];
    const float pz = pos[3ull * i + 2];
    const float d = dens[i];
    packed[i] = make_float4(px, py, pz, d);
    const float qx = cell_coord(px, sp.minx, sp.cell);
    const float qy = cell_coord(py, sp.miny, sp.cell);
    const float qz = cell_coord(pz, sp.minz, sp.cell);
    const uint32_t C = sp.rx * sp.ry * sp.rz;
    const bool ok = qx > -1.0f && qx < (float)sp.rx && qy > -1.0f && qy < (float)sp.ry &&
                    qz > -1.0f && qz < (float)sp.rz;
    uint32_t key = C;
    if (ok) {
        // truncation toward zero, as int3(float3) does
        const uint32_t ix = (uint32_t)(int)qx, iy = (uint32_t)(int)qy, iz = (uint32_t)(int)qz;
        key = iz * sp.rx * sp.ry + iy * sp.rx + ix;
    }
    keys[i] = key;
}

template <bool kFirst>
__global__ __launch_bounds__(kSortThreads) void k_radix_hist(const uint32_t* __restrict__ keys,
                                                             uint32_t n, uint32_t shift,
                                                             uint32_t* __restrict__ hist,
                                                             uint32_t tiles) {
    __shared__ uint32_t cnt[256];
    cnt[threadIdx.x] = 0;
    __syncthreads();
    const uint32_t base = blockIdx.x * kSortTile;
#pragma unroll
    for (int r = 0; r < kSortRounds; ++r) {
        const uint32_t i = base + r * kSortThreads + threadIdx.x;
        if (i < n) atomicAdd(&cnt[(keys[i] >> shift) & 255u], 1u);
    }
    __syncthreads();
    hist[threadIdx.x * tiles + blockIdx.x] = cnt[threadIdx.x];
}

template <bool kFirst>
__global__ __launch_bounds__(kSortThreads) void k_radix_scatter(
    const uint32_t* __restrict__ keys_in, const uint32_t* __restrict__ vals_in,
    uint32_t* __restrict__ keys_out, uint32_t* __restrict__ vals_out, uint32_t n,
    uint32_t shift, const uint32_t* __restrict__ offsets, uint32_t tiles) {
    __shared__ uint32_t digit_base[256];
    __shared__ uint32_t run[256];
    __shared__ uint32_t wave_cnt[4][256];
    __shared__ uint32_t wave_off[4][256];
    const uint32_t t = threadIdx.x;
    const uint32_t wave = t >> 6;
    digit_base[t] = offsets[t * tiles + blockIdx.x];
    run[t] = 0;
#pragma unroll
    for (int w = 0; w < 4; ++w) wave_cnt[w][t] = 0;
    __syncthreads();
    const uint32_t base = blockIdx.x * kSortTile;
    for (int r = 0; r < kSortRounds; ++r) {
        const uint32_t i = base + r * kSortThreads + t;
        const bool valid = i < n;
        const uint32_t key = valid ? keys_in[i] : 0u;
        const uint32_t val = kFirst ? i : (valid ? vals_in[i] : 0u);
        const uint32_t d = (key >> shift) & 255u;
        // lanes of this wave holding the same digit
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            const bool bit = (d >> b) & 1u;
            const uint64_t m = __ballot(bit);
            peers &= bit ? m : ~m;
        }
        const uint32_t rank = mask_rank(peers);
        if (valid && rank == 0) wave_cnt[wave][d] = (uint32_t)__popcll(peers);
        __syncthreads();
        {
            const uint32_t c0 = wave_cnt[0][t], c1 = wave_cnt[1][t], c2 = wave_cnt[2][t],
                           c3 = wave_cnt[3][t];
            const uint32_t r0 = run[t];
            wave_off[0][t] = r0;
            wave_off[1][t] = r0 + c0;
            wave_off[2][t] = r0 + c0 + c1;
            wave_off[3][t] = r0 + c0 + c1 + c2;
            run[t] = r0 + c0 + c1 + c2 + c3;
            wave_cnt[0][t] = 0; wave_cnt[1][t] = 0; wave_cnt[2][t] = 0; wave_cnt[3][t] = 0;
        }
        __syncthreads();
        if (valid) {
            const uint32_t p = digit_base[d] + wave_off[wave][d] + rank;
            keys_out[p] = key;
            vals_out[p] = val;
        }
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
    FSR_TRY(arena_typed(ctx, A_RADIX_HIST, 256ull * (tiles ? tiles : 1), &hist));

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
    hipLaunchKernelGGL(k_hash, dim3(blocks), dim3(256), 0, s, pos, dens, n, sp, packed, k0);
    FSR_LAUNCH_CHECK("k_hash");

    // key C (out of grid) must be representable
    int bits = 0;
    while (bits < 32 && (C >> bits) != 0) ++bits;
    const int passes = (bits + 7) / 8;
    ctx->last_sort_passes = passes;
    uint32_t *kin = k0, *kout = k1, *vin = v0, *vout = v1;
    for (int p = 0; p < passes; ++p) {
        const uint32_t shift = 8u * p;
        if (p == 0)
            hipLaunchKernelGGL(k_radix_hist<true>, dim3(tiles), dim3(kSortThreads), 0, s, kin, n,
                               shift, hist, tiles);
        else
            hipLaunchKernelGGL(k_radix_hist<false>, dim3(tiles), dim3(kSortThreads), 0, s, kin, n,
                               shift, hist, tiles);
        FSR_LAUNCH_CHECK("k_radix_hist");
        FSR_TRY(scan_exclusive(ctx, SCAN_U32, hist, hist, 256ull * tiles, nullptr, nullptr, s));
        if (p == 0)
            hipLaunchKernelGGL(k_radix_scatter<true>, dim3(tiles), dim3(kSortThreads), 0, s, kin,
                               vin, kout, vout, n, shift, hist, tiles);
        else
            hipLaunchKernelGGL(k_radix_scatter<false>, dim3(tiles), dim3(kSortThreads), 0, s, kin,
                               vin, kout, vout, n, shift, hist, tiles);
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
