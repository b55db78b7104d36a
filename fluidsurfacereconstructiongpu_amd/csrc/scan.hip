// Exclusive prefix sums over u32 arrays whose live length may sit in device
// memory.  Replaces launchThrustExclusivePrefixSumScan
// (ReconstructionCUDA.cu:1154-1172), which ran thrust::exclusive_scan and then
// copied the total back to the host with two blocking copies; here the total
// stays on the device, so later stages size themselves from it without a host
// round trip.
//
// Two launches, no atomics and no waiting: each of G <= 1024 workgroups
// reduces a contiguous segment of whole 4096-element tiles; then each
// workgroup sums the partials of the segments before its own (at most 1023
// words, from L2) and rescans its segment from that offset in tiles (uint4
// loads, wave scans, one LDS exchange per tile); the last workgroup writes the
// total.  HBM traffic: read n twice, write n words.  (A single-pass decoupled
// look-back form measured no faster here: its per-tile ticket is a
// same-address atomic, and those serialise at roughly 30 ns each.)
#include <algorithm>
#include <mutex>

#include "common.hpp"

// Scratch of the allocation-free scan (fsr_exclusive_scan_noalloc): the
// partial sums of its at most 1024 workgroups (4 KiB).  One copy per device,
// held for the life of the process.  Host code never passes this name to a
// kernel (on the host it is only a stand-in): it uses the device address that
// hipGetSymbolAddress returns, kept per device in g_scratch_ptr.
constexpr uint32_t kNoAllocScratchWords = 1024;
static __device__ __align__(256) uint32_t scratch[kNoAllocScratchWords];

namespace fsr {

namespace {

constexpr int kScanThreads = 256;
constexpr int kScanItems = 16;
constexpr uint32_t kScanTile = kScanThreads * kScanItems;
constexpr uint32_t kMaxScanBlocks = 1024;

template <int F>
__device__ __forceinline__ uint32_t xform(uint32_t v) {
    if (F == SCAN_POPC) return (uint32_t)__popc(v);
    if (F == SCAN_CHUNKS64) return (v + 63u) >> 6;
    if (F == SCAN_NONZERO) return v ? 1u : 0u;
    return v;
}

__device__ __forceinline__ uint64_t live_length(uint64_t n_max, const uint32_t* d_count) {
    if (!d_count) return n_max;
    uint64_t c = *d_count;
    return c < n_max ? c : n_max;
}

// Block-wide exclusive scan of one value per thread; returns the block total.
__device__ __forceinline__ uint32_t block_exclusive(uint32_t v, uint32_t* lds_wave,
                                                    uint32_t* out_excl) {
    const uint32_t lane = __lane_id();
    const uint32_t wave = threadIdx.x >> 6;
    uint32_t inc = wave_inclusive_sum(v);
    if (lane == 63) lds_wave[wave] = inc;
    __syncthreads();
    uint32_t wave_off = 0, total = 0;
#pragma unroll
    for (int w = 0; w < kScanThreads / 64; ++w) {
        uint32_t t = lds_wave[w];
        if ((uint32_t)w < wave) wave_off += t;
        total += t;
    }
    __syncthreads();
    *out_excl = wave_off + inc - v;
    return total;
}

__device__ __forceinline__ uint64_t segment_len(uint64_t n, uint32_t g) {
    const uint64_t seg = (n + g - 1) / g;
    return (seg + kScanTile - 1) / kScanTile * kScanTile;
}

// Two-pass form, pass 1: partials[b] = sum of segment b.
template <int F>
__global__ __launch_bounds__(kScanThreads) void k_scan_reduce(
    const uint32_t* __restrict__ in, uint64_t n_max, const uint32_t* d_count,
    uint32_t* __restrict__ partials, int vec_ok) {
    __shared__ uint32_t lds[kScanThreads / 64];
    const uint64_t n = live_length(n_max, d_count);
    const uint64_t seg = segment_len(n, gridDim.x);
    const uint64_t lo = (uint64_t)blockIdx.x * seg;
    const uint64_t hi = lo + seg < n ? lo + seg : n;
    uint32_t sum = 0;
    for (uint64_t base = lo; base < hi; base += kScanTile) {
        const uint64_t i0 = base + (uint64_t)kScanItems * threadIdx.x;
        if (base + kScanTile <= hi && vec_ok) {
#pragma unroll
            for (int k = 0; k < kScanItems; k += 4) {
                const uint4 q = *reinterpret_cast<const uint4*>(in + i0 + k);
                sum += xform<F>(q.x) + xform<F>(q.y) + xform<F>(q.z) + xform<F>(q.w);
            }
        } else {
#pragma unroll
            for (int k = 0; k < kScanItems; ++k)
                if (i0 + k < hi) sum += xform<F>(in[i0 + k]);
        }
    }
    uint32_t excl;
    const uint32_t total = block_exclusive(sum, lds, &excl);
    if (threadIdx.x == 0) partials[blockIdx.x] = total;
}

// Two-pass form, pass 2: offset = sum of the partials before this segment, then
// the segment rescanned in tiles; the last workgroup writes the total.
template <int F>
__global__ __launch_bounds__(kScanThreads) void k_scan_down(
    const uint32_t* in, uint32_t* out, uint64_t n_max, const uint32_t* d_count,
    const uint32_t* __restrict__ partials, uint32_t* d_total, int vec_ok) {
    __shared__ uint32_t lds[kScanThreads / 64];
    const bool last = blockIdx.x == gridDim.x - 1;
    if (!out && !last) return;
    const uint64_t n = live_length(n_max, d_count);
    const uint64_t seg = segment_len(n, gridDim.x);
    const uint64_t lo = (uint64_t)blockIdx.x * seg;
    const uint64_t hi = lo + seg < n ? lo + seg : n;
    // the partials of the blocks before this one (at most kMaxScanBlocks),
    // loaded together
    uint32_t s = 0;
    {
        uint32_t pv[kMaxScanBlocks / kScanThreads];
#pragma unroll
        for (uint32_t k = 0; k < kMaxScanBlocks / kScanThreads; ++k) {
            const uint32_t i = threadIdx.x + kScanThreads * k;
            pv[k] = i < blockIdx.x ? partials[i] : 0u;
        }
#pragma unroll
        for (uint32_t k = 0; k < kMaxScanBlocks / kScanThreads; ++k) s += pv[k];
    }
    uint32_t excl;
    uint32_t running = block_exclusive(s, lds, &excl);
    if (!out) {
        if (threadIdx.x == 0 && d_total) *d_total = running + partials[blockIdx.x];
        return;
    }
    for (uint64_t base = lo; base < hi; base += kScanTile) {
        const uint64_t i0 = base + (uint64_t)kScanItems * threadIdx.x;
        const bool full = base + kScanTile <= hi;
        uint32_t v[kScanItems];
        if (full && vec_ok) {
#pragma unroll
            for (int k = 0; k < kScanItems; k += 4) {
                const uint4 q = *reinterpret_cast<const uint4*>(in + i0 + k);
                v[k] = xform<F>(q.x);
                v[k + 1] = xform<F>(q.y);
                v[k + 2] = xform<F>(q.z);
                v[k + 3] = xform<F>(q.w);
            }
        } else {
#pragma unroll
            for (int k = 0; k < kScanItems; ++k) v[k] = (i0 + k < hi) ? xform<F>(in[i0 + k]) : 0u;
        }
        uint32_t tsum = 0;
#pragma unroll
        for (int k = 0; k < kScanItems; ++k) tsum += v[k];
        const uint32_t total = block_exclusive(tsum, lds, &excl);
        uint32_t r = running + excl;
        uint32_t o[kScanItems];
#pragma unroll
        for (int k = 0; k < kScanItems; ++k) {
            o[k] = r;
            r += v[k];
        }
        if (full && vec_ok) {
#pragma unroll
            for (int k = 0; k < kScanItems; k += 4)
                *reinterpret_cast<uint4*>(out + i0 + k) =
                    make_uint4(o[k], o[k + 1], o[k + 2], o[k + 3]);
        } else {
#pragma unroll
            for (int k = 0; k < kScanItems; ++k)
                if (i0 + k < hi) out[i0 + k] = o[k];
        }
        running += total;
    }
    if (last && threadIdx.x == 0 && d_total) *d_total = running;
}

// Marching-cubes offsets in two passes over the per-cube vertex counts nv[b]:
// vert_off = exclusive scan of nv, and the ascending list of the cubes with
// nv > 0 (the valid cubes) compacted in the same pass.  Pass 1: partials[b] =
// sum of nv over segment b, partials[G + b] = its valid cubes.
__global__ __launch_bounds__(kScanThreads) void k_mc_scan_reduce(
    const uint32_t* __restrict__ nv, uint64_t n_max, const uint32_t* d_count,
    uint32_t* __restrict__ partials) {
    __shared__ uint32_t lds[kScanThreads / 64];
    const uint64_t n = live_length(n_max, d_count);
    const uint64_t seg = segment_len(n, gridDim.x);
    const uint64_t lo = (uint64_t)blockIdx.x * seg;
    const uint64_t hi = lo + seg < n ? lo + seg : n;
    uint32_t sum = 0, cnt = 0;
    for (uint64_t base = lo; base < hi; base += kScanTile) {
        const uint64_t i0 = base + (uint64_t)kScanItems * threadIdx.x;
        if (base + kScanTile <= hi) {
#pragma unroll
            for (int k = 0; k < kScanItems; k += 4) {
                const uint4 q = *reinterpret_cast<const uint4*>(nv + i0 + k);
                sum += q.x + q.y + q.z + q.w;
                cnt += (q.x ? 1u : 0u) + (q.y ? 1u : 0u) + (q.z ? 1u : 0u) + (q.w ? 1u : 0u);
            }
        } else {
#pragma unroll
            for (int k = 0; k < kScanItems; ++k) {
                const uint32_t v = (i0 + k < hi) ? nv[i0 + k] : 0u;
                sum += v;
                cnt += v ? 1u : 0u;
            }
        }
    }
    uint32_t excl;
    const uint32_t total = block_exclusive(sum, lds, &excl);
    const uint32_t total_cnt = block_exclusive(cnt, lds, &excl);
    if (threadIdx.x == 0) {
        partials[blockIdx.x] = total;
        partials[gridDim.x + blockIdx.x] = total_cnt;
    }
}

// Pass 2: each segment from the sums of the partials before it.
__global__ __launch_bounds__(kScanThreads) void k_mc_scan_down(
    const uint32_t* __restrict__ nv, uint64_t n_max, const uint32_t* d_count,
    const uint32_t* __restrict__ partials, uint32_t* __restrict__ vert_off,
    uint32_t* __restrict__ valid_idx, uint32_t* d_total_verts, uint32_t* d_total_valid) {
    __shared__ uint32_t lds[kScanThreads / 64];
    const uint64_t n = live_length(n_max, d_count);
    const uint64_t seg = segment_len(n, gridDim.x);
    const uint64_t lo = (uint64_t)blockIdx.x * seg;
    const uint64_t hi = lo + seg < n ? lo + seg : n;
    // the two partials of the blocks before this one (at most kMaxScanBlocks),
    // loaded together
    uint32_t s0 = 0, s1 = 0;
    {
        uint32_t p0[kMaxScanBlocks / kScanThreads], p1[kMaxScanBlocks / kScanThreads];
#pragma unroll
        for (uint32_t k = 0; k < kMaxScanBlocks / kScanThreads; ++k) {
            const uint32_t i = threadIdx.x + kScanThreads * k;
            p0[k] = i < blockIdx.x ? partials[i] : 0u;
            p1[k] = i < blockIdx.x ? partials[gridDim.x + i] : 0u;
        }
#pragma unroll
        for (uint32_t k = 0; k < kMaxScanBlocks / kScanThreads; ++k) {
            s0 += p0[k];
            s1 += p1[k];
        }
    }
    uint32_t excl;
    uint32_t run_v = block_exclusive(s0, lds, &excl);
    uint32_t run_c = block_exclusive(s1, lds, &excl);
    for (uint64_t base = lo; base < hi; base += kScanTile) {
        const uint64_t i0 = base + (uint64_t)kScanItems * threadIdx.x;
        const bool full = base + kScanTile <= hi;
        uint32_t v[kScanItems];
        if (full) {
#pragma unroll
            for (int k = 0; k < kScanItems; k += 4) {
                const uint4 q = *reinterpret_cast<const uint4*>(nv + i0 + k);
                v[k] = q.x;
                v[k + 1] = q.y;
                v[k + 2] = q.z;
                v[k + 3] = q.w;
            }
        } else {
#pragma unroll
            for (int k = 0; k < kScanItems; ++k) v[k] = (i0 + k < hi) ? nv[i0 + k] : 0u;
        }
        uint32_t tv = 0, tc = 0;
#pragma unroll
        for (int k = 0; k < kScanItems; ++k) {
            tv += v[k];
            tc += v[k] ? 1u : 0u;
        }
        uint32_t ev, ec;
        const uint32_t totv = block_exclusive(tv, lds, &ev);
        const uint32_t totc = block_exclusive(tc, lds, &ec);
        uint32_t rv = run_v + ev, rc = run_c + ec;
        uint32_t o[kScanItems];
#pragma unroll
        for (int k = 0; k < kScanItems; ++k) {
            o[k] = rv;
            if (v[k] && i0 + k < hi) valid_idx[rc] = (uint32_t)(i0 + k);
            rv += v[k];
            rc += v[k] ? 1u : 0u;
        }
        if (full) {
#pragma unroll
            for (int k = 0; k < kScanItems; k += 4)
                *reinterpret_cast<uint4*>(vert_off + i0 + k) =
                    make_uint4(o[k], o[k + 1], o[k + 2], o[k + 3]);
        } else {
#pragma unroll
            for (int k = 0; k < kScanItems; ++k)
                if (i0 + k < hi) vert_off[i0 + k] = o[k];
        }
        run_v += totv;
        run_c += totc;
    }
    if (blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) {
        *d_total_verts = run_v;
        *d_total_valid = run_c;
    }
}

// Cell-major work list offsets in two passes over the per-cell band counts:
// cell_off = exclusive scan of cnt, chunk_off = exclusive scan of the cells'
// 64-vertex chunk counts, and chunk_cell (chunk -> cell) written in the same
// pass.  One cell per thread in 256-cell tiles and up to 4096 workgroups, so
// that the cells with many chunks (the ring cells outside the hash grid) are
// spread over many workgroups.  Pass 1: partials[b] / partials[G + b] = the
// vertices / chunks of segment b.
constexpr uint32_t kCellTile = kScanThreads;
constexpr uint32_t kMaxCellBlocks = 4096;

__device__ __forceinline__ uint32_t chunks64(uint32_t c) { return (c + 63u) >> 6; }

__device__ __forceinline__ uint64_t cell_segment(uint64_t n, uint32_t g) {
    const uint64_t seg = (n + g - 1) / g;
    return (seg + kCellTile - 1) / kCellTile * kCellTile;
}

__global__ __launch_bounds__(kScanThreads) void k_cells_scan_reduce(
    const uint32_t* __restrict__ cnt, uint64_t n, uint32_t* __restrict__ partials) {
    __shared__ uint32_t lds[kScanThreads / 64];
    const uint64_t seg = cell_segment(n, gridDim.x);
    const uint64_t lo = (uint64_t)blockIdx.x * seg;
    const uint64_t hi = lo + seg < n ? lo + seg : n;
    uint32_t sum = 0, ch = 0;
    for (uint64_t i = lo + threadIdx.x; i < hi; i += kScanThreads) {
        const uint32_t c = cnt[i];
        sum += c;
        ch += chunks64(c);
    }
    uint32_t excl;
    const uint32_t tsum = block_exclusive(sum, lds, &excl);
    const uint32_t tch = block_exclusive(ch, lds, &excl);
    if (threadIdx.x == 0) {
        partials[blockIdx.x] = tsum;
        partials[gridDim.x + blockIdx.x] = tch;
    }
}

__global__ __launch_bounds__(kScanThreads) void k_cells_scan_down(
    const uint32_t* __restrict__ cnt, uint64_t n, const uint32_t* __restrict__ partials,
    uint32_t* __restrict__ cell_off, uint32_t* __restrict__ chunk_off,
    uint32_t* __restrict__ chunk_cell, uint64_t max_chunks, uint32_t* d_num_chunks) {
    __shared__ uint32_t lds[kScanThreads / 64];
    const uint64_t seg = cell_segment(n, gridDim.x);
    const uint64_t lo = (uint64_t)blockIdx.x * seg;
    const uint64_t hi = lo + seg < n ? lo + seg : n;
    // the two partials of the blocks before this one (at most kMaxCellBlocks),
    // loaded together
    uint32_t s0 = 0, s1 = 0;
    {
        uint32_t p0[kMaxCellBlocks / kScanThreads], p1[kMaxCellBlocks / kScanThreads];
#pragma unroll
        for (uint32_t k = 0; k < kMaxCellBlocks / kScanThreads; ++k) {
            const uint32_t i = threadIdx.x + kScanThreads * k;
            p0[k] = i < blockIdx.x ? partials[i] : 0u;
            p1[k] = i < blockIdx.x ? partials[gridDim.x + i] : 0u;
        }
#pragma unroll
        for (uint32_t k = 0; k < kMaxCellBlocks / kScanThreads; ++k) {
            s0 += p0[k];
            s1 += p1[k];
        }
    }
    uint32_t excl;
    uint32_t run_v = block_exclusive(s0, lds, &excl);
    uint32_t run_c = block_exclusive(s1, lds, &excl);
    for (uint64_t base = lo; base < hi; base += kCellTile) {
        const uint64_t i = base + threadIdx.x;
        const uint32_t c = i < hi ? cnt[i] : 0u;
        const uint32_t nch = chunks64(c);
        uint32_t ev, ec;
        const uint32_t totv = block_exclusive(c, lds, &ev);
        const uint32_t totc = block_exclusive(nch, lds, &ec);
        if (i < hi) {
            const uint32_t rc = run_c + ec;
            cell_off[i] = run_v + ev;
            chunk_off[i] = rc;
            for (uint32_t j = 0; j < nch; ++j)
                if ((uint64_t)rc + j < max_chunks) chunk_cell[rc + j] = (uint32_t)i;
        }
        run_v += totv;
        run_c += totc;
    }
    if (blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) *d_num_chunks = run_c;
}

__global__ void k_fill_u32(uint32_t* p, uint32_t v, uint64_t n) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

__global__ void k_fill_f32(float* p, float v, uint64_t n) {
    // float4 stores for the bulk; the tail element by element
    const uint64_t n4 = n / 4;
    float4* p4 = reinterpret_cast<float4*>(p);
    const float4 q = make_float4(v, v, v, v);
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n4; i += stride)
        p4[i] = q;
    for (uint64_t i = 4 * n4 + (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
        p[i] = v;
}

struct FillArgs {
    uint32_t* p[kMaxFillJobs];
    uint64_t n[kMaxFillJobs];
    uint32_t value[kMaxFillJobs];
    uint32_t blk_end[kMaxFillJobs];  // cumulative block counts
    int count;
};

// Blocks [blk_end[j-1], blk_end[j]) fill job j (block-uniform job lookup):
// 16-byte stores from the first 16-byte boundary, single words around them.
__global__ __launch_bounds__(256) void k_fill_multi(FillArgs a) {
    int j = 0;
    while (j + 1 < a.count && blockIdx.x >= a.blk_end[j]) ++j;
    const uint32_t b0 = j ? a.blk_end[j - 1] : 0u;
    const uint64_t nb = a.blk_end[j] - b0;
    uint32_t* p = a.p[j];
    const uint32_t v = a.value[j];
    const uint64_t n = a.n[j];
    const uint64_t t = (uint64_t)(blockIdx.x - b0) * blockDim.x + threadIdx.x;
    const uint64_t stride = nb * blockDim.x;
    const uint64_t head = min<uint64_t>(((16u - ((uintptr_t)p & 15u)) & 15u) >> 2, n);
    const uint64_t n4 = (n - head) >> 2;
    uint4* p4 = reinterpret_cast<uint4*>(p + head);
    const uint4 q = make_uint4(v, v, v, v);
    for (uint64_t i = t; i < n4; i += stride) p4[i] = q;
    if (t < head) p[t] = v;
    const uint64_t tail = head + 4 * n4;
    if (t < n - tail) p[tail + t] = v;
}

// Workgroups of a scan over n_max elements (each owns whole 4096-element tiles).
uint32_t scan_blocks(uint64_t n_max) {
    const uint64_t want = (n_max + kScanTile - 1) / kScanTile;
    return (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(want, kMaxScanBlocks));
}

// The two launches of a scan whose workgroup partials go to `partials`
// (scan_blocks(n_max) words).
template <int F>
fsr_status scan_launch(const uint32_t* in, uint32_t* out, uint64_t n_max, const uint32_t* d_count,
                       uint32_t* d_total, uint32_t* partials, hipStream_t s) {
    const uint32_t g = scan_blocks(n_max);
    const int vec_ok = ((((uintptr_t)in) | ((uintptr_t)out)) & 15u) == 0 ? 1 : 0;
    hipLaunchKernelGGL(k_scan_reduce<F>, dim3(g), dim3(kScanThreads), 0, s, in, n_max, d_count,
                       partials, vec_ok);
    FSR_LAUNCH_CHECK("k_scan_reduce");
    hipLaunchKernelGGL(k_scan_down<F>, dim3(g), dim3(kScanThreads), 0, s, in, out, n_max, d_count,
                       partials, d_total, vec_ok);
    FSR_LAUNCH_CHECK("k_scan_down");
    return FSR_OK;
}

template <int F>
fsr_status scan_impl(fsr_context* ctx, const uint32_t* in, uint32_t* out, uint64_t n_max,
                     const uint32_t* d_count, uint32_t* d_total, hipStream_t s) {
    uint32_t* partials = nullptr;
    FSR_TRY(arena_typed(ctx, A_SCAN_PARTIALS, 2ull * kMaxCellBlocks, &partials));
    return scan_launch<F>(in, out, n_max, d_count, d_total, partials, s);
}

static_assert(kNoAllocScratchWords >= kMaxScanBlocks, "scratch holds every workgroup partial");

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
uint32_t* g_scratch_ptr[kMaxDevices] = {};

}  // namespace

fsr_status noalloc_scratch_init(int device) {
    if (device < 0 || device >= kMaxDevices) {
        set_error("noalloc_scratch_init: device index beyond the scratch table");
        return FSR_ERR_UNSUPPORTED;
    }
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_ptr[device]) return FSR_OK;
    int cur = -1;
    FSR_HIP(hipGetDevice(&cur));
    if (cur != device) FSR_HIP(hipSetDevice(device));
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (cur != device) (void)hipSetDevice(cur);
    if (e != hipSuccess) return hip_fail(e, "hipGetSymbolAddress(scratch)", __FILE__, __LINE__);
    if (!p) {
        set_error("hipGetSymbolAddress(scratch) returned a null address");
        return FSR_ERR_HIP;
    }
    g_scratch_ptr[device] = static_cast<uint32_t*>(p);
    return FSR_OK;
}

fsr_status scan_exclusive_noalloc(int device, const uint32_t* in, uint32_t* out, uint64_t n,
                                  uint32_t* d_total, hipStream_t s) {
    if (n > FSR_NOALLOC_SCAN_MAX_ELEMENTS) {
        set_error("fsr_exclusive_scan_noalloc: more than FSR_NOALLOC_SCAN_MAX_ELEMENTS (2^30) "
                  "elements");
        return FSR_ERR_CAPACITY;
    }
    uint32_t* partials = nullptr;
    {
        std::lock_guard<std::mutex> lock(g_scratch_mu);
        if (device >= 0 && device < kMaxDevices) partials = g_scratch_ptr[device];
    }
    if (!partials) {
        set_error("fsr_exclusive_scan_noalloc: the scratch of this device was not set up "
                  "(fsr_create on the device does it)");
        return FSR_ERR_INVALID_ARG;
    }
    if (scan_blocks(n) > kNoAllocScratchWords) {
        set_error("fsr_exclusive_scan_noalloc: scan needs more partials than the scratch holds");
        return FSR_ERR_CAPACITY;
    }
    return scan_launch<SCAN_U32>(in, out, n, nullptr, d_total, partials, s);
}

fsr_status scan_exclusive(fsr_context* ctx, ScanInput f, const uint32_t* in, uint32_t* out,
                          uint64_t n_max, const uint32_t* d_count, uint32_t* d_total,
                          hipStream_t s) {
    switch (f) {
        case SCAN_U32: return scan_impl<SCAN_U32>(ctx, in, out, n_max, d_count, d_total, s);
        case SCAN_POPC: return scan_impl<SCAN_POPC>(ctx, in, out, n_max, d_count, d_total, s);
        case SCAN_CHUNKS64:
            return scan_impl<SCAN_CHUNKS64>(ctx, in, out, n_max, d_count, d_total, s);
        case SCAN_NONZERO:
            return scan_impl<SCAN_NONZERO>(ctx, in, out, n_max, d_count, d_total, s);
    }
    set_error("scan_exclusive: unknown input transform");
    return FSR_ERR_INVALID_ARG;
}

fsr_status scan_mc(fsr_context* ctx, const uint32_t* nv, uint64_t n_max, const uint32_t* d_count,
                   uint32_t* vert_off, uint32_t* valid_idx, uint32_t* d_total_verts,
                   uint32_t* d_total_valid, hipStream_t s) {
    if (((uintptr_t)nv | (uintptr_t)vert_off) & 15u) {
        set_error("scan_mc: the vertex counts and offsets must be 16-byte aligned");
        return FSR_ERR_INVALID_ARG;
    }
    const uint64_t want = (n_max + kScanTile - 1) / kScanTile;
    const uint32_t g = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(want, kMaxScanBlocks));
    uint32_t* partials = nullptr;
    FSR_TRY(arena_typed(ctx, A_SCAN_PARTIALS, 2ull * kMaxCellBlocks, &partials));
    hipLaunchKernelGGL(k_mc_scan_reduce, dim3(g), dim3(kScanThreads), 0, s, nv, n_max, d_count,
                       partials);
    FSR_LAUNCH_CHECK("k_mc_scan_reduce");
    hipLaunchKernelGGL(k_mc_scan_down, dim3(g), dim3(kScanThreads), 0, s, nv, n_max, d_count,
                       partials, vert_off, valid_idx, d_total_verts, d_total_valid);
    FSR_LAUNCH_CHECK("k_mc_scan_down");
    return FSR_OK;
}

fsr_status scan_cells(fsr_context* ctx, const uint32_t* cnt, uint64_t n, uint32_t* cell_off,
                      uint32_t* chunk_off, uint32_t* chunk_cell, uint64_t max_chunks,
                      uint32_t* d_num_chunks, hipStream_t s) {
    const uint64_t want = (n + kCellTile - 1) / kCellTile;
    const uint32_t g = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(want, kMaxCellBlocks));
    uint32_t* partials = nullptr;
    FSR_TRY(arena_typed(ctx, A_SCAN_PARTIALS, 2ull * kMaxCellBlocks, &partials));
    hipLaunchKernelGGL(k_cells_scan_reduce, dim3(g), dim3(kScanThreads), 0, s, cnt, n, partials);
    FSR_LAUNCH_CHECK("k_cells_scan_reduce");
    hipLaunchKernelGGL(k_cells_scan_down, dim3(g), dim3(kScanThreads), 0, s, cnt, n, partials,
                       cell_off, chunk_off, chunk_cell, max_chunks, d_num_chunks);
    FSR_LAUNCH_CHECK("k_cells_scan_down");
    return FSR_OK;
}

fsr_status fill_u32(uint32_t* p, uint32_t value, uint64_t n, hipStream_t s) {
    if (n == 0) return FSR_OK;
    uint64_t blocks = (n + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(k_fill_u32, dim3((uint32_t)blocks), dim3(256), 0, s, p, value, n);
    FSR_LAUNCH_CHECK("k_fill_u32");
    return FSR_OK;
}

fsr_status fill_multi(const FillList& l, hipStream_t s) {
    FillArgs a;
    a.count = 0;
    uint32_t blocks = 0;
    for (int j = 0; j < l.count; ++j) {
        if (l.n[j] == 0) continue;
        const int k = a.count++;
        a.p[k] = l.p[j];
        a.n[k] = l.n[j];
        a.value[k] = l.value[j];
        blocks += (uint32_t)std::min<uint64_t>((l.n[j] + 1023) / 1024, 4096);
        a.blk_end[k] = blocks;
    }
    if (!a.count) return FSR_OK;
    hipLaunchKernelGGL(k_fill_multi, dim3(blocks), dim3(256), 0, s, a);
    FSR_LAUNCH_CHECK("k_fill_multi");
    return FSR_OK;
}

fsr_status fill_f32(float* p, float value, uint64_t n, hipStream_t s) {
    if (n == 0) return FSR_OK;
    if (((uintptr_t)p & 15u) != 0) {
        uint32_t bits;
        __builtin_memcpy(&bits, &value, 4);
        return fill_u32(reinterpret_cast<uint32_t*>(p), bits, n, s);
    }
    uint64_t blocks = (n / 4 + 255) / 256;
    if (blocks < 1) blocks = 1;
    if (blocks > 8192) blocks = 8192;
    hipLaunchKernelGGL(k_fill_f32, dim3((uint32_t)blocks), dim3(256), 0, s, p, value, n);
    FSR_LAUNCH_CHECK("k_fill_f32");
    return FSR_OK;
}

}  // namespace fsr

extern "C" fsr_status fsr_exclusive_scan_noalloc(fsr_context* ctx, uint32_t* output,
                                                 const uint32_t* input, uint32_t num_elements,
                                                 uint32_t* d_total, void* stream) {
    if (!ctx || (num_elements && (!output || !input))) {
        fsr::set_error("fsr_exclusive_scan_noalloc: null context or buffer");
        return FSR_ERR_INVALID_ARG;
    }
    FSR_HIP(hipSetDevice(ctx->device));
    return fsr::scan_exclusive_noalloc(ctx->device, input, output, num_elements, d_total,
                                       (hipStream_t)stream);
}
