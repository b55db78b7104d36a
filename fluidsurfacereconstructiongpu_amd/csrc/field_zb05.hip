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
// contiguous in the sorted particle array).  The wave stages 64 candidates at a
// time in LDS (one coalesced 16-byte load per lane) and every lane then reads
// them as broadcasts, so the inner loop is uniform, branch-free and has no
// per-lane memory traffic:
//   t = 1 - |x - p|^2/R^2 = c_x + a_x . p' - q_p      (4 VALU ops)
// with p' = p - o and x' = x - o relative to the cell corner o, a = 2 x'/R^2,
// c_x = 1 - |x'|^2/R^2 and q_p = |p'|^2/R^2 precomputed at staging; then
// w = max(t, 0)^3 and the sums of w and w p' (7 ops): 11 VALU ops per pair.
// Coordinates relative to the cell keep the expansion well conditioned and
// make |x - xbar| free of the cancellation the absolute-coordinate sum has.
// Candidates are visited in the reference's order (z, y, x cells, then range
// order).  Work is distributed dynamically, one chunk per atomic fetch, each XCD
// starting on its own contiguous eighth of the chunk list.
//
// Culling: a chunk's vertices are a compact slab of its cell (the cell-major
// list is in ascending vertex id), so many block candidates lie farther than R
// from the chunk's bounding box.  At staging every candidate is tested against
// that box with R^2 (1 + 1e-4); survivors are compacted in order (ballot rank)
// into a 128-entry LDS queue and the inner loop runs over full groups of 64.
// A culled candidate is at least 1.00005 R from every vertex of the chunk, so
// its computed t is negative (the rounding of t is below 1e-5 there) and its
// weight exactly 0: the sums, and so phi, are bit-identical to evaluating the
// whole block, in the same order.
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

__device__ __forceinline__ float wave_min(float v) {
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) v = fminf(v, __shfl_xor(v, d, 64));
    return v;
}

__device__ __forceinline__ float wave_max(float v) {
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) v = fmaxf(v, __shfl_xor(v, d, 64));
    return v;
}

struct Sums {
    float w, x, y, z;
};

// n staged candidates (a multiple of 16) against this lane's vertex.
__device__ __forceinline__ void accumulate(const float4* __restrict__ q, uint32_t n, float ax,
                                           float ay, float az, float c0, Sums& s) {
    for (uint32_t t0 = 0; t0 < n; t0 += 16u) {
#pragma unroll 8
        for (int t = 0; t < 16; ++t) {
            const float4 p = q[t0 + t];
            float tt = __builtin_fmaf(az, p.z, c0 - p.w);
            tt = __builtin_fmaf(ay, p.y, tt);
            tt = __builtin_fmaf(ax, p.x, tt);
            const float w = fmaxf(tt, 0.0f);
            const float w3 = w * w * w;
            s.w += w3;
            s.x = __builtin_fmaf(w3, p.x, s.x);
            s.y = __builtin_fmaf(w3, p.y, s.y);
            s.z = __builtin_fmaf(w3, p.z, s.z);
        }
    }
}

__global__ __launch_bounds__(kFieldThreads, 5) void k_field_zb05(FieldArgs a, BandCompactOut bc) {
    __shared__ float4 queue[kFieldThreads / 64][128];
    __shared__ uint32_t row_start[kFieldThreads / 64][9];
    __shared__ uint32_t row_pref[kFieldThreads / 64][10];
    const uint32_t lane = __lane_id();
    const uint32_t wave = threadIdx.x >> 6;
    const uint32_t num_chunks = a.counters->num_chunks;
    const float inv_r2 = 1.0f / a.radius_sq;
    const float cull_r2 = a.radius_sq * 1.0001f;
    const int r = bc.radius;
    float4* q = queue[wave];

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
            uint32_t s = 0xFFFFFFFFu, e = 0;
            for (int xx = x0; xx <= x1; ++xx) {
                const fsr_index_range rg = a.ranges[rowc + xx];
                if (rg.start != 0xFFFFFFFFu) {
                    if (s == 0xFFFFFFFFu) s = rg.start;
                    e = rg.end;
                }
            }
            if (s != 0xFFFFFFFFu) { rs = s; len = e - s; }
        }
        const uint32_t incl = wave_inclusive_sum(len);
        const uint32_t total = __shfl(incl, 63, 64);
        if ((int)lane < nrows) {
            row_start[wave][lane] = rs;
            row_pref[wave][lane] = incl - len;
        }
        wave_lds_sync();

        // cell corner as the local origin
        const float ox = __builtin_fmaf((float)cx, a.sp.cell, a.sp.minx);
        const float oy = __builtin_fmaf((float)cy, a.sp.cell, a.sp.miny);
        const float oz = __builtin_fmaf((float)cz, a.sp.cell, a.sp.minz);

        uint32_t vid = 0;
        float vx = 0.f, vy = 0.f, vz = 0.f;
        if (lane < cnt) {
            vid = bc.cell_band[first + lane];
            const uint32_t ix = vid % a.sc.rx;
            const uint32_t iy = (vid / a.sc.rx) % a.sc.ry;
            const uint32_t iz = vid / (a.sc.rx * a.sc.ry);
            vx = vertex_coord(ix, a.sc.minx, a.sc.cell) - ox;
            vy = vertex_coord(iy, a.sc.miny, a.sc.cell) - oy;
            vz = vertex_coord(iz, a.sc.minz, a.sc.cell) - oz;
        }
        const float ax = 2.0f * vx * inv_r2, ay = 2.0f * vy * inv_r2, az = 2.0f * vz * inv_r2;
        const float c0 = 1.0f - __builtin_fmaf(vz, vz, __builtin_fmaf(vy, vy, vx * vx)) * inv_r2;
        // bounding box of the chunk's vertices
        const bool live = lane < cnt;
        const float bx0 = wave_min(live ? vx : 3.0e38f), bx1 = wave_max(live ? vx : -3.0e38f);
        const float by0 = wave_min(live ? vy : 3.0e38f), by1 = wave_max(live ? vy : -3.0e38f);
        const float bz0 = wave_min(live ? vz : 3.0e38f), bz1 = wave_max(live ? vz : -3.0e38f);

        Sums sum = {0.f, 0.f, 0.f, 0.f};
        uint32_t fill = 0;
        // this lane's candidate of the next round is loaded one round ahead, so
        // the load overlaps the culling and the accumulation of the current one
        int j = 0;  // row of this lane's candidate; g only grows, so the search resumes
        float4 next = make_float4(0.f, 0.f, 0.f, 0.f);
        if (lane < total) {
            while (j + 1 < nrows && row_pref[wave][j + 1] <= lane) ++j;
            next = a.sorted[row_start[wave][j] + (lane - row_pref[wave][j])];
        }
        for (uint32_t base = 0; base < total; base += 64u) {
            const uint32_t g = base + lane;
            const float4 p = next;
            if (g + 64u < total) {
                while (j + 1 < nrows && row_pref[wave][j + 1] <= g + 64u) ++j;
                next = a.sorted[row_start[wave][j] + (g + 64u - row_pref[wave][j])];
            }
            bool keep = false;
            float4 st;
            if (g < total) {
                const float px = p.x - ox, py = p.y - oy, pz = p.z - oz;
                const float dx = fmaxf(fmaxf(bx0 - px, px - bx1), 0.0f);
                const float dy = fmaxf(fmaxf(by0 - py, py - by1), 0.0f);
                const float dz = fmaxf(fmaxf(bz0 - pz, pz - bz1), 0.0f);
                keep = __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, dx * dx)) < cull_r2;
                st = make_float4(px, py, pz,
                                 __builtin_fmaf(pz, pz, __builtin_fmaf(py, py, px * px)) * inv_r2);
            }
            const uint64_t m = __ballot(keep);
            if (keep) q[fill + mask_rank(m)] = st;
            fill += (uint32_t)__popcll(m);
            if (fill >= 64u) {
                wave_lds_sync();
                accumulate(q, 64u, ax, ay, az, c0, sum);
                wave_lds_sync();
                const uint32_t rem = fill - 64u;
                if (lane < rem) q[lane] = q[64u + lane];
                wave_lds_sync();
                fill = rem;
            }
        }
        if (fill) {
            const uint32_t n = (fill + 15u) & ~15u;
            if (lane >= fill && lane < n) q[lane] = make_float4(0.f, 0.f, 0.f, 1.0e30f);
            wave_lds_sync();
            accumulate(q, n, ax, ay, az, c0, sum);
            wave_lds_sync();
        }
        if (lane < cnt) {
            float4 keep_out = make_float4(0.f, 0.f, 0.f, 0.f);
            if (sum.w > 0.0f) {
                const float inv = 1.0f / sum.w;
                const float dx = vx - sum.x * inv, dy = vy - sum.y * inv, dz = vz - sum.z * inv;
                a.phi[vid] =
                    __builtin_sqrtf(__builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, dx * dx))) -
                    a.particle_radius;
                keep_out = make_float4(dx, dy, dz, sum.w);
            }
            // for the adjoint: x - xbar and sum w, at the vertex's cell-major slot
            if (a.saved) a.saved[first + lane] = keep_out;
        }
    }
}

}  // namespace

fsr_status field_zb05(fsr_context* ctx, const FieldArgs& a, hipStream_t s) {
    // a.counters was cleared when the frame (or the stage call) began, and no
    // other kernel fetches through xcd_field
    const int blocks = persistent_blocks(ctx, kFieldThreads, 8, TUNE_FIELD);
    hipLaunchKernelGGL(k_field_zb05, dim3(blocks), dim3(kFieldThreads), 0, s, a, *a.bc);
    FSR_LAUNCH_CHECK("k_field_zb05");
    return FSR_OK;
}

}  // namespace fsr
