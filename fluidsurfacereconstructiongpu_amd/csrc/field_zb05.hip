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
// order).  Work is distributed dynamically, one chunk per atomic fetch.
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

__global__ __launch_bounds__(kFieldThreads) void k_field_zb05(FieldArgs a, BandCompactOut bc) {
    __shared__ float4 stage[kFieldThreads / 64][64];
    __shared__ uint32_t row_start[kFieldThreads / 64][9];
    __shared__ uint32_t row_pref[kFieldThreads / 64][10];
    const uint32_t lane = __lane_id();
    const uint32_t wave = threadIdx.x >> 6;
    const uint32_t num_chunks = a.counters->num_chunks;
    const float inv_r2 = 1.0f / a.radius_sq;
    const int r = bc.radius;

    for (;;) {
        uint32_t ch = 0;
        if (lane == 0) ch = atomicAdd(&a.counters->field_next_chunk, 1u);
        ch = __shfl(ch, 0, 64);
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

        float sw = 0.f, sx = 0.f, sy = 0.f, sz = 0.f;
        for (uint32_t base = 0; base < total; base += 64u) {
            const uint32_t g = base + lane;
            float4 st = make_float4(0.f, 0.f, 0.f, 1.0e30f);
            if (g < total) {
                int j = 0;
                while (j + 1 < nrows && row_pref[wave][j + 1] <= g) ++j;
                const float4 p = a.sorted[row_start[wave][j] + (g - row_pref[wave][j])];
                const float px = p.x - ox, py = p.y - oy, pz = p.z - oz;
                st = make_float4(px, py, pz,
                                 __builtin_fmaf(pz, pz, __builtin_fmaf(py, py, px * px)) * inv_r2);
            }
            stage[wave][lane] = st;
            wave_lds_sync();
#pragma unroll 16
            for (int t = 0; t < 64; ++t) {
                const float4 p = stage[wave][t];
                float tt = __builtin_fmaf(az, p.z, c0 - p.w);
                tt = __builtin_fmaf(ay, p.y, tt);
                tt = __builtin_fmaf(ax, p.x, tt);
                const float w = fmaxf(tt, 0.0f);
                const float w3 = w * w * w;
                sw += w3;
                sx = __builtin_fmaf(w3, p.x, sx);
                sy = __builtin_fmaf(w3, p.y, sy);
                sz = __builtin_fmaf(w3, p.z, sz);
            }
            wave_lds_sync();
        }
        if (lane < cnt) {
            float4 keep = make_float4(0.f, 0.f, 0.f, 0.f);
            if (sw > 0.0f) {
                const float inv = 1.0f / sw;
                const float dx = vx - sx * inv, dy = vy - sy * inv, dz = vz - sz * inv;
                a.phi[vid] =
                    __builtin_sqrtf(__builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, dx * dx))) -
                    a.particle_radius;
                keep = make_float4(dx, dy, dz, sw);
            }
            // for the adjoint: x - xbar and sum w, at the vertex's cell-major slot
            if (a.saved) a.saved[first + lane] = keep;
        }
    }
}

}  // namespace

fsr_status field_zb05(fsr_context* ctx, const FieldArgs& a, hipStream_t s) {
    const int blocks = persistent_blocks(ctx, kFieldThreads, 8, TUNE_FIELD);
    hipLaunchKernelGGL(k_field_zb05, dim3(blocks), dim3(kFieldThreads), 0, s, a, *a.bc);
    FSR_LAUNCH_CHECK("k_field_zb05");
    return FSR_OK;
}

}  // namespace fsr
