// Surface-particle detection and band seeding.
//
//   k_cell_surface      extractionOfSurfaceParticles (ReconstructionCUDA.cu:70-120)
//                       and, with `involve`, extractionOfSurfaceAndInvolveParticles
//                       (ReconstructionCUDA.cu:466-551): an occupied cell that is
//                       not on the last layer of the grid and has an empty cell in
//                       its clamped 3^3 block is a surface cell; with `involve`,
//                       every cell of the clamped 5^3 block of a surface cell gets
//                       its particle count.
//   k_flags_from_cells  marks the particles of surface cells (the reference's
//                       per-cell loop over the cell's range, written per particle).
//   k_colorfield        extractionOfSurfaceParticlesUsingColorField
//                       (ReconstructionCUDA.cu:909-971): surface iff fewer than 25
//                       particles with |r|^2 > 0 in the clamped 3^3 cell block, or
//                       |sum m/rho_j gradW_spiky(r) r_hat| > 3.
// Every surface particle then seeds the band: its scalar cell is set in a bit
// mask of "surface scalar cells" that band.hip dilates, or, when the dilation
// is not exact for it, it is appended to an exception list that band.hip
// treats with the reference's per-particle box loop.
#include "common.hpp"

namespace fsr {

namespace {

__device__ __forceinline__ int clampi(int v, int lo, int hi) {
    return v < lo ? lo : (v > hi ? hi : v);
}

// Truncating conversion of a cell coordinate that may lie outside the grid;
// NaN maps to 0 (what v_cvt_i32_f32 returns) and the value is clamped so the
// conversion stays defined.
__device__ __forceinline__ int trunc_cell(float q, int res) {
    if (!(q == q)) return 0;
    const float lo = -4.0f, hi = (float)res + 4.0f;
    q = q < lo ? lo : (q > hi ? hi : q);
    return (int)q;
}

__global__ __launch_bounds__(256) void k_cell_surface(const float* __restrict__ occupancy,
                                                      const fsr_index_range* __restrict__ ranges,
                                                      GridDev sp, uint8_t* __restrict__ surf_cell,
                                                      uint32_t* __restrict__ involve) {
    const uint32_t C = sp.rx * sp.ry * sp.rz;
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= C) return;
    const uint32_t x = c % sp.rx, y = (c / sp.rx) % sp.ry, z = c / (sp.rx * sp.ry);
    uint8_t surf = 0;
    if (x < sp.rx - 1 && y < sp.ry - 1 && z < sp.rz - 1 && occupancy[c] >= 1.0f) {
        const int x0 = clampi((int)x - 1, 0, sp.rx - 1), x1 = clampi((int)x + 1, 0, sp.rx - 1);
        const int y0 = clampi((int)y - 1, 0, sp.ry - 1), y1 = clampi((int)y + 1, 0, sp.ry - 1);
        const int z0 = clampi((int)z - 1, 0, sp.rz - 1), z1 = clampi((int)z + 1, 0, sp.rz - 1);
        for (int zz = z0; zz <= z1 && !surf; ++zz)
            for (int yy = y0; yy <= y1 && !surf; ++yy)
                for (int xx = x0; xx <= x1; ++xx)
                    if (occupancy[((uint32_t)zz * sp.ry + yy) * sp.rx + xx] < 1.0f) {
                        surf = 1;
                        break;
                    }
    }
    surf_cell[c] = surf;
    if (involve && surf) {
        const int x0 = clampi((int)x - 2, 0, sp.rx - 1), x1 = clampi((int)x + 2, 0, sp.rx - 1);
        const int y0 = clampi((int)y - 2, 0, sp.ry - 1), y1 = clampi((int)y + 2, 0, sp.ry - 1);
        const int z0 = clampi((int)z - 2, 0, sp.rz - 1), z1 = clampi((int)z + 2, 0, sp.rz - 1);
        for (int zz = z0; zz <= z1; ++zz)
            for (int yy = y0; yy <= y1; ++yy)
                for (int xx = x0; xx <= x1; ++xx) {
                    const uint32_t cc = ((uint32_t)zz * sp.ry + yy) * sp.rx + xx;
                    const fsr_index_range r = ranges[cc];
                    if (r.start != 0xFFFFFFFFu) involve[cc] = r.end - r.start;  // same value from every writer
                }
    }
}

// Band seeding for one particle (called by every thread of the block; `surf` =
// its flag).
__device__ __forceinline__ void seed_band(const SurfaceArgs& a, uint32_t i, float px, float py,
                                          float pz, bool surf) {
    if (!a.sbits) {  // detection only (uniform across the launch)
        if (a.st)
            block_count_add2(&a.st->num_surface_particles, surf ? 1u : 0u,
                             &a.st->band_exceptions, 0u);
        return;
    }
    const float qx = cell_coord(px, a.sc.minx, a.sc.cell);
    const float qy = cell_coord(py, a.sc.miny, a.sc.cell);
    const float qz = cell_coord(pz, a.sc.minz, a.sc.cell);
    const bool in_grid = qx > -1.0f && qx < (float)a.sc.rx && qy > -1.0f &&
                         qy < (float)a.sc.ry && qz > -1.0f && qz < (float)a.sc.rz;
    const bool exc = surf && (a.all_exceptions || !in_grid);
    if (surf && !exc) {
        const uint32_t ix = (uint32_t)(int)qx, iy = (uint32_t)(int)qy, iz = (uint32_t)(int)qz;
        const uint32_t wx = row_words(a.sc.rx);
        const uint64_t w = ((uint64_t)iz * a.sc.ry + iy) * wx + (ix >> 5);
        atomicOr(&a.sbits[w], 1u << (ix & 31u));
    }
    const uint64_t m = __ballot(exc);
    if (m) {
        uint32_t base = 0;
        if (__lane_id() == (uint32_t)(__ffsll((unsigned long long)m) - 1))
            base = atomicAdd(&a.counters->num_exceptions, (uint32_t)__popcll(m));
        base = __shfl(base, __ffsll((unsigned long long)m) - 1, 64);
        if (exc) a.exceptions[base + mask_rank(m)] = i;
    }
    block_count_add2(&a.st->num_surface_particles, surf ? 1u : 0u, &a.st->band_exceptions,
                     exc ? 1u : 0u);
}

__global__ __launch_bounds__(256) void k_flags_from_cells(SurfaceArgs a,
                                                          const uint8_t* __restrict__ surf_cell) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const bool live = i < a.n;
    bool surf = false;
    float4 p = make_float4(0.f, 0.f, 0.f, 0.f);
    if (live) {
        const uint32_t key = a.keys[i];
        const uint32_t C = a.sp.rx * a.sp.ry * a.sp.rz;
        surf = key < C && surf_cell[key] != 0;
        a.flags[i] = surf ? 1u : 0u;
        p = a.sorted[i];
    }
    seed_band(a, i, p.x, p.y, p.z, surf);
}

__global__ __launch_bounds__(256) void k_colorfield(SurfaceArgs a, float mass) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const bool live = i < a.n;
    bool surf = false;
    float4 p = make_float4(0.f, 0.f, 0.f, 0.f);
    if (live) {
        p = a.sorted[i];
        const int cx = trunc_cell(cell_coord(p.x, a.sp.minx, a.sp.cell), a.sp.rx);
        const int cy = trunc_cell(cell_coord(p.y, a.sp.miny, a.sp.cell), a.sp.ry);
        const int cz = trunc_cell(cell_coord(p.z, a.sp.minz, a.sp.cell), a.sp.rz);
        const int x0 = clampi(cx - 1, 0, a.sp.rx - 1), x1 = clampi(cx + 1, 0, a.sp.rx - 1);
        const int y0 = clampi(cy - 1, 0, a.sp.ry - 1), y1 = clampi(cy + 1, 0, a.sp.ry - 1);
        const int z0 = clampi(cz - 1, 0, a.sp.rz - 1), z1 = clampi(cz + 1, 0, a.sp.rz - 1);
        const float h = a.h;
        const float inv_h = 1.0f / h;
        const float gcoef = -45.0f / (3.14159265358979f * (h * h) * (h * h));
        uint32_t count = 0;
        float gx = 0.f, gy = 0.f, gz = 0.f;
        for (int zz = z0; zz <= z1; ++zz)
            for (int yy = y0; yy <= y1; ++yy) {
                const uint32_t row = ((uint32_t)zz * a.sp.ry + yy) * a.sp.rx;
                uint32_t s = 0xFFFFFFFFu, e = 0;
                for (int xx = x0; xx <= x1; ++xx) {
                    const fsr_index_range r = a.ranges[row + xx];
                    if (r.start != 0xFFFFFFFFu) {
                        if (s == 0xFFFFFFFFu) s = r.start;
                        e = r.end;
                    }
                }
                if (s == 0xFFFFFFFFu) continue;
                for (uint32_t j = s; j < e; ++j) {
                    const float4 q = a.sorted[j];
                    const float rx = p.x - q.x, ry = p.y - q.y, rz = p.z - q.z;
                    const float d2 = __builtin_fmaf(rz, rz, __builtin_fmaf(ry, ry, rx * rx));
                    if (d2 > 0.0f) {
                        ++count;
                        const float dist = __builtin_sqrtf(d2);
                        if (dist < h) {
                            const float x = 1.0f - dist * inv_h;
                            const float c = mass / q.w * gcoef * x * x * __builtin_amdgcn_rsqf(d2);
                            gx = __builtin_fmaf(c, rx, gx);
                            gy = __builtin_fmaf(c, ry, gy);
                            gz = __builtin_fmaf(c, rz, gz);
                        }
                    }
                }
            }
        const float g2 = __builtin_fmaf(gz, gz, __builtin_fmaf(gy, gy, gx * gx));
        surf = count < 25u || __builtin_sqrtf(g2) > 3.0f;
        a.flags[i] = surf ? 1u : 0u;
    }
    seed_band(a, i, p.x, p.y, p.z, surf);
}

// band seeding from given flags (stage entry point of the band)
__global__ __launch_bounds__(256) void k_seed_flags(SurfaceArgs a) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const bool live = i < a.n;
    bool surf = false;
    float4 p = make_float4(0.f, 0.f, 0.f, 0.f);
    if (live) {
        surf = a.flags[i] != 0;
        p = a.sorted[i];
    }
    seed_band(a, i, p.x, p.y, p.z, surf);
}

}  // namespace

fsr_status surface_seed_from_flags(fsr_context* ctx, const SurfaceArgs& a, hipStream_t s) {
    (void)ctx;
    if (a.n) {
        hipLaunchKernelGGL(k_seed_flags, dim3((a.n + 255) / 256), dim3(256), 0, s, a);
        FSR_LAUNCH_CHECK("k_seed_flags");
    }
    return FSR_OK;
}

fsr_status surface_occupancy(fsr_context* ctx, const SurfaceArgs& a, uint8_t* surf_cell,
                             uint32_t* involve, hipStream_t s) {
    (void)ctx;
    const uint32_t C = a.sp.rx * a.sp.ry * a.sp.rz;
    if (involve) FSR_TRY(fill_u32(involve, 0u, C, s));
    hipLaunchKernelGGL(k_cell_surface, dim3((C + 255) / 256), dim3(256), 0, s, a.occupancy,
                       a.ranges, a.sp, surf_cell, involve);
    FSR_LAUNCH_CHECK("k_cell_surface");
    if (a.n) {
        hipLaunchKernelGGL(k_flags_from_cells, dim3((a.n + 255) / 256), dim3(256), 0, s, a,
                           (const uint8_t*)surf_cell);
        FSR_LAUNCH_CHECK("k_flags_from_cells");
    }
    return FSR_OK;
}

fsr_status surface_colorfield(fsr_context* ctx, const SurfaceArgs& a, float mass,
                              hipStream_t s) {
    (void)ctx;
    if (a.n) {
        hipLaunchKernelGGL(k_colorfield, dim3((a.n + 255) / 256), dim3(256), 0, s, a, mass);
        FSR_LAUNCH_CHECK("k_colorfield");
    }
    return FSR_OK;
}

}  // namespace fsr
