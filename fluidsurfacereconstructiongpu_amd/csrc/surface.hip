// Surface-particle detection and band seeding.
//
//   k_cell_surface      extractionOfSurfaceParticles (ReconstructionCUDA.cu:70-120)
//                       and the surface half of extractionOfSurfaceAndInvolveParticles
//                       (ReconstructionCUDA.cu:466-551): an occupied cell that is
//                       not on the last layer of the grid and has an empty cell in
//                       its clamped 3^3 block is a surface cell.
//   k_involve_window    the involve half of the latter: every cell of the clamped
//                       5^3 block of a surface cell gets its particle count (the
//                       block test as three separable 5-wide passes).
//   k_flags_from_cells  marks the particles of surface cells (the reference's
//                       per-cell loop over the cell's range, written per particle).
//   k_colorfield        extractionOfSurfaceParticlesUsingColorField
//                       (ReconstructionCUDA.cu:909-971): surface iff fewer than 25
//                       particles with |r|^2 > 0 in the clamped 3^3 cell block, or
//                       |sum m/rho_j gradW_spiky(r) r_hat| > 3.  With the radius-count
//                       surface mode (either method): surface iff fewer than 25
//                       particles at 0 < |r|^2 < h^2 (FP32 squared distances).
// Every surface particle then seeds the band: its scalar cell is set in a bit
// mask of "surface scalar cells" that band.hip dilates, or, when the dilation
// is not exact for it, it is appended to an exception list that band.hip
// treats with the reference's per-particle box loop.
#include <algorithm>

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
                                                      GridDev sp, uint8_t* __restrict__ surf_cell) {
    const uint32_t C = sp.rx * sp.ry * sp.rz;
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= C) return;
    const uint32_t x = c % sp.rx, y = (c / sp.rx) % sp.ry, z = c / (sp.rx * sp.ry);
    uint8_t surf = 0;
    if (x < sp.rx - 1 && y < sp.ry - 1 && z < sp.rz - 1 && occupancy[c] >= 1.0f) {
        // the clamped 3^3 block as 27 clamped offsets (a clamped offset repeats
        // a cell of the block, which leaves the test unchanged), all read with
        // no early exit, so that their loads issue together
#pragma unroll
        for (int dz = -1; dz <= 1; ++dz) {
            const uint32_t zz = (uint32_t)clampi((int)z + dz, 0, sp.rz - 1);
#pragma unroll
            for (int dy = -1; dy <= 1; ++dy) {
                const uint32_t yy = (uint32_t)clampi((int)y + dy, 0, sp.ry - 1);
#pragma unroll
                for (int dx = -1; dx <= 1; ++dx) {
                    const uint32_t xx = (uint32_t)clampi((int)x + dx, 0, sp.rx - 1);
                    surf |= occupancy[(zz * sp.ry + yy) * sp.rx + xx] < 1.0f ? 1 : 0;
                }
            }
        }
    }
    surf_cell[c] = surf;
}

// Involve counts (K4, the Yu13 half): an occupied cell within the clamped 5^3
// block of some surface cell gets its particle count, every other cell 0.  The
// block test is separable: three passes of a 5-wide OR along x, y and z over
// the surface-cell bytes (clamped to the grid, as the block is); the z pass
// writes the counts.  Every cell is written, so no clearing pass is needed.
template <int kAxis>
__global__ __launch_bounds__(256) void k_involve_window(const uint8_t* __restrict__ in,
                                                       uint8_t* __restrict__ out, GridDev sp,
                                                       const fsr_index_range* __restrict__ ranges,
                                                       uint32_t* __restrict__ involve) {
    const uint32_t C = sp.rx * sp.ry * sp.rz;
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= C) return;
    const uint32_t x = c % sp.rx, y = (c / sp.rx) % sp.ry, z = c / (sp.rx * sp.ry);
    const uint32_t pos = kAxis == 0 ? x : (kAxis == 1 ? y : z);
    const uint32_t res = kAxis == 0 ? sp.rx : (kAxis == 1 ? sp.ry : sp.rz);
    const uint32_t stride = kAxis == 0 ? 1u : (kAxis == 1 ? sp.rx : sp.rx * sp.ry);
    const uint32_t lo = pos >= 2u ? pos - 2u : 0u, hi = min(pos + 2u, res - 1u);
    uint8_t any = 0;
    for (uint32_t p = lo; p <= hi; ++p) any |= in[c + (p - pos) * stride];
    if (kAxis < 2) {
        out[c] = any;
    } else {
        const fsr_index_range r = ranges[c];
        involve[c] = (any && r.start != 0xFFFFFFFFu) ? r.end - r.start : 0u;
    }
}

// Band seeding of one particle: its scalar cell is set in the surface-cell
// byte mask, or the particle is appended to the exception list (exact path).
// Called with all 64 lanes of the wave active (lanes without a particle pass
// surf = false); returns whether the particle became an exception.
__device__ __forceinline__ bool seed_particle(const SurfaceArgs& a, uint32_t i, float px,
                                              float py, float pz, bool surf) {
    const float qx = cell_coord(px, a.sc.minx, a.sc.cell);
    const float qy = cell_coord(py, a.sc.miny, a.sc.cell);
    const float qz = cell_coord(pz, a.sc.minz, a.sc.cell);
    const bool in_grid = qx > -1.0f && qx < (float)a.sc.rx && qy > -1.0f &&
                         qy < (float)a.sc.ry && qz > -1.0f && qz < (float)a.sc.rz;
    const bool exc = surf && (a.all_exceptions || !in_grid);
    if (surf && !exc) {
        // one byte per scalar cell, rows padded to whole mask words: a plain
        // byte store (every writer stores 1), no read-modify-write
        const uint32_t ix = (uint32_t)(int)qx, iy = (uint32_t)(int)qy, iz = (uint32_t)(int)qz;
        const uint32_t wx = row_words(a.sc.rx);
        reinterpret_cast<uint8_t*>(a.sbits)[(((uint64_t)iz * a.sc.ry + iy) * wx << 5) + ix] = 1u;
    }
    const uint64_t m = __ballot(exc);
    if (m) {
        uint32_t base = 0;
        if (__lane_id() == (uint32_t)(__ffsll((unsigned long long)m) - 1))
            base = atomicAdd(&a.counters->num_exceptions, (uint32_t)__popcll(m));
        base = __shfl(base, __ffsll((unsigned long long)m) - 1, 64);
        if (exc) a.exceptions[base + mask_rank(m)] = i;
    }
    return exc;
}

// Elements per thread and grid-stride step of the per-particle kernels below:
// their loads are issued together, so one step waits for memory once per
// dependent load rather than once per element.
constexpr uint32_t kPerThread = 4;

// Surface flags from the surface hash cells (ZB05 occupancy method), band
// seeding, and the frame's surface / exception counts.  Grid-stride with a
// block-uniform trip count over a bounded grid: the two counters take one
// atomic per block, and thousands of atomics on one address serialise.
__global__ __launch_bounds__(256) void k_flags_from_cells(SurfaceArgs a,
                                                          const uint8_t* __restrict__ surf_cell) {
    const uint32_t C = a.sp.rx * a.sp.ry * a.sp.rz;
    const uint32_t span = blockDim.x * kPerThread;
    uint32_t n_surf = 0, n_exc = 0;
    for (uint32_t base = blockIdx.x * span; base < a.n; base += gridDim.x * span) {
        uint32_t key[kPerThread];
#pragma unroll
        for (uint32_t k = 0; k < kPerThread; ++k) {
            const uint32_t i = base + k * blockDim.x + threadIdx.x;
            key[k] = i < a.n ? a.keys[i] : C;
        }
        bool surf[kPerThread];
#pragma unroll
        for (uint32_t k = 0; k < kPerThread; ++k) {
            const uint32_t i = base + k * blockDim.x + threadIdx.x;
            surf[k] = key[k] < C && surf_cell[key[k]] != 0;
            if (i < a.n) a.flags[i] = surf[k] ? 1u : 0u;
        }
        // positions only of the surface particles (the band seeds)
        float4 p[kPerThread];
#pragma unroll
        for (uint32_t k = 0; k < kPerThread; ++k) {
            const uint32_t i = base + k * blockDim.x + threadIdx.x;
            p[k] = (a.sbits && surf[k]) ? a.sorted[i] : make_float4(0.f, 0.f, 0.f, 0.f);
        }
#pragma unroll
        for (uint32_t k = 0; k < kPerThread; ++k) {
            const uint32_t i = base + k * blockDim.x + threadIdx.x;
            const bool exc = a.sbits ? seed_particle(a, i, p[k].x, p[k].y, p[k].z, surf[k]) : false;
            n_surf += surf[k] ? 1u : 0u;
            n_exc += exc ? 1u : 0u;
        }
    }
    if (a.st) block_count_add2(&a.st->num_surface_particles, n_surf, &a.st->band_exceptions, n_exc);
}

// Particles outside the hash grid (sorted last, in no cell): one thread each,
// over the clamped block around their own clamped cell.  Grid-stride over the
// device-side count; the surface / exception counts are added once per block.
__global__ __launch_bounds__(256) void k_colorfield_tail(SurfaceArgs a, float mass) {
    const uint32_t first = a.n - a.st->num_out_of_grid;
    const float h = a.h;
    const float inv_h = 1.0f / h;
    const float gcoef = -45.0f / (3.14159265358979f * (h * h) * (h * h));
    uint32_t n_surf = 0, n_exc = 0;
    // block-uniform trip count: seed_particle needs whole waves
    for (uint32_t base = first + blockIdx.x * blockDim.x; base < a.n;
         base += gridDim.x * blockDim.x) {
        const uint32_t i = base + threadIdx.x;
        bool surf = false;
        float4 p = make_float4(0.f, 0.f, 0.f, 0.f);
        if (i < a.n) {
            p = a.sorted[i];
            const int cx = trunc_cell(cell_coord(p.x, a.sp.minx, a.sp.cell), a.sp.rx);
            const int cy = trunc_cell(cell_coord(p.y, a.sp.miny, a.sp.cell), a.sp.ry);
            const int cz = trunc_cell(cell_coord(p.z, a.sp.minz, a.sp.cell), a.sp.rz);
            const int x0 = clampi(cx - 1, 0, a.sp.rx - 1), x1 = clampi(cx + 1, 0, a.sp.rx - 1);
            const int y0 = clampi(cy - 1, 0, a.sp.ry - 1), y1 = clampi(cy + 1, 0, a.sp.ry - 1);
            const int z0 = clampi(cz - 1, 0, a.sp.rz - 1), z1 = clampi(cz + 1, 0, a.sp.rz - 1);
            uint32_t count = 0, count_r = 0;
            const float h2 = h * h;
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
                            if (d2 < h2) {
                                ++count_r;
                                const float ri = __builtin_amdgcn_rsqf(d2);
                                const float x = 1.0f - d2 * ri * inv_h;
                                const float c = mass / q.w * gcoef * x * x * ri;
                                gx = __builtin_fmaf(c, rx, gx);
                                gy = __builtin_fmaf(c, ry, gy);
                                gz = __builtin_fmaf(c, rz, gz);
                            }
                        }
                    }
                }
            const float g2 = __builtin_fmaf(gz, gz, __builtin_fmaf(gy, gy, gx * gx));
            surf = a.radius_count ? count_r < 25u : (count < 25u || __builtin_sqrtf(g2) > 3.0f);
            a.flags[i] = surf ? 1u : 0u;
        }
        const bool exc = a.sbits ? seed_particle(a, i, p.x, p.y, p.z, surf) : false;
        n_surf += surf ? 1u : 0u;
        n_exc += exc ? 1u : 0u;
    }
    block_count_add2(&a.st->num_surface_particles, n_surf, &a.st->band_exceptions, n_exc);
}

__device__ __forceinline__ float wave_minf(float v) {
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) v = fminf(v, __shfl_xor(v, d, 64));
    return v;
}

__device__ __forceinline__ float wave_maxf(float v) {
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) v = fmaxf(v, __shfl_xor(v, d, 64));
    return v;
}

__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

struct ColorChunks {
    const uint32_t* chunk_cell;
    const uint32_t* chunk_off;
    const uint32_t* num_chunks;
    uint32_t* next;  // per-XCD fetch counters
    uint32_t* work;  // per wave: candidates staged, kept by the box test, pairs within h
};

// The colour-field test of the particles of the hash cells, a wave per chunk of
// up to 64 particles of one cell: the 9 x-runs of the cell's 3^3 block are
// staged through LDS and read as broadcasts.  Candidates farther than
// h * (1 + 5e-5) from the bounding box of the chunk's particles can only add to
// the neighbour count (d^2 > 0, dist >= h): they are counted and dropped, the
// others are compacted in order and evaluated with the per-thread kernel's
// arithmetic, so counts, gradients and flags are identical to it.  Both take
// the reference's test |r| < h as |r|^2 < h^2 and the distance as
// |r|^2 rsqrt(|r|^2), one reciprocal square root per pair within h and no
// square root: the distance enters the gradient through (1 - |r|/h)^2, which
// vanishes where the two forms of the test could differ.
__global__ __launch_bounds__(256) void k_colorfield(SurfaceArgs a, float mass, ColorChunks cl) {
    __shared__ float4 queue[4][128];
    const uint32_t lane = __lane_id();
    const uint32_t wave = threadIdx.x >> 6;
    float4* q = queue[wave];
    const uint32_t nch = *cl.num_chunks;
    const float h = a.h;
    const float inv_h = 1.0f / h;
    const float gcoef = -45.0f / (3.14159265358979f * (h * h) * (h * h));
    const float cull2 = (h * h) * 1.0001f;
    const float h2 = h * h;
    uint32_t n_surf = 0, n_exc = 0;
    // work of this wave: candidates staged and kept (wave-uniform), and this
    // lane's pairs within h
    uint32_t n_staged = 0, n_kept = 0, n_inh = 0;
    uint32_t step = 0;
    for (;;) {
        const uint32_t ch = next_item_xcd(cl.next, nch, step);
        if (ch >= nch) break;
        const uint32_t c = cl.chunk_cell[ch];
        const uint32_t k = ch - cl.chunk_off[c];
        const fsr_index_range own = a.ranges[c];
        const uint32_t i = own.start + 64u * k + lane;
        const bool live = i < own.end;
        const int cx = (int)(c % a.sp.rx), cy = (int)((c / a.sp.rx) % a.sp.ry);
        const int cz = (int)(c / (a.sp.rx * a.sp.ry));
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
            uint32_t s0 = 0xFFFFFFFFu, e0 = 0;
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                if (rg[d].start != 0xFFFFFFFFu) {
                    if (s0 == 0xFFFFFFFFu) s0 = rg[d].start;
                    e0 = rg[d].end;
                }
            }
            if (s0 != 0xFFFFFFFFu) { rs = s0; len = e0 - s0; }
        }
        // lane j < nrows holds row j's first candidate index (rp) and the
        // offset from candidate index to sorted index in that row (ro)
        const uint32_t incl = wave_inclusive_sum(len);
        const uint32_t total = (uint32_t)__builtin_amdgcn_readlane((int)incl, 63);
        const uint32_t rp = incl - len;
        const uint32_t ro = rs - rp;
        // Row of candidate index gg of the staging round that starts at nb:
        // jc, a row whose first index is at most nb, only grows; the rows
        // that start inside the round are visited with wave-uniform reads, and
        // a lane takes the last one that starts at or before gg (an empty row
        // shares its first index with the next row, visited after it).
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
        float4 p = make_float4(0.f, 0.f, 0.f, 0.f);
        if (live) p = a.sorted[i];
        const float bx0 = wave_minf(live ? p.x : 3.0e38f), bx1 = wave_maxf(live ? p.x : -3.0e38f);
        const float by0 = wave_minf(live ? p.y : 3.0e38f), by1 = wave_maxf(live ? p.y : -3.0e38f);
        const float bz0 = wave_minf(live ? p.z : 3.0e38f), bz1 = wave_maxf(live ? p.z : -3.0e38f);
        uint32_t count = 0, count_r = 0, kept = 0, fill = 0;
        float gx = 0.f, gy = 0.f, gz = 0.f;
        auto run = [&](uint32_t n) {
            for (uint32_t t = 0; t < n; ++t) {
                const float4 qq = q[t];
                const float rx = p.x - qq.x, ry = p.y - qq.y, rz = p.z - qq.z;
                const float d2 = __builtin_fmaf(rz, rz, __builtin_fmaf(ry, ry, rx * rx));
                if (d2 > 0.0f) {
                    ++count;
                    if (d2 < h2) {
                        ++count_r;
                        const float ri = __builtin_amdgcn_rsqf(d2);
                        const float x = 1.0f - d2 * ri * inv_h;
                        const float cc = qq.w * x * x * ri;
                        gx = __builtin_fmaf(cc, rx, gx);
                        gy = __builtin_fmaf(cc, ry, gy);
                        gz = __builtin_fmaf(cc, rz, gz);
                    }
                }
            }
        };
        // each lane's candidate is loaded one staging round ahead
        float4 next = make_float4(0.f, 0.f, 0.f, 0.f);
        {
            const uint32_t off = row_offset(0u, lane);
            if (lane < total) next = a.sorted[off + lane];
        }
        for (uint32_t base = 0; base < total; base += 64u) {
            const uint32_t g = base + lane;
            const float4 cand = next;
            if (base + 64u < total) {
                const uint32_t off = row_offset(base + 64u, g + 64u);
                if (g + 64u < total) next = a.sorted[off + g + 64u];
            }
            bool keep = false;
            float4 st;
            if (g < total) {
                // c - clamp(c, lo, hi): its square equals that of max(lo - c, c - hi, 0)
                const float dx = cand.x - __builtin_amdgcn_fmed3f(cand.x, bx0, bx1);
                const float dy = cand.y - __builtin_amdgcn_fmed3f(cand.y, by0, by1);
                const float dz = cand.z - __builtin_amdgcn_fmed3f(cand.z, bz0, bz1);
                keep = __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, dx * dx)) < cull2;
                // (m / rho_j) * gcoef, in the per-thread kernel's operation order
                st = make_float4(cand.x, cand.y, cand.z, mass / cand.w * gcoef);
            }
            const uint64_t m = __ballot(keep);
            if (keep) q[fill + mask_rank(m)] = st;
            const uint32_t added = (uint32_t)__popcll(m);
            fill += added;
            kept += added;
            if (fill >= 64u) {
                wave_lds_sync();
                run(64u);
                wave_lds_sync();
                const uint32_t rem = fill - 64u;
                if (lane < rem) q[lane] = q[64u + lane];
                wave_lds_sync();
                fill = rem;
            }
        }
        if (fill) {
            wave_lds_sync();
            run(fill);
            wave_lds_sync();
        }
        // every dropped candidate has d^2 > 0: it counts as a neighbour
        count += total - kept;
        n_inh += count_r;
        n_staged += total;
        n_kept += kept;
        bool surf = false;
        if (live) {
            const float g2 = __builtin_fmaf(gz, gz, __builtin_fmaf(gy, gy, gx * gx));
            // dropped candidates lie beyond h: they never count in count_r
            surf = a.radius_count ? count_r < 25u : (count < 25u || __builtin_sqrtf(g2) > 3.0f);
            a.flags[i] = surf ? 1u : 0u;
        }
        bool exc = false;
        if (a.sbits) exc = seed_particle(a, i, p.x, p.y, p.z, surf);
        n_surf += surf ? 1u : 0u;
        n_exc += exc ? 1u : 0u;
    }
    const uint32_t inh = wave_sum(n_inh);
    if (lane == 0) {
        uint32_t* w = cl.work + 3u * (blockIdx.x * (blockDim.x >> 6) + wave);
        w[0] = n_staged;
        w[1] = n_kept;
        w[2] = inh;
    }
    block_count_add2(&a.st->num_surface_particles, n_surf, &a.st->band_exceptions, n_exc);
}

// band seeding from given flags (stage entry point of the band)
__global__ __launch_bounds__(256) void k_seed_flags(SurfaceArgs a) {
    const uint32_t span = blockDim.x * kPerThread;
    uint32_t n_surf = 0, n_exc = 0;
    for (uint32_t base = blockIdx.x * span; base < a.n; base += gridDim.x * span) {
        bool surf[kPerThread];
        float4 p[kPerThread];
#pragma unroll
        for (uint32_t k = 0; k < kPerThread; ++k) {
            const uint32_t i = base + k * blockDim.x + threadIdx.x;
            surf[k] = false;
            p[k] = make_float4(0.f, 0.f, 0.f, 0.f);
            if (i < a.n) {
                surf[k] = a.flags[i] != 0;
                p[k] = a.sorted[i];
            }
        }
#pragma unroll
        for (uint32_t k = 0; k < kPerThread; ++k) {
            const uint32_t i = base + k * blockDim.x + threadIdx.x;
            const bool exc = a.sbits ? seed_particle(a, i, p[k].x, p[k].y, p[k].z, surf[k]) : false;
            n_surf += surf[k] ? 1u : 0u;
            n_exc += exc ? 1u : 0u;
        }
    }
    if (a.st) block_count_add2(&a.st->num_surface_particles, n_surf, &a.st->band_exceptions, n_exc);
}

}  // namespace

// Grid of the per-particle kernels that end in block counter atomics.
static uint32_t count_blocks(fsr_context* ctx, uint32_t n) {
    const uint32_t cap = (uint32_t)persistent_blocks(ctx, 256, 4, TUNE_BAND);
    return std::max(1u, std::min((n + 256u * kPerThread - 1u) / (256u * kPerThread), cap));
}

fsr_status surface_seed_from_flags(fsr_context* ctx, const SurfaceArgs& a, hipStream_t s) {
    if (a.n) {
        hipLaunchKernelGGL(k_seed_flags, dim3(count_blocks(ctx, a.n)), dim3(256), 0, s, a);
        FSR_LAUNCH_CHECK("k_seed_flags");
    }
    return FSR_OK;
}

fsr_status surface_occupancy(fsr_context* ctx, const SurfaceArgs& a, uint8_t* surf_cell,
                             uint32_t* involve, hipStream_t s) {
    const uint32_t C = a.sp.rx * a.sp.ry * a.sp.rz;
    hipLaunchKernelGGL(k_cell_surface, dim3((C + 255) / 256), dim3(256), 0, s, a.occupancy,
                       a.sp, surf_cell);
    FSR_LAUNCH_CHECK("k_cell_surface");
    if (involve) {
        uint8_t* tmp;
        FSR_TRY(arena_typed(ctx, A_INVOLVE_TMP, 2ull * C, &tmp));
        const dim3 g((C + 255) / 256);
        hipLaunchKernelGGL(k_involve_window<0>, g, dim3(256), 0, s, (const uint8_t*)surf_cell,
                           tmp, a.sp, a.ranges, involve);
        hipLaunchKernelGGL(k_involve_window<1>, g, dim3(256), 0, s, (const uint8_t*)tmp,
                           tmp + C, a.sp, a.ranges, involve);
        hipLaunchKernelGGL(k_involve_window<2>, g, dim3(256), 0, s, (const uint8_t*)(tmp + C),
                           tmp, a.sp, a.ranges, involve);
        FSR_LAUNCH_CHECK("k_involve_window");
    }
    if (a.n) {
        hipLaunchKernelGGL(k_flags_from_cells, dim3(count_blocks(ctx, a.n)), dim3(256), 0, s, a,
                           (const uint8_t*)surf_cell);
        FSR_LAUNCH_CHECK("k_flags_from_cells");
    }
    return FSR_OK;
}

fsr_status surface_colorfield(fsr_context* ctx, const SurfaceArgs& a, float mass,
                              hipStream_t s) {
    if (!a.n) return FSR_OK;
    const uint32_t C = a.sp.rx * a.sp.ry * a.sp.rz;
    const uint64_t max_chunks = (uint64_t)a.n / 64 + C + 1;
    uint32_t* scratch;
    FSR_TRY(arena_typed(ctx, A_SURF_CHUNKS, 2ull * C + max_chunks, &scratch));
    uint32_t* total = &a.counters->aux[4];
    FSR_TRY(cell_chunk_list(ctx, a.ranges, nullptr, C, scratch, scratch + C, scratch + 2ull * C,
                            total, s));
    const int blocks = persistent_blocks(ctx, 256, 8, TUNE_FIELD);
    uint32_t* work;
    FSR_TRY(arena_typed(ctx, A_COLOR_WORK, 3ull * blocks * 4, &work));
    ColorChunks cl{scratch + 2ull * C, scratch + C, total, a.counters->xcd_next, work};
    FSR_TRY(reset_xcd(a.counters, s));
    hipLaunchKernelGGL(k_colorfield, dim3(blocks), dim3(256), 0, s, a, mass, cl);
    FSR_LAUNCH_CHECK("k_colorfield");
    // particles outside the grid (a device-side count)
    hipLaunchKernelGGL(k_colorfield_tail, dim3(64), dim3(256), 0, s, a, mass);
    FSR_LAUNCH_CHECK("k_colorfield_tail");
    return FSR_OK;
}

}  // namespace fsr
