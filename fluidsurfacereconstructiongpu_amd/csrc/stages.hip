// Stage-level entry points with the reference launchers' buffers and meaning
// (ReconstructionCUDA.cuh:17-248), for callers that drive the stages
// themselves.  They run the same kernels as fsr_reconstruct_frame; the
// adapters here convert between the reference's buffer formats (dense u32
// flag grids, float3 arrays, index lists) and the internal ones (bit masks,
// float4 records, cell-major work lists).  The launch geometry arguments of the
// reference (dim3 grid, block) are not taken: every kernel chooses its own.
#include <hip/hip_runtime.h>

#include <cstring>

#include "common.hpp"
#include "mc_tables.h"

namespace fsr {

namespace {

__device__ __forceinline__ int clampi(int v, int lo, int hi) {
    return v < lo ? lo : (v > hi ? hi : v);
}

__global__ void k_pack_f4(const float* __restrict__ p3, const float* __restrict__ w, uint32_t n,
                          float4* __restrict__ out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    out[i] = make_float4(p3[3ull * i], p3[3ull * i + 1], p3[3ull * i + 2], w ? w[i] : 1.0f);
}

__global__ void k_f4_to_f3(const float4* __restrict__ in, uint32_t n, float* __restrict__ out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4 q = in[i];
    out[3ull * i] = q.x;
    out[3ull * i + 1] = q.y;
    out[3ull * i + 2] = q.z;
}

// particles of surface cells get flag 1 (extractionOfSurfaceParticles marks
// range [start, end) of each surface cell, ReconstructionCUDA.cu:113-119)
__global__ void k_flags_of_cells(const uint8_t* __restrict__ surf_cell,
                                 const fsr_index_range* __restrict__ ranges, uint32_t C,
                                 uint32_t* __restrict__ flags) {
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= C || !surf_cell[c]) return;
    const fsr_index_range r = ranges[c];
    if (r.start == 0xFFFFFFFFu) return;
    for (uint32_t i = r.start; i < r.end; ++i) flags[i] = 1u;
}

__global__ void k_bits_to_dense(const uint32_t* __restrict__ bits, uint32_t rx, uint32_t wx,
                                uint64_t V, uint32_t* __restrict__ dense) {
    const uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= V) return;
    const uint64_t row = v / rx;
    const uint32_t x = (uint32_t)(v % rx);
    dense[v] = (bits[row * wx + (x >> 5)] >> (x & 31u)) & 1u;
}

__global__ void k_bits_from_dense(const uint32_t* __restrict__ dense, uint32_t rx, uint32_t wx,
                                  uint64_t nwords, uint32_t* __restrict__ bits) {
    const uint64_t w = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= nwords) return;
    const uint64_t row = w / wx;
    const uint32_t x0 = (uint32_t)(w % wx) * 32u;
    uint32_t m = 0;
    for (uint32_t b = 0; b < 32u && x0 + b < rx; ++b)
        m |= (dense[row * rx + x0 + b] == 1u ? 1u : 0u) << b;
    bits[w] = m;
}

__global__ void k_bits_from_list(const uint32_t* __restrict__ list, uint32_t count, uint32_t rx,
                                 uint32_t wx, uint32_t* __restrict__ bits) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= count) return;
    const uint32_t v = list[k];
    const uint32_t row = v / rx, x = v % rx;
    atomicOr(&bits[(uint64_t)row * wx + (x >> 5)], 1u << (x & 31u));
}

// out[scan[t]] = t for flags[t] == 1 (compactSurfaceVertex,
// ReconstructionCUDA.cu:293-304; compactValidSurfaceCubes, :367-380)
__global__ void k_compact_flags(uint32_t* __restrict__ out, const uint32_t* __restrict__ scan,
                                const uint32_t* __restrict__ flags, uint64_t n) {
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n && flags[t] == 1u) out[scan[t]] = (uint32_t)t;
}

// involved particles of cells with num_involve > 0 (ReconstructionCUDA.cu:677-685)
__global__ void k_involved_of_cells(const uint32_t* __restrict__ involve,
                                    const uint32_t* __restrict__ involve_scan,
                                    const fsr_index_range* __restrict__ ranges, uint64_t C,
                                    uint32_t* __restrict__ out) {
    const uint64_t c = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= C || involve[c] == 0) return;
    const fsr_index_range r = ranges[c];
    const uint32_t o = involve_scan[c];
    for (uint32_t i = 0; i < r.end - r.start; ++i) out[o + i] = r.start + i;
}

// per-cell involve counts and offsets recovered from an involved-particle list
// (cell of each listed particle by getIndex3D of its position)
__global__ void k_involve_from_list(const uint32_t* __restrict__ list, uint32_t count,
                                    const float4* __restrict__ p4,
                                    const fsr_index_range* __restrict__ ranges, GridDev sp,
                                    uint32_t* __restrict__ involve,
                                    uint32_t* __restrict__ involve_off) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= count) return;
    auto cell_of = [&](uint32_t i) -> uint32_t {
        const float4 p = p4[i];
        const float qx = cell_coord(p.x, sp.minx, sp.cell);
        const float qy = cell_coord(p.y, sp.miny, sp.cell);
        const float qz = cell_coord(p.z, sp.minz, sp.cell);
        if (!(qx > -1.0f && qx < (float)sp.rx && qy > -1.0f && qy < (float)sp.ry &&
              qz > -1.0f && qz < (float)sp.rz))
            return 0xFFFFFFFFu;
        return ((uint32_t)(int)qz * sp.ry + (uint32_t)(int)qy) * sp.rx + (uint32_t)(int)qx;
    };
    const uint32_t c = cell_of(list[k]);
    if (c == 0xFFFFFFFFu) return;
    if (k == 0 || cell_of(list[k - 1]) != c) {
        involve_off[c] = k;
        const fsr_index_range r = ranges[c];
        involve[c] = r.end - r.start;
    }
}

// field records (M = G^T G, coefficient) of every particle from the G list;
// the caller's G carries no support bound, so no record is culled (w = 3e38)
__global__ void k_rec_from_g(const fsr_index_range* __restrict__ ranges,
                             const uint32_t* __restrict__ involve,
                             const uint32_t* __restrict__ involve_off,
                             const fsr_matrix* __restrict__ g, const float* __restrict__ dens,
                             float mass, uint32_t C, float4* __restrict__ rec) {
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= C) return;
    const fsr_index_range r = ranges[c];
    if (r.start == 0xFFFFFFFFu) return;
    const float sigma = 1.5666814710608448f;
    for (uint32_t j = r.start; j < r.end; ++j) {
        const float mr = mass / dens[j];
        if (involve[c] > 0) {
            const fsr_matrix G = g[involve_off[c] + (j - r.start)];
            const float a[3][3] = {{G.a11, G.a12, G.a13}, {G.a21, G.a22, G.a23}, {G.a31, G.a32, G.a33}};
            float M[3][3];
            for (int p = 0; p < 3; ++p)
                for (int q = 0; q < 3; ++q)
                    M[p][q] = a[0][p] * a[0][q] + a[1][p] * a[1][q] + a[2][p] * a[2][q];
            const float det = a[0][0] * a[1][1] * a[2][2] + a[0][1] * a[1][2] * a[2][0] +
                              a[0][2] * a[1][0] * a[2][1] - a[0][0] * a[1][2] * a[2][1] -
                              a[0][1] * a[1][0] * a[2][2] - a[0][2] * a[1][1] * a[2][0];
            rec[2ull * j] = make_float4(M[0][0], M[1][1], M[2][2], M[0][1]);
            rec[2ull * j + 1] = make_float4(M[0][2], M[1][2], sigma * det * mr, 3.0e38f);
        } else {
            rec[2ull * j] = make_float4(1.0f, 1.0f, 1.0f, 0.0f);
            rec[2ull * j + 1] = make_float4(0.0f, 0.0f, sigma * mr, 3.0e38f);
        }
    }
}

inline uint32_t blocks_for(uint64_t n) { return (uint32_t)((n + 255) / 256 ? (n + 255) / 256 : 1); }

__device__ __forceinline__ int clampi_s(int v, int lo, int hi) {
    return v < lo ? lo : (v > hi ? hi : v);
}

// Largest range end over the cells (the particles in cells form a prefix of the
// cell-sorted array) into *max_end (zeroed by the caller).
__global__ void k_out_of_grid(const fsr_index_range* __restrict__ ranges, uint32_t C, uint32_t n,
                              uint32_t* max_end) {
    uint32_t m = 0;
    for (uint32_t c = blockIdx.x * blockDim.x + threadIdx.x; c < C; c += gridDim.x * blockDim.x) {
        const fsr_index_range r = ranges[c];
        if (r.start != 0xFFFFFFFFu && r.end <= n) m = max(m, r.end);
    }
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) m = max(m, (uint32_t)__shfl_xor((int)m, d, 64));
    if (__lane_id() == 0 && m) atomicMax(max_end, m);
}

__global__ void k_out_of_grid_finish(const uint32_t* max_end, uint32_t n, fsr_frame_stats* st) {
    if (threadIdx.x == 0) st->num_out_of_grid = n - *max_end;
}

// estimateSurfaceVertices (ReconstructionCUDA.cu:123-175): a density-grid cell
// (not on the last layer) whose 8 corners straddle 0.5 (getVertexFlag,
// ReconstructionCUDA.cuh:491-511: bit i set when corner i <= 0.5) marks every
// scalar vertex of [c * ratio - ext, (c + 1) * ratio + ext], clamped per axis.
__global__ void k_estimate_surface_vertices(const float* __restrict__ dens, GridDev dg,
                                            uint32_t* __restrict__ is_surface, GridDev sg,
                                            int ratio, int ext) {
    const uint32_t C = dg.rx * dg.ry * dg.rz;
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= C) return;
    const uint32_t x = c % dg.rx, y = (c / dg.rx) % dg.ry, z = c / (dg.rx * dg.ry);
    if (x >= dg.rx - 1 || y >= dg.ry - 1 || z >= dg.rz - 1) return;
    const uint64_t rxy = (uint64_t)dg.rx * dg.ry;
    const float v[8] = {dens[c], dens[c + rxy], dens[c + rxy + dg.rx], dens[c + dg.rx],
                        dens[c + 1], dens[c + 1 + rxy], dens[c + 1 + rxy + dg.rx],
                        dens[c + 1 + dg.rx]};
    uint32_t flag = 0;
    for (int i = 0; i < 8; ++i) flag |= (v[i] <= 0.5f) ? (1u << i) : 0u;
    if (flag == 0 || flag == 255) return;
    const int x0 = clampi_s((int)x * ratio - ext, 0, (int)sg.rx - 1);
    const int x1 = clampi_s(((int)x + 1) * ratio + ext, 0, (int)sg.rx - 1);
    const int y0 = clampi_s((int)y * ratio - ext, 0, (int)sg.ry - 1);
    const int y1 = clampi_s(((int)y + 1) * ratio + ext, 0, (int)sg.ry - 1);
    const int z0 = clampi_s((int)z * ratio - ext, 0, (int)sg.rz - 1);
    const int z1 = clampi_s(((int)z + 1) * ratio + ext, 0, (int)sg.rz - 1);
    for (int zz = z0; zz <= z1; ++zz)
        for (int yy = y0; yy <= y1; ++yy)
            for (int xx = x0; xx <= x1; ++xx)
                is_surface[((uint64_t)zz * sg.ry + yy) * sg.rx + xx] = 1u;
}

// updateScalarFieldValueTC01 (ReconstructionCUDA.cu:178-227) with kernelTC01
// (ReconstructionCUDA.cuh:252-260): phi(x) = sum over the clamped 3^3 block of
// (d - 1/2)^2 with d = |x - p|^2 / (2 R^2), for |x - p|^2 <= R^2.  One thread per
// listed vertex; the value is written unconditionally.
__global__ void k_field_tc01(const uint32_t* __restrict__ sv_index, uint32_t num_sv,
                             const fsr_index_range* __restrict__ ranges,
                             const float* __restrict__ particles, float* __restrict__ phi,
                             GridDev sp, GridDev sc, float radius_sq) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= num_sv) return;
    const uint32_t vid = sv_index[k];
    const uint32_t ix = vid % sc.rx, iy = (vid / sc.rx) % sc.ry, iz = vid / (sc.rx * sc.ry);
    const float vx = vertex_coord(ix, sc.minx, sc.cell);
    const float vy = vertex_coord(iy, sc.miny, sc.cell);
    const float vz = vertex_coord(iz, sc.minz, sc.cell);
    auto cell = [](float q, uint32_t res) {
        if (!(q == q)) return 0;
        const float hi = (float)res + 4.0f;
        q = q < -4.0f ? -4.0f : (q > hi ? hi : q);
        return (int)q;
    };
    const int cx = cell(cell_coord(vx, sp.minx, sp.cell), sp.rx);
    const int cy = cell(cell_coord(vy, sp.miny, sp.cell), sp.ry);
    const int cz = cell(cell_coord(vz, sp.minz, sp.cell), sp.rz);
    const int x0 = clampi_s(cx - 1, 0, sp.rx - 1), x1 = clampi_s(cx + 1, 0, sp.rx - 1);
    const int y0 = clampi_s(cy - 1, 0, sp.ry - 1), y1 = clampi_s(cy + 1, 0, sp.ry - 1);
    const int z0 = clampi_s(cz - 1, 0, sp.rz - 1), z1 = clampi_s(cz + 1, 0, sp.rz - 1);
    float val = 0.f;
    for (int zz = z0; zz <= z1; ++zz)
        for (int yy = y0; yy <= y1; ++yy)
            for (int xx = x0; xx <= x1; ++xx) {
                const fsr_index_range r = ranges[((uint32_t)zz * sp.ry + yy) * sp.rx + xx];
                if (r.start == 0xFFFFFFFFu) continue;
                for (uint32_t i = r.start; i < r.end; ++i) {
                    const float dx = vx - particles[3ull * i];
                    const float dy = vy - particles[3ull * i + 1];
                    const float dz = vz - particles[3ull * i + 2];
                    const float d2 = __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, dx * dx));
                    if (d2 > radius_sq) continue;
                    const float d = d2 / radius_sq * 0.5f;
                    val += d * d - d + 0.25f;
                }
            }
    phi[vid] = val;
}

}  // namespace

// launchers used by the entry points below (defined in the other units)
fsr_status surface_seed_from_flags(fsr_context* ctx, const SurfaceArgs& a, hipStream_t s);

}  // namespace fsr

using namespace fsr;

namespace {

fsr_status need(bool ok, const char* what) {
    if (ok) return FSR_OK;
    set_error(std::string(what) + ": invalid argument");
    return FSR_ERR_INVALID_ARG;
}

fsr_status grid_ok(const fsr_grid_info* g, const char* what) {
    if (!g || !g->resolution.x || !g->resolution.y || !g->resolution.z || !(g->cell_size > 0.0f) ||
        grid_count(*g) >= 0xFFFFFFFFull) {
        set_error(std::string(what) + ": invalid grid");
        return FSR_ERR_INVALID_ARG;
    }
    return FSR_OK;
}

// device stats + counters for one stage call, zeroed
fsr_status stage_state(fsr_context* ctx, fsr_frame_stats** st, Counters** ct, hipStream_t s) {
    FSR_TRY(arena_typed(ctx, A_STATS, 1, st));
    FSR_TRY(arena_typed(ctx, A_COUNTERS, 1, ct));
    FSR_TRY(fill_u32(reinterpret_cast<uint32_t*>(*st), 0u, sizeof(fsr_frame_stats) / 4, s));
    FSR_TRY(fill_u32(reinterpret_cast<uint32_t*>(*ct), 0u, sizeof(Counters) / 4, s));
    return FSR_OK;
}

// Band bit mask from a list, then the cell-major work list (radius 1 or 2).
fsr_status band_worklist_from_list(fsr_context* ctx, const uint32_t* list, uint32_t count,
                                   const GridDev& sc, const GridDev& sp, int radius,
                                   Counters* ct, fsr_frame_stats* st, BandCompactOut* bc,
                                   hipStream_t s) {
    const uint32_t wx = row_words(sc.rx);
    const uint64_t W = (uint64_t)wx * sc.ry * sc.rz;
    uint32_t* bits;
    FSR_TRY(arena_typed(ctx, A_BAND_BITS, W, &bits));
    FSR_TRY(fill_u32(bits, 0u, W, s));
    if (count) {
        hipLaunchKernelGGL(k_bits_from_list, dim3(blocks_for(count)), dim3(256), 0, s, list, count,
                           sc.rx, wx, bits);
        FSR_LAUNCH_CHECK("k_bits_from_list");
    }
    return band_compact(ctx, bits, sc, sp, radius, ct, st, bc, s);
}

}  // namespace

extern "C" {

fsr_status fsr_bind_textures(fsr_context* ctx, const uint32_t* edge_table, const int32_t* tri_table,
                             const uint32_t* num_vertices_table,
                             const uint32_t* vertex_indices_of_edge_table) {
    // The case tables are compiled into the kernels.  When the caller passes
    // its device tables they are read back once and compared, so a caller with
    // different tables is told rather than silently getting ours.
    FSR_TRY(need(ctx != nullptr, "fsr_bind_textures"));
    if (!tri_table) return FSR_OK;
    FSR_HIP(hipSetDevice(ctx->device));
    int32_t tri[256 * 16];
    FSR_HIP(hipMemcpy(tri, tri_table, sizeof(tri), hipMemcpyDeviceToHost));
    static const uint64_t packed[256] = {FSR_MC_TRI_PACKED_VALUES};
    for (int c = 0; c < 256; ++c) {
        for (int k = 0; k < 16; ++k) {
            const uint32_t e = (uint32_t)(packed[c] >> (4 * k)) & 0xFu;
            const int32_t want = e == 0xF ? -1 : (int32_t)e;
            if (tri[16 * c + k] != want) {
                set_error("fsr_bind_textures: triangle table differs from the built-in table");
                return FSR_ERR_INVALID_ARG;
            }
        }
    }
    (void)edge_table;
    (void)num_vertices_table;
    (void)vertex_indices_of_edge_table;
    return FSR_OK;
}

fsr_status fsr_spatial_grid_building(fsr_context* ctx, float* particles, float* densities,
                                     uint32_t num_particles, fsr_index_range* ranges,
                                     float* flag_grid, const fsr_grid_info* sp, void* stream) {
    FSR_TRY(need(ctx && (num_particles == 0 || (particles && densities)) && ranges && flag_grid,
                 "fsr_spatial_grid_building"));
    FSR_TRY(grid_ok(sp, "fsr_spatial_grid_building"));
    FSR_HIP(hipSetDevice(ctx->device));
    hipStream_t s = (hipStream_t)stream;
    fsr_frame_stats* st;
    Counters* ct;
    FSR_TRY(stage_state(ctx, &st, &ct, s));
    HashOut ho;
    FSR_TRY(spatial_hash(ctx, particles, densities, num_particles, *sp, st, &ho, s));
    const uint64_t C = grid_count(*sp);
    FSR_HIP(hipMemcpyAsync(ranges, ho.ranges, C * sizeof(fsr_index_range), hipMemcpyDeviceToDevice, s));
    FSR_HIP(hipMemcpyAsync(flag_grid, ho.occupancy, C * sizeof(float), hipMemcpyDeviceToDevice, s));
    return FSR_OK;
}

static fsr_status surface_cells_stage(fsr_context* ctx, const float* flag_grid,
                                      const fsr_index_range* ranges, const fsr_grid_info* sp,
                                      uint32_t* flags, uint32_t n, uint32_t* involve,
                                      hipStream_t s) {
    const GridDev g = to_dev(*sp);
    const uint32_t C = (uint32_t)grid_count(*sp);
    uint8_t* surf;
    FSR_TRY(arena_typed(ctx, A_SURF_CELL, C, &surf));
    FSR_TRY(fill_u32(flags, 0u, n, s));
    SurfaceArgs a;
    std::memset(&a, 0, sizeof(a));
    a.ranges = ranges;
    a.occupancy = flag_grid;
    a.sp = g;
    FSR_TRY(surface_occupancy(ctx, a, surf, involve, s));  // a.n == 0: cells only
    hipLaunchKernelGGL(k_flags_of_cells, dim3(blocks_for(C)), dim3(256), 0, s, surf, ranges, C,
                       flags);
    FSR_LAUNCH_CHECK("k_flags_of_cells");
    return FSR_OK;
}

fsr_status fsr_extraction_of_surface_particles(fsr_context* ctx, const fsr_sim_param* sim,
                                               const float* flag_grid, uint32_t* surface_flags,
                                               uint32_t num_particles,
                                               const fsr_index_range* ranges,
                                               const fsr_grid_info* sp, void* stream) {
    FSR_TRY(need(ctx && sim && flag_grid && ranges && (num_particles == 0 || surface_flags),
                 "fsr_extraction_of_surface_particles"));
    FSR_TRY(grid_ok(sp, "fsr_extraction_of_surface_particles"));
    FSR_HIP(hipSetDevice(ctx->device));
    return surface_cells_stage(ctx, flag_grid, ranges, sp, surface_flags, num_particles, nullptr,
                               (hipStream_t)stream);
}

fsr_status fsr_extraction_of_surface_and_involve_particles(
    fsr_context* ctx, const fsr_sim_param* sim, const float* flag_grid, uint32_t* num_involve,
    uint32_t* surface_flags, uint32_t num_particles, const fsr_index_range* ranges,
    const fsr_grid_info* sp, void* stream) {
    FSR_TRY(need(ctx && sim && flag_grid && ranges && num_involve &&
                     (num_particles == 0 || surface_flags),
                 "fsr_extraction_of_surface_and_involve_particles"));
    FSR_TRY(grid_ok(sp, "fsr_extraction_of_surface_and_involve_particles"));
    FSR_HIP(hipSetDevice(ctx->device));
    return surface_cells_stage(ctx, flag_grid, ranges, sp, surface_flags, num_particles,
                               num_involve, (hipStream_t)stream);
}

fsr_status fsr_extraction_of_surface_particles_for_akinci(
    fsr_context* ctx, const fsr_sim_param* sim, const float* particles, const float* densities,
    uint32_t num_particles, const fsr_index_range* ranges, const fsr_grid_info* sp,
    uint32_t* surface_flags, void* stream) {
    FSR_TRY(need(ctx && sim && ranges &&
                     (num_particles == 0 || (particles && densities && surface_flags)),
                 "fsr_extraction_of_surface_particles_for_akinci"));
    FSR_TRY(grid_ok(sp, "fsr_extraction_of_surface_particles_for_akinci"));
    FSR_HIP(hipSetDevice(ctx->device));
    hipStream_t s = (hipStream_t)stream;
    if (num_particles == 0) return FSR_OK;
    float4* p4;
    FSR_TRY(arena_typed(ctx, A_SORTED, num_particles, &p4));
    hipLaunchKernelGGL(k_pack_f4, dim3(blocks_for(num_particles)), dim3(256), 0, s, particles,
                       densities, num_particles, p4);
    FSR_LAUNCH_CHECK("k_pack_f4");
    fsr_frame_stats* st;
    Counters* ct;
    FSR_TRY(stage_state(ctx, &st, &ct, s));
    // particles past the last cell range are outside the grid
    const uint32_t C = (uint32_t)grid_count(*sp);
    hipLaunchKernelGGL(k_out_of_grid, dim3(256), dim3(256), 0, s, ranges, C, num_particles,
                       &ct->aux[5]);
    FSR_LAUNCH_CHECK("k_out_of_grid");
    hipLaunchKernelGGL(k_out_of_grid_finish, dim3(1), dim3(64), 0, s, &ct->aux[5], num_particles,
                       st);
    FSR_LAUNCH_CHECK("k_out_of_grid_finish");
    SurfaceArgs a;
    std::memset(&a, 0, sizeof(a));
    a.sorted = p4;
    a.ranges = ranges;
    a.n = num_particles;
    a.sp = to_dev(*sp);
    a.h = sim->smoothing_radius;
    a.flags = surface_flags;
    a.counters = ct;
    a.st = st;
    return surface_colorfield(ctx, a, sim->particle_mass, s);
}

fsr_status fsr_estimation_of_surface_vertices(fsr_context* ctx, const fsr_sim_param* sim,
                                              const fsr_grid_info* sc, const float* particles,
                                              const uint32_t* surface_flags,
                                              uint32_t num_particles,
                                              const fsr_index_range* ranges,
                                              uint32_t* is_surface_grid, void* stream) {
    (void)ranges;
    FSR_TRY(need(ctx && sim && is_surface_grid &&
                     (num_particles == 0 || (particles && surface_flags)),
                 "fsr_estimation_of_surface_vertices"));
    FSR_TRY(grid_ok(sc, "fsr_estimation_of_surface_vertices"));
    FSR_HIP(hipSetDevice(ctx->device));
    hipStream_t s = (hipStream_t)stream;
    const GridDev g = to_dev(*sc);
    const uint32_t wx = row_words(g.rx);
    const uint64_t W = (uint64_t)wx * g.ry * g.rz, V = grid_count(*sc);
    fsr_frame_stats* st;
    Counters* ct;
    FSR_TRY(stage_state(ctx, &st, &ct, s));
    uint32_t *sbits, *tmp, *bits, *exc;
    float4* p4;
    const uint64_t nn = num_particles ? num_particles : 1;
    FSR_TRY(arena_typed(ctx, A_SBITS, 8ull * W, &sbits));  // byte seed mask
    FSR_TRY(arena_typed(ctx, A_BAND_TMP, W, &tmp));
    FSR_TRY(arena_typed(ctx, A_BAND_BITS, W, &bits));
    FSR_TRY(arena_typed(ctx, A_EXCEPTIONS, nn, &exc));
    FSR_TRY(arena_typed(ctx, A_SORTED, nn, &p4));
    FSR_TRY(fill_u32(sbits, 0u, 8ull * W, s));
    if (num_particles) {
        hipLaunchKernelGGL(k_pack_f4, dim3(blocks_for(num_particles)), dim3(256), 0, s, particles,
                           (const float*)nullptr, num_particles, p4);
        FSR_LAUNCH_CHECK("k_pack_f4");
        SurfaceArgs a;
        std::memset(&a, 0, sizeof(a));
        a.sorted = p4;
        a.n = num_particles;
        a.sc = g;
        a.h = sim->smoothing_radius;
        a.flags = const_cast<uint32_t*>(surface_flags);
        a.sbits = sbits;
        a.exceptions = exc;
        a.counters = ct;
        a.st = st;
        a.all_exceptions = band_dilation_exact(*sc, sim->smoothing_radius) ? 0 : 1;
        FSR_TRY(surface_seed_from_flags(ctx, a, s));
    }
    BandArgs ba;
    ba.sorted = p4;
    ba.exceptions = exc;
    ba.max_exceptions = num_particles;
    ba.sbits = sbits;
    ba.tmp = tmp;
    ba.bits = bits;
    ba.sc = g;
    ba.h = sim->smoothing_radius;
    ba.counters = ct;
    ba.st = st;
    ba.flags = surface_flags;
    ba.num_particles = num_particles;
    ba.all_exceptions = band_dilation_exact(*sc, sim->smoothing_radius) ? 0 : 1;
    FSR_TRY(band_build(ctx, ba, s));
    hipLaunchKernelGGL(k_bits_to_dense, dim3(blocks_for(V)), dim3(256), 0, s, bits, g.rx, wx, V,
                       is_surface_grid);
    FSR_LAUNCH_CHECK("k_bits_to_dense");
    return FSR_OK;
}

fsr_status fsr_thrust_exclusive_prefix_sum_scan(fsr_context* ctx, uint32_t* output,
                                                const uint32_t* input, uint32_t num_elements,
                                                uint32_t* total, void* stream) {
    FSR_TRY(need(ctx && (num_elements == 0 || (output && input)),
                 "fsr_thrust_exclusive_prefix_sum_scan"));
    FSR_HIP(hipSetDevice(ctx->device));
    hipStream_t s = (hipStream_t)stream;
    fsr_frame_stats* st;
    Counters* ct;
    FSR_TRY(stage_state(ctx, &st, &ct, s));
    FSR_TRY(scan_exclusive(ctx, SCAN_U32, input, output, num_elements, nullptr, &ct->aux[0], s));
    if (total) {
        // blocking, as the reference's two copies of the last elements are
        FSR_HIP(hipMemcpyAsync(total, &ct->aux[0], sizeof(uint32_t), hipMemcpyDeviceToHost, s));
        FSR_HIP(hipStreamSynchronize(s));
    }
    return FSR_OK;
}

fsr_status fsr_compact_surface_vertex(fsr_context* ctx, uint32_t* sv_index,
                                      const uint32_t* is_surface_scan, const uint32_t* is_surface,
                                      uint64_t num_vertices, void* stream) {
    FSR_TRY(need(ctx && (num_vertices == 0 || (sv_index && is_surface_scan && is_surface)),
                 "fsr_compact_surface_vertex"));
    FSR_HIP(hipSetDevice(ctx->device));
    if (!num_vertices) return FSR_OK;
    hipLaunchKernelGGL(k_compact_flags, dim3(blocks_for(num_vertices)), dim3(256), 0,
                       (hipStream_t)stream, sv_index, is_surface_scan, is_surface, num_vertices);
    FSR_LAUNCH_CHECK("k_compact_flags");
    return FSR_OK;
}

fsr_status fsr_update_scalar_grid_values_compacted(
    fsr_context* ctx, const uint32_t* sv_index, uint32_t num_sv, const fsr_index_range* ranges,
    const float* particles, uint32_t num_particles, float* scalar_field, const fsr_grid_info* sp,
    const fsr_grid_info* sc, const fsr_sim_param* sim, void* stream) {
    FSR_TRY(need(ctx && sim && ranges && scalar_field && (num_sv == 0 || sv_index) &&
                     (num_particles == 0 || particles),
                 "fsr_update_scalar_grid_values_compacted"));
    FSR_TRY(grid_ok(sp, "fsr_update_scalar_grid_values_compacted"));
    FSR_TRY(grid_ok(sc, "fsr_update_scalar_grid_values_compacted"));
    FSR_HIP(hipSetDevice(ctx->device));
    hipStream_t s = (hipStream_t)stream;
    fsr_frame_stats* st;
    Counters* ct;
    FSR_TRY(stage_state(ctx, &st, &ct, s));
    float4* p4;
    FSR_TRY(arena_typed(ctx, A_SORTED, num_particles ? num_particles : 1, &p4));
    if (num_particles) {
        hipLaunchKernelGGL(k_pack_f4, dim3(blocks_for(num_particles)), dim3(256), 0, s, particles,
                           (const float*)nullptr, num_particles, p4);
        FSR_LAUNCH_CHECK("k_pack_f4");
    }
    const GridDev gsp = to_dev(*sp), gsc = to_dev(*sc);
    BandCompactOut bc;
    FSR_TRY(band_worklist_from_list(ctx, sv_index, num_sv, gsc, gsp, 1, ct, st, &bc, s));
    FieldArgs fa;
    fa.sorted = p4;
    fa.ranges = ranges;
    fa.bc = &bc;
    fa.sp = gsp;
    fa.sc = gsc;
    fa.radius_sq = sim->smoothing_radius_sq;
    fa.particle_radius = sim->particle_radius;
    fa.phi = scalar_field;
    fa.counters = ct;
    fa.saved = nullptr;
    fa.work = nullptr;
    return field_zb05(ctx, fa, s);
}

static fsr_status classify_stage(fsr_context* ctx, const uint32_t* sv_index, uint32_t num_sv,
                                 const float* scalar_field, const fsr_grid_info* sc,
                                 uint32_t* is_valid, uint32_t* num_vertices,
                                 const uint32_t* is_surface_grid, float iso, hipStream_t s) {
    const GridDev g = to_dev(*sc);
    const uint32_t wx = row_words(g.rx);
    const uint64_t W = (uint64_t)wx * g.ry * g.rz;
    fsr_frame_stats* st;
    Counters* ct;
    FSR_TRY(stage_state(ctx, &st, &ct, s));
    uint32_t *bits, *nv;
    FSR_TRY(arena_typed(ctx, A_BAND_BITS, W, &bits));
    hipLaunchKernelGGL(k_bits_from_dense, dim3(blocks_for(W)), dim3(256), 0, s, is_surface_grid,
                       g.rx, wx, W, bits);
    FSR_LAUNCH_CHECK("k_bits_from_dense");
    FSR_TRY(fill_u32(&st->num_band_vertices, num_sv, 1, s));
    if (!num_vertices) FSR_TRY(arena_typed(ctx, A_NVERTS, num_sv ? num_sv : 1, &nv));
    McArgs ma;
    std::memset(&ma, 0, sizeof(ma));
    ma.band_index = sv_index;
    ma.bits = bits;
    ma.phi = scalar_field;
    ma.sc = g;
    ma.iso = iso;
    ma.nverts = num_vertices ? num_vertices : nv;
    ma.valid = is_valid;
    ma.st = st;
    return mc_classify(ctx, ma, s);
}

fsr_status fsr_detect_valid_surface_cubes(fsr_context* ctx, const uint32_t* sv_index,
                                          uint32_t num_sv, const float* scalar_field,
                                          const fsr_grid_info* sc, uint32_t* is_valid,
                                          uint32_t* num_vertices, const uint32_t* is_surface_grid,
                                          const fsr_sim_param* sim, void* stream) {
    FSR_TRY(need(ctx && sim && scalar_field && is_surface_grid &&
                     (num_sv == 0 || (sv_index && is_valid && num_vertices)),
                 "fsr_detect_valid_surface_cubes"));
    FSR_TRY(grid_ok(sc, "fsr_detect_valid_surface_cubes"));
    FSR_HIP(hipSetDevice(ctx->device));
    return classify_stage(ctx, sv_index, num_sv, scalar_field, sc, is_valid, num_vertices,
                          is_surface_grid, sim->iso_value, (hipStream_t)stream);
}

fsr_status fsr_detection_of_valid_surface_cubes_for_akinci(
    fsr_context* ctx, const uint32_t* sv_index, uint32_t num_sv, const float* scalar_field,
    const fsr_grid_info* sc, uint32_t* is_valid, const uint32_t* is_surface_grid,
    const fsr_sim_param* sim, void* stream) {
    FSR_TRY(need(ctx && sim && scalar_field && is_surface_grid &&
                     (num_sv == 0 || (sv_index && is_valid)),
                 "fsr_detection_of_valid_surface_cubes_for_akinci"));
    FSR_TRY(grid_ok(sc, "fsr_detection_of_valid_surface_cubes_for_akinci"));
    FSR_HIP(hipSetDevice(ctx->device));
    return classify_stage(ctx, sv_index, num_sv, scalar_field, sc, is_valid, nullptr,
                          is_surface_grid, sim->iso_value, (hipStream_t)stream);
}

fsr_status fsr_compact_valid_surface_cubes(fsr_context* ctx, uint32_t* valid_index,
                                           const uint32_t* valid_scan, const uint32_t* valid,
                                           uint32_t num_sv, void* stream) {
    FSR_TRY(need(ctx && (num_sv == 0 || (valid_index && valid_scan && valid)),
                 "fsr_compact_valid_surface_cubes"));
    FSR_HIP(hipSetDevice(ctx->device));
    if (!num_sv) return FSR_OK;
    hipLaunchKernelGGL(k_compact_flags, dim3(blocks_for(num_sv)), dim3(256), 0,
                       (hipStream_t)stream, valid_index, valid_scan, valid, (uint64_t)num_sv);
    FSR_LAUNCH_CHECK("k_compact_flags");
    return FSR_OK;
}

fsr_status fsr_generate_triangles(fsr_context* ctx, const uint32_t* sv_index,
                                  const uint32_t* valid_index, uint32_t num_valid,
                                  const fsr_grid_info* sc, const uint32_t* num_vertices_scan,
                                  const float* scalar_field, float* positions, float* normals,
                                  uint32_t capacity, const fsr_sim_param* sim, void* stream) {
    FSR_TRY(need(ctx && sim && scalar_field &&
                     (num_valid == 0 || (sv_index && valid_index && num_vertices_scan && positions)),
                 "fsr_generate_triangles"));
    FSR_TRY(grid_ok(sc, "fsr_generate_triangles"));
    FSR_HIP(hipSetDevice(ctx->device));
    hipStream_t s = (hipStream_t)stream;
    fsr_frame_stats* st;
    Counters* ct;
    FSR_TRY(stage_state(ctx, &st, &ct, s));
    FSR_TRY(fill_u32(&st->num_valid_cubes, num_valid, 1, s));
    McArgs ma;
    std::memset(&ma, 0, sizeof(ma));
    ma.band_index = sv_index;
    ma.phi = scalar_field;
    ma.sc = to_dev(*sc);
    ma.iso = sim->iso_value;
    ma.vert_off = const_cast<uint32_t*>(num_vertices_scan);
    ma.valid_idx = const_cast<uint32_t*>(valid_index);
    ma.tri_pos = positions;
    ma.tri_nrm = normals;
    ma.capacity = capacity;
    ma.st = st;
    return mc_emit(ctx, ma, s);
}

fsr_status fsr_calculation_of_mean_and_smooth_particles(fsr_context* ctx,
                                                        const fsr_sim_param* sim,
                                                        const float* particles,
                                                        uint32_t num_particles, float* mean,
                                                        float* smoothed,
                                                        const fsr_index_range* ranges,
                                                        const fsr_grid_info* sp, void* stream) {
    FSR_TRY(need(ctx && sim && ranges &&
                     (num_particles == 0 || (particles && mean && smoothed)),
                 "fsr_calculation_of_mean_and_smooth_particles"));
    FSR_TRY(grid_ok(sp, "fsr_calculation_of_mean_and_smooth_particles"));
    FSR_HIP(hipSetDevice(ctx->device));
    hipStream_t s = (hipStream_t)stream;
    if (!num_particles) return FSR_OK;
    fsr_frame_stats* st;
    Counters* ct;
    FSR_TRY(stage_state(ctx, &st, &ct, s));
    AnisoArgs a;
    std::memset(&a, 0, sizeof(a));
    FSR_TRY(arena_typed(ctx, A_SORTED, num_particles, const_cast<float4**>(&a.sorted)));
    FSR_TRY(arena_typed(ctx, A_MEAN, num_particles, &a.mean));
    FSR_TRY(arena_typed(ctx, A_SMOOTH, num_particles, &a.smooth));
    hipLaunchKernelGGL(k_pack_f4, dim3(blocks_for(num_particles)), dim3(256), 0, s, particles,
                       (const float*)nullptr, num_particles, const_cast<float4*>(a.sorted));
    FSR_LAUNCH_CHECK("k_pack_f4");
    a.ranges = ranges;
    a.n = num_particles;
    a.sp = to_dev(*sp);
    a.aniso_radius = sim->anisotropic_radius;
    a.lambda = sim->lambda_for_smoothed;
    a.counters = ct;
    a.st = st;
    FSR_TRY(aniso_smooth(ctx, a, s));
    hipLaunchKernelGGL(k_f4_to_f3, dim3(blocks_for(num_particles)), dim3(256), 0, s, a.mean,
                       num_particles, mean);
    FSR_LAUNCH_CHECK("k_f4_to_f3");
    hipLaunchKernelGGL(k_f4_to_f3, dim3(blocks_for(num_particles)), dim3(256), 0, s, a.smooth,
                       num_particles, smoothed);
    FSR_LAUNCH_CHECK("k_f4_to_f3");
    return FSR_OK;
}

fsr_status fsr_compactation_of_surface_vertices_and_particles(
    fsr_context* ctx, const uint32_t* is_surface, const uint32_t* is_surface_scan,
    uint64_t num_vertices, const uint32_t* num_involve, const uint32_t* num_involve_scan,
    const fsr_index_range* ranges, uint64_t num_cells, uint32_t* sv_index,
    uint32_t* involved_index, void* stream) {
    FSR_TRY(need(ctx && (num_vertices == 0 || (is_surface && is_surface_scan && sv_index)) &&
                     (num_cells == 0 || (num_involve && num_involve_scan && ranges && involved_index)),
                 "fsr_compactation_of_surface_vertices_and_particles"));
    FSR_HIP(hipSetDevice(ctx->device));
    hipStream_t s = (hipStream_t)stream;
    if (num_vertices) {
        hipLaunchKernelGGL(k_compact_flags, dim3(blocks_for(num_vertices)), dim3(256), 0, s,
                           sv_index, is_surface_scan, is_surface, num_vertices);
        FSR_LAUNCH_CHECK("k_compact_flags");
    }
    if (num_cells) {
        hipLaunchKernelGGL(k_involved_of_cells, dim3(blocks_for(num_cells)), dim3(256), 0, s,
                           num_involve, num_involve_scan, ranges, num_cells, involved_index);
        FSR_LAUNCH_CHECK("k_involved_of_cells");
    }
    return FSR_OK;
}

fsr_status fsr_calculation_of_transform_matrices_for_particles(
    fsr_context* ctx, const fsr_sim_param* sim, const fsr_aniso_param* aniso,
    const fsr_grid_info* sp, const float* mean, const float* particles, uint32_t num_particles,
    const fsr_index_range* ranges, const uint32_t* involved_index, uint32_t num_involved,
    fsr_matrix* g, void* stream) {
    FSR_TRY(need(ctx && sim && ranges &&
                     (num_involved == 0 || (mean && particles && involved_index && g)),
                 "fsr_calculation_of_transform_matrices_for_particles"));
    FSR_TRY(grid_ok(sp, "fsr_calculation_of_transform_matrices_for_particles"));
    FSR_HIP(hipSetDevice(ctx->device));
    hipStream_t s = (hipStream_t)stream;
    if (!num_involved || !num_particles) return FSR_OK;
    fsr_frame_stats* st;
    Counters* ct;
    FSR_TRY(stage_state(ctx, &st, &ct, s));
    const uint32_t C = (uint32_t)grid_count(*sp);
    AnisoArgs a;
    std::memset(&a, 0, sizeof(a));
    const uint32_t n = num_particles;
    FSR_TRY(arena_typed(ctx, A_SORTED, n, const_cast<float4**>(&a.sorted)));
    FSR_TRY(arena_typed(ctx, A_MEAN, n, &a.mean));
    FSR_TRY(arena_typed(ctx, A_INVOLVE, C, const_cast<uint32_t**>(&a.involve)));
    FSR_TRY(arena_typed(ctx, A_INVOLVE_OFF, C, &a.involve_off));
    FSR_TRY(arena_typed(ctx, A_SIGMA, num_involved, &a.sigma));
    FSR_TRY(arena_typed(ctx, A_REC, 2ull * n, &a.rec));
    hipLaunchKernelGGL(k_pack_f4, dim3(blocks_for(n)), dim3(256), 0, s, particles,
                       (const float*)nullptr, n, const_cast<float4*>(a.sorted));
    FSR_LAUNCH_CHECK("k_pack_f4");
    hipLaunchKernelGGL(k_pack_f4, dim3(blocks_for(n)), dim3(256), 0, s, mean,
                       (const float*)nullptr, n, a.mean);
    FSR_LAUNCH_CHECK("k_pack_f4");
    FSR_TRY(fill_u32(const_cast<uint32_t*>(a.involve), 0u, C, s));
    hipLaunchKernelGGL(k_involve_from_list, dim3(blocks_for(num_involved)), dim3(256), 0, s,
                       involved_index, num_involved, a.sorted, ranges, to_dev(*sp),
                       const_cast<uint32_t*>(a.involve), a.involve_off);
    FSR_LAUNCH_CHECK("k_involve_from_list");
    a.ranges = ranges;
    a.n = n;
    a.sp = to_dev(*sp);
    a.aniso_radius = sim->anisotropic_radius;
    a.h2 = sim->smoothing_radius_sq;
    a.inv_h = sim->smoothing_radius_inv;
    a.kr = aniso ? aniso->kr : 4.0f;
    a.kn = aniso ? aniso->kn : 1.0f;
    a.mass = sim->particle_mass;
    a.min_neighbors = sim->min_num_neighbors;
    a.gmat = g;
    a.counters = ct;
    a.st = st;
    return aniso_cov(ctx, a, s);
}

fsr_status fsr_computation_of_scalar_field_grid(
    fsr_context* ctx, const fsr_sim_param* sim, uint32_t num_sv, const fsr_grid_info* sc,
    const fsr_grid_info* sp, const fsr_matrix* g, const float* smoothed, const float* densities,
    uint32_t num_particles, const fsr_index_range* ranges, const uint32_t* sv_index,
    const uint32_t* num_involve, const uint32_t* num_involve_scan, float* scalar_field,
    void* stream) {
    FSR_TRY(need(ctx && sim && ranges && num_involve && num_involve_scan && scalar_field &&
                     (num_sv == 0 || sv_index) &&
                     (num_particles == 0 || (smoothed && densities)),
                 "fsr_computation_of_scalar_field_grid"));
    FSR_TRY(grid_ok(sp, "fsr_computation_of_scalar_field_grid"));
    FSR_TRY(grid_ok(sc, "fsr_computation_of_scalar_field_grid"));
    FSR_HIP(hipSetDevice(ctx->device));
    hipStream_t s = (hipStream_t)stream;
    fsr_frame_stats* st;
    Counters* ct;
    FSR_TRY(stage_state(ctx, &st, &ct, s));
    const uint32_t C = (uint32_t)grid_count(*sp);
    const uint32_t n = num_particles;
    AnisoArgs a;
    std::memset(&a, 0, sizeof(a));
    FSR_TRY(arena_typed(ctx, A_SMOOTH, n ? n : 1, &a.smooth));
    FSR_TRY(arena_typed(ctx, A_REC, 2ull * (n ? n : 1), &a.rec));
    if (n) {
        hipLaunchKernelGGL(k_pack_f4, dim3(blocks_for(n)), dim3(256), 0, s, smoothed, densities, n,
                           a.smooth);
        FSR_LAUNCH_CHECK("k_pack_f4");
        hipLaunchKernelGGL(k_rec_from_g, dim3(blocks_for(C)), dim3(256), 0, s, ranges, num_involve,
                           num_involve_scan, g, densities, sim->particle_mass, C, a.rec);
        FSR_LAUNCH_CHECK("k_rec_from_g");
    }
    a.n = n;
    a.ranges = ranges;
    a.sp = to_dev(*sp);
    a.sc = to_dev(*sc);
    a.counters = ct;
    BandCompactOut bc;
    FSR_TRY(band_worklist_from_list(ctx, sv_index, num_sv, a.sc, a.sp, 2, ct, st, &bc, s));
    return aniso_field(ctx, a, bc, scalar_field, s);
}

fsr_status fsr_estimate_surface_vertices(fsr_context* ctx, const fsr_sim_param* sim,
                                         const float* density_grid, const fsr_grid_info* sp,
                                         uint32_t sc_search_ext, uint32_t* is_surface_grid,
                                         const fsr_grid_info* sc, void* stream) {
    FSR_TRY(need(ctx && sim && density_grid && is_surface_grid, "fsr_estimate_surface_vertices"));
    FSR_TRY(grid_ok(sp, "fsr_estimate_surface_vertices"));
    FSR_TRY(grid_ok(sc, "fsr_estimate_surface_vertices"));
    FSR_HIP(hipSetDevice(ctx->device));
    const uint64_t C = grid_count(*sp);
    hipLaunchKernelGGL(k_estimate_surface_vertices, dim3(blocks_for(C)), dim3(256), 0,
                       (hipStream_t)stream, density_grid, to_dev(*sp), is_surface_grid,
                       to_dev(*sc), (int)sim->sc_sp_grid_res_ratio, (int)sc_search_ext);
    FSR_LAUNCH_CHECK("k_estimate_surface_vertices");
    return FSR_OK;
}

fsr_status fsr_update_scalar_grid_values_tc01(fsr_context* ctx, const uint32_t* sv_index,
                                              uint32_t num_sv, const fsr_index_range* ranges,
                                              const float* particles, uint32_t num_particles,
                                              float* scalar_field, const fsr_grid_info* sp,
                                              const fsr_grid_info* sc, const fsr_sim_param* sim,
                                              void* stream) {
    FSR_TRY(need(ctx && sim && ranges && scalar_field && (num_sv == 0 || sv_index) &&
                     (num_particles == 0 || particles),
                 "fsr_update_scalar_grid_values_tc01"));
    FSR_TRY(grid_ok(sp, "fsr_update_scalar_grid_values_tc01"));
    FSR_TRY(grid_ok(sc, "fsr_update_scalar_grid_values_tc01"));
    FSR_HIP(hipSetDevice(ctx->device));
    if (!num_sv) return FSR_OK;
    hipLaunchKernelGGL(k_field_tc01, dim3(blocks_for(num_sv)), dim3(256), 0,
                       (hipStream_t)stream, sv_index, num_sv, ranges, particles, scalar_field,
                       to_dev(*sp), to_dev(*sc), sim->smoothing_radius_sq);
    FSR_LAUNCH_CHECK("k_field_tc01");
    return FSR_OK;
}

}  // extern "C"
