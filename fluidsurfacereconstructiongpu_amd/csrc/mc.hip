// Marching cubes over the band: classify, scan + compact, emit a triangle soup.
//
//   k_mc_classify  detectValidSurfaceCubes (ReconstructionCUDA.cu:328-364) and
//                  detectionOfValidSurfaceCubesForAkinci (:974-1008): the cube
//                  whose corner 0 is band vertex b is skipped on the last grid
//                  layer; cube index bit i is set iff phi(corner i) <= iso
//                  (ReconstructionCUDA.cuh:534-544); the cube is valid iff all 8
//                  corners are band vertices (isAllSfVertex, .cuh:602-609) and its
//                  case emits vertices.
//   scan_mc        the reference's two thrust scans (valid flags and vertex
//                  counts, GPU/ReconstructorGPUOurs.cpp:195-235) and
//                  compactValidSurfaceCubes (ReconstructionCUDA.cu:367-380) in
//                  one two-pass scan: a cube is valid iff it emits vertices, so
//                  the vertex counts alone give the offsets and the valid-cube
//                  list; totals stay on the device.
//   k_mc_emit      generateTriangles (ReconstructionCUDA.cu:383-463): per crossed
//                  edge t = (iso - phi0)/(phi1 - phi0) (0.5 if |delta| < 1e-7),
//                  position lerp of the corner positions, normal lerp of the
//                  normalised central-difference corner normals (out-of-grid
//                  samples read 0), negated when iso < 0; triangles written at the
//                  cube's scanned offset, in table order.
// The `cpu` flavour reproduces the reference's host marching cubes used by the
// Akinci12 method (GPU/MarchingCubesCPU.cpp:3-89, MarchingCubesCPU.h:142-181):
// positions min + i*cell without fusion, normals normalised only when their
// length exceeds 1e-7, no negation.  (Its out-of-range neighbour read is
// undefined in the reference; here it reads 0 like the GPU path.)
#include <algorithm>

#include "common.hpp"
#include "mc_tables.h"

namespace fsr {

namespace {

__constant__ uint64_t cMcTri[256] = {FSR_MC_TRI_PACKED_VALUES};

__device__ __forceinline__ uint32_t case_vertices(uint32_t flag) {
    const uint64_t x = cMcTri[flag];
    // nibbles equal to 0xF mark unused entries
    const uint64_t f = x & (x >> 1) & (x >> 2) & (x >> 3) & 0x1111111111111111ull;
    return 16u - (uint32_t)__popcll(f);
}

__device__ __forceinline__ uint32_t cube_flag(const float* __restrict__ phi, uint64_t v,
                                              uint32_t rx, uint64_t rxy, float iso) {
    uint32_t f = 0;
    f |= (phi[v] <= iso) ? 1u : 0u;
    f |= (phi[v + rxy] <= iso) ? 2u : 0u;
    f |= (phi[v + rxy + rx] <= iso) ? 4u : 0u;
    f |= (phi[v + rx] <= iso) ? 8u : 0u;
    f |= (phi[v + 1] <= iso) ? 16u : 0u;
    f |= (phi[v + 1 + rxy] <= iso) ? 32u : 0u;
    f |= (phi[v + 1 + rxy + rx] <= iso) ? 64u : 0u;
    f |= (phi[v + 1 + rx] <= iso) ? 128u : 0u;
    return f;
}

// Soup vertices of the cube whose corner 0 is band vertex vid (0: not valid).
__device__ __forceinline__ uint32_t classify_cube(const McArgs& a, uint32_t vid) {
    const uint32_t rx = a.sc.rx, ry = a.sc.ry, rz = a.sc.rz;
    const uint64_t rxy = (uint64_t)rx * ry;
    const uint32_t wx = row_words(rx);
    const uint32_t x = vid % rx, y = (vid / rx) % ry, z = (uint32_t)(vid / rxy);
    uint32_t nv = 0;
    if (x < rx - 1 && y < ry - 1 && z < rz - 1) {
        // the corners' band bits first: the field is read only for cubes
        // whose 8 corners are band vertices.  The 8 words (x and x + 1 of the
        // 4 rows, both inside the grid) are loaded together, not one after
        // another behind a short-circuit test.
        uint32_t all = 1u;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const uint32_t dy = (r == 1 || r == 2) ? 1u : 0u, dz = (r >= 2) ? 1u : 0u;
            const uint64_t row = ((uint64_t)(z + dz) * ry + (y + dy)) * wx;
            const uint32_t w0 = a.bits[row + (x >> 5)];
            const uint32_t w1 = a.bits[row + ((x + 1u) >> 5)];
            all &= (w0 >> (x & 31u)) & (w1 >> ((x + 1u) & 31u)) & 1u;
        }
        if (all) nv = case_vertices(cube_flag(a.phi, vid, rx, rxy, a.iso));
    }
    return nv;
}

__global__ __launch_bounds__(256) void k_mc_classify(McArgs a) {
    const uint32_t B = a.st->num_band_vertices;
    for (uint32_t b = blockIdx.x * blockDim.x + threadIdx.x; b < B; b += gridDim.x * blockDim.x) {
        const uint32_t nv = classify_cube(a, a.band_index[b]);
        a.nverts[b] = nv;
        if (a.valid) a.valid[b] = nv ? 1u : 0u;
    }
}

__device__ __forceinline__ float sample(const float* __restrict__ phi, int x, int y, int z,
                                        uint32_t rx, uint32_t ry, uint32_t rz) {
    if (x < 0 || y < 0 || z < 0 || x >= (int)rx || y >= (int)ry || z >= (int)rz) return 0.0f;
    return phi[((uint64_t)z * ry + y) * rx + x];
}

// OR of a value over the 8 lanes of an aligned lane group (ds_swizzle, bit
// mode, xor masks 1, 2, 4).
__device__ __forceinline__ uint32_t group8_or(uint32_t v) {
    v |= (uint32_t)__builtin_amdgcn_ds_swizzle((int)v, 0x1F | (1 << 10));
    v |= (uint32_t)__builtin_amdgcn_ds_swizzle((int)v, 0x1F | (2 << 10));
    v |= (uint32_t)__builtin_amdgcn_ds_swizzle((int)v, 0x1F | (4 << 10));
    return v;
}

// Eight lanes per valid cube: lane c of the group loads corner c's value and
// its six central-difference samples and forms the corner position and normal
// (the arithmetic of the reference's per-cube loop, corner by corner); the
// group ORs the corner bits into the cube index and exchanges the corner data
// through LDS; lane c then writes the cube's vertices c and c + 8.  `corner`
// is the block's LDS exchange (two float4 per thread); the caller keeps the 8
// lanes of a group in step (they sit in one wave).
template <bool kCpu>
__device__ __forceinline__ void emit_cube(const McArgs& a, uint32_t vid, uint32_t off,
                                          float4 (*corner)[2]) {
    const uint32_t rx = a.sc.rx, ry = a.sc.ry, rz = a.sc.rz;
    const uint64_t rxy = (uint64_t)rx * ry;
    const float sign = (!kCpu && a.iso < 0.0f) ? -1.0f : 1.0f;
    const uint32_t c = threadIdx.x & 7u;
    const uint32_t g0 = threadIdx.x & ~7u;
    const int dx = (c >= 4) ? 1 : 0;
    const int dy = (c == 2 || c == 3 || c == 6 || c == 7) ? 1 : 0;
    const int dz = (c == 1 || c == 2 || c == 5 || c == 6) ? 1 : 0;
    {
        const int x = (int)(vid % rx), y = (int)((vid / rx) % ry), z = (int)(vid / rxy);
        const int X = x + dx, Y = y + dy, Z = z + dz;
        const float val = a.phi[((uint64_t)Z * ry + Y) * rx + X];
        float cpx, cpy, cpz;
        if (kCpu) {
            cpx = a.sc.minx + (float)X * a.sc.cell;
            cpy = a.sc.miny + (float)Y * a.sc.cell;
            cpz = a.sc.minz + (float)Z * a.sc.cell;
        } else {
            cpx = vertex_coord((uint32_t)X, a.sc.minx, a.sc.cell);
            cpy = vertex_coord((uint32_t)Y, a.sc.miny, a.sc.cell);
            cpz = vertex_coord((uint32_t)Z, a.sc.minz, a.sc.cell);
        }
        float nx = sample(a.phi, X - 1, Y, Z, rx, ry, rz) - sample(a.phi, X + 1, Y, Z, rx, ry, rz);
        float ny = sample(a.phi, X, Y - 1, Z, rx, ry, rz) - sample(a.phi, X, Y + 1, Z, rx, ry, rz);
        float nz = sample(a.phi, X, Y, Z - 1, rx, ry, rz) - sample(a.phi, X, Y, Z + 1, rx, ry, rz);
        if (kCpu) {
            const float len = __builtin_sqrtf(nx * nx + ny * ny + nz * nz);
            if (len > 1.0e-7f) {
                const float fac = 1.0f / len;
                nx *= fac; ny *= fac; nz *= fac;
            }
        } else {
            const float rs = __builtin_amdgcn_rsqf(
                __builtin_fmaf(nz, nz, __builtin_fmaf(ny, ny, nx * nx)));
            nx *= rs; ny *= rs; nz *= rs;
        }
        corner[threadIdx.x][0] = make_float4(val, cpx, cpy, cpz);
        corner[threadIdx.x][1] = make_float4(nx, ny, nz, 0.0f);
        const uint32_t f = group8_or((val <= a.iso) ? (1u << c) : 0u);
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        const uint32_t nv = case_vertices(f);
        if ((uint64_t)off + nv <= a.capacity) {
            const uint64_t tri = cMcTri[f];
            for (uint32_t k = c; k < nv; k += 8u) {
                const uint32_t e = (uint32_t)(tri >> (4 * k)) & 0xFu;
                // corners of edge e
                const uint32_t s0 = (e < 8) ? (e & 3u) + (e & 4u) : e - 8u;
                const uint32_t s1 = (e < 8) ? (((e & 3u) + 1u) & 3u) + (e & 4u) : e - 4u;
                const float4 p0 = corner[g0 + s0][0], p1 = corner[g0 + s1][0];
                const float4 n0 = corner[g0 + s0][1], n1 = corner[g0 + s1][1];
                const float v0 = p0.x, v1 = p1.x;
                const float delta = v1 - v0;
                float t;
                if (kCpu) t = ((double)delta > -1.0e-7 && (double)delta < 1.0e-7) ? 0.5f : (a.iso - v0) / delta;
                else t = (delta > -1.0e-7f && delta < 1.0e-7f) ? 0.5f : (a.iso - v0) / delta;
                float px, py, pz, mx, my, mz;
                if (kCpu) {
                    px = p0.y + (p1.y - p0.y) * t;
                    py = p0.z + (p1.z - p0.z) * t;
                    pz = p0.w + (p1.w - p0.w) * t;
                    mx = n0.x + (n1.x - n0.x) * t;
                    my = n0.y + (n1.y - n0.y) * t;
                    mz = n0.z + (n1.z - n0.z) * t;
                    const float len = __builtin_sqrtf(mx * mx + my * my + mz * mz);
                    if (len > 1.0e-7f) {
                        const float fac = 1.0f / len;
                        mx *= fac; my *= fac; mz *= fac;
                    }
                } else {
                    px = __builtin_fmaf(t, p1.y - p0.y, p0.y);
                    py = __builtin_fmaf(t, p1.z - p0.z, p0.z);
                    pz = __builtin_fmaf(t, p1.w - p0.w, p0.w);
                    mx = __builtin_fmaf(t, n1.x - n0.x, n0.x);
                    my = __builtin_fmaf(t, n1.y - n0.y, n0.y);
                    mz = __builtin_fmaf(t, n1.z - n0.z, n0.z);
                    const float rs = __builtin_amdgcn_rsqf(
                        __builtin_fmaf(mz, mz, __builtin_fmaf(my, my, mx * mx)));
                    mx = sign * (mx * rs); my = sign * (my * rs); mz = sign * (mz * rs);
                }
                const uint64_t o = 3ull * (off + k);
                a.tri_pos[o + 0] = px;
                a.tri_pos[o + 1] = py;
                a.tri_pos[o + 2] = pz;
                if (a.tri_nrm) {
                    a.tri_nrm[o + 0] = mx;
                    a.tri_nrm[o + 1] = my;
                    a.tri_nrm[o + 2] = mz;
                }
            }
        }
        // the group's corner slots are rewritten by its next cube
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
}

template <bool kCpu>
__global__ __launch_bounds__(256) void k_mc_emit(McArgs a) {
    __shared__ float4 corner[256][2];  // per lane: (phi, x, y, z), (nx, ny, nz, 0)
    const uint32_t Q = a.st->num_valid_cubes;
    if (blockIdx.x == 0 && threadIdx.x == 0 && a.st->num_mesh_vertices > a.capacity)
        a.st->overflow = 1u;
    const uint32_t groups = gridDim.x * (blockDim.x / 8u);
    for (uint32_t q = blockIdx.x * (blockDim.x / 8u) + threadIdx.x / 8u; q < Q; q += groups) {
        const uint32_t b = a.valid_idx[q];
        emit_cube<kCpu>(a, a.band_index[b], a.vert_off[b], corner);
    }
}

// ---------------------------------------------------------------------------
// Deduplicated mesh: one vertex per grid edge (v, v + e_axis) that is crossed by
// the iso-surface and belongs to an emitting cube (the case tables place a
// vertex on exactly the crossed edges of a cube, so these are the soup's
// distinct vertices), owned by its lower end v.  Vertices are numbered in
// ascending (band rank of v, axis x, y, z) order; each soup slot becomes an
// index into them.  The vertex is interpolated from the lower end to the upper
// one, so a soup vertex whose cube walks the edge the other way differs from
// it only by the rounding of the lerp.
// ---------------------------------------------------------------------------

__device__ __forceinline__ uint32_t band_rank(const uint32_t* __restrict__ bits,
                                              const uint32_t* __restrict__ word_off,
                                              uint32_t wx, uint32_t ry, uint32_t x, uint32_t y,
                                              uint32_t z, bool* in) {
    const uint64_t w = ((uint64_t)z * ry + y) * wx + (x >> 5);
    const uint32_t word = bits[w];
    *in = (word >> (x & 31u)) & 1u;
    return word_off[w] + __popc(word & ((1u << (x & 31u)) - 1u));
}

template <bool kCpu>
__device__ __forceinline__ float corner_coord(uint32_t i, float mn, float cell) {
    return kCpu ? mn + (float)i * cell : vertex_coord(i, mn, cell);
}

// Central-difference normal at a grid vertex, normalised as the emitter does.
template <bool kCpu>
__device__ __forceinline__ void corner_normal(const float* __restrict__ phi, int X, int Y, int Z,
                                              uint32_t rx, uint32_t ry, uint32_t rz, float* n) {
    float nx = sample(phi, X - 1, Y, Z, rx, ry, rz) - sample(phi, X + 1, Y, Z, rx, ry, rz);
    float ny = sample(phi, X, Y - 1, Z, rx, ry, rz) - sample(phi, X, Y + 1, Z, rx, ry, rz);
    float nz = sample(phi, X, Y, Z - 1, rx, ry, rz) - sample(phi, X, Y, Z + 1, rx, ry, rz);
    if (kCpu) {
        const float len = __builtin_sqrtf(nx * nx + ny * ny + nz * nz);
        if (len > 1.0e-7f) {
            const float fac = 1.0f / len;
            nx *= fac; ny *= fac; nz *= fac;
        }
    } else {
        const float rs = __builtin_amdgcn_rsqf(__builtin_fmaf(nz, nz, __builtin_fmaf(ny, ny, nx * nx)));
        nx *= rs; ny *= rs; nz *= rs;
    }
    n[0] = nx; n[1] = ny; n[2] = nz;
}

__global__ __launch_bounds__(256) void k_mc_edges(McArgs a) {
    if (blockIdx.x == 0 && threadIdx.x == 0 && a.st->num_mesh_vertices > a.capacity)
        a.st->overflow = 1u;
    const uint32_t B = a.st->num_band_vertices;
    const uint32_t res[3] = {a.sc.rx, a.sc.ry, a.sc.rz};
    const uint64_t rxy = (uint64_t)res[0] * res[1];
    const uint64_t stride[3] = {1u, res[0], rxy};
    const uint32_t wx = row_words(res[0]);
    for (uint32_t b = blockIdx.x * blockDim.x + threadIdx.x; b < B; b += gridDim.x * blockDim.x) {
        const uint32_t vid = a.band_index[b];
        const uint32_t c[3] = {vid % res[0], (uint32_t)((vid / res[0]) % res[1]),
                               (uint32_t)(vid / rxy)};
        const bool below = a.phi[vid] <= a.iso;
        uint32_t mask = 0;
        for (int ax = 0; ax < 3; ++ax) {
            if (c[ax] + 1 >= res[ax]) continue;
            if ((a.phi[vid + stride[ax]] <= a.iso) == below) continue;
            const int p = ax == 0 ? 1 : 0, q = ax == 2 ? 1 : 2;
            bool used = false;
            for (uint32_t d = 0; d < 4 && !used; ++d) {
                uint32_t u[3] = {c[0], c[1], c[2]};
                const uint32_t dp = d & 1u, dq = d >> 1;
                if (u[p] < dp || u[q] < dq) continue;
                u[p] -= dp;
                u[q] -= dq;
                if (u[0] + 1 >= res[0] || u[1] + 1 >= res[1] || u[2] + 1 >= res[2]) continue;
                bool in;
                const uint32_t cb = band_rank(a.bits, a.word_off, wx, res[1], u[0], u[1], u[2], &in);
                used = in && a.nverts[cb] > 0;
            }
            if (used) mask |= 1u << ax;
        }
        a.edge_mask[b] = mask;
    }
}

template <bool kCpu>
__global__ __launch_bounds__(256) void k_mc_vertices(McArgs a) {
    const uint32_t B = a.st->num_band_vertices;
    const uint32_t U = a.st->num_unique_vertices;
    const uint32_t rx = a.sc.rx, ry = a.sc.ry, rz = a.sc.rz;
    const uint64_t rxy = (uint64_t)rx * ry;
    const float sign = (!kCpu && a.iso < 0.0f) ? -1.0f : 1.0f;
    if (blockIdx.x == 0 && threadIdx.x == 0 && U > a.vtx_capacity) a.st->overflow = 1u;
    for (uint32_t b = blockIdx.x * blockDim.x + threadIdx.x; b < B; b += gridDim.x * blockDim.x) {
        const uint32_t mask = a.edge_mask[b];
        if (!mask) continue;
        const uint32_t base = a.edge_off[b];
        const uint32_t vid = a.band_index[b];
        const int x = (int)(vid % rx), y = (int)((vid / rx) % ry), z = (int)(vid / rxy);
        const float va = a.phi[vid];
        const float pa[3] = {corner_coord<kCpu>(x, a.sc.minx, a.sc.cell),
                             corner_coord<kCpu>(y, a.sc.miny, a.sc.cell),
                             corner_coord<kCpu>(z, a.sc.minz, a.sc.cell)};
        float na[3];
        corner_normal<kCpu>(a.phi, x, y, z, rx, ry, rz, na);
        uint32_t slot = base;
        for (int ax = 0; ax < 3; ++ax) {
            if (!((mask >> ax) & 1u)) continue;
            const uint32_t k = slot++;
            if (k >= a.vtx_capacity) continue;
            const int X = x + (ax == 0), Y = y + (ax == 1), Z = z + (ax == 2);
            const float vb = a.phi[((uint64_t)Z * ry + Y) * rx + X];
            const float pb[3] = {corner_coord<kCpu>(X, a.sc.minx, a.sc.cell),
                                 corner_coord<kCpu>(Y, a.sc.miny, a.sc.cell),
                                 corner_coord<kCpu>(Z, a.sc.minz, a.sc.cell)};
            float nb[3];
            corner_normal<kCpu>(a.phi, X, Y, Z, rx, ry, rz, nb);
            const float delta = vb - va;
            float t;
            if (kCpu) t = ((double)delta > -1.0e-7 && (double)delta < 1.0e-7) ? 0.5f : (a.iso - va) / delta;
            else t = (delta > -1.0e-7f && delta < 1.0e-7f) ? 0.5f : (a.iso - va) / delta;
            float p[3], n[3];
            for (int i = 0; i < 3; ++i) {
                if (kCpu) {
                    p[i] = pa[i] + (pb[i] - pa[i]) * t;
                    n[i] = na[i] + (nb[i] - na[i]) * t;
                } else {
                    p[i] = __builtin_fmaf(t, pb[i] - pa[i], pa[i]);
                    n[i] = __builtin_fmaf(t, nb[i] - na[i], na[i]);
                }
            }
            if (kCpu) {
                const float len = __builtin_sqrtf(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
                if (len > 1.0e-7f) {
                    const float fac = 1.0f / len;
                    n[0] *= fac; n[1] *= fac; n[2] *= fac;
                }
            } else {
                const float rs = __builtin_amdgcn_rsqf(
                    __builtin_fmaf(n[2], n[2], __builtin_fmaf(n[1], n[1], n[0] * n[0])));
                n[0] = sign * (n[0] * rs); n[1] = sign * (n[1] * rs); n[2] = sign * (n[2] * rs);
            }
            const uint64_t o = 3ull * k;
            a.vtx_pos[o + 0] = p[0];
            a.vtx_pos[o + 1] = p[1];
            a.vtx_pos[o + 2] = p[2];
            if (a.vtx_nrm) {
                a.vtx_nrm[o + 0] = n[0];
                a.vtx_nrm[o + 1] = n[1];
                a.vtx_nrm[o + 2] = n[2];
            }
        }
    }
}

__global__ __launch_bounds__(256) void k_mc_indices(McArgs a) {
    const uint32_t Q = a.st->num_valid_cubes;
    const uint32_t rx = a.sc.rx, ry = a.sc.ry;
    const uint64_t rxy = (uint64_t)rx * ry;
    const uint32_t wx = row_words(rx);
    for (uint32_t q = blockIdx.x * blockDim.x + threadIdx.x; q < Q; q += gridDim.x * blockDim.x) {
        const uint32_t b = a.valid_idx[q];
        const uint32_t vid = a.band_index[b];
        const uint32_t off = a.vert_off[b];
        const uint32_t x = vid % rx, y = (uint32_t)((vid / rx) % ry), z = (uint32_t)(vid / rxy);
        const uint32_t f = cube_flag(a.phi, vid, rx, rxy, a.iso);
        const uint32_t nv = case_vertices(f);
        if ((uint64_t)off + nv > a.capacity) continue;
        const uint64_t tri = cMcTri[f];
        for (uint32_t k = 0; k < nv; ++k) {
            const uint32_t e = (uint32_t)(tri >> (4 * k)) & 0xFu;
            const uint32_t s0 = (e < 8) ? (e & 3u) + (e & 4u) : e - 8u;
            const uint32_t s1 = (e < 8) ? (((e & 3u) + 1u) & 3u) + (e & 4u) : e - 4u;
            const uint32_t o0[3] = {s0 >= 4 ? 1u : 0u,
                                    (s0 == 2 || s0 == 3 || s0 == 6 || s0 == 7) ? 1u : 0u,
                                    (s0 == 1 || s0 == 2 || s0 == 5 || s0 == 6) ? 1u : 0u};
            const uint32_t o1[3] = {s1 >= 4 ? 1u : 0u,
                                    (s1 == 2 || s1 == 3 || s1 == 6 || s1 == 7) ? 1u : 0u,
                                    (s1 == 1 || s1 == 2 || s1 == 5 || s1 == 6) ? 1u : 0u};
            const int ax = (o0[0] != o1[0]) ? 0 : ((o0[1] != o1[1]) ? 1 : 2);
            const bool lo0 = o0[ax] < o1[ax];
            const uint32_t* lo = lo0 ? o0 : o1;
            bool in;
            const uint32_t ob = band_rank(a.bits, a.word_off, wx, ry, x + lo[0], y + lo[1],
                                          z + lo[2], &in);
            a.indices[off + k] = a.edge_off[ob] + __popc(a.edge_mask[ob] & ((1u << ax) - 1u));
        }
    }
}

}  // namespace

fsr_status mc_run_indexed(fsr_context* ctx, const McArgs& a, hipStream_t s) {
    FSR_TRY(mc_classify(ctx, a, s));
    FSR_TRY(scan_mc(ctx, a.nverts, a.max_band, &a.st->num_band_vertices, a.vert_off, a.valid_idx,
                    &a.st->num_mesh_vertices, &a.st->num_valid_cubes, s));
    const int blocks = persistent_blocks(ctx, 256, 8, TUNE_MC);
    hipLaunchKernelGGL(k_mc_edges, dim3(blocks), dim3(256), 0, s, a);
    FSR_LAUNCH_CHECK("k_mc_edges");
    FSR_TRY(scan_exclusive(ctx, SCAN_POPC, a.edge_mask, a.edge_off, a.max_band,
                           &a.st->num_band_vertices, &a.st->num_unique_vertices, s));
    if (a.cpu_flavor)
        hipLaunchKernelGGL(k_mc_vertices<true>, dim3(blocks), dim3(256), 0, s, a);
    else
        hipLaunchKernelGGL(k_mc_vertices<false>, dim3(blocks), dim3(256), 0, s, a);
    FSR_LAUNCH_CHECK("k_mc_vertices");
    hipLaunchKernelGGL(k_mc_indices, dim3(blocks), dim3(256), 0, s, a);
    FSR_LAUNCH_CHECK("k_mc_indices");
    return FSR_OK;
}

fsr_status mc_classify(fsr_context* ctx, const McArgs& a, hipStream_t s) {
    const int blocks = persistent_blocks(ctx, 256, 8, TUNE_MC);
    hipLaunchKernelGGL(k_mc_classify, dim3(blocks), dim3(256), 0, s, a);
    FSR_LAUNCH_CHECK("k_mc_classify");
    return FSR_OK;
}

fsr_status mc_emit(fsr_context* ctx, const McArgs& a, hipStream_t s) {
    const int blocks = persistent_blocks(ctx, 256, 8, TUNE_MC);
    if (a.cpu_flavor)
        hipLaunchKernelGGL(k_mc_emit<true>, dim3(blocks), dim3(256), 0, s, a);
    else
        hipLaunchKernelGGL(k_mc_emit<false>, dim3(blocks), dim3(256), 0, s, a);
    FSR_LAUNCH_CHECK("k_mc_emit");
    return FSR_OK;
}

fsr_status mc_run(fsr_context* ctx, const McArgs& a, hipStream_t s) {
    FSR_TRY(mc_classify(ctx, a, s));
    FSR_TRY(scan_mc(ctx, a.nverts, a.max_band, &a.st->num_band_vertices, a.vert_off, a.valid_idx,
                    &a.st->num_mesh_vertices, &a.st->num_valid_cubes, s));
    return mc_emit(ctx, a, s);
}

}  // namespace fsr
