// Adjoint of the ZB05 field and of the marching-cubes vertex interpolation with
// respect to the particle positions, marching-cubes topology held fixed.  The
// reference has no backward pass; the forward semantics differentiated here are
// those of updateScalarFieldValueZB05 (ReconstructionCUDA.cu:230-289,
// kernelZB05 ReconstructionCUDA.cuh:262-268) and of generateTriangles
// (ReconstructionCUDA.cu:383-463, vertex lerp only).
//
// Vertex term (mc_vertex_backward).  A soup vertex on the edge (a, b) of a cube
// sits at c_a + t (c_b - c_a), t = (iso - phi_a) / (phi_b - phi_a), so
//   dt/dphi_a = (iso - phi_b) / D^2,  dt/dphi_b = -(iso - phi_a) / D^2,  D = phi_b - phi_a
// and an edge with |D| < 1e-7 (t fixed at 0.5) carries no gradient.  Normals
// are treated as constants.  Each emitted cube sums, per corner, the terms of
// its soup vertices on the edges at that corner (8 lanes per cube); each band
// vertex then adds the sums of the (up to) 8 cubes it is a corner of, in a
// fixed order: no floating-point atomics, the result is deterministic.
//
// Field term (field_zb05_backward).  With w_i = t_i^3, t_i = 1 - |x - p_i|^2/R^2,
// W = sum w_i, xbar = sum w_i p_i / W and phi = |x - xbar| - r_p:
//   dphi/dp_j = -(1/W) [ w_j n + (n . (p_j - xbar)) (6/R^2) t_j^2 (x - p_j) ],
//   n = (x - xbar) / |x - xbar|.
// A per-vertex pass folds the upstream gradient g into u = (g/W) n and
// c = (g/W) |x - xbar| (the forward kept x - xbar and W per vertex), so that
//   g dphi/dp_j = -[ w_j u + (c - u . (x - p_j)) (6/R^2) t_j^2 (x - p_j) ].
// The particle gradient is then a gather: a wave owns up to 64 particles of one
// hash cell and walks the band vertices of the 3^3 virtual cells around it (the
// cell-major band list keeps x-adjacent cells contiguous: 9 runs), staging 64
// vertices at a time in LDS as the forward stages particles.  A vertex in
// virtual cell v reads the hash cells clamp(v - 1 .. v + 1), so particle cell q
// is read exactly by the vertices of virtual cells q - 1 .. q + 1.  Neighbour
// lists are not stored: the gather walks the same cell ranges the forward used.
// The pairs are evaluated on the matrix cores (k_field_zb05_backward_mc).
// Results are written in the caller's particle order through the sort
// permutation; particles outside the hash grid get 0.
#include "common.hpp"
#include <type_traits>

#include "mc_tables.h"

namespace fsr {

namespace {

constexpr int kAdjThreads = 256;

__constant__ uint64_t cAdjTri[256] = {FSR_MC_TRI_PACKED_VALUES};

__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ uint32_t corner_dx(uint32_t c) { return c >= 4 ? 1u : 0u; }
__device__ __forceinline__ uint32_t corner_dy(uint32_t c) {
    return (c == 2 || c == 3 || c == 6 || c == 7) ? 1u : 0u;
}
__device__ __forceinline__ uint32_t corner_dz(uint32_t c) {
    return (c == 1 || c == 2 || c == 5 || c == 6) ? 1u : 0u;
}

// ---------------------------------------------------------------------------
// vertex term: g_phi[v] += sum over soup vertices on edges at v
// ---------------------------------------------------------------------------
struct McBackArgs {
    const uint32_t* band_index;
    const uint32_t* valid_idx;
    uint32_t* touched;          // bit mask over the grid (band mask layout)
    const uint32_t* bits;
    const uint32_t* word_off;
    const uint32_t* nverts;
    const uint32_t* vert_off;
    const float* phi;
    const float* g_tri;
    float* g_phi;
    float* corner;              // [8 (capacity / 3)] corner sums per emitted cube
    const fsr_frame_stats* st;
    GridDev sc;
    float iso;
    uint32_t capacity;
    int cpu_flavor;
};

// Pass 1, one 8-lane group per emitted cube (lane c = corner c): the terms of
// the cube's soup vertices (table order) on the edges at corner c, summed into
// corner[8 t0 + c], t0 = the cube's first triangle (vert_off / 3, unique per
// cube); and the touched bits of the cube's corners, so that pass 2 skips,
// after one bit test, the band vertices no soup vertex depends on.
__global__ __launch_bounds__(kAdjThreads) void k_mc_corner_terms(McBackArgs a) {
    const uint32_t Q = a.st->num_valid_cubes;
    const uint32_t rx = a.sc.rx, ry = a.sc.ry;
    const uint64_t rxy = (uint64_t)rx * ry;
    const uint32_t wx = row_words(rx);
    const uint32_t c = threadIdx.x & 7u;
    const int g0 = (int)(__lane_id() & ~7u);
    const uint32_t groups = gridDim.x * (blockDim.x / 8u);
    for (uint32_t q = blockIdx.x * (blockDim.x / 8u) + threadIdx.x / 8u; q < Q; q += groups) {
        const uint32_t cb = a.valid_idx[q];
        const uint32_t off = a.vert_off[cb], nv = a.nverts[cb];
        if ((uint64_t)off + nv > a.capacity) continue;  // the forward did not emit it
        const uint32_t vid = a.band_index[cb];
        const uint32_t x = vid % rx, y = (vid / rx) % ry, z = (uint32_t)(vid / rxy);
        const uint32_t dx = corner_dx(c), dy = corner_dy(c), dz = corner_dz(c);
        if (dx == 0u) {
            // corners x and x + 1 of this lane's (y, z) row; both bits may share a word
            const uint64_t row = (uint64_t)(z + dz) * ry + (y + dy);
            const uint64_t w0 = row * wx + (x >> 5);
            const uint32_t bit = x & 31u;
            if (bit < 31u) {
                atomicOr(&a.touched[w0], 3u << bit);
            } else {
                atomicOr(&a.touched[w0], 1u << 31);
                atomicOr(&a.touched[w0 + 1], 1u);
            }
        }
        const uint64_t u = ((uint64_t)z * ry + y) * rx + x;
        const float val = a.phi[u + dx + (uint64_t)dy * rx + (uint64_t)dz * rxy];
        uint32_t f = (val <= a.iso) ? (1u << c) : 0u;
        f |= (uint32_t)__builtin_amdgcn_ds_swizzle((int)f, 0x1F | (1 << 10));
        f |= (uint32_t)__builtin_amdgcn_ds_swizzle((int)f, 0x1F | (2 << 10));
        f |= (uint32_t)__builtin_amdgcn_ds_swizzle((int)f, 0x1F | (4 << 10));
        const uint64_t tri = cAdjTri[f];
        float acc = 0.0f;
        for (uint32_t k = 0; k < nv; ++k) {
            const uint32_t e = (uint32_t)(tri >> (4 * k)) & 0xFu;
            const uint32_t s0 = (e < 8) ? (e & 3u) + (e & 4u) : e - 8u;
            const uint32_t s1 = (e < 8) ? (((e & 3u) + 1u) & 3u) + (e & 4u) : e - 4u;
            // the edge's corner values from their lanes (the whole group runs
            // the same k: nv is the group's)
            const float v0 = __shfl(val, g0 + (int)s0, 64);
            const float v1 = __shfl(val, g0 + (int)s1, 64);
            if (s0 != c && s1 != c) continue;
            const float delta = v1 - v0;
            const bool flat = a.cpu_flavor
                                  ? ((double)delta > -1.0e-7 && (double)delta < 1.0e-7)
                                  : (delta > -1.0e-7f && delta < 1.0e-7f);
            if (flat) continue;
            // d(position)/dt: corner s1 minus corner s0 as the forward evaluates them
            const uint32_t X0 = x + corner_dx(s0), Y0 = y + corner_dy(s0), Z0 = z + corner_dz(s0);
            const uint32_t X1 = x + corner_dx(s1), Y1 = y + corner_dy(s1), Z1 = z + corner_dz(s1);
            float ex, ey, ez;
            if (a.cpu_flavor) {
                ex = (a.sc.minx + (float)X1 * a.sc.cell) - (a.sc.minx + (float)X0 * a.sc.cell);
                ey = (a.sc.miny + (float)Y1 * a.sc.cell) - (a.sc.miny + (float)Y0 * a.sc.cell);
                ez = (a.sc.minz + (float)Z1 * a.sc.cell) - (a.sc.minz + (float)Z0 * a.sc.cell);
            } else {
                ex = vertex_coord(X1, a.sc.minx, a.sc.cell) - vertex_coord(X0, a.sc.minx, a.sc.cell);
                ey = vertex_coord(Y1, a.sc.miny, a.sc.cell) - vertex_coord(Y0, a.sc.miny, a.sc.cell);
                ez = vertex_coord(Z1, a.sc.minz, a.sc.cell) - vertex_coord(Z0, a.sc.minz, a.sc.cell);
            }
            const uint64_t o = 3ull * (off + k);
            const float gt = __builtin_fmaf(a.g_tri[o + 2], ez,
                                            __builtin_fmaf(a.g_tri[o + 1], ey, a.g_tri[o] * ex));
            const float inv_d2 = 1.0f / (delta * delta);
            acc += (s0 == c) ? gt * (a.iso - v1) * inv_d2 : -gt * (a.iso - v0) * inv_d2;
        }
        a.corner[8ull * (off / 3u) + c] = acc;
    }
}

// Pass 2, one thread per band vertex: the corner sums of the (up to) 8 emitted
// cubes it is a corner of, in corner order.
__global__ __launch_bounds__(kAdjThreads) void k_mc_vertex_backward(McBackArgs a) {
    const uint32_t B = a.st->num_band_vertices;
    const uint32_t rx = a.sc.rx, ry = a.sc.ry, rz = a.sc.rz;
    const uint64_t rxy = (uint64_t)rx * ry;
    const uint32_t wx = row_words(rx);
    for (uint32_t b = blockIdx.x * blockDim.x + threadIdx.x; b < B; b += gridDim.x * blockDim.x) {
        const uint32_t vid = a.band_index[b];
        const uint32_t x = vid % rx, y = (vid / rx) % ry, z = (uint32_t)(vid / rxy);
        if (!((a.touched[((uint64_t)z * ry + y) * wx + (x >> 5)] >> (x & 31u)) & 1u)) continue;
        // the cubes whose corner cv is this vertex, level by level over the 8
        // (each level's loads issued together): mask word, band rank, vertex
        // count, offset, corner sum
        uint64_t word[8];
        uint32_t bw[8], sh[8];
        bool ok[8];
#pragma unroll
        for (uint32_t cv = 0; cv < 8; ++cv) {
            const uint32_t ox = corner_dx(cv), oy = corner_dy(cv), oz = corner_dz(cv);
            const uint32_t ux = x - ox, uy = y - oy, uz = z - oz;
            ok[cv] = x >= ox && y >= oy && z >= oz && ux + 1 < rx && uy + 1 < ry && uz + 1 < rz;
            word[cv] = ((uint64_t)uz * ry + uy) * wx + (ux >> 5);
            sh[cv] = ux & 31u;
            bw[cv] = ok[cv] ? a.bits[word[cv]] : 0u;
        }
        uint32_t cb[8];
#pragma unroll
        for (uint32_t cv = 0; cv < 8; ++cv) {
            ok[cv] = ok[cv] && ((bw[cv] >> sh[cv]) & 1u);
            cb[cv] = ok[cv] ? a.word_off[word[cv]] + __popc(bw[cv] & ((1u << sh[cv]) - 1u)) : 0u;
        }
        uint32_t nv[8];
#pragma unroll
        for (uint32_t cv = 0; cv < 8; ++cv) nv[cv] = ok[cv] ? a.nverts[cb[cv]] : 0u;
        uint32_t off[8];
#pragma unroll
        for (uint32_t cv = 0; cv < 8; ++cv) {
            off[cv] = nv[cv] ? a.vert_off[cb[cv]] : 0u;
        }
        float acc = 0.0f;
#pragma unroll
        for (uint32_t cv = 0; cv < 8; ++cv) {
            // a cube the forward did not emit (beyond the soup capacity) adds nothing
            if (nv[cv] && (uint64_t)off[cv] + nv[cv] <= a.capacity)
                acc += a.corner[8ull * (off[cv] / 3u) + cv];
        }
        if (acc != 0.0f) a.g_phi[vid] += acc;
    }
}

// ---------------------------------------------------------------------------
// per band vertex (cell-major slot): u = (g/W) n, c = (g/W) |x - xbar|
// ---------------------------------------------------------------------------
// Per cell-major slot: (u, c) (zero when the vertex carries no gradient) and
// the vertex position, so that the gather below stages a vertex with two
// 16-byte loads and no index arithmetic.
__global__ __launch_bounds__(kAdjThreads) void k_field_coef(const uint32_t* __restrict__ cell_band,
                                                            const float4* __restrict__ saved,
                                                            const float* __restrict__ g_phi,
                                                            const fsr_frame_stats* st, GridDev sc,
                                                            float4* __restrict__ coef,
                                                            float4* __restrict__ vpos) {
    const uint32_t B = st->num_band_vertices;
    const uint32_t rxy = sc.rx * sc.ry;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < B; i += gridDim.x * blockDim.x) {
        const float4 s = saved[i];
        const uint32_t vid = cell_band[i];
        const float g = g_phi[vid];
        const uint32_t ix = vid % sc.rx, iy = (vid / sc.rx) % sc.ry, iz = vid / rxy;
        vpos[i] = make_float4(vertex_coord(ix, sc.minx, sc.cell), vertex_coord(iy, sc.miny, sc.cell),
                              vertex_coord(iz, sc.minz, sc.cell), 0.0f);
        float4 c = make_float4(0.f, 0.f, 0.f, 0.f);
        const float len = __builtin_sqrtf(__builtin_fmaf(s.z, s.z, __builtin_fmaf(s.y, s.y, s.x * s.x)));
        if (s.w > 0.0f && len > 0.0f && g != 0.0f) {
            const float gw = g / s.w;
            const float k = gw / len;
            c = make_float4(k * s.x, k * s.y, k * s.z, gw * len);
        }
        coef[i] = c;
    }
}

// ---------------------------------------------------------------------------
// particle gather
// ---------------------------------------------------------------------------
struct FieldBackArgs {
    const float4* sorted;
    const fsr_index_range* ranges;
    const uint32_t* perm;
    const uint32_t* chunk_cell;
    const uint32_t* chunk_off;
    const uint32_t* num_chunks;
    uint32_t* next;               // per-XCD fetch counters (next_item_xcd)
    const uint32_t* cell_off;     // virtual cell -> first cell-major slot
    const uint32_t* cell_cnt;
    const uint32_t* cell_band;    // cell-major slot -> vertex id
    const float4* coef;           // cell-major slot -> (u, c)
    const float4* vpos;           // cell-major slot -> vertex position
    GridDev sp, sc;
    uint32_t vcx, vcy;
    int radius;
    float radius_sq;
    float* g_pos;                 // float3[N], caller order
};

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

// The gather evaluates its pairs in 16 x 16 particle/vertex tiles on the
// matrix cores (as the forward field): per 16 staged vertices and 16
// particles of the chunk
//   1. v_mfma_f32_16x16x4_f32: t = c0_p + (2 p'/R^2).x' - |x'|^2/R^2 with A =
//      (x', y', z', 1) of the vertices, B = (2p'/R^2, c0_p) of the particles,
//      C = -|x'|^2/R^2; a second one gives kd = u.p' + (c - u.x') with A =
//      (u, c - u.x') of the vertices and B = (p', 1);
//   2. t clamped to [0, 1], w = t^3 and kk = kd t^2 (4 VALU ops per pair);
//   3. rank-1 v_mfma_f32_4x4x1_16b_f32 updates: sum w (u, c - u.x') (the
//      last component unused) and sum kk (x', 1).
// Vertices whose coefficient is zero, or that lie at least 1.00005 R from the
// bounding box of the chunk's particles, add exactly 0 to every particle and
// are dropped at staging (order-preserving compaction, as in the forward).
constexpr int kAdjQueue = 128;
constexpr int kAdjStride = 144;
// rows: x', y', z', 1, ux, uy, uz, c - u.x', -|x'|^2/R^2
constexpr int kAdjOne = 3, kAdjU = 4, kAdjQ = 8, kAdjRows = 9;
typedef float v4f __attribute__((ext_vector_type(4)));

template <int M>
__device__ __forceinline__ float adj_swz_xor(float v) {
    return __int_as_float(__builtin_amdgcn_ds_swizzle(__float_as_int(v), 0x1F | (M << 10)));
}

template <int NT>
__device__ __forceinline__ void adj_tiles(const float* __restrict__ rows, uint32_t n,
                                          const float (&bt)[NT], const float (&bk)[NT],
                                          v4f (&gu)[NT], v4f (&gx)[NT]) {
    const uint32_t lane = __lane_id();
    const float* a1p = rows + (lane >> 4) * kAdjStride + (lane & 15u);
    const float* k1p = rows + (kAdjU + (lane >> 4)) * kAdjStride + (lane & 15u);
    const float* a2p = rows + (lane & 3u) * kAdjStride + 4u * (lane >> 4);
    const float* u2p = rows + (kAdjU + (lane & 3u)) * kAdjStride + 4u * (lane >> 4);
    const float* cqp = rows + kAdjQ * kAdjStride + 4u * (lane >> 4);
    for (uint32_t tb = 0; tb < n; tb += 16u) {
        const float a1 = a1p[tb];
        const float k1 = k1p[tb];
        const float4 cq = *reinterpret_cast<const float4*>(cqp + tb);
        const v4f c1 = v4f{cq.x, cq.y, cq.z, cq.w};
        const v4f zero = v4f{0.f, 0.f, 0.f, 0.f};
        v4f t[NT], kd[NT];
#pragma unroll
        for (int T = 0; T < NT; ++T) t[T] = __builtin_amdgcn_mfma_f32_16x16x4f32(a1, bt[T], c1, 0, 0, 0);
#pragma unroll
        for (int T = 0; T < NT; ++T) kd[T] = __builtin_amdgcn_mfma_f32_16x16x4f32(k1, bk[T], zero, 0, 0, 0);
#pragma unroll
        for (int T = 0; T < NT; ++T) {
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const float tp = __builtin_amdgcn_fmed3f(t[T][r], 0.0f, 1.0f);
                const float t2 = tp * tp;
                t[T][r] = t2 * tp;
                kd[T][r] = kd[T][r] * t2;
            }
        }
        const float4 au = *reinterpret_cast<const float4*>(u2p + tb);
        const float4 ax = *reinterpret_cast<const float4*>(a2p + tb);
#pragma unroll
        for (int T = 0; T < NT; ++T) gu[T] = __builtin_amdgcn_mfma_f32_4x4x1f32(au.x, t[T][0], gu[T], 0, 0, 0);
#pragma unroll
        for (int T = 0; T < NT; ++T) gu[T] = __builtin_amdgcn_mfma_f32_4x4x1f32(au.y, t[T][1], gu[T], 0, 0, 0);
#pragma unroll
        for (int T = 0; T < NT; ++T) gu[T] = __builtin_amdgcn_mfma_f32_4x4x1f32(au.z, t[T][2], gu[T], 0, 0, 0);
#pragma unroll
        for (int T = 0; T < NT; ++T) gu[T] = __builtin_amdgcn_mfma_f32_4x4x1f32(au.w, t[T][3], gu[T], 0, 0, 0);
#pragma unroll
        for (int T = 0; T < NT; ++T) gx[T] = __builtin_amdgcn_mfma_f32_4x4x1f32(ax.x, kd[T][0], gx[T], 0, 0, 0);
#pragma unroll
        for (int T = 0; T < NT; ++T) gx[T] = __builtin_amdgcn_mfma_f32_4x4x1f32(ax.y, kd[T][1], gx[T], 0, 0, 0);
#pragma unroll
        for (int T = 0; T < NT; ++T) gx[T] = __builtin_amdgcn_mfma_f32_4x4x1f32(ax.z, kd[T][2], gx[T], 0, 0, 0);
#pragma unroll
        for (int T = 0; T < NT; ++T) gx[T] = __builtin_amdgcn_mfma_f32_4x4x1f32(ax.w, kd[T][3], gx[T], 0, 0, 0);
    }
}

__global__ __launch_bounds__(kAdjThreads, 4) void k_field_zb05_backward_mc(FieldBackArgs a) {
    struct WaveLds {
        float queue[kAdjRows * kAdjStride];
        float bt[4][64];   // per lane: (2p'/R^2, c0)
    };
    __shared__ __align__(16) WaveLds lds[kAdjThreads / 64];
    const uint32_t lane = __lane_id();
    const uint32_t wave = threadIdx.x >> 6;
    const uint32_t nch = *a.num_chunks;
    const float inv_r2 = 1.0f / a.radius_sq;
    const float cull_r2 = a.radius_sq * 1.0001f;
    float* q = lds[wave].queue;
    float* vb = &lds[wave].bt[0][0];
    q[kAdjOne * kAdjStride + lane] = 1.0f;
    q[kAdjOne * kAdjStride + 64 + lane] = 1.0f;
    uint32_t step = 0;
    for (;;) {
        const uint32_t ch = next_item_xcd(a.next, nch, step);
        if (ch >= nch) break;
        const uint32_t c = a.chunk_cell[ch];
        const uint32_t k = ch - a.chunk_off[c];
        const fsr_index_range own = a.ranges[c];
        const uint32_t i = own.start + 64u * k + lane;
        const bool live = i < own.end;
        const uint32_t cnt = min(64u, own.end - (own.start + 64u * k));
        const uint32_t qcx = c % a.sp.rx, qcy = (c / a.sp.rx) % a.sp.ry, qcz = c / (a.sp.rx * a.sp.ry);

        // 9 runs of band slots: virtual cells (q + radius) + {-1, 0, 1}^3
        uint32_t len = 0, rs = 0;
        if (lane < 9) {
            const uint32_t vy = qcy + a.radius + (lane % 3) - 1;
            const uint32_t vz = qcz + a.radius + (lane / 3) - 1;
            const uint32_t v0 = (vz * a.vcy + vy) * a.vcx + qcx + a.radius - 1;
            rs = a.cell_off[v0];
            len = a.cell_off[v0 + 2] + a.cell_cnt[v0 + 2] - rs;
        }
        const uint32_t incl = wave_inclusive_sum(len);
        const uint32_t total = (uint32_t)__builtin_amdgcn_readlane((int)incl, 63);
        const uint32_t rp = incl - len;
        const uint32_t ro = rs - rp;
        uint32_t jc = 0;
        auto run_offset = [&](uint32_t nb, uint32_t gg) {
            uint32_t off = (uint32_t)__builtin_amdgcn_readlane((int)ro, (int)jc);
            for (uint32_t jj = jc + 1u; jj < 9u; ++jj) {
                const uint32_t pj = (uint32_t)__builtin_amdgcn_readlane((int)rp, (int)jj);
                if (pj >= nb + 64u) break;
                if (gg >= pj) off = (uint32_t)__builtin_amdgcn_readlane((int)ro, (int)jj);
                jc = jj;
            }
            return off;
        };

        const float ox = __builtin_fmaf((float)qcx, a.sp.cell, a.sp.minx);
        const float oy = __builtin_fmaf((float)qcy, a.sp.cell, a.sp.miny);
        const float oz = __builtin_fmaf((float)qcz, a.sp.cell, a.sp.minz);
        float px = 0.f, py = 0.f, pz = 0.f;
        if (live) {
            const float4 p = a.sorted[i];
            px = p.x - ox; py = p.y - oy; pz = p.z - oz;
        }
        vb[lane] = 2.0f * px * inv_r2;
        vb[64 + lane] = 2.0f * py * inv_r2;
        vb[128 + lane] = 2.0f * pz * inv_r2;
        vb[192 + lane] = 1.0f - __builtin_fmaf(pz, pz, __builtin_fmaf(py, py, px * px)) * inv_r2;
        const float bx0 = wave_min(live ? px : 3.0e38f), bx1 = wave_max(live ? px : -3.0e38f);
        const float by0 = wave_min(live ? py : 3.0e38f), by1 = wave_max(live ? py : -3.0e38f);
        const float bz0 = wave_min(live ? pz : 3.0e38f), bz1 = wave_max(live ? pz : -3.0e38f);
        wave_lds_sync();

        auto chunk = [&](auto ntc) {
            constexpr int NT = decltype(ntc)::value;
            float bt[NT], bk[NT];
            v4f gu[NT], gx[NT];
            const float* bsrc = vb + 64 * (lane >> 4) + (lane & 15u);
            // (p', 1) of particle 16T + l%16, component l/16, from the LDS copy
            // of 2p'/R^2 (times R^2 / 2) and a 1 for the fourth
#pragma unroll
            for (int T = 0; T < NT; ++T) {
                bt[T] = bsrc[16 * T];
                bk[T] = (lane >> 4) == 3u ? 1.0f : bt[T] * (0.5f * a.radius_sq);
                gu[T] = v4f{0.f, 0.f, 0.f, 0.f};
                gx[T] = v4f{0.f, 0.f, 0.f, 0.f};
            }
            uint32_t head = 0, fill = 0;
            float4 ncf = make_float4(0.f, 0.f, 0.f, 0.f), nvp = ncf;
            {
                const uint32_t off = run_offset(0u, lane);
                if (lane < total) {
                    const uint32_t slot = off + lane;
                    ncf = a.coef[slot];
                    nvp = a.vpos[slot];
                }
            }
            for (uint32_t base = 0; base < total; base += 64u) {
                const uint32_t gi = base + lane;
                const float4 cf = ncf, vp = nvp;
                if (base + 64u < total) {
                    const uint32_t off = run_offset(base + 64u, gi + 64u);
                    if (gi + 64u < total) {
                        const uint32_t slot = off + gi + 64u;
                        ncf = a.coef[slot];
                        nvp = a.vpos[slot];
                    }
                }
                bool keep = false;
                float xx = 0.f, xy = 0.f, xz = 0.f;
                if (gi < total && (cf.x != 0.0f || cf.y != 0.0f || cf.z != 0.0f || cf.w != 0.0f)) {
                    xx = vp.x - ox;
                    xy = vp.y - oy;
                    xz = vp.z - oz;
                    const float dx = xx - __builtin_amdgcn_fmed3f(xx, bx0, bx1);
                    const float dy = xy - __builtin_amdgcn_fmed3f(xy, by0, by1);
                    const float dz = xz - __builtin_amdgcn_fmed3f(xz, bz0, bz1);
                    keep = __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, dx * dx)) < cull_r2;
                }
                const uint64_t m = __ballot(keep);
                if (keep) {
                    const uint32_t slot = (head + fill + mask_rank(m)) & (kAdjQueue - 1);
                    q[slot] = xx;
                    q[kAdjStride + slot] = xy;
                    q[2 * kAdjStride + slot] = xz;
                    q[kAdjU * kAdjStride + slot] = cf.x;
                    q[(kAdjU + 1) * kAdjStride + slot] = cf.y;
                    q[(kAdjU + 2) * kAdjStride + slot] = cf.z;
                    q[(kAdjU + 3) * kAdjStride + slot] =
                        cf.w - __builtin_fmaf(cf.z, xz, __builtin_fmaf(cf.y, xy, cf.x * xx));
                    q[kAdjQ * kAdjStride + slot] =
                        -(__builtin_fmaf(xz, xz, __builtin_fmaf(xy, xy, xx * xx)) * inv_r2);
                }
                fill += (uint32_t)__popcll(m);
                if (fill >= 64u) {
                    wave_lds_sync();
                    adj_tiles<NT>(q + head, 64u, bt, bk, gu, gx);
                    wave_lds_sync();
                    head ^= 64u;
                    fill -= 64u;
                }
            }
            if (fill) {
                // pad to 16 with entries that add 0: t < 0, u = 0
                const uint32_t n = (fill + 15u) & ~15u;
                if (lane >= fill && lane < n) {
#pragma unroll
                    for (int rr = 0; rr < kAdjQ; ++rr)
                        if (rr != kAdjOne) q[rr * kAdjStride + head + lane] = 0.f;
                    q[kAdjQ * kAdjStride + head + lane] = -1.0e30f;
                }
                wave_lds_sync();
                adj_tiles<NT>(q + head, n, bt, bk, gu, gx);
                wave_lds_sync();
            }
            // the four partial sums of each particle; lane l keeps particle l
            const uint32_t my = lane >> 4;
            v4f su = gu[0], sx = gx[0];
#pragma unroll
            for (int T = 0; T < NT; ++T) {
                v4f u4 = gu[T], x4 = gx[T];
#pragma unroll
                for (int cc = 0; cc < 4; ++cc) {
                    float y = u4[cc];
                    y += adj_swz_xor<16>(y);
                    u4[cc] = y + __shfl_xor(y, 32, 64);
                    float z = x4[cc];
                    z += adj_swz_xor<16>(z);
                    x4[cc] = z + __shfl_xor(z, 32, 64);
                }
                if (my == (uint32_t)T) {
                    su = u4;
                    sx = x4;
                }
            }
            if (live) {
                const float s6 = 6.0f * inv_r2;
                const uint64_t o = 3ull * a.perm[i];
                a.g_pos[o + 0] = -__builtin_fmaf(s6, __builtin_fmaf(-sx[3], px, sx[0]), su[0]);
                a.g_pos[o + 1] = -__builtin_fmaf(s6, __builtin_fmaf(-sx[3], py, sx[1]), su[1]);
                a.g_pos[o + 2] = -__builtin_fmaf(s6, __builtin_fmaf(-sx[3], pz, sx[2]), su[2]);
            }
        };
        switch ((cnt + 15u) >> 4) {
            case 1: chunk(std::integral_constant<int, 1>{}); break;
            case 2: chunk(std::integral_constant<int, 2>{}); break;
            case 3: chunk(std::integral_constant<int, 3>{}); break;
            default: chunk(std::integral_constant<int, 4>{}); break;
        }
    }
}

}  // namespace

fsr_status mc_vertex_backward(fsr_context* ctx, const float* g_tri_pos, float* g_phi,
                              hipStream_t s) {
    const AdjointState& st = ctx->adj;
    if (!st.tri_capacity) return FSR_OK;  // no soup vertices were emitted
    McBackArgs a;
    a.band_index = st.bc.band_index;
    a.valid_idx = st.valid_idx;
    const uint64_t nwords = (uint64_t)row_words(st.sc.rx) * st.sc.ry * st.sc.rz;
    FSR_TRY(arena_typed(ctx, A_ADJ_TOUCH, nwords, &a.touched));
    FSR_TRY(fill_u32(a.touched, 0u, nwords, s));
    a.bits = st.bits;
    a.word_off = st.word_off;
    a.nverts = st.nverts;
    a.vert_off = st.vert_off;
    a.phi = st.phi;
    a.g_tri = g_tri_pos;
    a.g_phi = g_phi;
    a.st = st.stats;
    a.sc = st.sc;
    a.iso = st.iso;
    a.capacity = st.tri_capacity;
    a.cpu_flavor = st.cpu_flavor;
    FSR_TRY(arena_typed(ctx, A_ADJ_CORNER, 8ull * (st.tri_capacity / 3u) + 8u, &a.corner));
    const int blocks = persistent_blocks(ctx, kAdjThreads, 8, TUNE_MC);
    hipLaunchKernelGGL(k_mc_corner_terms, dim3(blocks), dim3(kAdjThreads), 0, s, a);
    FSR_LAUNCH_CHECK("k_mc_corner_terms");
    hipLaunchKernelGGL(k_mc_vertex_backward, dim3(blocks), dim3(kAdjThreads), 0, s, a);
    FSR_LAUNCH_CHECK("k_mc_vertex_backward");
    return FSR_OK;
}

fsr_status field_zb05_backward(fsr_context* ctx, const float* g_phi, float* g_pos,
                               hipStream_t s) {
    const AdjointState& st = ctx->adj;
    if (st.n == 0) return FSR_OK;
    FSR_TRY(fill_f32(g_pos, 0.0f, 3ull * st.n, s));
    const uint64_t V = (uint64_t)st.sc.rx * st.sc.ry * st.sc.rz;
    const uint32_t C = st.sp.rx * st.sp.ry * st.sp.rz;
    float4* coef;
    FSR_TRY(arena_typed(ctx, A_ADJ_COEF, 2 * V, &coef));  // (u, c), then vertex positions
    float4* vpos = coef + V;
    {
        const int blocks = persistent_blocks(ctx, kAdjThreads, 4, TUNE_MC);
        hipLaunchKernelGGL(k_field_coef, dim3(blocks), dim3(kAdjThreads), 0, s, st.bc.cell_band,
                           st.saved, g_phi, st.stats, st.sc, coef, vpos);
        FSR_LAUNCH_CHECK("k_field_coef");
    }
    // particle chunks per hash cell
    const uint64_t max_chunks = (uint64_t)st.n / 64 + C + 1;
    uint32_t* aux;
    FSR_TRY(arena_typed(ctx, A_ADJ_AUX, 2ull * C + max_chunks, &aux));
    uint32_t* cnt = aux;
    uint32_t* chunk_off = aux + C;
    uint32_t* chunk_cell = aux + 2ull * C;
    uint32_t* total = &st.counters->aux[8];
    uint32_t* next = st.counters->xcd_next;
    FSR_TRY(reset_xcd(st.counters, s));
    FSR_TRY(cell_chunk_list(ctx, st.ranges, nullptr, C, cnt, chunk_off, chunk_cell, total, s));
    FieldBackArgs a;
    a.sorted = st.sorted;
    a.ranges = st.ranges;
    a.perm = st.perm;
    a.chunk_cell = chunk_cell;
    a.chunk_off = chunk_off;
    a.num_chunks = total;
    a.next = next;
    a.cell_off = st.bc.cell_off;
    a.cell_cnt = st.bc.cell_cnt;
    a.cell_band = st.bc.cell_band;
    a.coef = coef;
    a.vpos = vpos;
    a.sp = st.sp;
    a.sc = st.sc;
    a.vcx = st.bc.vcx;
    a.vcy = st.bc.vcy;
    a.radius = st.bc.radius;
    a.radius_sq = st.radius_sq;
    a.g_pos = g_pos;
    const int blocks = persistent_blocks(ctx, kAdjThreads, 4, TUNE_FIELD);
    hipLaunchKernelGGL(k_field_zb05_backward_mc, dim3(blocks), dim3(kAdjThreads), 0, s, a);
    FSR_LAUNCH_CHECK("k_field_zb05_backward_mc");
    return FSR_OK;
}

}  // namespace fsr
