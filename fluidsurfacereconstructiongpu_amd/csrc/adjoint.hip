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
// are treated as constants.  Each band vertex gathers the terms of the soup
// vertices on its edges from the (up to) 8 cubes it is a corner of, in a fixed
// order: no atomics, the result is deterministic.
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
// Per pair: t by the cell-local expansion (4 ops), t^2, t^3, the dot product
// (3), and three accumulations (7): 18 VALU ops.  Vertices that contribute
// exactly zero to every particle of the chunk are culled at staging.  Results
// are written in the caller's particle order through the sort permutation;
// particles outside the hash grid get 0.
#include "common.hpp"
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
    const fsr_frame_stats* st;
    GridDev sc;
    float iso;
    uint32_t capacity;
    int cpu_flavor;
};

// Marks the 8 corners of every emitted cube, so that the gather below skips,
// after one bit test, the band vertices no soup vertex depends on.
__global__ __launch_bounds__(kAdjThreads) void k_mc_touch(McBackArgs a) {
    const uint32_t Q = a.st->num_valid_cubes;
    const uint32_t rx = a.sc.rx, ry = a.sc.ry;
    const uint64_t rxy = (uint64_t)rx * ry;
    const uint32_t wx = row_words(rx);
    for (uint32_t q = blockIdx.x * blockDim.x + threadIdx.x; q < Q; q += gridDim.x * blockDim.x) {
        const uint32_t cb = a.valid_idx[q];
        if ((uint64_t)a.vert_off[cb] + a.nverts[cb] > a.capacity) continue;
        const uint32_t vid = a.band_index[cb];
        const uint32_t x = vid % rx, y = (vid / rx) % ry, z = (uint32_t)(vid / rxy);
        // corners x and x + 1 of the 4 (y, z) rows; both bits may share a word
        for (uint32_t d = 0; d < 4; ++d) {
            const uint64_t row = (uint64_t)(z + (d >> 1)) * ry + (y + (d & 1u));
            const uint64_t w0 = row * wx + (x >> 5);
            const uint32_t bit = x & 31u;
            if (bit < 31u) {
                atomicOr(&a.touched[w0], 3u << bit);
            } else {
                atomicOr(&a.touched[w0], 1u << 31);
                atomicOr(&a.touched[w0 + 1], 1u);
            }
        }
    }
}

__global__ __launch_bounds__(kAdjThreads) void k_mc_vertex_backward(McBackArgs a) {
    const uint32_t B = a.st->num_band_vertices;
    const uint32_t rx = a.sc.rx, ry = a.sc.ry, rz = a.sc.rz;
    const uint64_t rxy = (uint64_t)rx * ry;
    const uint32_t wx = row_words(rx);
    for (uint32_t b = blockIdx.x * blockDim.x + threadIdx.x; b < B; b += gridDim.x * blockDim.x) {
        const uint32_t vid = a.band_index[b];
        const uint32_t x = vid % rx, y = (vid / rx) % ry, z = (uint32_t)(vid / rxy);
        if (!((a.touched[((uint64_t)z * ry + y) * wx + (x >> 5)] >> (x & 31u)) & 1u)) continue;
        float acc = 0.0f;
        for (uint32_t cv = 0; cv < 8; ++cv) {
            // the cube whose corner cv is this vertex
            const uint32_t ox = corner_dx(cv), oy = corner_dy(cv), oz = corner_dz(cv);
            if (x < ox || y < oy || z < oz) continue;
            const uint32_t ux = x - ox, uy = y - oy, uz = z - oz;
            if (ux + 1 >= rx || uy + 1 >= ry || uz + 1 >= rz) continue;
            const uint64_t word = ((uint64_t)uz * ry + uy) * wx + (ux >> 5);
            const uint32_t bw = a.bits[word];
            if (!((bw >> (ux & 31u)) & 1u)) continue;
            const uint32_t cb = a.word_off[word] + __popc(bw & ((1u << (ux & 31u)) - 1u));
            const uint32_t nv = a.nverts[cb];
            if (nv == 0) continue;
            const uint32_t off = a.vert_off[cb];
            if ((uint64_t)off + nv > a.capacity) continue;  // the forward did not emit it
            const uint64_t u = ((uint64_t)uz * ry + uy) * rx + ux;
            float val[8];
            uint32_t f = 0;
#pragma unroll
            for (uint32_t c = 0; c < 8; ++c) {
                val[c] = a.phi[u + corner_dx(c) + corner_dy(c) * rx + corner_dz(c) * rxy];
                f |= (val[c] <= a.iso) ? (1u << c) : 0u;
            }
            const uint64_t tri = cAdjTri[f];
            for (uint32_t k = 0; k < nv; ++k) {
                const uint32_t e = (uint32_t)(tri >> (4 * k)) & 0xFu;
                const uint32_t s0 = (e < 8) ? (e & 3u) + (e & 4u) : e - 8u;
                const uint32_t s1 = (e < 8) ? (((e & 3u) + 1u) & 3u) + (e & 4u) : e - 4u;
                if (s0 != cv && s1 != cv) continue;
                const float v0 = val[s0], v1 = val[s1];
                const float delta = v1 - v0;
                const bool flat = a.cpu_flavor
                                      ? ((double)delta > -1.0e-7 && (double)delta < 1.0e-7)
                                      : (delta > -1.0e-7f && delta < 1.0e-7f);
                if (flat) continue;
                // d(position)/dt: corner s1 minus corner s0 as the forward evaluates them
                const uint32_t X0 = ux + corner_dx(s0), Y0 = uy + corner_dy(s0), Z0 = uz + corner_dz(s0);
                const uint32_t X1 = ux + corner_dx(s1), Y1 = uy + corner_dy(s1), Z1 = uz + corner_dz(s1);
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
                acc += (s0 == cv) ? gt * (a.iso - v1) * inv_d2 : -gt * (a.iso - v0) * inv_d2;
            }
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

struct GradSums {
    float g1x, g1y, g1z, g2x, g2y, g2z, sk;
};

// n staged vertices (a multiple of 8) against this lane's particle.
__device__ __forceinline__ void gather_pairs(const float4* __restrict__ qx,
                                             const float4* __restrict__ qu, uint32_t n,
                                             float px, float py, float pz, float ax, float ay,
                                             float az, float c0, GradSums& g) {
    for (uint32_t t0 = 0; t0 < n; t0 += 8u) {
#pragma unroll
        for (int t = 0; t < 8; ++t) {
            const float4 xv = qx[t0 + t];
            const float4 uv = qu[t0 + t];
            float tt = __builtin_fmaf(az, xv.z, c0 - xv.w);
            tt = __builtin_fmaf(ay, xv.y, tt);
            tt = __builtin_fmaf(ax, xv.x, tt);
            // max(t, 0) as a clamp to [0, 1], folded into the last fma (as the
            // forward kernel clamps t^3): t exceeds 1 only by rounding
            const float tp = __builtin_amdgcn_fmed3f(tt, 0.0f, 1.0f);
            const float t2 = tp * tp;
            const float w = t2 * tp;
            const float kd = __builtin_fmaf(uv.x, px, __builtin_fmaf(uv.y, py, __builtin_fmaf(uv.z, pz, uv.w)));
            const float kk = kd * t2;
            g.g1x = __builtin_fmaf(w, uv.x, g.g1x);
            g.g1y = __builtin_fmaf(w, uv.y, g.g1y);
            g.g1z = __builtin_fmaf(w, uv.z, g.g1z);
            g.g2x = __builtin_fmaf(kk, xv.x, g.g2x);
            g.g2y = __builtin_fmaf(kk, xv.y, g.g2y);
            g.g2z = __builtin_fmaf(kk, xv.z, g.g2z);
            g.sk += kk;
        }
    }
}

// Vertices whose coefficient is zero, or that lie at least 1.00005 R from the
// bounding box of the chunk's particles, add exactly 0 to every lane and are
// dropped at staging (order-preserving compaction, as in the forward kernel).
__global__ __launch_bounds__(kAdjThreads) void k_field_zb05_backward(FieldBackArgs a) {
    __shared__ float4 queue_x[kAdjThreads / 64][128];
    __shared__ float4 queue_u[kAdjThreads / 64][128];
    const uint32_t lane = __lane_id();
    const uint32_t wave = threadIdx.x >> 6;
    const uint32_t nch = *a.num_chunks;
    const float inv_r2 = 1.0f / a.radius_sq;
    const float cull_r2 = a.radius_sq * 1.0001f;
    float4* qx = queue_x[wave];
    float4* qu = queue_u[wave];
    uint32_t step = 0;
    for (;;) {
        const uint32_t ch = next_item_xcd(a.next, nch, step);
        if (ch >= nch) break;
        const uint32_t c = a.chunk_cell[ch];
        const uint32_t k = ch - a.chunk_off[c];
        const fsr_index_range own = a.ranges[c];
        const uint32_t i = own.start + 64u * k + lane;
        const bool live = i < own.end;
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
        // lane j < 9 holds run j's first index (rp) and the offset from index
        // to band slot in that run (ro)
        const uint32_t incl = wave_inclusive_sum(len);
        const uint32_t total = (uint32_t)__builtin_amdgcn_readlane((int)incl, 63);
        const uint32_t rp = incl - len;
        const uint32_t ro = rs - rp;
        // Run of index gg of the staging round that starts at nb: jc, a run
        // whose first index is at most nb, only grows; the runs that start
        // inside the round are visited with wave-uniform reads, and a lane
        // takes the last one that starts at or before gg (an empty run shares
        // its first index with the next run, visited after it).
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
        const float ax = 2.0f * px * inv_r2, ay = 2.0f * py * inv_r2, az = 2.0f * pz * inv_r2;
        const float c0 = 1.0f - __builtin_fmaf(pz, pz, __builtin_fmaf(py, py, px * px)) * inv_r2;
        const float bx0 = wave_min(live ? px : 3.0e38f), bx1 = wave_max(live ? px : -3.0e38f);
        const float by0 = wave_min(live ? py : 3.0e38f), by1 = wave_max(live ? py : -3.0e38f);
        const float bz0 = wave_min(live ? pz : 3.0e38f), bz1 = wave_max(live ? pz : -3.0e38f);

        GradSums g = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
        uint32_t fill = 0;
        // this lane's vertex of the next round is loaded one round ahead
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
            float4 sx, su;
            if (gi < total) {
                if (cf.x != 0.0f || cf.y != 0.0f || cf.z != 0.0f || cf.w != 0.0f) {
                    const float xx = vp.x - ox;
                    const float xy = vp.y - oy;
                    const float xz = vp.z - oz;
                    const float dx = xx - __builtin_amdgcn_fmed3f(xx, bx0, bx1);
                    const float dy = xy - __builtin_amdgcn_fmed3f(xy, by0, by1);
                    const float dz = xz - __builtin_amdgcn_fmed3f(xz, bz0, bz1);
                    keep = __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, dx * dx)) < cull_r2;
                    sx = make_float4(xx, xy, xz,
                                     __builtin_fmaf(xz, xz, __builtin_fmaf(xy, xy, xx * xx)) * inv_r2);
                    // c - u . x' (the pair adds u . p')
                    su = make_float4(cf.x, cf.y, cf.z,
                                     cf.w - __builtin_fmaf(cf.z, xz, __builtin_fmaf(cf.y, xy, cf.x * xx)));
                }
            }
            const uint64_t m = __ballot(keep);
            if (keep) {
                const uint32_t o = fill + mask_rank(m);
                qx[o] = sx;
                qu[o] = su;
            }
            fill += (uint32_t)__popcll(m);
            if (fill >= 64u) {
                wave_lds_sync();
                gather_pairs(qx, qu, 64u, px, py, pz, ax, ay, az, c0, g);
                wave_lds_sync();
                const uint32_t rem = fill - 64u;
                if (lane < rem) {
                    qx[lane] = qx[64u + lane];
                    qu[lane] = qu[64u + lane];
                }
                wave_lds_sync();
                fill = rem;
            }
        }
        if (fill) {
            const uint32_t n = (fill + 7u) & ~7u;
            if (lane >= fill && lane < n) {
                qx[lane] = make_float4(0.f, 0.f, 0.f, 1.0e30f);
                qu[lane] = make_float4(0.f, 0.f, 0.f, 0.f);
            }
            wave_lds_sync();
            gather_pairs(qx, qu, n, px, py, pz, ax, ay, az, c0, g);
            wave_lds_sync();
        }
        if (live) {
            const float s6 = 6.0f * inv_r2;
            const uint64_t o = 3ull * a.perm[i];
            a.g_pos[o + 0] = -__builtin_fmaf(s6, __builtin_fmaf(-g.sk, px, g.g2x), g.g1x);
            a.g_pos[o + 1] = -__builtin_fmaf(s6, __builtin_fmaf(-g.sk, py, g.g2y), g.g1y);
            a.g_pos[o + 2] = -__builtin_fmaf(s6, __builtin_fmaf(-g.sk, pz, g.g2z), g.g1z);
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
    const int blocks = persistent_blocks(ctx, kAdjThreads, 8, TUNE_MC);
    hipLaunchKernelGGL(k_mc_touch, dim3(blocks), dim3(kAdjThreads), 0, s, a);
    FSR_LAUNCH_CHECK("k_mc_touch");
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
    hipLaunchKernelGGL(k_field_zb05_backward, dim3(blocks), dim3(kAdjThreads), 0, s, a);
    FSR_LAUNCH_CHECK("k_field_zb05_backward");
    return FSR_OK;
}

}  // namespace fsr
