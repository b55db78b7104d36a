// Anisotropic kernel of Yu & Turk 2013 (the reference's "Yu13" method, TC01).
//
//   k_smooth_mc     calculationOfMeanAndSmoothParticles (ReconstructionCUDA.cu:598-658):
//                   over the clamped 5^3 cell block, w = 1 - (d/r)^3 for d < r
//                   (wij, ReconstructionCUDA.cuh:282-292), r = anisotropic radius;
//                   mean = sum w p / sum w; smoothed = (1 - lambda) x + lambda mean.
//   k_involved      compactationOfSurfaceVerticesAndParticles, particle half
//                   (ReconstructionCUDA.cu:661-686): particles of cells with a
//                   nonzero involve count, in cell order.
//   k_covariance_mc calculationOfTransformMatricesForParticles (ReconstructionCUDA.cu:689-827):
//                   C = (h^2 I + sum w v v^T) / sum w with v = p_j - mean_i over the
//                   clamped 5^3 block; #{|v| <= r} < Ne gives G = I / (kn h); otherwise
//                   the singular values of C (|sigma|) are clamped to sigma_max / kr and
//                   G = cbrt(s1 s2 s3) / h * R diag(1/s) R^T.
//   k_field_aniso   computationOfScalarFieldGrid (ReconstructionCUDA.cu:830-906):
//                   phi = sum m/rho_j W - 0.5 over the clamped 5^3 block, with
//                   W = sigma det(G) (1 - |G r|^2)^3 for particles of involved cells
//                   (anisotropicW, .cuh:295-310) and W = sigma (1 - |r|^2)^3 with r
//                   in world units for the others (anisotropicWOpt, .cuh:313-322).
//
// SVD: C is symmetric positive definite (it starts at h^2 I), so its singular
// values are its eigenvalues and U = V = R; G depends only on that
// eigen-decomposition.  It is computed with the cyclic Jacobi method for
// symmetric matrices (Golub & Van Loan, Matrix Computations, sec. 8.5: the
// symmetric Schur 2x2 rotation, 6 sweeps) in registers.
//
// MI355X mapping: as for the ZB05 field, a wave owns 64 particles (or band
// vertices) of one hash cell, all lanes share the candidate block, and the 25
// x-runs of the 5^3 block are staged through LDS 64 candidates at a time.
// Smoothing and covariance evaluate their pairs in 16 x 16 tiles on the matrix
// cores (see k_smooth_mc).  The field kernel evaluates |G r|^2 as the quadratic form
// r^T M r with M = G^T G stored per particle (6 values): with v the vertex and
// p the particle relative to the cell corner, |G(p - v)|^2 =
// v^T M v - 2 v.(M p) + p^T M p, where M p and p^T M p are formed once per staged
// candidate; eight of its nine vertex terms are two K = 4 matrix products per
// 16 x 16 tile, the ninth one multiply-add (see aniso_tiles).  Particles of
// non-involved cells carry M = I and the unscaled kernel, as the reference has
// it; the field sums them through per-block moments (noninv_poly).
#include <cmath>
#include <type_traits>
#include <limits>

#include "common.hpp"

namespace fsr {

namespace {

// Workgroups of 4 waves.  Resident workgroups per CU, i.e. waves per SIMD:
// smoothing 5 (96 VGPRs, no spills), covariance 4 (128 VGPRs and 34 KB of LDS
// per workgroup), field 5 (96 VGPRs, a few spilled; the 4-wave build of the
// field, chosen by a field launch limit of at most 4, has none and measured
// the same).
constexpr int kAnisoThreads = 256;
// rows of a 5^3 cell block, one per lane
constexpr int kMaxRows = 25;
static_assert(kMaxRows <= 64, "one block row per lane");
constexpr float kSigma = 1.5666814710608448f;  // 315 / (64 pi)
// Candidate culling of the field kernel: a particle whose support ellipsoid
// (radius 1 / sigma_min(G)) stays 1% (in squared radius) away from the bounding
// box of a chunk's vertices has |G r|^2 >= 1.01 at every vertex of the chunk;
// the quadratic form's FP32 rounding there is below 1e-4, so its weight is
// exactly 0 and dropping it leaves the sums bit-identical.
constexpr float kCullMargin = 1.01f;
constexpr float kNoCull = 3.0e38f;

__device__ __forceinline__ int clampi(int v, int lo, int hi) {
    return v < lo ? lo : (v > hi ? hi : v);
}

__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Rows (x-runs) of the clamped (2R+1)^3 block around cell (cx, cy, cz), lane
// k holding row k: rp = first candidate index of the row, ro = offset from
// candidate index to sorted index.  Candidate indices are staged in rounds of
// 64 in ascending order; offset(nb, gg) gives the offset of the row holding
// index gg of the round that starts at nb.  jc, a row whose first index is at
// most nb, only grows; the rows that start inside the round are visited with
// wave-uniform reads, each lane taking the last one that starts at or before
// its index (an empty row shares its first index with the next row, visited
// after it).
struct BlockRows {
    uint32_t rp, ro, total, jc;
    int nrows;
    __device__ __forceinline__ uint32_t offset(uint32_t nb, uint32_t gg) {
        uint32_t off = (uint32_t)__builtin_amdgcn_readlane((int)ro, (int)jc);
        for (uint32_t jj = jc + 1u; jj < (uint32_t)nrows; ++jj) {
            const uint32_t pj = (uint32_t)__builtin_amdgcn_readlane((int)rp, (int)jj);
            if (pj >= nb + 64u) break;
            if (gg >= pj) off = (uint32_t)__builtin_amdgcn_readlane((int)ro, (int)jj);
            jc = jj;
        }
        return off;
    }
};

// Cell-level culling (cell_lo / cell_hi may be null): a cell counts only when
// the bounding box of its particles (cell_lo.xyz, cell_hi.xyz), grown by their
// largest support radius (cell_lo.w = its square, negative for a cell with
// nothing to stage), reaches the box [b0, b1]; a row's run then spans its first
// to its last such cell (cells between them stay in it, their particles are
// culled one by one).
struct CellCull {
    const float4* lo;
    const float4* hi;
    float b0[3], b1[3];
};

// The test on a cell's box (lo, hi) as loaded.
__device__ __forceinline__ bool cell_reaches(const CellCull& cc, const float4& lo,
                                             const float4& hi) {
    const float dx = fmaxf(fmaxf(lo.x - cc.b1[0], cc.b0[0] - hi.x), 0.0f);
    const float dy = fmaxf(fmaxf(lo.y - cc.b1[1], cc.b0[1] - hi.y), 0.0f);
    const float dz = fmaxf(fmaxf(lo.z - cc.b1[2], cc.b0[2] - hi.z), 0.0f);
    return lo.w > 0.0f && __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, dx * dx)) < lo.w;
}

template <int R>
__device__ __forceinline__ BlockRows block_rows(const fsr_index_range* __restrict__ ranges,
                                                const GridDev& sp, int cx, int cy, int cz,
                                                const CellCull* cull = nullptr) {
    const uint32_t lane = __lane_id();
    const int x0 = clampi(cx - R, 0, sp.rx - 1), x1 = clampi(cx + R, 0, sp.rx - 1);
    const int y0 = clampi(cy - R, 0, sp.ry - 1), y1 = clampi(cy + R, 0, sp.ry - 1);
    const int z0 = clampi(cz - R, 0, sp.rz - 1), z1 = clampi(cz + R, 0, sp.rz - 1);
    const int ny = y1 - y0 + 1;
    const int nrows = ny * (z1 - z0 + 1);
    uint32_t len = 0, rs = 0;
    if ((int)lane < nrows) {
        const int yy = y0 + (int)lane % ny, zz = z0 + (int)lane / ny;
        const uint32_t rowc = ((uint32_t)zz * sp.ry + yy) * sp.rx;
        // the row's 2R + 1 cells as clamped offsets (a repeated edge cell
        // leaves the run unchanged), their loads issued together
        constexpr int kW = 2 * R + 1;
        uint32_t cell[kW];
        fsr_index_range rg[kW];
        bool in[kW];
#pragma unroll
        for (int d = 0; d < kW; ++d) {
            cell[d] = rowc + (uint32_t)clampi(cx - R + d, x0, x1);
            rg[d] = ranges[cell[d]];
        }
        if (cull) {
            // every cell has a box entry (empty cells w = -1)
            float4 lo[kW], hi[kW];
#pragma unroll
            for (int d = 0; d < kW; ++d) {
                lo[d] = cull->lo[cell[d]];
                hi[d] = cull->hi[cell[d]];
            }
#pragma unroll
            for (int d = 0; d < kW; ++d)
                in[d] = rg[d].start != 0xFFFFFFFFu && cell_reaches(*cull, lo[d], hi[d]);
        } else {
#pragma unroll
            for (int d = 0; d < kW; ++d) in[d] = rg[d].start != 0xFFFFFFFFu;
        }
        uint32_t s = 0xFFFFFFFFu, e = 0;
#pragma unroll
        for (int d = 0; d < kW; ++d) {
            if (in[d]) {
                if (s == 0xFFFFFFFFu) s = rg[d].start;
                e = rg[d].end;
            }
        }
        if (s != 0xFFFFFFFFu) { rs = s; len = e - s; }
    }
    const uint32_t incl = wave_inclusive_sum(len);
    BlockRows b;
    b.total = (uint32_t)__builtin_amdgcn_readlane((int)incl, 63);
    b.rp = incl - len;
    b.ro = rs - b.rp;
    b.jc = 0;
    b.nrows = nrows;
    return b;
}

struct ChunkList {
    const uint32_t* chunk_cell;  // chunk -> cell
    const uint32_t* chunk_off;   // cell -> first chunk
    const uint32_t* num_chunks;  // device count
    uint32_t* next;              // per-XCD fetch counters (next_item_xcd)
};

// Work records of the neighbour kernels: one section per kernel (smoothing,
// covariance, field) of 1 + 2 W words, W = kWorkWavesPerCu waves per CU: the
// launch's wave count, then per wave (candidates staged, candidates kept by
// the chunk-box test).  The counts are wave-uniform.
constexpr uint32_t kWorkWavesPerCu = 32;
enum WorkSection { WORK_SMOOTH = 0, WORK_COV = 1, WORK_FIELD = 2, WORK_SECTIONS = 3 };

__device__ __forceinline__ void write_work(uint32_t* __restrict__ sec, uint32_t staged,
                                           uint32_t kept) {
    const uint32_t gw = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    if (__lane_id() == 0) {
        sec[1 + 2 * gw] = staged;
        sec[2 + 2 * gw] = kept;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) sec[0] = gridDim.x * (blockDim.x >> 6);
}

// ---------------------------------------------------------------- K10 -------
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

// Squared distance from (x, y, z) to the box [b0, b1].
__device__ __forceinline__ float box_dist2(float x, float y, float z, const float* b0,
                                           const float* b1) {
    // x - clamp(x, lo, hi): its square equals that of max(lo - x, x - hi, 0)
    const float dx = x - __builtin_amdgcn_fmed3f(x, b0[0], b1[0]);
    const float dy = y - __builtin_amdgcn_fmed3f(y, b0[1], b1[1]);
    const float dz = z - __builtin_amdgcn_fmed3f(z, b0[2], b1[2]);
    return __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, dx * dx));
}

// Bounding box of the live lanes' points.
__device__ __forceinline__ void wave_box(bool live, float x, float y, float z, float* b0,
                                         float* b1) {
    b0[0] = wave_minf(live ? x : 3.0e38f); b1[0] = wave_maxf(live ? x : -3.0e38f);
    b0[1] = wave_minf(live ? y : 3.0e38f); b1[1] = wave_maxf(live ? y : -3.0e38f);
    b0[2] = wave_minf(live ? z : 3.0e38f); b1[2] = wave_maxf(live ? z : -3.0e38f);
}

// Candidates farther than r * (1 + 5e-4) from the bounding box of a chunk's
// points have weight 0 (and lie outside the neighbour radius) at every lane:
// they are dropped at staging, the rest compacted in order into the LDS queue.
constexpr float kRadiusCull = 1.001f;


// ---------------------------------------------------------------- K8 --------
// Involved list and the per-particle field records of non-involved particles.
__global__ __launch_bounds__(256) void k_involved(
    const uint32_t* __restrict__ keys, const fsr_index_range* __restrict__ ranges,
    const uint32_t* __restrict__ involve, const uint32_t* __restrict__ involve_off,
    const float4* __restrict__ sorted, uint32_t n, uint32_t C, float mass,
    uint32_t* __restrict__ involved_idx, float4* __restrict__ rec) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t c = keys[i];
    if (c < C && involve[c] > 0) {
        involved_idx[involve_off[c] + (i - ranges[c].start)] = i;
    } else {
        // M = I, unscaled kernel (anisotropicWOpt): its support (|r| < 1 in world
        // units) covers the whole candidate block, so it is never culled
        rec[2ull * i + 0] = make_float4(1.0f, 1.0f, 1.0f, 0.0f);
        rec[2ull * i + 1] = make_float4(0.0f, 0.0f, kSigma * (mass / sorted[i].w), kNoCull);
    }
}

// ------------------------------------------------------ 3x3 Jacobi eigen -----
// Symmetric A (a00 a01 a02 / a11 a12 / a22) -> eigenvalues d and rotation V
// (columns = eigenvectors), A = V diag(d) V^T.
__device__ __forceinline__ void jacobi_rotate(float& app, float& aqq, float& apq, float& arp,
                                              float& arq, float* vp, float* vq) {
    if (fabsf(apq) <= 1.0e-30f) return;
    const float theta = (aqq - app) / (2.0f * apq);
    const float t = copysignf(1.0f, theta) / (fabsf(theta) + __builtin_sqrtf(__builtin_fmaf(theta, theta, 1.0f)));
    const float c = 1.0f / __builtin_sqrtf(__builtin_fmaf(t, t, 1.0f));
    const float s = t * c;
    app = __builtin_fmaf(-t, apq, app);
    aqq = __builtin_fmaf(t, apq, aqq);
    apq = 0.0f;
    const float rp = arp, rq = arq;
    arp = c * rp - s * rq;
    arq = s * rp + c * rq;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const float a = vp[k], b = vq[k];
        vp[k] = c * a - s * b;
        vq[k] = s * a + c * b;
    }
}

__device__ __forceinline__ void sym_eigen3(float a00, float a01, float a02, float a11, float a12,
                                           float a22, float d[3], float V[3][3]) {
    // V stored by columns: V[k] is eigenvector k
    float v0[3] = {1.f, 0.f, 0.f}, v1[3] = {0.f, 1.f, 0.f}, v2[3] = {0.f, 0.f, 1.f};
#pragma unroll 1
    for (int sweep = 0; sweep < 6; ++sweep) {
        jacobi_rotate(a00, a11, a01, a02, a12, v0, v1);   // (p,q) = (0,1), r = 2
        jacobi_rotate(a00, a22, a02, a01, a12, v0, v2);   // (0,2), r = 1
        jacobi_rotate(a11, a22, a12, a01, a02, v1, v2);   // (1,2), r = 0
    }
    d[0] = a00; d[1] = a11; d[2] = a22;
#pragma unroll
    for (int k = 0; k < 3; ++k) { V[0][k] = v0[k]; V[1][k] = v1[k]; V[2][k] = v2[k]; }
}

// ---------------------------------------------------------------- K11 -------
struct CovArgs {
    const float4* sorted;
    const float4* mean;
    const fsr_index_range* ranges;
    const uint32_t* involve_off;
    GridDev sp;
    float radius;           // anisotropic radius r
    // the reference's tests on the correctly rounded distance, as exact tests
    // on d^2: sqrt(d2) <= r  <=>  d2 <= d2_le_r;  sqrt(d2) < r  <=>  d2 <= d2_lt_r
    float d2_le_r, d2_lt_r;
    float h2;               // smoothing radius squared
    float inv_h;            // 1 / h
    float kr, kn;
    uint32_t min_neighbors;
    float mass;
    fsr_matrix* gmat;       // [I]
    float4* sigma;          // [I] eigenvalues, descending
    float4* rec;            // [2N] field records
    uint32_t* work;         // work records (may be null)
    const float4* cell_box; // per cell lo[C], hi[C] (k_cell_box)
};

// G and the field record of involved particle i (its involved index gi) from
// its neighbour count and weighted sums: C = (h^2 I + sum w v v^T) / sum w.
__device__ __forceinline__ void cov_finish(const CovArgs& a, uint32_t gi, uint32_t i,
                                           uint32_t cnt, float sw, float c00, float c01,
                                           float c02, float c11, float c12, float c22) {
    float G[3][3];
    float4 sig;
    float gmin;  // smallest singular value of G: the support is |r| < 1 / gmin
    if (cnt < a.min_neighbors) {
        const float gv = a.inv_h / a.kn;
        for (int p = 0; p < 3; ++p)
            for (int q = 0; q < 3; ++q) G[p][q] = p == q ? gv : 0.0f;
        sig = make_float4(0.f, 0.f, 0.f, 0.f);
        gmin = gv;
    } else {
        const float iw = 1.0f / sw;
        const float a00 = (a.h2 + c00) * iw, a11 = (a.h2 + c11) * iw, a22 = (a.h2 + c22) * iw;
        const float a01 = c01 * iw, a02 = c02 * iw, a12 = c12 * iw;
        float d[3], V[3][3];
        sym_eigen3(a00, a01, a02, a11, a12, a22, d, V);
        float s0 = fabsf(d[0]), s1 = fabsf(d[1]), s2 = fabsf(d[2]);
        // report sorted descending
        float e0 = fmaxf(s0, fmaxf(s1, s2)), e2 = fminf(s0, fminf(s1, s2));
        sig = make_float4(e0, s0 + s1 + s2 - e0 - e2, e2, 0.f);
        const float floor_v = e0 / a.kr;
        s0 = fmaxf(s0, floor_v);
        s1 = fmaxf(s1, floor_v);
        s2 = fmaxf(s2, floor_v);
        const float scale = cbrtf(s0 * s1 * s2) * a.inv_h;
        const float g0 = scale / s0, g1 = scale / s1, g2 = scale / s2;
        gmin = fminf(g0, fminf(g1, g2));
        for (int p = 0; p < 3; ++p)
            for (int q = 0; q < 3; ++q)
                G[p][q] = V[0][p] * g0 * V[0][q] + V[1][p] * g1 * V[1][q] + V[2][p] * g2 * V[2][q];
    }
    fsr_matrix out;
    out.a11 = G[0][0]; out.a12 = G[0][1]; out.a13 = G[0][2];
    out.a21 = G[1][0]; out.a22 = G[1][1]; out.a23 = G[1][2];
    out.a31 = G[2][0]; out.a32 = G[2][1]; out.a33 = G[2][2];
    a.gmat[gi] = out;
    a.sigma[gi] = sig;
    // determinant (ReconstructionCUDA.cuh:690-700) and M = G^T G
    const float det = G[0][0] * G[1][1] * G[2][2] + G[0][1] * G[1][2] * G[2][0] +
                      G[0][2] * G[1][0] * G[2][1] - G[0][0] * G[1][2] * G[2][1] -
                      G[0][1] * G[1][0] * G[2][2] - G[0][2] * G[1][1] * G[2][0];
    float M[3][3];
    for (int p = 0; p < 3; ++p)
        for (int q = 0; q < 3; ++q)
            M[p][q] = G[0][p] * G[0][q] + G[1][p] * G[1][q] + G[2][p] * G[2][q];
    const float coef = kSigma * det * (a.mass / a.sorted[i].w);
    a.rec[2ull * i + 0] = make_float4(M[0][0], M[1][1], M[2][2], M[0][1]);
    // squared culling radius of the field kernel, with a 1% margin
    a.rec[2ull * i + 1] = make_float4(M[0][2], M[1][2], coef, kCullMargin / (gmin * gmin));
}


// ------------------------------------------- K10 / K11 on the matrix cores ---
// Smoothing and covariance as pair tiles on the matrix cores, as the ZB05
// field does: a wave owns a chunk of up to 64 particles of one hash cell (the
// "lanes"), candidates of the clamped 5^3 block that pass the chunk-box test
// are staged in LDS, and per 16 candidates and 16 lanes
//   1. one v_mfma_f32_16x16x4_f32 gives s2 = |q - x|^2 / r^2 = q''.(-2 x''/r^2)
//      + |x''|^2/r^2 + |q''|^2/r^2 for the 16 x 16 pairs, with q'' and x''
//      relative to the chunk's origin (the centre of its box): A = (x'', y'',
//      z'', 1) of the candidates, B = (-2x''/r^2, -2y''/r^2, -2z''/r^2,
//      |x''|^2/r^2) of the lanes, C = |q''|^2/r^2;
//   2. w = max(1 - s2^(3/2), 0) (the reference's 1 - (d/r)^3 for d < r, 0
//      beyond), and for the covariance the neighbour count s2 <= le;
//   3. rank-1 updates of 4 x 4 blocks (v_mfma_f32_4x4x1_16b_f32) accumulate
//      sum w (q'', 1) and, for the covariance, sum w (q''q''^T) in two more
//      groups (xx, xy, xz, yy) and (yz, zz, 0, 0).
// x is the particle (smoothing) or its weighted mean (covariance).  The
// covariance is then sum w (q - m)(q - m)^T = S2 - S1 m''^T - m'' S1^T +
// S0 m'' m''^T with m'' the mean relative to the chunk's origin (small: the
// box centre lies among the chunk's means).
typedef float v4f __attribute__((ext_vector_type(4)));

constexpr int kNbQueue = 128;
constexpr int kNbStride = 144;
// queue rows: x'', y'', z'', 1, then (covariance) xx, xy, xz, yy, yz, zz, 0, 0,
// then |q''|^2 / r^2
constexpr int kNbOne = 3, kNbXX = 4, kNbYZ = 8, kNbZero = 10;
template <bool kCov>
struct NbRows {
    static constexpr int kQ = kCov ? 12 : 4;
    static constexpr int kRows = kQ + 1;
};

template <int M>
__device__ __forceinline__ float nb_swz_xor(float v) {
    return __int_as_float(__builtin_amdgcn_ds_swizzle(__float_as_int(v), 0x1F | (M << 10)));
}

template <int M>
__device__ __forceinline__ uint32_t nb_swz_xor_u(uint32_t v) {
    return (uint32_t)__builtin_amdgcn_ds_swizzle((int)v, 0x1F | (M << 10));
}

// Sum over the four partial sums of a lane-particle (lanes l%16 + 16m).
__device__ __forceinline__ float nb_reduce4(float x) {
    x += nb_swz_xor<16>(x);
    return x + __shfl_xor(x, 32, 64);
}

__device__ __forceinline__ uint32_t nb_reduce4_u(uint32_t x) {
    x += nb_swz_xor_u<16>(x);
    return x + (uint32_t)__shfl_xor((int)x, 32, 64);
}

// The tiles of n staged candidates (a multiple of 16) against the NT lane
// tiles.  rows: the queue at the first staged entry.
template <bool kCov, int NT>
__device__ __forceinline__ void nb_tiles(const float* __restrict__ rows, uint32_t n,
                                         const float (&b1)[NT], v4f (&acc)[kCov ? 3 : 1][NT],
                                         uint32_t (&cnt)[NT], float le_s2) {
    const uint32_t lane = __lane_id();
    const float* a1p = rows + (lane >> 4) * kNbStride + (lane & 15u);
    const float* a2p = rows + (lane & 3u) * kNbStride + 4u * (lane >> 4);
    const float* cqp = rows + NbRows<kCov>::kQ * kNbStride + 4u * (lane >> 4);
    for (uint32_t tb = 0; tb < n; tb += 16u) {
        const float a1 = a1p[tb];
        const float4 cq = *reinterpret_cast<const float4*>(cqp + tb);
        const v4f c1 = v4f{cq.x, cq.y, cq.z, cq.w};
        v4f w[NT];
#pragma unroll
        for (int T = 0; T < NT; ++T) w[T] = __builtin_amdgcn_mfma_f32_16x16x4f32(a1, b1[T], c1, 0, 0, 0);
#pragma unroll
        for (int T = 0; T < NT; ++T) {
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                // s2 may lie a rounding error (|s2| < 1e-6) below 0 for a
                // candidate at the particle itself: it counts (s2 <= le_s2), and
                // 1 + |s2|^1.5 clamps to 1 exactly as s2 = 0 would.  (No max
                // with 0: on a matrix result it costs a canonicalisation too;
                // the absolute value is a free source modifier of the square
                // root, and the clamp one of the multiply-add.)
                const float s2 = w[T][r];
                if (kCov) cnt[T] += s2 <= le_s2 ? 1u : 0u;
                w[T][r] = __builtin_amdgcn_fmed3f(
                    __builtin_fmaf(-s2, __builtin_amdgcn_sqrtf(fabsf(s2)), 1.0f), 0.0f, 1.0f);
            }
        }
        const float4 a2 = *reinterpret_cast<const float4*>(a2p + tb);
#pragma unroll
        for (int T = 0; T < NT; ++T) acc[0][T] = __builtin_amdgcn_mfma_f32_4x4x1f32(a2.x, w[T][0], acc[0][T], 0, 0, 0);
#pragma unroll
        for (int T = 0; T < NT; ++T) acc[0][T] = __builtin_amdgcn_mfma_f32_4x4x1f32(a2.y, w[T][1], acc[0][T], 0, 0, 0);
#pragma unroll
        for (int T = 0; T < NT; ++T) acc[0][T] = __builtin_amdgcn_mfma_f32_4x4x1f32(a2.z, w[T][2], acc[0][T], 0, 0, 0);
#pragma unroll
        for (int T = 0; T < NT; ++T) acc[0][T] = __builtin_amdgcn_mfma_f32_4x4x1f32(a2.w, w[T][3], acc[0][T], 0, 0, 0);
        if (kCov) {
#pragma unroll
            for (int g = 1; g < 3; ++g) {
                const float4 ag = *reinterpret_cast<const float4*>(a2p + (4 * g) * kNbStride + tb);
#pragma unroll
                for (int T = 0; T < NT; ++T) acc[g][T] = __builtin_amdgcn_mfma_f32_4x4x1f32(ag.x, w[T][0], acc[g][T], 0, 0, 0);
#pragma unroll
                for (int T = 0; T < NT; ++T) acc[g][T] = __builtin_amdgcn_mfma_f32_4x4x1f32(ag.y, w[T][1], acc[g][T], 0, 0, 0);
#pragma unroll
                for (int T = 0; T < NT; ++T) acc[g][T] = __builtin_amdgcn_mfma_f32_4x4x1f32(ag.z, w[T][2], acc[g][T], 0, 0, 0);
#pragma unroll
                for (int T = 0; T < NT; ++T) acc[g][T] = __builtin_amdgcn_mfma_f32_4x4x1f32(ag.w, w[T][3], acc[g][T], 0, 0, 0);
            }
        }
    }
}

// Sums of one chunk over the candidates of its 5^3 block: staging (box test,
// ordered compaction into the LDS ring), the tiles, the reduction.  vt: the
// lanes' B operands (4 rows of 64).  On return lane l holds, for its own
// particle, s1 = (sum w q''x, y, z, sum w), s2a = (xx, xy, xz, yy), s2b = (yz,
// zz, -, -) and the count (covariance).
template <bool kCov, int NT>
__device__ __forceinline__ void nb_chunk(float* __restrict__ q, const float* __restrict__ vt,
                                         const float4* __restrict__ sorted, BlockRows& br,
                                         uint32_t total, float ox, float oy, float oz,
                                         const float* bb0, const float* bb1, float inv_r2,
                                         float cull2, float le_s2, uint32_t& n_kept, v4f& s1,
                                         v4f& s2a, v4f& s2b, uint32_t& count) {
    const uint32_t lane = __lane_id();
    constexpr int G = kCov ? 3 : 1;
    float b1[NT];
    v4f acc[G][NT];
    uint32_t cnt[NT];
    const float* bsrc = vt + 64 * (lane >> 4) + (lane & 15u);
#pragma unroll
    for (int T = 0; T < NT; ++T) {
        b1[T] = bsrc[16 * T];
        cnt[T] = 0;
#pragma unroll
        for (int g = 0; g < G; ++g) acc[g][T] = v4f{0.f, 0.f, 0.f, 0.f};
    }
    uint32_t head = 0, fill = 0;
    float4 next = make_float4(0.f, 0.f, 0.f, 0.f);
    {
        const uint32_t off = br.offset(0u, lane);
        if (lane < total) next = sorted[off + lane];
    }
    for (uint32_t base = 0; base < total; base += 64u) {
        const uint32_t g = base + lane;
        const float4 cand = next;
        if (base + 64u < total) {
            const uint32_t off = br.offset(base + 64u, g + 64u);
            if (g + 64u < total) next = sorted[off + g + 64u];
        }
        bool keep = false;
        float px = 0.f, py = 0.f, pz = 0.f;
        if (g < total) {
            px = cand.x - ox;
            py = cand.y - oy;
            pz = cand.z - oz;
            keep = box_dist2(px, py, pz, bb0, bb1) < cull2;
        }
        const uint64_t m = __ballot(keep);
        if (keep) {
            const uint32_t slot = (head + fill + mask_rank(m)) & (kNbQueue - 1);
            q[slot] = px;
            q[kNbStride + slot] = py;
            q[2 * kNbStride + slot] = pz;
            q[NbRows<kCov>::kQ * kNbStride + slot] =
                __builtin_fmaf(pz, pz, __builtin_fmaf(py, py, px * px)) * inv_r2;
            if (kCov) {
                q[(kNbXX + 0) * kNbStride + slot] = px * px;
                q[(kNbXX + 1) * kNbStride + slot] = px * py;
                q[(kNbXX + 2) * kNbStride + slot] = px * pz;
                q[(kNbXX + 3) * kNbStride + slot] = py * py;
                q[(kNbYZ + 0) * kNbStride + slot] = py * pz;
                q[(kNbYZ + 1) * kNbStride + slot] = pz * pz;
            }
        }
        fill += (uint32_t)__popcll(m);
        n_kept += (uint32_t)__popcll(m);
        if (fill >= 64u) {
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            nb_tiles<kCov, NT>(q + head, 64u, b1, acc, cnt, le_s2);
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            head ^= 64u;
            fill -= 64u;
        }
    }
    if (fill) {
        // pad to a multiple of 16 with entries of weight 0 and no count
        // (s2 = 1e30); head + 64 <= kNbQueue, so the pad does not wrap
        const uint32_t n = (fill + 15u) & ~15u;
        if (lane >= fill && lane < n) {
            q[head + lane] = 0.f;
            q[kNbStride + head + lane] = 0.f;
            q[2 * kNbStride + head + lane] = 0.f;
            q[NbRows<kCov>::kQ * kNbStride + head + lane] = 1.0e30f;
            if (kCov) {
#pragma unroll
                for (int k = 0; k < 6; ++k) q[(kNbXX + k) * kNbStride + head + lane] = 0.f;
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        nb_tiles<kCov, NT>(q + head, n, b1, acc, cnt, le_s2);
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
    // the four partial sums of each lane-particle; lane l keeps those of
    // particle 16 T + l % 16 with T = l / 16, i.e. particle l
    const uint32_t my = lane >> 4;
    s1 = v4f{0.f, 0.f, 0.f, 0.f};
    s2a = s1;
    s2b = s1;
    count = 0;
#pragma unroll
    for (int T = 0; T < NT; ++T) {
        v4f r0 = acc[0][T];
#pragma unroll
        for (int c = 0; c < 4; ++c) r0[c] = nb_reduce4(r0[c]);
        if (my == (uint32_t)T) s1 = r0;
        if (kCov) {
            v4f r1 = acc[1][T], r2 = acc[2][T];
#pragma unroll
            for (int c = 0; c < 4; ++c) r1[c] = nb_reduce4(r1[c]);
            r2[0] = nb_reduce4(r2[0]);
            r2[1] = nb_reduce4(r2[1]);
            const uint32_t ct = nb_reduce4_u(cnt[T]);
            if (my == (uint32_t)T) {
                s2a = r1;
                s2b = r2;
                count = ct;
            }
        }
    }
}

// The chunk's origin (centre of the box of the lanes' points, absolute) and
// the box relative to it; the lanes' B operands into vt.
__device__ __forceinline__ void nb_setup(bool live, float x, float y, float z, float inv_r2,
                                         float* vt, float* o, float* bb0, float* bb1,
                                         float* rel) {
    float b0[3], b1[3];
    wave_box(live, x, y, z, b0, b1);
    // a chunk without live lanes does not occur (chunks list live particles)
#pragma unroll
    for (int d = 0; d < 3; ++d) o[d] = 0.5f * (b0[d] + b1[d]);
    const float px = x - o[0], py = y - o[1], pz = z - o[2];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        bb0[d] = b0[d] - o[d];
        bb1[d] = b1[d] - o[d];
    }
    const uint32_t lane = __lane_id();
    vt[lane] = -2.0f * px * inv_r2;
    vt[64 + lane] = -2.0f * py * inv_r2;
    vt[128 + lane] = -2.0f * pz * inv_r2;
    vt[192 + lane] = __builtin_fmaf(pz, pz, __builtin_fmaf(py, py, px * px)) * inv_r2;
    rel[0] = px;
    rel[1] = py;
    rel[2] = pz;
}

// Per hash cell, the box of its particles (sorted positions) for the cell test
// of the smoothing and covariance kernels (CellCull): widened by 1e-6 of the
// largest coordinate (the test runs in absolute coordinates, the particle
// test relative to the chunk's centre), w = r^2 kRadiusCull (1.002) + 1e-12
// for an occupied cell, -1 for an empty one.  One wave per cell.
__global__ __launch_bounds__(64) void k_cell_box(const fsr_index_range* __restrict__ ranges,
                                                 const float4* __restrict__ sorted, float w,
                                                 float4* __restrict__ lo,
                                                 float4* __restrict__ hi) {
    const uint32_t c = blockIdx.x;
    const uint32_t lane = threadIdx.x;
    const fsr_index_range r = ranges[c];
    float4 l = make_float4(3.0e38f, 3.0e38f, 3.0e38f, -1.0f);
    float4 u = make_float4(-3.0e38f, -3.0e38f, -3.0e38f, 0.0f);
    if (r.start != 0xFFFFFFFFu) {
        for (uint32_t j = r.start + lane; j < r.end; j += 64u) {
            const float4 p = sorted[j];
            l.x = fminf(l.x, p.x); l.y = fminf(l.y, p.y); l.z = fminf(l.z, p.z);
            u.x = fmaxf(u.x, p.x); u.y = fmaxf(u.y, p.y); u.z = fmaxf(u.z, p.z);
        }
    }
    l.x = wave_minf(l.x); l.y = wave_minf(l.y); l.z = wave_minf(l.z);
    u.x = wave_maxf(u.x); u.y = wave_maxf(u.y); u.z = wave_maxf(u.z);
    if (lane == 0) {
        if (r.start != 0xFFFFFFFFu && r.end > r.start) {
            l.w = w;
            const float e = 1.0e-6f * fmaxf(fmaxf(fmaxf(fabsf(l.x), fabsf(u.x)),
                                                  fmaxf(fabsf(l.y), fabsf(u.y))),
                                            fmaxf(fabsf(l.z), fabsf(u.z)));
            l.x -= e; l.y -= e; l.z -= e;
            u.x += e; u.y += e; u.z += e;
        }
        lo[c] = l;
        hi[c] = u;
    }
}

// The cell test of a chunk: the cells of its block whose particle box lies
// within the radius of the chunk's box [o + bb0, o + bb1].
__device__ __forceinline__ CellCull chunk_cull(const float4* box, uint32_t C, const float* o,
                                               const float* bb0, const float* bb1) {
    CellCull cc;
    cc.lo = box;
    cc.hi = box + C;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        cc.b0[d] = bb0[d] + o[d];
        cc.b1[d] = bb1[d] + o[d];
    }
    return cc;
}

template <bool kCov>
struct NbLds {
    float queue[NbRows<kCov>::kRows * kNbStride];
    float vtx[4][64];
};

__global__ __launch_bounds__(kAnisoThreads, 5) void k_smooth_mc(
    const float4* __restrict__ sorted, const fsr_index_range* __restrict__ ranges, GridDev sp,
    float radius, float lambda, ChunkList cl, const float4* __restrict__ cell_box,
    float4* __restrict__ mean_out, float4* __restrict__ smooth_out, uint32_t* __restrict__ work) {
    __shared__ __align__(16) NbLds<false> lds[kAnisoThreads / 64];
    const uint32_t lane = __lane_id();
    const uint32_t wave = threadIdx.x >> 6;
    float* q = lds[wave].queue;
    float* vt = &lds[wave].vtx[0][0];
    q[kNbOne * kNbStride + lane] = 1.0f;
    q[kNbOne * kNbStride + 64 + lane] = 1.0f;
    const uint32_t nch = *cl.num_chunks;
    const float inv_r2 = 1.0f / (radius * radius);
    uint32_t step = 0;
    uint32_t n_staged = 0, n_kept = 0;
    for (;;) {
        const uint32_t ch = next_item_xcd(cl.next, nch, step);
        if (ch >= nch) break;
        const uint32_t c = cl.chunk_cell[ch];
        const uint32_t k = ch - cl.chunk_off[c];
        const fsr_index_range own = ranges[c];
        const uint32_t i = own.start + 64u * k + lane;
        const bool live = i < own.end;
        const int cx = (int)(c % sp.rx), cy = (int)((c / sp.rx) % sp.ry), cz = (int)(c / (sp.rx * sp.ry));
        float4 self = make_float4(0.f, 0.f, 0.f, 0.f);
        if (live) self = sorted[i];
        float o[3], bb0[3], bb1[3], rel[3];
        nb_setup(live, self.x, self.y, self.z, inv_r2, vt, o, bb0, bb1, rel);
        // the block's rows, without the cells that cannot reach the chunk
        const CellCull cc = chunk_cull(cell_box, sp.rx * sp.ry * sp.rz, o, bb0, bb1);
        BlockRows br = block_rows<2>(ranges, sp, cx, cy, cz, &cc);
        const uint32_t total = br.total;
        n_staged += total;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        v4f s1, s2a, s2b;
        uint32_t count;
        const uint32_t cnt = min(64u, own.end - (own.start + 64u * k));
        const float c2 = radius * radius * kRadiusCull;
        switch ((cnt + 15u) >> 4) {
            case 1: nb_chunk<false, 1>(q, vt, sorted, br, total, o[0], o[1], o[2], bb0, bb1, inv_r2, c2, 0.f, n_kept, s1, s2a, s2b, count); break;
            case 2: nb_chunk<false, 2>(q, vt, sorted, br, total, o[0], o[1], o[2], bb0, bb1, inv_r2, c2, 0.f, n_kept, s1, s2a, s2b, count); break;
            case 3: nb_chunk<false, 3>(q, vt, sorted, br, total, o[0], o[1], o[2], bb0, bb1, inv_r2, c2, 0.f, n_kept, s1, s2a, s2b, count); break;
            default: nb_chunk<false, 4>(q, vt, sorted, br, total, o[0], o[1], o[2], bb0, bb1, inv_r2, c2, 0.f, n_kept, s1, s2a, s2b, count); break;
        }
        if (live && s1[3] > 0.0f) {
            const float inv = 1.0f / s1[3];
            const float mx = o[0] + s1[0] * inv, my = o[1] + s1[1] * inv, mz = o[2] + s1[2] * inv;
            mean_out[i] = make_float4(mx, my, mz, 0.0f);
            const float a = 1.0f - lambda;
            smooth_out[i] = make_float4(a * self.x + lambda * mx, a * self.y + lambda * my,
                                        a * self.z + lambda * mz, 0.0f);
        }
        (void)rel;
    }
    if (work) write_work(work, n_staged, n_kept);
}

__global__ __launch_bounds__(kAnisoThreads, 4) void k_covariance_mc(CovArgs a, ChunkList cl) {
    __shared__ __align__(16) NbLds<true> lds[kAnisoThreads / 64];
    const uint32_t lane = __lane_id();
    const uint32_t wave = threadIdx.x >> 6;
    float* q = lds[wave].queue;
    float* vt = &lds[wave].vtx[0][0];
    q[kNbOne * kNbStride + lane] = 1.0f;
    q[kNbOne * kNbStride + 64 + lane] = 1.0f;
    q[kNbZero * kNbStride + lane] = 0.0f;
    q[kNbZero * kNbStride + 64 + lane] = 0.0f;
    q[(kNbZero + 1) * kNbStride + lane] = 0.0f;
    q[(kNbZero + 1) * kNbStride + 64 + lane] = 0.0f;
    const uint32_t nch = *cl.num_chunks;
    const float r = a.radius;
    const float inv_r2 = 1.0f / (r * r);
    const float c2 = r * r * kRadiusCull;
    // the neighbour test |q - m| <= r on s2 = |q - m|^2 / r^2
    const float le_s2 = a.d2_le_r * inv_r2;
    uint32_t step = 0;
    uint32_t n_staged = 0, n_kept = 0;
    for (;;) {
        const uint32_t ch = next_item_xcd(cl.next, nch, step);
        if (ch >= nch) break;
        const uint32_t c = cl.chunk_cell[ch];
        const uint32_t k = ch - cl.chunk_off[c];
        const fsr_index_range own = a.ranges[c];
        const uint32_t i = own.start + 64u * k + lane;
        const bool live = i < own.end;
        const int cx = (int)(c % a.sp.rx), cy = (int)((c / a.sp.rx) % a.sp.ry),
                  cz = (int)(c / (a.sp.rx * a.sp.ry));
        float4 m = make_float4(0.f, 0.f, 0.f, 0.f);
        if (live) m = a.mean[i];
        float o[3], bb0[3], bb1[3], rel[3];
        nb_setup(live, m.x, m.y, m.z, inv_r2, vt, o, bb0, bb1, rel);
        // the block's rows, without the cells that cannot reach the chunk
        const CellCull cc = chunk_cull(a.cell_box, a.sp.rx * a.sp.ry * a.sp.rz, o, bb0, bb1);
        BlockRows br = block_rows<2>(a.ranges, a.sp, cx, cy, cz, &cc);
        const uint32_t total = br.total;
        n_staged += total;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        v4f s1, s2a, s2b;
        uint32_t count;
        const uint32_t cnt = min(64u, own.end - (own.start + 64u * k));
        switch ((cnt + 15u) >> 4) {
            case 1: nb_chunk<true, 1>(q, vt, a.sorted, br, total, o[0], o[1], o[2], bb0, bb1, inv_r2, c2, le_s2, n_kept, s1, s2a, s2b, count); break;
            case 2: nb_chunk<true, 2>(q, vt, a.sorted, br, total, o[0], o[1], o[2], bb0, bb1, inv_r2, c2, le_s2, n_kept, s1, s2a, s2b, count); break;
            case 3: nb_chunk<true, 3>(q, vt, a.sorted, br, total, o[0], o[1], o[2], bb0, bb1, inv_r2, c2, le_s2, n_kept, s1, s2a, s2b, count); break;
            default: nb_chunk<true, 4>(q, vt, a.sorted, br, total, o[0], o[1], o[2], bb0, bb1, inv_r2, c2, le_s2, n_kept, s1, s2a, s2b, count); break;
        }
        if (live) {
            // sum w v v^T with v = q'' - m'' from the moments about the origin
            const float mx = rel[0], my = rel[1], mz = rel[2];
            const float sw = s1[3];
            const float sx = s1[0], sy = s1[1], sz = s1[2];
            const float c00 = s2a[0] - 2.0f * sx * mx + sw * mx * mx;
            const float c01 = s2a[1] - sx * my - sy * mx + sw * mx * my;
            const float c02 = s2a[2] - sx * mz - sz * mx + sw * mx * mz;
            const float c11 = s2a[3] - 2.0f * sy * my + sw * my * my;
            const float c12 = s2b[0] - sy * mz - sz * my + sw * my * mz;
            const float c22 = s2b[1] - 2.0f * sz * mz + sw * mz * mz;
            cov_finish(a, a.involve_off[c] + (i - own.start), i, count, sw, c00, c01, c02, c11,
                       c12, c22);
        }
    }
    if (a.work) write_work(a.work, n_staged, n_kept);
}

// ---------------------------------------------------------------- K12 -------
struct AnisoFieldArgs {
    const float4* smooth;
    const float4* rec;
    const fsr_index_range* ranges;
    BandCompactOut bc;
    GridDev sp, sc;
    float* phi;
    uint32_t* next;
    const uint32_t* num_chunks;
    // when set: kNonInvMoments moments per virtual cell of the non-involved
    // particles of its clamped 5^3 block (the field kernel then skips those
    // particles and evaluates their sum as a polynomial), about the origin
    // (ox, oy, oz)
    const double* block_mom;
    double ox, oy, oz;
    // with block_mom: per cell, the box of its involved particles' smoothed
    // positions and their largest squared support radius (see CellCull)
    const float4* cell_lo;
    const float4* cell_hi;
    uint32_t* work;              // work records (may be null)
};

// The non-involved part of the field as a polynomial.  For the particles p of
// non-involved cells the reference adds c (1 - |x - p|^2)^3 with c = sigma m /
// rho in world units; over a block whose every |x - p| < 1 (checked on the
// host) the sum is, with A = 1 - |x|^2, B = 2 x.p - |p|^2 and P = |p|^2,
//   sum c (A + B)^3 = A^3 S0 + 3 A^2 sum cB + 3 A sum cB^2 + sum cB^3,
//   sum cB   = 2 x.M1 - S_P
//   sum cB^2 = 4 x^T M2 x - 4 x.M1P + S_P2
//   sum cB^3 = 8 sum c (x.p)^3 - 12 x^T M2P x + 6 x.M1P2 - S_P3,
// exactly, in terms of 35 weighted moments of the particles (about one origin
// for the whole grid, in FP64):
//   0 S0  1-3 M1  4 S_P  5-10 M2 (xx xy xz yy yz zz)  11-13 M1P  14 S_P2
//   15-24 M3 (xxx xxy xxz xyy xyz xzz yyy yyz yzz zzz)  25-30 M2P  31-33 M1P2
//   34 S_P3
// A block's moments are the sums of its cells' moments.
constexpr int kNonInvMoments = 35;

__device__ __forceinline__ void moment_terms(double x, double y, double z,
                                             double (&t)[kNonInvMoments]) {
    const double P = x * x + y * y + z * z;
    const double xx = x * x, xy = x * y, xz = x * z, yy = y * y, yz = y * z, zz = z * z;
    t[0] = 1.0;
    t[1] = x; t[2] = y; t[3] = z;
    t[4] = P;
    t[5] = xx; t[6] = xy; t[7] = xz; t[8] = yy; t[9] = yz; t[10] = zz;
    t[11] = x * P; t[12] = y * P; t[13] = z * P;
    t[14] = P * P;
    t[15] = xx * x; t[16] = xx * y; t[17] = xx * z; t[18] = x * yy; t[19] = xy * z;
    t[20] = x * zz; t[21] = yy * y; t[22] = yy * z; t[23] = y * zz; t[24] = zz * z;
    t[25] = xx * P; t[26] = xy * P; t[27] = xz * P; t[28] = yy * P; t[29] = yz * P;
    t[30] = zz * P;
    t[31] = x * P * P; t[32] = y * P * P; t[33] = z * P * P;
    t[34] = P * P * P;
}

__device__ __forceinline__ double noninv_poly(const double* __restrict__ m, double x, double y,
                                              double z) {
    const double A = 1.0 - (x * x + y * y + z * z);
    const double xx = x * x, yy = y * y, zz = z * z, xy = x * y, xz = x * z, yz = y * z;
    const double sB = 2.0 * (x * m[1] + y * m[2] + z * m[3]) - m[4];
    const double xM2x = xx * m[5] + yy * m[8] + zz * m[10] + 2.0 * (xy * m[6] + xz * m[7] + yz * m[9]);
    const double sB2 = 4.0 * xM2x - 4.0 * (x * m[11] + y * m[12] + z * m[13]) + m[14];
    const double T3 = x * (xx * m[15] + 3.0 * (xy * m[16] + xz * m[17] + yy * m[18]) +
                           6.0 * yz * m[19] + 3.0 * zz * m[20]) +
                      y * (yy * m[21] + 3.0 * (yz * m[22] + zz * m[23])) + z * zz * m[24];
    const double xM2Px =
        xx * m[25] + yy * m[28] + zz * m[30] + 2.0 * (xy * m[26] + xz * m[27] + yz * m[29]);
    const double sB3 = 8.0 * T3 - 12.0 * xM2Px + 6.0 * (x * m[31] + y * m[32] + z * m[33]) - m[34];
    return A * (A * (A * m[0] + 3.0 * sB) + 3.0 * sB2) + sB3;
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

// LDS queue of one field wave: a ring of kAfQueue staged candidates, stored
// as rows of kAfStride floats (structure of arrays).  Rows 0-3 and 4-7 are the
// A operands of the two matrix products, (M00, M11, M22, M01) and (M02, M12,
// bx, by) with b = M p; row 8 holds 1 - p^T M p, row 9 bz, row 10 the
// coefficient c (p relative to the chunk's cell corner).  At a chunk's start
// the ring is empty, and the upper half of rows 0-8 holds the chunk's vertex
// operands until they are in registers.
constexpr int kAfStride = 144;
constexpr int kAfQueue = 128;
constexpr int kAfOneMinusC = 8, kAfBz = 9, kAfCoef = 10, kAfRows = 11;

// Sums c max(1 - |G (p - x)|^2, 0)^3 of n staged candidates (a multiple of 16)
// for the NT 16-vertex tiles of a chunk.  With x relative to the same corner,
//   |G (p - x)|^2 = p^T M p + [M00 M11 M22 M01].[xx yy zz 2xy]
//                 + [M02 M12 bx by].[2xz 2yz -2x -2y] - 2 bz z,
// so per 16 candidates and vertex tile two v_mfma_f32_16x16x4_f32 give
// D = 1 - p^T M p - (the first eight terms) (C = the candidates' 1 - p^T M p,
// B = the vertex terms negated), and e = D + bz (2z) takes one multiply-add
// whose clamp bit gives max(., 0) (e exceeds 1 only by rounding, for a vertex
// at the particle).  Lane l holds the pairs of candidates 4(l/16)+r (r = 0..3)
// with vertex l%16 of each tile and adds c e^3 to part[T]; the caller adds
// the partial sums of the four lane groups of a vertex.
template <int NT>
__device__ __forceinline__ void aniso_tiles(const float* __restrict__ rows, uint32_t n,
                                            const float (&b1)[NT], const float (&b2)[NT],
                                            const float (&z2)[NT], float (&part)[NT]) {
    const uint32_t lane = __lane_id();
    const float* a1p = rows + (lane >> 4) * kAfStride + (lane & 15u);
    const float* a2p = rows + (4u + (lane >> 4)) * kAfStride + (lane & 15u);
    const float* cp = rows + kAfOneMinusC * kAfStride + 4u * (lane >> 4);
    const float* zp = rows + kAfBz * kAfStride + 4u * (lane >> 4);
    const float* kp = rows + kAfCoef * kAfStride + 4u * (lane >> 4);
    for (uint32_t tb = 0; tb < n; tb += 16u) {
        const float a1 = a1p[tb];
        const float a2 = a2p[tb];
        const float4 c4 = *reinterpret_cast<const float4*>(cp + tb);
        const float4 z4 = *reinterpret_cast<const float4*>(zp + tb);
        const float4 k4 = *reinterpret_cast<const float4*>(kp + tb);
        const v4f c = v4f{c4.x, c4.y, c4.z, c4.w};
        const float bzr[4] = {z4.x, z4.y, z4.z, z4.w};
        const float kr[4] = {k4.x, k4.y, k4.z, k4.w};
        v4f d[NT];
#pragma unroll
        for (int T = 0; T < NT; ++T) d[T] = __builtin_amdgcn_mfma_f32_16x16x4f32(a1, b1[T], c, 0, 0, 0);
#pragma unroll
        for (int T = 0; T < NT; ++T) d[T] = __builtin_amdgcn_mfma_f32_16x16x4f32(a2, b2[T], d[T], 0, 0, 0);
#pragma unroll
        for (int T = 0; T < NT; ++T) {
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const float e = __builtin_amdgcn_fmed3f(__builtin_fmaf(bzr[r], z2[T], d[T][r]),
                                                        0.0f, 1.0f);
                part[T] = __builtin_fmaf(kr[r], e * e * e, part[T]);
            }
        }
    }
}

// kWaves: the resident waves per SIMD the register allocation is bounded for
// (5: 96 VGPRs; 4: 128 VGPRs).
template <int kWaves>
__global__ __launch_bounds__(kAnisoThreads, kWaves) void k_field_aniso(AnisoFieldArgs a) {
    __shared__ __align__(16) float queue[kAnisoThreads / 64][kAfRows * kAfStride];
    const uint32_t lane = __lane_id();
    const uint32_t wave = threadIdx.x >> 6;
    const uint32_t nch = *a.num_chunks;
    const int R = a.bc.radius;
    float* q = queue[wave];
    uint32_t step = 0;
    uint32_t n_staged = 0, n_kept = 0;
    for (;;) {
        const uint32_t ch = next_item_xcd(a.next, nch, step);
        if (ch >= nch) break;
        const uint32_t v = a.bc.chunk_cell[ch];
        const int cx = (int)(v % a.bc.vcx) - R;
        const int cy = (int)((v / a.bc.vcx) % a.bc.vcy) - R;
        const int cz = (int)(v / (a.bc.vcx * a.bc.vcy)) - R;
        const uint32_t k = ch - a.bc.chunk_off[v];
        const uint32_t first = a.bc.cell_off[v] + 64u * k;
        const uint32_t cnt = min(64u, a.bc.cell_cnt[v] - 64u * k);
        const float ox = __builtin_fmaf((float)cx, a.sp.cell, a.sp.minx);
        const float oy = __builtin_fmaf((float)cy, a.sp.cell, a.sp.miny);
        const float oz = __builtin_fmaf((float)cz, a.sp.cell, a.sp.minz);
        uint32_t vid = 0;
        float vx = 0.f, vy = 0.f, vz = 0.f;
        const bool live = lane < cnt;
        if (live) {
            vid = a.bc.cell_band[first + lane];
            const uint32_t ix = vid % a.sc.rx;
            const uint32_t iy = (vid / a.sc.rx) % a.sc.ry;
            const uint32_t iz = vid / (a.sc.rx * a.sc.ry);
            vx = vertex_coord(ix, a.sc.minx, a.sc.cell) - ox;
            vy = vertex_coord(iy, a.sc.miny, a.sc.cell) - oy;
            vz = vertex_coord(iz, a.sc.minz, a.sc.cell) - oz;
        }
        // the vertex operands (negated terms, and 2z) into the ring's upper half
        float* vt = q + kAfQueue / 2;
        vt[0 * kAfStride + lane] = -(vx * vx);
        vt[1 * kAfStride + lane] = -(vy * vy);
        vt[2 * kAfStride + lane] = -(vz * vz);
        vt[3 * kAfStride + lane] = -(2.0f * vx * vy);
        vt[4 * kAfStride + lane] = -(2.0f * vx * vz);
        vt[5 * kAfStride + lane] = -(2.0f * vy * vz);
        vt[6 * kAfStride + lane] = 2.0f * vx;
        vt[7 * kAfStride + lane] = 2.0f * vy;
        vt[8 * kAfStride + lane] = 2.0f * vz;
        const float bx0 = wave_min(live ? vx : 3.0e38f), bx1 = wave_max(live ? vx : -3.0e38f);
        const float by0 = wave_min(live ? vy : 3.0e38f), by1 = wave_max(live ? vy : -3.0e38f);
        const float bz0 = wave_min(live ? vz : 3.0e38f), bz1 = wave_max(live ? vz : -3.0e38f);
        // the block's candidate rows; with the non-involved moments only the
        // cells whose involved particles can reach the chunk's vertices
        BlockRows br;
        if (a.block_mom) {
            CellCull cc;
            cc.lo = a.cell_lo;
            cc.hi = a.cell_hi;
            cc.b0[0] = bx0 + ox; cc.b0[1] = by0 + oy; cc.b0[2] = bz0 + oz;
            cc.b1[0] = bx1 + ox; cc.b1[1] = by1 + oy; cc.b1[2] = bz1 + oz;
            br = block_rows<2>(a.ranges, a.sp, cx, cy, cz, &cc);
        } else {
            br = block_rows<2>(a.ranges, a.sp, cx, cy, cz);
        }
        const uint32_t total = br.total;
        n_staged += total;
        wave_lds_sync();
        const int nt = (int)((cnt + 15u) >> 4);
        float sum = 0.f;
        // The rest of the chunk, compiled once per tile count NT, so that the
        // vertex operands and sums keep fixed registers across the staging.
        auto run = [&](auto ntc) {
            constexpr int NT = decltype(ntc)::value;
            // lane l: component l/16 of the operands of vertex 16 T + l%16
            float b1[NT], b2[NT], z2[NT], tot[NT];
            {
                const float* src = vt + (lane >> 4) * kAfStride + (lane & 15u);
#pragma unroll
                for (int T = 0; T < NT; ++T) {
                    b1[T] = src[16 * T];
                    b2[T] = src[4 * kAfStride + 16 * T];
                    z2[T] = vt[8 * kAfStride + 16 * T + (lane & 15u)];
                    tot[T] = 0.f;
                }
            }
            wave_lds_sync();
            // Staging in two steps.  Each lane's candidate position (with its
            // squared culling radius in w) is loaded one round ahead; after
            // the box test, the two record words are loaded for the survivors
            // only (a dropped candidate reads record 0, one shared line) and
            // enter the queue one round later, so their load overlaps the next
            // round's test.  Queue order stays candidate order.  Entries
            // [head, head + fill) of the ring wait for evaluation; head is 0
            // or 64.
            uint32_t head = 0, fill = 0;
            float4 np4 = make_float4(0.f, 0.f, 0.f, 0.f);
            uint32_t nj;
            {
                nj = br.offset(0u, lane) + lane;
                if (lane < total) np4 = a.smooth[nj];
            }
            bool keep_p = false;
            float px_p = 0.f, py_p = 0.f, pz_p = 0.f;
            float4 r0_p = make_float4(0.f, 0.f, 0.f, 0.f), r1_p = r0_p;
            for (uint32_t base = 0; base < total + 64u; base += 64u) {
                const uint32_t g = base + lane;
                bool keep = false;
                float px = 0.f, py = 0.f, pz = 0.f;
                float4 r0 = make_float4(0.f, 0.f, 0.f, 0.f), r1 = r0;
                if (base < total) {
                    const float4 p4 = np4;
                    const uint32_t j = nj;
                    if (base + 64u < total) {
                        nj = br.offset(base + 64u, g + 64u) + g + 64u;
                        if (g + 64u < total) np4 = a.smooth[nj];
                    }
                    if (g < total) {
                        px = p4.x - ox;
                        py = p4.y - oy;
                        pz = p4.z - oz;
                        const float dx = px - __builtin_amdgcn_fmed3f(px, bx0, bx1);
                        const float dy = py - __builtin_amdgcn_fmed3f(py, by0, by1);
                        const float dz = pz - __builtin_amdgcn_fmed3f(pz, bz0, bz1);
                        keep = __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, dx * dx)) < p4.w;
                    }
                    const uint32_t jr = keep ? j : 0u;
                    r0 = a.rec[2ull * jr + 0];
                    r1 = a.rec[2ull * jr + 1];
                }
                if (base == 0) {
                    keep_p = keep;
                    px_p = px; py_p = py; pz_p = pz;
                    r0_p = r0; r1_p = r1;
                    continue;
                }
                // the previous round's survivors enter the queue
                const uint64_t m = __ballot(keep_p);
                if (keep_p) {
                    const float M00 = r0_p.x, M11 = r0_p.y, M22 = r0_p.z, M01 = r0_p.w;
                    const float M02 = r1_p.x, M12 = r1_p.y;
                    const float bx = M00 * px_p + M01 * py_p + M02 * pz_p;
                    const float by = M01 * px_p + M11 * py_p + M12 * pz_p;
                    const float bz = M02 * px_p + M12 * py_p + M22 * pz_p;
                    const float cc = px_p * bx + py_p * by + pz_p * bz;
                    const uint32_t o = (head + fill + mask_rank(m)) & (kAfQueue - 1);
                    q[0 * kAfStride + o] = M00;
                    q[1 * kAfStride + o] = M11;
                    q[2 * kAfStride + o] = M22;
                    q[3 * kAfStride + o] = M01;
                    q[4 * kAfStride + o] = M02;
                    q[5 * kAfStride + o] = M12;
                    q[6 * kAfStride + o] = bx;
                    q[7 * kAfStride + o] = by;
                    q[kAfOneMinusC * kAfStride + o] = 1.0f - cc;
                    q[kAfBz * kAfStride + o] = bz;
                    q[kAfCoef * kAfStride + o] = r1_p.z;
                }
                keep_p = keep;
                px_p = px; py_p = py; pz_p = pz;
                r0_p = r0; r1_p = r1;
                fill += (uint32_t)__popcll(m);
                n_kept += (uint32_t)__popcll(m);
                if (fill >= 64u) {
                    // one batch of 64: its own partial sums, added to the
                    // running totals (a vertex sums many similar weights; a
                    // batch sums them at its own, smaller, magnitude)
                    float part[NT];
#pragma unroll
                    for (int T = 0; T < NT; ++T) part[T] = 0.f;
                    wave_lds_sync();
                    aniso_tiles<NT>(q + head, 64u, b1, b2, z2, part);
                    wave_lds_sync();
#pragma unroll
                    for (int T = 0; T < NT; ++T) tot[T] += part[T];
                    head ^= 64u;
                    fill -= 64u;
                }
            }
            if (fill) {
                // pad to a multiple of 16 with entries of weight 0 (1 - p^T M p
                // = -1, every other row 0); head + 64 <= kAfQueue, so the pad
                // does not wrap
                const uint32_t n = (fill + 15u) & ~15u;
                if (lane >= fill && lane < n) {
#pragma unroll
                    for (int rr = 0; rr < kAfRows; ++rr)
                        q[rr * kAfStride + head + lane] = rr == kAfOneMinusC ? -1.0f : 0.0f;
                }
                float part[NT];
#pragma unroll
                for (int T = 0; T < NT; ++T) part[T] = 0.f;
                wave_lds_sync();
                aniso_tiles<NT>(q + head, n, b1, b2, z2, part);
                wave_lds_sync();
#pragma unroll
                for (int T = 0; T < NT; ++T) tot[T] += part[T];
            }
            // add the four lane groups' sums of each vertex; lane l keeps those
            // of vertex 16 (l/16) + l%16 = l
            const uint32_t my = lane >> 4;
#pragma unroll
            for (int T = 0; T < NT; ++T) {
                float x = tot[T];
                x += __int_as_float(__builtin_amdgcn_ds_swizzle(__float_as_int(x), 0x1F | (16 << 10)));
                x += __shfl_xor(x, 32, 64);
                if (my == (uint32_t)T) sum = x;
            }
        };
        switch (nt) {
            case 1: run(std::integral_constant<int, 1>{}); break;
            case 2: run(std::integral_constant<int, 2>{}); break;
            case 3: run(std::integral_constant<int, 3>{}); break;
            default: run(std::integral_constant<int, 4>{}); break;
        }
        if (live) {
            if (a.block_mom) {
                const double px = (double)(vx + ox) - a.ox, py = (double)(vy + oy) - a.oy,
                             pz = (double)(vz + oz) - a.oz;
                const double pn = noninv_poly(a.block_mom + (uint64_t)kNonInvMoments * v, px, py, pz);
                a.phi[vid] = (float)((double)sum + pn - 0.5);
            } else {
                a.phi[vid] = sum - 0.5f;
            }
        }
    }
    if (a.work) write_work(a.work, n_staged, n_kept);
}


// One wave per hash cell.
//  * Moments: for a non-involved cell the kNonInvMoments moments of its
//    particles (smoothed positions about the origin (ox, oy, oz), weights
//    sigma m / rho), zeros otherwise.  The terms of 64 particles at a time go
//    through LDS and lane k sums term k over them, so the order is fixed.
//  * Box: for the cell test of the field kernel, the box of the smoothed
//    positions of the particles the field stages (w > 0: their squared culling
//    radius) with the largest w, widened by 0.1% in radius and a few ulps in
//    position (the cell test runs in absolute coordinates, the particle test
//    relative to the chunk's cell); w = -1 when the cell stages none.
__global__ __launch_bounds__(64) void k_cell_moments_boxes(
    const fsr_index_range* __restrict__ ranges, const uint32_t* __restrict__ involve,
    const float4* __restrict__ smooth, const float4* __restrict__ sorted, float mass, double ox,
    double oy, double oz, double* __restrict__ mom, float4* __restrict__ lo,
    float4* __restrict__ hi) {
    // stride kNonInvMoments + 1 doubles between particles
    __shared__ double terms[64 * (kNonInvMoments + 1)];
    const uint32_t c = blockIdx.x;
    const uint32_t lane = threadIdx.x;
    const fsr_index_range r = ranges[c];
    const bool occupied = r.start != 0xFFFFFFFFu;
    const bool inv = occupied && involve[c] > 0;
    double acc = 0.0;
    float4 l = make_float4(3.0e38f, 3.0e38f, 3.0e38f, -1.0f);
    float4 u = make_float4(-3.0e38f, -3.0e38f, -3.0e38f, 0.0f);
    if (occupied) {
        for (uint32_t base = r.start; base < r.end; base += 64u) {
            const uint32_t j = base + lane;
            const bool live = j < r.end;
            float4 p = make_float4(0.f, 0.f, 0.f, -1.0f);
            if (live) p = smooth[j];
            if (inv) {
                if (p.w > 0.0f) {
                    l.x = fminf(l.x, p.x); l.y = fminf(l.y, p.y); l.z = fminf(l.z, p.z);
                    u.x = fmaxf(u.x, p.x); u.y = fmaxf(u.y, p.y); u.z = fmaxf(u.z, p.z);
                    l.w = fmaxf(l.w, p.w);
                }
            } else {
                const float cj = live ? kSigma * (mass / sorted[j].w) : 0.0f;
                double t[kNonInvMoments];
                moment_terms((double)p.x - ox, (double)p.y - oy, (double)p.z - oz, t);
#pragma unroll
                for (int k = 0; k < kNonInvMoments; ++k)
                    terms[lane * (kNonInvMoments + 1) + k] = (double)cj * t[k];
                wave_lds_sync();
                if (lane < (uint32_t)kNonInvMoments) {
                    for (uint32_t q = 0; q < 64u; ++q) acc += terms[q * (kNonInvMoments + 1) + lane];
                }
                wave_lds_sync();
            }
        }
    }
    if (lane < (uint32_t)kNonInvMoments) mom[(uint64_t)c * kNonInvMoments + lane] = acc;
    l.x = wave_minf(l.x); l.y = wave_minf(l.y); l.z = wave_minf(l.z);
    u.x = wave_maxf(u.x); u.y = wave_maxf(u.y); u.z = wave_maxf(u.z);
    l.w = wave_maxf(l.w);
    if (lane == 0) {
        if (l.w > 0.0f) {
            l.w = l.w * 1.002f + 1.0e-12f;
            const float e = 1.0e-6f * fmaxf(fmaxf(fmaxf(fabsf(l.x), fabsf(u.x)),
                                                  fmaxf(fabsf(l.y), fabsf(u.y))),
                                            fmaxf(fabsf(l.z), fabsf(u.z)));
            l.x -= e; l.y -= e; l.z -= e;
            u.x += e; u.y += e; u.z += e;
        } else {
            l.w = -1.0f;
        }
        lo[c] = l;
        hi[c] = u;
    }
}

// One pass of the 5^3 block sums: along one axis, out(e) = sum of in over the
// clamped window [c - 2, c + 2] of the real cells, where the output index along
// the axis is a virtual cell (c + 2, c in [-2, res + 1]).  Elements are
// (a, i, b) with i along the axis; in has nin cells along it, out nout.
__global__ __launch_bounds__(256) void k_window5(const double* __restrict__ in,
                                                 double* __restrict__ out, uint32_t na,
                                                 uint32_t nin, uint32_t nout, uint32_t nb) {
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t total = (uint64_t)na * nout * nb * kNonInvMoments;
    if (t >= total) return;
    const uint32_t k = (uint32_t)(t % kNonInvMoments);
    const uint64_t e = t / kNonInvMoments;
    const uint32_t b = (uint32_t)(e % nb);
    const uint32_t i = (uint32_t)((e / nb) % nout);
    const uint32_t a = (uint32_t)(e / ((uint64_t)nb * nout));
    const int c = (int)i - 2;
    const int lo = clampi(c - 2, 0, (int)nin - 1), hi = clampi(c + 2, 0, (int)nin - 1);
    double acc = 0.0;
    for (int j = lo; j <= hi; ++j)
        acc += in[(((uint64_t)a * nin + (uint32_t)j) * nb + b) * kNonInvMoments + k];
    out[t] = acc;
}

// The field kernel tests a candidate's box distance against the candidate's
// squared culling radius (w of its second record word) before it reads the
// records: a copy of that radius goes into the w of the smoothed position.
// With the non-involved moments, those particles (never culled otherwise)
// get a negative radius: the box test then drops them.
__global__ __launch_bounds__(256) void k_cull_radius_to_w(const float4* __restrict__ rec,
                                                          uint32_t n, float4* __restrict__ smooth,
                                                          int drop_noninv) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float w = rec[2ull * i + 1].w;
    if (drop_noninv && w == kNoCull) w = -1.0f;
    reinterpret_cast<float*>(smooth)[4ull * i + 3] = w;
}

__global__ __launch_bounds__(256) void k_cell_counts(const fsr_index_range* __restrict__ ranges,
                                                     const uint32_t* __restrict__ filter,
                                                     uint32_t C, uint32_t* __restrict__ cnt) {
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= C) return;
    const fsr_index_range r = ranges[c];
    uint32_t n = r.start == 0xFFFFFFFFu ? 0u : r.end - r.start;
    if (filter && filter[c] == 0) n = 0;
    cnt[c] = n;
}

__global__ __launch_bounds__(256) void k_chunks_of_cells(const uint32_t* __restrict__ cnt,
                                                         const uint32_t* __restrict__ chunk_off,
                                                         uint32_t C, uint32_t* __restrict__ chunk_cell) {
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= C) return;
    const uint32_t k = (cnt[c] + 63u) >> 6;
    const uint32_t o = chunk_off[c];
    for (uint32_t j = 0; j < k; ++j) chunk_cell[o + j] = c;
}

}  // namespace

// Largest float t with sqrtf(t) <= r (strict: sqrtf(t) < r), sqrtf correctly
// rounded; r > 0.
static float sqrt_bound(float r, bool strict) {
    auto ok = [&](float t) { return strict ? std::sqrt(t) < r : std::sqrt(t) <= r; };
    const float inf = std::numeric_limits<float>::infinity();
    float t = r * r;
    while (t > 0.0f && !ok(t)) t = std::nextafter(t, 0.0f);
    while (ok(std::nextafter(t, inf))) t = std::nextafter(t, inf);
    return t;
}

fsr_status cell_chunk_list(fsr_context* ctx, const fsr_index_range* ranges,
                           const uint32_t* filter, uint32_t C, uint32_t* cnt,
                           uint32_t* chunk_off, uint32_t* chunk_cell, uint32_t* d_total,
                           hipStream_t s) {
    (void)ctx;
    hipLaunchKernelGGL(k_cell_counts, dim3((C + 255) / 256), dim3(256), 0, s, ranges, filter, C,
                       cnt);
    FSR_LAUNCH_CHECK("k_cell_counts");
    FSR_TRY(scan_exclusive(ctx, SCAN_CHUNKS64, cnt, chunk_off, C, nullptr, d_total, s));
    hipLaunchKernelGGL(k_chunks_of_cells, dim3((C + 255) / 256), dim3(256), 0, s, cnt, chunk_off,
                       C, chunk_cell);
    FSR_LAUNCH_CHECK("k_chunks_of_cells");
    return FSR_OK;
}

// Scratch of the two particle passes: per-cell counts and chunk offsets, and
// the chunk -> cell lists.
// The work-record sections (see write_work), one arena slot for the three
// kernels.
static fsr_status work_section(fsr_context* ctx, WorkSection w, uint32_t** out) {
    const uint64_t per = 1ull + 2ull * kWorkWavesPerCu * (uint64_t)ctx->num_cus;
    uint32_t* base;
    FSR_TRY(arena_typed(ctx, A_ANISO_WORK, WORK_SECTIONS * per, &base));
    *out = base + (uint64_t)w * per;
    return FSR_OK;
}

struct PassScratch {
    uint32_t *cnt, *chunk_off, *cnt2, *chunk_off2, *chunk_cell, *chunk_cell2;
};

static fsr_status pass_scratch(fsr_context* ctx, uint32_t C, uint32_t n, PassScratch* ps) {
    FSR_TRY(arena_typed(ctx, A_AUX0, 4ull * C, &ps->cnt));
    ps->chunk_off = ps->cnt + C;
    ps->cnt2 = ps->chunk_off + C;
    ps->chunk_off2 = ps->cnt2 + C;
    const uint64_t max_chunks = (uint64_t)n / 64 + C + 1;
    FSR_TRY(arena_typed(ctx, A_AUX1, 2ull * max_chunks, &ps->chunk_cell));
    ps->chunk_cell2 = ps->chunk_cell + max_chunks;
    return FSR_OK;
}

// The cell boxes of the smoothing and covariance: a.cell_box, or computed here.
static fsr_status cell_boxes(fsr_context* ctx, const AnisoArgs& a, hipStream_t s,
                             const float4** out) {
    if (a.cell_box) {
        *out = a.cell_box;
        return FSR_OK;
    }
    const uint32_t C = a.sp.rx * a.sp.ry * a.sp.rz;
    float4* box;
    FSR_TRY(arena_typed(ctx, A_CELL_BOX_RAW, 2ull * C, &box));
    const float w = a.aniso_radius * a.aniso_radius * kRadiusCull * 1.002f + 1.0e-12f;
    hipLaunchKernelGGL(k_cell_box, dim3(C), dim3(64), 0, s, a.ranges, a.sorted, w, box,
                       box + C);
    FSR_LAUNCH_CHECK("k_cell_box");
    *out = box;
    return FSR_OK;
}

fsr_status aniso_smooth(fsr_context* ctx, const AnisoArgs& a, hipStream_t s) {
    const uint32_t C = a.sp.rx * a.sp.ry * a.sp.rz;
    const float4* box;
    FSR_TRY(cell_boxes(ctx, a, s, &box));
    const int blocks = persistent_blocks(ctx, kAnisoThreads, 5, TUNE_ANISO);
    PassScratch ps;
    FSR_TRY(pass_scratch(ctx, C, a.n, &ps));
    hipLaunchKernelGGL(k_cell_counts, dim3((C + 255) / 256), dim3(256), 0, s, a.ranges,
                       (const uint32_t*)nullptr, C, ps.cnt);
    FSR_LAUNCH_CHECK("k_cell_counts");
    FSR_TRY(scan_exclusive(ctx, SCAN_CHUNKS64, ps.cnt, ps.chunk_off, C, nullptr,
                           &a.counters->aux[0], s));
    hipLaunchKernelGGL(k_chunks_of_cells, dim3((C + 255) / 256), dim3(256), 0, s, ps.cnt,
                       ps.chunk_off, C, ps.chunk_cell);
    FSR_LAUNCH_CHECK("k_chunks_of_cells");
    ChunkList cl{ps.chunk_cell, ps.chunk_off, &a.counters->aux[0], a.counters->xcd_next};
    FSR_TRY(reset_xcd(a.counters, s));
    uint32_t* work;
    FSR_TRY(work_section(ctx, WORK_SMOOTH, &work));
    hipLaunchKernelGGL(k_smooth_mc, dim3(blocks), dim3(kAnisoThreads), 0, s, a.sorted,
                       a.ranges, a.sp, a.aniso_radius, a.lambda, cl, box, a.mean, a.smooth,
                       work);
    FSR_LAUNCH_CHECK("k_smooth_mc");
    return FSR_OK;
}

fsr_status aniso_involved(fsr_context* ctx, const AnisoArgs& a, hipStream_t s) {
    const uint32_t C = a.sp.rx * a.sp.ry * a.sp.rz;
    FSR_TRY(scan_exclusive(ctx, SCAN_U32, a.involve, a.involve_off, C, nullptr,
                           &a.st->num_involved_particles, s));
    if (a.n) {
        hipLaunchKernelGGL(k_involved, dim3((a.n + 255) / 256), dim3(256), 0, s, a.keys, a.ranges,
                           a.involve, a.involve_off, a.sorted, a.n, C, a.mass, a.involved_idx,
                           a.rec);
        FSR_LAUNCH_CHECK("k_involved");
    }
    return FSR_OK;
}

fsr_status aniso_cov(fsr_context* ctx, const AnisoArgs& a, hipStream_t s) {
    const uint32_t C = a.sp.rx * a.sp.ry * a.sp.rz;
    // 4 resident workgroups per CU (LDS: 34 KB each)
    const int blocks_cov = persistent_blocks(ctx, kAnisoThreads, 4, TUNE_ANISO);
    PassScratch ps;
    FSR_TRY(pass_scratch(ctx, C, a.n, &ps));
    hipLaunchKernelGGL(k_cell_counts, dim3((C + 255) / 256), dim3(256), 0, s, a.ranges,
                       a.involve, C, ps.cnt2);
    FSR_LAUNCH_CHECK("k_cell_counts");
    FSR_TRY(scan_exclusive(ctx, SCAN_CHUNKS64, ps.cnt2, ps.chunk_off2, C, nullptr,
                           &a.counters->aux[2], s));
    hipLaunchKernelGGL(k_chunks_of_cells, dim3((C + 255) / 256), dim3(256), 0, s, ps.cnt2,
                       ps.chunk_off2, C, ps.chunk_cell2);
    FSR_LAUNCH_CHECK("k_chunks_of_cells");
    ChunkList cl2{ps.chunk_cell2, ps.chunk_off2, &a.counters->aux[2], a.counters->xcd_next};
    CovArgs ca;
    ca.sorted = a.sorted;
    ca.mean = a.mean;
    ca.ranges = a.ranges;
    ca.involve_off = a.involve_off;
    ca.sp = a.sp;
    ca.radius = a.aniso_radius;
    ca.d2_le_r = sqrt_bound(a.aniso_radius, false);
    ca.d2_lt_r = sqrt_bound(a.aniso_radius, true);
    ca.h2 = a.h2;
    ca.inv_h = a.inv_h;
    ca.kr = a.kr;
    ca.kn = a.kn;
    ca.min_neighbors = a.min_neighbors;
    ca.mass = a.mass;
    ca.gmat = a.gmat;
    ca.sigma = a.sigma;
    ca.rec = a.rec;
    FSR_TRY(work_section(ctx, WORK_COV, &ca.work));
    FSR_TRY(cell_boxes(ctx, a, s, &ca.cell_box));
    FSR_TRY(reset_xcd(a.counters, s));
    hipLaunchKernelGGL(k_covariance_mc, dim3(blocks_cov), dim3(kAnisoThreads), 0, s, ca, cl2);
    FSR_LAUNCH_CHECK("k_covariance_mc");
    return FSR_OK;
}

fsr_status aniso_run(fsr_context* ctx, const AnisoArgs& a, hipStream_t s) {
    // the cell boxes once for both neighbour passes
    AnisoArgs b = a;
    FSR_TRY(cell_boxes(ctx, a, s, &b.cell_box));
    FSR_TRY(aniso_smooth(ctx, b, s));
    FSR_TRY(aniso_involved(ctx, b, s));
    return aniso_cov(ctx, b, s);
}

// Whether every particle of a clamped 5^3 block lies within distance 1 of
// every vertex whose cell owns the block, smoothing displacement included: the
// reference's non-involved kernel then never reaches its cutoff |r| >= 1 and
// the block sum is the polynomial of noninv_poly.
static bool noninv_polynomial_exact(const GridDev& sp, float smoothing_radius) {
    const double reach = std::sqrt(3.0) * 6.0 * (double)sp.cell + 2.0 * (double)smoothing_radius;
    return reach < 0.9;
}

fsr_status aniso_field(fsr_context* ctx, const AnisoArgs& a, const BandCompactOut& bc,
                       float* phi, hipStream_t s) {
    // a field limit of at most 4 workgroups per CU runs the build bounded for
    // 4 waves per SIMD (no spills); the default the one for 5
    const int tuned = ctx->tune[TUNE_FIELD];
    const int field_waves = (tuned > 0 && tuned <= 4) ? 4 : 5;
    const int blocks = persistent_blocks(ctx, kAnisoThreads, field_waves, TUNE_FIELD);
    AnisoFieldArgs fa;
    fa.block_mom = nullptr;
    fa.cell_lo = nullptr;
    fa.cell_hi = nullptr;
    FSR_TRY(work_section(ctx, WORK_FIELD, &fa.work));
    fa.ox = a.sp.minx;
    fa.oy = a.sp.miny;
    fa.oz = a.sp.minz;
    const bool moments = a.noninv_moments && noninv_polynomial_exact(a.sp, a.aniso_radius);
    if (moments) {
        const uint32_t C = a.sp.rx * a.sp.ry * a.sp.rz;
        const uint32_t rx = a.sp.rx, ry = a.sp.ry, rz = a.sp.rz;
        const uint32_t vx = bc.vcx, vy = bc.vcy, vz = bc.vcz;
        double *cm, *t1, *t2, *bm;
        FSR_TRY(arena_typed(ctx, A_MOM_CELL, (uint64_t)C * kNonInvMoments, &cm));
        FSR_TRY(arena_typed(ctx, A_MOM_T1, (uint64_t)rz * ry * vx * kNonInvMoments, &t1));
        FSR_TRY(arena_typed(ctx, A_MOM_T2, (uint64_t)rz * vy * vx * kNonInvMoments, &t2));
        FSR_TRY(arena_typed(ctx, A_MOM_BLOCK, (uint64_t)vz * vy * vx * kNonInvMoments, &bm));
        float4* boxes;
        FSR_TRY(arena_typed(ctx, A_CELL_BOX, 2ull * C, &boxes));
        fa.cell_lo = boxes;
        fa.cell_hi = boxes + C;
        // the boxes need the culling radii in the w of the smoothed positions
        if (a.n) {
            hipLaunchKernelGGL(k_cull_radius_to_w, dim3((a.n + 255) / 256), dim3(256), 0, s,
                               a.rec, a.n, a.smooth, 1);
            FSR_LAUNCH_CHECK("k_cull_radius_to_w");
        }
        hipLaunchKernelGGL(k_cell_moments_boxes, dim3(C), dim3(64), 0, s, a.ranges, a.involve,
                           (const float4*)a.smooth, a.sorted, a.mass, fa.ox, fa.oy, fa.oz, cm,
                           boxes, boxes + C);
        FSR_LAUNCH_CHECK("k_cell_moments_boxes");
        auto window = [&](const double* in, double* out, uint32_t na, uint32_t nin, uint32_t nout,
                          uint32_t nb) -> fsr_status {
            const uint64_t total = (uint64_t)na * nout * nb * kNonInvMoments;
            hipLaunchKernelGGL(k_window5, dim3((uint32_t)((total + 255) / 256)), dim3(256), 0, s,
                               in, out, na, nin, nout, nb);
            FSR_LAUNCH_CHECK("k_window5");
            return FSR_OK;
        };
        FSR_TRY(window(cm, t1, rz * ry, rx, vx, 1));   // x: (z, y) x vx
        FSR_TRY(window(t1, t2, rz, ry, vy, vx));        // y: z x vy x vx
        FSR_TRY(window(t2, bm, 1, rz, vz, vy * vx));    // z: vz x (vy, vx)
        fa.block_mom = bm;
    }
    fa.smooth = a.smooth;
    fa.rec = a.rec;
    fa.ranges = a.ranges;
    fa.bc = bc;
    fa.sp = a.sp;
    fa.sc = a.sc;
    fa.phi = phi;
    fa.next = a.counters->xcd_next;
    fa.num_chunks = &a.counters->num_chunks;
    if (a.n && !moments) {
        hipLaunchKernelGGL(k_cull_radius_to_w, dim3((a.n + 255) / 256), dim3(256), 0, s, a.rec,
                           a.n, a.smooth, 0);
        FSR_LAUNCH_CHECK("k_cull_radius_to_w");
    }
    FSR_TRY(reset_xcd(a.counters, s));
    if (field_waves == 4)
        hipLaunchKernelGGL(k_field_aniso<4>, dim3(blocks), dim3(kAnisoThreads), 0, s, fa);
    else
        hipLaunchKernelGGL(k_field_aniso<5>, dim3(blocks), dim3(kAnisoThreads), 0, s, fa);
    FSR_LAUNCH_CHECK("k_field_aniso");
    return FSR_OK;
}

}  // namespace fsr
