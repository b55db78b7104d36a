// CPU oracle of the particle -> field -> mesh path.  TEST INFRASTRUCTURE ONLY
// (the checker of the parity tests and the CPU baseline of bench.py).
//
// Every function restates one reference kernel; the citation is given at the
// function.  Reference files are under FluidSurfaceReconstruction/GPU/.
// Build: g++ -O2 -fopenmp -ffp-contract=off (oracle/Makefile), so FP32
// expressions are evaluated exactly as written, one IEEE operation at a time,
// and fused multiply-adds appear only where fmaf() is written.
#include "fsr_oracle.h"

#include <omp.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <vector>

#include "../fluidsurfacereconstructiongpu_amd/csrc/mc_tables.h"

namespace {

const uint64_t kTri[256] = {FSR_MC_TRI_PACKED_VALUES};

inline int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

// int3((p - min) / cell) of ReconstructionCUDA.cuh:354-358 in FP32, with the
// quotient clamped first so that the truncation stays defined for far-away or
// NaN positions (NaN -> 0).
inline float cell_q(float p, float mn, float cell) { return (p - mn) / cell; }
inline int trunc_q(float q, int res) {
    if (!(q == q)) return 0;
    const float lo = -4.0f, hi = (float)res + 4.0f;
    q = q < lo ? lo : (q > hi ? hi : q);
    return (int)q;
}

// getVertexPos (ReconstructionCUDA.cuh:366-370) as one FP32 fused multiply-add.
inline float vpos(uint32_t i, float mn, float cell) { return fmaf((float)i, cell, mn); }

inline uint32_t case_nverts(uint32_t flag) {
    uint32_t n = 0;
    for (int k = 0; k < 16; ++k) {
        if (((kTri[flag] >> (4 * k)) & 0xF) == 0xF) break;
        ++n;
    }
    return n;
}

}  // namespace

extern "C" {

int orc_num_threads(void) { return omp_get_max_threads(); }
void orc_set_threads(int n) {
    if (n > 0) omp_set_num_threads(n);
}

uint64_t orc_mc_tri_packed(uint32_t c) { return c < 256 ? kTri[c] : ~0ull; }

// calcParticlesHashKernel (ReconstructionCUDA.cu:1379-1396), thrust::sort_by_key
// (:1461-1466, a stable sort by key) and findCellRangeKernel (:1399-1435).
int orc_hash_sort(const float* pos, const float* dens, uint32_t n, orc_grid sp, uint32_t* keys,
                  uint32_t* perm, float* pos_out, float* dens_out, uint32_t* ranges, float* occ) {
    const uint64_t C = (uint64_t)sp.rx * sp.ry * sp.rz;
    std::vector<uint32_t> key(n);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < (int64_t)n; ++i) {
        const float qx = cell_q(pos[3 * i + 0], sp.minx, sp.cell);
        const float qy = cell_q(pos[3 * i + 1], sp.miny, sp.cell);
        const float qz = cell_q(pos[3 * i + 2], sp.minz, sp.cell);
        const bool ok = qx > -1.0f && qx < (float)sp.rx && qy > -1.0f && qy < (float)sp.ry &&
                        qz > -1.0f && qz < (float)sp.rz;
        key[i] = ok ? (uint32_t)(int)qz * sp.rx * sp.ry + (uint32_t)(int)qy * sp.rx +
                          (uint32_t)(int)qx
                    : (uint32_t)C;
    }
    // stable counting sort by key
    std::vector<uint32_t> start(C + 2, 0);
    for (uint32_t i = 0; i < n; ++i) start[key[i] + 1]++;
    for (uint64_t c = 0; c <= C; ++c) start[c + 1] += start[c];
    std::vector<uint32_t> cursor(start.begin(), start.end() - 1);
    std::vector<uint32_t> order(n);
    for (uint32_t i = 0; i < n; ++i) order[cursor[key[i]]++] = i;
    for (uint64_t c = 0; c < 2 * C; ++c) ranges[c] = 0xFFFFFFFFu;
    for (uint64_t c = 0; c < C; ++c) occ[c] = 0.0f;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < (int64_t)n; ++i) {
        const uint32_t j = order[i];
        keys[i] = key[j];
        if (perm) perm[i] = j;
        pos_out[3 * i + 0] = pos[3 * j + 0];
        pos_out[3 * i + 1] = pos[3 * j + 1];
        pos_out[3 * i + 2] = pos[3 * j + 2];
        dens_out[i] = dens[j];
    }
    for (uint64_t c = 0; c < C; ++c) {
        if (start[c + 1] > start[c]) {
            ranges[2 * c] = start[c];
            ranges[2 * c + 1] = start[c + 1];
            occ[c] = 1.0f;
        }
    }
    return 0;
}

// extractionOfSurfaceParticles (ReconstructionCUDA.cu:70-120); with involve,
// extractionOfSurfaceAndInvolveParticles (:466-551).
int orc_surface_occupancy(const float* occ, const uint32_t* ranges, orc_grid sp,
                          const uint32_t* keys, uint32_t n, uint8_t* surf_cell, uint32_t* flags,
                          uint32_t* involve) {
    const uint64_t C = (uint64_t)sp.rx * sp.ry * sp.rz;
    const int rx = sp.rx, ry = sp.ry, rz = sp.rz;
#pragma omp parallel for schedule(static)
    for (int64_t c = 0; c < (int64_t)C; ++c) {
        const int x = (int)(c % rx), y = (int)((c / rx) % ry), z = (int)(c / ((int64_t)rx * ry));
        uint8_t s = 0;
        if (x < rx - 1 && y < ry - 1 && z < rz - 1 && !(occ[c] < 1.0f)) {
            for (int zz = clampi(z - 1, 0, rz - 1); zz <= clampi(z + 1, 0, rz - 1); ++zz)
                for (int yy = clampi(y - 1, 0, ry - 1); yy <= clampi(y + 1, 0, ry - 1); ++yy)
                    for (int xx = clampi(x - 1, 0, rx - 1); xx <= clampi(x + 1, 0, rx - 1); ++xx)
                        if (occ[((int64_t)zz * ry + yy) * rx + xx] < 1.0f) s = 1;
        }
        surf_cell[c] = s;
    }
    if (involve) {
        for (uint64_t c = 0; c < C; ++c) involve[c] = 0;
        for (uint64_t c = 0; c < C; ++c) {
            if (!surf_cell[c]) continue;
            const int x = (int)(c % rx), y = (int)((c / rx) % ry), z = (int)(c / ((int64_t)rx * ry));
            for (int zz = clampi(z - 2, 0, rz - 1); zz <= clampi(z + 2, 0, rz - 1); ++zz)
                for (int yy = clampi(y - 2, 0, ry - 1); yy <= clampi(y + 2, 0, ry - 1); ++yy)
                    for (int xx = clampi(x - 2, 0, rx - 1); xx <= clampi(x + 2, 0, rx - 1); ++xx) {
                        const uint64_t cc = ((uint64_t)zz * ry + yy) * rx + xx;
                        if (ranges[2 * cc] != 0xFFFFFFFFu)
                            involve[cc] = ranges[2 * cc + 1] - ranges[2 * cc];
                    }
        }
    }
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < (int64_t)n; ++i) flags[i] = (keys[i] < C && surf_cell[keys[i]]) ? 1u : 0u;
    return 0;
}

// extractionOfSurfaceParticlesUsingColorField (ReconstructionCUDA.cu:909-971)
// with wSpikyGrad (ReconstructionCUDA.cuh:325-334); the gradient sum in FP64.
int orc_surface_colorfield(const float* pos, const float* dens, uint32_t n,
                           const uint32_t* ranges, orc_grid sp, float h, float mass,
                           uint32_t* flags, double* gnorm, uint32_t* count) {
    const int rx = sp.rx, ry = sp.ry, rz = sp.rz;
    const double H = h;
    const double coef = -45.0 / (M_PI * H * H * H * H);
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t i = 0; i < (int64_t)n; ++i) {
        const float px = pos[3 * i], py = pos[3 * i + 1], pz = pos[3 * i + 2];
        const int cx = trunc_q(cell_q(px, sp.minx, sp.cell), rx);
        const int cy = trunc_q(cell_q(py, sp.miny, sp.cell), ry);
        const int cz = trunc_q(cell_q(pz, sp.minz, sp.cell), rz);
        uint32_t cnt = 0;
        double gx = 0, gy = 0, gz = 0;
        for (int zz = clampi(cz - 1, 0, rz - 1); zz <= clampi(cz + 1, 0, rz - 1); ++zz)
            for (int yy = clampi(cy - 1, 0, ry - 1); yy <= clampi(cy + 1, 0, ry - 1); ++yy)
                for (int xx = clampi(cx - 1, 0, rx - 1); xx <= clampi(cx + 1, 0, rx - 1); ++xx) {
                    const uint64_t c = ((uint64_t)zz * ry + yy) * rx + xx;
                    if (ranges[2 * c] == 0xFFFFFFFFu) continue;
                    for (uint32_t j = ranges[2 * c]; j < ranges[2 * c + 1]; ++j) {
                        // |r|^2 > 0 decided in FP32 as the kernel computes it
                        const float fx = px - pos[3 * j], fy = py - pos[3 * j + 1],
                                    fz = pz - pos[3 * j + 2];
                        const float d2f = fmaf(fz, fz, fmaf(fy, fy, fx * fx));
                        if (!(d2f > 0.0f)) continue;
                        ++cnt;
                        const double dx = (double)px - pos[3 * j], dy = (double)py - pos[3 * j + 1],
                                     dz = (double)pz - pos[3 * j + 2];
                        const double dist = std::sqrt(dx * dx + dy * dy + dz * dz);
                        if (dist >= H || dist == 0.0) continue;
                        const double x = 1.0 - dist / H;
                        const double g = (double)mass / (double)dens[j] * coef * x * x / dist;
                        gx += g * dx;
                        gy += g * dy;
                        gz += g * dz;
                    }
                }
        const double gn = std::sqrt(gx * gx + gy * gy + gz * gz);
        flags[i] = (cnt < 25u || gn > 3.0) ? 1u : 0u;
        if (gnorm) gnorm[i] = gn;
        if (count) count[i] = cnt;
    }
    return 0;
}

// Radius-count surface test (no reference counterpart: the "fewer than 25
// neighbours within the radius" band of the dam-break configuration): i is a
// surface particle iff fewer than 25 particles j of its clamped 3^3 cell block
// have 0 < d2 < h * h, with d2 = fmaf(dz, dz, fmaf(dy, dy, dx * dx)) in FP32 of
// d = p_i - p_j (the kernel's order).  The hash cell is at least h, so the
// block holds every such j.
int orc_surface_radius_count(const float* pos, uint32_t n, const uint32_t* ranges, orc_grid sp,
                             float h, uint32_t* flags, uint32_t* count) {
    const int rx = sp.rx, ry = sp.ry, rz = sp.rz;
    const float h2 = h * h;
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t i = 0; i < (int64_t)n; ++i) {
        const float px = pos[3 * i], py = pos[3 * i + 1], pz = pos[3 * i + 2];
        const int cx = trunc_q(cell_q(px, sp.minx, sp.cell), rx);
        const int cy = trunc_q(cell_q(py, sp.miny, sp.cell), ry);
        const int cz = trunc_q(cell_q(pz, sp.minz, sp.cell), rz);
        uint32_t cnt = 0;
        for (int zz = clampi(cz - 1, 0, rz - 1); zz <= clampi(cz + 1, 0, rz - 1); ++zz)
            for (int yy = clampi(cy - 1, 0, ry - 1); yy <= clampi(cy + 1, 0, ry - 1); ++yy)
                for (int xx = clampi(cx - 1, 0, rx - 1); xx <= clampi(cx + 1, 0, rx - 1); ++xx) {
                    const uint64_t c = ((uint64_t)zz * ry + yy) * rx + xx;
                    if (ranges[2 * c] == 0xFFFFFFFFu) continue;
                    for (uint32_t j = ranges[2 * c]; j < ranges[2 * c + 1]; ++j) {
                        const float fx = px - pos[3 * j], fy = py - pos[3 * j + 1],
                                    fz = pz - pos[3 * j + 2];
                        const float d2 = fmaf(fz, fz, fmaf(fy, fy, fx * fx));
                        if (d2 > 0.0f && d2 < h2) ++cnt;
                    }
                }
        flags[i] = cnt < 25u ? 1u : 0u;
        if (count) count[i] = cnt;
    }
    return 0;
}

// estimationOfSurfaceVertices (ReconstructionCUDA.cu:554-595): box [i-3, i+3)
// x [j-3, j+3] x [k-3, k+3] clamped to the grid, FP32 |v - p| compared in double
// with 1.5 h.
int orc_band(const float* pos, const uint32_t* flags, uint32_t n, orc_grid sc, float h,
             uint8_t* band) {
    const int rx = sc.rx, ry = sc.ry, rz = sc.rz;
    const uint64_t V = (uint64_t)rx * ry * rz;
    std::memset(band, 0, V);
    const double limit = (double)h * 1.5;
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t i = 0; i < (int64_t)n; ++i) {
        if (!flags[i]) continue;
        const float px = pos[3 * i], py = pos[3 * i + 1], pz = pos[3 * i + 2];
        const int ix = trunc_q(cell_q(px, sc.minx, sc.cell), rx);
        const int iy = trunc_q(cell_q(py, sc.miny, sc.cell), ry);
        const int iz = trunc_q(cell_q(pz, sc.minz, sc.cell), rz);
        const int x0 = clampi(ix - 3, 0, rx - 1), x1 = clampi(ix + 3, 0, rx - 1);
        const int y0 = clampi(iy - 3, 0, ry - 1), y1 = clampi(iy + 3, 0, ry - 1);
        const int z0 = clampi(iz - 3, 0, rz - 1), z1 = clampi(iz + 3, 0, rz - 1);
        for (int z = z0; z <= z1; ++z)
            for (int y = y0; y <= y1; ++y)
                for (int x = x0; x < x1; ++x) {
                    const float dx = vpos(x, sc.minx, sc.cell) - px;
                    const float dy = vpos(y, sc.miny, sc.cell) - py;
                    const float dz = vpos(z, sc.minz, sc.cell) - pz;
                    const float dist = sqrtf(fmaf(dz, dz, fmaf(dy, dy, dx * dx)));
                    if ((double)dist < limit) {
#pragma omp atomic write
                        band[((uint64_t)z * ry + y) * rx + x] = 1;
                    }
                }
    }
    return 0;
}

uint32_t orc_compact(const uint8_t* band, uint64_t V, uint32_t* index) {
    uint32_t b = 0;
    for (uint64_t v = 0; v < V; ++v)
        if (band[v]) index[b++] = (uint32_t)v;
    return b;
}

}  // extern "C"

namespace {
// updateScalarFieldValueZB05 (ReconstructionCUDA.cu:230-289) with kernelZB05
// (ReconstructionCUDA.cuh:262-268), in FP64.  The vertex position and its hash
// cell are the FP32 values the kernel computes.  Positions are FP32 (the
// frame's) or FP64 (the adjoint's perturbed ones).
template <typename P>
void field_zb05_impl(const uint32_t* band_index, uint32_t B, const P* pos,
                     const uint32_t* ranges, orc_grid sp, orc_grid sc, float radius_sq,
                     float particle_radius, double* phi, double* wsum) {
    const int rx = sp.rx, ry = sp.ry, rz = sp.rz;
    const double R2 = radius_sq;
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t b = 0; b < (int64_t)B; ++b) {
        const uint32_t vid = band_index[b];
        const uint32_t ix = vid % sc.rx, iy = (vid / sc.rx) % sc.ry,
                       iz = vid / (sc.rx * sc.ry);
        const float vxf = vpos(ix, sc.minx, sc.cell), vyf = vpos(iy, sc.miny, sc.cell),
                    vzf = vpos(iz, sc.minz, sc.cell);
        const int cx = trunc_q(cell_q(vxf, sp.minx, sp.cell), rx);
        const int cy = trunc_q(cell_q(vyf, sp.miny, sp.cell), ry);
        const int cz = trunc_q(cell_q(vzf, sp.minz, sp.cell), rz);
        const double vx = vxf, vy = vyf, vz = vzf;
        double sw = 0, sx = 0, sy = 0, sz = 0;
        for (int zz = clampi(cz - 1, 0, rz - 1); zz <= clampi(cz + 1, 0, rz - 1); ++zz)
            for (int yy = clampi(cy - 1, 0, ry - 1); yy <= clampi(cy + 1, 0, ry - 1); ++yy)
                for (int xx = clampi(cx - 1, 0, rx - 1); xx <= clampi(cx + 1, 0, rx - 1); ++xx) {
                    const uint64_t c = ((uint64_t)zz * ry + yy) * rx + xx;
                    if (ranges[2 * c] == 0xFFFFFFFFu) continue;
                    for (uint32_t j = ranges[2 * c]; j < ranges[2 * c + 1]; ++j) {
                        const double px = pos[3 * j], py = pos[3 * j + 1], pz = pos[3 * j + 2];
                        const double dx = vx - px, dy = vy - py, dz = vz - pz;
                        const double d2 = dx * dx + dy * dy + dz * dz;
                        if (d2 >= R2) continue;
                        const double t = 1.0 - d2 / R2;
                        const double w = t * t * t;
                        sw += w;
                        sx += w * px;
                        sy += w * py;
                        sz += w * pz;
                    }
                }
        if (wsum) wsum[b] = sw;
        if (sw > 0.0) {
            const double ax = vx - sx / sw, ay = vy - sy / sw, az = vz - sz / sw;
            phi[vid] = std::sqrt(ax * ax + ay * ay + az * az) - (double)particle_radius;
        }
    }
}
}  // namespace

extern "C" {

int orc_field_zb05(const uint32_t* band_index, uint32_t B, const float* pos,
                   const uint32_t* ranges, orc_grid sp, orc_grid sc, float radius_sq,
                   float particle_radius, double* phi, double* wsum) {
    field_zb05_impl(band_index, B, pos, ranges, sp, sc, radius_sq, particle_radius, phi, wsum);
    return 0;
}

// Corner normal of the reference's marching cubes: central differences
// normalised as helper_math's normalize, v * rsqrtf(dot(v, v))
// (getVertexNorm, ReconstructionCUDA.cuh:419-430; NaN for a zero vector), or,
// for the host flavour, as fVector3::normalize (GPU/MathVector.cpp:5-13: only
// above length 1e-7).  Both square in FP32: a squared length above FLT_MAX is
// +inf there (a central difference that reads the non-band fill 0x64646464,
// about 1.7e22, makes one), and v * rsqrt(inf) = v / inf is the zero vector.
static void normalise_corner(double n[3], int cpu) {
    const double d = n[0] * n[0] + n[1] * n[1] + n[2] * n[2];
    if (d > (double)FLT_MAX) {
        n[0] = n[1] = n[2] = 0.0;
        return;
    }
    const double len = std::sqrt(d);
    if (cpu) {
        if (len > 1e-7) for (int a = 0; a < 3; ++a) n[a] /= len;
    } else {
        const double r = 1.0 / len;
        for (int a = 0; a < 3; ++a) n[a] *= r;
    }
}

// detectValidSurfaceCubes + compactValidSurfaceCubes + generateTriangles
// (ReconstructionCUDA.cu:328-463, .cuh:366-430, 534-609); cpu = 1 follows the
// host marching cubes of GPU/MarchingCubesCPU.cpp:3-89 instead (no normal
// negation, normals normalised only above length 1e-7, positions min + i*cell).
uint64_t orc_mc(const uint32_t* band_index, uint32_t B, const uint8_t* band, const double* phi,
                orc_grid sc, float iso, int cpu, uint32_t* nverts, double* tri_pos,
                double* tri_nrm, uint64_t capacity, double* nrm_len) {
    const int rx = sc.rx, ry = sc.ry, rz = sc.rz;
    const uint64_t rxy = (uint64_t)rx * ry;
    static const int off[8][3] = {{0, 0, 0}, {0, 0, 1}, {0, 1, 1}, {0, 1, 0},
                                  {1, 0, 0}, {1, 0, 1}, {1, 1, 1}, {1, 1, 0}};
    static const int edge[12][2] = {{0, 1}, {1, 2}, {2, 3}, {3, 0}, {4, 5}, {5, 6},
                                    {6, 7}, {7, 4}, {0, 4}, {1, 5}, {2, 6}, {3, 7}};
    const double isod = iso;
#pragma omp parallel for schedule(static)
    for (int64_t b = 0; b < (int64_t)B; ++b) {
        const uint32_t vid = band_index[b];
        const int x = (int)(vid % rx), y = (int)((vid / rx) % ry), z = (int)(vid / rxy);
        uint32_t nv = 0;
        if (x < rx - 1 && y < ry - 1 && z < rz - 1) {
            uint32_t f = 0;
            bool all = true;
            for (int c = 0; c < 8; ++c) {
                const uint64_t v = ((uint64_t)(z + off[c][2]) * ry + (y + off[c][1])) * rx +
                                   (x + off[c][0]);
                if (phi[v] <= isod) f |= 1u << c;
                if (!band[v]) all = false;
            }
            nv = all ? case_nverts(f) : 0;
        }
        nverts[b] = nv;
    }
    std::vector<uint64_t> offs(B + 1, 0);
    for (uint32_t b = 0; b < B; ++b) offs[b + 1] = offs[b] + nverts[b];
    const uint64_t total = offs[B];
    if (!tri_pos || total > capacity) return total;
    auto sample = [&](int X, int Y, int Z) -> double {
        if (X < 0 || Y < 0 || Z < 0 || X >= rx || Y >= ry || Z >= rz) return 0.0;
        return phi[((uint64_t)Z * ry + Y) * rx + X];
    };
    const double sign = (!cpu && iso < 0.0f) ? -1.0 : 1.0;
#pragma omp parallel for schedule(static)
    for (int64_t b = 0; b < (int64_t)B; ++b) {
        if (!nverts[b]) continue;
        const uint32_t vid = band_index[b];
        const int x = (int)(vid % rx), y = (int)((vid / rx) % ry), z = (int)(vid / rxy);
        double val[8], cp[8][3], cn[8][3];
        uint32_t f = 0;
        for (int c = 0; c < 8; ++c) {
            const int X = x + off[c][0], Y = y + off[c][1], Z = z + off[c][2];
            val[c] = phi[((uint64_t)Z * ry + Y) * rx + X];
            if (val[c] <= isod) f |= 1u << c;
            if (cpu) {
                cp[c][0] = sc.minx + (float)X * sc.cell;
                cp[c][1] = sc.miny + (float)Y * sc.cell;
                cp[c][2] = sc.minz + (float)Z * sc.cell;
            } else {
                cp[c][0] = vpos(X, sc.minx, sc.cell);
                cp[c][1] = vpos(Y, sc.miny, sc.cell);
                cp[c][2] = vpos(Z, sc.minz, sc.cell);
            }
            cn[c][0] = sample(X - 1, Y, Z) - sample(X + 1, Y, Z);
            cn[c][1] = sample(X, Y - 1, Z) - sample(X, Y + 1, Z);
            cn[c][2] = sample(X, Y, Z - 1) - sample(X, Y, Z + 1);
            normalise_corner(cn[c], cpu);
        }
        const uint32_t nv = nverts[b];
        for (uint32_t k = 0; k < nv; ++k) {
            const int e = (int)((kTri[f] >> (4 * k)) & 0xF);
            const int s0 = edge[e][0], s1 = edge[e][1];
            const double delta = val[s1] - val[s0];
            const double t = (delta > -1e-7 && delta < 1e-7) ? 0.5 : (isod - val[s0]) / delta;
            double p[3], nn[3];
            for (int a = 0; a < 3; ++a) {
                p[a] = cp[s0][a] + t * (cp[s1][a] - cp[s0][a]);
                nn[a] = cn[s0][a] + t * (cn[s1][a] - cn[s0][a]);
            }
            const double len = std::sqrt(nn[0] * nn[0] + nn[1] * nn[1] + nn[2] * nn[2]);
            // length of the interpolated normal before its normalisation: the
            // normalised direction's sensitivity to rounding is 1 / len
            if (nrm_len) nrm_len[offs[b] + k] = len;
            if (cpu) {
                if (len > 1e-7) for (int a = 0; a < 3; ++a) nn[a] /= len;
            } else {
                const double r = 1.0 / len;
                for (int a = 0; a < 3; ++a) nn[a] = sign * (nn[a] * r);
            }
            const uint64_t o = 3 * (offs[b] + k);
            for (int a = 0; a < 3; ++a) {
                tri_pos[o + a] = p[a];
                if (tri_nrm) tri_nrm[o + a] = nn[a];
            }
        }
    }
    return total;
}

// ------------------------------------------------------------- Yu13 ---------

// calculationOfMeanAndSmoothParticles (ReconstructionCUDA.cu:598-658) with wij
// (ReconstructionCUDA.cuh:282-292), FP64: clamped 5^3 block of the particle's
// cell, w = 1 - (d/r)^3 for d < r; mean = sum w p / sum w; smoothed =
// (1 - lambda) x + lambda mean.
int orc_smooth(const float* pos, uint32_t n, const uint32_t* ranges, orc_grid sp, float radius,
               float lambda, double* mean, double* smooth) {
    const int rx = sp.rx, ry = sp.ry, rz = sp.rz;
    const double r = radius, lam = lambda;
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t i = 0; i < (int64_t)n; ++i) {
        const float px = pos[3 * i], py = pos[3 * i + 1], pz = pos[3 * i + 2];
        const int cx = trunc_q(cell_q(px, sp.minx, sp.cell), rx);
        const int cy = trunc_q(cell_q(py, sp.miny, sp.cell), ry);
        const int cz = trunc_q(cell_q(pz, sp.minz, sp.cell), rz);
        double sw = 0, sx = 0, sy = 0, sz = 0;
        for (int zz = clampi(cz - 2, 0, rz - 1); zz <= clampi(cz + 2, 0, rz - 1); ++zz)
            for (int yy = clampi(cy - 2, 0, ry - 1); yy <= clampi(cy + 2, 0, ry - 1); ++yy)
                for (int xx = clampi(cx - 2, 0, rx - 1); xx <= clampi(cx + 2, 0, rx - 1); ++xx) {
                    const uint64_t c = ((uint64_t)zz * ry + yy) * rx + xx;
                    if (ranges[2 * c] == 0xFFFFFFFFu) continue;
                    for (uint32_t j = ranges[2 * c]; j < ranges[2 * c + 1]; ++j) {
                        const double qx = pos[3 * j], qy = pos[3 * j + 1], qz = pos[3 * j + 2];
                        const double dx = px - qx, dy = py - qy, dz = pz - qz;
                        const double d = std::sqrt(dx * dx + dy * dy + dz * dz);
                        if (d >= r) continue;
                        const double s = d / r;
                        const double w = 1.0 - s * s * s;
                        sw += w;
                        sx += w * qx;
                        sy += w * qy;
                        sz += w * qz;
                    }
                }
        const double m[3] = {sx / sw, sy / sw, sz / sw};
        const double x[3] = {px, py, pz};
        for (int a = 0; a < 3; ++a) {
            mean[3 * i + a] = m[a];
            smooth[3 * i + a] = (1.0 - lam) * x[a] + lam * m[a];
        }
    }
    return 0;
}

// Cyclic Jacobi eigen-decomposition of a symmetric 3x3 matrix in FP64
// (Golub & Van Loan, Matrix Computations, sec. 8.5), iterated to convergence.
// A is overwritten; d = eigenvalues, V columns = eigenvectors.
static void jacobi_eigen3(double A[3][3], double d[3], double V[3][3]) {
    for (int p = 0; p < 3; ++p)
        for (int q = 0; q < 3; ++q) V[p][q] = p == q ? 1.0 : 0.0;
    for (int sweep = 0; sweep < 50; ++sweep) {
        const double off = A[0][1] * A[0][1] + A[0][2] * A[0][2] + A[1][2] * A[1][2];
        const double dia = A[0][0] * A[0][0] + A[1][1] * A[1][1] + A[2][2] * A[2][2];
        if (off <= 1e-34 * dia) break;
        for (int p = 0; p < 2; ++p)
            for (int q = p + 1; q < 3; ++q) {
                if (A[p][q] == 0.0) continue;
                const double theta = (A[q][q] - A[p][p]) / (2.0 * A[p][q]);
                const double t = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
                const double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
                for (int k = 0; k < 3; ++k) {  // A <- A J (columns p, q)
                    const double akp = A[k][p], akq = A[k][q];
                    A[k][p] = c * akp - s * akq;
                    A[k][q] = s * akp + c * akq;
                }
                for (int k = 0; k < 3; ++k) {  // A <- J^T A (rows p, q)
                    const double apk = A[p][k], aqk = A[q][k];
                    A[p][k] = c * apk - s * aqk;
                    A[q][k] = s * apk + c * aqk;
                }
                for (int k = 0; k < 3; ++k) {
                    const double vkp = V[k][p], vkq = V[k][q];
                    V[k][p] = c * vkp - s * vkq;
                    V[k][q] = s * vkp + c * vkq;
                }
            }
    }
    for (int a = 0; a < 3; ++a) d[a] = A[a][a];
}

int orc_eigen3(const double* a, double* d, double* v) {
    double A[3][3], V[3][3];
    for (int p = 0; p < 3; ++p)
        for (int q = 0; q < 3; ++q) A[p][q] = a[3 * p + q];
    jacobi_eigen3(A, d, V);
    for (int p = 0; p < 3; ++p)
        for (int q = 0; q < 3; ++q) v[3 * p + q] = V[p][q];
    return 0;
}

// calculationOfTransformMatricesForParticles (ReconstructionCUDA.cu:689-827) in
// FP64 for the involved particles `index[I]` (sorted ids): G (9 per particle,
// row-major), singular values descending (3), and the neighbour count bounds
// count_lo/hi = #{|v| <= r (1 -/+ 1e-5)}, which flag particles whose FP32 count
// could fall on the other side of Ne.
int orc_aniso_matrices(const float* pos, const double* mean, const float* dens, uint32_t I,
                       const uint32_t* index, const uint32_t* ranges, orc_grid sp, float h,
                       float radius, float kr, float kn, uint32_t min_neighbors, double* G,
                       double* sigma, uint32_t* count_lo, uint32_t* count_hi) {
    (void)dens;
    const int rx = sp.rx, ry = sp.ry, rz = sp.rz;
    const double r = radius, H = h;
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t k = 0; k < (int64_t)I; ++k) {
        const uint32_t i = index[k];
        const float px = pos[3 * i], py = pos[3 * i + 1], pz = pos[3 * i + 2];
        const int cx = trunc_q(cell_q(px, sp.minx, sp.cell), rx);
        const int cy = trunc_q(cell_q(py, sp.miny, sp.cell), ry);
        const int cz = trunc_q(cell_q(pz, sp.minz, sp.cell), rz);
        const double m[3] = {mean[3 * i], mean[3 * i + 1], mean[3 * i + 2]};
        double C[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
        double sw = 0;
        uint32_t cnt = 0, lo = 0, hi = 0;
        for (int zz = clampi(cz - 2, 0, rz - 1); zz <= clampi(cz + 2, 0, rz - 1); ++zz)
            for (int yy = clampi(cy - 2, 0, ry - 1); yy <= clampi(cy + 2, 0, ry - 1); ++yy)
                for (int xx = clampi(cx - 2, 0, rx - 1); xx <= clampi(cx + 2, 0, rx - 1); ++xx) {
                    const uint64_t c = ((uint64_t)zz * ry + yy) * rx + xx;
                    if (ranges[2 * c] == 0xFFFFFFFFu) continue;
                    for (uint32_t j = ranges[2 * c]; j < ranges[2 * c + 1]; ++j) {
                        const double v[3] = {pos[3 * j] - m[0], pos[3 * j + 1] - m[1],
                                             pos[3 * j + 2] - m[2]};
                        const double d = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
                        cnt += d <= r;
                        lo += d <= r * (1.0 - 1e-5);
                        hi += d <= r * (1.0 + 1e-5);
                        if (d >= r) continue;
                        const double s = d / r;
                        const double w = 1.0 - s * s * s;
                        sw += w;
                        for (int p = 0; p < 3; ++p)
                            for (int q = 0; q < 3; ++q) C[p][q] += w * v[p] * v[q];
                    }
                }
        count_lo[k] = lo;
        count_hi[k] = hi;
        double g[3][3];
        if (cnt < min_neighbors) {
            for (int p = 0; p < 3; ++p)
                for (int q = 0; q < 3; ++q) g[p][q] = p == q ? 1.0 / ((double)kn * H) : 0.0;
            sigma[3 * k] = sigma[3 * k + 1] = sigma[3 * k + 2] = 0.0;
        } else {
            for (int p = 0; p < 3; ++p) {
                C[p][p] += (double)(float)(h * h);
                for (int q = 0; q < 3; ++q) C[p][q] /= sw;
            }
            double d[3], V[3][3];
            jacobi_eigen3(C, d, V);
            double s[3] = {std::fabs(d[0]), std::fabs(d[1]), std::fabs(d[2])};
            double srt[3] = {s[0], s[1], s[2]};
            std::sort(srt, srt + 3);
            sigma[3 * k] = srt[2];
            sigma[3 * k + 1] = srt[1];
            sigma[3 * k + 2] = srt[0];
            const double fl = srt[2] / kr;
            for (int a = 0; a < 3; ++a) s[a] = std::max(s[a], fl);
            const double scale = std::cbrt(s[0] * s[1] * s[2]) / H;
            for (int p = 0; p < 3; ++p)
                for (int q = 0; q < 3; ++q) {
                    double acc = 0;
                    for (int a = 0; a < 3; ++a) acc += V[p][a] * (scale / s[a]) * V[q][a];
                    g[p][q] = acc;
                }
        }
        for (int p = 0; p < 3; ++p)
            for (int q = 0; q < 3; ++q) G[9 * k + 3 * p + q] = g[p][q];
    }
    return 0;
}

// computationOfScalarFieldGrid (ReconstructionCUDA.cu:830-906) with
// anisotropicW / anisotropicWOpt (ReconstructionCUDA.cuh:295-322), FP64.
// G is indexed by involve_off[c] + (j - start_c) for involved cells.
// ambiguous[N] (may be NULL): vertices that receive weight from a flagged
// particle get flag_out[b] = 1.
int orc_field_aniso(const uint32_t* band_index, uint32_t B, const double* smooth,
                    const float* dens, const uint32_t* ranges, const uint32_t* involve,
                    const uint32_t* involve_off, const double* G, orc_grid sp, orc_grid sc,
                    float mass, const uint8_t* ambiguous, double* phi, uint8_t* flag_out) {
    const int rx = sp.rx, ry = sp.ry, rz = sp.rz;
    const double sigma = 315.0 / (64.0 * M_PI);
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t b = 0; b < (int64_t)B; ++b) {
        const uint32_t vid = band_index[b];
        const uint32_t ix = vid % sc.rx, iy = (vid / sc.rx) % sc.ry, iz = vid / (sc.rx * sc.ry);
        const float vxf = vpos(ix, sc.minx, sc.cell), vyf = vpos(iy, sc.miny, sc.cell),
                    vzf = vpos(iz, sc.minz, sc.cell);
        const int cx = trunc_q(cell_q(vxf, sp.minx, sp.cell), rx);
        const int cy = trunc_q(cell_q(vyf, sp.miny, sp.cell), ry);
        const int cz = trunc_q(cell_q(vzf, sp.minz, sp.cell), rz);
        const double v[3] = {vxf, vyf, vzf};
        double sum = 0;
        uint8_t flag = 0;
        for (int zz = clampi(cz - 2, 0, rz - 1); zz <= clampi(cz + 2, 0, rz - 1); ++zz)
            for (int yy = clampi(cy - 2, 0, ry - 1); yy <= clampi(cy + 2, 0, ry - 1); ++yy)
                for (int xx = clampi(cx - 2, 0, rx - 1); xx <= clampi(cx + 2, 0, rx - 1); ++xx) {
                    const uint64_t c = ((uint64_t)zz * ry + yy) * rx + xx;
                    if (ranges[2 * c] == 0xFFFFFFFFu) continue;
                    const uint32_t s0 = ranges[2 * c], e0 = ranges[2 * c + 1];
                    for (uint32_t j = s0; j < e0; ++j) {
                        const double r[3] = {smooth[3 * j] - v[0], smooth[3 * j + 1] - v[1],
                                             smooth[3 * j + 2] - v[2]};
                        const double mr = (double)mass / (double)dens[j];
                        double d2, det = 1.0;
                        if (involve[c] > 0) {
                            const double* g = G + 9ull * (involve_off[c] + (j - s0));
                            double t[3];
                            for (int p = 0; p < 3; ++p)
                                t[p] = g[3 * p] * r[0] + g[3 * p + 1] * r[1] + g[3 * p + 2] * r[2];
                            d2 = t[0] * t[0] + t[1] * t[1] + t[2] * t[2];
                            det = g[0] * (g[4] * g[8] - g[5] * g[7]) -
                                  g[1] * (g[3] * g[8] - g[5] * g[6]) +
                                  g[2] * (g[3] * g[7] - g[4] * g[6]);
                        } else {
                            d2 = r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
                        }
                        if (d2 >= 1.0) continue;
                        const double x = 1.0 - d2;
                        sum += mr * sigma * det * x * x * x;
                        if (ambiguous && ambiguous[j]) flag = 1;
                    }
                }
        phi[vid] = sum - 0.5;
        if (flag_out) flag_out[b] = flag;
    }
    return 0;
}

// One frame of the ZB05 (method 0) or Akinci12 (method 1) reconstructor
// (GPU/ReconstructorGPUOursZB05.cpp:22-78, GPU/ReconstructorGPUAkinci12.cpp:34-84);
// method | 2: surface particles by the radius-count test instead.
int orc_frame(int method, float* pos, float* dens, uint32_t n, orc_grid sp, orc_grid sc, float h,
              float radius_sq, float particle_radius, float mass, float iso, float* phi_out,
              uint64_t* counts) {
    const uint64_t C = (uint64_t)sp.rx * sp.ry * sp.rz;
    const uint64_t V = (uint64_t)sc.rx * sc.ry * sc.rz;
    std::vector<uint32_t> keys(n), ranges(2 * C), flags(n);
    std::vector<float> occ(C), pos_s(3ull * n), dens_s(n);
    orc_hash_sort(pos, dens, n, sp, keys.data(), nullptr, pos_s.data(), dens_s.data(),
                  ranges.data(), occ.data());
    std::memcpy(pos, pos_s.data(), sizeof(float) * 3ull * n);
    std::memcpy(dens, dens_s.data(), sizeof(float) * n);
    const bool akinci = (method & 1) != 0;
    if (method & 2) {
        orc_surface_radius_count(pos, n, ranges.data(), sp, h, flags.data(), nullptr);
    } else if (akinci) {
        orc_surface_colorfield(pos, dens, n, ranges.data(), sp, h, mass, flags.data(), nullptr,
                               nullptr);
    } else {
        std::vector<uint8_t> surf(C);
        orc_surface_occupancy(occ.data(), ranges.data(), sp, keys.data(), n, surf.data(),
                              flags.data(), nullptr);
    }
    uint64_t nsurf = 0, nout = 0;
    for (uint32_t i = 0; i < n; ++i) {
        nsurf += flags[i];
        nout += keys[i] == C;
    }
    std::vector<uint8_t> band(V);
    orc_band(pos, flags.data(), n, sc, h, band.data());
    std::vector<uint32_t> index(V);
    const uint32_t B = orc_compact(band.data(), V, index.data());
    std::vector<double> phi(V, 0.0);
    orc_field_zb05(index.data(), B, pos, ranges.data(), sp, sc, radius_sq, particle_radius,
                   phi.data(), nullptr);
    std::vector<uint32_t> nverts(B);
    const uint64_t total = orc_mc(index.data(), B, band.data(), phi.data(), sc, iso, akinci,
                                  nverts.data(), nullptr, nullptr, 0, nullptr);
    std::vector<double> tp(3 * total), tn(3 * total);
    orc_mc(index.data(), B, band.data(), phi.data(), sc, iso, akinci, nverts.data(),
           tp.data(), tn.data(), total, nullptr);
    uint64_t q = 0;
    for (uint32_t b = 0; b < B; ++b) q += nverts[b] > 0;
    if (phi_out)
        for (uint64_t v = 0; v < V; ++v) phi_out[v] = (float)phi[v];
    if (counts) {
        counts[0] = nsurf;
        counts[1] = B;
        counts[2] = q;
        counts[3] = total;
        counts[4] = nout;
        counts[5] = 0;
    }
    return 0;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Adjoint of the ZB05 field and of the marching-cubes vertex lerp, topology
// held fixed.  The reference has no backward pass: these are the derivatives of
// the FP64 restatements above, written out by hand, plus the loss of the
// adjoint workload with incremental central differences that pin them.
// ---------------------------------------------------------------------------
namespace {

const int kCornerOff[8][3] = {{0, 0, 0}, {0, 0, 1}, {0, 1, 1}, {0, 1, 0},
                              {1, 0, 0}, {1, 0, 1}, {1, 1, 1}, {1, 1, 0}};
const int kEdgeCorner[12][2] = {{0, 1}, {1, 2}, {2, 3}, {3, 0}, {4, 5}, {5, 6},
                                {6, 7}, {7, 4}, {0, 4}, {1, 5}, {2, 6}, {3, 7}};

// Corner position as the frame's marching cubes evaluates it (orc_mc).
inline double corner_pos(int cpu, uint32_t X, float mn, float cell) {
    return cpu ? (double)(mn + (float)X * cell) : (double)vpos(X, mn, cell);
}

// One band vertex of the ZB05 field: sum w, sum w p (FP64), same candidate
// cells as field_zb05_impl.
template <typename P>
void zb05_sums(const P* pos, const uint32_t* ranges, orc_grid sp, orc_grid sc, double R2,
               uint32_t vid, double x[3], double* sw, double s[3]) {
    const int rx = sp.rx, ry = sp.ry, rz = sp.rz;
    const uint32_t ix = vid % sc.rx, iy = (vid / sc.rx) % sc.ry, iz = vid / (sc.rx * sc.ry);
    const float vxf = vpos(ix, sc.minx, sc.cell), vyf = vpos(iy, sc.miny, sc.cell),
                vzf = vpos(iz, sc.minz, sc.cell);
    const int cx = trunc_q(cell_q(vxf, sp.minx, sp.cell), rx);
    const int cy = trunc_q(cell_q(vyf, sp.miny, sp.cell), ry);
    const int cz = trunc_q(cell_q(vzf, sp.minz, sp.cell), rz);
    x[0] = vxf; x[1] = vyf; x[2] = vzf;
    *sw = 0; s[0] = s[1] = s[2] = 0;
    for (int zz = clampi(cz - 1, 0, rz - 1); zz <= clampi(cz + 1, 0, rz - 1); ++zz)
        for (int yy = clampi(cy - 1, 0, ry - 1); yy <= clampi(cy + 1, 0, ry - 1); ++yy)
            for (int xx = clampi(cx - 1, 0, rx - 1); xx <= clampi(cx + 1, 0, rx - 1); ++xx) {
                const uint64_t c = ((uint64_t)zz * ry + yy) * rx + xx;
                if (ranges[2 * c] == 0xFFFFFFFFu) continue;
                for (uint32_t j = ranges[2 * c]; j < ranges[2 * c + 1]; ++j) {
                    const double px = pos[3 * j], py = pos[3 * j + 1], pz = pos[3 * j + 2];
                    const double dx = x[0] - px, dy = x[1] - py, dz = x[2] - pz;
                    const double d2 = dx * dx + dy * dy + dz * dz;
                    if (d2 >= R2) continue;
                    const double t = 1.0 - d2 / R2;
                    const double w = t * t * t;
                    *sw += w;
                    s[0] += w * px;
                    s[1] += w * py;
                    s[2] += w * pz;
                }
            }
}

// g_pos (sorted order, += ) of sum_v g_phi[v] phi_v over the band.
template <typename P>
void field_zb05_backward_impl(const uint32_t* band_index, uint32_t B, const P* pos,
                              const uint32_t* ranges, orc_grid sp, orc_grid sc, float radius_sq,
                              const double* g_phi, double* g_pos) {
    const int rx = sp.rx, ry = sp.ry, rz = sp.rz;
    const double R2 = radius_sq;
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t b = 0; b < (int64_t)B; ++b) {
        const uint32_t vid = band_index[b];
        const double g = g_phi[vid];
        if (g == 0.0) continue;
        double x[3], sw, s[3];
        zb05_sums(pos, ranges, sp, sc, R2, vid, x, &sw, s);
        if (!(sw > 0.0)) continue;
        const double xb[3] = {s[0] / sw, s[1] / sw, s[2] / sw};
        const double d[3] = {x[0] - xb[0], x[1] - xb[1], x[2] - xb[2]};
        const double len = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
        if (!(len > 0.0)) continue;
        const double nrm[3] = {d[0] / len, d[1] / len, d[2] / len};
        const double gw = g / sw;
        const int cx = trunc_q(cell_q((float)x[0], sp.minx, sp.cell), rx);
        const int cy = trunc_q(cell_q((float)x[1], sp.miny, sp.cell), ry);
        const int cz = trunc_q(cell_q((float)x[2], sp.minz, sp.cell), rz);
        for (int zz = clampi(cz - 1, 0, rz - 1); zz <= clampi(cz + 1, 0, rz - 1); ++zz)
            for (int yy = clampi(cy - 1, 0, ry - 1); yy <= clampi(cy + 1, 0, ry - 1); ++yy)
                for (int xx = clampi(cx - 1, 0, rx - 1); xx <= clampi(cx + 1, 0, rx - 1); ++xx) {
                    const uint64_t c = ((uint64_t)zz * ry + yy) * rx + xx;
                    if (ranges[2 * c] == 0xFFFFFFFFu) continue;
                    for (uint32_t j = ranges[2 * c]; j < ranges[2 * c + 1]; ++j) {
                        const double p[3] = {(double)pos[3 * j], (double)pos[3 * j + 1],
                                             (double)pos[3 * j + 2]};
                        const double e[3] = {x[0] - p[0], x[1] - p[1], x[2] - p[2]};
                        const double d2 = e[0] * e[0] + e[1] * e[1] + e[2] * e[2];
                        if (d2 >= R2) continue;
                        const double t = 1.0 - d2 / R2;
                        const double w = t * t * t;
                        // d phi / d p_j = -(1/W)[w n + (n.(p - xbar)) (6/R^2) t^2 (x - p)]
                        const double m = nrm[0] * (p[0] - xb[0]) + nrm[1] * (p[1] - xb[1]) +
                                         nrm[2] * (p[2] - xb[2]);
                        const double k = m * 6.0 / R2 * t * t;
                        for (int a = 0; a < 3; ++a) {
                            const double v = -gw * (w * nrm[a] + k * e[a]);
#pragma omp atomic
                            g_pos[3 * j + a] += v;
                        }
                    }
                }
    }
}

// Soup vertices of the marching cubes over `phi` in soup order (cube band
// position, then table order), with the per-cube capacity rule of the device
// emitter: a = grid id of corner s0, b = of corner s1 of each vertex's edge.
struct Soup {
    std::vector<uint32_t> a, b;
    std::vector<double> ca, dc;  // corner s0 position, corner s1 - corner s0
};

uint64_t soup_edges(const uint32_t* band_index, uint32_t B, const uint8_t* band,
                    const double* phi, orc_grid sc, float iso, int cpu, uint64_t capacity,
                    Soup* out) {
    const uint64_t rx = sc.rx, ry = sc.ry, rz = sc.rz, rxy = rx * ry;
    std::vector<uint32_t> nverts(B);
    orc_mc(band_index, B, band, phi, sc, iso, cpu, nverts.data(), nullptr, nullptr, 0, nullptr);
    uint64_t off = 0;
    for (uint32_t bb = 0; bb < B; ++bb) {
        const uint32_t nv = nverts[bb];
        if (!nv) continue;
        const uint64_t first = off;
        off += nv;
        if (first + nv > capacity) continue;
        const uint32_t vid = band_index[bb];
        const uint32_t x = vid % rx, y = (vid / rx) % ry, z = (uint32_t)(vid / rxy);
        (void)rz;
        uint32_t f = 0;
        for (int c = 0; c < 8; ++c) {
            const uint64_t v = ((uint64_t)(z + kCornerOff[c][2]) * ry + (y + kCornerOff[c][1])) * rx +
                               (x + kCornerOff[c][0]);
            if (phi[v] <= (double)iso) f |= 1u << c;
        }
        for (uint32_t k = 0; k < nv; ++k) {
            const int e = (int)((kTri[f] >> (4 * k)) & 0xF);
            uint32_t id[2];
            double cp[2][3];
            for (int q = 0; q < 2; ++q) {
                const int c = kEdgeCorner[e][q];
                const uint32_t X = x + kCornerOff[c][0], Y = y + kCornerOff[c][1],
                               Z = z + kCornerOff[c][2];
                id[q] = (uint32_t)(((uint64_t)Z * ry + Y) * rx + X);
                cp[q][0] = corner_pos(cpu, X, sc.minx, sc.cell);
                cp[q][1] = corner_pos(cpu, Y, sc.miny, sc.cell);
                cp[q][2] = corner_pos(cpu, Z, sc.minz, sc.cell);
            }
            out->a.push_back(id[0]);
            out->b.push_back(id[1]);
            for (int t = 0; t < 3; ++t) {
                out->ca.push_back(cp[0][t]);
                out->dc.push_back(cp[1][t] - cp[0][t]);
            }
        }
    }
    return off;
}

inline double lerp_t(double pa, double pb, double iso) {
    const double delta = pb - pa;
    return (delta > -1e-7 && delta < 1e-7) ? 0.5 : (iso - pa) / delta;
}

}  // namespace

extern "C" {

int orc_mc_vertex_backward(const uint32_t* band_index, uint32_t B, const uint8_t* band,
                           const double* phi, orc_grid sc, float iso, int cpu,
                           const double* g_tri, uint64_t capacity, double* g_phi) {
    Soup soup;
    soup_edges(band_index, B, band, phi, sc, iso, cpu, capacity, &soup);
    const double isod = iso;
    for (size_t k = 0; k < soup.a.size(); ++k) {
        const double pa = phi[soup.a[k]], pb = phi[soup.b[k]];
        const double delta = pb - pa;
        if (delta > -1e-7 && delta < 1e-7) continue;
        // the soup slot of this vertex: soup order is the emitter's
        const double gt = g_tri[3 * k] * soup.dc[3 * k] + g_tri[3 * k + 1] * soup.dc[3 * k + 1] +
                          g_tri[3 * k + 2] * soup.dc[3 * k + 2];
        g_phi[soup.a[k]] += gt * (isod - pb) / (delta * delta);
        g_phi[soup.b[k]] -= gt * (isod - pa) / (delta * delta);
    }
    return 0;
}

int orc_field_zb05_backward(const uint32_t* band_index, uint32_t B, const float* pos,
                            uint32_t n, const uint32_t* ranges, orc_grid sp, orc_grid sc,
                            float radius_sq, const double* g_phi, double* g_pos) {
    std::fill(g_pos, g_pos + 3ull * n, 0.0);
    field_zb05_backward_impl(band_index, B, pos, ranges, sp, sc, radius_sq, g_phi, g_pos);
    return 0;
}

}  // extern "C"

struct orc_adj {
    int cpu = 0;
    orc_grid sp{}, sc{};
    float radius_sq = 0.f, particle_radius = 0.f, iso = 0.f;
    double R0 = 0, wg = 0, ws = 0;
    uint32_t n = 0, B = 0;
    uint64_t V = 0, K = 0;
    std::vector<double> pos;  // hash-sorted, FP64
    std::vector<uint32_t> perm, inv, ranges, band_index;
    std::vector<uint8_t> band;
    std::vector<double> phi, sdf;
    Soup soup;
    std::vector<uint32_t> vs_off, vs_list;  // band position -> soup vertices on its edges
};

namespace {

uint32_t band_pos(const orc_adj* A, uint32_t vid) {
    return (uint32_t)(std::lower_bound(A->band_index.begin(), A->band_index.end(), vid) -
                      A->band_index.begin());
}

double phi_at(const orc_adj* A, uint32_t vid) {
    double x[3], sw, s[3];
    zb05_sums(A->pos.data(), A->ranges.data(), A->sp, A->sc, (double)A->radius_sq, vid, x, &sw,
              s);
    if (!(sw > 0.0)) return 0.0;
    const double d0 = x[0] - s[0] / sw, d1 = x[1] - s[1] / sw, d2 = x[2] - s[2] / sw;
    return std::sqrt(d0 * d0 + d1 * d1 + d2 * d2) - (double)A->particle_radius;
}

double soup_term(const orc_adj* A, uint64_t k, double pa, double pb) {
    const double t = lerp_t(pa, pb, (double)A->iso);
    double r2 = 0;
    for (int a = 0; a < 3; ++a) {
        const double y = A->soup.ca[3 * k + a] + t * A->soup.dc[3 * k + a];
        r2 += y * y;
    }
    const double d = std::sqrt(r2) - A->R0;
    return d * d;
}

}  // namespace

extern "C" {

orc_adj* orc_adj_create(int method, const float* pos, const float* dens, uint32_t n, orc_grid sp,
                        orc_grid sc, float h, float radius_sq, float particle_radius, float mass,
                        float iso, double sphere_radius, double w_grid, double w_soup) {
    orc_adj* A = new orc_adj();
    A->cpu = method == 1;
    A->sp = sp;
    A->sc = sc;
    A->radius_sq = radius_sq;
    A->particle_radius = particle_radius;
    A->iso = iso;
    A->R0 = sphere_radius;
    A->wg = w_grid;
    A->ws = w_soup;
    A->n = n;
    const uint64_t C = (uint64_t)sp.rx * sp.ry * sp.rz;
    A->V = (uint64_t)sc.rx * sc.ry * sc.rz;
    std::vector<uint32_t> keys(n), flags(n);
    std::vector<float> occ(C), ps(3ull * n), ds(n);
    A->perm.resize(n);
    A->ranges.resize(2 * C);
    orc_hash_sort(pos, dens, n, sp, keys.data(), A->perm.data(), ps.data(), ds.data(),
                  A->ranges.data(), occ.data());
    A->inv.resize(n);
    for (uint32_t i = 0; i < n; ++i) A->inv[A->perm[i]] = i;
    if (method == 1) {
        orc_surface_colorfield(ps.data(), ds.data(), n, A->ranges.data(), sp, h, mass,
                               flags.data(), nullptr, nullptr);
    } else {
        std::vector<uint8_t> surf(C);
        orc_surface_occupancy(occ.data(), A->ranges.data(), sp, keys.data(), n, surf.data(),
                              flags.data(), nullptr);
    }
    A->band.resize(A->V);
    orc_band(ps.data(), flags.data(), n, sc, h, A->band.data());
    A->band_index.resize(A->V);
    A->B = orc_compact(A->band.data(), A->V, A->band_index.data());
    A->band_index.resize(A->B);
    A->pos.assign(ps.begin(), ps.end());
    A->phi.assign(A->V, 0.0);
    field_zb05_impl(A->band_index.data(), A->B, A->pos.data(), A->ranges.data(), sp, sc,
                    radius_sq, particle_radius, A->phi.data(), nullptr);
    A->sdf.resize(A->V);
    for (uint64_t v = 0; v < A->V; ++v) {
        const uint32_t ix = v % sc.rx, iy = (v / sc.rx) % sc.ry, iz = v / ((uint64_t)sc.rx * sc.ry);
        const double x = vpos(ix, sc.minx, sc.cell), y = vpos(iy, sc.miny, sc.cell),
                     z = vpos(iz, sc.minz, sc.cell);
        A->sdf[v] = std::sqrt(x * x + y * y + z * z) - sphere_radius;
    }
    soup_edges(A->band_index.data(), A->B, A->band.data(), A->phi.data(), sc, iso, A->cpu,
               ~0ull, &A->soup);
    A->K = A->soup.a.size();
    A->vs_off.assign(A->B + 1, 0);
    for (uint64_t k = 0; k < A->K; ++k) {
        ++A->vs_off[band_pos(A, A->soup.a[k]) + 1];
        ++A->vs_off[band_pos(A, A->soup.b[k]) + 1];
    }
    for (uint32_t b = 0; b < A->B; ++b) A->vs_off[b + 1] += A->vs_off[b];
    A->vs_list.resize(A->vs_off[A->B]);
    std::vector<uint32_t> cur(A->vs_off.begin(), A->vs_off.end() - 1);
    for (uint64_t k = 0; k < A->K; ++k) {
        A->vs_list[cur[band_pos(A, A->soup.a[k])]++] = (uint32_t)k;
        A->vs_list[cur[band_pos(A, A->soup.b[k])]++] = (uint32_t)k;
    }
    return A;
}

void orc_adj_destroy(orc_adj* A) { delete A; }

// info[6] = {n, B, K (soup vertices), V, 0, 0}
void orc_adj_info(const orc_adj* A, uint64_t* info) {
    info[0] = A->n;
    info[1] = A->B;
    info[2] = A->K;
    info[3] = A->V;
    info[4] = 0;
    info[5] = 0;
}

// phi (double[V]) and soup vertex positions (double[3K]) of the base state;
// band index (u32[B]); permutation (u32[n], sorted -> caller index).
void orc_adj_state(const orc_adj* A, double* phi, double* soup_pos, uint32_t* band_index,
                   uint32_t* perm) {
    if (phi) std::copy(A->phi.begin(), A->phi.end(), phi);
    if (soup_pos)
        for (uint64_t k = 0; k < A->K; ++k) {
            const double t = lerp_t(A->phi[A->soup.a[k]], A->phi[A->soup.b[k]], (double)A->iso);
            for (int a = 0; a < 3; ++a)
                soup_pos[3 * k + a] = A->soup.ca[3 * k + a] + t * A->soup.dc[3 * k + a];
        }
    if (band_index) std::copy(A->band_index.begin(), A->band_index.end(), band_index);
    if (perm) std::copy(A->perm.begin(), A->perm.end(), perm);
}

// L = w_grid/V sum_v (phi_v - sdf_v)^2 + w_soup/K sum_k (|y_k| - R0)^2
double orc_adj_loss(const orc_adj* A) {
    double lg = 0, ls = 0;
    for (uint64_t v = 0; v < A->V; ++v) {
        const double d = A->phi[v] - A->sdf[v];
        lg += d * d;
    }
    for (uint64_t k = 0; k < A->K; ++k)
        ls += soup_term(A, k, A->phi[A->soup.a[k]], A->phi[A->soup.b[k]]);
    return A->wg * lg / (double)A->V + (A->K ? A->ws * ls / (double)A->K : 0.0);
}

// Upstream gradients of the loss: g_phi (double[V]), g_soup (double[3K]).
void orc_adj_upstream(const orc_adj* A, double* g_phi, double* g_soup) {
    for (uint64_t v = 0; v < A->V; ++v)
        g_phi[v] = 2.0 * A->wg * (A->phi[v] - A->sdf[v]) / (double)A->V;
    for (uint64_t k = 0; k < A->K; ++k) {
        const double t = lerp_t(A->phi[A->soup.a[k]], A->phi[A->soup.b[k]], (double)A->iso);
        double y[3], r2 = 0;
        for (int a = 0; a < 3; ++a) {
            y[a] = A->soup.ca[3 * k + a] + t * A->soup.dc[3 * k + a];
            r2 += y[a] * y[a];
        }
        const double r = std::sqrt(r2);
        const double c = r > 0 ? 2.0 * A->ws * (r - A->R0) / (double)A->K / r : 0.0;
        for (int a = 0; a < 3; ++a) g_soup[3 * k + a] = c * y[a];
    }
}

// Analytic dL/dx (double[3n], caller order).
int orc_adj_grad(const orc_adj* A, double* grad) {
    std::vector<double> g_phi(A->V), g_soup(3 * A->K), gs(3ull * A->n, 0.0);
    orc_adj_upstream(A, g_phi.data(), g_soup.data());
    const double isod = A->iso;
    for (uint64_t k = 0; k < A->K; ++k) {
        const double pa = A->phi[A->soup.a[k]], pb = A->phi[A->soup.b[k]];
        const double delta = pb - pa;
        if (delta > -1e-7 && delta < 1e-7) continue;
        const double gt = g_soup[3 * k] * A->soup.dc[3 * k] +
                          g_soup[3 * k + 1] * A->soup.dc[3 * k + 1] +
                          g_soup[3 * k + 2] * A->soup.dc[3 * k + 2];
        g_phi[A->soup.a[k]] += gt * (isod - pb) / (delta * delta);
        g_phi[A->soup.b[k]] -= gt * (isod - pa) / (delta * delta);
    }
    // only band vertices carry a field
    std::vector<double> g_band(A->V, 0.0);
    for (uint32_t b = 0; b < A->B; ++b) g_band[A->band_index[b]] = g_phi[A->band_index[b]];
    field_zb05_backward_impl(A->band_index.data(), A->B, A->pos.data(), A->ranges.data(), A->sp,
                             A->sc, A->radius_sq, g_band.data(), gs.data());
    for (uint32_t i = 0; i < A->n; ++i)
        for (int a = 0; a < 3; ++a) grad[3ull * A->perm[i] + a] = gs[3ull * i + a];
    return 0;
}

// Central difference (L(x + eps e) - L(x - eps e)) / (2 eps) for the caller's
// particle j, coordinate axis, topology and cells frozen.  Only the band
// vertices within reach of the particle and the soup vertices on their edges
// are re-evaluated; all other terms cancel exactly.
double orc_adj_fd(orc_adj* A, uint32_t j, int axis, double eps) {
    const uint32_t s = A->inv[j];
    const double base = A->pos[3ull * s + axis];
    const double R = std::sqrt((double)A->radius_sq) + 2.0 * eps;
    const orc_grid& sc = A->sc;
    const float mins[3] = {sc.minx, sc.miny, sc.minz};
    const uint32_t res[3] = {sc.rx, sc.ry, sc.rz};
    int lo[3], hi[3];
    for (int a = 0; a < 3; ++a) {
        const double p = A->pos[3ull * s + a];
        lo[a] = std::max(0, (int)std::floor((p - R - mins[a]) / sc.cell) - 1);
        hi[a] = std::min((int)res[a] - 1, (int)std::ceil((p + R - mins[a]) / sc.cell) + 1);
    }
    std::vector<uint32_t> verts;
    for (int z = lo[2]; z <= hi[2]; ++z)
        for (int y = lo[1]; y <= hi[1]; ++y)
            for (int x = lo[0]; x <= hi[0]; ++x) {
                const uint32_t vid = (uint32_t)(((uint64_t)z * sc.ry + y) * sc.rx + x);
                if (A->band[vid]) verts.push_back(vid);
            }
    std::vector<uint32_t> soups;
    for (uint32_t vid : verts) {
        const uint32_t b = band_pos(A, vid);
        for (uint32_t q = A->vs_off[b]; q < A->vs_off[b + 1]; ++q) soups.push_back(A->vs_list[q]);
    }
    std::sort(soups.begin(), soups.end());
    soups.erase(std::unique(soups.begin(), soups.end()), soups.end());
    double lg[2] = {0, 0}, ls[2] = {0, 0};
    std::vector<double> local(verts.size());
    for (int side = 0; side < 2; ++side) {
        A->pos[3ull * s + axis] = base + (side == 0 ? eps : -eps);
        for (size_t i = 0; i < verts.size(); ++i) {
            local[i] = phi_at(A, verts[i]);
            const double d = local[i] - A->sdf[verts[i]];
            lg[side] += d * d;
        }
        auto phi_of = [&](uint32_t vid) {
            const auto it = std::lower_bound(verts.begin(), verts.end(), vid);
            return (it != verts.end() && *it == vid) ? local[it - verts.begin()] : A->phi[vid];
        };
        for (uint32_t k : soups) ls[side] += soup_term(A, k, phi_of(A->soup.a[k]), phi_of(A->soup.b[k]));
    }
    A->pos[3ull * s + axis] = base;
    const double dl = A->wg * (lg[0] - lg[1]) / (double)A->V +
                      (A->K ? A->ws * (ls[0] - ls[1]) / (double)A->K : 0.0);
    return dl / (2.0 * eps);
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Deduplicated marching-cubes mesh (no reference counterpart; the reference
// emits a soup).  One vertex per grid edge that is crossed and used by a valid
// cube, numbered by (lower grid vertex ascending, axis x, y, z), interpolated
// from the lower end in FP64; indices in the soup's order.
// ---------------------------------------------------------------------------
extern "C" {

uint64_t orc_mc_indexed(const uint32_t* band_index, uint32_t B, const uint8_t* band,
                        const double* phi, orc_grid sc, float iso, int cpu, double* vtx_pos,
                        double* vtx_nrm, uint64_t vtx_capacity, uint32_t* indices,
                        uint64_t idx_capacity, uint64_t* num_indices) {
    const uint64_t rx = sc.rx, ry = sc.ry, rz = sc.rz, rxy = rx * ry;
    const uint64_t res[3] = {rx, ry, rz};
    const uint64_t stride[3] = {1, rx, rxy};
    const double isod = iso;
    std::vector<uint32_t> nverts(B);
    orc_mc(band_index, B, band, phi, sc, iso, cpu, nverts.data(), nullptr, nullptr, 0, nullptr);
    auto bpos = [&](uint64_t vid) -> int64_t {
        const uint32_t* it = std::lower_bound(band_index, band_index + B, (uint32_t)vid);
        return (it != band_index + B && *it == vid) ? (int64_t)(it - band_index) : -1;
    };
    auto cube_valid = [&](uint64_t x, uint64_t y, uint64_t z) {
        if (x + 1 >= rx || y + 1 >= ry || z + 1 >= rz) return false;
        const int64_t b = bpos((z * ry + y) * rx + x);
        return b >= 0 && nverts[b] > 0;
    };
    // vertex ids per (band position, axis)
    std::vector<int64_t> id(3ull * B, -1);
    uint64_t U = 0;
    for (uint32_t b = 0; b < B; ++b) {
        const uint64_t vid = band_index[b];
        const uint64_t c[3] = {vid % rx, (vid / rx) % ry, vid / rxy};
        const bool below = phi[vid] <= isod;
        for (int ax = 0; ax < 3; ++ax) {
            if (c[ax] + 1 >= res[ax]) continue;
            if ((phi[vid + stride[ax]] <= isod) == below) continue;
            const int p = ax == 0 ? 1 : 0, q = ax == 2 ? 1 : 2;
            bool used = false;
            for (int d = 0; d < 4 && !used; ++d) {
                uint64_t u[3] = {c[0], c[1], c[2]};
                const uint64_t dp = d & 1, dq = d >> 1;
                if (u[p] < dp || u[q] < dq) continue;
                u[p] -= dp;
                u[q] -= dq;
                used = cube_valid(u[0], u[1], u[2]);
            }
            if (used) id[3ull * b + ax] = (int64_t)U++;
        }
    }
    auto sample = [&](int64_t X, int64_t Y, int64_t Z) -> double {
        if (X < 0 || Y < 0 || Z < 0 || X >= (int64_t)rx || Y >= (int64_t)ry || Z >= (int64_t)rz)
            return 0.0;
        return phi[((uint64_t)Z * ry + Y) * rx + X];
    };
    auto normal = [&](int64_t X, int64_t Y, int64_t Z, double n[3]) {
        n[0] = sample(X - 1, Y, Z) - sample(X + 1, Y, Z);
        n[1] = sample(X, Y - 1, Z) - sample(X, Y + 1, Z);
        n[2] = sample(X, Y, Z - 1) - sample(X, Y, Z + 1);
        normalise_corner(n, cpu);
    };
    const float mins[3] = {sc.minx, sc.miny, sc.minz};
    const double sign = (!cpu && iso < 0.0f) ? -1.0 : 1.0;
    if (vtx_pos && U <= vtx_capacity) {
        for (uint32_t b = 0; b < B; ++b) {
            const uint64_t vid = band_index[b];
            const int64_t c[3] = {(int64_t)(vid % rx), (int64_t)((vid / rx) % ry),
                                  (int64_t)(vid / rxy)};
            for (int ax = 0; ax < 3; ++ax) {
                const int64_t k = id[3ull * b + ax];
                if (k < 0) continue;
                int64_t e[3] = {c[0], c[1], c[2]};
                e[ax] += 1;
                const double va = phi[vid], vb = phi[vid + stride[ax]];
                const double delta = vb - va;
                const double t = (delta > -1e-7 && delta < 1e-7) ? 0.5 : (isod - va) / delta;
                double na[3], nb[3], nn[3];
                normal(c[0], c[1], c[2], na);
                normal(e[0], e[1], e[2], nb);
                for (int a = 0; a < 3; ++a) {
                    const double pa = cpu ? (double)(mins[a] + (float)c[a] * sc.cell)
                                          : (double)vpos((uint32_t)c[a], mins[a], sc.cell);
                    const double pb = cpu ? (double)(mins[a] + (float)e[a] * sc.cell)
                                          : (double)vpos((uint32_t)e[a], mins[a], sc.cell);
                    vtx_pos[3 * k + a] = pa + t * (pb - pa);
                    nn[a] = na[a] + t * (nb[a] - na[a]);
                }
                const double len = std::sqrt(nn[0] * nn[0] + nn[1] * nn[1] + nn[2] * nn[2]);
                if (cpu) {
                    if (len > 1e-7) for (int a = 0; a < 3; ++a) nn[a] /= len;
                } else {
                    for (int a = 0; a < 3; ++a) nn[a] = sign * (nn[a] * (1.0 / len));
                }
                if (vtx_nrm) for (int a = 0; a < 3; ++a) vtx_nrm[3 * k + a] = nn[a];
            }
        }
    }
    uint64_t total = 0;
    for (uint32_t b = 0; b < B; ++b) total += nverts[b];
    if (num_indices) *num_indices = total;
    if (indices && total <= idx_capacity) {
        uint64_t off = 0;
        for (uint32_t b = 0; b < B; ++b) {
            const uint32_t nv = nverts[b];
            if (!nv) continue;
            const uint64_t vid = band_index[b];
            const uint64_t x = vid % rx, y = (vid / rx) % ry, z = vid / rxy;
            uint32_t f = 0;
            for (int c = 0; c < 8; ++c) {
                const uint64_t v = ((z + kCornerOff[c][2]) * ry + (y + kCornerOff[c][1])) * rx +
                                   (x + kCornerOff[c][0]);
                if (phi[v] <= isod) f |= 1u << c;
            }
            for (uint32_t k = 0; k < nv; ++k) {
                const int e = (int)((kTri[f] >> (4 * k)) & 0xF);
                const int s0 = kEdgeCorner[e][0], s1 = kEdgeCorner[e][1];
                int ax = 0;
                while (kCornerOff[s0][ax] == kCornerOff[s1][ax]) ++ax;
                const int lo = kCornerOff[s0][ax] < kCornerOff[s1][ax] ? s0 : s1;
                const uint64_t lv = ((z + kCornerOff[lo][2]) * ry + (y + kCornerOff[lo][1])) * rx +
                                    (x + kCornerOff[lo][0]);
                indices[off + k] = (uint32_t)id[3ull * bpos(lv) + ax];
            }
            off += nv;
        }
    }
    return U;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Launchers the reference defines but never calls.
// ---------------------------------------------------------------------------
extern "C" {

// estimateSurfaceVertices (ReconstructionCUDA.cu:123-175) with getVertexFlag
// (ReconstructionCUDA.cuh:491-511) on the density (occupancy) grid, iso 0.5.
int orc_estimate_surface_vertices(const float* dens, orc_grid dg, uint32_t ratio, uint32_t ext,
                                  orc_grid sg, uint32_t* is_surface) {
    const int64_t rx = dg.rx, ry = dg.ry, rz = dg.rz;
    for (int64_t z = 0; z + 1 < rz; ++z)
        for (int64_t y = 0; y + 1 < ry; ++y)
            for (int64_t x = 0; x + 1 < rx; ++x) {
                uint32_t flag = 0;
                for (int c = 0; c < 8; ++c) {
                    const int64_t X = x + kCornerOff[c][0], Y = y + kCornerOff[c][1],
                                  Z = z + kCornerOff[c][2];
                    if (dens[(Z * ry + Y) * rx + X] <= 0.5f) flag |= 1u << c;
                }
                if (flag == 0 || flag == 255) continue;
                const int64_t lo[3] = {x * (int64_t)ratio - (int64_t)ext,
                                       y * (int64_t)ratio - (int64_t)ext,
                                       z * (int64_t)ratio - (int64_t)ext};
                const int64_t hi[3] = {(x + 1) * (int64_t)ratio + (int64_t)ext,
                                       (y + 1) * (int64_t)ratio + (int64_t)ext,
                                       (z + 1) * (int64_t)ratio + (int64_t)ext};
                const int64_t res[3] = {sg.rx, sg.ry, sg.rz};
                int64_t a[3], b[3];
                for (int k = 0; k < 3; ++k) {
                    a[k] = std::min(std::max(lo[k], (int64_t)0), res[k] - 1);
                    b[k] = std::min(std::max(hi[k], (int64_t)0), res[k] - 1);
                }
                for (int64_t zz = a[2]; zz <= b[2]; ++zz)
                    for (int64_t yy = a[1]; yy <= b[1]; ++yy)
                        for (int64_t xx = a[0]; xx <= b[0]; ++xx)
                            is_surface[(zz * sg.ry + yy) * sg.rx + xx] = 1;
            }
    return 0;
}

// updateScalarFieldValueTC01 (ReconstructionCUDA.cu:178-227) with kernelTC01
// (ReconstructionCUDA.cuh:252-260) in FP64: (d - 1/2)^2, d = |x - p|^2 / (2 R^2),
// over the clamped 3^3 block, for |x - p|^2 <= R^2.
int orc_field_tc01(const uint32_t* band_index, uint32_t B, const float* pos,
                   const uint32_t* ranges, orc_grid sp, orc_grid sc, float radius_sq,
                   double* phi) {
    const int rx = sp.rx, ry = sp.ry, rz = sp.rz;
    const double R2 = radius_sq;
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t b = 0; b < (int64_t)B; ++b) {
        const uint32_t vid = band_index[b];
        const uint32_t ix = vid % sc.rx, iy = (vid / sc.rx) % sc.ry, iz = vid / (sc.rx * sc.ry);
        const float vxf = vpos(ix, sc.minx, sc.cell), vyf = vpos(iy, sc.miny, sc.cell),
                    vzf = vpos(iz, sc.minz, sc.cell);
        const int cx = trunc_q(cell_q(vxf, sp.minx, sp.cell), rx);
        const int cy = trunc_q(cell_q(vyf, sp.miny, sp.cell), ry);
        const int cz = trunc_q(cell_q(vzf, sp.minz, sp.cell), rz);
        double val = 0;
        for (int zz = clampi(cz - 1, 0, rz - 1); zz <= clampi(cz + 1, 0, rz - 1); ++zz)
            for (int yy = clampi(cy - 1, 0, ry - 1); yy <= clampi(cy + 1, 0, ry - 1); ++yy)
                for (int xx = clampi(cx - 1, 0, rx - 1); xx <= clampi(cx + 1, 0, rx - 1); ++xx) {
                    const uint64_t c = ((uint64_t)zz * ry + yy) * rx + xx;
                    if (ranges[2 * c] == 0xFFFFFFFFu) continue;
                    for (uint32_t j = ranges[2 * c]; j < ranges[2 * c + 1]; ++j) {
                        const double dx = (double)vxf - pos[3 * j];
                        const double dy = (double)vyf - pos[3 * j + 1];
                        const double dz = (double)vzf - pos[3 * j + 2];
                        const double d2 = dx * dx + dy * dy + dz * dz;
                        if (d2 > R2) continue;
                        const double d = d2 / R2 * 0.5;
                        val += d * d - d + 0.25;
                    }
                }
        phi[vid] = val;
    }
    return 0;
}

}  // extern "C"
