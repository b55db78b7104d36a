// Host-side frame I/O of the reference's on-disk formats, and the constants
// its reconstructor classes set, so that a frame sequence can be driven the way
// ReconstructorGPU::reconstruct does (GPU/ReconstructorGPU.cpp:25-79).
//
//   .xyz  (FSR/data_format.txt; reader GPU/ReconstructorGPU.cpp:200-290): five
//         header lines (AABB min, AABB max, kernel radius, particle radius,
//         particle mass), then one "x y z density" line per particle.  Numbers
//         are parsed as the reference's `stream >> float` does (correctly
//         rounded decimal to binary32).  Blank lines are skipped; the reference
//         would push an uninitialised particle for a trailing empty line.
//   .obj  (GPU/ReconstructorGPUOurs.cpp:20-100): all "v x y z" lines, then all
//         "vn" lines, then "f i//i j//j k//k" with 1-based indices, numbers in
//         the default iostream format (6 significant digits).
//   .vis  (saveMiddleDataToVisFile, GPU/ReconstructorGPUOursZB05.cpp:104-231 and
//         its Akinci12 / Yu13 versions): the frame's intermediates for the
//         reference's Visualizer, same number format; index lists are written
//         "i j k " on one line, followed by a line break only when the list is
//         not empty (the occupied-cell list: when the hash grid is not empty).
#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "common.hpp"

using namespace fsr;

namespace {

bool parse_floats(const std::string& line, float* out, int n) {
    const char* p = line.c_str();
    for (int i = 0; i < n; ++i) {
        char* end = nullptr;
        errno = 0;
        const float v = std::strtof(p, &end);
        if (end == p) return false;
        out[i] = v;
        p = end;
    }
    return true;
}

bool blank(const std::string& line) {
    for (char c : line)
        if (c != ' ' && c != '\t' && c != '\r' && c != '\n') return false;
    return true;
}

fsr_status read_header(std::ifstream& f, const char* path, fsr_xyz_header* h) {
    std::string line;
    float v[3];
    const char* what[5] = {"AABB min", "AABB max", "kernel radius", "particle radius",
                           "particle mass"};
    const int count[5] = {3, 3, 1, 1, 1};
    float* dst[5] = {&h->aabb_min.x, &h->aabb_max.x, &h->kernel_radius, &h->particle_radius,
                     &h->particle_mass};
    for (int k = 0; k < 5; ++k) {
        if (!std::getline(f, line) || !parse_floats(line, v, count[k])) {
            set_error(std::string(path) + ": cannot parse the " + what[k] + " header line");
            return FSR_ERR_INVALID_ARG;
        }
        if (count[k] == 3) {
            dst[k][0] = v[0];
            dst[k][1] = v[1];
            dst[k][2] = v[2];
        } else {
            *dst[k] = v[0];
        }
    }
    return FSR_OK;
}

// Default iostream formatting (%g with 6 significant digits), as the
// reference's `file << float` writes.
inline int put_float(char* buf, float x) { return std::snprintf(buf, 32, "%g", (double)x); }

}  // namespace

extern "C" {

fsr_status fsr_xyz_info(const char* path, fsr_xyz_header* header, uint32_t* num_particles) {
    if (!path || !header || !num_particles) {
        set_error("fsr_xyz_info: null argument");
        return FSR_ERR_INVALID_ARG;
    }
    std::ifstream f(path);
    if (!f) {
        set_error(std::string("fsr_xyz_info: cannot open ") + path);
        return FSR_ERR_INVALID_ARG;
    }
    FSR_TRY(read_header(f, path, header));
    std::string line;
    uint64_t n = 0;
    while (std::getline(f, line))
        if (!blank(line)) ++n;
    if (n > 0xFFFFFFFFull) {
        set_error("fsr_xyz_info: more than 2^32 - 1 particles");
        return FSR_ERR_INVALID_ARG;
    }
    *num_particles = (uint32_t)n;
    return FSR_OK;
}

fsr_status fsr_xyz_read(const char* path, fsr_xyz_header* header, float* pos, float* dens,
                        uint32_t capacity, uint32_t* num_particles) {
    if (!path || !header || !num_particles || (capacity && (!pos || !dens))) {
        set_error("fsr_xyz_read: null argument");
        return FSR_ERR_INVALID_ARG;
    }
    std::ifstream f(path);
    if (!f) {
        set_error(std::string("fsr_xyz_read: cannot open ") + path);
        return FSR_ERR_INVALID_ARG;
    }
    FSR_TRY(read_header(f, path, header));
    std::string line;
    uint32_t n = 0;
    uint64_t lineno = 5;
    while (std::getline(f, line)) {
        ++lineno;
        if (blank(line)) continue;
        if (n == capacity) {
            set_error("fsr_xyz_read: more particles than the capacity");
            return FSR_ERR_CAPACITY;
        }
        float v[4];
        if (!parse_floats(line, v, 4)) {
            set_error(std::string(path) + ": cannot parse particle line " + std::to_string(lineno));
            return FSR_ERR_INVALID_ARG;
        }
        pos[3ull * n + 0] = v[0];
        pos[3ull * n + 1] = v[1];
        pos[3ull * n + 2] = v[2];
        dens[n] = v[3];
        ++n;
    }
    *num_particles = n;
    return FSR_OK;
}

fsr_status fsr_obj_write_soup(const char* path, const float* pos, const float* nrm,
                              uint32_t num_vertices) {
    if (!path || (num_vertices && !pos) || num_vertices % 3) {
        set_error("fsr_obj_write_soup: null buffer or a vertex count that is not 3T");
        return FSR_ERR_INVALID_ARG;
    }
    std::FILE* f = std::fopen(path, "w");
    if (!f) {
        set_error(std::string("fsr_obj_write_soup: cannot write ") + path);
        return FSR_ERR_INVALID_ARG;
    }
    char a[32], b[32], c[32];
    for (uint32_t i = 0; i < num_vertices; ++i) {
        put_float(a, pos[3ull * i]);
        put_float(b, pos[3ull * i + 1]);
        put_float(c, pos[3ull * i + 2]);
        std::fprintf(f, "v %s %s %s\n", a, b, c);
    }
    if (nrm)
        for (uint32_t i = 0; i < num_vertices; ++i) {
            put_float(a, nrm[3ull * i]);
            put_float(b, nrm[3ull * i + 1]);
            put_float(c, nrm[3ull * i + 2]);
            std::fprintf(f, "vn %s %s %s\n", a, b, c);
        }
    for (uint64_t i = 1; i <= num_vertices; i += 3)
        std::fprintf(f, "f %llu//%llu %llu//%llu %llu//%llu \n", (unsigned long long)i,
                     (unsigned long long)i, (unsigned long long)(i + 1),
                     (unsigned long long)(i + 1), (unsigned long long)(i + 2),
                     (unsigned long long)(i + 2));
    const bool ok = std::fclose(f) == 0;
    if (!ok) {
        set_error(std::string("fsr_obj_write_soup: write error on ") + path);
        return FSR_ERR_INVALID_ARG;
    }
    return FSR_OK;
}

fsr_status fsr_obj_write_indexed(const char* path, const float* pos, const float* nrm,
                                 uint32_t num_vertices, const uint32_t* indices,
                                 uint32_t num_indices) {
    if (!path || (num_vertices && !pos) || (num_indices && !indices) || num_indices % 3) {
        set_error("fsr_obj_write_indexed: null buffer or an index count that is not 3T");
        return FSR_ERR_INVALID_ARG;
    }
    for (uint32_t i = 0; i < num_indices; ++i)
        if (indices[i] >= num_vertices) {
            set_error("fsr_obj_write_indexed: index out of range");
            return FSR_ERR_INVALID_ARG;
        }
    std::FILE* f = std::fopen(path, "w");
    if (!f) {
        set_error(std::string("fsr_obj_write_indexed: cannot write ") + path);
        return FSR_ERR_INVALID_ARG;
    }
    char a[32], b[32], c[32];
    for (uint32_t i = 0; i < num_vertices; ++i) {
        put_float(a, pos[3ull * i]);
        put_float(b, pos[3ull * i + 1]);
        put_float(c, pos[3ull * i + 2]);
        std::fprintf(f, "v %s %s %s\n", a, b, c);
    }
    if (nrm)
        for (uint32_t i = 0; i < num_vertices; ++i) {
            put_float(a, nrm[3ull * i]);
            put_float(b, nrm[3ull * i + 1]);
            put_float(c, nrm[3ull * i + 2]);
            std::fprintf(f, "vn %s %s %s\n", a, b, c);
        }
    for (uint32_t t = 0; t + 2 < num_indices; t += 3) {
        const unsigned long long i0 = indices[t] + 1ull, i1 = indices[t + 1] + 1ull,
                                 i2 = indices[t + 2] + 1ull;
        std::fprintf(f, "f %llu//%llu %llu//%llu %llu//%llu \n", i0, i0, i1, i1, i2, i2);
    }
    const bool ok = std::fclose(f) == 0;
    if (!ok) {
        set_error(std::string("fsr_obj_write_indexed: write error on ") + path);
        return FSR_ERR_INVALID_ARG;
    }
    return FSR_OK;
}

fsr_status fsr_vis_write(fsr_context* ctx, const char* path, fsr_method method,
                         const fsr_grid_info* spg, const fsr_grid_info* scg,
                         const fsr_sim_param* sim, const float* sorted_pos,
                         uint32_t n, const fsr_frame_stats* st, const char* obj_name) {
    if (!ctx || !path || !spg || !scg || !sim || !st || !obj_name || (n && !sorted_pos)) {
        set_error("fsr_vis_write: null argument");
        return FSR_ERR_INVALID_ARG;
    }
    if (method != FSR_METHOD_ZB05 && method != FSR_METHOD_AKINCI12 && method != FSR_METHOD_YU13) {
        set_error("fsr_vis_write: unknown method");
        return FSR_ERR_INVALID_ARG;
    }
    const bool yu13 = method == FSR_METHOD_YU13;
    const uint64_t C = (uint64_t)spg->resolution.x * spg->resolution.y * spg->resolution.z;
    FSR_HIP(hipSetDevice(ctx->device));
    FSR_HIP(hipDeviceSynchronize());
    auto fetch = [&](fsr_buffer which, uint64_t count, size_t elem, std::vector<uint8_t>& out)
        -> fsr_status {
        out.resize(count * elem);
        if (!count) return FSR_OK;
        void* p = nullptr;
        uint64_t have = 0;
        FSR_TRY(fsr_frame_buffer(ctx, which, &p, &have));
        if (!p || have < count) {
            set_error("fsr_vis_write: the context holds no frame of this size");
            return FSR_ERR_INVALID_ARG;
        }
        FSR_HIP(hipMemcpy(out.data(), p, count * elem, hipMemcpyDeviceToHost));
        return FSR_OK;
    };
    std::vector<uint8_t> occ, band, flags, valid, smooth, involved;
    std::vector<float> pos(3ull * n);
    if (n) FSR_HIP(hipMemcpy(pos.data(), sorted_pos, 12ull * n, hipMemcpyDeviceToHost));
    FSR_TRY(fetch(FSR_BUF_OCCUPANCY, C, 4, occ));
    FSR_TRY(fetch(FSR_BUF_BAND_INDEX, st->num_band_vertices, 4, band));
    FSR_TRY(fetch(FSR_BUF_SURFACE_FLAGS, n, 4, flags));
    FSR_TRY(fetch(FSR_BUF_VALID_CUBES, st->num_valid_cubes, 4, valid));
    if (yu13) {
        FSR_TRY(fetch(FSR_BUF_SMOOTH_POS, n, 16, smooth));
        FSR_TRY(fetch(FSR_BUF_INVOLVED_INDEX, st->num_involved_particles, 4, involved));
    }
    std::FILE* f = std::fopen(path, "w");
    if (!f) {
        set_error(std::string("fsr_vis_write: cannot write ") + path);
        return FSR_ERR_INVALID_ARG;
    }
    char a[32], b[32], c[32];
    auto put3 = [&](float x, float y, float z) {
        put_float(a, x);
        put_float(b, y);
        put_float(c, z);
        std::fprintf(f, "%s %s %s\n", a, b, c);
    };
    auto put1 = [&](float x) {
        put_float(a, x);
        std::fprintf(f, "%s\n", a);
    };
    auto grid = [&](const fsr_grid_info* g) {
        put3(g->min_pos.x, g->min_pos.y, g->min_pos.z);
        put3(g->max_pos.x, g->max_pos.y, g->max_pos.z);
        std::fprintf(f, "%u %u %u\n", g->resolution.x, g->resolution.y, g->resolution.z);
        put1(g->cell_size);
    };
    auto u32s = [](const std::vector<uint8_t>& v) {
        return reinterpret_cast<const uint32_t*>(v.data());
    };
    auto f32s = [](const std::vector<uint8_t>& v) {
        return reinterpret_cast<const float*>(v.data());
    };
    // index list: the count, then "i j k " and a line break when not empty
    auto list = [&](const uint32_t* idx, uint64_t count) {
        std::fprintf(f, "%llu\n", (unsigned long long)count);
        for (uint64_t i = 0; i < count; ++i) std::fprintf(f, "%u ", idx[i]);
        if (count) std::fputc('\n', f);
    };
    grid(spg);
    grid(scg);
    put1(sim->particle_radius);
    std::fprintf(f, "%u\n", n);
    for (uint32_t i = 0; i < n; ++i) put3(pos[3ull * i], pos[3ull * i + 1], pos[3ull * i + 2]);
    // occupied hash cells
    uint64_t nocc = 0;
    for (uint64_t i = 0; i < C; ++i) nocc += f32s(occ)[i] > 0.0f ? 1u : 0u;
    std::fprintf(f, "%llu\n", (unsigned long long)nocc);
    for (uint64_t i = 0; i < C; ++i)
        if (f32s(occ)[i] > 0.0f) std::fprintf(f, "%llu ", (unsigned long long)i);
    if (C) std::fputc('\n', f);
    // smoothed particles (Yu13 only)
    if (yu13) {
        std::fprintf(f, "%u\n", n);
        for (uint32_t i = 0; i < n; ++i)
            put3(f32s(smooth)[4ull * i], f32s(smooth)[4ull * i + 1], f32s(smooth)[4ull * i + 2]);
    } else {
        std::fprintf(f, "0\n");
    }
    list(u32s(band), st->num_band_vertices);
    // surface particles: the ids whose flag is 1
    std::vector<uint32_t> surf;
    for (uint32_t i = 0; i < n; ++i)
        if (u32s(flags)[i] == 1u) surf.push_back(i);
    list(surf.data(), surf.size());
    if (yu13)
        list(u32s(involved), st->num_involved_particles);
    else
        std::fprintf(f, "0\n");
    list(u32s(valid), st->num_valid_cubes);
    std::fprintf(f, "%s\n", obj_name);
    put1(sim->smoothing_radius);
    if (std::fclose(f) != 0) {
        set_error(std::string("fsr_vis_write: write error on ") + path);
        return FSR_ERR_INVALID_ARG;
    }
    return FSR_OK;
}

fsr_status fsr_sim_param_for(fsr_method method, const fsr_xyz_header* h, fsr_sim_param* out) {
    if (!h || !out) {
        set_error("fsr_sim_param_for: null argument");
        return FSR_ERR_INVALID_ARG;
    }
    if (method != FSR_METHOD_ZB05 && method != FSR_METHOD_AKINCI12 && method != FSR_METHOD_YU13) {
        set_error("fsr_sim_param_for: unknown method");
        return FSR_ERR_INVALID_ARG;
    }
    fsr_sim_param p;
    std::memset(&p, 0, sizeof(p));
    // readParticlesFromFile (GPU/ReconstructorGPU.cpp:226-245)
    p.smoothing_radius = h->kernel_radius;
    p.smoothing_radius_inv = (float)(1.0 / (double)p.smoothing_radius);
    p.smoothing_radius_sq = p.smoothing_radius * p.smoothing_radius;
    p.particle_radius = h->particle_radius;
    p.particle_mass = h->particle_mass;
    if (method == FSR_METHOD_YU13) {
        // ReconstructorGPUOursYu13.cpp:21 and :78-99
        p.anisotropic_radius = p.smoothing_radius * 2.0f;
        p.iso_value = 0.0f;
        p.sc_sp_grid_res_ratio = 3;
        p.expand_extent = 3;
        p.spatial_cell_size_scale = 1.3f;
        p.lambda_for_smoothed = 0.5f;
        p.min_num_neighbors = 25;
    } else {
        // ReconstructorGPUOursZB05.cpp:85-97, ReconstructorGPUAkinci12.cpp:93-106
        p.iso_value = -0.0001f;
        p.sc_sp_grid_res_ratio = 2;
        p.expand_extent = 4;
        p.spatial_cell_size_scale = 1.0f;
    }
    *out = p;
    return FSR_OK;
}

}  // extern "C"
