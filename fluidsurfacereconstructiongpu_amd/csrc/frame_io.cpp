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
#include <algorithm>
#include <cerrno>
#include <charconv>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <thread>
#include <vector>

#include "common.hpp"

using namespace fsr;

namespace {

// ---- .xyz body: the whole file is read into memory and its particle lines
// are parsed by several host threads, each over a range of whole lines (a
// count pass, then a parse pass into the rows its count assigns).  Numbers go
// through std::from_chars, which rounds correctly like `stream >> float`.

bool is_space(char c) { return c == ' ' || c == '\t' || c == '\r'; }

// A float at p (leading blanks and one '+' skipped); false when there is none.
bool next_float(const char*& p, const char* end, float* out) {
    while (p < end && is_space(*p)) ++p;
    if (p < end && *p == '+') ++p;
    const std::from_chars_result r = std::from_chars(p, end, *out);
    if (r.ec == std::errc::invalid_argument || r.ptr == p) return false;
    if (r.ec == std::errc::result_out_of_range) {
        // from_chars leaves the value unset here: take strtof's (+-HUGE_VALF,
        // or the nearest denormal or +-0 for an underflow)
        const std::string token(p, r.ptr);
        *out = std::strtof(token.c_str(), nullptr);
    }
    p = r.ptr;
    return true;
}

bool blank_range(const char* a, const char* b) {
    for (; a < b; ++a)
        if (!is_space(*a)) return false;
    return true;
}

fsr_status slurp(const char* path, const char* who, std::vector<char>* buf) {
    std::FILE* f = std::fopen(path, "rb");
    if (!f) {
        set_error(std::string(who) + ": cannot open " + path);
        return FSR_ERR_INVALID_ARG;
    }
    buf->clear();
    long size = -1;
    if (std::fseek(f, 0, SEEK_END) == 0) size = std::ftell(f);
    if (size < 0 || std::fseek(f, 0, SEEK_SET) != 0) {
        std::fclose(f);
        set_error(std::string(who) + ": cannot size " + path);
        return FSR_ERR_INVALID_ARG;
    }
    buf->resize((size_t)size);
    const size_t got = size ? std::fread(buf->data(), 1, (size_t)size, f) : 0;
    std::fclose(f);
    if (got != (size_t)size) {
        set_error(std::string(who) + ": short read of " + path);
        return FSR_ERR_INVALID_ARG;
    }
    return FSR_OK;
}

// The five header lines; *body = offset of the first particle line.
fsr_status header_from(const std::vector<char>& buf, const char* path, fsr_xyz_header* h,
                       size_t* body) {
    const char* what[5] = {"AABB min", "AABB max", "kernel radius", "particle radius",
                           "particle mass"};
    const int count[5] = {3, 3, 1, 1, 1};
    float* dst[5] = {&h->aabb_min.x, &h->aabb_max.x, &h->kernel_radius, &h->particle_radius,
                     &h->particle_mass};
    size_t at = 0;
    for (int k = 0; k < 5; ++k) {
        const char* b = buf.data() + at;
        const char* e = static_cast<const char*>(std::memchr(b, '\n', buf.size() - at));
        const char* le = e ? e : buf.data() + buf.size();
        bool ok = true;
        float v[3];
        const char* p = b;
        for (int j = 0; j < count[k] && ok; ++j) ok = next_float(p, le, &v[j]);
        if (!ok || at >= buf.size()) {
            set_error(std::string(path) + ": cannot parse the " + what[k] + " header line");
            return FSR_ERR_INVALID_ARG;
        }
        for (int j = 0; j < count[k]; ++j) dst[k][j] = v[j];
        at = e ? (size_t)(e - buf.data()) + 1 : buf.size();
    }
    *body = at;
    return FSR_OK;
}

// Line-aligned ranges of the body, one per thread.
std::vector<size_t> line_ranges(const std::vector<char>& buf, size_t body) {
    const size_t len = buf.size() - body;
    unsigned t = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    if (len < (1u << 20)) t = 1;
    std::vector<size_t> cut{body};
    for (unsigned k = 1; k < t; ++k) {
        size_t at = std::max(cut.back(), body + len * k / t);
        const void* nl = at < buf.size() ? std::memchr(buf.data() + at, '\n', buf.size() - at)
                                         : nullptr;
        at = nl ? (size_t)(static_cast<const char*>(nl) - buf.data()) + 1 : buf.size();
        cut.push_back(at);
    }
    cut.push_back(buf.size());
    return cut;
}

template <class F>
void parallel_ranges(size_t n, F f) {
    std::vector<std::thread> th;
    for (size_t k = 1; k < n; ++k) th.emplace_back(f, k);
    f(0);
    for (auto& x : th) x.join();
}

// Lines and non-blank lines of [a, b).
void count_lines(const char* a, const char* b, uint64_t* lines, uint64_t* particles) {
    uint64_t nl = 0, np = 0;
    while (a < b) {
        const char* e = static_cast<const char*>(std::memchr(a, '\n', (size_t)(b - a)));
        const char* le = e ? e : b;
        ++nl;
        if (!blank_range(a, le)) ++np;
        a = e ? e + 1 : b;
    }
    *lines = nl;
    *particles = np;
}

// Default iostream formatting (%g with 6 significant digits), as the
// reference's `file << float` writes.
inline int put_float(char* buf, float x) { return std::snprintf(buf, 32, "%g", (double)x); }

// The same text without the printf machinery: std::to_chars in the general
// format with precision 6 is specified as printf's "%.6g" (correctly rounded
// from the float's exact value, which (double)x keeps).
inline char* put_g(char* p, float x) {
    return std::to_chars(p, p + 32, x, std::chars_format::general, 6).ptr;
}
inline char* put_u(char* p, unsigned long long v) { return std::to_chars(p, p + 24, v).ptr; }

// Lines 0 .. n-1, line i written by line(i, p) (returning the end), at most
// max_len bytes each: formatted on up to 16 threads in contiguous ranges and
// written in order.  Returns false on a write error.
template <typename L>
bool write_lines(std::FILE* f, uint64_t n, size_t max_len, L line) {
    if (!n) return true;
    unsigned t = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    if (n < 4096) t = 1;
    std::vector<std::vector<char>> out(t);
    parallel_ranges(t, [&](size_t k) {
        const uint64_t lo = n * k / t, hi = n * (k + 1) / t;
        std::vector<char>& b = out[k];
        b.resize((size_t)(hi - lo) * max_len);
        char* p = b.data();
        for (uint64_t i = lo; i < hi; ++i) p = line(i, p);
        b.resize((size_t)(p - b.data()));
    });
    for (const auto& b : out)
        if (!b.empty() && std::fwrite(b.data(), 1, b.size(), f) != b.size()) return false;
    return true;
}

// "v x y z" / "vn x y z" lines of n float3s.
bool write_vec_lines(std::FILE* f, const char* tag, const float* v, uint64_t n) {
    const size_t tl = std::strlen(tag);
    return write_lines(f, n, tl + 3 * 16 + 1, [&](uint64_t i, char* p) {
        std::memcpy(p, tag, tl);
        p += tl;
        for (int c = 0; c < 3; ++c) {
            *p++ = ' ';
            p = put_g(p, v[3 * i + c]);
        }
        *p++ = '\n';
        return p;
    });
}

// "f a//a b//b c//c " lines (1-based vertex numbers).
template <typename V>
bool write_face_lines(std::FILE* f, uint64_t num_tris, V vertex) {
    return write_lines(f, num_tris, 2 + 3 * (2 * 20 + 3) + 1, [&](uint64_t t, char* p) {
        *p++ = 'f';
        for (int c = 0; c < 3; ++c) {
            const unsigned long long v = vertex(3 * t + c);
            *p++ = ' ';
            p = put_u(p, v);
            *p++ = '/';
            *p++ = '/';
            p = put_u(p, v);
        }
        *p++ = ' ';
        *p++ = '\n';
        return p;
    });
}

}  // namespace

extern "C" {

fsr_status fsr_xyz_info(const char* path, fsr_xyz_header* header, uint32_t* num_particles) {
    if (!path || !header || !num_particles) {
        set_error("fsr_xyz_info: null argument");
        return FSR_ERR_INVALID_ARG;
    }
    std::vector<char> buf;
    FSR_TRY(slurp(path, "fsr_xyz_info", &buf));
    size_t body = 0;
    FSR_TRY(header_from(buf, path, header, &body));
    const std::vector<size_t> cut = line_ranges(buf, body);
    std::vector<uint64_t> lines(cut.size() - 1), parts(cut.size() - 1);
    parallel_ranges(cut.size() - 1, [&](size_t k) {
        count_lines(buf.data() + cut[k], buf.data() + cut[k + 1], &lines[k], &parts[k]);
    });
    uint64_t n = 0;
    for (uint64_t p : parts) n += p;
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
    std::vector<char> buf;
    FSR_TRY(slurp(path, "fsr_xyz_read", &buf));
    size_t body = 0;
    FSR_TRY(header_from(buf, path, header, &body));
    const std::vector<size_t> cut = line_ranges(buf, body);
    const size_t T = cut.size() - 1;
    std::vector<uint64_t> lines(T), parts(T);
    parallel_ranges(T, [&](size_t k) {
        count_lines(buf.data() + cut[k], buf.data() + cut[k + 1], &lines[k], &parts[k]);
    });
    uint64_t n = 0;
    std::vector<uint64_t> row0(T), line0(T);
    uint64_t ln = 6;  // the first particle line is line 6 of the file
    for (size_t k = 0; k < T; ++k) {
        row0[k] = n;
        line0[k] = ln;
        n += parts[k];
        ln += lines[k];
    }
    if (n > capacity) {
        set_error("fsr_xyz_read: more particles than the capacity");
        return FSR_ERR_CAPACITY;
    }
    // the first line (in file order) that does not parse, per range
    std::vector<uint64_t> bad(T, 0);
    parallel_ranges(T, [&](size_t k) {
        const char* a = buf.data() + cut[k];
        const char* b = buf.data() + cut[k + 1];
        uint64_t r = row0[k], l = line0[k];
        while (a < b) {
            const char* e = static_cast<const char*>(std::memchr(a, '\n', (size_t)(b - a)));
            const char* le = e ? e : b;
            if (!blank_range(a, le)) {
                float v[4];
                const char* p = a;
                bool ok = true;
                for (int j = 0; j < 4 && ok; ++j) ok = next_float(p, le, &v[j]);
                if (!ok) {
                    bad[k] = l;
                    return;
                }
                pos[3 * r + 0] = v[0];
                pos[3 * r + 1] = v[1];
                pos[3 * r + 2] = v[2];
                dens[r] = v[3];
                ++r;
            }
            ++l;
            a = e ? e + 1 : b;
        }
    });
    for (size_t k = 0; k < T; ++k)
        if (bad[k]) {
            set_error(std::string(path) + ": cannot parse particle line " + std::to_string(bad[k]));
            return FSR_ERR_INVALID_ARG;
        }
    *num_particles = (uint32_t)n;
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
    bool ok = write_vec_lines(f, "v", pos, num_vertices);
    if (ok && nrm) ok = write_vec_lines(f, "vn", nrm, num_vertices);
    if (ok)
        ok = write_face_lines(f, num_vertices / 3,
                              [](uint64_t k) { return (unsigned long long)k + 1ull; });
    ok = (std::fclose(f) == 0) && ok;
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
    bool ok = write_vec_lines(f, "v", pos, num_vertices);
    if (ok && nrm) ok = write_vec_lines(f, "vn", nrm, num_vertices);
    if (ok)
        ok = write_face_lines(f, num_indices / 3,
                              [&](uint64_t k) { return indices[k] + 1ull; });
    ok = (std::fclose(f) == 0) && ok;
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
