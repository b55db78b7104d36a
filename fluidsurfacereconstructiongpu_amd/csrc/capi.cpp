// C-ABI of libfsr: context, scratch arena, error reporting, the reference's
// grid rule and the whole-frame pipeline.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <string>

#include "common.hpp"

namespace fsr {

static thread_local std::string g_last_error;

void set_error(const std::string& msg) { g_last_error = msg; }

fsr_status hip_fail(hipError_t e, const char* what, const char* file, int line) {
    g_last_error = std::string(what) + " failed: " + hipGetErrorString(e) + " (" + file + ":" +
                   std::to_string(line) + ")";
    return e == hipErrorOutOfMemory ? FSR_ERR_OUT_OF_MEMORY : FSR_ERR_HIP;
}

fsr_status arena_get(fsr_context* ctx, ArenaSlot s, size_t bytes, void** out,
                     uint64_t count_for_report) {
    Arena& a = ctx->arena;
    if (bytes > a.bytes[s] && ctx->capturing) {
        set_error("scratch arena too small inside a graph capture (run the frames eagerly first)");
        return FSR_ERR_CAPACITY;
    }
    if (bytes > a.bytes[s]) {
        // grow (never shrinks); other slots may be in use by queued work
        FSR_HIP(hipDeviceSynchronize());
        if (a.ptr[s]) FSR_HIP(hipFree(a.ptr[s]));
        a.ptr[s] = nullptr;
        a.bytes[s] = 0;
        size_t want = bytes + bytes / 8;  // headroom so small growth does not reallocate
        want = (want + 255) & ~(size_t)255;
        hipError_t e = hipMalloc(&a.ptr[s], want);
        if (e != hipSuccess) {
            a.ptr[s] = nullptr;
            return hip_fail(e, "hipMalloc(arena)", __FILE__, __LINE__);
        }
        a.bytes[s] = want;
    }
    a.count[s] = count_for_report;
    *out = a.ptr[s];
    return FSR_OK;
}

void prof_mark(fsr_context* ctx, int k, hipStream_t s) {
    if (!ctx->prof_ring) return;
    const uint32_t slot = ctx->prof_frames % ctx->prof_ring;
    (void)hipEventRecord(ctx->prof_events[slot * (FSR_NUM_STAGES + 1) + k], s);
    if (k == FSR_NUM_STAGES) ctx->prof_frames++;
}

static void prof_release(fsr_context* ctx) {
    for (hipEvent_t e : ctx->prof_events) (void)hipEventDestroy(e);
    ctx->prof_events.clear();
    ctx->prof_ring = 0;
    ctx->prof_frames = 0;
}

int persistent_blocks(fsr_context* ctx, int threads, int per_cu, TuneClass cls) {
    (void)threads;
    // a tuned value limits the workgroups per CU below the kernel's resident
    // count (per_cu); it never raises it: workgroups beyond the resident ones
    // would only start late
    const int t = ctx->tune[cls];
    return std::max(1, ctx->num_cus * (t > 0 ? std::min(t, per_cu) : per_cu));
}

static fsr_status check_grid(const fsr_grid_info* g, const char* name) {
    if (!g) {
        set_error(std::string(name) + " is null");
        return FSR_ERR_INVALID_ARG;
    }
    if (g->resolution.x == 0 || g->resolution.y == 0 || g->resolution.z == 0 ||
        !(g->cell_size > 0.0f)) {
        set_error(std::string(name) + ": resolution and cell size must be positive");
        return FSR_ERR_INVALID_ARG;
    }
    const uint64_t n = grid_count(*g);
    if (n >= 0xFFFFFFFFull) {
        set_error(std::string(name) + ": more than 2^32-1 vertices (32-bit linear ids)");
        return FSR_ERR_INVALID_ARG;
    }
    return FSR_OK;
}

}  // namespace fsr

using namespace fsr;

extern "C" {

const char* fsr_last_error(void) { return g_last_error.c_str(); }

const char* fsr_status_string(fsr_status s) {
    switch (s) {
        case FSR_OK: return "ok";
        case FSR_ERR_INVALID_ARG: return "invalid argument";
        case FSR_ERR_HIP: return "HIP runtime error";
        case FSR_ERR_CAPACITY: return "output capacity exceeded";
        case FSR_ERR_OUT_OF_MEMORY: return "out of device memory";
        case FSR_ERR_NO_DEVICE: return "no HIP device";
        case FSR_ERR_UNSUPPORTED: return "unsupported";
    }
    return "unknown status";
}

const char* fsr_version(void) { return "fsr 0.1.0 (gfx950)"; }

fsr_status fsr_create(fsr_context** out_ctx, int device) {
    if (!out_ctx) {
        set_error("fsr_create: out_ctx is null");
        return FSR_ERR_INVALID_ARG;
    }
    *out_ctx = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) {
        set_error("fsr_create: no HIP device visible");
        return FSR_ERR_NO_DEVICE;
    }
    if (device < 0 || device >= n) {
        set_error("fsr_create: device index out of range");
        return FSR_ERR_INVALID_ARG;
    }
    FSR_HIP(hipSetDevice(device));
    FSR_TRY(noalloc_scratch_init(device));
    fsr_context* ctx = new fsr_context();
    ctx->device = device;

    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) == hipSuccess &&
        cus > 0)
        ctx->num_cus = cus;
    *out_ctx = ctx;
    return FSR_OK;
}

fsr_status fsr_destroy(fsr_context* ctx) {
    if (!ctx) return FSR_OK;
    (void)hipSetDevice(ctx->device);
    (void)hipDeviceSynchronize();
    prof_release(ctx);
    for (int i = 0; i < A_NUM_SLOTS; ++i)
        if (ctx->arena.ptr[i]) (void)hipFree(ctx->arena.ptr[i]);
    delete ctx;
    return FSR_OK;
}

fsr_status fsr_grid_from_aabb(fsr_float3 aabb_min, fsr_float3 aabb_max, float smoothing_radius,
                              float cell_scale, uint32_t sc_sp_ratio, fsr_grid_info* out_spatial,
                              fsr_grid_info* out_scalar) {
    if (!out_spatial || !out_scalar || !(smoothing_radius > 0.0f) || !(cell_scale > 0.0f) ||
        sc_sp_ratio == 0) {
        set_error("fsr_grid_from_aabb: invalid argument");
        return FSR_ERR_INVALID_ARG;
    }
    // GPU/ReconstructorGPU.cpp:239 and :254-267, all in float
    fsr_grid_info sp;
    sp.cell_size = smoothing_radius * cell_scale;
    sp.min_pos.x = aabb_min.x - sp.cell_size;
    sp.min_pos.y = aabb_min.y - sp.cell_size;
    sp.min_pos.z = aabb_min.z - sp.cell_size;
    sp.max_pos.x = aabb_max.x + sp.cell_size;
    sp.max_pos.y = aabb_max.y + sp.cell_size;
    sp.max_pos.z = aabb_max.z + sp.cell_size;
    volatile float qx = (sp.max_pos.x - sp.min_pos.x) / sp.cell_size + 1.0f;
    volatile float qy = (sp.max_pos.y - sp.min_pos.y) / sp.cell_size + 1.0f;
    volatile float qz = (sp.max_pos.z - sp.min_pos.z) / sp.cell_size + 1.0f;
    sp.resolution.x = (uint32_t)qx;
    sp.resolution.y = (uint32_t)qy;
    sp.resolution.z = (uint32_t)qz;
    // GPU/ReconstructorGPU.cpp:99-105
    fsr_grid_info sc;
    sc.cell_size = sp.cell_size / (float)sc_sp_ratio;
    sc.min_pos = sp.min_pos;
    sc.max_pos = sp.max_pos;
    volatile float sx = (sc.max_pos.x - sc.min_pos.x) / sc.cell_size + 1.0f;
    volatile float sy = (sc.max_pos.y - sc.min_pos.y) / sc.cell_size + 1.0f;
    volatile float sz = (sc.max_pos.z - sc.min_pos.z) / sc.cell_size + 1.0f;
    sc.resolution.x = (uint32_t)sx;
    sc.resolution.y = (uint32_t)sy;
    sc.resolution.z = (uint32_t)sz;
    *out_spatial = sp;
    *out_scalar = sc;
    return FSR_OK;
}

}  // extern "C"

namespace {

// Output of the marching-cubes stage: a triangle soup, or a deduplicated mesh.
struct MeshOut {
    bool indexed = false;
    float* tri_pos = nullptr;    // soup
    float* tri_nrm = nullptr;
    uint32_t tri_capacity = 0;   // soup vertices, or indices in indexed mode
    float* vtx_pos = nullptr;    // indexed
    float* vtx_nrm = nullptr;
    uint32_t vtx_capacity = 0;
    uint32_t* indices = nullptr;
};

fsr_status frame_impl(const char* who, fsr_context* ctx, fsr_method method,
                      const fsr_sim_param* sim, const fsr_aniso_param* aniso,
                      const fsr_grid_info* sp_grid, const fsr_grid_info* sc_grid, float* pos,
                      float* dens, uint32_t num_particles, float* phi, const MeshOut& mo,
                      fsr_frame_stats* stats, void* stream) {
    if (!ctx || !sim || !stats || !phi) {
        set_error(std::string(who) + ": ctx, sim, phi and stats are required");
        return FSR_ERR_INVALID_ARG;
    }
    if (num_particles && (!pos || !dens)) {
        set_error(std::string(who) + ": particle buffers are null");
        return FSR_ERR_INVALID_ARG;
    }
    if (method != FSR_METHOD_ZB05 && method != FSR_METHOD_AKINCI12 && method != FSR_METHOD_YU13) {
        set_error(std::string(who) + ": unknown method");
        return FSR_ERR_INVALID_ARG;
    }
    const uint32_t tri_capacity = mo.tri_capacity;
    const bool yu13 = method == FSR_METHOD_YU13;
    fsr_aniso_param ap;
    ap.kr = 4.0f;
    ap.ks = 1.0f;
    ap.kn = 1.0f;
    ap.reserved = 0;
    if (aniso) ap = *aniso;
    if (yu13 && !(ap.kr >= 1.0f && ap.kn > 0.0f && sim->anisotropic_radius > 0.0f)) {
        set_error("fsr_reconstruct_frame: anisotropic parameters need kr >= 1, kn > 0, r > 0");
        return FSR_ERR_INVALID_ARG;
    }
    FSR_TRY(check_grid(sp_grid, "sp_grid"));
    FSR_TRY(check_grid(sc_grid, "sc_grid"));
    FSR_HIP(hipSetDevice(ctx->device));
    hipStream_t s = (hipStream_t)stream;

    const fsr_grid_info& spg = *sp_grid;
    const fsr_grid_info& scg = *sc_grid;
    const GridDev sp = to_dev(spg), sc = to_dev(scg);
    const uint32_t n = num_particles;
    const uint64_t C = grid_count(spg);
    const uint64_t V = grid_count(scg);
    const uint64_t W = (uint64_t)row_words(sc.rx) * sc.ry * sc.rz;

    Counters* counters;
    FSR_TRY(arena_typed(ctx, A_COUNTERS, 1, &counters));
    uint32_t *flags, *sbits, *tmp, *bits, *exc;
    FSR_TRY(arena_typed(ctx, A_SURF_FLAGS, n, &flags));
    FSR_TRY(arena_typed(ctx, A_SBITS, 8ull * W, &sbits));  // byte seed mask
    FSR_TRY(arena_typed(ctx, A_BAND_TMP, W, &tmp));
    FSR_TRY(arena_typed(ctx, A_BAND_BITS, W, &bits));
    FSR_TRY(arena_typed(ctx, A_EXCEPTIONS, n, &exc));
    // counters, stats and the band seed mask start at zero; cleared in the
    // launch that initialises the hash grid
    FillList fills;
    fills.add(counters, 0u, sizeof(Counters) / 4);
    fills.add(stats, 0u, sizeof(fsr_frame_stats) / 4);
    fills.add(sbits, 0u, 8ull * W);
    // the dense field outside the band keeps the fill: 0, or the byte 100 for
    // Yu13 (the reference's memset of the Yu13 grid)
    fills.add(phi, yu13 ? 0x64646464u : 0u, V);

    // 1. spatial hashing
    prof_mark(ctx, FSR_STAGE_HASH, s);
    HashOut ho;
    FSR_TRY(spatial_hash(ctx, pos, dens, n, spg, stats, &ho, s, &fills));
    prof_mark(ctx, FSR_STAGE_SURFACE, s);

    // 2. surface particles and band seeds
    SurfaceArgs sa;
    sa.sorted = ho.sorted;
    sa.keys = ho.keys;
    sa.ranges = ho.ranges;
    sa.occupancy = ho.occupancy;
    sa.n = n;
    sa.sp = sp;
    sa.sc = sc;
    sa.h = sim->smoothing_radius;
    sa.flags = flags;
    sa.sbits = sbits;
    sa.exceptions = exc;
    sa.counters = counters;
    sa.st = stats;
    sa.all_exceptions = band_dilation_exact(scg, sim->smoothing_radius) ? 0 : 1;
    sa.radius_count = (!yu13 && ctx->surface_mode == FSR_SURFACE_RADIUS_COUNT) ? 1 : 0;
    uint32_t* involve = nullptr;
    if (yu13) FSR_TRY(arena_typed(ctx, A_INVOLVE, C, &involve));
    if (method == FSR_METHOD_AKINCI12 || sa.radius_count) {
        FSR_TRY(surface_colorfield(ctx, sa, sim->particle_mass, s));
    } else {
        // the occupancy test per cell, then flags and band seeds per particle
        uint8_t* surf_cell;
        FSR_TRY(arena_typed(ctx, A_SURF_CELL, C, &surf_cell));
        FSR_TRY(surface_occupancy(ctx, sa, surf_cell, involve, s));
    }

    // 3. band
    prof_mark(ctx, FSR_STAGE_BAND, s);
    BandArgs ba;
    ba.sorted = ho.sorted;
    ba.exceptions = exc;
    ba.max_exceptions = n;
    ba.sbits = sbits;
    ba.tmp = tmp;
    ba.bits = bits;
    ba.sc = sc;
    ba.h = sim->smoothing_radius;
    ba.counters = counters;
    ba.st = stats;
    ba.flags = flags;
    ba.num_particles = n;
    ba.all_exceptions = sa.all_exceptions;
    FSR_TRY(band_build(ctx, ba, s));
    BandCompactOut bc;
    // the cell-major list itself is needed by the anisotropic field and the
    // adjoint; the ZB05 field lists its chunks from the mask
    const bool materialize = yu13 || ctx->save_adjoint;
    FSR_TRY(band_compact(ctx, bits, sc, sp, yu13 ? 2 : 1, counters, stats, &bc, s, materialize));

    // 4. field
    // (the dense grid was filled in the frame's first launch)
    prof_mark(ctx, FSR_STAGE_PREP, s);
    if (yu13) {
        AnisoArgs aa;
        const uint64_t nn = n ? n : 1;
        aa.sorted = ho.sorted;
        aa.keys = ho.keys;
        aa.ranges = ho.ranges;
        aa.involve = involve;
        FSR_TRY(arena_typed(ctx, A_INVOLVE_OFF, C, &aa.involve_off));
        aa.n = n;
        aa.sp = sp;
        aa.sc = sc;
        aa.aniso_radius = sim->anisotropic_radius;
        aa.lambda = sim->lambda_for_smoothed;
        aa.h2 = sim->smoothing_radius_sq;
        aa.inv_h = sim->smoothing_radius_inv;
        aa.kr = ap.kr;
        aa.kn = ap.kn;
        aa.mass = sim->particle_mass;
        aa.min_neighbors = sim->min_num_neighbors;
        FSR_TRY(arena_typed(ctx, A_MEAN, nn, &aa.mean));
        FSR_TRY(arena_typed(ctx, A_SMOOTH, nn, &aa.smooth));
        FSR_TRY(arena_typed(ctx, A_INVOLVED_IDX, nn, &aa.involved_idx));
        FSR_TRY(arena_typed(ctx, A_GMAT, nn, &aa.gmat));
        FSR_TRY(arena_typed(ctx, A_SIGMA, nn, &aa.sigma));
        FSR_TRY(arena_typed(ctx, A_REC, 2 * nn, &aa.rec));
        aa.counters = counters;
        aa.st = stats;
        aa.noninv_moments = 1;
        FSR_TRY(aniso_run(ctx, aa, s));
        prof_mark(ctx, FSR_STAGE_FIELD, s);
        FSR_TRY(aniso_field(ctx, aa, bc, phi, s));
    } else {
        prof_mark(ctx, FSR_STAGE_FIELD, s);
        FieldArgs fa;
        fa.sorted = ho.sorted;
        fa.ranges = ho.ranges;
        fa.bc = &bc;
        fa.sp = sp;
        fa.sc = sc;
        fa.radius_sq = sim->smoothing_radius_sq;
        fa.particle_radius = sim->particle_radius;
        fa.phi = phi;
        fa.counters = counters;
        fa.saved = nullptr;
        fa.work = nullptr;
        if (ctx->save_adjoint) FSR_TRY(arena_typed(ctx, A_ADJ_VERT, V, &fa.saved));
        FSR_TRY(field_zb05(ctx, fa, s));
    }

    // 5. marching cubes
    prof_mark(ctx, FSR_STAGE_MC, s);
    McArgs ma;
    uint32_t *nverts, *vert_off, *valid_idx;
    FSR_TRY(arena_typed(ctx, A_NVERTS, V, &nverts));
    FSR_TRY(arena_typed(ctx, A_VERT_OFF, V, &vert_off));
    FSR_TRY(arena_typed(ctx, A_VALID_IDX, V, &valid_idx));
    ma.band_index = bc.band_index;
    ma.bits = bits;
    ma.phi = phi;
    ma.sc = sc;
    ma.iso = sim->iso_value;
    ma.cpu_flavor = method == FSR_METHOD_AKINCI12 ? 1 : 0;
    ma.nverts = nverts;
    ma.valid = nullptr;      // the valid-cube list comes straight from the vertex counts
    ma.valid_off = nullptr;
    ma.vert_off = vert_off;
    ma.valid_idx = valid_idx;
    ma.tri_pos = mo.tri_pos;
    ma.tri_nrm = mo.tri_nrm;
    ma.capacity = tri_capacity;
    ma.max_band = V;
    ma.st = stats;
    ma.word_off = static_cast<const uint32_t*>(ctx->arena.ptr[A_WORD_OFF]);
    ma.edge_mask = nullptr;
    ma.edge_off = nullptr;
    ma.vtx_pos = mo.vtx_pos;
    ma.vtx_nrm = mo.vtx_nrm;
    ma.vtx_capacity = mo.vtx_capacity;
    ma.indices = mo.indices;
    if (mo.indexed) {
        FSR_TRY(arena_typed(ctx, A_EDGE_MASK, V, &ma.edge_mask));
        FSR_TRY(arena_typed(ctx, A_EDGE_OFF, V, &ma.edge_off));
        FSR_TRY(mc_run_indexed(ctx, ma, s));
    } else {
        FSR_TRY(mc_run(ctx, ma, s));
    }
    prof_mark(ctx, FSR_NUM_STAGES, s);
    AdjointState& st = ctx->adj;
    st = AdjointState();
    if (ctx->save_adjoint && !yu13 && !mo.indexed) {
        st.valid = true;
        st.cpu_flavor = ma.cpu_flavor;
        st.n = n;
        st.sp = sp;
        st.sc = sc;
        st.radius_sq = sim->smoothing_radius_sq;
        st.iso = sim->iso_value;
        st.sorted = ho.sorted;
        st.ranges = ho.ranges;
        st.perm = ho.perm;
        st.bc = bc;
        st.saved = static_cast<const float4*>(ctx->arena.ptr[A_ADJ_VERT]);
        st.phi = phi;
        st.bits = bits;
        st.word_off = static_cast<const uint32_t*>(ctx->arena.ptr[A_WORD_OFF]);
        st.nverts = nverts;
        st.vert_off = vert_off;
        st.valid_idx = valid_idx;
        st.tri_capacity = tri_capacity;
        st.stats = stats;
        st.counters = counters;
    }
    return FSR_OK;
}

}  // namespace

extern "C" {

fsr_status fsr_reconstruct_frame(fsr_context* ctx, fsr_method method, const fsr_sim_param* sim,
                                 const fsr_aniso_param* aniso, const fsr_grid_info* sp_grid,
                                 const fsr_grid_info* sc_grid, float* pos, float* dens,
                                 uint32_t num_particles, float* phi, float* tri_pos,
                                 float* tri_nrm, uint32_t tri_capacity, fsr_frame_stats* stats,
                                 void* stream) {
    if (tri_capacity && !tri_pos) {
        set_error("fsr_reconstruct_frame: tri_pos is null but tri_capacity > 0");
        return FSR_ERR_INVALID_ARG;
    }
    MeshOut mo;
    mo.tri_pos = tri_pos;
    mo.tri_nrm = tri_nrm;
    mo.tri_capacity = tri_capacity;
    return frame_impl("fsr_reconstruct_frame", ctx, method, sim, aniso, sp_grid, sc_grid, pos,
                      dens, num_particles, phi, mo, stats, stream);
}

fsr_status fsr_reconstruct_frame_indexed(fsr_context* ctx, fsr_method method,
                                         const fsr_sim_param* sim, const fsr_aniso_param* aniso,
                                         const fsr_grid_info* sp_grid,
                                         const fsr_grid_info* sc_grid, float* pos, float* dens,
                                         uint32_t num_particles, float* phi, float* vtx_pos,
                                         float* vtx_nrm, uint32_t vtx_capacity,
                                         uint32_t* indices, uint32_t idx_capacity,
                                         fsr_frame_stats* stats, void* stream) {
    if ((vtx_capacity && !vtx_pos) || (idx_capacity && !indices)) {
        set_error("fsr_reconstruct_frame_indexed: null mesh buffer with a nonzero capacity");
        return FSR_ERR_INVALID_ARG;
    }
    MeshOut mo;
    mo.indexed = true;
    mo.tri_capacity = idx_capacity;
    mo.vtx_pos = vtx_pos;
    mo.vtx_nrm = vtx_nrm;
    mo.vtx_capacity = vtx_capacity;
    mo.indices = indices;
    return frame_impl("fsr_reconstruct_frame_indexed", ctx, method, sim, aniso, sp_grid,
                      sc_grid, pos, dens, num_particles, phi, mo, stats, stream);
}

fsr_status fsr_set_profiling(fsr_context* ctx, uint32_t ring_frames) {
    if (!ctx) {
        set_error("fsr_set_profiling: null context");
        return FSR_ERR_INVALID_ARG;
    }
    FSR_HIP(hipSetDevice(ctx->device));
    FSR_HIP(hipDeviceSynchronize());
    prof_release(ctx);
    if (ring_frames == 0) return FSR_OK;
    const size_t n = (size_t)ring_frames * (FSR_NUM_STAGES + 1);
    ctx->prof_events.resize(n);
    for (size_t i = 0; i < n; ++i) {
        hipError_t e = hipEventCreate(&ctx->prof_events[i]);
        if (e != hipSuccess) {
            ctx->prof_events.resize(i);
            prof_release(ctx);
            return hip_fail(e, "hipEventCreate", __FILE__, __LINE__);
        }
    }
    ctx->prof_ring = ring_frames;
    ctx->prof_frames = 0;
    return FSR_OK;
}

fsr_status fsr_stage_times(fsr_context* ctx, float* ms, uint32_t max_frames, uint32_t* num_frames) {
    if (!ctx || !ms || !num_frames) {
        set_error("fsr_stage_times: null argument");
        return FSR_ERR_INVALID_ARG;
    }
    *num_frames = 0;
    if (!ctx->prof_ring) {
        set_error("fsr_stage_times: profiling is not enabled");
        return FSR_ERR_INVALID_ARG;
    }
    uint32_t n = std::min(ctx->prof_frames, ctx->prof_ring);
    n = std::min(n, max_frames);
    const uint32_t first = ctx->prof_frames - n;
    for (uint32_t f = 0; f < n; ++f) {
        const uint32_t slot = (first + f) % ctx->prof_ring;
        hipEvent_t* ev = &ctx->prof_events[(size_t)slot * (FSR_NUM_STAGES + 1)];
        FSR_HIP(hipEventSynchronize(ev[FSR_NUM_STAGES]));
        for (int k = 0; k < FSR_NUM_STAGES; ++k) {
            float t = 0.f;
            FSR_HIP(hipEventElapsedTime(&t, ev[k], ev[k + 1]));
            ms[(size_t)f * FSR_NUM_STAGES + k] = t;
        }
    }
    *num_frames = n;
    ctx->prof_frames = 0;
    return FSR_OK;
}

fsr_status fsr_set_surface_mode(fsr_context* ctx, fsr_surface_mode mode) {
    if (!ctx || (mode != FSR_SURFACE_REFERENCE && mode != FSR_SURFACE_RADIUS_COUNT)) {
        set_error("fsr_set_surface_mode: need a context and a known mode");
        return FSR_ERR_INVALID_ARG;
    }
    ctx->surface_mode = mode;
    return FSR_OK;
}

fsr_status fsr_set_adjoint(fsr_context* ctx, int enable) {
    if (!ctx) {
        set_error("fsr_set_adjoint: null context");
        return FSR_ERR_INVALID_ARG;
    }
    ctx->save_adjoint = enable != 0;
    if (!ctx->save_adjoint) ctx->adj = AdjointState();
    return FSR_OK;
}

static fsr_status adjoint_ready(fsr_context* ctx, const char* who) {
    if (!ctx) {
        set_error(std::string(who) + ": null context");
        return FSR_ERR_INVALID_ARG;
    }
    if (!ctx->adj.valid) {
        set_error(std::string(who) +
                  ": no adjoint state; enable it with fsr_set_adjoint before a ZB05 or "
                  "Akinci12 frame (Yu13 frames have no adjoint)");
        return FSR_ERR_UNSUPPORTED;
    }
    return FSR_OK;
}

fsr_status fsr_mc_vertex_backward(fsr_context* ctx, const float* g_tri_pos, float* g_phi,
                                  void* stream) {
    FSR_TRY(adjoint_ready(ctx, "fsr_mc_vertex_backward"));
    if (!g_phi || (!g_tri_pos && ctx->adj.tri_capacity)) {
        set_error("fsr_mc_vertex_backward: g_tri_pos and g_phi are required");
        return FSR_ERR_INVALID_ARG;
    }
    FSR_HIP(hipSetDevice(ctx->device));
    return mc_vertex_backward(ctx, g_tri_pos, g_phi, (hipStream_t)stream);
}

fsr_status fsr_field_zb05_backward(fsr_context* ctx, const float* g_phi, float* g_pos,
                                   void* stream) {
    FSR_TRY(adjoint_ready(ctx, "fsr_field_zb05_backward"));
    if (!g_phi || (!g_pos && ctx->adj.n)) {
        set_error("fsr_field_zb05_backward: g_phi and g_pos are required");
        return FSR_ERR_INVALID_ARG;
    }
    if (ctx->adj.bc.radius != 1) {
        set_error("fsr_field_zb05_backward: the frame's band work list is not a ZB05 one");
        return FSR_ERR_UNSUPPORTED;
    }
    FSR_HIP(hipSetDevice(ctx->device));
    return field_zb05_backward(ctx, g_phi, g_pos, (hipStream_t)stream);
}

fsr_status fsr_frame_buffer(fsr_context* ctx, fsr_buffer which, void** dev_ptr, uint64_t* count) {
    if (!ctx || !dev_ptr) {
        set_error("fsr_frame_buffer: null argument");
        return FSR_ERR_INVALID_ARG;
    }
    ArenaSlot slot;
    switch (which) {
        case FSR_BUF_CELL_KEYS: slot = A_KEYS0; break;
        case FSR_BUF_PERMUTATION: slot = A_VALS0; break;
        case FSR_BUF_CELL_RANGES: slot = A_RANGES; break;
        case FSR_BUF_OCCUPANCY: slot = A_OCCUPANCY; break;
        case FSR_BUF_SURFACE_FLAGS: slot = A_SURF_FLAGS; break;
        case FSR_BUF_BAND_BITS: slot = A_BAND_BITS; break;
        case FSR_BUF_BAND_INDEX: slot = A_BAND_INDEX; break;
        case FSR_BUF_CUBE_NVERTS: slot = A_NVERTS; break;
        case FSR_BUF_VALID_CUBES: slot = A_VALID_IDX; break;
        case FSR_BUF_INVOLVE_COUNTS: slot = A_INVOLVE; break;
        case FSR_BUF_MEAN_POS: slot = A_MEAN; break;
        case FSR_BUF_SMOOTH_POS: slot = A_SMOOTH; break;
        case FSR_BUF_G_MATRICES: slot = A_GMAT; break;
        case FSR_BUF_SINGULAR: slot = A_SIGMA; break;
        case FSR_BUF_INVOLVED_INDEX: slot = A_INVOLVED_IDX; break;
        case FSR_BUF_FIELD_WORK: slot = A_FIELD_WORK; break;
        case FSR_BUF_COLOR_WORK: slot = A_COLOR_WORK; break;
        case FSR_BUF_ANISO_WORK: slot = A_ANISO_WORK; break;
        default:
            set_error("fsr_frame_buffer: unknown buffer");
            return FSR_ERR_INVALID_ARG;
    }
    if (which == FSR_BUF_CELL_KEYS || which == FSR_BUF_PERMUTATION) {
        // the sorted keys/ids end in the ping-pong buffer the last pass wrote
        const int passes = ctx->last_sort_passes;
        if (passes % 2 == 1) slot = (which == FSR_BUF_CELL_KEYS) ? A_KEYS1 : A_VALS1;
    }
    *dev_ptr = ctx->arena.ptr[slot];
    if (count) *count = ctx->arena.count[slot];
    return FSR_OK;
}

}  // extern "C"
