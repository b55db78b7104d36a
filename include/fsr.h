/*
 * fsr.h -- C-ABI of the MI355X-native particle -> scalar field -> triangle mesh
 * library (libfsr.so).
 *
 * This header is the drop-in boundary.  It replaces the `extern "C"` launcher
 * layer of the reference (FluidSurfaceReconstruction/GPU/ReconstructionCUDA.cuh:17-248)
 * and the per-frame stage order of the reference's reconstructor classes
 * (GPU/ReconstructorGPUOursZB05.cpp:22-78, GPU/ReconstructorGPUAkinci12.cpp:34-84,
 * GPU/ReconstructorGPUOursYu13.cpp:24-64).
 *
 * Conventions (differences from the reference are deliberate and listed here):
 *   - every entry point returns fsr_status (0 = FSR_OK); nothing calls exit();
 *   - every entry point that launches GPU work takes a `stream` (a hipStream_t
 *     passed as void*, NULL = the null stream) and is asynchronous on it unless
 *     its comment says "blocking";
 *   - all buffers are device pointers owned by the caller unless noted; the
 *     context owns only its scratch arena;
 *   - POD layouts are binary compatible with the reference's SimParam, GridInfo,
 *     IndexRange and MatrixValue (GPU/Defines.h:18-68), so reference-side structs
 *     can be passed as they are;
 *   - particles are AoS float3 (12 bytes), densities float, as in the reference.
 *
 * Compiled for gfx950 only.
 */
#ifndef FSR_H_
#define FSR_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------------- */
/* POD types                                                                  */
/* ------------------------------------------------------------------------- */

typedef struct { float x, y, z; } fsr_float3;
typedef struct { uint32_t x, y, z; } fsr_uint3;

/* IndexRange (GPU/Defines.h:45): [start, end) of a cell's particles in sorted
 * order; an empty cell holds start = end = 0xFFFFFFFF. */
typedef struct { uint32_t start, end; } fsr_index_range;

/* MatrixValue (GPU/Defines.h:54-59): row-major 3x3, 36 bytes. */
typedef struct {
    float a11, a12, a13;
    float a21, a22, a23;
    float a31, a32, a33;
} fsr_matrix;

/* SimParam (GPU/Defines.h:18-32), same field order and size (48 bytes). */
typedef struct {
    float particle_radius;        /* r_p */
    float particle_mass;          /* m */
    float smoothing_radius;       /* h, the kernel radius */
    float smoothing_radius_sq;    /* h^2 */
    float smoothing_radius_inv;   /* 1/h */
    float lambda_for_smoothed;    /* Laplacian smoothing blend (Yu13) */
    float anisotropic_radius;     /* covariance radius, 2h in the reference */
    float iso_value;              /* iso-contour */
    uint32_t sc_sp_grid_res_ratio;/* spatial cell / scalar cell (reference grid rule) */
    uint32_t expand_extent;       /* unused on every path (the band extent is 3 cells) */
    uint32_t min_num_neighbors;   /* Ne (Yu13) */
    float spatial_cell_size_scale;/* hash cell = h * scale */
} fsr_sim_param;

/* GridInfo (GPU/Defines.h:62-68), 40 bytes. */
typedef struct {
    fsr_float3 min_pos;
    fsr_float3 max_pos;
    float cell_size;
    fsr_uint3 resolution;
} fsr_grid_info;

/* Anisotropy constants of the Yu & Turk kernel that the reference hard-codes.
 * Reference values: kr = 4 (singular values clamped to sigma_max / kr,
 * ReconstructionCUDA.cu:796), kn = 1 (isolated particles get G = I / (kn h),
 * ReconstructionCUDA.cu:772).  ks scales the covariance and cancels under the
 * volume normalisation (ReconstructionCUDA.cu:816-822); it is accepted and
 * ignored. */
typedef struct {
    float kr;
    float ks;
    float kn;
    uint32_t reserved;
} fsr_aniso_param;

typedef enum {
    FSR_OK = 0,
    FSR_ERR_INVALID_ARG = 1,
    FSR_ERR_HIP = 2,
    FSR_ERR_CAPACITY = 3,
    FSR_ERR_OUT_OF_MEMORY = 4,
    FSR_ERR_NO_DEVICE = 5,
    FSR_ERR_UNSUPPORTED = 6
} fsr_status;

typedef enum {
    FSR_METHOD_ZB05 = 0,      /* occupancy surface cells + ZB05 field (ReconstructorGPUOursZB05) */
    FSR_METHOD_AKINCI12 = 1,  /* colour-field surface particles + ZB05 field (ReconstructorGPUAkinci12) */
    FSR_METHOD_YU13 = 2       /* anisotropic kernel (ReconstructorGPUOursYu13), "TC01" */
} fsr_method;

/* Counters of one frame, written by the GPU into a caller-provided device
 * buffer (read it back after the stream is synchronised). */
typedef struct {
    uint32_t num_particles;
    uint32_t num_out_of_grid;        /* particles whose hash cell is outside the spatial grid */
    uint32_t num_surface_particles;
    uint32_t num_involved_particles; /* Yu13: particles that carry a G matrix */
    uint32_t num_band_vertices;      /* B: scalar vertices in the surface band */
    uint32_t num_valid_cubes;        /* Q: cubes that emit triangles */
    uint32_t num_mesh_vertices;      /* 3T: soup vertices required */
    uint32_t overflow;               /* nonzero: num_mesh_vertices exceeded the capacity */
    uint32_t band_exceptions;        /* surface particles handled by the exact band path */
    uint32_t num_unique_vertices;    /* U: vertices of the deduplicated mesh (indexed mode) */
    uint32_t reserved[6];
} fsr_frame_stats;

/* ------------------------------------------------------------------------- */
/* Context                                                                    */
/* ------------------------------------------------------------------------- */

typedef struct fsr_context fsr_context;

/* Creates a context on HIP device `device`: uploads the marching-cubes case
 * tables (replaces bindTextures, ReconstructionCUDA.cuh:17-22 /
 * ReconstructionCUDA.cu:50-65) and creates an empty scratch arena. */
fsr_status fsr_create(fsr_context** out_ctx, int device);
fsr_status fsr_destroy(fsr_context* ctx);

/* Thread-local text of the last error raised on this thread. */
const char* fsr_last_error(void);
const char* fsr_status_string(fsr_status s);
/* Library version string. */
const char* fsr_version(void);

/* ------------------------------------------------------------------------- */
/* Grid set-up                                                                */
/* ------------------------------------------------------------------------- */

/* Host helper: the reference's grid rule.  From the particle AABB, kernel
 * radius h, cell scale and scalar:spatial ratio computes the padded spatial
 * hashing grid and the scalar grid exactly as
 * GPU/ReconstructorGPU.cpp:237-267 and :99-105 do (FP32 arithmetic, truncating
 * conversion to uint). */
fsr_status fsr_grid_from_aabb(fsr_float3 aabb_min, fsr_float3 aabb_max,
                              float smoothing_radius, float cell_scale,
                              uint32_t sc_sp_ratio,
                              fsr_grid_info* out_spatial, fsr_grid_info* out_scalar);

/* ------------------------------------------------------------------------- */
/* Whole-frame entry point                                                    */
/* ------------------------------------------------------------------------- */

/* Runs one frame of `method` on the stream, with no host synchronisation:
 *   spatial hashing (positions and densities are sorted in place, as
 *   launchSpatialGridBuilding does, ReconstructionCUDA.cu:1461-1466) ->
 *   surface particles -> band -> field -> marching cubes.
 *
 *   phi        [out] float[V], V = product of sc_grid->resolution; non-band
 *                    vertices hold 0 (ZB05, Akinci12) or the byte pattern
 *                    0x64646464 (Yu13), as the reference leaves them.
 *   tri_pos    [out] float3[tri_capacity] triangle-soup positions;
 *   tri_nrm    [out] float3[tri_capacity] soup normals (may be NULL);
 *   stats      [out] device pointer to one fsr_frame_stats.
 * When 3T > tri_capacity nothing beyond the capacity is written, stats->overflow
 * is set and stats->num_mesh_vertices holds the size needed.
 * aniso may be NULL (reference constants). */
fsr_status fsr_reconstruct_frame(fsr_context* ctx, fsr_method method,
                                 const fsr_sim_param* sim, const fsr_aniso_param* aniso,
                                 const fsr_grid_info* sp_grid, const fsr_grid_info* sc_grid,
                                 float* pos, float* dens, uint32_t num_particles,
                                 float* phi, float* tri_pos, float* tri_nrm,
                                 uint32_t tri_capacity, fsr_frame_stats* stats,
                                 void* stream);

/* fsr_reconstruct_frame with a deduplicated mesh instead of the soup (no
 * counterpart in the reference, whose generateTriangles, ReconstructionCUDA.cu:
 * 383-463, emits a soup): one vertex per grid edge that is crossed by the
 * iso-surface and used by an emitted cube, numbered in ascending order of
 * (lower grid vertex, axis x, y, z), interpolated from the lower end of its
 * edge; and indices[3T] into those vertices in the soup's order (triangle k =
 * indices[3k..3k+2]).
 *   vtx_pos / vtx_nrm [out] float3[vtx_capacity] (vtx_nrm may be NULL);
 *   indices           [out] u32[idx_capacity].
 * stats->num_unique_vertices = U and stats->num_mesh_vertices = 3T; overflow is
 * set when U > vtx_capacity or 3T > idx_capacity (then nothing beyond the
 * capacities is written).  Frames of this entry point keep no adjoint state. */
fsr_status fsr_reconstruct_frame_indexed(fsr_context* ctx, fsr_method method,
                                         const fsr_sim_param* sim, const fsr_aniso_param* aniso,
                                         const fsr_grid_info* sp_grid,
                                         const fsr_grid_info* sc_grid, float* pos, float* dens,
                                         uint32_t num_particles, float* phi, float* vtx_pos,
                                         float* vtx_nrm, uint32_t vtx_capacity,
                                         uint32_t* indices, uint32_t idx_capacity,
                                         fsr_frame_stats* stats, void* stream);

/* Intermediate buffers of the last frame run on `ctx`, for inspection and
 * parity checks.  The pointer stays valid until the next call that grows the
 * arena.  `count` is the element capacity of the buffer (the live length is
 * in fsr_frame_stats). */
typedef enum {
    FSR_BUF_CELL_KEYS = 0,      /* u32[N]   sorted hash cell id of every particle */
    FSR_BUF_PERMUTATION = 1,    /* u32[N]   original index of every sorted particle */
    FSR_BUF_CELL_RANGES = 2,    /* fsr_index_range[C] */
    FSR_BUF_OCCUPANCY = 3,      /* float[C] 1.0 for occupied cells (the reference's flag grid) */
    FSR_BUF_SURFACE_FLAGS = 4,  /* u32[N]   1 for surface particles (sorted order) */
    FSR_BUF_BAND_BITS = 5,      /* u32[ceil(resx/32) * resy * resz] band bit mask */
    FSR_BUF_BAND_INDEX = 6,     /* u32[B]   band vertex linear ids, ascending */
    FSR_BUF_CUBE_NVERTS = 7,    /* u32[B]   soup vertices emitted by the cube at band entry b */
    FSR_BUF_VALID_CUBES = 8,    /* u32[Q]   band entries of valid cubes, ascending */
    FSR_BUF_INVOLVE_COUNTS = 9, /* u32[C]   Yu13 involved-particle count per cell */
    FSR_BUF_MEAN_POS = 10,      /* float4[N] Yu13 weighted mean positions (x, y, z, unused) */
    FSR_BUF_SMOOTH_POS = 11,    /* float4[N] Yu13 smoothed positions (x, y, z, unused) */
    FSR_BUF_G_MATRICES = 12,    /* fsr_matrix[I] Yu13 G of the involved particles */
    FSR_BUF_SINGULAR = 13,      /* float3[I] Yu13 covariance singular values (before clamping) */
    FSR_BUF_INVOLVED_INDEX = 14,/* u32[I]   Yu13 involved particle ids (sorted order) */
    FSR_BUF_FIELD_WORK = 15,    /* u32[3 * chunks] work of the last ZB05 field launch, per
                                 * 64-vertex chunk: candidates staged, candidates kept by the
                                 * chunk-box test, 16 x 16 vertex/candidate tiles evaluated;
                                 * after the last chunk, one word 0xFFFFFFFF */
    FSR_BUF_COLOR_WORK = 16,    /* u32[3 * waves] work of the last Akinci12 colour-field launch,
                                 * per wave: candidates staged, candidates kept by the
                                 * chunk-box test, pairs with 0 < |r| < h */
    FSR_BUF_ANISO_WORK = 17     /* u32[3 * (1 + 2 W)] work of the last Yu13 smoothing,
                                 * covariance and field launches (in that order), W = 32 x
                                 * the CU count: per launch its wave count w, then per wave
                                 * (w of them) candidates staged, candidates kept by the
                                 * chunk-box test */
} fsr_buffer;

fsr_status fsr_frame_buffer(fsr_context* ctx, fsr_buffer which,
                            void** dev_ptr, uint64_t* count);

/* ------------------------------------------------------------------------- */
/* Batched frames (no counterpart in the reference: it runs one frame at a    */
/* time, GPU/ReconstructorGPU.cpp:37-69)                                      */
/* ------------------------------------------------------------------------- */

/* One frame of a batch: the arguments of fsr_reconstruct_frame.  When src_pos
 * and src_dens are set, the frame first copies them into pos/dens (so a
 * replayed batch starts again from unsorted particles). */
typedef struct {
    const float* src_pos;
    const float* src_dens;
    float* pos;
    float* dens;
    uint32_t num_particles;
    uint32_t tri_capacity;
    fsr_grid_info sp_grid;
    fsr_grid_info sc_grid;
    fsr_sim_param sim;
    float* phi;
    float* tri_pos;
    float* tri_nrm;
    fsr_frame_stats* stats;
} fsr_frame_desc;

typedef struct fsr_batch fsr_batch;

/* A batch executor with `lanes` concurrent lanes, each with its own scratch
 * arena and stream on the device of ctx. */
fsr_status fsr_batch_create(fsr_context* ctx, uint32_t lanes, fsr_batch** out);
fsr_status fsr_batch_destroy(fsr_batch* batch);

/* Runs the frames, lanes concurrently (frames assigned largest first to the
 * least-loaded lane, by particles + scalar vertices / 16).  The first
 * call for a frame list runs it eagerly (sizing the lane arenas) and captures
 * the whole batch, forked over the lanes and joined on `stream`, into one HIP
 * graph; a later call with the same list (same descriptors) replays that graph
 * with a single launch.  timestamps (device u64[2 * num_frames], may be NULL):
 * start and end of every frame on the 100 MHz device clock. */
fsr_status fsr_batch_run(fsr_batch* batch, fsr_method method, const fsr_aniso_param* aniso,
                         const fsr_frame_desc* frames, uint32_t num_frames,
                         uint64_t* timestamps, void* stream);

/* Launch shapes of the persistent kernels: per kernel class (field, MC,
 * anisotropic, band) a limit on the workgroups per CU, which never exceeds a
 * kernel's resident count (0 = the resident count, the default).
 *
 * Autotuner: runs `frame` (blocking) with the default and the limits 2, 3 and
 * 4 for each class, times the class's stage with HIP events (median of
 * `repeats`), and keeps a limit in ctx only when it beats the default by more
 * than 1%.  choice (may be NULL) receives the chosen values, one per class. */
fsr_status fsr_autotune(fsr_context* ctx, fsr_method method, const fsr_aniso_param* aniso,
                        const fsr_frame_desc* frame, uint32_t repeats, uint32_t* choice,
                        void* stream);
/* Batch autotuner: runs the frames through the batch eagerly (blocking) with
 * each candidate limit of the lanes' field kernels (2, 1, 3, 4 workgroups per
 * CU, 0 = resident count), times the whole batch (median of `repeats`), and
 * keeps the fastest (a candidate must beat the first by more than 1%) in the
 * lanes.  The next fsr_batch_run captures a new graph.  choice (may be NULL)
 * receives the chosen value. */
fsr_status fsr_batch_autotune(fsr_batch* batch, fsr_method method, const fsr_aniso_param* aniso,
                              const fsr_frame_desc* frames, uint32_t num_frames,
                              uint32_t repeats, uint32_t* choice, void* stream);
/* Sets the launch shape directly (0 = default), one value per class. */
fsr_status fsr_set_launch_shape(fsr_context* ctx, const uint32_t* workgroups_per_cu);

/* ------------------------------------------------------------------------- */
/* Stage entry points: one per reference launcher                            */
/* ------------------------------------------------------------------------- */
/* Same buffers and meaning as the reference launcher named in each comment
 * (ReconstructionCUDA.cuh), so a caller that drives the stages itself (the
 * reconstructor classes) replaces each launchX(...) call by fsr_X(ctx, ...).
 * The reference's dim3 launch geometry arguments are dropped; buffer sizes the
 * reference read from Grid<T> are explicit.  Particles are float3[n] in the
 * cell order produced by fsr_spatial_grid_building.  Stream-ordered and
 * asynchronous unless marked blocking. */

/* bindTextures (ReconstructionCUDA.cuh:17-22): the case tables are built in;
 * when tri_table (device int[256*16]) is given it is read back (blocking) and
 * must equal the built-in table. */
fsr_status fsr_bind_textures(fsr_context* ctx, const uint32_t* edge_table,
                             const int32_t* tri_table, const uint32_t* num_vertices_table,
                             const uint32_t* vertex_indices_of_edge_table);

/* launchSpatialGridBuilding (ReconstructionCUDA.cuh:241-248): cell id, stable
 * sort of particles and densities in place, ranges[C] (empty = 0xFFFFFFFF) and
 * flag_grid[C] (1.0 occupied, 0.0 empty) of the spatial grid sp. */
fsr_status fsr_spatial_grid_building(fsr_context* ctx, float* particles, float* densities,
                                     uint32_t num_particles, fsr_index_range* ranges,
                                     float* flag_grid, const fsr_grid_info* sp, void* stream);

/* launchExtractionOfSurfaceParticles (ReconstructionCUDA.cuh:25-32):
 * surface_flags[n] = 1 for particles of surface cells, 0 otherwise. */
fsr_status fsr_extraction_of_surface_particles(fsr_context* ctx, const fsr_sim_param* sim,
                                               const float* flag_grid, uint32_t* surface_flags,
                                               uint32_t num_particles,
                                               const fsr_index_range* ranges,
                                               const fsr_grid_info* sp, void* stream);

/* launchExtractionOfSurfaceAndInvolveParticles (ReconstructionCUDA.cuh:134-142):
 * as above, plus num_involve[C] = particle count of every cell in the 5^3 block
 * of a surface cell (0 elsewhere). */
fsr_status fsr_extraction_of_surface_and_involve_particles(
    fsr_context* ctx, const fsr_sim_param* sim, const float* flag_grid, uint32_t* num_involve,
    uint32_t* surface_flags, uint32_t num_particles, const fsr_index_range* ranges,
    const fsr_grid_info* sp, void* stream);

/* launchExtractionOfSurfaceParticlesForAkinci (ReconstructionCUDA.cuh:216-225):
 * colour-field surface test (fewer than 25 neighbours in the 3^3 block or
 * gradient norm above 3). */
fsr_status fsr_extraction_of_surface_particles_for_akinci(
    fsr_context* ctx, const fsr_sim_param* sim, const float* particles, const float* densities,
    uint32_t num_particles, const fsr_index_range* ranges, const fsr_grid_info* sp,
    uint32_t* surface_flags, void* stream);

/* launchEstimationOfSurfaceVertices (ReconstructionCUDA.cuh:145-154):
 * is_surface_grid[V] = 1 for band vertices, 0 otherwise (ranges unused). */
fsr_status fsr_estimation_of_surface_vertices(fsr_context* ctx, const fsr_sim_param* sim,
                                              const fsr_grid_info* sc, const float* particles,
                                              const uint32_t* surface_flags,
                                              uint32_t num_particles,
                                              const fsr_index_range* ranges,
                                              uint32_t* is_surface_grid, void* stream);

/* launchThrustExclusivePrefixSumScan (ReconstructionCUDA.cuh:118-119):
 * exclusive scan; *total (host, may be NULL) receives the sum: blocking when
 * total is given, as the reference is. */
fsr_status fsr_thrust_exclusive_prefix_sum_scan(fsr_context* ctx, uint32_t* output,
                                                const uint32_t* input, uint32_t num_elements,
                                                uint32_t* total, void* stream);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): the exclusive scan of
 * fsr_thrust_exclusive_prefix_sum_scan with its workgroup partials in a fixed
 * scratch array of the library (4 KiB per device, set up by fsr_create and
 * held until the process ends) instead of the context arena.  Allocates
 * nothing.  d_total (device u32, may be NULL) receives the sum; nothing is
 * copied to the host and nothing blocks.  num_elements above
 * FSR_NOALLOC_SCAN_MAX_ELEMENTS returns FSR_ERR_CAPACITY.  All calls on one
 * device share its scratch: calls that overlap in time on one device (two
 * streams, two threads) race and must be ordered by the caller. */
#define FSR_NOALLOC_SCAN_MAX_ELEMENTS (1u << 30)
fsr_status fsr_exclusive_scan_noalloc(fsr_context* ctx, uint32_t* output, const uint32_t* input,
                                      uint32_t num_elements, uint32_t* d_total, void* stream);

/* launchCompactSurfaceVertex (ReconstructionCUDA.cuh:58-65):
 * sv_index[scan[v]] = v for is_surface[v] == 1, v < num_vertices. */
fsr_status fsr_compact_surface_vertex(fsr_context* ctx, uint32_t* sv_index,
                                      const uint32_t* is_surface_scan, const uint32_t* is_surface,
                                      uint64_t num_vertices, void* stream);

/* launchUpdateScalarGridValuesCompacted (ReconstructionCUDA.cuh:68-79): ZB05
 * field at the num_sv vertices of sv_index; vertices with zero weight keep the
 * caller's value (the reference's caller zeroes the grid first). */
fsr_status fsr_update_scalar_grid_values_compacted(
    fsr_context* ctx, const uint32_t* sv_index, uint32_t num_sv, const fsr_index_range* ranges,
    const float* particles, uint32_t num_particles, float* scalar_field, const fsr_grid_info* sp,
    const fsr_grid_info* sc, const fsr_sim_param* sim, void* stream);

/* launchDetectValidSurfaceCubes (ReconstructionCUDA.cuh:82-92): is_valid[b],
 * num_vertices[b] of the cube at band entry b. */
fsr_status fsr_detect_valid_surface_cubes(fsr_context* ctx, const uint32_t* sv_index,
                                          uint32_t num_sv, const float* scalar_field,
                                          const fsr_grid_info* sc, uint32_t* is_valid,
                                          uint32_t* num_vertices, const uint32_t* is_surface_grid,
                                          const fsr_sim_param* sim, void* stream);

/* launchDetectionOfValidSurfaceCubesForAkinci (ReconstructionCUDA.cuh:228-237). */
fsr_status fsr_detection_of_valid_surface_cubes_for_akinci(
    fsr_context* ctx, const uint32_t* sv_index, uint32_t num_sv, const float* scalar_field,
    const fsr_grid_info* sc, uint32_t* is_valid, const uint32_t* is_surface_grid,
    const fsr_sim_param* sim, void* stream);

/* launchCompactValidSurfaceCubes (ReconstructionCUDA.cuh:95-102). */
fsr_status fsr_compact_valid_surface_cubes(fsr_context* ctx, uint32_t* valid_index,
                                           const uint32_t* valid_scan, const uint32_t* valid,
                                           uint32_t num_sv, void* stream);

/* launchGenerateTriangles (ReconstructionCUDA.cuh:105-116): soup positions and
 * normals (float3) at num_vertices_scan[b]; capacity = soup vertices allocated. */
fsr_status fsr_generate_triangles(fsr_context* ctx, const uint32_t* sv_index,
                                  const uint32_t* valid_index, uint32_t num_valid,
                                  const fsr_grid_info* sc, const uint32_t* num_vertices_scan,
                                  const float* scalar_field, float* positions, float* normals,
                                  uint32_t capacity, const fsr_sim_param* sim, void* stream);

/* launchCalculationOfMeanAndSmoothParticles (ReconstructionCUDA.cuh:157-166):
 * mean and smoothed positions (float3[n]). */
fsr_status fsr_calculation_of_mean_and_smooth_particles(fsr_context* ctx,
                                                        const fsr_sim_param* sim,
                                                        const float* particles,
                                                        uint32_t num_particles, float* mean,
                                                        float* smoothed,
                                                        const fsr_index_range* ranges,
                                                        const fsr_grid_info* sp, void* stream);

/* launchCompactationOfSurfaceVerticesAndParticles (ReconstructionCUDA.cuh:169-180):
 * band vertex list and involved particle list (particles of cells with
 * num_involve > 0, in cell order). */
fsr_status fsr_compactation_of_surface_vertices_and_particles(
    fsr_context* ctx, const uint32_t* is_surface, const uint32_t* is_surface_scan,
    uint64_t num_vertices, const uint32_t* num_involve, const uint32_t* num_involve_scan,
    const fsr_index_range* ranges, uint64_t num_cells, uint32_t* sv_index,
    uint32_t* involved_index, void* stream);

/* launchCalculationOfTransformMatricesForParticles (ReconstructionCUDA.cuh:183-193):
 * G of every listed involved particle (aniso may be NULL: reference constants). */
fsr_status fsr_calculation_of_transform_matrices_for_particles(
    fsr_context* ctx, const fsr_sim_param* sim, const fsr_aniso_param* aniso,
    const fsr_grid_info* sp, const float* mean, const float* particles, uint32_t num_particles,
    const fsr_index_range* ranges, const uint32_t* involved_index, uint32_t num_involved,
    fsr_matrix* g, void* stream);

/* launchComputationOfScalarFieldGrid (ReconstructionCUDA.cuh:196-211):
 * anisotropic field at the band vertices (sum m/rho W - 0.5). */
fsr_status fsr_computation_of_scalar_field_grid(
    fsr_context* ctx, const fsr_sim_param* sim, uint32_t num_sv, const fsr_grid_info* sc,
    const fsr_grid_info* sp, const fsr_matrix* g, const float* smoothed, const float* densities,
    uint32_t num_particles, const fsr_index_range* ranges, const uint32_t* sv_index,
    const uint32_t* num_involve, const uint32_t* num_involve_scan, float* scalar_field,
    void* stream);

/* Launchers of the reference that its reconstructor classes never call, kept
 * for a complete launcher set: */

/* launchEstimateSurfaceVertices (ReconstructionCUDA.cuh:36-42, kernel
 * ReconstructionCUDA.cu:123-175): every density-grid cell (sp, not on the last
 * layer) whose corners straddle 0.5 sets is_surface_grid[V] = 1 on the scalar
 * vertices [c * ratio - ext, (c + 1) * ratio + ext] (clamped), ratio =
 * sim->sc_sp_grid_res_ratio.  Writes ones only (the caller zeroes). */
fsr_status fsr_estimate_surface_vertices(fsr_context* ctx, const fsr_sim_param* sim,
                                         const float* density_grid, const fsr_grid_info* sp,
                                         uint32_t sc_search_ext, uint32_t* is_surface_grid,
                                         const fsr_grid_info* sc, void* stream);

/* The isotropic TC01 field (updateScalarFieldValueTC01, ReconstructionCUDA.cu:
 * 178-227, kernelTC01 ReconstructionCUDA.cuh:252-260; the reference defines it
 * but launches it nowhere): phi = sum over the clamped 3^3 block of
 * (|x - p|^2 / (2 R^2) - 1/2)^2 for |x - p|^2 <= R^2, written at the num_sv
 * listed vertices.  Particles are float3[n] in cell order. */
fsr_status fsr_update_scalar_grid_values_tc01(fsr_context* ctx, const uint32_t* sv_index,
                                              uint32_t num_sv, const fsr_index_range* ranges,
                                              const float* particles, uint32_t num_particles,
                                              float* scalar_field, const fsr_grid_info* sp,
                                              const fsr_grid_info* sc, const fsr_sim_param* sim,
                                              void* stream);

/* ------------------------------------------------------------------------- */
/* Stage timing (the reference's optional per-stage timers,                   */
/* GPU/ReconstructorGPUOursZB05.cpp:26-76, as device events)                  */
/* ------------------------------------------------------------------------- */

/* Stages of one frame, in order. */
typedef enum {
    FSR_STAGE_HASH = 0,     /* cell ids, radix sort, gather, cell ranges */
    FSR_STAGE_SURFACE = 1,  /* surface particles and band seeds */
    FSR_STAGE_BAND = 2,     /* band dilation, compaction, cell-major work list */
    FSR_STAGE_PREP = 3,     /* Yu13: smoothing, involved list, covariance -> G (empty for
                             * ZB05 / Akinci12; the dense fill runs in the hash stage) */
    FSR_STAGE_FIELD = 4,    /* field kernel on the band */
    FSR_STAGE_MC = 5,       /* marching cubes classify, scans, compaction, emit */
    FSR_NUM_STAGES = 6
} fsr_stage;

/* Surface particles of the ZB05 and Akinci12 frames.
 *   FSR_SURFACE_REFERENCE (default): each method's own test, as the reference
 *     has it: the 27-cell occupancy test for ZB05
 *     (extractionOfSurfaceParticles, ReconstructionCUDA.cu:70-120), the
 *     27-cell count and colour-field gradient for Akinci12
 *     (ReconstructionCUDA.cu:909-971).
 *   FSR_SURFACE_RADIUS_COUNT: a particle is a surface particle iff fewer than
 *     25 other particles lie closer than the smoothing radius (squared
 *     distances in FP32, fmaf(dz, dz, fmaf(dy, dy, dx * dx)) < h * h).  No
 *     reference counterpart: it is the "fewer than 25 neighbours within the
 *     radius" band of the dam-break configuration.
 * Yu13 frames always use the reference's test.  Applies to the following
 * frames on ctx. */
typedef enum {
    FSR_SURFACE_REFERENCE = 0,
    FSR_SURFACE_RADIUS_COUNT = 1
} fsr_surface_mode;
fsr_status fsr_set_surface_mode(fsr_context* ctx, fsr_surface_mode mode);

/* ring_frames > 0: every following frame records HIP events at its stage
 * boundaries on its stream, into a ring of ring_frames slots; 0 disables. */
fsr_status fsr_set_profiling(fsr_context* ctx, uint32_t ring_frames);

/* Blocking.  Writes the stage durations (ms) of the frames recorded since the
 * last call, oldest first, at most max_frames: ms[f * FSR_NUM_STAGES + s].
 * *num_frames receives the number written; the record is then cleared. */
fsr_status fsr_stage_times(fsr_context* ctx, float* ms, uint32_t max_frames,
                           uint32_t* num_frames);

/* ------------------------------------------------------------------------- */
/* Frame files (host only, no device work): the reference's .xyz input and    */
/* .obj output, and the constants its reconstructor classes set               */
/* ------------------------------------------------------------------------- */

/* Header of a .xyz frame (FSR/data_format.txt). */
typedef struct {
    fsr_float3 aabb_min;
    fsr_float3 aabb_max;
    float kernel_radius;
    float particle_radius;
    float particle_mass;
} fsr_xyz_header;

/* Reads the header and counts the particle lines of a .xyz file
 * (readParticlesFromFile, GPU/ReconstructorGPU.cpp:200-290). */
fsr_status fsr_xyz_info(const char* path, fsr_xyz_header* header, uint32_t* num_particles);

/* Reads the header and up to `capacity` particles into host buffers:
 * pos float3[capacity], dens float[capacity].  FSR_ERR_CAPACITY when the file
 * holds more. */
fsr_status fsr_xyz_read(const char* path, fsr_xyz_header* header, float* pos, float* dens,
                        uint32_t capacity, uint32_t* num_particles);

/* Writes a triangle soup (host float3[num_vertices] positions and normals,
 * normals may be NULL) as the reference's .obj
 * (saveFluidSurfaceObjToFile, GPU/ReconstructorGPUOurs.cpp:20-100). */
fsr_status fsr_obj_write_soup(const char* path, const float* pos, const float* nrm,
                              uint32_t num_vertices);

/* Writes a deduplicated mesh (host buffers from fsr_reconstruct_frame_indexed)
 * in the same .obj dialect: v, vn, then "f a//a b//b c//c" (1-based). */
fsr_status fsr_obj_write_indexed(const char* path, const float* pos, const float* nrm,
                                 uint32_t num_vertices, const uint32_t* indices,
                                 uint32_t num_indices);

/* Writes the intermediates of the last frame run on ctx as the reference's
 * .vis file for its Visualizer (saveMiddleDataToVisFile:
 * GPU/ReconstructorGPUOursZB05.cpp:104-231, ReconstructorGPUAkinci12.cpp:290-413,
 * ReconstructorGPUOursYu13.cpp:285-431): the two grids, the particle radius,
 * the sorted particles, the occupied hash cells, the smoothed particles (Yu13),
 * the band vertices, the surface particles, the involved particles (Yu13), the
 * valid cubes (band entries), the mesh file name and the smoothing radius.
 * `sorted_pos` is the frame's (device) position buffer after the frame, and
 * `stats` a host copy of its fsr_frame_stats.  Blocking: copies to the host
 * with the device synchronised, as the reference does. */
fsr_status fsr_vis_write(fsr_context* ctx, const char* path, fsr_method method,
                         const fsr_grid_info* sp_grid, const fsr_grid_info* sc_grid,
                         const fsr_sim_param* sim, const float* sorted_pos,
                         uint32_t num_particles, const fsr_frame_stats* stats,
                         const char* obj_name);

/* The simulation parameters a reconstructor class derives from a .xyz header
 * and its own constants (readParticlesFromFile, GPU/ReconstructorGPU.cpp:226-245;
 * onInitialization of ReconstructorGPUOursZB05.cpp:85-97,
 * ReconstructorGPUAkinci12.cpp:93-106, ReconstructorGPUOursYu13.cpp:21,78-99). */
fsr_status fsr_sim_param_for(fsr_method method, const fsr_xyz_header* header,
                             fsr_sim_param* out);

/* ------------------------------------------------------------------------- */
/* Adjoint (no counterpart in the reference, which has no backward pass):     */
/* dL/d(particle positions) through the ZB05 field                            */
/* (updateScalarFieldValueZB05, ReconstructionCUDA.cu:230-289) and the        */
/* marching-cubes vertex interpolation (generateTriangles,                    */
/* ReconstructionCUDA.cu:383-463), topology held fixed.  ZB05 and Akinci12    */
/* frames only (both use the ZB05 field).                                     */
/* ------------------------------------------------------------------------- */

/* enable != 0: every following fsr_reconstruct_frame on ctx keeps what the
 * adjoint needs (16 bytes per band vertex beyond the forward).  The state of
 * the last frame stays valid until the next frame on ctx; the particle,
 * phi and arena buffers of that frame must not change in between. */
fsr_status fsr_set_adjoint(fsr_context* ctx, int enable);

/* Adds the vertex term of the last frame to g_phi (float[V], dL/dphi on the
 * scalar grid, updated in place at band vertices only):
 *   g_phi[a] += sum over emitted soup vertices k on an edge (a, b) of
 *               g_tri_pos[k] . (c_b - c_a) * dt/dphi_a.
 * g_tri_pos: float[3 * num_mesh_vertices] of that frame (soup order).
 * Deterministic (gather, no atomics).  Enqueued on `stream`; no host sync. */
fsr_status fsr_mc_vertex_backward(fsr_context* ctx, const float* g_tri_pos, float* g_phi,
                                  void* stream);

/* g_pos (float[3 * num_particles], the caller's particle order of the last
 * frame) = sum over band vertices v of g_phi[v] * dphi_v/dx_p.  Particles
 * outside the hash grid get 0.  Deterministic.  Enqueued on `stream`. */
fsr_status fsr_field_zb05_backward(fsr_context* ctx, const float* g_phi, float* g_pos,
                                   void* stream);

#ifdef __cplusplus
}
#endif

#endif /* FSR_H_ */
