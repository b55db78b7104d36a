/*
 * CPU oracle of the particle -> field -> mesh path.  TEST INFRASTRUCTURE ONLY:
 * it is the checker the parity tests compare libfsr against, and the CPU
 * baseline timed by bench.py; no product code links or calls it.
 *
 * FP64 restatement of the reference's kernels (FluidSurfaceReconstruction/GPU/
 * ReconstructionCUDA.cu and .cuh), OpenMP-parallel.  Discrete decisions (cell
 * ids, band membership, cube indices of a given field) follow the reference's
 * FP32 operation order exactly; continuous quantities (fields, gradients,
 * covariances, eigenvalues, mesh vertices) are computed in FP64.
 */
#ifndef FSR_ORACLE_H_
#define FSR_ORACLE_H_
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    float minx, miny, minz;
    float cell;
    uint32_t rx, ry, rz;
} orc_grid;

int orc_num_threads(void);
void orc_set_threads(int n);

/* Spatial hashing (ReconstructionCUDA.cu:1379-1481).  Outputs in sorted order.
 * ranges: u32[2C] (start, end) pairs, 0xFFFFFFFF for empty cells; occ: f32[C].
 * Particles outside the grid get key C and sort last. */
int orc_hash_sort(const float* pos, const float* dens, uint32_t n, orc_grid sp,
                  uint32_t* keys, uint32_t* perm, float* pos_out, float* dens_out,
                  uint32_t* ranges, float* occ);

/* extractionOfSurfaceParticles (ReconstructionCUDA.cu:70-120) and, with
 * involve != NULL, extractionOfSurfaceAndInvolveParticles (:466-551). */
int orc_surface_occupancy(const float* occ, const uint32_t* ranges, orc_grid sp,
                          const uint32_t* keys, uint32_t n, uint8_t* surf_cell,
                          uint32_t* flags, uint32_t* involve);

/* extractionOfSurfaceParticlesUsingColorField (ReconstructionCUDA.cu:909-971),
 * gradient in FP64.  gnorm (may be NULL): |sum|; count (may be NULL). */
int orc_surface_colorfield(const float* pos, const float* dens, uint32_t n,
                           const uint32_t* ranges, orc_grid sp, float h, float mass,
                           uint32_t* flags, double* gnorm, uint32_t* count);

/* estimationOfSurfaceVertices (ReconstructionCUDA.cu:554-595): band[V] = 1. */
int orc_band(const float* pos, const uint32_t* flags, uint32_t n, orc_grid sc, float h,
             uint8_t* band);

/* Ascending compaction of band[V]; returns B. */
uint32_t orc_compact(const uint8_t* band, uint64_t V, uint32_t* index);

/* ZB05 field (ReconstructionCUDA.cu:230-289, .cuh:262-268) on the band, FP64.
 * phi[V] is written at band vertices with sum w > 0 only; wsum[B] (may be NULL). */
int orc_field_zb05(const uint32_t* band_index, uint32_t B, const float* pos,
                   const uint32_t* ranges, orc_grid sp, orc_grid sc, float radius_sq,
                   float particle_radius, double* phi, double* wsum);

/* Marching cubes over the band (ReconstructionCUDA.cu:328-463; with cpu = 1 the
 * host MC of GPU/MarchingCubesCPU.cpp).  Cube index from phi (FP64 compare
 * against iso), vertices in FP64.  nverts[B] out; tri_pos/tri_nrm (3 doubles
 * per soup vertex) written when not NULL and the total fits `capacity`;
 * nrm_len (may be NULL) receives each soup normal's length before its
 * normalisation.  Returns the total soup vertex count (3T). */
uint64_t orc_mc(const uint32_t* band_index, uint32_t B, const uint8_t* band,
                const double* phi, orc_grid sc, float iso, int cpu, uint32_t* nverts,
                double* tri_pos, double* tri_nrm, uint64_t capacity, double* nrm_len);

/* Yu13: smoothing (ReconstructionCUDA.cu:598-658); mean/smooth: 3 doubles each. */
int orc_smooth(const float* pos, uint32_t n, const uint32_t* ranges, orc_grid sp, float radius,
               float lambda, double* mean, double* smooth);
/* FP64 cyclic Jacobi eigen-decomposition of a symmetric 3x3 (row-major a[9]);
 * d[3] eigenvalues, v[9] eigenvectors as columns. */
int orc_eigen3(const double* a, double* d, double* v);
/* Yu13: covariance -> G (ReconstructionCUDA.cu:689-827) for involved particles. */
int orc_aniso_matrices(const float* pos, const double* mean, const float* dens, uint32_t I,
                       const uint32_t* index, const uint32_t* ranges, orc_grid sp, float h,
                       float radius, float kr, float kn, uint32_t min_neighbors, double* G,
                       double* sigma, uint32_t* count_lo, uint32_t* count_hi);
/* Yu13: anisotropic field (ReconstructionCUDA.cu:830-906). */
int orc_field_aniso(const uint32_t* band_index, uint32_t B, const double* smooth,
                    const float* dens, const uint32_t* ranges, const uint32_t* involve,
                    const uint32_t* involve_off, const double* G, orc_grid sp, orc_grid sc,
                    float mass, const uint8_t* ambiguous, double* phi, uint8_t* flag_out);

/* Radius-count surface test: flags[i] = 1 iff fewer than 25 particles of the
 * clamped 3^3 block lie at 0 < d2 < h*h (FP32 d2).  count (may be NULL). */
int orc_surface_radius_count(const float* pos, uint32_t n, const uint32_t* ranges, orc_grid sp,
                             float h, uint32_t* flags, uint32_t* count);

/* Whole frame (method 0 = ZB05 occupancy band, 1 = Akinci12 colour-field band;
 * | 2 = radius-count surface particles instead), as the reconstructors run it.  Positions and densities are sorted in place.
 * phi_out (float[V], may be NULL).  counts[6] = {surface particles, B, Q, 3T,
 * out of grid, 0}.  Returns 0 on success. */
int orc_frame(int method, float* pos, float* dens, uint32_t n, orc_grid sp, orc_grid sc,
              float h, float radius_sq, float particle_radius, float mass, float iso,
              float* phi_out, uint64_t* counts);

/* Case tables (for the table parity test). */
uint64_t orc_mc_tri_packed(uint32_t c);

/* estimateSurfaceVertices (ReconstructionCUDA.cu:123-175), never launched by
 * the reference: ones on is_surface[V] (the caller zeroes). */
int orc_estimate_surface_vertices(const float* dens, orc_grid dg, uint32_t ratio, uint32_t ext,
                                  orc_grid sg, uint32_t* is_surface);
/* The isotropic TC01 field (ReconstructionCUDA.cu:178-227), never launched by
 * the reference, FP64, at the listed vertices. */
int orc_field_tc01(const uint32_t* band_index, uint32_t B, const float* pos,
                   const uint32_t* ranges, orc_grid sp, orc_grid sc, float radius_sq,
                   double* phi);

/* Deduplicated mesh (no reference counterpart): one vertex per crossed grid
 * edge used by a valid cube, ordered by (lower grid vertex, axis), interpolated
 * from the lower end in FP64; indices in the soup's order.  Returns U; writes
 * the vertices when U <= vtx_capacity and the indices when 3T <= idx_capacity;
 * *num_indices = 3T. */
uint64_t orc_mc_indexed(const uint32_t* band_index, uint32_t B, const uint8_t* band,
                        const double* phi, orc_grid sc, float iso, int cpu, double* vtx_pos,
                        double* vtx_nrm, uint64_t vtx_capacity, uint32_t* indices,
                        uint64_t idx_capacity, uint64_t* num_indices);

/* ---- Adjoint (no reference counterpart), FP64, topology held fixed ------- */

/* g_phi[V] += vertex term of the marching cubes over phi (soup order and the
 * per-cube capacity rule of the device emitter): for a soup vertex on the edge
 * (a, b), g . (c_b - c_a) * dt/dphi_a and dt/dphi_b; flat edges skipped. */
int orc_mc_vertex_backward(const uint32_t* band_index, uint32_t B, const uint8_t* band,
                           const double* phi, orc_grid sc, float iso, int cpu,
                           const double* g_tri, uint64_t capacity, double* g_phi);

/* g_pos[3n] (hash-sorted order) = sum over band vertices of g_phi[v] dphi_v/dp. */
int orc_field_zb05_backward(const uint32_t* band_index, uint32_t B, const float* pos,
                            uint32_t n, const uint32_t* ranges, orc_grid sp, orc_grid sc,
                            float radius_sq, const double* g_phi, double* g_pos);

/* The adjoint workload: a frame (method 0 ZB05 / 1 Akinci12) reconstructed in
 * FP64 from FP32 particles, then the loss
 *   L = w_grid/V sum_v (phi_v - (|x_v| - R0))^2 + w_soup/K sum_k (|y_k| - R0)^2
 * with band, cells and soup topology frozen at the base positions. */
typedef struct orc_adj orc_adj;
orc_adj* orc_adj_create(int method, const float* pos, const float* dens, uint32_t n,
                        orc_grid sp, orc_grid sc, float h, float radius_sq,
                        float particle_radius, float mass, float iso, double sphere_radius,
                        double w_grid, double w_soup);
void orc_adj_destroy(orc_adj* A);
void orc_adj_info(const orc_adj* A, uint64_t* info);
void orc_adj_state(const orc_adj* A, double* phi, double* soup_pos, uint32_t* band_index,
                   uint32_t* perm);
double orc_adj_loss(const orc_adj* A);
void orc_adj_upstream(const orc_adj* A, double* g_phi, double* g_soup);
int orc_adj_grad(const orc_adj* A, double* grad);
double orc_adj_fd(orc_adj* A, uint32_t j, int axis, double eps);

#ifdef __cplusplus
}
#endif
#endif
