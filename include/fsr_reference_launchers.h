/*
 * fsr_reference_launchers.h -- the reference's launcher layer, bound to libfsr.
 *
 * A drop-in replacement for FluidSurfaceReconstruction/GPU/ReconstructionCUDA.cuh
 * (:17-248): the reconstructor classes include this header in place of that one
 * and keep every call site, name and argument order.  Each launcher forwards to
 * its fsr_* entry point (include/fsr.h), which does the work on the GPU.
 *
 * Requirements at the point of inclusion (a HIP build of the reference's host
 * code, e.g. hipcc -I<this dir>):
 *   - the reference's types from GPU/Defines.h: SimParam, GridInfo, IndexRange,
 *     MatrixValue, ParticlePosition, ScalarValue, Grid<T> and its typedefs
 *     (ParticleArray, DensityGrid, IsSurfaceGrid, ...), uint;
 *   - dim3 and float3/uint3 (hip/hip_runtime.h).
 *
 * Semantics kept from the reference:
 *   - every launcher blocks until the device is idle, as each reference
 *     launcher ends with cudaDeviceSynchronize (ReconstructionCUDA.cu:1012-1481);
 *   - launchThrustExclusivePrefixSumScan returns the total to the host
 *     (ReconstructionCUDA.cu:1154-1172);
 *   - launchSpatialGridBuilding sorts the caller's particles and densities in
 *     place (ReconstructionCUDA.cu:1461-1466);
 *   - the dim3 launch geometry arguments are accepted and ignored (libfsr
 *     picks its own launch shapes).
 * Differences: errors throw std::runtime_error with fsr_last_error() instead
 * of exit(); bindTextures creates the library context (the MC tables are built
 * into libfsr; a table passed here is checked against them).
 *
 * launchUpdateScalarGridValues is declared by the reference
 * (ReconstructionCUDA.cuh:45-55) but defined nowhere, so a reference build that
 * called it would not link; here it throws std::logic_error.
 */
#ifndef FSR_REFERENCE_LAUNCHERS_H_
#define FSR_REFERENCE_LAUNCHERS_H_

#include <hip/hip_runtime_api.h>

#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>

#include "fsr.h"

/* The POD types cross the boundary by pointer cast, so their layouts must be
 * the reference's field by field (GPU/Defines.h:18-68). */
static_assert(sizeof(SimParam) == sizeof(fsr_sim_param), "SimParam layout");
static_assert(offsetof(SimParam, particleRadius) == offsetof(fsr_sim_param, particle_radius), "");
static_assert(offsetof(SimParam, particleMass) == offsetof(fsr_sim_param, particle_mass), "");
static_assert(offsetof(SimParam, smoothingRadius) == offsetof(fsr_sim_param, smoothing_radius), "");
static_assert(offsetof(SimParam, smoothingRadiusSq) ==
              offsetof(fsr_sim_param, smoothing_radius_sq), "");
static_assert(offsetof(SimParam, smoothingRadiusInv) ==
              offsetof(fsr_sim_param, smoothing_radius_inv), "");
static_assert(offsetof(SimParam, lambdaForSmoothed) ==
              offsetof(fsr_sim_param, lambda_for_smoothed), "");
static_assert(offsetof(SimParam, anisotropicRadius) ==
              offsetof(fsr_sim_param, anisotropic_radius), "");
static_assert(offsetof(SimParam, isoValue) == offsetof(fsr_sim_param, iso_value), "");
static_assert(offsetof(SimParam, scSpGridResRatio) ==
              offsetof(fsr_sim_param, sc_sp_grid_res_ratio), "");
static_assert(offsetof(SimParam, expandExtent) == offsetof(fsr_sim_param, expand_extent), "");
static_assert(offsetof(SimParam, minNumNeighbors) ==
              offsetof(fsr_sim_param, min_num_neighbors), "");
static_assert(offsetof(SimParam, spatialCellSizeScale) ==
              offsetof(fsr_sim_param, spatial_cell_size_scale), "");
static_assert(sizeof(GridInfo) == sizeof(fsr_grid_info), "GridInfo layout");
static_assert(offsetof(GridInfo, minPos) == offsetof(fsr_grid_info, min_pos), "");
static_assert(offsetof(GridInfo, maxPos) == offsetof(fsr_grid_info, max_pos), "");
static_assert(offsetof(GridInfo, cellSize) == offsetof(fsr_grid_info, cell_size), "");
static_assert(offsetof(GridInfo, resolution) == offsetof(fsr_grid_info, resolution), "");
static_assert(sizeof(IndexRange) == sizeof(fsr_index_range), "IndexRange layout");
static_assert(offsetof(IndexRange, start) == offsetof(fsr_index_range, start), "");
static_assert(offsetof(IndexRange, end) == offsetof(fsr_index_range, end), "");
static_assert(sizeof(MatrixValue) == sizeof(fsr_matrix), "MatrixValue layout");
static_assert(offsetof(MatrixValue, a21) == offsetof(fsr_matrix, a21), "");
static_assert(offsetof(MatrixValue, a33) == offsetof(fsr_matrix, a33), "");
static_assert(sizeof(ParticlePosition) == 3 * sizeof(float), "ParticlePosition is a float3");
static_assert(sizeof(ScalarValue) == sizeof(float), "ScalarValue is a float");

namespace fsr_reference {

/* The library context of the launchers (created by bindTextures, which the
 * reference calls once in ReconstructorGPU's constructor). */
inline fsr_context*& context() {
    static fsr_context* ctx = nullptr;
    return ctx;
}

inline void check(fsr_status s) {
    if (s != FSR_OK) throw std::runtime_error(std::string("libfsr: ") + fsr_last_error());
}

/* Every launcher ends like the reference's: the device is idle afterwards. */
inline void finish(fsr_status s) {
    check(s);
    if (hipDeviceSynchronize() != hipSuccess)
        throw std::runtime_error("libfsr: hipDeviceSynchronize failed");
}

inline fsr_context* ctx() {
    if (!context())
        throw std::runtime_error("libfsr: bindTextures must run before the other launchers");
    return context();
}

inline const fsr_sim_param* sim(const SimParam& p) {
    return reinterpret_cast<const fsr_sim_param*>(&p);
}
inline const fsr_grid_info* grid(const GridInfo& g) {
    return reinterpret_cast<const fsr_grid_info*>(&g);
}
template <typename T>
inline const fsr_index_range* ranges(const Grid<T>& r) {
    return reinterpret_cast<const fsr_index_range*>(r.grid);
}
template <typename T>
inline float* floats(const Grid<T>& g) {
    return reinterpret_cast<float*>(g.grid);
}

/* A grid description that carries only a resolution, for the stages that read
 * nothing else (the reference kernels read Grid<T>::resolution there). */
inline fsr_grid_info resolution_only(uint3 r) {
    fsr_grid_info g{};
    g.cell_size = 1.0f;
    g.resolution.x = r.x;
    g.resolution.y = r.y;
    g.resolution.z = r.z;
    return g;
}

}  // namespace fsr_reference

/* bindTextures (ReconstructionCUDA.cuh:17-22). */
inline void bindTextures(uint* d_edgeTable, int* d_edgeIndicesOfTriangleTable,
                         uint* d_numVerticesTable, uint* d_vertexIndicesOfEdgeTable) {
    if (!fsr_reference::context()) {
        int device = 0;
        if (hipGetDevice(&device) != hipSuccess)
            throw std::runtime_error("libfsr: hipGetDevice failed");
        fsr_reference::check(fsr_create(&fsr_reference::context(), device));
    }
    fsr_reference::check(fsr_bind_textures(fsr_reference::context(), d_edgeTable,
                                           d_edgeIndicesOfTriangleTable, d_numVerticesTable,
                                           d_vertexIndicesOfEdgeTable));
}

/* launchExtractionOfSurfaceParticles (ReconstructionCUDA.cuh:25-32). */
inline void launchExtractionOfSurfaceParticles(dim3, dim3, SimParam simParam,
                                               DensityGrid flagGrid,
                                               IsSurfaceGrid surfaceParticlesFlagGrid,
                                               ParticleIndexRangeGrid particleIndexRangeGrid) {
    const fsr_grid_info sp = fsr_reference::resolution_only(particleIndexRangeGrid.resolution);
    fsr_reference::finish(fsr_extraction_of_surface_particles(
        fsr_reference::ctx(), fsr_reference::sim(simParam), flagGrid.grid,
        surfaceParticlesFlagGrid.grid, (uint32_t)surfaceParticlesFlagGrid.size,
        fsr_reference::ranges(particleIndexRangeGrid), &sp, nullptr));
}

/* launchEstimateSurfaceVertices (ReconstructionCUDA.cuh:35-42). */
inline void launchEstimateSurfaceVertices(dim3, dim3, DensityGrid densityGrid,
                                          IsSurfaceGrid isSurfaceGrid, uint scSearchExt,
                                          SimParam params) {
    const fsr_grid_info sp = fsr_reference::resolution_only(densityGrid.resolution);
    const fsr_grid_info sc = fsr_reference::resolution_only(isSurfaceGrid.resolution);
    fsr_reference::finish(fsr_estimate_surface_vertices(fsr_reference::ctx(),
                                                        fsr_reference::sim(params),
                                                        densityGrid.grid, &sp, scSearchExt,
                                                        isSurfaceGrid.grid, &sc, nullptr));
}

/* launchUpdateScalarGridValues (ReconstructionCUDA.cuh:45-55): declared by the
 * reference, defined nowhere. */
inline void launchUpdateScalarGridValues(dim3, dim3, IsSurfaceGrid, ParticleIndexRangeGrid,
                                         ParticleArray, ScalarFieldGrid, GridInfo, GridInfo,
                                         SimParam) {
    throw std::logic_error(
        "launchUpdateScalarGridValues has no definition in the reference; "
        "use launchUpdateScalarGridValuesCompacted");
}

/* launchCompactSurfaceVertex (ReconstructionCUDA.cuh:58-65). */
inline void launchCompactSurfaceVertex(dim3, dim3, SurfaceVerticesIndexArray svIndexArray,
                                       IsSurfaceGrid isSurfaceGridScan,
                                       IsSurfaceGrid isSurfaceGrid, SimParam) {
    fsr_reference::finish(fsr_compact_surface_vertex(fsr_reference::ctx(), svIndexArray.grid,
                                                     isSurfaceGridScan.grid, isSurfaceGrid.grid,
                                                     (uint64_t)isSurfaceGrid.size, nullptr));
}

/* launchUpdateScalarGridValuesCompacted (ReconstructionCUDA.cuh:68-79). */
inline void launchUpdateScalarGridValuesCompacted(dim3, dim3,
                                                  SurfaceVerticesIndexArray svIndexArray,
                                                  uint numSurfaceVertices,
                                                  ParticleIndexRangeGrid particleIndexRangeGrid,
                                                  ParticleArray particleArray,
                                                  ScalarFieldGrid scalarFieldGrid,
                                                  GridInfo spatialGridInfo,
                                                  GridInfo scalarGridInfo, SimParam params) {
    fsr_reference::finish(fsr_update_scalar_grid_values_compacted(
        fsr_reference::ctx(), svIndexArray.grid, numSurfaceVertices,
        fsr_reference::ranges(particleIndexRangeGrid), fsr_reference::floats(particleArray),
        (uint32_t)particleArray.size, fsr_reference::floats(scalarFieldGrid),
        fsr_reference::grid(spatialGridInfo), fsr_reference::grid(scalarGridInfo),
        fsr_reference::sim(params), nullptr));
}

/* launchDetectValidSurfaceCubes (ReconstructionCUDA.cuh:82-92). */
inline void launchDetectValidSurfaceCubes(dim3, dim3, SurfaceVerticesIndexArray svIndexArray,
                                          uint numSurfaceVertices, ScalarFieldGrid vGrid,
                                          IsValidSurfaceGrid isValidSurfaceGrid,
                                          NumVerticesGrid numVerticesGrid,
                                          IsSurfaceGrid isSfGrid, SimParam params) {
    const fsr_grid_info sc = fsr_reference::resolution_only(vGrid.resolution);
    fsr_reference::finish(fsr_detect_valid_surface_cubes(
        fsr_reference::ctx(), svIndexArray.grid, numSurfaceVertices,
        fsr_reference::floats(vGrid), &sc, isValidSurfaceGrid.grid, numVerticesGrid.grid,
        isSfGrid.grid, fsr_reference::sim(params), nullptr));
}

/* launchCompactValidSurfaceCubes (ReconstructionCUDA.cuh:95-102). */
inline void launchCompactValidSurfaceCubes(dim3, dim3, ValidSurfaceIndexArray indexArray,
                                           IsValidSurfaceGrid validSurfaceGridScan,
                                           IsValidSurfaceGrid validSurafceGrid, SimParam) {
    fsr_reference::finish(fsr_compact_valid_surface_cubes(
        fsr_reference::ctx(), indexArray.grid, validSurfaceGridScan.grid, validSurafceGrid.grid,
        (uint32_t)validSurafceGrid.size, nullptr));
}

/* launchGenerateTriangles (ReconstructionCUDA.cuh:105-116). */
inline void launchGenerateTriangles(dim3, dim3, SurfaceVerticesIndexArray surfaceIndexInGridArray,
                                    ValidSurfaceIndexArray validIndexInSurfaceArray,
                                    GridInfo scGridInfo, NumVerticesGrid numVerticesGridScan,
                                    ScalarFieldGrid scalarFieldGrid, Float3Grid posGrid,
                                    Float3Grid norGrid, SimParam params) {
    fsr_reference::finish(fsr_generate_triangles(
        fsr_reference::ctx(), surfaceIndexInGridArray.grid, validIndexInSurfaceArray.grid,
        (uint32_t)validIndexInSurfaceArray.size, fsr_reference::grid(scGridInfo),
        numVerticesGridScan.grid, fsr_reference::floats(scalarFieldGrid),
        fsr_reference::floats(posGrid), fsr_reference::floats(norGrid), (uint32_t)posGrid.size,
        fsr_reference::sim(params), nullptr));
}

/* launchThrustExclusivePrefixSumScan (ReconstructionCUDA.cuh:118-119):
 * returns the total, blocking. */
inline uint launchThrustExclusivePrefixSumScan(uint* output, uint* input, uint numElements) {
    uint32_t total = 0;
    fsr_reference::finish(fsr_thrust_exclusive_prefix_sum_scan(fsr_reference::ctx(), output,
                                                                input, numElements, &total,
                                                                nullptr));
    return total;
}

/* launchExtractionOfSurfaceAndInvolveParticles (ReconstructionCUDA.cuh:134-142). */
inline void launchExtractionOfSurfaceAndInvolveParticles(
    dim3, dim3, SimParam simParam, DensityGrid flagGrid,
    NumInvolveParticlesGrid numInvolveParticlesGrid, IsSurfaceGrid surfaceParticlesFlagGrid,
    ParticleIndexRangeGrid particleIndexRangeGrid) {
    const fsr_grid_info sp = fsr_reference::resolution_only(particleIndexRangeGrid.resolution);
    fsr_reference::finish(fsr_extraction_of_surface_and_involve_particles(
        fsr_reference::ctx(), fsr_reference::sim(simParam), flagGrid.grid,
        numInvolveParticlesGrid.grid, surfaceParticlesFlagGrid.grid,
        (uint32_t)surfaceParticlesFlagGrid.size, fsr_reference::ranges(particleIndexRangeGrid),
        &sp, nullptr));
}

/* launchEstimationOfSurfaceVertices (ReconstructionCUDA.cuh:145-154). */
inline void launchEstimationOfSurfaceVertices(dim3, dim3, SimParam simParam,
                                              GridInfo scalarGridInfo,
                                              ParticleArray particleArray,
                                              IsSurfaceGrid surfaceParticlesFlagGrid,
                                              ParticleIndexRangeGrid particleIndexRangeGrid,
                                              IsSurfaceGrid isSurfaceGrid) {
    fsr_reference::finish(fsr_estimation_of_surface_vertices(
        fsr_reference::ctx(), fsr_reference::sim(simParam), fsr_reference::grid(scalarGridInfo),
        fsr_reference::floats(particleArray), surfaceParticlesFlagGrid.grid,
        (uint32_t)particleArray.size, fsr_reference::ranges(particleIndexRangeGrid),
        isSurfaceGrid.grid, nullptr));
}

/* launchCalculationOfMeanAndSmoothParticles (ReconstructionCUDA.cuh:157-166). */
inline void launchCalculationOfMeanAndSmoothParticles(
    dim3, dim3, SimParam simParam, ParticleArray particleArray, ParticleArray meanParticleArray,
    ParticleArray smoothedParticleArray, ParticleIndexRangeGrid particleIndexRangeGrid,
    GridInfo spatialGridInfo) {
    fsr_reference::finish(fsr_calculation_of_mean_and_smooth_particles(
        fsr_reference::ctx(), fsr_reference::sim(simParam), fsr_reference::floats(particleArray),
        (uint32_t)particleArray.size, fsr_reference::floats(meanParticleArray),
        fsr_reference::floats(smoothedParticleArray),
        fsr_reference::ranges(particleIndexRangeGrid), fsr_reference::grid(spatialGridInfo),
        nullptr));
}

/* launchCompactationOfSurfaceVerticesAndParticles (ReconstructionCUDA.cuh:169-180). */
inline void launchCompactationOfSurfaceVerticesAndParticles(
    dim3, dim3, SimParam, IsSurfaceGrid isSurfaceGrid, IsSurfaceGrid isSurfaceGridScan,
    NumInvolveParticlesGrid numInvolveParticlesGrid,
    NumInvolveParticlesGrid numInvolveParticlesGridScan,
    ParticleIndexRangeGrid particleIndexRangeGrid,
    SurfaceVerticesIndexArray surfaceVerticesIndexArray,
    SurfaceParticlesIndexArray surfaceParticlesIndexArray) {
    fsr_reference::finish(fsr_compactation_of_surface_vertices_and_particles(
        fsr_reference::ctx(), isSurfaceGrid.grid, isSurfaceGridScan.grid,
        (uint64_t)isSurfaceGrid.size, numInvolveParticlesGrid.grid,
        numInvolveParticlesGridScan.grid, fsr_reference::ranges(particleIndexRangeGrid),
        (uint64_t)particleIndexRangeGrid.size, surfaceVerticesIndexArray.grid,
        surfaceParticlesIndexArray.grid, nullptr));
}

/* launchCalculationOfTransformMatricesForParticles (ReconstructionCUDA.cuh:183-193),
 * with the reference's anisotropy constants (kr 4, kn 1). */
inline void launchCalculationOfTransformMatricesForParticles(
    dim3, dim3, SimParam simParam, GridInfo spatialGridInfo, ParticleArray meanParticleArray,
    ParticleArray particleArray, ParticleIndexRangeGrid particleIndexRangeGrid,
    SurfaceParticlesIndexArray surfaceParticlesIndexArray, MatrixArray svdMatricesArray) {
    fsr_reference::finish(fsr_calculation_of_transform_matrices_for_particles(
        fsr_reference::ctx(), fsr_reference::sim(simParam), nullptr,
        fsr_reference::grid(spatialGridInfo), fsr_reference::floats(meanParticleArray),
        fsr_reference::floats(particleArray), (uint32_t)particleArray.size,
        fsr_reference::ranges(particleIndexRangeGrid), surfaceParticlesIndexArray.grid,
        (uint32_t)surfaceParticlesIndexArray.size,
        reinterpret_cast<fsr_matrix*>(svdMatricesArray.grid), nullptr));
}

/* launchComputationOfScalarFieldGrid (ReconstructionCUDA.cuh:196-211). */
inline void launchComputationOfScalarFieldGrid(
    dim3, dim3, SimParam simParam, uint numSurfaceVertices, GridInfo scalarGridInfo,
    GridInfo spatialGridInfo, MatrixArray svdMatricesArray, ParticleArray smoothedParticleArray,
    ScalarFieldGrid particlesDensityArray, ParticleIndexRangeGrid particleIndexRangeGrid,
    SurfaceVerticesIndexArray surfaceVerticesIndexArray,
    NumInvolveParticlesGrid numInvolveParticlesGrid,
    NumInvolveParticlesGrid numInvolveParticlesGridScan, ScalarFieldGrid scalarFieldGrid) {
    fsr_reference::finish(fsr_computation_of_scalar_field_grid(
        fsr_reference::ctx(), fsr_reference::sim(simParam), numSurfaceVertices,
        fsr_reference::grid(scalarGridInfo), fsr_reference::grid(spatialGridInfo),
        reinterpret_cast<const fsr_matrix*>(svdMatricesArray.grid),
        fsr_reference::floats(smoothedParticleArray), fsr_reference::floats(particlesDensityArray),
        (uint32_t)smoothedParticleArray.size, fsr_reference::ranges(particleIndexRangeGrid),
        surfaceVerticesIndexArray.grid, numInvolveParticlesGrid.grid,
        numInvolveParticlesGridScan.grid, fsr_reference::floats(scalarFieldGrid), nullptr));
}

/* launchExtractionOfSurfaceParticlesForAkinci (ReconstructionCUDA.cuh:216-225). */
inline void launchExtractionOfSurfaceParticlesForAkinci(
    dim3, dim3, SimParam simParam, ParticleArray particlesArray,
    ScalarFieldGrid particlesDensityArray, ParticleIndexRangeGrid particleIndexRangeGrid,
    GridInfo spatialGridInfo, IsSurfaceGrid surfaceParticlesFlagGrid) {
    fsr_reference::finish(fsr_extraction_of_surface_particles_for_akinci(
        fsr_reference::ctx(), fsr_reference::sim(simParam), fsr_reference::floats(particlesArray),
        fsr_reference::floats(particlesDensityArray), (uint32_t)particlesArray.size,
        fsr_reference::ranges(particleIndexRangeGrid), fsr_reference::grid(spatialGridInfo),
        surfaceParticlesFlagGrid.grid, nullptr));
}

/* launchDetectionOfValidSurfaceCubesForAkinci (ReconstructionCUDA.cuh:228-237). */
inline void launchDetectionOfValidSurfaceCubesForAkinci(
    dim3, dim3, SurfaceVerticesIndexArray svIndexArray, uint numSurfaceVertices,
    ScalarFieldGrid vGrid, IsValidSurfaceGrid isValidSurfaceGrid, IsSurfaceGrid isSfGrid,
    SimParam params) {
    const fsr_grid_info sc = fsr_reference::resolution_only(vGrid.resolution);
    fsr_reference::finish(fsr_detection_of_valid_surface_cubes_for_akinci(
        fsr_reference::ctx(), svIndexArray.grid, numSurfaceVertices,
        fsr_reference::floats(vGrid), &sc, isValidSurfaceGrid.grid, isSfGrid.grid,
        fsr_reference::sim(params), nullptr));
}

/* launchSpatialGridBuilding (ReconstructionCUDA.cuh:241-248): sorts the
 * particles and densities in place. */
inline void launchSpatialGridBuilding(ParticleArray particlesArray, ScalarFieldGrid densitiesArray,
                                      uint numParticles,
                                      ParticleIndexRangeGrid particlesIndexRangerArray,
                                      DensityGrid densityGrid, GridInfo spatialGridInfo) {
    fsr_reference::finish(fsr_spatial_grid_building(
        fsr_reference::ctx(), fsr_reference::floats(particlesArray),
        fsr_reference::floats(densitiesArray), numParticles,
        reinterpret_cast<fsr_index_range*>(particlesIndexRangerArray.grid), densityGrid.grid,
        fsr_reference::grid(spatialGridInfo), nullptr));
}

#endif /* FSR_REFERENCE_LAUNCHERS_H_ */
