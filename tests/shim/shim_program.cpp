// Compiles include/fsr_reference_launchers.h against the reference's types
// and links every launcher to libfsr.so: a table of all 19 launcher names of
// ReconstructionCUDA.cuh:17-248, each taken with its reference signature.
#include "reference_types.h"
#include "fsr_reference_launchers.h"

#include <cstdio>

int main() {
    const void* launchers[] = {
        reinterpret_cast<const void*>(&bindTextures),
        reinterpret_cast<const void*>(&launchExtractionOfSurfaceParticles),
        reinterpret_cast<const void*>(&launchEstimateSurfaceVertices),
        reinterpret_cast<const void*>(&launchUpdateScalarGridValues),
        reinterpret_cast<const void*>(&launchCompactSurfaceVertex),
        reinterpret_cast<const void*>(&launchUpdateScalarGridValuesCompacted),
        reinterpret_cast<const void*>(&launchDetectValidSurfaceCubes),
        reinterpret_cast<const void*>(&launchCompactValidSurfaceCubes),
        reinterpret_cast<const void*>(&launchGenerateTriangles),
        reinterpret_cast<const void*>(&launchThrustExclusivePrefixSumScan),
        reinterpret_cast<const void*>(&launchExtractionOfSurfaceAndInvolveParticles),
        reinterpret_cast<const void*>(&launchEstimationOfSurfaceVertices),
        reinterpret_cast<const void*>(&launchCalculationOfMeanAndSmoothParticles),
        reinterpret_cast<const void*>(&launchCompactationOfSurfaceVerticesAndParticles),
        reinterpret_cast<const void*>(&launchCalculationOfTransformMatricesForParticles),
        reinterpret_cast<const void*>(&launchComputationOfScalarFieldGrid),
        reinterpret_cast<const void*>(&launchExtractionOfSurfaceParticlesForAkinci),
        reinterpret_cast<const void*>(&launchDetectionOfValidSurfaceCubesForAkinci),
        reinterpret_cast<const void*>(&launchSpatialGridBuilding),
    };
    int n = 0;
    for (const void* p : launchers) n += p != nullptr;
    std::printf("%d launchers\n", n);
    return n == 19 ? 0 : 1;
}
