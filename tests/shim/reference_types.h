// Declarations of the reference's host types (FluidSurfaceReconstruction/GPU/
// Defines.h:14-92: SimParam, Grid<T> and its typedefs, IndexRange,
// ParticlePosition, ScalarValue, MatrixValue, GridInfo), written here so that
// include/fsr_reference_launchers.h can be compiled without the reference
// tree.  Field order and types follow Defines.h; dim3 / float3 / uint3 come
// from hip/hip_runtime.h.
#pragma once
#include <hip/hip_runtime.h>

typedef unsigned int uint;
typedef unsigned long long ulonglong;

struct SimParam {
    float particleRadius, particleMass, smoothingRadius, smoothingRadiusSq, smoothingRadiusInv;
    float lambdaForSmoothed, anisotropicRadius, isoValue;
    uint scSpGridResRatio, expandExtent, minNumNeighbors;
    float spatialCellSizeScale;
};

template <typename T>
struct Grid {
    T* grid;
    ulonglong size;
    ulonglong count;
    uint3 resolution;
};

struct IndexRange { uint start, end; };
struct ParticlePosition { float3 pos; };
struct ScalarValue { float value; };
struct MatrixValue { float a11, a12, a13, a21, a22, a23, a31, a32, a33; };
struct GridInfo {
    float3 minPos;
    float3 maxPos;
    float cellSize;
    uint3 resolution;
};

typedef Grid<uint> UintGrid;
typedef Grid<float> FloatGrid;
typedef Grid<float3> Float3Grid;
typedef Grid<float> DensityGrid;
typedef Grid<ParticlePosition> ParticleArray;
typedef Grid<IndexRange> ParticleIndexRangeGrid;
typedef Grid<ScalarValue> ScalarFieldGrid;
typedef Grid<uint> IsSurfaceGrid;
typedef Grid<uint> NumInvolveParticlesGrid;
typedef Grid<uint> NumVerticesGrid;
typedef Grid<uint> IsValidSurfaceGrid;
typedef Grid<uint> SurfaceParticlesIndexArray;
typedef Grid<uint> SurfaceVerticesIndexArray;
typedef Grid<uint> ValidSurfaceIndexArray;
typedef Grid<MatrixValue> MatrixArray;
