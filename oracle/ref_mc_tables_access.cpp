// Accessor for the reference's marching-cubes case tables, compiled together
// with the reference's own GPU/MarchingCubesHelper.cpp (read in place under the
// reference tree) into oracle/_ref/libref_mctables.so.  Test infrastructure
// only: it lets the CPU tests compare this project's packed case tables with
// the reference's values entry by entry.
#include "MarchingCubesHelper.h"

extern "C" {
const unsigned int* ref_mc_edge_flags() { return MarchingCubesHelper::edgeFlags; }
const int* ref_mc_tri_table() { return &MarchingCubesHelper::edgeIndexesOfTriangle[0][0]; }
const unsigned int* ref_mc_num_vertices() { return MarchingCubesHelper::numVertices; }
const unsigned int* ref_mc_edge_vertices() { return &MarchingCubesHelper::vertexIndexesOfEdge[0][0]; }
const int* ref_mc_corner_offsets() { return &MarchingCubesHelper::cubeVerticesOffset[0][0]; }
}
