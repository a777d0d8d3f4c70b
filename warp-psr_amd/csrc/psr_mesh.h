// Marching cubes on a uniform model-space grid (meshing.py:150-280): the
// library's case table and the slab-wise extraction state (psr_mesh.hip).
#pragma once
#include "psr_common.h"

#include <cstdint>

namespace psr {

constexpr int MC_MAX_TRIANGLES = 12;

// Triangles of each of the 256 corner-sign cases: count[code] triangles, each
// three cube edges (edges[code][3 t + k]).
struct McTable {
  int8_t count[256];
  int8_t edges[256][3 * MC_MAX_TRIANGLES];
};
const McTable& mc_table();

struct McState;
McState* mc_create(const int shape[3], const float lower[3], const float upper[3], hipStream_t s);
void mc_free(McState* m);
// nodes of planes [start, start + width] (all y, z): width + 1 planes
int64_t mc_slab_nodes(const McState& m, int start, int width);
void mc_slab_points(McState& m, int start, int width, float* out);
// classify the cells of planes [start, start + width) from the slab's values
void mc_slab_values(McState& m, int start, int width, const float* values);
// count and list the crossing cells (ascending index); *invalid != 0 when a
// classified value was not finite
int mc_select(McState& m, int* invalid);
void mc_corner_points(McState& m, float* out);
void mc_triangulate(McState& m, const float* corner_values, int* V, int* T);
void mc_copy(const McState& m, float* vertices, int* indices);
void mc_vertex_vjp(const McState& m, const float* corner_values, const float* grad_vertices, float* grad_values);

}  // namespace psr
