#pragma once
#include "cloth_mesh.h"

#include <cstdint>

// Generate a regular square cloth mesh (include/mesh_generator.h:1-15).
// Vertices are (j*size, 0, i*size): the cloth lies in the x-z plane.
//
// nrows, ncols : number of vertex rows and columns
// size         : edge length of each grid cell
// type         : triangulation pattern
//                  0 = uniform diagonal "\"
//                  1 = checkerboard (alternating "\" and "/")
//                  2 = uniform diagonal "/"
//                  3 = cross pattern: one centre vertex per cell, 4 triangles
void generate_square_cloth(int nrows, int ncols, float size, int type,
                           ClothMesh& mesh);

// Irregular grid cloth (no counterpart in the reference): the diagonal of
// every quad is drawn at random ("\" or "/"), and every interior rest vertex
// is moved in the x-z plane by a uniform offset in [-jitter, +jitter] * size
// on each axis. Draws come from SplitMix64 seeded with `seed`, in the order:
// per interior vertex (row-major) x then z, then per quad (row-major) one
// diagonal bit.
void generate_irregular_cloth(int nrows, int ncols, float size, float jitter,
                              uint64_t seed, ClothMesh& mesh);
