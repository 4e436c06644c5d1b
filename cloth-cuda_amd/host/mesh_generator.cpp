// Grid cloth generators. generate_square_cloth follows
// src/mesh_generator.cpp:6-91 of the reference: the vertex numbering and the
// triangle corner order are part of the interface (topology must match the
// reference bit for bit), so they are kept identical.
#include "mesh_generator.h"
#include "cloth_rng.h"

#include <cassert>
#include <cstdio>
#include <cstdlib>

void generate_square_cloth(int nrows, int ncols, float size, int type,
                           ClothMesh& mesh)
{
    if (nrows < 2 || ncols < 2) {
        std::fprintf(stderr, "generate_square_cloth: need nrows, ncols >= 2\n");
        std::exit(EXIT_FAILURE);
    }
    mesh.rest_pos.clear();
    mesh.triangles.clear();

    const int cr = nrows - 1;   // cells per column
    const int cc = ncols - 1;   // cells per row
    const bool cross = (type == 3);

    mesh.rest_pos.reserve((size_t)nrows * ncols + (cross ? (size_t)cr * cc : 0));
    for (int i = 0; i < nrows; ++i)
        for (int j = 0; j < ncols; ++j)
            mesh.rest_pos.emplace_back((float)j * size, 0.0f, (float)i * size);
    if (cross) {
        // cell-centre vertex of cell (i, j) has index nrows*ncols + i*cc + j
        for (int i = 0; i < cr; ++i)
            for (int j = 0; j < cc; ++j)
                mesh.rest_pos.emplace_back(((float)j + 0.5f) * size, 0.0f,
                                           ((float)i + 0.5f) * size);
    }
    mesh.num_verts = (int)mesh.rest_pos.size();

    mesh.triangles.reserve((size_t)cr * cc * (cross ? 4 : 2));
    for (int i = 0; i < cr; ++i) {
        for (int j = 0; j < cc; ++j) {
            const int a = i * ncols + j;          // top-left
            const int b = a + 1;                  // top-right
            const int c = a + ncols;              // bottom-left
            const int d = c + 1;                  // bottom-right
            if (cross) {
                const int m = nrows * ncols + i * cc + j;
                mesh.triangles.emplace_back(a, b, m);
                mesh.triangles.emplace_back(b, d, m);
                mesh.triangles.emplace_back(d, c, m);
                mesh.triangles.emplace_back(c, a, m);
                continue;
            }
            bool back_slash = true;                         // type 0
            if (type == 1) back_slash = ((i + j) % 2 == 0);  // checkerboard
            else if (type == 2) back_slash = false;
            if (back_slash) {
                mesh.triangles.emplace_back(a, b, d);
                mesh.triangles.emplace_back(a, d, c);
            } else {
                mesh.triangles.emplace_back(a, b, c);
                mesh.triangles.emplace_back(b, d, c);
            }
        }
    }
    mesh.num_tris = (int)mesh.triangles.size();
}

void generate_irregular_cloth(int nrows, int ncols, float size, float jitter,
                              uint64_t seed, ClothMesh& mesh)
{
    // Start from the regular grid (type 0 numbering), then jitter interior
    // rest vertices in-plane and redraw each quad's diagonal.
    generate_square_cloth(nrows, ncols, size, 0, mesh);
    cloth::SplitMix64 rng(seed);
    for (int i = 1; i < nrows - 1; ++i) {
        for (int j = 1; j < ncols - 1; ++j) {
            cloth::Vec3f& p = mesh.rest_pos[(size_t)i * ncols + j];
            const float dx = (float)rng.uniform(-jitter, jitter) * size;
            const float dz = (float)rng.uniform(-jitter, jitter) * size;
            p(0) += dx;
            p(2) += dz;
        }
    }
    mesh.triangles.clear();
    for (int i = 0; i < nrows - 1; ++i) {
        for (int j = 0; j < ncols - 1; ++j) {
            const int a = i * ncols + j, b = a + 1, c = a + ncols, d = c + 1;
            if (rng.next_u64() & 1ull) {          // "\" diagonal a-d
                mesh.triangles.emplace_back(a, b, d);
                mesh.triangles.emplace_back(a, d, c);
            } else {                              // "/" diagonal b-c
                mesh.triangles.emplace_back(a, b, c);
                mesh.triangles.emplace_back(b, d, c);
            }
        }
    }
    mesh.num_tris = (int)mesh.triangles.size();
}
