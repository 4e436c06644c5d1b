#pragma once
// MeshTopology: CPU-only edge and adjacency analysis.
// Same interface as the reference (include/mesh_topology.h:1-30).

#include "cloth_mesh.h"

#include <vector>

struct MeshTopology {
    struct EdgeTopo {
        int   v0, v1;      // vertex indices, always v0 < v1
        int   tri_a;       // first adjacent triangle
        int   tri_b;       // second adjacent triangle; -1 = boundary edge
        float rest_len;    // |rest_pos[v1] - rest_pos[v0]|
    };

    std::vector<EdgeTopo>          edges;        // all unique edges
    std::vector<std::vector<int>>  vert_to_tris; // [N] adjacency lists
    std::vector<cloth::Vec3i>      tri_to_tris;  // [T] neighbour across the edge opposite each corner

    // Number of inner (non-boundary) edges.
    int num_inner_edges() const;

    // Build topology from mesh.rest_pos and mesh.triangles.
    static MeshTopology build(const ClothMesh& mesh);
};
