#pragma once
// Pinned vertices (hard positional constraints).
// Same interface as the reference (include/constraints.h:1-36).
#include "cloth_mesh.h"

#include <vector>

struct Constraints {
    std::vector<int>   pinned_indices;    // fixed vertex indices
    std::vector<float> target_positions;  // desired positions [num_verts * 3]

    // --- Preset boundary conditions ---

    // Fix the first row (indices 0 .. ncols-1)
    void pin_top_row(const ClothMesh& mesh, int ncols);

    // Fix the two top corners (index 0 and ncols-1)
    void pin_corners(const ClothMesh& mesh, int ncols);

    // Fix an arbitrary list of vertex indices (targets = rest positions)
    void set_from_list(const std::vector<int>& indices, const ClothMesh& mesh);

    // Reset mesh CPU rest positions of pinned vertices to their targets.
    void reset_to_rest(ClothMesh& mesh) const;

    void clear();
    void print_stats() const;

    // ---- GPU support ----
    int*   d_pinned_indices = nullptr;  // device [num_pinned]
    float* d_target_pos     = nullptr;  // device [num_pinned * 3]
    int    num_pinned       = 0;

    void upload_to_gpu();
    void free_gpu();
    void apply_gpu(float* d_pos, float* d_vel) const;  // set pinned vertices
    // Like the reference struct, Constraints is copyable and does not free its
    // device buffers on destruction: call free_gpu() explicitly.
};
