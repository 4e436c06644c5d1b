#pragma once
// SimConstraints: PD constraint data and the precomputed Jacobi diagonal.
// Same interface as the reference (include/sim_constraints.h:1-78).
//
// Stretch:  per-triangle polar (Stiefel) projection of the deformation
//           gradient F = Ds * Dm_inv; geometry lives in ClothMesh.
// Bend:     cotangent-weighted quadratic bending. For hinge (v0,v1,v2,v3)
//           with shared edge (v0,v1):
//             e      = sum_i w_i x_i
//             p_proj = e / |e| * n_rest
//           w0 = cot02+cot03, w1 = cot12+cot13,
//           w2 = -(cot02+cot12), w3 = -(cot03+cot13),
//           n_rest = |sum_i w_i X_i| at rest.
// Jacobi diagonal:
//   diag_i = m_i + h^2 sum_stretch wA |D_col_i|^2 + h^2 sum_bend k w_i^2
//
// Call order: upload_to_gpu() frees every device buffer, d_jacobi_diag
// included, so precompute_jacobi_diag() must follow upload_to_gpu().

#include "cloth_mesh.h"
#include "mesh_topology.h"

#include <vector>

struct SimConstraints {
    // ---- GPU: triangle stretch ----
    float* d_tri_stretch_k = nullptr;  // [T]      per-triangle stiffness

    // ---- GPU: Jacobi system diagonal M_ii + h^2 sum w_c ----
    float* d_jacobi_diag   = nullptr;  // [N]

    // ---- GPU: cotangent bend constraints ----
    int*   d_bend_quads    = nullptr;  // [E*4]  vertex indices (v0,v1,v2,v3)
    float* d_bend_w        = nullptr;  // [E*4]  cotangent weights per vertex
    float* d_bend_n        = nullptr;  // [E]    rest curvature norm
    float* d_bend_k        = nullptr;  // [E]    stiffness per constraint

    // ---- Scalars ----
    int num_tris      = 0;
    int num_bend_cons = 0;

    // ---- CPU constraint data ----
    std::vector<float>        tri_stretch_k;  // [T]
    std::vector<cloth::Vec4i> bend_quads;     // [E] (v0,v1,v2,v3)
    std::vector<float>        bend_w;         // [E*4]
    std::vector<float>        bend_n;         // [E]
    std::vector<float>        bend_stiffness; // [E]

    // Per-triangle stretch stiffness (uniform).
    void build_stretch(const ClothMesh& mesh, float stiffness);

    // Cotangent-weighted bend constraints from the inner edges of topo.
    void build_bend(const ClothMesh& mesh, const MeshTopology& topo,
                    float stiffness);

    // Compute the Jacobi diagonal from the CPU data and upload it
    // (allocates d_jacobi_diag). Call after upload_to_gpu().
    void precompute_jacobi_diag(const ClothMesh& mesh, float dt);

    // ---- appended ----
    // The diagonal precompute_jacobi_diag uploads, computed on the host only.
    std::vector<float> compute_jacobi_diag(const ClothMesh& mesh, float dt) const;
    // Host copy of the diagonal from the last precompute_jacobi_diag call.
    std::vector<float> jacobi_diag;

    // ---- GPU lifecycle ----
    void upload_to_gpu();
    void free_gpu();

    SimConstraints()  = default;
    ~SimConstraints() { free_gpu(); }
    SimConstraints(const SimConstraints&)            = delete;
    SimConstraints& operator=(const SimConstraints&) = delete;
};
