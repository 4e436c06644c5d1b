#pragma once
// ClothMesh: cloth geometry and simulation state.
//
// Same members, member order and methods as the reference struct
// (include/cloth_mesh.h:18-67). Device buffers come from libclothhip
// (clh_alloc); CPU members use the small types of cloth_types.h in place of
// Eigen. All host code is built by one compiler, so no member is hidden from
// any translation unit.

#include "cloth_types.h"

#include <string>
#include <vector>

struct ClothMesh {
    // ---- GPU: simulation state ----
    float* d_pos       = nullptr;  // current positions  [N*3]
    float* d_vel       = nullptr;  // current velocities [N*3]
    float* d_mass      = nullptr;  // nodal masses       [N]

    // ---- GPU: FEM precomputed geometry ----
    int*   d_tris      = nullptr;  // triangle indices   [T*3]
    float* d_Dm_inv    = nullptr;  // Dm_inv col-major   [T*4] (2x2 per triangle)
    float* d_rest_area = nullptr;  // rest areas         [T]

    // ---- Scalars ----
    int num_verts = 0;
    int num_tris  = 0;

    // ---- CPU geometry (reference configuration) ----
    std::vector<cloth::Vec3f> rest_pos;    // [N]
    std::vector<cloth::Vec3i> triangles;   // [T]

    // ---- CPU FEM precomputed ----
    std::vector<cloth::Mat2f>   Dm_inv;    // [T] inverse reference edge matrix
    std::vector<cloth::Mat2f>   deltaUV;   // [T] rest 2-D edge matrix Dm
    std::vector<cloth::Mat6x9f> dF_dx;     // [T] deformation-gradient derivative
    std::vector<float>          rest_area; // [T]
    std::vector<float>          mass;      // [N] lumped nodal mass
    std::vector<cloth::Vec3f>   pos_cpu;   // [N] CPU positions (rest copy)
    std::vector<cloth::Vec3f>   vel_cpu;   // [N] CPU velocities (zero)

    // ---- I/O ----
    bool load_obj(const std::string& path);

    // Precompute Dm_inv, rest_area, mass from rest_pos + triangles.
    // Must be called before upload_to_gpu().
    void precompute_rest_state(float density = 0.1f);

    void print_stats() const;

    // ---- GPU lifecycle ----
    // Upload rest_pos (as the initial positions), zero vel, mass, tris,
    // Dm_inv, rest_area to the device. Frees any previous buffers first.
    void upload_to_gpu();
    void free_gpu();

    ClothMesh()  = default;
    ~ClothMesh() { free_gpu(); }
    ClothMesh(const ClothMesh&)            = delete;
    ClothMesh& operator=(const ClothMesh&) = delete;
};
