// Pinned-vertex constraints (reference: src/constraints.cpp:14-119).
#include "constraints.h"
#include "check.h"
#include "pd_solver.h"

#include <cstdio>
#include <cstdlib>

namespace {
void require_index(int idx, int n, const char* who)
{
    if (idx < 0 || idx >= n) {
        std::fprintf(stderr, "%s: vertex index %d out of range [0, %d)\n", who, idx, n);
        std::exit(EXIT_FAILURE);
    }
}
}  // namespace

void Constraints::pin_top_row(const ClothMesh& mesh, int ncols)
{
    if (ncols > mesh.num_verts) {
        std::fprintf(stderr, "pin_top_row: ncols %d > num_verts %d\n", ncols, mesh.num_verts);
        std::exit(EXIT_FAILURE);
    }
    std::vector<int> row((size_t)ncols);
    for (int j = 0; j < ncols; ++j) row[(size_t)j] = j;
    set_from_list(row, mesh);
}

void Constraints::pin_corners(const ClothMesh& mesh, int ncols)
{
    if (ncols > mesh.num_verts) {
        std::fprintf(stderr, "pin_corners: ncols %d > num_verts %d\n", ncols, mesh.num_verts);
        std::exit(EXIT_FAILURE);
    }
    set_from_list({0, ncols - 1}, mesh);
}

void Constraints::set_from_list(const std::vector<int>& indices, const ClothMesh& mesh)
{
    pinned_indices = indices;
    target_positions.assign((size_t)mesh.num_verts * 3, 0.0f);
    for (int idx : pinned_indices) {
        require_index(idx, mesh.num_verts, "set_from_list");
        for (int k = 0; k < 3; ++k) target_positions[(size_t)idx * 3 + k] = mesh.rest_pos[idx](k);
    }
}

void Constraints::reset_to_rest(ClothMesh& mesh) const
{
    for (int idx : pinned_indices) {
        require_index(idx, mesh.num_verts, "reset_to_rest");
        for (int k = 0; k < 3; ++k) mesh.rest_pos[idx](k) = target_positions[(size_t)idx * 3 + k];
    }
}

void Constraints::clear()
{
    pinned_indices.clear();
    target_positions.clear();
}

void Constraints::print_stats() const
{
    std::printf("=== Constraints stats ===\n");
    std::printf("  Pinned vertices: %d\n", (int)pinned_indices.size());
    if (!pinned_indices.empty()) {
        std::printf("  Indices:");
        for (int idx : pinned_indices) std::printf(" %d", idx);
        std::printf("\n");
    }
}

void Constraints::upload_to_gpu()
{
    free_gpu();
    num_pinned = (int)pinned_indices.size();
    if (num_pinned == 0) return;

    void* d = nullptr;
    CLH_CHECK(clh_alloc(&d, (size_t)num_pinned * sizeof(int)));
    CLH_CHECK(clh_memcpy(d, pinned_indices.data(), (size_t)num_pinned * sizeof(int), CLH_H2D));
    d_pinned_indices = static_cast<int*>(d);

    // compact targets: one xyz triple per pinned index, in list order
    std::vector<float> compact((size_t)num_pinned * 3);
    for (int i = 0; i < num_pinned; ++i)
        for (int k = 0; k < 3; ++k)
            compact[(size_t)i * 3 + k] = target_positions[(size_t)pinned_indices[i] * 3 + k];
    d = nullptr;
    CLH_CHECK(clh_alloc(&d, compact.size() * sizeof(float)));
    CLH_CHECK(clh_memcpy(d, compact.data(), compact.size() * sizeof(float), CLH_H2D));
    d_target_pos = static_cast<float*>(d);
}

void Constraints::free_gpu()
{
    if (d_pinned_indices) clh_free(d_pinned_indices);
    if (d_target_pos) clh_free(d_target_pos);
    d_pinned_indices = nullptr;
    d_target_pos = nullptr;
    num_pinned = 0;
}

void Constraints::apply_gpu(float* d_pos, float* d_vel) const
{
    if (num_pinned == 0 || d_pinned_indices == nullptr) return;
    launch_apply_constraints_kernel(d_pos, d_vel, d_pinned_indices, d_target_pos, num_pinned);
}
