#pragma once
// CPU reference stretch solver: drop-in for the reference's
// CpuStretchReferenceSolver (include/stretch_reference.h:11-47,
// src/stretch_reference.cpp:38-280), the exact-global-solve backend that
// sim_cloth selects with --stretch-backend cpu-ref.
//
// Same public interface and semantics as the reference:
//   - stretch only (bending is ignored), 50 local-global iterations per step;
//   - local step: the closest rotation to F = Ds Dm^-1 in the triangle's
//     Gram-Schmidt frame (the SVD polar factor with the determinant fix);
//   - global step: an exact solve of K x = M y + h^2 sum kA dF/dx^T vec(R),
//     K = M + h^2 sum kA dF/dx^T dF/dx, with the rows and columns of pinned
//     degrees of freedom zeroed, a unit diagonal and the target on the right
//     (so unpinned neighbours see the pins at the origin, as in the
//     reference, src/stretch_reference.cpp:225-255);
//   - it steps the host state mesh.pos_cpu / mesh.vel_cpu, and
//     v = (x_new - x_old)(1 - damping) / dt for every vertex.
//
// The global matrix does not change between iterations or frames: it is the
// same scalar matrix for x, y and z (dF/dx = kron(D, I3)). It is therefore
// assembled and factored once per pin set, in FP64 (Cholesky of the unpinned
// vertices' matrix in reverse Cuthill-McKee order, envelope storage), where
// the reference assembles and factors a 3N x 3N FP32 matrix on every
// iteration. Each iteration is one local step and three triangular solves.
#include "cloth_mesh.h"
#include "constraints.h"
#include "sim_constraints.h"

#include <string>
#include <vector>

enum class StretchBackend {
    GpuCurrent,
    CpuReference,
};

struct CpuStretchReferenceSolver {
    explicit CpuStretchReferenceSolver(const ClothMesh& mesh,
                                       const SimConstraints& sim_cons,
                                       const Constraints& pin_cons,
                                       float dt,
                                       float gravity,
                                       float damping);

    void step(ClothMesh& mesh, const Constraints& pin_cons);
    static const char* backend_name() { return "cpu-ref"; }

    // ---- appended ----
    static constexpr int kIterations = 50;  // local-global iterations per step
    // Re-assemble and re-factor the global matrix on every iteration, as the
    // reference does (for timing comparisons; the result is the same).
    bool refactor_every_iteration = false;
    // Factorizations done so far (one per pin set unless refactoring).
    int factorizations() const { return factorizations_; }

private:
    int num_verts_ = 0;
    int num_tris_  = 0;
    float dt_      = 0.01f;
    float gravity_ = -9.8f;
    float damping_ = 0.0f;
    std::string error_;  // set when the inputs cannot be solved; step() reports it

    std::vector<float>  masses_;   // [N]
    std::vector<double> kA_;       // [T] stretch_k * rest_area (0 skips the triangle)
    std::vector<int>    tris_;     // [T*3]
    std::vector<float>  dm_inv_;   // [T*4] column-major [m00, m10, m01, m11]
    std::vector<double> coef_;     // [T*6] D[r][a]: F column r = sum_a D[r][a] x_a

    // factor of the unpinned vertices' matrix
    std::vector<char>   pinned_;   // [N] pin set the factor was built for
    std::vector<int>    order_;    // elimination order -> vertex
    std::vector<int>    slot_;     // vertex -> elimination order (-1 pinned)
    std::vector<int>    first_;    // first column of each row's envelope
    std::vector<size_t> row_at_;   // offset of each row's envelope in factor_
    std::vector<double> factor_;   // L, row by row over the envelope
    bool factored_ = false;
    int  factorizations_ = 0;

    bool factor_for(const std::vector<char>& pinned);
    void project(const double* x, int t, double R[6]) const;
};
