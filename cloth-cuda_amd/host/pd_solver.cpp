// PDSolver: the reference's C++ solver class (include/pd_solver.h:30-71,
// src/pd_solver.cu:445-621) on top of libclothhip. The class keeps the
// reference's contract: it caches N, T and E at construction, takes the
// mesh, constraint and pin buffers per call without storing them, mutates
// mesh.d_pos / mesh.d_vel in place, and returns with the frame complete as
// seen by a following blocking copy.
#include "pd_solver.h"
#include "check.h"
#include "cloth_mesh.h"
#include "constraints.h"
#include "sim_constraints.h"

#include <cstdio>

PDSolver::PDSolver(const PDSolverConfig& config, const ClothMesh& mesh,
                   const SimConstraints& sim_cons)
    : config_(config),
      num_verts_(mesh.num_verts),
      num_tris_(mesh.num_tris),
      num_bend_cons_(sim_cons.num_bend_cons)
{
    clh_config c;
    CLH_CHECK(clh_config_default(&c));
    c.mode           = config.mode;
    c.max_iterations = config.max_iterations;
    c.use_chebyshev  = config.use_chebyshev ? 1 : 0;
    c.rho            = config.rho;
    c.gravity        = config.gravity;
    c.dt             = config.dt;
    c.damping        = config.damping;
    c.tolerance      = (config.mode == CLH_MODE_PD) ? config.tolerance : 0.0f;
    c.check_every    = config.check_every;
    c.cheb_delay     = config.cheb_delay;
    for (int k = 0; k < 3; ++k) c.ext_accel[k] = config.external_accel[k];
    c.kernel_variant = config.kernel_variant;
    c.use_graph      = config.use_graph ? 1 : 0;
    c.relax          = config.relax;
    c.restart_rule   = config.restart_rule;
    c.restart_factor = config.restart_factor;

    clh_mesh_desc d;
    d.num_verts    = num_verts_;
    d.num_tris     = num_tris_;
    d.num_bend     = num_bend_cons_;
    d.d_tris       = nullptr;  // topology is read at the first step
    d.d_bend_quads = nullptr;
    CLH_CHECK(clh_solver_create(&c, &d, &solver_));
}

PDSolver::~PDSolver()
{
    if (solver_) clh_solver_destroy(solver_);
    solver_ = nullptr;
}

void PDSolver::reset()
{
    CLH_CHECK(clh_solver_reset(solver_));
}

void PDSolver::step(ClothMesh& mesh, const SimConstraints& sim_cons, const Constraints& pin_cons)
{
    clh_step_args a = {};
    a.d_pos         = mesh.d_pos;
    a.d_vel         = mesh.d_vel;
    a.d_mass        = mesh.d_mass;
    a.d_tris        = mesh.d_tris;
    a.d_Dm_inv      = mesh.d_Dm_inv;
    a.d_rest_area   = mesh.d_rest_area;
    // stretch runs only when T > 0 and stiffness is on the device, bending
    // only when hinges exist (src/pd_solver.cu:534, 542)
    a.d_stretch_k   = (num_tris_ > 0) ? sim_cons.d_tri_stretch_k : nullptr;
    a.d_jacobi_diag = sim_cons.d_jacobi_diag;
    const bool bend = num_bend_cons_ > 0 && sim_cons.d_bend_quads != nullptr;
    a.d_bend_quads  = bend ? sim_cons.d_bend_quads : nullptr;
    a.d_bend_w      = bend ? sim_cons.d_bend_w : nullptr;
    a.d_bend_n      = bend ? sim_cons.d_bend_n : nullptr;
    a.d_bend_k      = bend ? sim_cons.d_bend_k : nullptr;
    a.d_pin_idx     = pin_cons.d_pinned_indices;
    a.d_pin_target  = pin_cons.d_target_pos;
    a.num_pinned    = pin_cons.d_pinned_indices ? pin_cons.num_pinned : 0;
    CLH_CHECK(clh_solver_step(solver_, &a));
}

int PDSolver::last_iterations() const
{
    clh_step_stats st;
    CLH_CHECK(clh_solver_get_stats(solver_, &st));
    return st.iterations;
}

double PDSolver::last_relative_residual() const
{
    clh_step_stats st;
    CLH_CHECK(clh_solver_get_stats(solver_, &st));
    return st.rhs_norm > 0.0 ? st.residual / st.rhs_norm : 0.0;
}

bool PDSolver::last_converged() const
{
    clh_step_stats st;
    CLH_CHECK(clh_solver_get_stats(solver_, &st));
    return st.converged != 0;
}

void launch_apply_constraints_kernel(float* d_pos, float* d_vel, const int* d_indices,
                                     const float* d_target, int num_pinned)
{
    if (num_pinned == 0) return;
    CLH_CHECK(clh_apply_constraints(d_pos, d_vel, d_indices, d_target, num_pinned));
}
