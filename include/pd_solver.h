#pragma once
// PD solver: drop-in for the reference's PDSolver (include/pd_solver.h:1-77).
//
// The fields of PDSolverConfig up to `damping` are the reference's, in its
// order and with its defaults; the fields after it are appended with
// defaults that keep the reference's behaviour (mode = reference mode).

struct PDSolverConfig {
    int   max_iterations     = 50;    // local-global iterations per frame
    float tolerance          = 1e-4f; // PD mode: relative FP64 residual to stop at (reference mode ignores it)
    bool  use_chebyshev      = true;  // PD mode: Chebyshev blending
    bool  use_cpu_stretch_reference = false; // ignored, as in the reference's device build
    float rho                = 0.9f;  // spectral radius estimate for Chebyshev
    float gravity            = -9.8f; // gravitational acceleration (m/s^2), along y
    float dt                 = 0.01f; // time step
    float stretch_stiffness  = 1.0f;  // unused by the solver (per-triangle k lives in SimConstraints)
    float bend_stiffness     = 0.0f;  // unused by the solver (per-hinge k lives in SimConstraints)
    float damping            = 0.0f;  // velocity damping (0 = none, 1 = full)

    // ---- appended ----
    int   mode               = 0;     // 0 = reference mode, 1 = PD mode (true Jacobi + Chebyshev)
    float external_accel[3]  = {0.0f, 0.0f, 0.0f}; // uniform extra acceleration (wind)
    int   check_every        = 10;    // PD mode: residual test period
    int   cheb_delay         = 0;     // PD mode: iterations before Chebyshev starts
    int   kernel_variant     = 0;     // 0 auto, 1 unfused, 2 fused
    bool  use_graph          = true;  // capture fixed-length frames in a hipGraph
    float relax              = 0.0f;  // PD mode: Jacobi step scale gamma (0 = 1; -1 = from the
                                      // Gershgorin bound of D^-1 K, CLH_RELAX_AUTO)
    int   restart_rule       = 0;     // PD mode with a tolerance: CLH_RESTART_* (0 = on any growth)
    float restart_factor     = 10.0f; // CLH_RESTART_SEQ_MIN: growth factor over the sequence's best
};

struct ClothMesh;
struct SimConstraints;
struct Constraints;
struct clh_solver;

// Usage:
//   PDSolver solver(config, mesh, sim_cons);
//   // per frame:
//   solver.step(mesh, sim_cons, pin_cons);
class PDSolver {
public:
    PDSolver(const PDSolverConfig& config,
             const ClothMesh& mesh,
             const SimConstraints& sim_cons);
    ~PDSolver();

    PDSolver(const PDSolver&)            = delete;
    PDSolver& operator=(const PDSolver&) = delete;

    // One frame: predict -> local/global iterations -> velocity update.
    // Returns with the results visible to a following blocking copy.
    void step(ClothMesh& mesh,
              const SimConstraints& sim_cons,
              const Constraints& pin_cons);

    // Reset per-frame state; call when constraints change at runtime.
    void reset();

    // ---- appended ----
    // Iterations used and FP64 residuals of the last frame (PD mode). A frame
    // run to a tolerance that reaches max_iterations first is still
    // completed (velocity update, pins) from its last iterate: callers that
    // need the tolerance check last_converged() after each frame.
    int    last_iterations() const;
    double last_relative_residual() const;
    bool   last_converged() const;
    clh_solver* native_handle() const { return solver_; }

private:
    PDSolverConfig config_;
    int num_verts_;
    int num_tris_;
    int num_bend_cons_;
    clh_solver* solver_ = nullptr;
};

// Kernel wrapper used by Constraints::apply_gpu.
void launch_apply_constraints_kernel(
    float* d_pos, float* d_vel,
    const int* d_indices, const float* d_target,
    int num_pinned);
