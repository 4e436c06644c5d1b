#pragma once
/*
 * pd_oracle.h — CPU restatement of the reference's projective-dynamics
 * frame, used ONLY as a checker (tests, the smoke check, and the CPU
 * baseline leg of the benchmark). It is test infrastructure, never linked
 * into or called by the solver libraries.
 *
 * Modes:
 *   OR_REF_CUDA  PDSolver::step of src/pd_solver.cu:496-621 (transposed 2-D
 *                rotation of pd_solver.cu:184-194, x = rhs/diag with no
 *                off-diagonal term, no Chebyshev blend).
 *   OR_PD        projective dynamics with a true Jacobi sweep, Chebyshev
 *                blending (the formula of the reference's DESIGN.md:50-55)
 *                and the relative residual ||b - K x|| / ||b||.
 *   OR_EXACT     CpuStretchReferenceSolver::step of
 *                src/stretch_reference.cpp:140-280: exact global solves of
 *                K x = rhs (dense Cholesky of the scalar N x N matrix; K is
 *                that matrix Kronecker I3), polar rotation from a 2x2 SVD
 *                with the determinant fix. Flags select the reference's
 *                semantics (stretch only, pinned rows and columns zeroed
 *                without moving K_fp x_p to the right-hand side) or the
 *                corrected ones (bending included, pins eliminated).
 *
 * Every routine is templated on the working precision; the _f32 entry
 * points compute in float (the reference's precision), the _f64 ones in
 * double from the same float inputs.
 */
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { OR_REF_CUDA = 0, OR_PD = 1, OR_EXACT = 2 };

typedef struct or_mesh {
    int N, T, E, P;
    const float* mass;         /* [N]   */
    const int*   tris;         /* [T*3] */
    const float* Dm_inv;       /* [T*4] column-major */
    const float* rest_area;    /* [T]   */
    const float* stretch_k;    /* [T] or NULL (stretch off) */
    const float* diag;         /* [N]   */
    const int*   quads;        /* [E*4] or NULL (bending off) */
    const float* bend_w;       /* [E*4] */
    const float* bend_n;       /* [E]   */
    const float* bend_k;       /* [E]   */
    const int*   pin_idx;      /* [P]   */
    const float* pin_target;   /* [P*3] */
} or_mesh;

typedef struct or_params {
    int   mode;            /* OR_*                                               */
    int   iterations;      /* maximum iterations of the frame                    */
    int   use_chebyshev;   /* OR_PD                                              */
    float rho;             /* OR_PD                                              */
    int   cheb_delay;      /* OR_PD: iterations before blending starts (>= 1)    */
    float gravity;
    float dt;
    float damping;
    float ext_accel[3];
    float tolerance;       /* OR_PD: stop when ||r_k|| <= tol ||b_k|| (0 = off)  */
    int   check_every;     /* OR_PD: residual test period                        */
    int   exact_bend;      /* OR_EXACT: include bending (reference: 0)            */
    int   exact_pins;      /* OR_EXACT: eliminate pins correctly (reference: 0)   */
    int   threads;         /* OpenMP threads for OR_PD / OR_REF_CUDA (<= 0: 1)    */
    float relax;           /* OR_PD: under-relaxation gamma of the Jacobi step    */
    int   restart_rule;    /* OR_PD with a tolerance: OR_RESTART_*                 */
    float restart_factor;  /* OR_RESTART_SEQ_MIN: growth factor over the minimum   */
    /* OR_PD with a tolerance restarts the Chebyshev sequence at a test when the
     * rule says so: OR_RESTART_PREV, the residual exceeds the one at the
     * previous test; OR_RESTART_SEQ_MIN, it exceeds restart_factor times the
     * smallest residual since the sequence (re)started; OR_RESTART_NONE, never.
     * Every third restart doubles 1 - rho, down to rho = 0. */
} or_params;

enum { OR_RESTART_PREV = 0, OR_RESTART_SEQ_MIN = 1, OR_RESTART_NONE = 2 };

void or_params_default(or_params* p);

/* One frame. pos/vel [N*3] in/out. init: optional first iterate [N*3]
 * (NULL = pos; in OR_PD pinned vertices are set to their targets). Optional
 * outputs (NULL to skip): trace [iterations*N*3] every iterate; resid
 * [2*iterations] FP64 ||r_k||^2 and ||b_k||^2 of the iterate each iteration
 * starts from (OR_PD); *iters_out the number of iterations run. Returns 0 on
 * success. */
int or_step_f32(const or_mesh* m, const or_params* p, float* pos, float* vel,
                const float* init, float* trace, double* resid, int* iters_out);
int or_step_f64(const or_mesh* m, const or_params* p, double* pos, double* vel,
                const double* init, double* trace, double* resid, int* iters_out);

/* Element kernels, exposed for unit checks. F and R are 3x2 as [col0 xyz,
 * col1 xyz]. */
void or_ref_cuda_rotation_f32(const float* F, float* R);
void or_polar_rotation_f64(const double* F, double* R);
void or_polar_rotation_f32(const float* F, float* R);

/* Rest-state quantities recomputed in FP64 from rest positions, as
 * ClothMesh::precompute_rest_state and SimConstraints::build_bend /
 * precompute_jacobi_diag define them (src/cloth_mesh.cpp:138-211,
 * src/sim_constraints.cpp:48-192). Outputs in double:
 * Dm_inv [T*4], area [T], mass [N]; bend_w [E*4], bend_n [E] for given
 * quads; diag [N]. */
void or_rest_state_f64(int N, int T, const float* rest_pos, const int* tris, float density,
                       double* Dm_inv, double* area, double* mass);
void or_bend_weights_f64(int E, const float* rest_pos, const int* quads, double* w,
                         double* n_rest);
void or_jacobi_diag_f64(const or_mesh* m, float dt, double* diag);

/* Residual of a state: r = b(x) - K x with b from the projections at x,
 * pinned vertices excluded; returns ||r||^2 and ||b||^2 in FP64 and, if
 * r_out is non-NULL, r [N*3]. y is the inertial prediction. */
void or_residual_f64(const or_mesh* m, float dt, const double* x, const double* y,
                     double* r_out, double* r2, double* b2);

/* Residual of base + disp in the form the PD iteration evaluates it with
 * FP32 storage (FP32 base, dy = y - base and element outputs; double
 * displacement and element arithmetic); r_out [N*3]. */
void or_residual_disp_f32(const or_mesh* m, float dt, const float* base, const double* disp,
                          const float* dy, double* r_out, double* r2, double* b2);

#ifdef __cplusplus
}
#endif
