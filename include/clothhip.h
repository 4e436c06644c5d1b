#pragma once
/*
 * clothhip.h — C ABI of libclothhip.so, the MI355X (gfx950) projective-
 * dynamics cloth solver.
 *
 * Every entry point returns an int status (CLH_OK == 0); on failure
 * clh_last_error() returns a message for the calling thread. Only plain
 * pointers, sizes and POD structs cross this boundary. Device pointers are
 * pointers returned by clh_alloc (or any other allocator of device memory).
 *
 * Each entry point names the reference interface it replaces
 * (Chongyao/cloth-cuda, file:line).
 */
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---------------------------------------------------- */
enum {
    CLH_OK              = 0,
    CLH_ERR_INVALID     = 1,  /* bad argument (null pointer, size, count) */
    CLH_ERR_DEVICE      = 2,  /* the HIP runtime reported an error       */
    CLH_ERR_ALLOC       = 3,  /* device allocation failed                 */
    CLH_ERR_UNSUPPORTED = 4,  /* request outside what the library handles */
    CLH_ERR_NUMERIC     = 5   /* non-finite state detected                */
};

/* ---- copy directions (same numbering as the runtime's memcpy kinds) -- */
enum { CLH_H2H = 0, CLH_H2D = 1, CLH_D2H = 2, CLH_D2D = 3 };

/* ---- solver modes ------------------------------------------------------
 * CLH_MODE_REF_CUDA  reproduces PDSolver::step of the CUDA reference
 *                    (src/pd_solver.cu:496-621): rhs = M y + h^2 sum w A^T p,
 *                    x = rhs / diag, the 2-D rotation of the stretch
 *                    projection as that kernel forms it (pd_solver.cu:184-194),
 *                    no Chebyshev blend (pd_solver.cu:623-634).
 * CLH_MODE_PD        projective dynamics with a true Jacobi sweep
 *                    x <- x + D^-1 (b(x) - K x), the polar rotation of F,
 *                    Chebyshev blending and an FP64 relative residual
 *                    ||b - K x|| / ||b|| used for early termination.
 */
enum { CLH_MODE_REF_CUDA = 0, CLH_MODE_PD = 1 };

/* ---- kernel variants --------------------------------------------------- */
enum {
    CLH_KERNEL_AUTO    = 0,  /* fused when available, else unfused          */
    CLH_KERNEL_UNFUSED = 1,  /* element kernels + per-vertex gather          */
    CLH_KERNEL_FUSED   = 2   /* one LDS-tiled kernel per iteration           */
};

/* ---- runtime ------------------------------------------------------------ */

/* Message describing the last failure on this thread ("" if none). */
const char* clh_last_error(void);

/* Replaces cuda_device_available() (src/utils/cuda_helper.h:19-22). */
int clh_device_count(int* count);

typedef struct clh_device_info {
    char   name[256];
    char   arch[64];          /* gcnArchName, e.g. "gfx950:sramecc+:xnack-" */
    int    compute_units;
    int    clock_khz;
    int    mem_clock_khz;
    int    mem_bus_width;
    size_t total_mem;
    size_t l2_cache_bytes;
} clh_device_info;

/* Replaces print_device_info() (src/utils/cuda_helper.h:24-37). */
int clh_get_device_info(int device, clh_device_info* out);

/* Device memory used by the upload / free bodies of ClothMesh, SimConstraints
 * and Constraints (src/cloth_mesh.cpp:18-83, src/sim_constraints.cpp:187-247,
 * src/constraints.cpp:77-101), which call cudaMalloc / cudaMemcpy /
 * cudaMemset / cudaFree directly. clh_memcpy and clh_memset are blocking with
 * respect to the host, like the runtime calls they replace. */
int clh_alloc(void** ptr, size_t bytes);
int clh_free(void* ptr);
int clh_memcpy(void* dst, const void* src, size_t bytes, int kind);
int clh_memset(void* ptr, int value, size_t bytes);
int clh_synchronize(void);

/* pos[idx[i]] = target[i]; vel[idx[i]] = 0 for i < num_pinned.
 * Replaces launch_apply_constraints_kernel (include/pd_solver.h:74-77,
 * src/pd_solver.cu:414-439), the kernel behind Constraints::apply_gpu
 * (src/constraints.cpp:112-119). Runs on the null stream. */
int clh_apply_constraints(float* d_pos, float* d_vel, const int* d_indices,
                          const float* d_target, int num_pinned);

/* ---- solver ------------------------------------------------------------- */

typedef struct clh_solver clh_solver;

/* Mirrors PDSolverConfig (include/pd_solver.h:4-15) plus appended fields. */
typedef struct clh_config {
    int   mode;             /* CLH_MODE_*                                       */
    int   max_iterations;   /* local-global iterations per frame                */
    int   use_chebyshev;    /* CLH_MODE_PD: blend with Chebyshev weights        */
    float rho;              /* spectral radius estimate for the weights         */
    float gravity;          /* acceleration along component 1 (the y axis)      */
    float dt;               /* time step                                        */
    float damping;          /* velocity damping in [0, 1]                       */
    float tolerance;        /* CLH_MODE_PD: stop at ||r||/||b|| <= tolerance; 0 = fixed iterations */
    int   check_every;      /* CLH_MODE_PD: residual test period (iterations)   */
    int   cheb_delay;       /* CLH_MODE_PD: iterations before blending starts   */
    float ext_accel[3];     /* uniform external acceleration (wind), m/s^2      */
    int   kernel_variant;   /* CLH_KERNEL_*                                     */
    int   use_graph;        /* capture fixed-length frames in a hipGraph        */
    int   tile_verts;       /* fused variant: vertices per tile (0 = default)   */
    int   block_threads;    /* fused variant: workgroup size, 256, 512 or 1024 (0 = default;
                               two-iteration launches need a workgroup of at least a
                               tile's own and ring-1 vertices) */
    int   persistent;       /* CLH_PERSIST_ON: frames to a tolerance in one cooperative
                               launch per chunk when every tile fits resident (256- or
                               1024-thread workgroups): each tile waits only for its
                               neighbour tiles between iterations, a grid barrier joins
                               them at the checks; same bits as launches; off by
                               default */
    int   iters_per_launch; /* fused variant, CLH_MODE_PD: 2 runs two iterations per
                               launch (each tile also updates its ring-1 halo in the
                               first; same bits as one per launch) and lets the
                               autotuner choose between 1 and 2; 1 or 0 (default) one
                               per launch: on MI355X two are slower per iteration */
    int   element_regions;  /* fused variant, CLH_MODE_PD: LDS regions of a tile's element
                               results: 2 = separate triangle and hinge regions
                               (evaluated in one phase, two barriers fewer) where they
                               fit, 1 = one shared region, 0 = the default; the
                               autotuner times both. Same bits either way. */
    int   stall_checks;     /* CLH_MODE_PD with tolerance > 0: a frame (in a batch: a
                               cloth) whose relative residual has not dropped by 10 %
                               (below 0.9 of its value at the last such drop) for this
                               many checks in a row stops, not converged, instead of
                               running to max_iterations; 0 = off */
    int   restart_rule;     /* CLH_MODE_PD with tolerance > 0: when a residual check
                               restarts the Chebyshev sequence (CLH_RESTART_*) */
    float restart_factor;   /* CLH_RESTART_SEQ_MIN: restart when the residual exceeds this
                               factor (> 1) times the smallest residual of the current
                               sequence */
    float relax;            /* CLH_MODE_PD: scale gamma of the Jacobi step,
                               x + gamma D^-1 (b - K x), in (0, 2); 0 means 1 (plain
                               Jacobi). Below 1 it damps the Jacobi modes with
                               eigenvalues below -1 that stiff bending produces.
                               CLH_RELAX_AUTO: gamma = min(1, (1 + rho) / mu), mu the
                               Gershgorin bound of D^-1 K computed on the device from
                               the constants, so every eigenvalue of the damped Jacobi
                               iteration lies in [-rho, 1). In a batch each cloth takes
                               gamma from its own bound, as it would alone */
} clh_config;

#define CLH_RELAX_AUTO (-1.0f)

/* Chebyshev restart rules (clh_config.restart_rule). Every third restart of a
 * frame (of a cloth, in a batch) doubles 1 - rho, down to rho = 0 (plain
 * Jacobi), under each rule that restarts. */
enum {
    CLH_RESTART_PREV    = 0,  /* the residual grew since the previous check (default) */
    CLH_RESTART_SEQ_MIN = 1,  /* the residual exceeds restart_factor times the smallest
                                 residual since the sequence (re)started */
    CLH_RESTART_NONE    = 2   /* no restarts */
};

/* Per-cloth state of a batch frame run to a tolerance (clh_batch_cloth_status). */
enum {
    CLH_CLOTH_RUNNING   = 0,  /* still iterating when the frame reached max_iterations */
    CLH_CLOTH_CONVERGED = 1,
    CLH_CLOTH_STALLED   = 2,  /* stopped by the stall rule (clh_config.stall_checks) */
    CLH_CLOTH_NONFINITE = 3   /* diverged: a non-finite residual, or a relative residual
                                 above 1000 times its best of the frame; the cloth is
                                 returned to x_n (zero velocity) for this frame and the
                                 other cloths' results stand */
};

enum { CLH_PERSIST_OFF = 0, CLH_PERSIST_ON = 1 };

/* Fills *cfg with the defaults of PDSolverConfig (include/pd_solver.h:5-14)
 * and CLH_MODE_REF_CUDA. */
int clh_config_default(clh_config* cfg);

/* Topology fixed for the lifetime of a solver. The library copies what it
 * needs; the pointers need not outlive the call. */
typedef struct clh_mesh_desc {
    int        num_verts;
    int        num_tris;
    int        num_bend;
    const int* d_tris;        /* [T*3] device */
    const int* d_bend_quads;  /* [E*4] device, may be NULL when num_bend == 0 */
} clh_mesh_desc;

/* Replaces PDSolver::PDSolver (src/pd_solver.cu:445-460), which caches N, T
 * and E and allocates the solver scratch (pd_solver.cu:470-482). */
int clh_solver_create(const clh_config* cfg, const clh_mesh_desc* mesh,
                      clh_solver** out);
/* Replaces PDSolver::~PDSolver (src/pd_solver.cu:462-468). */
int clh_solver_destroy(clh_solver* s);
/* Replaces PDSolver::reset (src/pd_solver.cu:491-494). Also drops cached
 * per-element constants so that values changed in place are re-read. */
int clh_solver_reset(clh_solver* s);

/* Per-frame arguments; the same arrays PDSolver::step reads from ClothMesh,
 * SimConstraints and Constraints (src/pd_solver.cu:496-621). A NULL
 * d_stretch_k disables stretch and a NULL d_bend_quads disables bending,
 * as in the reference (pd_solver.cu:534, 542). */
typedef struct clh_step_args {
    float*       d_pos;          /* [N*3] in/out                                */
    float*       d_vel;          /* [N*3] in/out                                */
    const float* d_mass;         /* [N]                                         */
    const int*   d_tris;         /* [T*3]                                       */
    const float* d_Dm_inv;       /* [T*4] column-major [m00, m10, m01, m11]     */
    const float* d_rest_area;    /* [T]                                         */
    const float* d_stretch_k;    /* [T]                                         */
    const float* d_jacobi_diag;  /* [N]                                         */
    const int*   d_bend_quads;   /* [E*4]                                       */
    const float* d_bend_w;       /* [E*4]                                       */
    const float* d_bend_n;       /* [E]                                         */
    const float* d_bend_k;       /* [E]                                         */
    const int*   d_pin_idx;      /* [P]                                         */
    const float* d_pin_target;   /* [P*3] compact targets                       */
    int          num_pinned;
    const float* d_init_guess;   /* optional [N*3] first iterate (warm start); NULL = x_n */
    float*       d_trace;        /* optional [max_iterations*N*3] every iterate */
    const float* d_cloth_accel;  /* optional [C*3] per-cloth acceleration added to
                                    ext_accel (C = 1 for a solver of one cloth) */
} clh_step_args;

/* Replaces PDSolver::step (src/pd_solver.cu:496-621). Work is enqueued on a
 * non-blocking stream the solver creates, ordered after earlier null-stream
 * work, and later null-stream work waits for it (events both ways), so a
 * following blocking copy observes the result, as in the reference. After
 * clh_solver_set_stream the work goes to the caller's stream, unfenced. */
int clh_solver_step(clh_solver* s, const clh_step_args* a);

typedef struct clh_step_stats {
    int    iterations;   /* iterations run in the last frame                */
    int    converged;    /* CLH_MODE_PD with tolerance > 0: reached it       */
    double residual;     /* last FP64 ||b - K x|| (CLH_MODE_PD)              */
    double residual0;    /* first FP64 ||b - K x|| of the frame              */
    double rhs_norm;     /* FP64 ||b|| at the last residual evaluation        */
    int    frames;       /* frames stepped since creation or reset           */
    int    restarts;     /* Chebyshev restarts of the last frame (one solve; 0 in a batch) */
    float  rho;          /* rho in use when the last frame ended (one solve) */
    float  relax;        /* Jacobi step scale gamma of the last frame (a batch under
                            CLH_RELAX_AUTO: the smallest over its cloths) */
    double jacobi_bound; /* CLH_RELAX_AUTO: the Gershgorin bound of D^-1 K it used, the
                            largest over a batch's cloths (else 0) */
} clh_step_stats;

/* FP64 relative residual ||b - K x_0|| / ||b|| of the frame's first iterate
 * (x_0 = a->d_init_guess, or x_n when it is NULL; pinned vertices at their
 * targets) without stepping: the mesh state is left as it is. CLH_MODE_PD.
 * Lets a caller compare candidate starts by the measure the frame stops on.
 * No counterpart in the reference. */
int clh_solver_start_residual(clh_solver* s, const clh_step_args* a, double* rel);

/* Synchronizes the solver's stream and returns statistics of the last frame. */
int clh_solver_get_stats(clh_solver* s, clh_step_stats* out);
/* Copies up to max_count FP64 relative residuals ||r_k||/||b_k|| of the last
 * frame (one per residual evaluation) into host memory; *count gets the
 * number available. */
int clh_solver_residual_history(clh_solver* s, double* out, int max_count,
                                int* count);
/* Sets the stream (a hipStream_t) that step and run enqueue on; NULL returns
 * to the solver's own stream with the null-stream fences of clh_solver_step. */
int clh_solver_set_stream(clh_solver* s, void* stream);

/* ---- measurement ---------------------------------------------------------- */

/* Steps num_frames frames back to back on the solver's stream and returns the
 * elapsed device time between events recorded before and after them. */
int clh_solver_run_frames(clh_solver* s, const clh_step_args* a,
                          int num_frames, float* elapsed_ms);

/* Launches the dominant iteration kernel num_launches times back to back on
 * the solver's stream, bracketed by events, and returns the average time per
 * launch. The launches are consecutive steady-phase iterations (x_k and
 * x_{k-1} read, each launch reading what the one before wrote, the Chebyshev
 * blend on in PD mode at a fixed weight; two iterations per launch when
 * clh_plan_info.iters_per_launch is 2) from iterates computed from the mesh
 * state; only solver scratch is written, the mesh state is left as it was. */
int clh_solver_time_iteration_kernel(clh_solver* s, const clh_step_args* a,
                                     int num_launches, float* avg_ms);

/* Bytes the plan of the fused kernel reads per iteration beyond the
 * algorithmic count (halo indices, adjacency, duplicated halo elements), and
 * the tile statistics, for the roofline report. */
typedef struct clh_plan_info {
    int    variant;          /* CLH_KERNEL_* actually used */
    int    num_tiles;
    int    tile_verts;
    double halo_tri_factor;  /* triangles processed / T */
    double halo_bend_factor; /* hinges processed / E    */
    double bytes_per_iter;   /* bytes the kernel streams per iteration (model) */
    double lds_bytes;        /* LDS per workgroup of the residual-check variant */
    int    block_threads;    /* workgroup size of the fused kernel */
    int    num_cloths;       /* cloths in the plan (1 unless batched) */
    int    padded_verts;     /* tile-order length: each cloth padded to whole tiles */
    int    persistent;       /* frames to a tolerance use the persistent kernel */
    int    tri_max;          /* largest triangle count of a tile (halo included) */
    int    hinge_max;        /* largest hinge count of a tile */
    int    halo_max;         /* largest halo of a tile */
    int    pair;             /* the plan has the two-iteration extension (PD mode) */
    int    iters_per_launch; /* iterations per launch of the fused kernel in use: 1 or 2 */
    int    pair_halo_max;    /* largest ring-1 + ring-2 halo of a tile */
    int    pair_tri_max;     /* largest triangle count of a two-iteration tile */
    int    pair_hinge_max;   /* largest hinge count of a two-iteration tile */
    double pair_bytes_per_launch; /* bytes a two-iteration launch streams (model) */
    int    pair_compute_max; /* largest own + ring-1 count: the workgroup size it needs */
    double pair_lds_bytes;   /* LDS per workgroup of the two-iteration kernel (check variant) */
    int    element_regions;  /* 2 when iteration launches use separate triangle and hinge
                                regions, else 1 */
    double split_lds_bytes;  /* LDS per workgroup with separate regions (check variant) */
    int    lds_limit;        /* LDS a workgroup of the fused kernels may use on this device */
    int    gather_lanes;     /* lanes per own vertex in the gathers of a one-iteration launch
                                (1, 2 or 4; the sums have the same bits for each) */
} clh_plan_info;
int clh_solver_plan_info(clh_solver* s, clh_plan_info* out);

/* Builds the fused kernel's tile plan for a mesh from host arrays (no device
 * needed) and checks every index in it: the tile order is a permutation of
 * the vertices plus padding, each cloth fills whole tiles of its own, local
 * ids lie inside each tile's LDS table, halo ids outside the tile and inside
 * its cloth, each local slot names the element's vertex, and each vertex's
 * gather list equals its incident element slots in ascending order. Returns
 * CLH_OK and the plan statistics, or an error naming the first violated
 * property. num_cloths = 0 means one cloth; otherwise cloth_vert_offsets
 * [num_cloths + 1] delimits the cloths as in clh_batch_create. */
int clh_fused_plan_check(int num_verts, int num_tris, int num_bend, const int* tris,
                         const int* bend_quads, const float* pos, int tile_verts, int mode,
                         int num_cloths, const int* cloth_vert_offsets, clh_plan_info* out);

/* ---- batched cloths -------------------------------------------------------
 * Many independent cloths of uneven size and material in one solver: the
 * arrays of clh_step_args are the concatenation of the cloths' arrays, with
 * vertex indices global (cloth c owns vertices [cloth_vert_offsets[c],
 * cloth_vert_offsets[c+1]) and every triangle, hinge and pin of it indexes
 * only those). One launch of the fused iteration kernel per iteration covers
 * every cloth; tiles never straddle two cloths. Per-cloth stiffness comes
 * from the per-element arrays, per-cloth wind from d_cloth_accel. A frame run
 * to a tolerance stops when every cloth's relative residual meets it.
 * No counterpart in the reference, whose PDSolver holds one mesh
 * (src/pd_solver.cu:445-460). */
int clh_batch_create(const clh_config* cfg, const clh_mesh_desc* mesh, int num_cloths,
                     const int* cloth_vert_offsets, clh_solver** out);

/* FP64 relative residual ||b_c - K_c x_c|| / ||b_c|| of each cloth at the last
 * residual evaluation of the last frame (CLH_MODE_PD); *count gets the number
 * of cloths. In a frame run to a tolerance, a cloth's value is the one at the
 * check where it stopped. */
int clh_batch_cloth_residuals(clh_solver* s, double* out, int max_count, int* count);

/* State (CLH_CLOTH_*) and iterations of each cloth in the last frame of a
 * batch run to a tolerance: in such frames every cloth has its own Chebyshev
 * sequence, restarts and convergence test, and is frozen once it stops, so a
 * cloth of a batch follows the frame it would follow alone, and the frame
 * ends when no cloth runs. A diverged cloth does not fail the frame (a single
 * cloth's does, with CLH_ERR_NUMERIC). Either output may be NULL; *count gets
 * the number of cloths. */
int clh_batch_cloth_status(clh_solver* s, int* state, int* iterations, int max_count, int* count);

/* Autotuner of the fused kernel's launch shape: times every (tile_verts in
 * {32, 64, 128, 256}, block_threads in {256, 512, 1024}; in CLH_MODE_PD
 * element regions in {1, 2} where separate regions fit, and with
 * iters_per_launch = 2 also two iterations per launch at every workgroup
 * size that holds a tile's own and ring-1 vertices; at most 36 candidates)
 * candidate on this
 * solver's mesh (the lower of two timings of num_launches steady-phase
 * launches each, as timed by clh_solver_time_iteration_kernel, after about
 * 20 ms of warm-up launches per tile size) and keeps the one with the
 * shortest time per iteration. Results are cached per process under a key of the vertex,
 * triangle, hinge and cloth counts, the mode, whether two-iteration launches
 * may be used and the histogram of vertex valences, so another solver on an
 * equally shaped mesh reuses the choice without timing. */
#define CLH_TUNE_MAX_CANDIDATES 40

typedef struct clh_tune_report {
    int   num_candidates;
    int   tile_verts[CLH_TUNE_MAX_CANDIDATES];
    int   block_threads[CLH_TUNE_MAX_CANDIDATES];
    int   iters_per_launch[CLH_TUNE_MAX_CANDIDATES];
    int   element_regions[CLH_TUNE_MAX_CANDIDATES];
    float us_per_iteration[CLH_TUNE_MAX_CANDIDATES]; /* launch time / iterations; 0 when the
                                                         candidate did not fit */
    int   best;               /* index of the chosen candidate */
    int   from_cache;         /* 1 when the choice came from the cache */
    unsigned long long key;   /* cache key of the mesh shape */
} clh_tune_report;
int clh_solver_autotune(clh_solver* s, const clh_step_args* a, int num_launches,
                        clh_tune_report* out);

/* Streams bytes of device memory through a copy kernel num_launches times and
 * returns the average time per launch (measured copy peak). */
int clh_measure_copy(void* d_dst, const void* d_src, size_t bytes,
                     int num_launches, float* avg_ms);

/* ---- learned warm start ----------------------------------------------------
 * A per-vertex message-passing MLP that predicts each frame's converged
 * positions from the three previous frames, used as the first iterate of
 * clh_solver_step (clh_step_args.d_init_guess); the unchanged iterations then
 * run to the same residual tolerance, so the plain method stays the arbiter.
 * guess = 3 x_n - 3 x_{n-1} + x_{n-2} + S(h^2 g0 s_o MLP(features)), with
 * g0 = 9.81 m/s^2, s_o the output scale, 16 BF16 features per vertex
 * (velocity and acceleration of the last frames in units of h^2 g0, the
 * external acceleration over g0 when ext_feature is set, the pin flag, and
 * 1-ring means of velocity
 * and acceleration minus the vertex's own), two SiLU hidden layers of 64 and
 * a linear head of 3, on BF16 MFMA with FP32 accumulation, and S optional
 * passes of 1-ring averaging of the correction; trained on the
 * device by the library's own forward and backward kernels with Adam on FP32
 * master weights. Pinned vertices keep x_n (the solver sets their targets).
 * No counterpart in the reference, which starts each frame from x_n
 * (src/pd_solver.cu:528). All device work goes to the handle's stream (the
 * null stream unless clh_ws_set_stream was called). */
typedef struct clh_ws clh_ws;

typedef struct clh_ws_desc {
    int          num_verts;
    int          num_tris;
    const int*   d_tris;              /* [T*3] device: the triangles' edges are the message graph */
    int          num_cloths;          /* 0 or 1: one cloth; else cloth_vert_offsets is read */
    const int*   cloth_vert_offsets;  /* host [num_cloths + 1], as clh_batch_create */
    const int*   d_pin_idx;           /* [P] device */
    int          num_pinned;
    float        dt;                  /* frame time step */
    float        gravity;             /* along component 1, as clh_config.gravity */
    float        ext_accel[3];        /* uniform external acceleration, as clh_config */
    unsigned long long seed;          /* initial weights and minibatch sampling */
    int          ext_feature;         /* 1: the external acceleration / g0 is among the
                                         features (6..8); 0: those features are zero */
} clh_ws_desc;

int clh_ws_create(const clh_ws_desc* desc, clh_ws** out);
int clh_ws_destroy(clh_ws* w);
int clh_ws_set_stream(clh_ws* w, void* stream);
/* Number of FP32 parameters: W1 [64x16], b1 [64], W2 [64x64], b2 [64],
 * W3 [16x64] (rows 3..15 zero), b3 [16] (entries 3..15 zero), row-major. */
int clh_ws_param_count(void);
int clh_ws_get_params(clh_ws* w, float* out, int count, float* out_scale);
/* out_scale >= 0 multiplies the network's correction; 0 turns it off (the
 * prediction is the extrapolation), and training frames then need a positive
 * scale set first. */
int clh_ws_set_params(clh_ws* w, const float* in, int count, float out_scale);
/* d_guess [N*3] from the positions of the three previous frames (device
 * [N*3] each) and the optional per-cloth acceleration [C*3];
 * use_network = 0 gives the extrapolation alone. */
int clh_ws_predict(clh_ws* w, const float* d_xm2, const float* d_xm1, const float* d_x,
                   const float* d_cloth_accel, float* d_guess, int use_network);
/* Passes of 1-ring averaging applied to the network's correction before it
 * is added to the extrapolation (default 0 = none). */
int clh_ws_set_smoothing(clh_ws* w, int passes);
/* Filter on the whole predicted step: the guess becomes
 * x_n + gain * S^passes(x_hat - x_n), with x_hat the prediction above (after
 * the correction's own smoothing) and S one pass of 1-ring averaging (pinned
 * vertices keep x_n). A first iterate that is wrong in the cloth's stiff,
 * rough modes raises the residual the frame must reduce; smoothing keeps the
 * step's smooth part, and gain scales it (0 gives x_n, the plain start).
 * Defaults: 0 passes, gain 1 (the prediction unchanged). */
int clh_ws_set_step_filter(clh_ws* w, int passes, float gain);
/* The 16 BF16 features per vertex [N*16] (device) of a state. */
int clh_ws_features(clh_ws* w, const float* d_xm2, const float* d_xm1, const float* d_x,
                    const float* d_cloth_accel, unsigned short* d_features);
/* Training set on the device: each call appends one frame (N samples padded
 * to a multiple of 16 with weight-0 samples): features of (x_{n-2}, x_{n-1},
 * x_n) and the target, the change of acceleration to the converged x_{n+1}
 * over s_o. A sample's weight is its cloth's entry of d_cloth_weight [C]
 * (device, NULL = 1; 0 leaves out a cloth whose frame did not converge);
 * pinned vertices have weight 0. */
int clh_ws_dataset_reserve(clh_ws* w, long long samples);
int clh_ws_dataset_add(clh_ws* w, const float* d_xm2, const float* d_xm1, const float* d_x,
                       const float* d_xnext, const float* d_cloth_accel,
                       const float* d_cloth_weight);
/* Appends count samples from host arrays: features [count*16] BF16 bits and
 * targets [count*4] (3 components over s_o, weight), padded to a multiple of
 * 16 with weight-0 samples (a stored training set, or test vectors). */
int clh_ws_dataset_append(clh_ws* w, const unsigned short* features, const float* targets,
                          int count);
int clh_ws_dataset_size(clh_ws* w, long long* samples);
int clh_ws_dataset_clear(clh_ws* w);
/* Host copies of samples [first, first + count): features [count*16] BF16
 * bits, targets [count*4] (3 components over s_o, weight). */
int clh_ws_dataset_read(clh_ws* w, long long first, int count, unsigned short* features,
                        float* targets);
/* Sets s_o to the RMS of the weighted targets and rescales them. */
int clh_ws_fit_output_scale(clh_ws* w, float* out_scale);
/* Loss sum_s w_s |o_s - t_s|^2 / (3 count) over samples [first, first + count)
 * (first and count multiples of 16) and its gradient [param_count] (host). */
int clh_ws_loss_grad(clh_ws* w, long long first, int count, float* loss, float* grad);
/* steps Adam steps (beta 0.9 / 0.999, eps 1e-8, learning rate decayed from lr
 * to lr / 10 on a cosine) on minibatches of batch samples (a multiple of 16)
 * drawn as random 16-sample blocks; losses [steps] (host, optional). */
int clh_ws_train(clh_ws* w, int steps, int batch, float lr, float* losses);
/* One Adam step at learning rate lr with a given gradient [param_count] (host). */
int clh_ws_adam_step(clh_ws* w, const float* grad, float lr);

/* ---- fixed-length frames without allocation --------------------------------
 * No counterpart in the reference: an entry point for callers that cannot
 * allocate device memory while they step (code under hipGraph stream capture,
 * tight serving loops, harnesses that forbid hipMalloc). One fixed-length
 * frame of cfg->max_iterations iterations in either mode, computed by the
 * same kernels, in the same order, as clh_solver_step with kernel_variant
 * CLH_KERNEL_UNFUSED, and so with the same bits. Every buffer it needs
 * beyond the caller's inputs and outputs lies in one static device array of
 * CLH_NOALLOC_SCRATCH_BYTES that each device the library runs on holds until
 * the process ends; the call makes no allocation (no hipMalloc, hipFree or
 * allocating library routine) and enqueues only kernels.
 *   - Limit: about 112 N + 84 T + 40 E bytes of scratch (each region rounded
 *     to 256 bytes) must fit; a 256x256 cloth (65,536 vertices, 130,050
 *     triangles, 194,565 hinges) needs 26 MB. Larger meshes, tolerance > 0
 *     and d_trace are rejected with CLH_ERR_UNSUPPORTED.
 *   - Indices are not range-checked on the host (that would need a copy): a
 *     triangle, hinge or pin that names a vertex outside [0, num_verts) is
 *     inert instead (no force, no pin), and never addresses memory.
 *   - Calls on one device share that device's array and leave their contents
 *     in it: calls that overlap on one device race, so serialize them (one
 *     stream, or events). Each call clears what it needs cleared.
 *   - Work is enqueued on `stream` (a hipStream_t; NULL = the null stream,
 *     after which a blocking copy observes the result) and the call returns
 *     without waiting. No statistics are recorded. */
#define CLH_NOALLOC_SCRATCH_BYTES (32u << 20)
int clh_step_noalloc(const clh_config* cfg, int num_verts, int num_tris, int num_bend,
                     const clh_step_args* a, void* stream);

#ifdef __cplusplus
}  /* extern "C" */
#endif
