#pragma once
/*
 * cloth_scene.h — C API of libcloth_core.so for building the seeded cloth
 * workloads through the C++ cloth API (generate_square_cloth,
 * precompute_rest_state, MeshTopology::build, build_stretch, build_bend,
 * Constraints, upload_to_gpu, precompute_jacobi_diag, PDSolver) and for
 * driving PDSolver from languages without C++ bindings. It has no
 * counterpart in the reference, whose drivers are C++ programs
 * (src/tools/sim_cloth.cpp:169-252, src/tests/test_pd_constraints.cpp).
 */
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { CC_PIN_NONE = 0, CC_PIN_TOP_ROW = 1, CC_PIN_CORNERS = 2, CC_PIN_RANDOM_BOUNDARY = 3 };
enum { CC_MESH_IRREGULAR = 4 };  /* mesh type beyond the reference's 0..3 */

typedef struct cc_scene_spec {
    int      nrows, ncols;
    float    size;          /* grid spacing                                       */
    int      type;          /* 0..3 as generate_square_cloth, 4 = irregular        */
    float    rest_jitter;   /* type 4: in-plane rest jitter, fraction of size       */
    uint64_t seed;
    float    density;       /* areal density, kg/m^2                               */
    float    stretch_k;     /* stretch stiffness (lower bound when a range)         */
    float    stretch_k_max; /* > stretch_k: per-triangle log-uniform in the range   */
    float    bend_k;        /* <= 0: no bend constraints                           */
    int      pin_mode;      /* CC_PIN_*                                            */
    int      pin_count;     /* CC_PIN_RANDOM_BOUNDARY: number of distinct pins      */
    float    pos_jitter;    /* initial positions: U(-a, a) added on y, unpinned    */
    float    dt;            /* time step for the Jacobi diagonal                   */
} cc_scene_spec;

typedef struct cc_scene cc_scene;

/* Fills *spec with a 32x32 type-0 unit cloth, config-1 material. */
void cc_scene_spec_default(cc_scene_spec* spec);
cc_scene* cc_scene_build(const cc_scene_spec* spec);
/* Same, with the mesh read by ClothMesh::load_obj (NULL on failure); the
 * grid fields of spec are used only for pin selection (ncols). */
cc_scene* cc_scene_build_obj(const char* path, const cc_scene_spec* spec);
void cc_scene_destroy(cc_scene* s);
/* Batched workload: each spec built as cc_scene_build builds it, then the
 * cloths concatenated into one scene (vertex and triangle ids shifted; the
 * specs must share dt). The per-cloth offsets give each cloth's range. */
cc_scene* cc_batch_build(const cc_scene_spec* specs, int count);
int  cc_scene_num_cloths(const cc_scene* s);
/* Offsets [C+1] of each cloth's vertices, triangles, hinges and pins in the
 * concatenated arrays; a null pointer skips that table. */
int  cc_scene_cloth_offsets(const cc_scene* s, int* vert, int* tri, int* bend, int* pin);
void cc_scene_counts(const cc_scene* s, int* N, int* T, int* E, int* P);
/* Host export of a CPU array by name into dst (bytes must match exactly);
 * returns 0 on success. Names: rest_pos pos0 tris Dm_inv rest_area mass
 * stretch_k jacobi_diag bend_quads bend_w bend_n bend_k pin_idx pin_target
 * edges (v0, v1, tri_a, tri_b per edge) edge_len vt_ptr vt_idx tri_to_tris
 * deltaUV dF_dx. num_edges gives the edge count. */
int  cc_scene_get(const cc_scene* s, const char* name, void* dst, size_t bytes);
int  cc_scene_num_edges(const cc_scene* s);
/* Uploads mesh, constraints and pins, computes the Jacobi diagonal for the
 * spec's dt, and sets the device positions to pos0. */
int  cc_scene_upload(cc_scene* s);
/* Device pointer by name: pos vel mass tris Dm_inv rest_area stretch_k
 * jacobi_diag bend_quads bend_w bend_n bend_k pin_idx pin_target. */
void* cc_scene_device(const cc_scene* s, const char* name);
/* Blocking copies of the device state from / to host arrays [N*3];
 * a null pointer skips that array. */
int  cc_scene_set_state(cc_scene* s, const float* pos, const float* vel);
int  cc_scene_get_state(const cc_scene* s, float* pos, float* vel);

/* PDSolverConfig as a C struct (include/pd_solver.h). */
typedef struct cc_pd_config {
    int   max_iterations;
    float tolerance;
    int   use_chebyshev;
    float rho;
    float gravity;
    float dt;
    float damping;
    int   mode;
    float external_accel[3];
    int   check_every;
    int   cheb_delay;
    int   kernel_variant;
    int   use_graph;
} cc_pd_config;

void  cc_pd_config_default(cc_pd_config* c);
void* cc_pdsolver_create(cc_scene* s, const cc_pd_config* c);
void  cc_pdsolver_step(void* solver, cc_scene* s);
void  cc_pdsolver_reset(void* solver);
void* cc_pdsolver_native(void* solver);  /* clh_solver* */
void  cc_pdsolver_destroy(void* solver);
/* Energies of a state given as host arrays pos, vel [N*3] (no device):
 * out[0..3] = kinetic, gravitational potential (gravity along y), stretch
 * sum 1/2 k A |F - R|^2, bend sum 1/2 k |e - p|^2 (host/cloth_energy.h). */
void  cc_scene_energy(const cc_scene* s, float gravity, const float* pos, const float* vel,
                      double* out);
/* Constraints::apply_gpu on the scene's device state. */
void  cc_scene_apply_pins(cc_scene* s);

/* The host state of the CPU reference backend, mesh.pos_cpu / vel_cpu
 * (host arrays [N*3]; a null pointer skips that array). */
int   cc_scene_get_cpu_state(const cc_scene* s, float* pos, float* vel);
int   cc_scene_set_cpu_state(cc_scene* s, const float* pos, const float* vel);
/* CpuStretchReferenceSolver (include/stretch_reference.h) on the scene's mesh,
 * stretch constraints and pins, with the spec's dt; refactor != 0 re-factors the global matrix
 * every iteration, as the reference does. step runs `steps` steps of the host
 * state. */
void* cc_cpuref_create(cc_scene* s, float gravity, float damping, int refactor);
void  cc_cpuref_step(void* solver, cc_scene* s, int steps);
int   cc_cpuref_factorizations(void* solver);
void  cc_cpuref_destroy(void* solver);

#ifdef __cplusplus
}
#endif
