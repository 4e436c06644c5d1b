// sim_cloth: command-line driver of the PD solver with the reference tool's
// arguments, defaults, output lines and ASCII PLY frames
// (src/tools/sim_cloth.cpp:20-268). Appended options select this build's PD
// mode, a residual tolerance, an OBJ mesh and the energy diagnostics.
#include "check.h"
#include "cloth_energy.h"
#include "cloth_mesh.h"
#include "constraints.h"
#include "mesh_generator.h"
#include "mesh_topology.h"
#include "pd_solver.h"
#include "sim_constraints.h"
#include "stretch_reference.h"

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

namespace {

// ASCII PLY of the device positions, or of the host positions of the CPU
// reference backend (format of src/tools/sim_cloth.cpp:20-61).
void export_ply(const ClothMesh& mesh, const char* filename, bool use_cpu_ref)
{
    FILE* fp = std::fopen(filename, "w");
    if (!fp) {
        std::fprintf(stderr, "Failed to open %s for writing\n", filename);
        return;
    }
    std::vector<float> pos, vel;
    if (use_cpu_ref) host_state(mesh, pos, vel);
    else download_state(mesh, pos, vel);
    std::fprintf(fp, "ply\nformat ascii 1.0\n");
    std::fprintf(fp, "element vertex %d\n", mesh.num_verts);
    std::fprintf(fp, "property float x\nproperty float y\nproperty float z\n");
    std::fprintf(fp, "element face %d\n", mesh.num_tris);
    std::fprintf(fp, "property list uchar int vertex_indices\nend_header\n");
    for (int i = 0; i < mesh.num_verts; ++i)
        std::fprintf(fp, "%f %f %f\n", pos[(size_t)i * 3], pos[(size_t)i * 3 + 1],
                     pos[(size_t)i * 3 + 2]);
    for (const auto& t : mesh.triangles) std::fprintf(fp, "3 %d %d %d\n", t(0), t(1), t(2));
    std::fclose(fp);
}

void print_usage(const char* prog)
{
    std::fprintf(stderr, "Usage: %s <nrows> <ncols> <size> [options]\n\n", prog);
    std::fprintf(stderr, "Options:\n");
    std::fprintf(stderr, "  --steps <n>      Simulation steps        (default: 100)\n");
    std::fprintf(stderr, "  --dt <t>         Time step               (default: 0.01)\n");
    std::fprintf(stderr, "  --pin <mode>     top | corners           (default: top)\n");
    std::fprintf(stderr, "  --stiffness <k>  Stretch stiffness       (default: 1.0)\n");
    std::fprintf(stderr, "  --bend <k>       Bend stiffness          (default: 0 = off)\n");
    std::fprintf(stderr, "  --iter <n>       PD iterations/frame     (default: 50)\n");
    std::fprintf(stderr, "  --type <n>       Mesh type 0-3           (default: 3)\n");
    std::fprintf(stderr, "  --gravity <g>    Gravitational accel     (default: -9.8)\n");
    std::fprintf(stderr, "  --damping <d>    Velocity damping 0-1    (default: 0)\n");
    std::fprintf(stderr, "  --stretch-backend <b>  gpu-current | cpu-ref (default: gpu-current)\n");
    std::fprintf(stderr, "  --export <dir>   Export PLY frames\n");
    std::fprintf(stderr, "  --verbose        Per-frame energy output\n");
    std::fprintf(stderr, "Appended:\n");
    std::fprintf(stderr, "  --mode <m>       ref | pd                (default: ref)\n");
    std::fprintf(stderr, "  --tol <r>        pd: stop at this relative FP64 residual (default: 0 = fixed)\n");
    std::fprintf(stderr, "  --rho <r>        pd: Chebyshev spectral radius (default: 0.9)\n");
    std::fprintf(stderr, "  --relax <g|auto> pd: Jacobi step scale; auto from the Gershgorin bound (default: auto)\n");
    std::fprintf(stderr, "  --cheb-delay <n> pd: plain Jacobi iterations before the blend (default: 10)\n");
    std::fprintf(stderr, "  --restart <r>    pd: prev | seqmin | none (default: seqmin, tenfold growth)\n");
    std::fprintf(stderr, "  --obj <path>     mesh from an OBJ file instead of the grid\n");
    std::fprintf(stderr, "  --energy         with --verbose: stretch, bend, kinetic, potential terms\n");
}

}  // namespace

int main(int argc, char* argv[])
{
    if (argc < 4) {
        print_usage(argv[0]);
        return 1;
    }
    int nrows = 0, ncols = 0;
    float size = 0.0f;
    try {
        nrows = std::stoi(argv[1]);
        ncols = std::stoi(argv[2]);
        size = std::stof(argv[3]);
    } catch (...) {
        print_usage(argv[0]);
        return 1;
    }

    int steps = 100;
    float dt = 0.01f;
    std::string pin_mode = "top";
    float stiffness = 1.0f, bend_stiffness = 0.0f;
    int iterations = 50, mesh_type = 3;
    float gravity = -9.8f, damping = 0.0f;
    std::string stretch_backend = "gpu-current", export_dir, mode = "ref", obj;
    bool verbose = false, energy = false;
    float tol = 0.0f, rho = 0.9f;
    // pd: the damped step from the Gershgorin bound of D^-1 K, the Chebyshev
    // blend after 10 plain Jacobi iterations (Wang 2015), and a restart only
    // when the residual has grown tenfold over the sequence's best
    std::string relax = "auto", restart = "seqmin";
    int cheb_delay = 10;
    for (int i = 4; i < argc; ++i) {
        const std::string a = argv[i];
        const bool more = i + 1 < argc;
        if (a == "--steps" && more) steps = std::stoi(argv[++i]);
        else if (a == "--dt" && more) dt = std::stof(argv[++i]);
        else if (a == "--pin" && more) pin_mode = argv[++i];
        else if (a == "--stiffness" && more) stiffness = std::stof(argv[++i]);
        else if (a == "--bend" && more) bend_stiffness = std::stof(argv[++i]);
        else if (a == "--iter" && more) iterations = std::stoi(argv[++i]);
        else if (a == "--type" && more) mesh_type = std::stoi(argv[++i]);
        else if (a == "--gravity" && more) gravity = std::stof(argv[++i]);
        else if (a == "--damping" && more) damping = std::stof(argv[++i]);
        else if (a == "--stretch-backend" && more) stretch_backend = argv[++i];
        else if (a == "--export" && more) export_dir = argv[++i];
        else if (a == "--verbose") verbose = true;
        else if (a == "--mode" && more) mode = argv[++i];
        else if (a == "--tol" && more) tol = std::stof(argv[++i]);
        else if (a == "--rho" && more) rho = std::stof(argv[++i]);
        else if (a == "--relax" && more) relax = argv[++i];
        else if (a == "--cheb-delay" && more) cheb_delay = std::stoi(argv[++i]);
        else if (a == "--restart" && more) restart = argv[++i];
        else if (a == "--obj" && more) obj = argv[++i];
        else if (a == "--energy") energy = true;
    }
    if (stretch_backend != "gpu-current" && stretch_backend != "cpu-ref") {
        std::fprintf(stderr, "unknown stretch backend '%s' (gpu-current | cpu-ref)\n",
                     stretch_backend.c_str());
        return 1;
    }
    const bool use_cpu_ref = stretch_backend == "cpu-ref";
    if (restart != "prev" && restart != "seqmin" && restart != "none") {
        std::fprintf(stderr, "unknown restart rule '%s' (prev | seqmin | none)\n", restart.c_str());
        return 1;
    }
    if (mode != "ref" && mode != "pd") {
        std::fprintf(stderr, "unknown mode '%s' (ref | pd)\n", mode.c_str());
        return 1;
    }

    std::printf("=== PD Cloth Simulation ===\n");
    std::printf("Grid: %d x %d, size: %.4f, type: %d\n", nrows, ncols, size, mesh_type);
    std::printf("Steps: %d, dt: %.4f, iter: %d\n", steps, dt, iterations);
    std::printf("stiffness: %.3f, bend: %.4f, damping: %.3f, pin: %s, backend: %s\n", stiffness,
                bend_stiffness, damping, pin_mode.c_str(), stretch_backend.c_str());

    ClothMesh mesh;
    if (!obj.empty()) {
        if (!mesh.load_obj(obj)) return 1;
    } else {
        generate_square_cloth(nrows, ncols, size, mesh_type, mesh);
    }
    mesh.precompute_rest_state(0.1f);
    std::printf("\n");
    MeshTopology topo = MeshTopology::build(mesh);
    SimConstraints sim_cons;
    sim_cons.build_stretch(mesh, stiffness);
    if (bend_stiffness > 0.0f) sim_cons.build_bend(mesh, topo, bend_stiffness);
    std::printf("\nMesh: %d vertices, %d triangles, %d bend constraints\n", mesh.num_verts,
                mesh.num_tris, sim_cons.num_bend_cons);

    Constraints pin_cons;
    const int pin_cols = std::min(ncols, mesh.num_verts);
    if (pin_mode == "top") pin_cons.pin_top_row(mesh, pin_cols);
    else if (pin_mode == "corners") pin_cons.pin_corners(mesh, pin_cols);
    std::printf("Pinned vertices: %d\n", (int)pin_cons.pinned_indices.size());

    // the CPU reference backend steps the host state only, so it needs no
    // device (the reference uploads the mesh and pins regardless)
    if (!use_cpu_ref) {
        mesh.upload_to_gpu();
        pin_cons.upload_to_gpu();
        sim_cons.upload_to_gpu();
        sim_cons.precompute_jacobi_diag(mesh, dt);
    }

    PDSolverConfig config;
    config.dt = dt;
    config.max_iterations = iterations;
    config.stretch_stiffness = stiffness;
    config.bend_stiffness = bend_stiffness;
    config.gravity = gravity;
    config.damping = damping;
    config.use_chebyshev = true;
    config.rho = rho;
    config.mode = mode == "pd" ? 1 : 0;
    config.tolerance = tol;
    config.cheb_delay = cheb_delay;
    config.relax = relax == "auto" ? CLH_RELAX_AUTO : std::stof(relax);
    config.restart_rule = restart == "prev" ? CLH_RESTART_PREV
                        : restart == "none" ? CLH_RESTART_NONE : CLH_RESTART_SEQ_MIN;
    config.restart_factor = 10.0f;
    PDSolver* gpu_solver = nullptr;
    CpuStretchReferenceSolver* cpu_ref_solver = nullptr;
    if (use_cpu_ref)
        cpu_ref_solver = new CpuStretchReferenceSolver(mesh, sim_cons, pin_cons, dt, gravity, damping);
    else
        gpu_solver = new PDSolver(config, mesh, sim_cons);

    auto report = [&](const char* head) {
        std::printf("%s%.6f\n", head, reference_energy(mesh, config, use_cpu_ref));
    };
    auto report_terms = [&](int step) {
        std::vector<float> pos, vel;
        if (use_cpu_ref) host_state(mesh, pos, vel);
        else download_state(mesh, pos, vel);
        const ClothEnergy e = cloth_energy(mesh, sim_cons, gravity, pos, vel);
        std::printf("Step %d: kinetic %.6e potential %.6e stretch %.6e bend %.6e total %.6e", step,
                    e.kinetic, e.potential, e.stretch, e.bend, e.total());
        if (gpu_solver && config.mode == 1)
            std::printf(" | %d iterations, residual %.3e", gpu_solver->last_iterations(),
                        gpu_solver->last_relative_residual());
        std::printf("\n");
    };

    if (!export_dir.empty()) {
        char fn[512];
        std::snprintf(fn, sizeof(fn), "%s/frame_0000.ply", export_dir.c_str());
        export_ply(mesh, fn, use_cpu_ref);
    }
    std::printf("\nSimulating...\n");
    report("Initial energy: ");
    const int export_interval = std::max(1, steps / 10);
    for (int step = 1; step <= steps; ++step) {
        if (use_cpu_ref) cpu_ref_solver->step(mesh, pin_cons);
        else gpu_solver->step(mesh, sim_cons, pin_cons);
        if (verbose && step % 10 == 0) {
            if (energy) report_terms(step);
            else std::printf("Step %d: Energy = %.6f\n", step,
                             reference_energy(mesh, config, use_cpu_ref));
        }
        if (!export_dir.empty() && step % export_interval == 0) {
            char fn[512];
            std::snprintf(fn, sizeof(fn), "%s/frame_%04d.ply", export_dir.c_str(),
                          step / export_interval);
            export_ply(mesh, fn, use_cpu_ref);
        }
    }
    report("Final energy: ");
    if (!export_dir.empty()) {
        char fn[512];
        std::snprintf(fn, sizeof(fn), "%s/frame_final.ply", export_dir.c_str());
        export_ply(mesh, fn, use_cpu_ref);
        std::printf("\nExported frames to %s/\n", export_dir.c_str());
    }
    delete gpu_solver;
    delete cpu_ref_solver;
    if (!use_cpu_ref) pin_cons.free_gpu();
    std::printf("\nSimulation complete.\n");
    return 0;
}
