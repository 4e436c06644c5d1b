#pragma once
// Energy diagnostics of a cloth state, computed on the host in double from
// positions and velocities copied off the device.
//
// The reference's driver prints kinetic + gravitational potential energy in
// FP32 (src/tools/sim_cloth.cpp:64-105); these add the elastic energies of
// the projective-dynamics constraints the solver minimises:
//   stretch  sum_t 1/2 k_t A_t |F_t - R(F_t)|_F^2   (R the closest rotation)
//   bend     sum_e 1/2 k_e |e_e - p(e_e)|^2          (e = sum_i w_i x_i,
//            p = e n_rest / |e|)
#include "cloth_mesh.h"
#include "pd_solver.h"
#include "sim_constraints.h"

#include <vector>

struct ClothEnergy {
    double kinetic   = 0.0;  // 1/2 sum m |v|^2
    double potential = 0.0;  // sum m (-g) y
    double stretch   = 0.0;
    double bend      = 0.0;
    double total() const { return kinetic + potential + stretch + bend; }
};

// The reference's compute_energy: kinetic + potential accumulated in FP32
// over the device state, or over the host state pos_cpu / vel_cpu of the CPU
// reference backend (src/tools/sim_cloth.cpp:64-105).
float reference_energy(const ClothMesh& mesh, const PDSolverConfig& cfg, bool use_cpu_ref = false);

// All four terms from host arrays pos, vel [N*3].
ClothEnergy cloth_energy(const ClothMesh& mesh, const SimConstraints& cons, float gravity,
                         const std::vector<float>& pos, const std::vector<float>& vel);

// Copies the device state of mesh into pos, vel.
void download_state(const ClothMesh& mesh, std::vector<float>& pos, std::vector<float>& vel);
// Copies the host state pos_cpu, vel_cpu of mesh into pos, vel.
void host_state(const ClothMesh& mesh, std::vector<float>& pos, std::vector<float>& vel);
