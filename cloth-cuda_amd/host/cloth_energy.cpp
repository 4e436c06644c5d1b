// Energy diagnostics (see cloth_energy.h).
#include "cloth_energy.h"

#include "check.h"

#include <cmath>

namespace {

struct V3 {
    double x, y, z;
};
V3 sub(V3 a, V3 b) { return V3{a.x - b.x, a.y - b.y, a.z - b.z}; }
V3 add(V3 a, V3 b) { return V3{a.x + b.x, a.y + b.y, a.z + b.z}; }
V3 mul(V3 a, double s) { return V3{a.x * s, a.y * s, a.z * s}; }
double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
V3 at(const std::vector<float>& p, int i)
{
    return V3{p[(size_t)i * 3], p[(size_t)i * 3 + 1], p[(size_t)i * 3 + 2]};
}

// Closest rotation (3x2, orthonormal columns) to F = [F0 F1]: Gram-Schmidt
// frame of F, then the 2-D rotation maximising tr(R^T F2d).
void closest_rotation(V3 F0, V3 F1, V3& R0, V3& R1)
{
    const double l0 = std::sqrt(dot(F0, F0));
    const V3 u1 = l0 > 1e-20 ? mul(F0, 1.0 / l0) : V3{1.0, 0.0, 0.0};
    V3 c = sub(F1, mul(u1, dot(F1, u1)));
    const double lc = std::sqrt(dot(c, c));
    V3 u2;
    if (lc > 1e-20) {
        u2 = mul(c, 1.0 / lc);
    } else {
        u2 = V3{-u1.y, u1.x, 0.0};
        if (dot(u2, u2) < 1e-20) u2 = V3{0.0, -u1.z, u1.y};
        u2 = mul(u2, 1.0 / std::sqrt(dot(u2, u2)));
    }
    const double a00 = dot(u1, F0), a01 = dot(u1, F1), a10 = dot(u2, F0), a11 = dot(u2, F1);
    double cs = a00 + a11, sn = a10 - a01;
    const double n = std::sqrt(cs * cs + sn * sn);
    if (n > 1e-30) {
        cs /= n;
        sn /= n;
    } else {
        cs = 1.0;
        sn = 0.0;
    }
    R0 = add(mul(u1, cs), mul(u2, sn));
    R1 = sub(mul(u2, cs), mul(u1, sn));
}

}  // namespace

void download_state(const ClothMesh& mesh, std::vector<float>& pos, std::vector<float>& vel)
{
    const size_t n = (size_t)mesh.num_verts * 3;
    pos.resize(n);
    vel.resize(n);
    CLH_CHECK(clh_memcpy(pos.data(), mesh.d_pos, n * sizeof(float), CLH_D2H));
    CLH_CHECK(clh_memcpy(vel.data(), mesh.d_vel, n * sizeof(float), CLH_D2H));
}

void host_state(const ClothMesh& mesh, std::vector<float>& pos, std::vector<float>& vel)
{
    const size_t n = (size_t)mesh.num_verts;
    pos.resize(n * 3);
    vel.resize(n * 3);
    for (size_t i = 0; i < n && i < mesh.pos_cpu.size() && i < mesh.vel_cpu.size(); ++i)
        for (int k = 0; k < 3; ++k) {
            pos[i * 3 + k] = mesh.pos_cpu[i](k);
            vel[i * 3 + k] = mesh.vel_cpu[i](k);
        }
}

float reference_energy(const ClothMesh& mesh, const PDSolverConfig& cfg, bool use_cpu_ref)
{
    std::vector<float> pos, vel;
    if (use_cpu_ref) host_state(mesh, pos, vel);
    else download_state(mesh, pos, vel);
    float kinetic = 0.0f, potential = 0.0f;
    for (int i = 0; i < mesh.num_verts; ++i) {
        const float m = mesh.mass[(size_t)i];
        const float vx = vel[(size_t)i * 3], vy = vel[(size_t)i * 3 + 1], vz = vel[(size_t)i * 3 + 2];
        kinetic += 0.5f * m * (vx * vx + vy * vy + vz * vz);
        potential += m * (-cfg.gravity) * pos[(size_t)i * 3 + 1];
    }
    return kinetic + potential;
}

ClothEnergy cloth_energy(const ClothMesh& mesh, const SimConstraints& cons, float gravity,
                         const std::vector<float>& pos, const std::vector<float>& vel)
{
    ClothEnergy e;
    for (int i = 0; i < mesh.num_verts; ++i) {
        const double m = mesh.mass[(size_t)i];
        const V3 v = at(vel, i);
        e.kinetic += 0.5 * m * dot(v, v);
        e.potential += m * (-(double)gravity) * pos[(size_t)i * 3 + 1];
    }
    for (int t = 0; t < mesh.num_tris; ++t) {
        if ((size_t)t >= cons.tri_stretch_k.size()) break;
        const auto& tri = mesh.triangles[(size_t)t];
        const auto& G = mesh.Dm_inv[(size_t)t];
        const V3 x0 = at(pos, tri(0));
        const V3 d0 = sub(at(pos, tri(1)), x0), d1 = sub(at(pos, tri(2)), x0);
        const V3 F0 = add(mul(d0, G(0, 0)), mul(d1, G(1, 0)));
        const V3 F1 = add(mul(d0, G(0, 1)), mul(d1, G(1, 1)));
        V3 R0, R1;
        closest_rotation(F0, F1, R0, R1);
        const V3 E0 = sub(F0, R0), E1 = sub(F1, R1);
        e.stretch += 0.5 * cons.tri_stretch_k[(size_t)t] * mesh.rest_area[(size_t)t] *
                     (dot(E0, E0) + dot(E1, E1));
    }
    for (int h = 0; h < cons.num_bend_cons; ++h) {
        const auto& q = cons.bend_quads[(size_t)h];
        V3 c{0.0, 0.0, 0.0};
        for (int i = 0; i < 4; ++i) c = add(c, mul(at(pos, q(i)), cons.bend_w[(size_t)h * 4 + i]));
        const double len = std::sqrt(dot(c, c));
        const double nr = cons.bend_n[(size_t)h];
        const V3 p = (len > 1e-10 && nr > 1e-10) ? mul(c, nr / len) : V3{0.0, 0.0, 0.0};
        const V3 d = sub(c, p);
        e.bend += 0.5 * cons.bend_stiffness[(size_t)h] * dot(d, d);
    }
    return e;
}
