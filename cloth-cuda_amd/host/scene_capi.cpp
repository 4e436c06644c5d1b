// C API for building seeded workloads through the C++ cloth API.
#include "cloth_scene.h"

#include "check.h"
#include "cloth_energy.h"
#include "cloth_mesh.h"
#include "cloth_rng.h"
#include "constraints.h"
#include "mesh_generator.h"
#include "mesh_topology.h"
#include "pd_solver.h"
#include "sim_constraints.h"
#include "stretch_reference.h"

#include <algorithm>
#include <cstring>
#include <set>
#include <string>
#include <vector>

struct cc_scene {
    cc_scene_spec  spec;
    ClothMesh      mesh;
    MeshTopology   topo;
    SimConstraints sim;
    Constraints    pins;
    std::vector<float> pos0;  // initial positions [N*3]
    // per-cloth offsets [C+1] into the vertex, triangle, hinge and pin arrays
    std::vector<int> cloth_vert{0}, cloth_tri{0}, cloth_bend{0}, cloth_pin{0};
    bool uploaded = false;
    ~cc_scene() { pins.free_gpu(); }
};

namespace {
// Independent SplitMix64 streams per purpose, derived from the spec seed.
constexpr uint64_t kStreamStretch = 0x53545245544348ull;  // stretch stiffness draws
constexpr uint64_t kStreamPins    = 0x50494E53ull;        // random boundary pins
constexpr uint64_t kStreamJitter  = 0x4A49545445520000ull; // initial-position jitter

std::vector<int> boundary_vertices(const cc_scene_spec& sp)
{
    std::vector<int> b;
    for (int i = 0; i < sp.nrows; ++i)
        for (int j = 0; j < sp.ncols; ++j)
            if (i == 0 || j == 0 || i == sp.nrows - 1 || j == sp.ncols - 1)
                b.push_back(i * sp.ncols + j);
    return b;
}
}  // namespace

extern "C" void cc_scene_spec_default(cc_scene_spec* sp)
{
    std::memset(sp, 0, sizeof(*sp));
    sp->nrows = sp->ncols = 32;
    sp->size = 1.0f / 31.0f;
    sp->type = 0;
    sp->seed = 1;
    sp->density = 0.1f;
    sp->stretch_k = 1e3f;
    sp->stretch_k_max = 0.0f;
    sp->bend_k = 1e-2f;
    sp->pin_mode = CC_PIN_CORNERS;
    sp->pin_count = 0;
    sp->pos_jitter = 1e-4f;
    sp->dt = 1.0f / 60.0f;
}

namespace {
// Rest state, topology, constraints, pins and initial positions of a scene
// whose mesh geometry is already in s->mesh.
void finish_scene(cc_scene* s);
}  // namespace

extern "C" cc_scene* cc_scene_build(const cc_scene_spec* spec)
{
    cc_scene* s = new cc_scene();
    s->spec = *spec;
    const cc_scene_spec& sp = s->spec;
    if (sp.type == CC_MESH_IRREGULAR)
        generate_irregular_cloth(sp.nrows, sp.ncols, sp.size, sp.rest_jitter, sp.seed, s->mesh);
    else
        generate_square_cloth(sp.nrows, sp.ncols, sp.size, sp.type, s->mesh);
    finish_scene(s);
    return s;
}

extern "C" cc_scene* cc_scene_build_obj(const char* path, const cc_scene_spec* spec)
{
    cc_scene* s = new cc_scene();
    s->spec = *spec;
    if (!s->mesh.load_obj(path)) {
        delete s;
        return nullptr;
    }
    finish_scene(s);
    return s;
}

namespace {
void finish_scene(cc_scene* s)
{
    const cc_scene_spec& sp = s->spec;
    s->mesh.precompute_rest_state(sp.density);
    s->topo = MeshTopology::build(s->mesh);

    s->sim.build_stretch(s->mesh, sp.stretch_k);
    if (sp.stretch_k_max > sp.stretch_k) {
        cloth::SplitMix64 rng(sp.seed ^ kStreamStretch);
        for (float& k : s->sim.tri_stretch_k)
            k = (float)rng.log_uniform(sp.stretch_k, sp.stretch_k_max);
    }
    if (sp.bend_k > 0.0f) s->sim.build_bend(s->mesh, s->topo, sp.bend_k);

    switch (sp.pin_mode) {
        case CC_PIN_TOP_ROW: s->pins.pin_top_row(s->mesh, std::min(sp.ncols, s->mesh.num_verts)); break;
        case CC_PIN_CORNERS: s->pins.pin_corners(s->mesh, std::min(sp.ncols, s->mesh.num_verts)); break;
        case CC_PIN_RANDOM_BOUNDARY: {
            std::vector<int> b = boundary_vertices(sp);
            b.erase(std::remove_if(b.begin(), b.end(), [&](int v) { return v >= s->mesh.num_verts; }),
                    b.end());
            cloth::SplitMix64 rng(sp.seed ^ kStreamPins);
            std::vector<int> chosen;
            std::set<int> seen;
            const int want = std::min<int>(std::max(sp.pin_count, 0), (int)b.size());
            while ((int)chosen.size() < want) {
                const int v = b[rng.below(b.size())];
                if (seen.insert(v).second) chosen.push_back(v);
            }
            s->pins.set_from_list(chosen, s->mesh);
            break;
        }
        default: s->pins.set_from_list({}, s->mesh); break;
    }

    s->sim.jacobi_diag = s->sim.compute_jacobi_diag(s->mesh, sp.dt);

    const int N = s->mesh.num_verts;
    s->cloth_vert = {0, N};
    s->cloth_tri = {0, s->mesh.num_tris};
    s->cloth_bend = {0, s->sim.num_bend_cons};
    s->cloth_pin = {0, (int)s->pins.pinned_indices.size()};
    s->pos0.resize((size_t)N * 3);
    for (int i = 0; i < N; ++i)
        for (int k = 0; k < 3; ++k) s->pos0[(size_t)i * 3 + k] = s->mesh.rest_pos[i](k);
    if (sp.pos_jitter > 0.0f) {
        std::vector<char> pinned((size_t)N, 0);
        for (int v : s->pins.pinned_indices) pinned[(size_t)v] = 1;
        cloth::SplitMix64 rng(sp.seed ^ kStreamJitter);
        for (int i = 0; i < N; ++i) {
            const float d = (float)rng.uniform(-sp.pos_jitter, sp.pos_jitter);
            if (!pinned[(size_t)i]) s->pos0[(size_t)i * 3 + 1] += d;
        }
    }
}
}  // namespace

extern "C" void cc_scene_destroy(cc_scene* s) { delete s; }

namespace {
// Appends cloth `c` to the batch `b`, shifting its vertex and triangle ids.
void append_cloth(cc_scene& b, const cc_scene& c)
{
    const int vo = b.mesh.num_verts, to = b.mesh.num_tris;
    ClothMesh& m = b.mesh;
    const ClothMesh& cm = c.mesh;
    m.rest_pos.insert(m.rest_pos.end(), cm.rest_pos.begin(), cm.rest_pos.end());
    for (const auto& t : cm.triangles) m.triangles.push_back(cloth::Vec3i(t(0) + vo, t(1) + vo, t(2) + vo));
    m.Dm_inv.insert(m.Dm_inv.end(), cm.Dm_inv.begin(), cm.Dm_inv.end());
    m.deltaUV.insert(m.deltaUV.end(), cm.deltaUV.begin(), cm.deltaUV.end());
    m.dF_dx.insert(m.dF_dx.end(), cm.dF_dx.begin(), cm.dF_dx.end());
    m.rest_area.insert(m.rest_area.end(), cm.rest_area.begin(), cm.rest_area.end());
    m.mass.insert(m.mass.end(), cm.mass.begin(), cm.mass.end());
    m.pos_cpu.insert(m.pos_cpu.end(), cm.pos_cpu.begin(), cm.pos_cpu.end());
    m.vel_cpu.insert(m.vel_cpu.end(), cm.vel_cpu.begin(), cm.vel_cpu.end());
    m.num_verts += cm.num_verts;
    m.num_tris += cm.num_tris;

    SimConstraints& sc = b.sim;
    const SimConstraints& cs = c.sim;
    sc.tri_stretch_k.insert(sc.tri_stretch_k.end(), cs.tri_stretch_k.begin(), cs.tri_stretch_k.end());
    for (const auto& q : cs.bend_quads)
        sc.bend_quads.push_back(cloth::Vec4i(q(0) + vo, q(1) + vo, q(2) + vo, q(3) + vo));
    sc.bend_w.insert(sc.bend_w.end(), cs.bend_w.begin(), cs.bend_w.end());
    sc.bend_n.insert(sc.bend_n.end(), cs.bend_n.begin(), cs.bend_n.end());
    sc.bend_stiffness.insert(sc.bend_stiffness.end(), cs.bend_stiffness.begin(), cs.bend_stiffness.end());
    sc.jacobi_diag.insert(sc.jacobi_diag.end(), cs.jacobi_diag.begin(), cs.jacobi_diag.end());
    sc.num_tris += cs.num_tris;
    sc.num_bend_cons += cs.num_bend_cons;

    for (auto e : c.topo.edges) {
        e.v0 += vo;
        e.v1 += vo;
        e.tri_a += to;
        if (e.tri_b >= 0) e.tri_b += to;
        b.topo.edges.push_back(e);
    }
    for (const auto& l : c.topo.vert_to_tris) {
        std::vector<int> s2;
        for (int t : l) s2.push_back(t + to);
        b.topo.vert_to_tris.push_back(s2);
    }
    for (const auto& t : c.topo.tri_to_tris)
        b.topo.tri_to_tris.push_back(cloth::Vec3i(t(0) >= 0 ? t(0) + to : t(0),
                                                  t(1) >= 0 ? t(1) + to : t(1),
                                                  t(2) >= 0 ? t(2) + to : t(2)));

    for (int v : c.pins.pinned_indices) b.pins.pinned_indices.push_back(v + vo);
    b.pins.target_positions.insert(b.pins.target_positions.end(), c.pins.target_positions.begin(),
                                   c.pins.target_positions.end());
    b.pos0.insert(b.pos0.end(), c.pos0.begin(), c.pos0.end());

    b.cloth_vert.push_back(m.num_verts);
    b.cloth_tri.push_back(m.num_tris);
    b.cloth_bend.push_back(sc.num_bend_cons);
    b.cloth_pin.push_back((int)b.pins.pinned_indices.size());
}
}  // namespace

extern "C" cc_scene* cc_batch_build(const cc_scene_spec* specs, int count)
{
    if (!specs || count < 1) return nullptr;
    for (int c = 1; c < count; ++c)
        if (specs[c].dt != specs[0].dt) return nullptr;  // one Jacobi diagonal time step
    cc_scene* b = new cc_scene();
    b->spec = specs[0];
    for (int c = 0; c < count; ++c) {
        cc_scene* one = cc_scene_build(&specs[c]);
        append_cloth(*b, *one);
        cc_scene_destroy(one);
    }
    return b;
}

extern "C" int cc_scene_num_cloths(const cc_scene* s) { return (int)s->cloth_vert.size() - 1; }

extern "C" int cc_scene_cloth_offsets(const cc_scene* s, int* vert, int* tri, int* bend, int* pin)
{
    const size_t n = s->cloth_vert.size();
    if (vert) std::memcpy(vert, s->cloth_vert.data(), n * sizeof(int));
    if (tri) std::memcpy(tri, s->cloth_tri.data(), n * sizeof(int));
    if (bend) std::memcpy(bend, s->cloth_bend.data(), n * sizeof(int));
    if (pin) std::memcpy(pin, s->cloth_pin.data(), n * sizeof(int));
    return 0;
}

extern "C" void cc_scene_counts(const cc_scene* s, int* N, int* T, int* E, int* P)
{
    if (N) *N = s->mesh.num_verts;
    if (T) *T = s->mesh.num_tris;
    if (E) *E = s->sim.num_bend_cons;
    if (P) *P = (int)s->pins.pinned_indices.size();
}

extern "C" int cc_scene_num_edges(const cc_scene* s) { return (int)s->topo.edges.size(); }

namespace {
template <typename T>
int put(void* dst, size_t bytes, const std::vector<T>& v)
{
    if (bytes != v.size() * sizeof(T)) return 1;
    if (bytes) std::memcpy(dst, v.data(), bytes);
    return 0;
}
}  // namespace

extern "C" int cc_scene_get(const cc_scene* s, const char* name, void* dst, size_t bytes)
{
    const std::string n(name);
    const ClothMesh& m = s->mesh;
    const SimConstraints& c = s->sim;
    if (n == "pos0") return put(dst, bytes, s->pos0);
    if (n == "rest_pos") {
        std::vector<float> v;
        for (const auto& p : m.rest_pos) for (int k = 0; k < 3; ++k) v.push_back(p(k));
        return put(dst, bytes, v);
    }
    if (n == "tris") {
        std::vector<int> v;
        for (const auto& t : m.triangles) for (int k = 0; k < 3; ++k) v.push_back(t(k));
        return put(dst, bytes, v);
    }
    if (n == "Dm_inv" || n == "deltaUV") {
        const auto& src = (n == "Dm_inv") ? m.Dm_inv : m.deltaUV;
        std::vector<float> v;
        for (const auto& g : src) { v.push_back(g(0, 0)); v.push_back(g(1, 0)); v.push_back(g(0, 1)); v.push_back(g(1, 1)); }
        return put(dst, bytes, v);
    }
    if (n == "dF_dx") {
        std::vector<float> v;  // row-major 6x9 per triangle
        for (const auto& d : m.dF_dx) for (int r = 0; r < 6; ++r) for (int cc = 0; cc < 9; ++cc) v.push_back(d(r, cc));
        return put(dst, bytes, v);
    }
    if (n == "rest_area") return put(dst, bytes, m.rest_area);
    if (n == "mass") return put(dst, bytes, m.mass);
    if (n == "stretch_k") return put(dst, bytes, c.tri_stretch_k);
    if (n == "jacobi_diag") return put(dst, bytes, c.jacobi_diag);
    if (n == "bend_quads") {
        std::vector<int> v;
        for (const auto& q : c.bend_quads) for (int k = 0; k < 4; ++k) v.push_back(q(k));
        return put(dst, bytes, v);
    }
    if (n == "bend_w") return put(dst, bytes, c.bend_w);
    if (n == "bend_n") return put(dst, bytes, c.bend_n);
    if (n == "bend_k") return put(dst, bytes, c.bend_stiffness);
    if (n == "pin_idx") return put(dst, bytes, s->pins.pinned_indices);
    if (n == "pin_target") {
        std::vector<float> v;
        for (int idx : s->pins.pinned_indices) for (int k = 0; k < 3; ++k) v.push_back(s->pins.target_positions[(size_t)idx * 3 + k]);
        return put(dst, bytes, v);
    }
    if (n == "edges") {
        std::vector<int> v;
        for (const auto& e : s->topo.edges) { v.push_back(e.v0); v.push_back(e.v1); v.push_back(e.tri_a); v.push_back(e.tri_b); }
        return put(dst, bytes, v);
    }
    if (n == "edge_len") {
        std::vector<float> v;
        for (const auto& e : s->topo.edges) v.push_back(e.rest_len);
        return put(dst, bytes, v);
    }
    if (n == "vt_ptr" || n == "vt_idx") {
        std::vector<int> ptr(1, 0), idx;
        for (const auto& l : s->topo.vert_to_tris) { idx.insert(idx.end(), l.begin(), l.end()); ptr.push_back((int)idx.size()); }
        return put(dst, bytes, n == "vt_ptr" ? ptr : idx);
    }
    if (n == "tri_to_tris") {
        std::vector<int> v;
        for (const auto& t : s->topo.tri_to_tris) for (int k = 0; k < 3; ++k) v.push_back(t(k));
        return put(dst, bytes, v);
    }
    return 2;
}

extern "C" int cc_scene_upload(cc_scene* s)
{
    s->mesh.upload_to_gpu();
    s->sim.upload_to_gpu();
    s->sim.precompute_jacobi_diag(s->mesh, s->spec.dt);  // after upload_to_gpu
    s->pins.upload_to_gpu();
    const size_t bytes = s->pos0.size() * sizeof(float);
    if (clh_memcpy(s->mesh.d_pos, s->pos0.data(), bytes, CLH_H2D) != CLH_OK) return 1;
    s->uploaded = true;
    return 0;
}

extern "C" void* cc_scene_device(const cc_scene* s, const char* name)
{
    const std::string n(name);
    const ClothMesh& m = s->mesh;
    const SimConstraints& c = s->sim;
    if (n == "pos") return m.d_pos;
    if (n == "vel") return m.d_vel;
    if (n == "mass") return m.d_mass;
    if (n == "tris") return m.d_tris;
    if (n == "Dm_inv") return m.d_Dm_inv;
    if (n == "rest_area") return m.d_rest_area;
    if (n == "stretch_k") return c.d_tri_stretch_k;
    if (n == "jacobi_diag") return c.d_jacobi_diag;
    if (n == "bend_quads") return c.d_bend_quads;
    if (n == "bend_w") return c.d_bend_w;
    if (n == "bend_n") return c.d_bend_n;
    if (n == "bend_k") return c.d_bend_k;
    if (n == "pin_idx") return s->pins.d_pinned_indices;
    if (n == "pin_target") return s->pins.d_target_pos;
    return nullptr;
}

extern "C" int cc_scene_set_state(cc_scene* s, const float* pos, const float* vel)
{
    const size_t bytes = (size_t)s->mesh.num_verts * 3 * sizeof(float);
    if (pos && clh_memcpy(s->mesh.d_pos, pos, bytes, CLH_H2D) != CLH_OK) return 1;
    if (vel && clh_memcpy(s->mesh.d_vel, vel, bytes, CLH_H2D) != CLH_OK) return 1;
    return 0;
}

extern "C" int cc_scene_get_state(const cc_scene* s, float* pos, float* vel)
{
    const size_t bytes = (size_t)s->mesh.num_verts * 3 * sizeof(float);
    if (pos && clh_memcpy(pos, s->mesh.d_pos, bytes, CLH_D2H) != CLH_OK) return 1;
    if (vel && clh_memcpy(vel, s->mesh.d_vel, bytes, CLH_D2H) != CLH_OK) return 1;
    return 0;
}

extern "C" void cc_pd_config_default(cc_pd_config* c)
{
    const PDSolverConfig d;
    c->max_iterations = d.max_iterations;
    c->tolerance = d.tolerance;
    c->use_chebyshev = d.use_chebyshev ? 1 : 0;
    c->rho = d.rho;
    c->gravity = d.gravity;
    c->dt = d.dt;
    c->damping = d.damping;
    c->mode = d.mode;
    for (int k = 0; k < 3; ++k) c->external_accel[k] = d.external_accel[k];
    c->check_every = d.check_every;
    c->cheb_delay = d.cheb_delay;
    c->kernel_variant = d.kernel_variant;
    c->use_graph = d.use_graph ? 1 : 0;
}

extern "C" void* cc_pdsolver_create(cc_scene* s, const cc_pd_config* c)
{
    PDSolverConfig cfg;
    cfg.max_iterations = c->max_iterations;
    cfg.tolerance = c->tolerance;
    cfg.use_chebyshev = c->use_chebyshev != 0;
    cfg.rho = c->rho;
    cfg.gravity = c->gravity;
    cfg.dt = c->dt;
    cfg.damping = c->damping;
    cfg.mode = c->mode;
    for (int k = 0; k < 3; ++k) cfg.external_accel[k] = c->external_accel[k];
    cfg.check_every = c->check_every;
    cfg.cheb_delay = c->cheb_delay;
    cfg.kernel_variant = c->kernel_variant;
    cfg.use_graph = c->use_graph != 0;
    return new PDSolver(cfg, s->mesh, s->sim);
}

extern "C" void cc_pdsolver_step(void* solver, cc_scene* s)
{
    static_cast<PDSolver*>(solver)->step(s->mesh, s->sim, s->pins);
}

extern "C" void cc_pdsolver_reset(void* solver) { static_cast<PDSolver*>(solver)->reset(); }

extern "C" void* cc_pdsolver_native(void* solver)
{
    return static_cast<PDSolver*>(solver)->native_handle();
}

extern "C" void cc_pdsolver_destroy(void* solver) { delete static_cast<PDSolver*>(solver); }

extern "C" void cc_scene_energy(const cc_scene* s, float gravity, const float* pos,
                                const float* vel, double* out)
{
    const size_t n = (size_t)s->mesh.num_verts * 3;
    const std::vector<float> p(pos, pos + n), v(vel, vel + n);
    const ClothEnergy e = cloth_energy(s->mesh, s->sim, gravity, p, v);
    out[0] = e.kinetic;
    out[1] = e.potential;
    out[2] = e.stretch;
    out[3] = e.bend;
}

extern "C" void cc_scene_apply_pins(cc_scene* s) { s->pins.apply_gpu(s->mesh.d_pos, s->mesh.d_vel); }

extern "C" int cc_scene_get_cpu_state(const cc_scene* s, float* pos, float* vel)
{
    const ClothMesh& m = s->mesh;
    if ((int)m.pos_cpu.size() != m.num_verts || (int)m.vel_cpu.size() != m.num_verts) return 1;
    for (int i = 0; i < m.num_verts; ++i)
        for (int k = 0; k < 3; ++k) {
            if (pos) pos[(size_t)i * 3 + k] = m.pos_cpu[(size_t)i](k);
            if (vel) vel[(size_t)i * 3 + k] = m.vel_cpu[(size_t)i](k);
        }
    return 0;
}

extern "C" int cc_scene_set_cpu_state(cc_scene* s, const float* pos, const float* vel)
{
    ClothMesh& m = s->mesh;
    m.pos_cpu.resize((size_t)m.num_verts);
    m.vel_cpu.resize((size_t)m.num_verts);
    for (int i = 0; i < m.num_verts; ++i)
        for (int k = 0; k < 3; ++k) {
            if (pos) m.pos_cpu[(size_t)i](k) = pos[(size_t)i * 3 + k];
            if (vel) m.vel_cpu[(size_t)i](k) = vel[(size_t)i * 3 + k];
        }
    return 0;
}

extern "C" void* cc_cpuref_create(cc_scene* s, float gravity, float damping, int refactor)
{
    auto* r = new CpuStretchReferenceSolver(s->mesh, s->sim, s->pins, s->spec.dt, gravity, damping);
    r->refactor_every_iteration = refactor != 0;
    return r;
}

extern "C" void cc_cpuref_step(void* solver, cc_scene* s, int steps)
{
    auto* r = static_cast<CpuStretchReferenceSolver*>(solver);
    for (int i = 0; i < steps; ++i) r->step(s->mesh, s->pins);
}

extern "C" int cc_cpuref_factorizations(void* solver)
{
    return static_cast<CpuStretchReferenceSolver*>(solver)->factorizations();
}

extern "C" void cc_cpuref_destroy(void* solver) { delete static_cast<CpuStretchReferenceSolver*>(solver); }
