// ClothMesh: rest-state precomputation, OBJ import and device upload.
// Follows src/cloth_mesh.cpp of the reference (free_gpu 18-29,
// upload_to_gpu 31-83, load_obj 89-132, precompute_rest_state 138-211).
#include "cloth_mesh.h"
#include "check.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <sstream>

namespace cloth {
bool verbose()
{
    static const bool v = [] {
        const char* q = std::getenv("CLOTH_QUIET");
        return !(q && q[0] == '1');
    }();
    return v;
}
}  // namespace cloth

using cloth::Mat2f;
using cloth::Mat3x2f;
using cloth::Mat6x9f;
using cloth::Vec3f;

void ClothMesh::free_gpu()
{
    float** fbufs[] = {&d_pos, &d_vel, &d_mass, &d_Dm_inv, &d_rest_area};
    for (float** p : fbufs) {
        if (*p) clh_free(*p);
        *p = nullptr;
    }
    if (d_tris) clh_free(d_tris);
    d_tris = nullptr;
}

namespace {
template <typename T>
T* upload_array(const T* host, size_t count)
{
    void* d = nullptr;
    if (count == 0) return nullptr;
    CLH_CHECK(clh_alloc(&d, count * sizeof(T)));
    CLH_CHECK(clh_memcpy(d, host, count * sizeof(T), CLH_H2D));
    return static_cast<T*>(d);
}
}  // namespace

void ClothMesh::upload_to_gpu()
{
    free_gpu();
    const size_t N = (size_t)num_verts, T = (size_t)num_tris;

    // rest positions are the initial positions
    std::vector<float> pos(N * 3);
    for (size_t i = 0; i < N; ++i)
        for (int k = 0; k < 3; ++k) pos[i * 3 + k] = rest_pos[i](k);
    d_pos = upload_array(pos.data(), pos.size());

    void* v = nullptr;
    CLH_CHECK(clh_alloc(&v, N * 3 * sizeof(float)));
    CLH_CHECK(clh_memset(v, 0, N * 3 * sizeof(float)));
    d_vel = static_cast<float*>(v);

    d_mass = upload_array(mass.data(), N);

    std::vector<int> tris(T * 3);
    for (size_t t = 0; t < T; ++t)
        for (int k = 0; k < 3; ++k) tris[t * 3 + k] = triangles[t](k);
    d_tris = upload_array(tris.data(), tris.size());

    // column-major 2x2: [m00, m10, m01, m11]
    std::vector<float> dminv(T * 4);
    for (size_t t = 0; t < T; ++t) {
        dminv[t * 4 + 0] = Dm_inv[t](0, 0);
        dminv[t * 4 + 1] = Dm_inv[t](1, 0);
        dminv[t * 4 + 2] = Dm_inv[t](0, 1);
        dminv[t * 4 + 3] = Dm_inv[t](1, 1);
    }
    d_Dm_inv    = upload_array(dminv.data(), dminv.size());
    d_rest_area = upload_array(rest_area.data(), T);
}

bool ClothMesh::load_obj(const std::string& path)
{
    std::ifstream in(path);
    if (!in) {
        std::fprintf(stderr, "ClothMesh::load_obj: cannot open '%s'\n", path.c_str());
        return false;
    }
    rest_pos.clear();
    triangles.clear();

    std::string line;
    while (std::getline(in, line)) {
        if (line.empty() || line[0] == '#') continue;
        std::istringstream ls(line);
        std::string tag;
        ls >> tag;
        if (tag == "v") {
            float x = 0, y = 0, z = 0;
            ls >> x >> y >> z;
            rest_pos.emplace_back(x, y, z);
        } else if (tag == "f") {
            // "i", "i/t", "i/t/n" or "i//n"; only the 1-based vertex index is used
            std::vector<int> poly;
            std::string tok;
            while (ls >> tok) {
                int idx = 0;
                std::sscanf(tok.c_str(), "%d", &idx);
                poly.push_back(idx - 1);
            }
            // fan triangulation around the first corner
            for (size_t k = 1; k + 1 < poly.size(); ++k)
                triangles.emplace_back(poly[0], poly[k], poly[k + 1]);
        }
    }
    num_verts = (int)rest_pos.size();
    num_tris  = (int)triangles.size();
    if (num_verts == 0 || num_tris == 0) {
        std::fprintf(stderr, "ClothMesh::load_obj: empty mesh in '%s'\n", path.c_str());
        return false;
    }
    return true;
}

void ClothMesh::precompute_rest_state(float density)
{
    const size_t T = (size_t)num_tris;
    Dm_inv.assign(T, Mat2f());
    deltaUV.assign(T, Mat2f());
    dF_dx.assign(T, Mat6x9f());
    rest_area.assign(T, 0.0f);
    mass.assign((size_t)num_verts, 0.0f);
    pos_cpu = rest_pos;
    vel_cpu.assign((size_t)num_verts, Vec3f());

    for (size_t t = 0; t < T; ++t) {
        const int i0 = triangles[t](0), i1 = triangles[t](1), i2 = triangles[t](2);
        const Vec3f e1 = rest_pos[i1] - rest_pos[i0];
        const Vec3f e2 = rest_pos[i2] - rest_pos[i0];

        // Orthonormal frame (p0, p1) of the triangle plane by Gram-Schmidt.
        const Vec3f p0 = e1.normalized();
        const Vec3f ortho = e2 - p0 * e2.dot(p0);
        Vec3f p1;
        if (ortho.norm() > 1e-10f) {
            p1 = ortho.normalized();
        } else {
            // degenerate triangle: any unit vector perpendicular to p0
            const float ax = std::abs(p0(0)), ay = std::abs(p0(1)), az = std::abs(p0(2));
            if (ax <= ay && ax <= az) p1 = Vec3f(0.0f, -p0(2), p0(1)).normalized();
            else                      p1 = Vec3f(-p0(1), p0(0), 0.0f).normalized();
        }

        // Dm = P^T [e1 e2]
        Mat2f Dm;
        Dm(0, 0) = p0.dot(e1); Dm(0, 1) = p0.dot(e2);
        Dm(1, 0) = p1.dot(e1); Dm(1, 1) = p1.dot(e2);
        deltaUV[t] = Dm;
        Dm_inv[t]  = Dm.inverse();

        const float area = 0.5f * std::abs(Dm.determinant());
        rest_area[t] = area;

        // dF/dx = kron(Dm_inv^T * [[-1,1,0],[-1,0,1]], I3): row block r (F column r),
        // column block v (vertex v) is D(r, v) * I3.
        const Mat2f& G = Dm_inv[t];
        float D[2][3];
        for (int r = 0; r < 2; ++r) {
            D[r][1] = G(0, r);
            D[r][2] = G(1, r);
            D[r][0] = -G(0, r) - G(1, r);
        }
        Mat6x9f deriv;
        for (int r = 0; r < 2; ++r)
            for (int v = 0; v < 3; ++v)
                for (int k = 0; k < 3; ++k) deriv(r * 3 + k, v * 3 + k) = D[r][v];
        dF_dx[t] = deriv;

        const float third = density * area / 3.0f;
        mass[i0] += third;
        mass[i1] += third;
        mass[i2] += third;
    }
}

void ClothMesh::print_stats() const
{
    std::printf("=== ClothMesh stats ===\n");
    std::printf("  Vertices : %d\n", num_verts);
    std::printf("  Triangles: %d\n", num_tris);
    if (!rest_area.empty()) {
        float sum = 0.0f, lo = rest_area[0], hi = rest_area[0];
        for (float a : rest_area) { sum += a; lo = std::min(lo, a); hi = std::max(hi, a); }
        std::printf("  Total rest area : %.6f (min=%.6f, max=%.6f)\n", sum, lo, hi);
    }
    if (!mass.empty()) {
        float sum = 0.0f, lo = mass[0], hi = mass[0];
        for (float m : mass) { sum += m; lo = std::min(lo, m); hi = std::max(hi, m); }
        std::printf("  Total mesh mass : %.6f (min=%.6f, max=%.6f)\n", sum, lo, hi);
    }
    for (int t = 0; t < std::min(num_tris, 3); ++t)
        std::printf("  tri[%d]: area=%.6f  Dm_inv=[[%.4f,%.4f],[%.4f,%.4f]]\n", t,
                    rest_area[t], Dm_inv[t](0, 0), Dm_inv[t](0, 1), Dm_inv[t](1, 0),
                    Dm_inv[t](1, 1));
    std::printf("  GPU buffers: %s\n", d_pos ? "allocated" : "not allocated");
}
