// MeshTopology::build (reference: src/mesh_topology.cpp:11-97).
//
// The edge list is the iteration order of a std::unordered_map keyed by the
// sorted vertex pair with hash (v0 << 32) | v1, reserved for 3T entries and
// filled triangle by triangle, corner edge (c, c+1 mod 3) for c = 0, 1, 2.
// The bend hinges inherit this order, and the reference defines its edge and
// hinge numbering by exactly these container operations, so they are
// reproduced as-is: with the same standard library the numbering is
// identical bit for bit.
#include "mesh_topology.h"
#include "check.h"

#include <cstdio>
#include <unordered_map>
#include <utility>

namespace {
struct EdgeKey {
    int lo, hi;
    bool operator==(const EdgeKey& o) const { return lo == o.lo && hi == o.hi; }
};
struct EdgeKeyHash {
    size_t operator()(const EdgeKey& k) const
    {
        return ((size_t)k.lo << 32) | (size_t)k.hi;
    }
};
// Local corner (0..2) of triangle `t` that is neither a nor b; -1 if none.
int opposite_corner(const ClothMesh& m, int t, int a, int b)
{
    for (int c = 0; c < 3; ++c) {
        const int v = m.triangles[t](c);
        if (v != a && v != b) return c;
    }
    return -1;
}
}  // namespace

MeshTopology MeshTopology::build(const ClothMesh& mesh)
{
    MeshTopology topo;
    const int N = mesh.num_verts, T = mesh.num_tris;

    topo.vert_to_tris.assign((size_t)N, std::vector<int>());
    for (int t = 0; t < T; ++t)
        for (int c = 0; c < 3; ++c) topo.vert_to_tris[mesh.triangles[t](c)].push_back(t);

    // sorted vertex pair -> (first triangle, second triangle or -1)
    std::unordered_map<EdgeKey, std::pair<int, int>, EdgeKeyHash> owners;
    owners.reserve((size_t)T * 3);
    for (int t = 0; t < T; ++t) {
        for (int c = 0; c < 3; ++c) {
            int a = mesh.triangles[t](c), b = mesh.triangles[t]((c + 1) % 3);
            if (a > b) std::swap(a, b);
            const EdgeKey key{a, b};
            auto hit = owners.find(key);
            if (hit == owners.end()) owners[key] = std::make_pair(t, -1);
            else                     hit->second.second = t;
        }
    }

    topo.edges.reserve(owners.size());
    for (const auto& kv : owners) {
        EdgeTopo e;
        e.v0 = kv.first.lo;
        e.v1 = kv.first.hi;
        e.tri_a = kv.second.first;
        e.tri_b = kv.second.second;
        e.rest_len = (mesh.rest_pos[e.v1] - mesh.rest_pos[e.v0]).norm();
        topo.edges.push_back(e);
    }

    topo.tri_to_tris.assign((size_t)T, cloth::Vec3i(-1, -1, -1));
    for (const EdgeTopo& e : topo.edges) {
        if (e.tri_b == -1) continue;
        const int ca = opposite_corner(mesh, e.tri_a, e.v0, e.v1);
        const int cb = opposite_corner(mesh, e.tri_b, e.v0, e.v1);
        if (ca != -1) topo.tri_to_tris[e.tri_a](ca) = e.tri_b;
        if (cb != -1) topo.tri_to_tris[e.tri_b](cb) = e.tri_a;
    }

    if (cloth::verbose()) {
        const int inner = topo.num_inner_edges();
        std::printf("=== MeshTopology::build ===\n");
        std::printf("  Vertices: %d  Triangles: %d\n", N, T);
        std::printf("  Total edges: %zu  (boundary: %d  inner: %d)\n", topo.edges.size(),
                    (int)topo.edges.size() - inner, inner);
    }
    return topo;
}

int MeshTopology::num_inner_edges() const
{
    int n = 0;
    for (const EdgeTopo& e : edges) n += (e.tri_b != -1);
    return n;
}
