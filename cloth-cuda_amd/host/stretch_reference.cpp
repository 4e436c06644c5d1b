// CpuStretchReferenceSolver (reference: src/stretch_reference.cpp:38-280):
// projective dynamics with stretch only, an exact global solve and the
// reference's treatment of pins, on the host state mesh.pos_cpu / vel_cpu.
//
// The global matrix K = M + h^2 sum kA dF/dx^T dF/dx is constant, and with
// dF/dx = kron(D, I3) it is one scalar N x N matrix shared by x, y and z.
// Pinned rows and columns are replaced by identity rows, so the unpinned
// vertices form their own system K_uu x_u = b_u. That system is factored
// once per pin set (FP64 Cholesky, reverse Cuthill-McKee order, envelope
// storage) and each iteration solves it for the three components.
#include "stretch_reference.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <queue>

namespace {

constexpr double kGuard = 1e-10;  // the reference's degeneracy threshold

struct V3 {
    double x, y, z;
};
inline V3 sub(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 add(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 scl(V3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
inline double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline double len(V3 a) { return std::sqrt(dot(a, a)); }

// Reverse Cuthill-McKee order of the graph adj restricted to vertices with
// keep[v]: each component is numbered breadth first from a vertex of
// smallest degree, neighbours by increasing degree, and the result reversed.
std::vector<int> rcm_order(const std::vector<std::vector<int>>& adj, const std::vector<char>& keep)
{
    const int N = (int)adj.size();
    std::vector<int> deg((size_t)N, 0), order;
    for (int v = 0; v < N; ++v)
        if (keep[(size_t)v])
            for (int w : adj[(size_t)v]) deg[(size_t)v] += keep[(size_t)w] ? 1 : 0;
    std::vector<int> seeds;
    for (int v = 0; v < N; ++v)
        if (keep[(size_t)v]) seeds.push_back(v);
    std::stable_sort(seeds.begin(), seeds.end(), [&](int a, int b) { return deg[(size_t)a] < deg[(size_t)b]; });
    std::vector<char> seen((size_t)N, 0);
    std::vector<int> nb;
    for (int s : seeds) {
        if (seen[(size_t)s]) continue;
        std::queue<int> q;
        q.push(s);
        seen[(size_t)s] = 1;
        while (!q.empty()) {
            const int v = q.front();
            q.pop();
            order.push_back(v);
            nb.clear();
            for (int w : adj[(size_t)v])
                if (keep[(size_t)w] && !seen[(size_t)w]) {
                    seen[(size_t)w] = 1;
                    nb.push_back(w);
                }
            std::stable_sort(nb.begin(), nb.end(), [&](int a, int b) { return deg[(size_t)a] < deg[(size_t)b]; });
            for (int w : nb) q.push(w);
        }
    }
    std::reverse(order.begin(), order.end());
    return order;
}

}  // namespace

CpuStretchReferenceSolver::CpuStretchReferenceSolver(const ClothMesh& mesh,
                                                     const SimConstraints& sim_cons,
                                                     const Constraints& /*pin_cons*/,
                                                     float dt, float gravity, float damping)
    : num_verts_(mesh.num_verts), num_tris_(mesh.num_tris), dt_(dt), gravity_(gravity),
      damping_(damping)
{
    const size_t T = (size_t)num_tris_;
    masses_ = mesh.mass;
    kA_.assign(T, 0.0);
    tris_.assign(T * 3, 0);
    dm_inv_.assign(T * 4, 0.0f);
    coef_.assign(T * 6, 0.0);
    if ((int)masses_.size() != num_verts_ || mesh.rest_area.size() != T || mesh.Dm_inv.size() != T ||
        mesh.dF_dx.size() != T || (int)mesh.triangles.size() != num_tris_) {
        error_ = "the mesh has no rest state (call precompute_rest_state first)";
        return;
    }
    for (size_t t = 0; t < T; ++t) {
        // the reference skips a triangle whose sqrt(max(0, k A)) is zero
        const float k = t < sim_cons.tri_stretch_k.size() ? sim_cons.tri_stretch_k[t] : 0.0f;
        const double kA = (double)k * (double)mesh.rest_area[t];
        kA_[t] = kA > 0.0 ? kA : 0.0;
        for (int c = 0; c < 3; ++c) {
            const int v = mesh.triangles[t](c);
            if (v < 0 || v >= num_verts_) {
                error_ = "triangle vertex index out of range";
                return;
            }
            tris_[t * 3 + c] = v;
        }
        const cloth::Mat2f& g = mesh.Dm_inv[t];
        dm_inv_[t * 4 + 0] = g(0, 0);
        dm_inv_[t * 4 + 1] = g(1, 0);
        dm_inv_[t * 4 + 2] = g(0, 1);
        dm_inv_[t * 4 + 3] = g(1, 1);
        // dF/dx must be kron(D, I3): row r*3+k (component k of F column r),
        // column a*3+k' (component k' of corner a) holds D[r][a] if k == k'
        const cloth::Mat6x9f& d = mesh.dF_dx[t];
        for (int r = 0; r < 2; ++r)
            for (int a = 0; a < 3; ++a) {
                const float D = d(r * 3, a * 3);
                for (int k = 0; k < 3; ++k)
                    for (int kk = 0; kk < 3; ++kk)
                        if (d(r * 3 + k, a * 3 + kk) != (k == kk ? D : 0.0f)) {
                            error_ = "dF_dx is not kron(D, I3)";
                            return;
                        }
                coef_[t * 6 + r * 3 + a] = D;
            }
    }
}

// Closest rotation R (3x2, columns R[0..2], R[3..5]) to F = [x1 - x0,
// x2 - x0] Dm^-1 in F's own plane (src/stretch_reference.cpp:38-68):
// Gram-Schmidt frame B of F's columns with the reference's fallbacks, F2d =
// B^T F, and the rotation maximising tr(R2d^T F2d), (cos, sin) proportional
// to (F2d00 + F2d11, F2d10 - F2d01): the SVD polar factor U V^T when
// det F2d > 0 and the reference's determinant-fixed U diag(1, -1) V^T when it
// is negative. F2d = 0 gives R2d = I, as an SVD of the zero matrix does.
void CpuStretchReferenceSolver::project(const double* x, int t, double R[6]) const
{
    const int* tv = &tris_[(size_t)t * 3];
    auto ld = [&](int v) { return V3{x[3 * (size_t)v], x[3 * (size_t)v + 1], x[3 * (size_t)v + 2]}; };
    const V3 p0 = ld(tv[0]);
    const V3 e1 = sub(ld(tv[1]), p0), e2 = sub(ld(tv[2]), p0);
    const float* g = &dm_inv_[(size_t)t * 4];
    const V3 F0 = add(scl(e1, g[0]), scl(e2, g[1]));
    const V3 F1 = add(scl(e1, g[2]), scl(e2, g[3]));

    V3 b0 = len(F0) > kGuard ? scl(F0, 1.0 / len(F0)) : V3{1.0, 0.0, 0.0};
    const V3 c1 = sub(F1, scl(b0, dot(F1, b0)));
    V3 b1;
    if (len(c1) > kGuard) {
        b1 = scl(c1, 1.0 / len(c1));
    } else {
        V3 f = {-b0.y, b0.x, 0.0};
        if (len(f) < kGuard) f = {0.0, -b0.z, b0.y};
        b1 = scl(f, 1.0 / len(f));
    }
    const double a00 = dot(b0, F0), a01 = dot(b0, F1), a10 = dot(b1, F0), a11 = dot(b1, F1);
    double c = a00 + a11, s = a10 - a01;
    const double n = std::sqrt(c * c + s * s);
    if (n > 0.0) {
        c /= n;
        s /= n;
    } else {
        c = 1.0;
        s = 0.0;
    }
    // R = B R2d, R2d = [[c, -s], [s, c]]
    const V3 r0 = add(scl(b0, c), scl(b1, s));
    const V3 r1 = add(scl(b0, -s), scl(b1, c));
    R[0] = r0.x; R[1] = r0.y; R[2] = r0.z;
    R[3] = r1.x; R[4] = r1.y; R[5] = r1.z;
}

bool CpuStretchReferenceSolver::factor_for(const std::vector<char>& pinned)
{
    const int N = num_verts_;
    const double h2 = (double)dt_ * (double)dt_;
    std::vector<std::vector<int>> adj((size_t)N);
    for (int t = 0; t < num_tris_; ++t) {
        if (kA_[(size_t)t] == 0.0) continue;
        const int* tv = &tris_[(size_t)t * 3];
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b)
                if (a != b) adj[(size_t)tv[a]].push_back(tv[b]);
    }
    std::vector<char> free_v((size_t)N);
    for (int v = 0; v < N; ++v) free_v[(size_t)v] = pinned[(size_t)v] ? 0 : 1;
    order_ = rcm_order(adj, free_v);
    const int n = (int)order_.size();
    slot_.assign((size_t)N, -1);
    for (int i = 0; i < n; ++i) slot_[(size_t)order_[(size_t)i]] = i;

    // envelope: row i spans columns [first_[i], i]
    first_.assign((size_t)n, 0);
    for (int i = 0; i < n; ++i) {
        int f = i;
        for (int w : adj[(size_t)order_[(size_t)i]]) {
            const int j = slot_[(size_t)w];
            if (j >= 0) f = std::min(f, j);
        }
        first_[(size_t)i] = f;
    }
    row_at_.assign((size_t)n + 1, 0);
    for (int i = 0; i < n; ++i) row_at_[(size_t)i + 1] = row_at_[(size_t)i] + (size_t)(i - first_[(size_t)i] + 1);
    factor_.assign(row_at_[(size_t)n], 0.0);
    auto at = [&](int i, int j) -> double& { return factor_[row_at_[(size_t)i] + (size_t)(j - first_[(size_t)i])]; };

    // K_uu (lower triangle): masses, then h^2 kA sum_r D[r][a] D[r][b]
    for (int i = 0; i < n; ++i) at(i, i) = masses_[(size_t)order_[(size_t)i]];
    for (int t = 0; t < num_tris_; ++t) {
        const double kA = kA_[(size_t)t];
        if (kA == 0.0) continue;
        const int* tv = &tris_[(size_t)t * 3];
        const double* D = &coef_[(size_t)t * 6];
        for (int a = 0; a < 3; ++a) {
            const int i = slot_[(size_t)tv[a]];
            if (i < 0) continue;
            for (int b = 0; b < 3; ++b) {
                const int j = slot_[(size_t)tv[b]];
                if (j < 0 || j > i) continue;
                at(i, j) += h2 * kA * (D[a] * D[b] + D[3 + a] * D[3 + b]);
            }
        }
    }

    // Cholesky in place over the envelope, row by row: L_ij for j < i, then L_ii
    for (int i = 0; i < n; ++i) {
        const int fi = first_[(size_t)i];
        double* Li = &factor_[row_at_[(size_t)i]];
        for (int j = fi; j < i; ++j) {
            const int fj = first_[(size_t)j];
            const double* Lj = &factor_[row_at_[(size_t)j]];
            const int k0 = std::max(fi, fj);
            double s = Li[j - fi];
            for (int k = k0; k < j; ++k) s -= Li[k - fi] * Lj[k - fj];
            Li[j - fi] = s / Lj[j - fj];
        }
        double d = Li[i - fi];
        for (int k = fi; k < i; ++k) d -= Li[k - fi] * Li[k - fi];
        if (!(d > 0.0)) return false;
        Li[i - fi] = std::sqrt(d);
    }
    pinned_ = pinned;
    factored_ = true;
    ++factorizations_;
    return true;
}

void CpuStretchReferenceSolver::step(ClothMesh& mesh, const Constraints& pin_cons)
{
    if (!error_.empty()) {
        std::fprintf(stderr, "CPU solver: %s\n", error_.c_str());
        return;
    }
    const int N = num_verts_;
    if ((int)mesh.pos_cpu.size() != N || (int)mesh.vel_cpu.size() != N) {
        std::fprintf(stderr, "CPU solver: the mesh's host state does not have %d vertices\n", N);
        return;
    }
    if ((int)pin_cons.target_positions.size() < 3 * N && !pin_cons.pinned_indices.empty()) {
        std::fprintf(stderr, "CPU solver: pin targets do not cover the mesh\n");
        return;
    }
    std::vector<char> pinned((size_t)N, 0);
    for (int v : pin_cons.pinned_indices) {
        if (v < 0 || v >= N) {
            std::fprintf(stderr, "CPU solver: pinned index %d out of range\n", v);
            return;
        }
        pinned[(size_t)v] = 1;
    }
    if (!factored_ || pinned != pinned_ || refactor_every_iteration) {
        if (!factor_for(pinned)) {
            std::fprintf(stderr, "CPU solver: factorization failed\n");
            factored_ = false;
            return;
        }
    }

    const double h = dt_, h2 = h * h;
    const size_t n3 = (size_t)N * 3;
    std::vector<double> x_old(n3), y(n3), x(n3);
    for (int v = 0; v < N; ++v)
        for (int k = 0; k < 3; ++k) {
            const double p = mesh.pos_cpu[(size_t)v](k), u = mesh.vel_cpu[(size_t)v](k);
            x_old[(size_t)v * 3 + k] = p;
            y[(size_t)v * 3 + k] = p + h * u + (k == 1 ? h2 * (double)gravity_ : 0.0);
        }
    x = x_old;

    const int n = (int)order_.size();
    std::vector<double> rot((size_t)num_tris_ * 6), rhs(n3), b((size_t)n);
    for (int iter = 0; iter < kIterations; ++iter) {
        if (iter > 0 && refactor_every_iteration && !factor_for(pinned)) {
            std::fprintf(stderr, "CPU solver: factorization failed at iter %d\n", iter);
            return;
        }
        // local step (per triangle, independent)
#pragma omp parallel for schedule(static)
        for (int t = 0; t < num_tris_; ++t)
            if (kA_[(size_t)t] != 0.0) project(x.data(), t, &rot[(size_t)t * 6]);
        // right-hand side M y + h^2 sum kA D^T R, in triangle order
        for (size_t i = 0; i < n3; ++i) rhs[i] = (double)masses_[i / 3] * y[i];
        for (int t = 0; t < num_tris_; ++t) {
            const double kA = kA_[(size_t)t];
            if (kA == 0.0) continue;
            const int* tv = &tris_[(size_t)t * 3];
            const double* D = &coef_[(size_t)t * 6];
            const double* R = &rot[(size_t)t * 6];
            for (int a = 0; a < 3; ++a)
                for (int k = 0; k < 3; ++k)
                    rhs[(size_t)tv[a] * 3 + k] += h2 * kA * (D[a] * R[k] + D[3 + a] * R[3 + k]);
        }
        // pinned degrees of freedom take their targets; the others solve
        // K_uu x_u = rhs_u (pinned columns dropped, as the reference zeroes them)
        for (int v : pin_cons.pinned_indices)
            for (int k = 0; k < 3; ++k)
                x[(size_t)v * 3 + k] = pin_cons.target_positions[(size_t)v * 3 + k];
        for (int k = 0; k < 3; ++k) {
            for (int i = 0; i < n; ++i) b[(size_t)i] = rhs[(size_t)order_[(size_t)i] * 3 + k];
            // L z = b
            for (int i = 0; i < n; ++i) {
                const int fi = first_[(size_t)i];
                const double* Li = &factor_[row_at_[(size_t)i]];
                double s = b[(size_t)i];
                for (int j = fi; j < i; ++j) s -= Li[j - fi] * b[(size_t)j];
                b[(size_t)i] = s / Li[i - fi];
            }
            // L^T x = z
            for (int i = n - 1; i >= 0; --i) {
                const int fi = first_[(size_t)i];
                const double* Li = &factor_[row_at_[(size_t)i]];
                const double xi = b[(size_t)i] / Li[i - fi];
                b[(size_t)i] = xi;
                for (int j = fi; j < i; ++j) b[(size_t)j] -= Li[j - fi] * xi;
            }
            for (int i = 0; i < n; ++i) x[(size_t)order_[(size_t)i] * 3 + k] = b[(size_t)i];
        }
    }

    const double scale = (1.0 - (double)damping_) / h;
    for (int v = 0; v < N; ++v)
        for (int k = 0; k < 3; ++k) {
            const size_t i = (size_t)v * 3 + k;
            mesh.vel_cpu[(size_t)v](k) = (float)((x[i] - x_old[i]) * scale);
            mesh.pos_cpu[(size_t)v](k) = (float)x[i];
        }
}
