// CPU restatement of the reference's projective-dynamics frame. Test
// infrastructure only: the checker of the HIP solver, never part of it.
// See pd_oracle.h for the modes; each routine cites the reference lines it
// restates (Chongyao/cloth-cuda).
#include "pd_oracle.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

template <typename R>
struct V3 {
    R x, y, z;
};
template <typename R> inline V3<R> v3(R x, R y, R z) { return V3<R>{x, y, z}; }
template <typename R> inline V3<R> add(V3<R> a, V3<R> b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
template <typename R> inline V3<R> sub(V3<R> a, V3<R> b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
template <typename R> inline V3<R> scl(V3<R> a, R s) { return v3(a.x * s, a.y * s, a.z * s); }
template <typename R> inline R dot3(V3<R> a, V3<R> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <typename R> inline V3<R> ld(const R* p, size_t i) { return v3(p[3 * i], p[3 * i + 1], p[3 * i + 2]); }
template <typename R> inline void st(R* p, size_t i, V3<R> a) { p[3 * i] = a.x; p[3 * i + 1] = a.y; p[3 * i + 2] = a.z; }

template <typename R> inline R thr() { return (R)1e-10f; }

// ---------------------------------------------------------------------------
// Stretch projections
// ---------------------------------------------------------------------------

// Restates tri_stretch_project_kernel (src/pd_solver.cu:79-204): Gram-Schmidt
// frame (79-107), F2d = U^T F (113-116), eigenvectors of F2d^T F2d with the
// diagonal / isotropic branch (120-149), W = F2d V Sigma^-1 (151-182), and
// R2d = V W^T (184-194), R = U R2d (199-204).
template <typename R>
void ref_cuda_rot(V3<R> F0, V3<R> F1, V3<R>& R0, V3<R>& R1)
{
    const R eps = thr<R>();
    const R lenF0 = std::sqrt(dot3(F0, F0));
    V3<R> u1 = (lenF0 > eps) ? v3(F0.x / lenF0, F0.y / lenF0, F0.z / lenF0) : v3<R>(1, 0, 0);
    const R f1u1 = dot3(F1, u1);
    const V3<R> perp = v3(F1.x - f1u1 * u1.x, F1.y - f1u1 * u1.y, F1.z - f1u1 * u1.z);
    const R lp = std::sqrt(dot3(perp, perp));
    V3<R> u2;
    if (lp > eps) {
        u2 = v3(perp.x / lp, perp.y / lp, perp.z / lp);
    } else {
        if (std::fabs(u1.x) > (R)0.9f) u2 = v3(-u1.y, u1.x, (R)0);
        else                           u2 = v3((R)0, -u1.z, u1.y);
        const R l2 = std::sqrt(dot3(u2, u2));
        u2 = v3(u2.x / l2, u2.y / l2, u2.z / l2);
    }
    const R f00 = dot3(u1, F0), f01 = dot3(u1, F1), f10 = dot3(u2, F0), f11 = dot3(u2, F1);
    const R A00 = f00 * f00 + f10 * f10;
    const R A01 = f00 * f01 + f10 * f11;
    const R A11 = f01 * f01 + f11 * f11;
    const R tr = (A00 + A11) * (R)0.5f;
    const R disc = std::sqrt(std::max((R)0, (A00 - A11) * (A00 - A11) * (R)0.25f + A01 * A01));
    const R l1 = tr + disc, l2 = tr - disc;
    R v1x, v1y, v2x, v2y;
    if (disc < eps || std::fabs(A01) < eps) {
        v1x = 1; v1y = 0; v2x = 0; v2y = 1;
    } else {
        const R tmp = l1 - A00;
        const R len = std::sqrt(A01 * A01 + tmp * tmp);
        if (len > eps) { v1x = A01 / len; v1y = tmp / len; }
        else           { v1x = 1; v1y = 0; }
        v2x = -v1y;
        v2y = v1x;
    }
    const R s1 = std::sqrt(std::max((R)0, l1));
    const R s2 = std::sqrt(std::max((R)0, l2));
    R w1x, w1y, w2x, w2y;
    if (s1 > eps) { w1x = (f00 * v1x + f01 * v1y) / s1; w1y = (f10 * v1x + f11 * v1y) / s1; }
    else          { w1x = 1; w1y = 0; }
    if (s2 > eps) {
        w2x = (f00 * v2x + f01 * v2y) / s2;
        w2y = (f10 * v2x + f11 * v2y) / s2;
    } else {
        const R fx = f00 * v2x + f01 * v2y, fy = f10 * v2x + f11 * v2y;
        const R d = w1x * fx + w1y * fy;
        const R wx = fx - d * w1x, wy = fy - d * w1y;
        const R len = std::sqrt(wx * wx + wy * wy);
        if (len > eps) { w2x = wx / len; w2y = wy / len; }
        else           { w2x = -w1y; w2y = w1x; }
    }
    const R r00 = v1x * w1x + v2x * w2x, r01 = v1x * w1y + v2x * w2y;
    const R r10 = v1y * w1x + v2y * w2x, r11 = v1y * w1y + v2y * w2y;
    R0 = v3(u1.x * r00 + u2.x * r10, u1.y * r00 + u2.y * r10, u1.z * r00 + u2.z * r10);
    R1 = v3(u1.x * r01 + u2.x * r11, u1.y * r01 + u2.y * r11, u1.z * r01 + u2.z * r11);
}

// Restates project_to_manifold (src/stretch_reference.cpp:38-68): frame with
// the CPU fallbacks (43-54), F2d = basis^T F (56), polar factor U V^T of the
// SVD of F2d with the determinant fix (57-65), R = basis R2d (67). The 2x2
// SVD is computed here from the symmetric eigenproblem of F2d^T F2d with a
// Jacobi rotation angle theta = atan2(2 A01, A00 - A11) / 2.
template <typename R>
void polar_rot(V3<R> F0, V3<R> F1, V3<R>& R0, V3<R>& R1)
{
    const R eps = thr<R>();
    const R n0 = std::sqrt(dot3(F0, F0));
    const V3<R> b0 = (n0 > eps) ? v3(F0.x / n0, F0.y / n0, F0.z / n0) : v3<R>(1, 0, 0);
    const R a = dot3(F1, b0);
    const V3<R> c1 = v3(F1.x - a * b0.x, F1.y - a * b0.y, F1.z - a * b0.z);
    const R n1 = std::sqrt(dot3(c1, c1));
    V3<R> b1;
    if (n1 > eps) {
        b1 = v3(c1.x / n1, c1.y / n1, c1.z / n1);
    } else {
        V3<R> fb = v3(-b0.y, b0.x, (R)0);
        if (std::sqrt(dot3(fb, fb)) < eps) fb = v3((R)0, -b0.z, b0.y);
        const R nf = std::sqrt(dot3(fb, fb));
        b1 = v3(fb.x / nf, fb.y / nf, fb.z / nf);
    }
    // F2d = [[m00, m01], [m10, m11]]
    const R m00 = dot3(b0, F0), m01 = dot3(b0, F1), m10 = dot3(b1, F0), m11 = dot3(b1, F1);
    // right singular vectors: eigenvectors of A = F2d^T F2d
    const R A00 = m00 * m00 + m10 * m10, A01 = m00 * m01 + m10 * m11, A11 = m01 * m01 + m11 * m11;
    const R th = (R)0.5 * std::atan2((R)2 * A01, A00 - A11);
    const R ct = std::cos(th), sth = std::sin(th);
    // columns v1 = (ct, sth) (largest eigenvalue), v2 = (-sth, ct)
    R V[2][2] = {{ct, -sth}, {sth, ct}};
    R U[2][2];
    const R l1 = ct * (A00 * ct + A01 * sth) + sth * (A01 * ct + A11 * sth);
    const R l2 = sth * (A00 * sth - A01 * ct) - ct * (A01 * sth - A11 * ct);
    const R s[2] = {std::sqrt(std::max((R)0, l1)), std::sqrt(std::max((R)0, l2))};
    for (int j = 0; j < 2; ++j) {
        const R ux = m00 * V[0][j] + m01 * V[1][j];
        const R uy = m10 * V[0][j] + m11 * V[1][j];
        if (s[j] > eps * std::max(s[0], (R)1)) {
            U[0][j] = ux / s[j];
            U[1][j] = uy / s[j];
        } else if (j == 0) {
            U[0][0] = 1; U[1][0] = 0;
        } else {
            U[0][1] = -U[1][0];
            U[1][1] = U[0][0];
        }
    }
    // re-orthonormalise the second left vector against the first
    {
        const R d = U[0][0] * U[0][1] + U[1][0] * U[1][1];
        R gx = U[0][1] - d * U[0][0], gy = U[1][1] - d * U[1][0];
        const R ng = std::sqrt(gx * gx + gy * gy);
        if (ng > eps) { gx /= ng; gy /= ng; }
        else          { gx = -U[1][0]; gy = U[0][0]; }
        U[0][1] = gx;
        U[1][1] = gy;
    }
    auto uvt = [&](R out[2][2]) {
        for (int i = 0; i < 2; ++i)
            for (int j = 0; j < 2; ++j) out[i][j] = U[i][0] * V[j][0] + U[i][1] * V[j][1];
    };
    R R2[2][2];
    uvt(R2);
    if (R2[0][0] * R2[1][1] - R2[0][1] * R2[1][0] < (R)0) {
        U[0][1] = -U[0][1];
        U[1][1] = -U[1][1];
        uvt(R2);
    }
    R0 = v3(b0.x * R2[0][0] + b1.x * R2[1][0], b0.y * R2[0][0] + b1.y * R2[1][0],
            b0.z * R2[0][0] + b1.z * R2[1][0]);
    R1 = v3(b0.x * R2[0][1] + b1.x * R2[1][1], b0.y * R2[0][1] + b1.y * R2[1][1],
            b0.z * R2[0][1] + b1.z * R2[1][1]);
}

// ---------------------------------------------------------------------------
// Frame
// ---------------------------------------------------------------------------

// vertex -> element slots (element * arity + corner) in ascending order
void incidence(const int* elems, int ne, int arity, int N, std::vector<int>& ptr,
               std::vector<int>& idx)
{
    ptr.assign((size_t)N + 1, 0);
    const size_t ns = (size_t)ne * arity;
    for (size_t s = 0; s < ns; ++s) ptr[(size_t)elems[s] + 1]++;
    for (int v = 0; v < N; ++v) ptr[(size_t)v + 1] += ptr[(size_t)v];
    idx.assign(ns, 0);
    std::vector<int> fill(ptr.begin(), ptr.end() - 1);
    for (size_t s = 0; s < ns; ++s) idx[(size_t)fill[(size_t)elems[s]]++] = (int)s;
}

int nthreads(const or_params* p) { return p->threads > 0 ? p->threads : 1; }

template <typename R>
struct Solver {
    const or_mesh& m;
    const or_params& p;
    int N, T, E, P;
    bool stretch, bend;
    R h2;
    std::vector<int> vt_ptr, vt_idx, ve_ptr, ve_idx, pin_map;
    std::vector<R> fr, fd, bp, bd;  // per-element outputs

    Solver(const or_mesh& mm, const or_params& pp) : m(mm), p(pp)
    {
        N = m.N; T = m.T; E = m.quads ? m.E : 0; P = m.P;
        stretch = T > 0 && m.stretch_k != nullptr;
        bend = E > 0;
        h2 = (R)p.dt * (R)p.dt;
        incidence(m.tris, T, 3, N, vt_ptr, vt_idx);
        if (bend) incidence(m.quads, E, 4, N, ve_ptr, ve_idx);
        pin_map.assign((size_t)N, -1);
        for (int i = 0; i < P; ++i) pin_map[(size_t)m.pin_idx[i]] = i;
        fr.assign((size_t)T * 9, 0);
        fd.assign((size_t)T * 9, 0);
        bp.assign((size_t)E * 3, 0);
        bd.assign((size_t)E * 3, 0);
    }

    V3<R> target(int slot) const
    {
        return v3((R)m.pin_target[3 * slot], (R)m.pin_target[3 * slot + 1],
                  (R)m.pin_target[3 * slot + 2]);
    }

    // Reference mode: corner forces h^2 wA D^T R of every triangle from
    // absolute positions x (src/pd_solver.cu:48-298).
    void triangles_ref(const R* x)
    {
        if (!stretch) return;
#pragma omp parallel for schedule(static) num_threads(nthreads(&p))
        for (int t = 0; t < T; ++t) {
            const int* tv = m.tris + (size_t)t * 3;
            const V3<R> d0 = sub(ld(x, tv[1]), ld(x, tv[0])), d1 = sub(ld(x, tv[2]), ld(x, tv[0]));
            const R g00 = m.Dm_inv[4 * (size_t)t], g10 = m.Dm_inv[4 * (size_t)t + 1];
            const R g01 = m.Dm_inv[4 * (size_t)t + 2], g11 = m.Dm_inv[4 * (size_t)t + 3];
            const V3<R> F0 = v3(d0.x * g00 + d1.x * g10, d0.y * g00 + d1.y * g10, d0.z * g00 + d1.z * g10);
            const V3<R> F1 = v3(d0.x * g01 + d1.x * g11, d0.y * g01 + d1.y * g11, d0.z * g01 + d1.z * g11);
            V3<R> R0, R1;
            ref_cuda_rot(F0, F1, R0, R1);
            const R wA = (R)m.stretch_k[t] * (R)m.rest_area[t] * h2;
            const R s0 = g00 + g10, s1 = g01 + g11;
            R* o = &fr[(size_t)t * 9];
            st(o, 0, v3(-(s0 * R0.x + s1 * R1.x) * wA, -(s0 * R0.y + s1 * R1.y) * wA, -(s0 * R0.z + s1 * R1.z) * wA));
            st(o, 1, v3((g00 * R0.x + g01 * R1.x) * wA, (g00 * R0.y + g01 * R1.y) * wA, (g00 * R0.z + g01 * R1.z) * wA));
            st(o, 2, v3((g10 * R0.x + g11 * R1.x) * wA, (g10 * R0.y + g11 * R1.y) * wA, (g10 * R0.z + g11 * R1.z) * wA));
        }
    }

    // PD mode: positions are base + x with base the frame-start positions
    // (null = 0) in the storage precision R and x the displacement, kept in
    // double. Whatever R is, the element arithmetic (edge vectors, F, the
    // polar rotation, R - F, the corner forces) runs in double and only the
    // forces are stored in R: R - F cancels to the strain (1e-4..1e-3 of its
    // O(1) operands), which single precision resolves only to ~1e-3.
    void triangles_pd(const R* base, const double* x)
    {
        if (!stretch) return;
        using D = V3<double>;
#pragma omp parallel for schedule(static) num_threads(nthreads(&p))
        for (int t = 0; t < T; ++t) {
            const int* tv = m.tris + (size_t)t * 3;
            D d0 = sub(ld(x, tv[1]), ld(x, tv[0])), d1 = sub(ld(x, tv[2]), ld(x, tv[0]));
            if (base) {
                d0 = add(sub(ldd(base, tv[1]), ldd(base, tv[0])), d0);
                d1 = add(sub(ldd(base, tv[2]), ldd(base, tv[0])), d1);
            }
            const double g00 = m.Dm_inv[4 * (size_t)t], g10 = m.Dm_inv[4 * (size_t)t + 1];
            const double g01 = m.Dm_inv[4 * (size_t)t + 2], g11 = m.Dm_inv[4 * (size_t)t + 3];
            const D F0 = v3(d0.x * g00 + d1.x * g10, d0.y * g00 + d1.y * g10, d0.z * g00 + d1.z * g10);
            const D F1 = v3(d0.x * g01 + d1.x * g11, d0.y * g01 + d1.y * g11, d0.z * g01 + d1.z * g11);
            D R0, R1;
            polar_rot(F0, F1, R0, R1);
            const double wA = (double)((R)m.stretch_k[t] * (R)m.rest_area[t] * h2);
            const double s0 = g00 + g10, s1 = g01 + g11;
            auto corners = [&](D A, D B, R* o) {
                st(o, 0, v3((R)(-(s0 * A.x + s1 * B.x) * wA), (R)(-(s0 * A.y + s1 * B.y) * wA), (R)(-(s0 * A.z + s1 * B.z) * wA)));
                st(o, 1, v3((R)((g00 * A.x + g01 * B.x) * wA), (R)((g00 * A.y + g01 * B.y) * wA), (R)((g00 * A.z + g01 * B.z) * wA)));
                st(o, 2, v3((R)((g10 * A.x + g11 * B.x) * wA), (R)((g10 * A.y + g11 * B.y) * wA), (R)((g10 * A.z + g11 * B.z) * wA)));
            };
            corners(R0, R1, &fr[(size_t)t * 9]);
            corners(sub(R0, F0), sub(R1, F1), &fd[(size_t)t * 9]);
        }
    }

    static V3<double> ldd(const R* a, int i) { return v3<double>(a[3 * i], a[3 * i + 1], a[3 * i + 2]); }

    static V3<R> project_hinge(V3<R> acc, R nr)
    {
        const R len = std::sqrt(dot3(acc, acc));
        if (len > thr<R>() && nr > thr<R>()) return scl(acc, nr / len);
        return v3<R>(0, 0, 0);
    }

    // Reference mode hinge projection p = e n_rest / |e| or 0
    // (src/pd_solver.cu:304-335) from absolute positions.
    void hinges_ref(const R* x)
    {
        if (!bend) return;
#pragma omp parallel for schedule(static) num_threads(nthreads(&p))
        for (int e = 0; e < E; ++e) {
            const int* q = m.quads + (size_t)e * 4;
            const float* w = m.bend_w + (size_t)e * 4;
            V3<R> acc = v3<R>(0, 0, 0);
            for (int i = 0; i < 4; ++i) {
                const R wi = w[i];
                const V3<R> pi = ld(x, q[i]);
                acc.x += wi * pi.x;
                acc.y += wi * pi.y;
                acc.z += wi * pi.z;
            }
            st(bp.data(), e, project_hinge(acc, (R)m.bend_n[e]));
        }
    }

    // PD hinge: e = sum_i w_i (base_i + x_i) in double, evaluated as
    // sum_{i>0} w_i ((base_i - base_0) + (x_i - x_0)) + (sum_i w_i)(base_0 + x_0)
    // (sum_i w_i, zero up to the FP32 rounding of the weights, summed in
    // double); p and p - e stored in R.
    void hinges_pd(const R* base, const double* x)
    {
        if (!bend) return;
        using D = V3<double>;
#pragma omp parallel for schedule(static) num_threads(nthreads(&p))
        for (int e = 0; e < E; ++e) {
            const int* q = m.quads + (size_t)e * 4;
            const float* w = m.bend_w + (size_t)e * 4;
            const D b0 = base ? ldd(base, q[0]) : v3<double>(0, 0, 0), x0 = ld(x, q[0]);
            D acc = v3<double>(0, 0, 0);
            for (int i = 1; i < 4; ++i) {
                const D bi = base ? ldd(base, q[i]) : v3<double>(0, 0, 0);
                acc = add(acc, scl(add(sub(bi, b0), sub(ld(x, q[i]), x0)), (double)w[i]));
            }
            const double W = (((double)w[0] + (double)w[1]) + (double)w[2]) + (double)w[3];
            acc = add(acc, scl(add(b0, x0), W));
            const double nr = m.bend_n[e];
            const double len = std::sqrt(dot3(acc, acc));
            D pr = v3<double>(0, 0, 0);
            if (len > 1e-10 && nr > 1e-10) pr = scl(acc, nr / len);
            st(bp.data(), e, v3((R)pr.x, (R)pr.y, (R)pr.z));
            const D dq = sub(pr, acc);
            st(bd.data(), e, v3((R)dq.x, (R)dq.y, (R)dq.z));
        }
    }

    // reference mode vertex update (src/pd_solver.cu:364-398 + pins 414-428)
    void vertex_ref(const R* y, const R* x, R* out)
    {
#pragma omp parallel for schedule(static) num_threads(nthreads(&p))
        for (int v = 0; v < N; ++v) {
            if (pin_map[v] >= 0) { st(out, v, target(pin_map[v])); continue; }
            V3<R> acc = v3<R>(0, 0, 0);
            if (stretch)
                for (int j = vt_ptr[v]; j < vt_ptr[v + 1]; ++j) acc = add(acc, ld(fr.data(), vt_idx[j]));
            if (bend)
                for (int j = ve_ptr[v]; j < ve_ptr[v + 1]; ++j) {
                    const int s = ve_idx[j];
                    const R kw = ((R)m.bend_k[s >> 2] * h2) * (R)m.bend_w[s];
                    const V3<R> q = ld(bp.data(), s >> 2);
                    acc = add(acc, v3(kw * q.x, kw * q.y, kw * q.z));
                }
            const R mv = m.mass[v];
            const V3<R> yv = ld(y, v);
            const V3<R> rhs = v3(acc.x + mv * yv.x, acc.y + mv * yv.y, acc.z + mv * yv.z);
            const R d = m.diag[v];
            const R inv = (d > thr<R>()) ? (R)1 / d : (R)0;
            st(out, v, scl(rhs, inv));
        }
    }

    // PD vertex update on displacements x (double) from the frame-start
    // positions base (null = 0), dy = y - base. The corner and hinge terms
    // are summed in R; then, in double,
    // r = sum h^2 wA D^T (R - F) + sum h^2 k w (p - e) + m (dy - x),
    // x_J = x + gamma r / diag, the Chebyshev blend with x_{k-1}; pinned
    // vertices at target - base. Per-vertex |r|^2 and |b|^2 with
    // b = sum h^2 wA D^T R + sum h^2 k w p + m y.
    void vertex_pd(const R* base, const R* dy, const double* x, const double* xkm1, double* out,
                   double omega, bool blend, std::vector<double>& rr, std::vector<double>& bb,
                   double* r_out = nullptr, double gamma = 1.0)
    {
        using D = V3<double>;
#pragma omp parallel for schedule(static) num_threads(nthreads(&p))
        for (int v = 0; v < N; ++v) {
            rr[v] = 0; bb[v] = 0;
            if (r_out) st(r_out, v, v3<double>(0, 0, 0));
            const D bv = base ? ldd(base, v) : v3<double>(0, 0, 0);
            if (pin_map[v] >= 0) {
                const V3<R> t = target(pin_map[v]);
                st(out, v, sub(v3<double>(t.x, t.y, t.z), bv));
                continue;
            }
            V3<R> ar = v3<R>(0, 0, 0), ab = v3<R>(0, 0, 0);
            if (stretch)
                for (int j = vt_ptr[v]; j < vt_ptr[v + 1]; ++j) {
                    ar = add(ar, ld(fd.data(), vt_idx[j]));
                    ab = add(ab, ld(fr.data(), vt_idx[j]));
                }
            if (bend)
                for (int j = ve_ptr[v]; j < ve_ptr[v + 1]; ++j) {
                    const int s = ve_idx[j];
                    const R kw = ((R)m.bend_k[s >> 2] * h2) * (R)m.bend_w[s];
                    const V3<R> dq = ld(bd.data(), s >> 2), pq = ld(bp.data(), s >> 2);
                    ar = add(ar, v3(kw * dq.x, kw * dq.y, kw * dq.z));
                    ab = add(ab, v3(kw * pq.x, kw * pq.y, kw * pq.z));
                }
            const double mv = m.mass[v];
            const D xv = ld(x, v), yv = ldd(dy, v);
            const D r = v3((double)ar.x + mv * (yv.x - xv.x), (double)ar.y + mv * (yv.y - xv.y),
                           (double)ar.z + mv * (yv.z - xv.z));
            const V3<R> ya = v3((R)(bv.x + yv.x), (R)(bv.y + yv.y), (R)(bv.z + yv.z));
            const R mr = m.mass[v];
            const V3<R> b = v3(ab.x + mr * ya.x, ab.y + mr * ya.y, ab.z + mr * ya.z);
            const double d = m.diag[v];
            const double inv = (d > 1e-10) ? gamma / d : 0.0;
            D xn = v3(xv.x + r.x * inv, xv.y + r.y * inv, xv.z + r.z * inv);
            if (blend) {
                const D xp = ld(xkm1, v);
                xn = v3(omega * (xn.x - xp.x) + xp.x, omega * (xn.y - xp.y) + xp.y, omega * (xn.z - xp.z) + xp.z);
            }
            st(out, v, xn);
            if (r_out) st(r_out, v, r);
            rr[v] = r.x * r.x + r.y * r.y + r.z * r.z;
            bb[v] = (double)b.x * b.x + (double)b.y * b.y + (double)b.z * b.z;
        }
    }
};

// Dense Cholesky K = L L^T in place (lower triangle); returns false if K is
// not positive definite.
template <typename R>
bool cholesky(std::vector<R>& K, int n)
{
    for (int j = 0; j < n; ++j) {
        R d = K[(size_t)j * n + j];
        for (int k = 0; k < j; ++k) d -= K[(size_t)j * n + k] * K[(size_t)j * n + k];
        if (!(d > (R)0)) return false;
        const R l = std::sqrt(d);
        K[(size_t)j * n + j] = l;
        for (int i = j + 1; i < n; ++i) {
            R s = K[(size_t)i * n + j];
            for (int k = 0; k < j; ++k) s -= K[(size_t)i * n + k] * K[(size_t)j * n + k];
            K[(size_t)i * n + j] = s / l;
        }
    }
    return true;
}

template <typename R>
void chol_solve(const std::vector<R>& L, int n, std::vector<R>& b)
{
    for (int i = 0; i < n; ++i) {
        R s = b[i];
        for (int k = 0; k < i; ++k) s -= L[(size_t)i * n + k] * b[k];
        b[i] = s / L[(size_t)i * n + i];
    }
    for (int i = n - 1; i >= 0; --i) {
        R s = b[i];
        for (int k = i + 1; k < n; ++k) s -= L[(size_t)k * n + i] * b[k];
        b[i] = s / L[(size_t)i * n + i];
    }
}

// Chebyshev weight of update k (Wang 2015, delay S): 1 for k < S,
// 2/(2 - rho^2) at S, 4/(4 - rho^2 w) after.
template <typename R>
R cheb_omega(const or_params& p, int k, R prev)
{
    const int S = std::max(1, p.cheb_delay);
    const R r2 = (R)p.rho * (R)p.rho;
    if (k < S) return (R)1;
    if (k == S) return (R)2 / ((R)2 - r2);
    return (R)4 / ((R)4 - r2 * prev);
}

template <typename R>
int step(const or_mesh* mp, const or_params* pp, R* pos, R* vel, const R* init, R* trace,
         double* resid, int* iters_out)
{
    if (!mp || !pp || !pos || !vel) return 1;
    const or_mesh& m = *mp;
    const or_params& p = *pp;
    const int N = m.N;
    Solver<R> S(m, p);
    const R dt = (R)p.dt;
    const R h2 = S.h2;
    // y = x + h v + h^2 a, with gravity on component 1 (src/pd_solver.cu:18-30)
    const R acc[3] = {(R)p.ext_accel[0], (R)p.gravity + (R)p.ext_accel[1], (R)p.ext_accel[2]};
    std::vector<R> xprev(pos, pos + (size_t)N * 3), y((size_t)N * 3);
    for (size_t i = 0; i < (size_t)N; ++i)
        for (int k = 0; k < 3; ++k) y[3 * i + k] = pos[3 * i + k] + dt * vel[3 * i + k] + h2 * acc[k];

    std::vector<R> buf[3];
    for (auto& b : buf) b.assign((size_t)N * 3, 0);
    buf[0] = xprev;
    if (init) buf[0].assign(init, init + (size_t)N * 3);
    int iters = 0;

    if (p.mode == OR_REF_CUDA) {
        for (int k = 0; k < p.iterations; ++k) {
            const R* x = buf[k % 3].data();
            R* out = buf[(k + 1) % 3].data();
            S.triangles_ref(x);
            S.hinges_ref(x);
            S.vertex_ref(y.data(), x, out);
            if (trace) std::memcpy(trace + (size_t)k * N * 3, out, sizeof(R) * N * 3);
            ++iters;
        }
    } else if (p.mode == OR_PD) {
        // iterate on displacements d_k = x_k - x_n, kept in double; dy = y - x_n
        std::vector<R> dy((size_t)N * 3);
        for (size_t i = 0; i < (size_t)N; ++i)
            for (int k = 0; k < 3; ++k) dy[3 * i + k] = dt * vel[3 * i + k] + h2 * acc[k];
        std::vector<double> db[3];
        for (auto& b : db) b.assign((size_t)N * 3, 0.0);
        if (init)
            for (size_t i = 0; i < (size_t)N * 3; ++i) db[0][i] = (double)init[i] - (double)xprev[i];
        for (int i = 0; i < m.P; ++i) {
            const int v = m.pin_idx[i];
            for (int k = 0; k < 3; ++k)
                db[0][3 * (size_t)v + k] = (double)m.pin_target[3 * i + k] - (double)xprev[3 * (size_t)v + k];
        }
        std::vector<double> rr((size_t)N), bb((size_t)N);
        float omf = 1.0f;         // the weights are computed in FP32, as the solver does
        const double tol2 = (double)p.tolerance * (double)p.tolerance;
        const int every = std::max(1, p.check_every);
        int ck = 0;               // iterations since the last Chebyshev (re)start
        int restarts = 0;
        double last_r2 = -1.0;    // residual at the previous check
        double seq_min = -1.0;    // smallest residual of the current sequence
        or_params q = p;          // rho is lowered after every third restart
        for (int k = 0; k < p.iterations; ++k) {
            omf = cheb_omega<float>(q, ck, omf);
            const double omega = omf;
            const bool blend = p.use_chebyshev && ck > 0 && omf != 1.0f;
            const double* x = db[k % 3].data();
            S.triangles_pd(xprev.data(), x);
            S.hinges_pd(xprev.data(), x);
            S.vertex_pd(xprev.data(), dy.data(), x, db[(k + 2) % 3].data(), db[(k + 1) % 3].data(),
                        omega, blend, rr, bb, nullptr, p.relax > 0 ? (double)p.relax : 1.0);
            ++ck;
            double r2 = 0, b2 = 0;
            for (int v = 0; v < N; ++v) { r2 += rr[v]; b2 += bb[v]; }
            if (resid) { resid[2 * k] = r2; resid[2 * k + 1] = b2; }
            if (trace) {
                const double* d = db[(k + 1) % 3].data();
                R* o = trace + (size_t)k * N * 3;
                for (size_t i = 0; i < (size_t)N * 3; ++i) o[i] = (R)((double)xprev[i] + d[i]);
            }
            ++iters;
            if (p.tolerance > 0 && ((k + 1) % every == 0)) {
                if (r2 <= tol2 * b2) break;
                // Chebyshev restart when the rule calls for one (by default
                // when the residual grew since the last test); every third
                // restart doubles 1 - rho, down to rho = 0 (plain Jacobi)
                bool restart = false;
                if (p.restart_rule == OR_RESTART_SEQ_MIN) {
                    const double f = (double)p.restart_factor;
                    restart = seq_min >= 0.0 && r2 > f * f * seq_min;
                } else if (p.restart_rule != OR_RESTART_NONE) {
                    restart = last_r2 >= 0.0 && r2 > last_r2;
                }
                if (p.use_chebyshev && restart) {
                    ck = 0;
                    if (++restarts % 3 == 0) q.rho = std::max(0.0f, 1.0f - 2.0f * (1.0f - q.rho));
                    seq_min = r2;
                } else if (seq_min < 0.0 || r2 < seq_min) {
                    seq_min = r2;
                }
                last_r2 = r2;
            }
        }
        // x_{n+1} = x_n + d, v = d (1 - damping)/h, pins at target with v = 0
        const double* d = db[iters % 3].data();
        const double scale = (1.0 - (double)p.damping) / (double)dt;
        for (int v = 0; v < N; ++v) {
            if (S.pin_map[v] >= 0) {
                st(pos, v, S.target(S.pin_map[v]));
                st(vel, v, v3<R>(0, 0, 0));
                continue;
            }
            for (int k = 0; k < 3; ++k) {
                pos[3 * (size_t)v + k] = (R)((double)xprev[3 * (size_t)v + k] + d[3 * (size_t)v + k]);
                vel[3 * (size_t)v + k] = (R)(d[3 * (size_t)v + k] * scale);
            }
        }
        if (iters_out) *iters_out = iters;
        return 0;
    } else if (p.mode == OR_EXACT) {
        // K = M + h^2 (sum wA D^T D [+ sum k w w^T]) as an N x N scalar matrix
        std::vector<R> K((size_t)N * N, 0);
        for (int v = 0; v < N; ++v) K[(size_t)v * N + v] = m.mass[v];
        std::vector<R> wts((size_t)m.T, 0);
        if (S.stretch)
            for (int t = 0; t < m.T; ++t) {
                const R w = std::sqrt(std::max((R)0, (R)m.stretch_k[t] * (R)m.rest_area[t]));
                wts[t] = w;
                if (w == (R)0) continue;
                const R g00 = m.Dm_inv[4 * (size_t)t], g10 = m.Dm_inv[4 * (size_t)t + 1];
                const R g01 = m.Dm_inv[4 * (size_t)t + 2], g11 = m.Dm_inv[4 * (size_t)t + 3];
                const R D[2][3] = {{-g00 - g10, g00, g10}, {-g01 - g11, g01, g11}};
                for (int a = 0; a < 3; ++a)
                    for (int b = 0; b < 3; ++b) {
                        R s = 0;
                        for (int r = 0; r < 2; ++r) s += (w * D[r][a]) * (w * D[r][b]);
                        K[(size_t)m.tris[3 * t + a] * N + m.tris[3 * t + b]] += h2 * s;
                    }
            }
        const bool ebend = p.exact_bend && S.bend;
        if (ebend)
            for (int e = 0; e < S.E; ++e)
                for (int a = 0; a < 4; ++a)
                    for (int b = 0; b < 4; ++b)
                        K[(size_t)m.quads[4 * e + a] * N + m.quads[4 * e + b]] +=
                            h2 * (R)m.bend_k[e] * (R)m.bend_w[4 * e + a] * (R)m.bend_w[4 * e + b];
        std::vector<char> pinned((size_t)N, 0);
        for (int i = 0; i < m.P; ++i) pinned[(size_t)m.pin_idx[i]] = 1;
        // pinned rows and columns zeroed with a unit diagonal
        // (src/stretch_reference.cpp:225-255); the coupling K_fp is kept
        // aside for correct elimination
        std::vector<R> Kfp;
        if (p.exact_pins) Kfp = K;
        std::vector<R> L = K;
        for (int i = 0; i < N; ++i)
            for (int j = 0; j < N; ++j)
                if (pinned[i] || pinned[j]) L[(size_t)i * N + j] = (i == j) ? (R)1 : (R)0;
        if (!cholesky(L, N)) return 2;
        std::vector<R> x = buf[0], rhs[3];
        for (int k = 0; k < p.iterations; ++k) {
            for (int c = 0; c < 3; ++c) rhs[c].assign((size_t)N, 0);
            for (int v = 0; v < N; ++v)
                for (int c = 0; c < 3; ++c) rhs[c][v] += (R)m.mass[v] * y[3 * (size_t)v + c];
            if (S.stretch)
                for (int t = 0; t < m.T; ++t) {
                    const R w = wts[t];
                    if (w == (R)0) continue;
                    const int* tv = m.tris + 3 * (size_t)t;
                    const V3<R> p0 = ld(x.data(), tv[0]), p1 = ld(x.data(), tv[1]), p2 = ld(x.data(), tv[2]);
                    const V3<R> d0 = sub(p1, p0), d1 = sub(p2, p0);
                    const R g00 = m.Dm_inv[4 * (size_t)t], g10 = m.Dm_inv[4 * (size_t)t + 1];
                    const R g01 = m.Dm_inv[4 * (size_t)t + 2], g11 = m.Dm_inv[4 * (size_t)t + 3];
                    const V3<R> F0 = v3(d0.x * g00 + d1.x * g10, d0.y * g00 + d1.y * g10, d0.z * g00 + d1.z * g10);
                    const V3<R> F1 = v3(d0.x * g01 + d1.x * g11, d0.y * g01 + d1.y * g11, d0.z * g01 + d1.z * g11);
                    V3<R> R0, R1;
                    polar_rot(F0, F1, R0, R1);
                    const R D[2][3] = {{-g00 - g10, g00, g10}, {-g01 - g11, g01, g11}};
                    const R Rc[2][3] = {{w * R0.x, w * R0.y, w * R0.z}, {w * R1.x, w * R1.y, w * R1.z}};
                    for (int a = 0; a < 3; ++a)
                        for (int c = 0; c < 3; ++c) {
                            R s = 0;
                            for (int r = 0; r < 2; ++r) s += (w * D[r][a]) * Rc[r][c];
                            rhs[c][tv[a]] += h2 * s;
                        }
                }
            if (ebend) {
                S.hinges_ref(x.data());
                for (int e = 0; e < S.E; ++e)
                    for (int a = 0; a < 4; ++a) {
                        const R kw = h2 * (R)m.bend_k[e] * (R)m.bend_w[4 * e + a];
                        rhs[0][m.quads[4 * e + a]] += kw * S.bp[3 * (size_t)e];
                        rhs[1][m.quads[4 * e + a]] += kw * S.bp[3 * (size_t)e + 1];
                        rhs[2][m.quads[4 * e + a]] += kw * S.bp[3 * (size_t)e + 2];
                    }
            }
            for (int c = 0; c < 3; ++c) {
                if (p.exact_pins)
                    for (int i = 0; i < N; ++i) {
                        if (pinned[i]) continue;
                        for (int j = 0; j < m.P; ++j) {
                            const int pv = m.pin_idx[j];
                            rhs[c][i] -= Kfp[(size_t)i * N + pv] * (R)m.pin_target[3 * j + c];
                        }
                    }
                for (int j = 0; j < m.P; ++j) rhs[c][m.pin_idx[j]] = (R)m.pin_target[3 * j + c];
                chol_solve(L, N, rhs[c]);
            }
            for (int v = 0; v < N; ++v)
                for (int c = 0; c < 3; ++c) x[3 * (size_t)v + c] = rhs[c][v];
            if (trace) std::memcpy(trace + (size_t)k * N * 3, x.data(), sizeof(R) * N * 3);
            ++iters;
        }
        // velocity from the frame-start positions; pins hold via the solve
        // (src/stretch_reference.cpp:274-279)
        const R scale = ((R)1 - (R)p.damping) / dt;
        for (size_t i = 0; i < (size_t)N * 3; ++i) {
            vel[i] = (x[i] - xprev[i]) * scale;
            pos[i] = x[i];
        }
        if (iters_out) *iters_out = iters;
        return 0;
    } else {
        return 3;
    }

    // x_{n+1}, v = (x_{n+1} - x_n)(1 - d)/h, pins (src/pd_solver.cu:400-412, 602-618)
    const R* xf = buf[iters % 3].data();
    const R scale = ((R)1 - (R)p.damping) / dt;
    for (int v = 0; v < N; ++v) {
        if (S.pin_map[v] >= 0) {
            st(pos, v, S.target(S.pin_map[v]));
            st(vel, v, v3<R>(0, 0, 0));
            continue;
        }
        const V3<R> x = ld(xf, v), xp = ld(xprev.data(), v);
        st(pos, v, x);
        st(vel, v, v3((x.x - xp.x) * scale, (x.y - xp.y) * scale, (x.z - xp.z) * scale));
    }
    if (iters_out) *iters_out = iters;
    return 0;
}

}  // namespace

extern "C" void or_params_default(or_params* p)
{
    std::memset(p, 0, sizeof(*p));
    p->mode = OR_REF_CUDA;
    p->iterations = 50;
    p->use_chebyshev = 1;
    p->rho = 0.9f;
    p->cheb_delay = 1;
    p->gravity = -9.8f;
    p->dt = 0.01f;
    p->check_every = 1;
    p->threads = 1;
    p->relax = 1.0f;
}

extern "C" int or_step_f32(const or_mesh* m, const or_params* p, float* pos, float* vel,
                           const float* init, float* trace, double* resid, int* iters_out)
{
    return step<float>(m, p, pos, vel, init, trace, resid, iters_out);
}

extern "C" int or_step_f64(const or_mesh* m, const or_params* p, double* pos, double* vel,
                           const double* init, double* trace, double* resid, int* iters_out)
{
    return step<double>(m, p, pos, vel, init, trace, resid, iters_out);
}

extern "C" void or_ref_cuda_rotation_f32(const float* F, float* R)
{
    V3<float> R0, R1;
    ref_cuda_rot(v3(F[0], F[1], F[2]), v3(F[3], F[4], F[5]), R0, R1);
    R[0] = R0.x; R[1] = R0.y; R[2] = R0.z; R[3] = R1.x; R[4] = R1.y; R[5] = R1.z;
}

extern "C" void or_polar_rotation_f64(const double* F, double* R)
{
    V3<double> R0, R1;
    polar_rot(v3(F[0], F[1], F[2]), v3(F[3], F[4], F[5]), R0, R1);
    R[0] = R0.x; R[1] = R0.y; R[2] = R0.z; R[3] = R1.x; R[4] = R1.y; R[5] = R1.z;
}

extern "C" void or_polar_rotation_f32(const float* F, float* R)
{
    V3<float> R0, R1;
    polar_rot(v3(F[0], F[1], F[2]), v3(F[3], F[4], F[5]), R0, R1);
    R[0] = R0.x; R[1] = R0.y; R[2] = R0.z; R[3] = R1.x; R[4] = R1.y; R[5] = R1.z;
}

// src/cloth_mesh.cpp:148-210 in FP64.
extern "C" void or_rest_state_f64(int N, int T, const float* rest_pos, const int* tris,
                                  float density, double* Dm_inv, double* area, double* mass)
{
    for (int v = 0; v < N; ++v) mass[v] = 0.0;
    for (int t = 0; t < T; ++t) {
        const int i0 = tris[3 * t], i1 = tris[3 * t + 1], i2 = tris[3 * t + 2];
        V3<double> X0 = v3<double>(rest_pos[3 * i0], rest_pos[3 * i0 + 1], rest_pos[3 * i0 + 2]);
        V3<double> X1 = v3<double>(rest_pos[3 * i1], rest_pos[3 * i1 + 1], rest_pos[3 * i1 + 2]);
        V3<double> X2 = v3<double>(rest_pos[3 * i2], rest_pos[3 * i2 + 1], rest_pos[3 * i2 + 2]);
        const V3<double> e1 = sub(X1, X0), e2 = sub(X2, X0);
        const double n1 = std::sqrt(dot3(e1, e1));
        const V3<double> p0 = scl(e1, 1.0 / n1);
        const V3<double> o = sub(e2, scl(p0, dot3(e2, p0)));
        const double no = std::sqrt(dot3(o, o));
        const V3<double> p1 = scl(o, 1.0 / no);
        const double a = dot3(p0, e1), b = dot3(p0, e2), c = dot3(p1, e1), d = dot3(p1, e2);
        const double det = a * d - c * b;
        Dm_inv[4 * t + 0] = d / det;   // (0,0)
        Dm_inv[4 * t + 1] = -c / det;  // (1,0)
        Dm_inv[4 * t + 2] = -b / det;  // (0,1)
        Dm_inv[4 * t + 3] = a / det;   // (1,1)
        area[t] = 0.5 * std::fabs(det);
        const double third = (double)density * area[t] / 3.0;
        mass[i0] += third;
        mass[i1] += third;
        mass[i2] += third;
    }
}

// src/sim_constraints.cpp:71-111 in FP64 (cotangents from Heron areas).
extern "C" void or_bend_weights_f64(int E, const float* rest_pos, const int* quads, double* w,
                                    double* n_rest)
{
    for (int e = 0; e < E; ++e) {
        V3<double> X[4];
        for (int i = 0; i < 4; ++i) {
            const int v = quads[4 * e + i];
            X[i] = v3<double>(rest_pos[3 * v], rest_pos[3 * v + 1], rest_pos[3 * v + 2]);
        }
        auto len = [&](int a, int b) { const V3<double> d = sub(X[b], X[a]); return std::sqrt(dot3(d, d)); };
        const double l01 = len(0, 1), l02 = len(0, 2), l03 = len(0, 3), l12 = len(1, 2), l13 = len(1, 3);
        auto heron = [](double a, double b, double c) {
            const double s = 0.5 * (a + b + c);
            const double q = s * (s - a) * (s - b) * (s - c);
            return q > 0 ? std::sqrt(q) : 0.0;
        };
        const double A0 = heron(l01, l02, l12), A1 = heron(l01, l03, l13);
        const double c02 = (l01 * l01 + l12 * l12 - l02 * l02) / (4 * A0);
        const double c12 = (l01 * l01 + l02 * l02 - l12 * l12) / (4 * A0);
        const double c03 = (l01 * l01 + l13 * l13 - l03 * l03) / (4 * A1);
        const double c13 = (l01 * l01 + l03 * l03 - l13 * l13) / (4 * A1);
        const double ww[4] = {c02 + c03, c12 + c13, -(c02 + c12), -(c03 + c13)};
        V3<double> er = v3<double>(0, 0, 0);
        for (int i = 0; i < 4; ++i) {
            w[4 * e + i] = ww[i];
            er = add(er, scl(X[i], ww[i]));
        }
        n_rest[e] = std::sqrt(dot3(er, er));
    }
}

// src/sim_constraints.cpp:133-177 in FP64.
extern "C" void or_jacobi_diag_f64(const or_mesh* m, float dt, double* diag)
{
    const double h2 = (double)dt * (double)dt;
    for (int v = 0; v < m->N; ++v) diag[v] = m->mass[v];
    if (m->stretch_k)
        for (int t = 0; t < m->T; ++t) {
            const double wA = (double)m->stretch_k[t] * (double)m->rest_area[t];
            const double g00 = m->Dm_inv[4 * t], g10 = m->Dm_inv[4 * t + 1];
            const double g01 = m->Dm_inv[4 * t + 2], g11 = m->Dm_inv[4 * t + 3];
            const double s0 = g00 + g10, s1 = g01 + g11;
            diag[m->tris[3 * t]] += h2 * wA * (s0 * s0 + s1 * s1);
            diag[m->tris[3 * t + 1]] += h2 * wA * (g00 * g00 + g01 * g01);
            diag[m->tris[3 * t + 2]] += h2 * wA * (g10 * g10 + g11 * g11);
        }
    if (m->quads)
        for (int e = 0; e < m->E; ++e)
            for (int i = 0; i < 4; ++i) {
                const double wi = m->bend_w[4 * e + i];
                diag[m->quads[4 * e + i]] += h2 * (double)m->bend_k[e] * wi * wi;
            }
}

extern "C" void or_residual_f64(const or_mesh* m, float dt, const double* x, const double* y,
                                double* r_out, double* r2, double* b2)
{
    or_params p;
    or_params_default(&p);
    p.dt = dt;
    Solver<double> S(*m, p);
    S.triangles_pd(nullptr, x);
    S.hinges_pd(nullptr, x);
    std::vector<double> rr((size_t)m->N), bb((size_t)m->N);
    std::vector<double> out((size_t)m->N * 3);
    S.vertex_pd(nullptr, y, x, x, out.data(), 1.0, false, rr, bb, r_out);
    double sr = 0, sb = 0;
    for (int v = 0; v < m->N; ++v) { sr += rr[v]; sb += bb[v]; }
    *r2 = sr;
    *b2 = sb;
}

// FP32 residual of base + disp in the displacement form the PD iteration
// uses (dy = y - base); r_out [N*3] in float. For precision studies.
extern "C" void or_residual_disp_f32(const or_mesh* m, float dt, const float* base,
                                     const double* disp, const float* dy, double* r_out,
                                     double* r2, double* b2)
{
    or_params p;
    or_params_default(&p);
    p.dt = dt;
    Solver<float> S(*m, p);
    S.triangles_pd(base, disp);
    S.hinges_pd(base, disp);
    std::vector<double> rr((size_t)m->N), bb((size_t)m->N);
    std::vector<double> out((size_t)m->N * 3);
    S.vertex_pd(base, dy, disp, disp, out.data(), 1.0, false, rr, bb, r_out);
    double sr = 0, sb = 0;
    for (int v = 0; v < m->N; ++v) { sr += rr[v]; sb += bb[v]; }
    *r2 = sr;
    *b2 = sb;
}
