#pragma once
// Device-side element math shared by the solver kernels.
//
// Two stretch projections:
//  * ref_cuda_rotation reproduces the reference kernel
//    tri_stretch_project_kernel (src/pd_solver.cu:48-208) branch for branch:
//    Gram-Schmidt frame of F, eigen-decomposition of F2d^T F2d with the same
//    1e-10 guards, and the 2-D rotation it assembles as V W^T
//    (pd_solver.cu:184-194), which is the transpose of the polar factor.
//  * polar_rotation is the closest rotation to F on the Stiefel manifold:
//    same Gram-Schmidt frame, then the 2-D rotation maximising tr(R^T F2d),
//    (cos, sin) proportional to (F2d00 + F2d11, F2d10 - F2d01). For
//    det(F2d) > 0 this is the polar factor W V^T that the reference's CPU
//    path takes from an SVD (src/stretch_reference.cpp:38-68); for det < 0
//    it is the sign-corrected factor of the same code (lines 61-65).
#include <hip/hip_runtime.h>

#include "clh_internal.h"

namespace clh {

struct f3 {
    float x, y, z;
};
__device__ __forceinline__ f3 mk(float x, float y, float z) { return f3{x, y, z}; }
__device__ __forceinline__ f3 operator+(f3 a, f3 b) { return f3{a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ f3 operator-(f3 a, f3 b) { return f3{a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ f3 operator*(f3 a, float s) { return f3{a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ float dot(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ f3 ld3(const float* __restrict__ p, size_t i)
{
    return f3{p[i * 3 + 0], p[i * 3 + 1], p[i * 3 + 2]};
}
__device__ __forceinline__ void st3(float* __restrict__ p, size_t i, f3 v)
{
    p[i * 3 + 0] = v.x;
    p[i * 3 + 1] = v.y;
    p[i * 3 + 2] = v.z;
}

// ---------------------------------------------------------------------------
// Reference-mode arithmetic. It is evaluated without FMA contraction,
// operation for operation as written, so it rounds like the FP32
// restatement in the oracle: the reference dynamics are sensitive to the
// last bit (after one iteration its update shrinks interior triangles to
// ~1e-4 m, where the 1e-10 branch tests of the rotation flip on a single
// rounding). Every multiply-add of the reference path goes through these
// helpers; the f3 operators above would let inlined multiply/add pairs fuse.
// ---------------------------------------------------------------------------

__device__ __forceinline__ float dot_nc(f3 a, f3 b)
{
#pragma clang fp contract(off)
    return a.x * b.x + a.y * b.y + a.z * b.z;
}

// F = [d0 d1] G for G = [[g00, g01], [g10, g11]].
__device__ __forceinline__ void deformation_nc(f3 d0, f3 d1, float g00, float g10, float g01,
                                               float g11, f3& F0, f3& F1)
{
#pragma clang fp contract(off)
    F0 = f3{d0.x * g00 + d1.x * g10, d0.y * g00 + d1.y * g10, d0.z * g00 + d1.z * g10};
    F1 = f3{d0.x * g01 + d1.x * g11, d0.y * g01 + d1.y * g11, d0.z * g01 + d1.z * g11};
}

// Corner forces wA D^T [A0 A1] of the three corners (src/pd_solver.cu:279-293).
__device__ __forceinline__ void corner_forces_nc(float g00, float g10, float g01, float g11,
                                                 float wA, f3 A0, f3 A1, f3 c[3])
{
#pragma clang fp contract(off)
    const float s0 = g00 + g10, s1 = g01 + g11;
    c[0] = f3{-(s0 * A0.x + s1 * A1.x) * wA, -(s0 * A0.y + s1 * A1.y) * wA,
              -(s0 * A0.z + s1 * A1.z) * wA};
    c[1] = f3{(g00 * A0.x + g01 * A1.x) * wA, (g00 * A0.y + g01 * A1.y) * wA,
              (g00 * A0.z + g01 * A1.z) * wA};
    c[2] = f3{(g10 * A0.x + g11 * A1.x) * wA, (g10 * A0.y + g11 * A1.y) * wA,
              (g10 * A0.z + g11 * A1.z) * wA};
}

// (x + dt * u) + h2a, component-wise: the inertial prediction
// (src/pd_solver.cu:26-29).
__device__ __forceinline__ f3 predict_nc(f3 x, f3 u, float dt, float a0, float a1, float a2)
{
#pragma clang fp contract(off)
    return f3{x.x + dt * u.x + a0, x.y + dt * u.y + a1, x.z + dt * u.z + a2};
}

// acc + w * p, component-wise.
__device__ __forceinline__ f3 madd_nc(f3 acc, float w, f3 p)
{
#pragma clang fp contract(off)
    return f3{acc.x + w * p.x, acc.y + w * p.y, acc.z + w * p.z};
}

// Hinge projection p = e * n_rest / |e| or 0 (src/pd_solver.cu:315-332).
__device__ __forceinline__ f3 hinge_projection_nc(f3 e, float nr)
{
#pragma clang fp contract(off)
    const float len = sqrtf(e.x * e.x + e.y * e.y + e.z * e.z);
    if (len > 1e-10f && nr > 1e-10f) {
        const float sc = nr / len;
        return f3{e.x * sc, e.y * sc, e.z * sc};
    }
    return f3{0.0f, 0.0f, 0.0f};
}

// Gram-Schmidt frame (u1, u2) of the columns of F with the reference GPU
// kernel's fallbacks (src/pd_solver.cu:79-107).
__device__ __forceinline__ void frame_ref_cuda(f3 F0, f3 F1, f3& u1, f3& u2)
{
#pragma clang fp contract(off)
    const float l0 = sqrtf(dot_nc(F0, F0));
    u1 = (l0 > 1e-10f) ? mk(F0.x / l0, F0.y / l0, F0.z / l0) : mk(1.0f, 0.0f, 0.0f);
    const float a = dot_nc(F1, u1);
    const f3 c = mk(F1.x - a * u1.x, F1.y - a * u1.y, F1.z - a * u1.z);
    const float lc = sqrtf(dot_nc(c, c));
    if (lc > 1e-10f) {
        u2 = mk(c.x / lc, c.y / lc, c.z / lc);
    } else {
        u2 = (fabsf(u1.x) > 0.9f) ? mk(-u1.y, u1.x, 0.0f) : mk(0.0f, -u1.z, u1.y);
        const float l2 = sqrtf(dot_nc(u2, u2));
        u2 = mk(u2.x / l2, u2.y / l2, u2.z / l2);
    }
}

// Rotation (R0, R1) = columns of U * R2d as the reference GPU kernel forms it.
__device__ __forceinline__ void ref_cuda_rotation(f3 F0, f3 F1, f3& R0, f3& R1)
{
#pragma clang fp contract(off)
    f3 u1, u2;
    frame_ref_cuda(F0, F1, u1, u2);
    const float a00 = dot_nc(u1, F0), a01 = dot_nc(u1, F1);
    const float a10 = dot_nc(u2, F0), a11 = dot_nc(u2, F1);

    // eigen-decomposition of A = F2d^T F2d
    const float A00 = a00 * a00 + a10 * a10;
    const float A01 = a00 * a01 + a10 * a11;
    const float A11 = a01 * a01 + a11 * a11;
    const float tr = (A00 + A11) * 0.5f;
    const float disc = sqrtf(fmaxf(0.0f, (A00 - A11) * (A00 - A11) * 0.25f + A01 * A01));
    const float lam1 = tr + disc, lam2 = tr - disc;

    float v1x = 1.0f, v1y = 0.0f, v2x = 0.0f, v2y = 1.0f;
    if (!(disc < 1e-10f || fabsf(A01) < 1e-10f)) {
        const float q = lam1 - A00;
        const float len = sqrtf(A01 * A01 + q * q);
        if (len > 1e-10f) {
            v1x = A01 / len;
            v1y = q / len;
        } else {
            v1x = 1.0f;
            v1y = 0.0f;
        }
        v2x = -v1y;
        v2y = v1x;
    }

    const float s1 = sqrtf(fmaxf(0.0f, lam1));
    const float s2 = sqrtf(fmaxf(0.0f, lam2));
    float w1x, w1y, w2x, w2y;
    if (s1 > 1e-10f) {
        w1x = (a00 * v1x + a01 * v1y) / s1;
        w1y = (a10 * v1x + a11 * v1y) / s1;
    } else {
        w1x = 1.0f;
        w1y = 0.0f;
    }
    if (s2 > 1e-10f) {
        w2x = (a00 * v2x + a01 * v2y) / s2;
        w2y = (a10 * v2x + a11 * v2y) / s2;
    } else {
        const float fx = a00 * v2x + a01 * v2y;
        const float fy = a10 * v2x + a11 * v2y;
        const float d = w1x * fx + w1y * fy;
        const float gx = fx - d * w1x, gy = fy - d * w1y;
        const float lg = sqrtf(gx * gx + gy * gy);
        if (lg > 1e-10f) {
            w2x = gx / lg;
            w2y = gy / lg;
        } else {
            w2x = -w1y;
            w2y = w1x;
        }
    }
    // R2d[i][j] = V[i,:] . W[j,:]
    const float r00 = v1x * w1x + v2x * w2x;
    const float r01 = v1x * w1y + v2x * w2y;
    const float r10 = v1y * w1x + v2y * w2x;
    const float r11 = v1y * w1y + v2y * w2y;
    R0 = mk(u1.x * r00 + u2.x * r10, u1.y * r00 + u2.y * r10, u1.z * r00 + u2.z * r10);
    R1 = mk(u1.x * r01 + u2.x * r11, u1.y * r01 + u2.y * r11, u1.z * r01 + u2.z * r11);
}

// ---------------------------------------------------------------------------
// PD-mode element arithmetic in double. Storage stays FP32; the element math
// (edge vectors, F, the polar rotation, R - F, the hinge curvature) runs in
// double because R - F and sum w_i x_i cancel to strain-sized values (1e-4
// of their O(1) operands), which FP32 resolves only to ~1e-3, and that
// rounding noise, amplified by the Chebyshev extrapolation, holds the
// relative residual near 1e-5 at 64x64 and above it at 256x256.
// ---------------------------------------------------------------------------

struct d3 {
    double x, y, z;
};
__device__ __forceinline__ d3 operator+(d3 a, d3 b) { return d3{a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ d3 operator-(d3 a, d3 b) { return d3{a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ d3 operator*(d3 a, double s) { return d3{a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ double dot(d3 a, d3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ d3 to_d(f3 a) { return d3{(double)a.x, (double)a.y, (double)a.z}; }
__device__ __forceinline__ f3 to_f(d3 a) { return f3{(float)a.x, (float)a.y, (float)a.z}; }
__device__ __forceinline__ d3 ld3d(const double* __restrict__ p, size_t i)
{
    return d3{p[i * 3 + 0], p[i * 3 + 1], p[i * 3 + 2]};
}
__device__ __forceinline__ void st3d(double* __restrict__ p, size_t i, d3 v)
{
    p[i * 3 + 0] = v.x;
    p[i * 3 + 1] = v.y;
    p[i * 3 + 2] = v.z;
}

// 1/sqrt(q) for q in the normal FP32 range: FP32 hardware estimate refined
// by one Newton step in double (relative error ~1e-14).
__device__ __forceinline__ double rsqrt_d(double q)
{
    const double y = (double)__builtin_amdgcn_rsqf((float)q);
    return y * (1.5 - 0.5 * q * y * y);
}

// Closest rotation on the Stiefel manifold, in double: the same frame and
// 2-D rotation as polar_rotation below.
__device__ __forceinline__ void polar_rotation_d(d3 F0, d3 F1, d3& R0, d3& R1)
{
    const double q0 = dot(F0, F0);
    const d3 u1 = (q0 > 1e-20) ? F0 * rsqrt_d(q0) : d3{1.0, 0.0, 0.0};
    const double a = dot(F1, u1);
    const d3 c = F1 - u1 * a;
    const double qc = dot(c, c);
    d3 u2;
    if (qc > 1e-20) {
        u2 = c * rsqrt_d(qc);
    } else {
        u2 = d3{-u1.y, u1.x, 0.0};
        if (dot(u2, u2) < 1e-20) u2 = d3{0.0, -u1.z, u1.y};
        u2 = u2 * rsqrt_d(dot(u2, u2));
    }
    const double a00 = dot(u1, F0), a01 = dot(u1, F1);
    const double a10 = dot(u2, F0), a11 = dot(u2, F1);
    double cs = a00 + a11, sn = a10 - a01;
    const double qn = cs * cs + sn * sn;
    if (qn > 1e-30) {
        const double r = rsqrt_d(qn);
        cs *= r;
        sn *= r;
    } else {
        cs = 1.0;
        sn = 0.0;
    }
    R0 = u1 * cs + u2 * sn;
    R1 = u2 * cs - u1 * sn;
}

// PD triangle: corner forces of R (fr) and of R - F (fd) from the corner
// positions X = x_n + (x - x_n), all in double, rounded to FP32 at the end.
// The edges X_j - X_0 are formed from the FP64 sums, so their rounding is
// 1e-16 of the cloth's coordinates, far below the FP32 rounding of the
// stored results.
__device__ __forceinline__ void pd_triangle(d3 X0, d3 X1, d3 X2, float4 G, float wAf, f3 fr[3],
                                            f3 fd[3], bool want_fr)
{
    const d3 d0 = X1 - X0;
    const d3 d1 = X2 - X0;
    const double g00 = G.x, g10 = G.y, g01 = G.z, g11 = G.w, wA = wAf;
    const d3 F0 = d0 * g00 + d1 * g10;
    const d3 F1 = d0 * g01 + d1 * g11;
    d3 R0, R1;
    polar_rotation_d(F0, F1, R0, R1);
    const double s0 = g00 + g10, s1 = g01 + g11;
    const d3 E0 = R0 - F0, E1 = R1 - F1;
    fd[0] = to_f((E0 * s0 + E1 * s1) * (-wA));
    fd[1] = to_f((E0 * g00 + E1 * g01) * wA);
    fd[2] = to_f((E0 * g10 + E1 * g11) * wA);
    if (want_fr) {
        fr[0] = to_f((R0 * s0 + R1 * s1) * (-wA));
        fr[1] = to_f((R0 * g00 + R1 * g01) * wA);
        fr[2] = to_f((R0 * g10 + R1 * g11) * wA);
    }
}

// PD hinge: e = sum_{i>0} w_i (X_i - X_0) + (sum w_i) X_0 in double (sum w_i
// is zero up to the FP32 rounding of the weights, and is summed in double so
// that no rounding of it enters e); p = e n_rest / |e| (or 0); returns p and
// p - e in FP32.
__device__ __forceinline__ void pd_hinge(const d3 X[4], float4 w, float nr, f3& p, f3& pe)
{
    const double ws[4] = {w.x, w.y, w.z, w.w};
    d3 e = d3{0.0, 0.0, 0.0};
#pragma unroll
    for (int i = 1; i < 4; ++i) e = e + (X[i] - X[0]) * ws[i];
    const double W = ((ws[0] + ws[1]) + ws[2]) + ws[3];
    e = e + X[0] * W;
    const double q = dot(e, e);
    d3 pd = d3{0.0, 0.0, 0.0};
    if (q > 1e-20 && nr > 1e-10f) pd = e * ((double)nr * rsqrt_d(q));
    p = to_f(pd);
    pe = to_f(pd - e);
}

// Corner forces wA D^T [A0 A1] (contraction allowed; PD mode).
__device__ __forceinline__ void corner_forces(float g00, float g10, float g01, float g11,
                                              float wA, f3 A0, f3 A1, f3 c[3])
{
    const float s0 = g00 + g10, s1 = g01 + g11;
    c[0] = f3{-(s0 * A0.x + s1 * A1.x) * wA, -(s0 * A0.y + s1 * A1.y) * wA,
              -(s0 * A0.z + s1 * A1.z) * wA};
    c[1] = f3{(g00 * A0.x + g01 * A1.x) * wA, (g00 * A0.y + g01 * A1.y) * wA,
              (g00 * A0.z + g01 * A1.z) * wA};
    c[2] = f3{(g10 * A0.x + g11 * A1.x) * wA, (g10 * A0.y + g11 * A1.y) * wA,
              (g10 * A0.z + g11 * A1.z) * wA};
}

// Closest rotation on the Stiefel manifold (the frame fallbacks follow the
// CPU path, src/stretch_reference.cpp:43-54).
__device__ __forceinline__ void polar_rotation(f3 F0, f3 F1, f3& R0, f3& R1)
{
    const float l0 = sqrtf(dot(F0, F0));
    const f3 u1 = (l0 > 1e-10f) ? mk(F0.x / l0, F0.y / l0, F0.z / l0) : mk(1.0f, 0.0f, 0.0f);
    const float a = dot(F1, u1);
    const f3 c = mk(F1.x - a * u1.x, F1.y - a * u1.y, F1.z - a * u1.z);
    const float lc = sqrtf(dot(c, c));
    f3 u2;
    if (lc > 1e-10f) {
        u2 = mk(c.x / lc, c.y / lc, c.z / lc);
    } else {
        u2 = mk(-u1.y, u1.x, 0.0f);
        if (sqrtf(dot(u2, u2)) < 1e-10f) u2 = mk(0.0f, -u1.z, u1.y);
        const float l2 = sqrtf(dot(u2, u2));
        u2 = mk(u2.x / l2, u2.y / l2, u2.z / l2);
    }
    const float a00 = dot(u1, F0), a01 = dot(u1, F1);
    const float a10 = dot(u2, F0), a11 = dot(u2, F1);
    float cs = a00 + a11, sn = a10 - a01;
    const float n = sqrtf(cs * cs + sn * sn);
    if (n > 1e-30f) {
        cs = cs / n;
        sn = sn / n;
    } else {
        cs = 1.0f;
        sn = 0.0f;
    }
    // R2d = [[cs, -sn], [sn, cs]]; R = [u1 u2] R2d
    R0 = mk(u1.x * cs + u2.x * sn, u1.y * cs + u2.y * sn, u1.z * cs + u2.z * sn);
    R1 = mk(u2.x * cs - u1.x * sn, u2.y * cs - u1.y * sn, u2.z * cs - u1.z * sn);
}

// ---------------------------------------------------------------------------
// Device-side Chebyshev control (FrameCtl): weights of Wang 2015 with delay
// S = max(1, delay): 1 before S, 2 / (2 - rho^2) at S, 4 / (4 - rho^2 w) after,
// in FP32 as the host computes them.
// ---------------------------------------------------------------------------

__device__ __forceinline__ float ctl_omega_rho(float rho, int delay, int ck, float prev)
{
    const int S = delay > 1 ? delay : 1;
    const float r2 = rho * rho;
    if (ck < S) return 1.0f;
    if (ck == S) return 2.0f / (2.0f - r2);
    return 4.0f / (4.0f - r2 * prev);
}

__device__ __forceinline__ float ctl_omega(const FrameCtl& c, int ck, float prev)
{
    return ctl_omega_rho(c.rho, c.delay, ck, prev);
}

// The state of the iteration after the one in `s`, for a sequence with
// spectral radius estimate rho.
__device__ __forceinline__ CtlSlot ctl_next_rho(float rho, int delay, int use_cheb,
                                                const CtlSlot& s)
{
    CtlSlot n;
    n.ck = s.ck + 1;
    n.omega = ctl_omega_rho(rho, delay, n.ck, s.omega);
    n.blend = (use_cheb && n.ck > 0 && n.omega != 1.0f) ? 1 : 0;
    n.k = s.k + 1;
    return n;
}

// Writes the state of the iteration after the one in `s` into slot par ^ 1
// (block 0 only: the one writer of that slot during a launch; readers use
// `par`).
__device__ __forceinline__ CtlSlot ctl_next(const FrameCtl& c, const CtlSlot& s)
{
    return ctl_next_rho(c.rho, c.delay, c.use_cheb, s);
}

// Stall rule: true when the relative residual (r2 / b2) has not dropped by
// 10 % for `limit` checks in a row (limit > 0): a drop is a value below 0.9
// of the one at the last drop (stall_ref, squared as rel2 is), so a residual
// that falls steadily, however slowly per check, keeps resetting the count.
// best_rel2, the frame's smallest relative residual, is tracked either way
// (the divergence guard compares with it).
__device__ __forceinline__ bool stall_update(double r2, double b2, double& best_rel2,
                                             double& stall_ref, int& since_drop, int limit)
{
    const double rel2 = b2 > 0.0 ? r2 / b2 : r2;
    if (rel2 < best_rel2) best_rel2 = rel2;
    if (rel2 < 0.81 * stall_ref) {
        stall_ref = rel2;
        since_drop = 0;
        return false;
    }
    return limit > 0 && ++since_drop >= limit;
}

// Restart test of a residual check (clh_config.restart_rule):
// CLH_RESTART_PREV, the residual grew since the previous check;
// CLH_RESTART_SEQ_MIN, it exceeds `factor` times the smallest residual of the
// current sequence (since the frame began or the last restart), so the
// oscillation of an early Chebyshev sequence below that factor is left alone;
// CLH_RESTART_NONE, never. r2 values are squared norms.
__device__ __forceinline__ bool ctl_restart_due(int rule, float factor, double r2, double last_r2,
                                                double seq_min_r2)
{
    if (rule == CLH_RESTART_NONE) return false;
    if (rule == CLH_RESTART_SEQ_MIN)
        return seq_min_r2 >= 0.0 && r2 > (double)factor * (double)factor * seq_min_r2;
    return last_r2 >= 0.0 && r2 > last_r2;
}

__device__ __forceinline__ void ctl_write_next(FrameCtl* c, int par, const CtlSlot& s)
{
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    c->slot[par ^ 1] = ctl_next(*c, s);
}

// Reads the state of iteration slot `par`; false when the frame has stopped.
__device__ __forceinline__ bool ctl_begin(FrameCtl* c, int par, float& omega, int& blend)
{
    if (c->stop) return false;
    const CtlSlot s = c->slot[par];
    omega = s.omega;
    blend = s.blend;
    ctl_write_next(c, par, s);
    return true;
}

// After the residual of a check iteration has been summed: history, stop on
// convergence (`conv`: every cloth meets the tolerance) or a non-finite
// residual, and a Chebyshev restart when the restart rule (ctl_restart_due)
// calls for one. Every third restart in a frame doubles 1 - rho (a less aggressive
// sequence, against the oscillation the nonlinear local step can feed; a rho
// below the Jacobi spectral radius is not repaired by it), down to rho = 0,
// where every weight is 1 and the iteration is plain Jacobi. Without the
// floor, rho from 0.999999 turns negative at the 20th doubling (60 restarts);
// the weights depend on rho^2, so they pass 2 only once rho < -1, at the
// 21st doubling (63 restarts), after which the iterate diverges to a
// non-finite residual. The check iteration
// is slot `par`'s iteration plus kofs (1 when a two-iteration launch checked
// its second). One thread.
__device__ __forceinline__ void ctl_check(FrameCtl* c, int par, double r2, double b2, double* hist,
                                          bool conv, int kofs = 0)
{
    const int slot = c->checks;
    hist[2 * slot + 0] = r2;
    hist[2 * slot + 1] = b2;
    c->checks = slot + 1;
    const int k = c->slot[par].k + kofs;
    if (!(r2 == r2) || r2 > 1.7e308) {
        c->nonfinite = 1;
        c->stop = 1;
        c->iters = k + 1;
    } else if (conv) {
        c->converged = 1;
        c->stop = 1;
        c->iters = k + 1;
    } else if (stall_update(r2, b2, c->best_rel2, c->stall_ref, c->since_drop, c->stall_checks)) {
        c->stalled = 1;
        c->stop = 1;
        c->iters = k + 1;
    } else {
        if (c->use_cheb &&
            ctl_restart_due(c->restart_rule, c->restart_factor, r2, c->last_r2, c->seq_min_r2)) {
            CtlSlot n;
            n.ck = 0;
            n.omega = 1.0f;
            n.blend = 0;
            n.k = k + 1;
            c->slot[par ^ 1] = n;
            c->restarts += 1;
            if (c->restarts % 3 == 0) c->rho = fmaxf(0.0f, 1.0f - 2.0f * (1.0f - c->rho));
            c->seq_min_r2 = r2;
        } else if (c->seq_min_r2 < 0.0 || r2 < c->seq_min_r2) {
            c->seq_min_r2 = r2;
        }
        c->last_r2 = r2;
    }
    if (c->mirror) c->mirror[slot & 1] = *c;
}

// The check of one cloth of a batch (per-cloth control): convergence,
// divergence and the stall rule freeze the cloth; otherwise the restart rule
// of ctl_check, on the cloth's own sequence and rho. A cloth diverges when its
// residual is non-finite or its relative residual exceeds 1000 times its best
// of the frame: where bending dominates, some modes of the Jacobi iteration
// have eigenvalues below -1 and grow whatever the Chebyshev weights. k is the
// check iteration, par the control slot of the check launch. One thread.
__device__ __forceinline__ void cloth_check(ClothCtl* q, const FrameCtl* c, int par, int k,
                                            double r2, double b2, bool conv)
{
    q->r2 = r2;
    q->b2 = b2;
    const double rel2 = b2 > 0.0 ? r2 / b2 : r2;
    if (!(r2 == r2) || r2 > 1.7e308 || (q->best_rel2 < 1.0e300 && rel2 > 1.0e6 * q->best_rel2)) {
        q->state = CLH_CLOTH_NONFINITE;
        q->iters = k + 1;
    } else if (conv) {
        q->state = CLH_CLOTH_CONVERGED;
        q->iters = k + 1;
    } else if (stall_update(r2, b2, q->best_rel2, q->stall_ref, q->since_drop, c->stall_checks)) {
        q->state = CLH_CLOTH_STALLED;
        q->iters = k + 1;
    } else {
        if (c->use_cheb &&
            ctl_restart_due(c->restart_rule, c->restart_factor, r2, q->last_r2, q->seq_min_r2)) {
            CtlSlot n;
            n.ck = 0;
            n.omega = 1.0f;
            n.blend = 0;
            n.k = k + 1;
            q->slot[par ^ 1] = n;
            q->restarts += 1;
            if (q->restarts % 3 == 0) q->rho = fmaxf(0.0f, 1.0f - 2.0f * (1.0f - q->rho));
            q->seq_min_r2 = r2;
        } else if (q->seq_min_r2 < 0.0 || r2 < q->seq_min_r2) {
            q->seq_min_r2 = r2;
        }
        q->last_r2 = r2;
    }
}

}  // namespace clh
