// Solver object of the C ABI: plan (per-vertex adjacency), the frame
// pipeline of both modes, hipGraph capture of fixed-length frames, FP64
// residual statistics and the measurement entry points.
//
// Frame pipeline (one HIP stream):
//   prologue     x_prev = x_n, y = x_n + h v + h^2 a, first iterate, pin map
//
// PD mode iterates on displacements d_k = x_k - x_n (x_n kept in x_prev) and
// forms edge vectors as (x_n,j - x_n,i) + (d_j - d_i). The global system is
// stiff (stiffness / mass ~ 1e5 at 256x256), so rounding absolute FP32
// positions (~6e-8 m) alone leaves a relative residual near 1.5e-4 at
// 256x256; with displacements the floor falls by about two orders of
// magnitude while storage and arithmetic stay FP32. Reference mode keeps
// absolute positions, as the reference does.
//   per iteration k:
//     element kernels  triangles -> corner forces, hinges -> bend vectors
//     vertex kernel    deterministic gather over the vertex's corners and
//                      hinges in ascending element order, Jacobi update,
//                      Chebyshev blend, pins, FP64 residual partials
//   epilogue     x_{n+1}, v_{n+1} = (x_{n+1} - x_n)(1 - d)/h, pins
// No atomics: every sum has a fixed order, so a frame is bitwise
// reproducible run to run.
#include "clh_internal.h"
#include "pd_kernels.h"
#include "pd_fused.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

using namespace clh;

namespace {

constexpr int kBlock = 256;
constexpr int kMaxIterations = 1 << 20;
// Iterations per fused PD launch when the configuration leaves it at 0. One:
// on MI355X the two-iteration kernel is slower per iteration at every size
// measured (7.7 against 7.2 us at 256x256, 139 against 101 us at 1024x1024,
// 735 against 599 us for 512 batched cloths), because its ring-1 elements
// and extra element passes cost more than the launch and memory trip saved.
constexpr int kDefaultItersPerLaunch = 1;
// LDS regions of the element results when the configuration leaves it at 0
constexpr int kDefaultElementRegions = 1;

inline unsigned grid_for(size_t n) { return (unsigned)((n + kBlock - 1) / kBlock); }

// ---------------------------------------------------------------------------
// kernels
// ---------------------------------------------------------------------------

// x_prev = x, first iterate, pin map cleared. Reference mode: y = x + h v +
// h^2 a and x0 = x (or the guess). PD mode: y holds dy = h v + h^2 a and x0
// holds the displacement of the guess (0 without one). With cloth_accel (one
// cloth) the acceleration is acc + cloth_accel[0..2], else h2a = h^2 acc.
__global__ void __launch_bounds__(kBlock)
k_prologue(const float* __restrict__ pos, const float* __restrict__ vel,
           const float* __restrict__ guess, const float* __restrict__ cloth_accel,
           float* __restrict__ xprev, float* __restrict__ y, void* __restrict__ x0,
           int* __restrict__ pin_map, int N, float dt, float h2, float acc0, float acc1,
           float acc2, float h2a0, float h2a1, float h2a2, int disp)
{
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= N) return;
    if (cloth_accel) {
        h2a0 = h2 * (acc0 + cloth_accel[0]);
        h2a1 = h2 * (acc1 + cloth_accel[1]);
        h2a2 = h2 * (acc2 + cloth_accel[2]);
    }
    const f3 x = ld3(pos, v);
    const f3 u = ld3(vel, v);
    st3(xprev, v, x);
    const f3 dy = mk(dt * u.x + h2a0, dt * u.y + h2a1, dt * u.z + h2a2);
    if (disp) {
        st3(y, v, dy);
        st3d(static_cast<double*>(x0), v,
             guess ? to_d(ld3(guess, v)) - to_d(x) : d3{0.0, 0.0, 0.0});
    } else {
        st3(y, v, predict_nc(x, u, dt, h2a0, h2a1, h2a2));
        st3(static_cast<float*>(x0), v, guess ? ld3(guess, v) : x);
    }
    pin_map[v] = -1;
}

// pin_map[idx[i]] = i; in PD mode the first displacement of a pinned vertex
// takes it to its target.
__global__ void __launch_bounds__(kBlock)
k_pin_scatter(const int* __restrict__ idx, const float* __restrict__ target,
              const float* __restrict__ xprev, int* __restrict__ pin_map,
              void* __restrict__ x0, int P, int N, int disp)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P) return;
    const int v = idx[i];
    if ((unsigned)v >= (unsigned)N) return;  // out of range: the pin is inert
    pin_map[v] = i;
    if (disp) st3d(static_cast<double*>(x0), v, to_d(ld3(target, i)) - to_d(ld3(xprev, v)));
}

// Per triangle: deformation gradient, rotation, and the three corner forces.
// Reference mode stores h^2 wA D^T R (src/pd_solver.cu:256-298, with the
// rotation of pd_solver.cu:48-208). PD mode stores h^2 wA D^T (R - F) in
// `fd` (residual) and h^2 wA D^T R in `fr` (right-hand side).
// PD mode reads positions as base + displacement x. A triangle with a vertex
// index outside [0, N) is inert: zero corner forces.
template <int kMode>
__global__ void __launch_bounds__(kBlock)
k_triangles(const float* __restrict__ base, const void* __restrict__ xv,
            const int* __restrict__ tris, const float* __restrict__ dminv,
            const float* __restrict__ area, const float* __restrict__ kst, float* __restrict__ fr,
            float* __restrict__ fd, int T, int N, float h2, const FrameCtl* __restrict__ ctl)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= T) return;
    const int i0 = tris[(size_t)t * 3 + 0], i1 = tris[(size_t)t * 3 + 1],
              i2 = tris[(size_t)t * 3 + 2];
    if ((unsigned)i0 >= (unsigned)N || (unsigned)i1 >= (unsigned)N ||
        (unsigned)i2 >= (unsigned)N) {
        for (int j = 0; j < 9; ++j) {
            fr[(size_t)t * 9 + j] = 0.0f;
            if (kMode == CLH_MODE_PD) fd[(size_t)t * 9 + j] = 0.0f;
        }
        return;
    }
    const float g00 = dminv[(size_t)t * 4 + 0], g10 = dminv[(size_t)t * 4 + 1];
    const float g01 = dminv[(size_t)t * 4 + 2], g11 = dminv[(size_t)t * 4 + 3];
    const float wA = kst[t] * area[t] * h2;
    f3 c[3], e[3];
    if (kMode == CLH_MODE_PD) {
        const double* x = static_cast<const double*>(xv);
        pd_triangle(to_d(ld3(base, i0)) + ld3d(x, i0), to_d(ld3(base, i1)) + ld3d(x, i1),
                    to_d(ld3(base, i2)) + ld3d(x, i2), make_float4(g00, g10, g01, g11), wA, c, e,
                    true);
        float* q = fd + (size_t)t * 9;
        st3(q, 0, e[0]);
        st3(q, 1, e[1]);
        st3(q, 2, e[2]);
    } else {
        const float* x = static_cast<const float*>(xv);
        const f3 p0 = ld3(x, i0), p1 = ld3(x, i1), p2 = ld3(x, i2);
        f3 F0, F1, R0, R1;
        deformation_nc(p1 - p0, p2 - p0, g00, g10, g01, g11, F0, F1);
        ref_cuda_rotation(F0, F1, R0, R1);
        corner_forces_nc(g00, g10, g01, g11, wA, R0, R1, c);
    }
    if (ctl && ctl->stop) return;
    float* o = fr + (size_t)t * 9;
    st3(o, 0, c[0]);
    st3(o, 1, c[1]);
    st3(o, 2, c[2]);
}

// Per hinge: curvature vector e = sum w_i x_i and its projection
// p = e * n_rest / |e| (0 when |e| or n_rest <= 1e-10), as in
// src/pd_solver.cu:304-335. Stores p, and in PD mode also p - e, with e
// evaluated from base + displacement as
// sum_{i>0} w_i ((b_i - b_0) + (x_i - x_0)) + (sum_i w_i)(b_0 + x_0).
// A hinge with a vertex index outside [0, N) is inert: zero vectors.
template <int kMode>
__global__ void __launch_bounds__(kBlock)
k_hinges(const float* __restrict__ base, const void* __restrict__ xv,
         const int* __restrict__ quads, const float* __restrict__ w,
         const float* __restrict__ nrest, float* __restrict__ bp, float* __restrict__ bd, int E,
         int N, const FrameCtl* __restrict__ ctl)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= E) return;
    bool inert = false;
#pragma unroll
    for (int i = 0; i < 4; ++i) inert |= (unsigned)quads[(size_t)e * 4 + i] >= (unsigned)N;
    if (inert) {
        st3(bp, e, mk(0.0f, 0.0f, 0.0f));
        if (kMode == CLH_MODE_PD) st3(bd, e, mk(0.0f, 0.0f, 0.0f));
        return;
    }
    if (kMode == CLH_MODE_PD) {
        const double* x = static_cast<const double*>(xv);
        d3 X[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int qi = quads[(size_t)e * 4 + i];
            X[i] = to_d(ld3(base, qi)) + ld3d(x, qi);
        }
        const float4 wv = make_float4(w[(size_t)e * 4], w[(size_t)e * 4 + 1], w[(size_t)e * 4 + 2],
                                      w[(size_t)e * 4 + 3]);
        f3 p, pe;
        pd_hinge(X, wv, nrest[e], p, pe);
        if (ctl && ctl->stop) return;
        st3(bp, e, p);
        st3(bd, e, pe);
        return;
    }
    const float* x = static_cast<const float*>(xv);
    f3 acc = mk(0.0f, 0.0f, 0.0f);
#pragma unroll
    for (int i = 0; i < 4; ++i)
        acc = madd_nc(acc, w[(size_t)e * 4 + i], ld3(x, quads[(size_t)e * 4 + i]));
    st3(bp, e, hinge_projection_nc(acc, nrest[e]));
}

struct VertexArgs {
    const float* base;    // x_n (PD: iterates are FP64 displacements from it)
    const void*  x;       // x_k     (float in reference mode, double in PD mode)
    const void*  xkm1;    // x_{k-1} (Chebyshev)
    void*        out;     // x_{k+1}
    const float* y;       // y (reference mode) or y - x_n (PD)
    const float* mass;
    const float* diag;
    const int*   pin_map;
    const float* pin_target;
    const int*   vt_ptr;  // vertex -> triangle corners (t*3+c), ascending
    const int*   vt_idx;
    const int*   ve_ptr;  // vertex -> hinge slots (e*4+i), ascending
    const int*   ve_idx;
    const float* fr;      // corner forces of R   [T*9]
    const float* fd;      // corner forces of R-F [T*9] (PD)
    const float* bp;      // hinge projections p  [E*3]
    const float* bd;      // hinge p - e          [E*3] (PD)
    const float* bk;      // hinge stiffness      [E]
    const float* bw;      // hinge weights        [E*4]
    double*      partials;  // [gridDim.x * 2] per-block sum |r|^2, |b|^2 (PD, when non-null)
    int   N;
    int   use_stretch;
    int   use_bend;
    float h2;
    float omega;          // Chebyshev weight of this update (PD)
    int   blend;          // apply the Chebyshev blend (PD)
    FrameCtl* ctl;        // when non-null, omega and blend come from ctl->slot[par]
    int   par;
    float relax;          // PD: scale gamma of the Jacobi step (0 means 1)
};

// Reference mode: rhs = sum corner forces + sum k h^2 w p + m y,
// x = rhs / diag (0 when diag <= 1e-10), pins (src/pd_solver.cu:364-398,
// 414-428).
__global__ void __launch_bounds__(kBlock) k_vertex_ref(VertexArgs a)
{
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= a.N) return;
    const int slot = a.pin_map[v];
    float* out = static_cast<float*>(a.out);
    if (slot >= 0) {
        st3(out, v, ld3(a.pin_target, slot));
        return;
    }
    // the fused kernels' order: four interleaved partial sums (entry offset
    // mod 4 in the triangle list and in the hinge list), (P0 + P1) + (P2 + P3)
    f3 P[4] = {mk(0.0f, 0.0f, 0.0f), mk(0.0f, 0.0f, 0.0f), mk(0.0f, 0.0f, 0.0f),
               mk(0.0f, 0.0f, 0.0f)};
    if (a.use_stretch)
        for (int j = a.vt_ptr[v], e = a.vt_ptr[v + 1]; j < e; j += 4)
#pragma unroll
            for (int u = 0; u < 4; ++u)
                if (j + u < e) P[u] = P[u] + ld3(a.fr, a.vt_idx[j + u]);
    if (a.use_bend)
        for (int j = a.ve_ptr[v], e = a.ve_ptr[v + 1]; j < e; j += 4)
#pragma unroll
            for (int u = 0; u < 4; ++u)
                if (j + u < e) {
                    const int s = a.ve_idx[j + u];
                    const float kw = (a.bk[s >> 2] * a.h2) * a.bw[s];
                    P[u] = madd_nc(P[u], kw, ld3(a.bp, s >> 2));
                }
    const f3 acc = (P[0] + P[1]) + (P[2] + P[3]);
    const float m = a.mass[v];
    const f3 rhs = madd_nc(acc, m, ld3(a.y, v));
    const float d = a.diag[v];
    const float inv = (d > 1e-10f) ? (1.0f / d) : 0.0f;
    st3(out, v, rhs * inv);
}

// Block-wide sum of two doubles; thread 0 writes them to out[2*blockIdx.x].
__device__ void block_sum2(double r, double b, double* out)
{
    __shared__ double sr[kBlock / 64], sb[kBlock / 64];
    for (int off = 32; off > 0; off >>= 1) {
        r += __shfl_down(r, off, 64);
        b += __shfl_down(b, off, 64);
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) {
        sr[wave] = r;
        sb[wave] = b;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double tr = 0.0, tb = 0.0;
        for (int i = 0; i < kBlock / 64; ++i) {
            tr += sr[i];
            tb += sb[i];
        }
        out[2 * blockIdx.x + 0] = tr;
        out[2 * blockIdx.x + 1] = tb;
    }
}

// PD mode on displacements: r = m (dy - x) + sum h^2 wA D^T (R - F) +
// sum h^2 k w (p - e), x_J = x + r / diag, x_{k+1} = omega (x_J - x_{k-1}) +
// x_{k-1} when blending, pinned vertices held at target - base.
__global__ void __launch_bounds__(kBlock) k_vertex_pd(VertexArgs a)
{
    if (a.ctl && !ctl_begin(a.ctl, a.par, a.omega, a.blend)) return;
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    const double* x = static_cast<const double*>(a.x);
    double* out = static_cast<double*>(a.out);
    double rr = 0.0, bb = 0.0;
    if (v < a.N) {
        const int slot = a.pin_map[v];
        const f3 bv = ld3(a.base, v);
        if (slot >= 0) {
            st3d(out, v, to_d(ld3(a.pin_target, slot)) - to_d(bv));
        } else {
            // the fused kernels' order of the gather sums (four partials)
            f3 PR[4], PB[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) PR[u] = PB[u] = mk(0.0f, 0.0f, 0.0f);
            if (a.use_stretch)
                for (int j = a.vt_ptr[v], e = a.vt_ptr[v + 1]; j < e; j += 4)
#pragma unroll
                    for (int u = 0; u < 4; ++u)
                        if (j + u < e) {
                            const int s = a.vt_idx[j + u];
                            PR[u] = PR[u] + ld3(a.fd, s);
                            PB[u] = PB[u] + ld3(a.fr, s);
                        }
            if (a.use_bend)
                for (int j = a.ve_ptr[v], e = a.ve_ptr[v + 1]; j < e; j += 4)
#pragma unroll
                    for (int u = 0; u < 4; ++u)
                        if (j + u < e) {
                            const int s = a.ve_idx[j + u];
                            const float kw = (a.bk[s >> 2] * a.h2) * a.bw[s];
                            const f3 d = ld3(a.bd, s >> 2), p = ld3(a.bp, s >> 2);
                            PR[u] = PR[u] + mk(kw * d.x, kw * d.y, kw * d.z);
                            PB[u] = PB[u] + mk(kw * p.x, kw * p.y, kw * p.z);
                        }
            const f3 ar = (PR[0] + PR[1]) + (PR[2] + PR[3]);
            const f3 ab = (PB[0] + PB[1]) + (PB[2] + PB[3]);
            const float m = a.mass[v];
            const f3 yv = ld3(a.y, v);
            const d3 xv = ld3d(x, v);
            const double md = m;
            const d3 r = d3{(double)ar.x + md * ((double)yv.x - xv.x),
                            (double)ar.y + md * ((double)yv.y - xv.y),
                            (double)ar.z + md * ((double)yv.z - xv.z)};
            const float d = a.diag[v];
            const float gamma = a.relax > 0.0f ? a.relax : 1.0f;
            const double inv = (d > 1e-10f) ? (double)gamma / (double)d : 0.0;
            d3 xn = xv + r * inv;
            if (a.blend) {
                const d3 xp = ld3d(static_cast<const double*>(a.xkm1), v);
                xn = (xn - xp) * (double)a.omega + xp;
            }
            st3d(out, v, xn);
            if (a.partials) {
                const f3 ya = bv + yv;
                const f3 b = mk(ab.x + m * ya.x, ab.y + m * ya.y, ab.z + m * ya.z);
                rr = r.x * r.x + r.y * r.y + r.z * r.z;
                bb = (double)b.x * b.x + (double)b.y * b.y + (double)b.z * b.z;
            }
        }
    }
    if (a.partials) block_sum2(rr, bb, a.partials);
}

// Sums the per-block partials in a fixed order into hist[2*slot .. +1].
__global__ void __launch_bounds__(kBlock)
k_reduce_partials(const double* __restrict__ partials, int nblocks, double* __restrict__ hist,
                  int slot, FrameCtl* __restrict__ ctl, int par, int kofs)
{
    if (ctl && ctl->stop) return;
    double r = 0.0, b = 0.0;
    for (int i = threadIdx.x; i < nblocks; i += kBlock) {
        r += partials[2 * i + 0];
        b += partials[2 * i + 1];
    }
    __shared__ double sr[kBlock], sb[kBlock];
    sr[threadIdx.x] = r;
    sb[threadIdx.x] = b;
    __syncthreads();
    for (int s = kBlock / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            sr[threadIdx.x] += sr[threadIdx.x + s];
            sb[threadIdx.x] += sb[threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        if (ctl) {
            ctl_check(ctl, par, sr[0], sb[0], hist, sr[0] <= ctl->tol2 * sb[0], kofs);
        } else {
            hist[2 * slot + 0] = sr[0];
            hist[2 * slot + 1] = sb[0];
        }
    }
}

// Batched cloths: per-cloth sums of the tile partials (cloth c owns tiles
// [cloth_tiles[c], cloth_tiles[c+1])), one wave per cloth with lanes striding
// over its tiles and a fixed-order butterfly, so every sum is reproducible.
// Writes (r^2, b^2) of each cloth to cloth_res, the batch totals to the
// history, and converges only when every cloth meets the tolerance.
// With per-cloth control (cctl), each cloth is checked on its own
// (cloth_check): a cloth that has stopped keeps the residual it stopped at,
// and the frame stops when no cloth runs (converged when every cloth did).
constexpr int kReduceCloths = 1024;
__global__ void __launch_bounds__(kReduceCloths)
k_reduce_cloths(const double* __restrict__ partials, const int* __restrict__ cloth_tiles, int C,
                double* __restrict__ cloth_res, double* __restrict__ hist, int slot,
                FrameCtl* __restrict__ ctl, int par, int kofs, ClothCtl* __restrict__ cctl)
{
    if (ctl && ctl->stop) return;
    constexpr int kWaves = kReduceCloths / 64;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const double tol2 = ctl ? ctl->tol2 : 0.0;
    const int kchk = ctl ? ctl->slot[par].k + kofs : 0;
    double wr = 0.0, wb = 0.0;
    int wconv = 1, wrun = 0, wall = 1, wnan = 0;
    for (int c = wave; c < C; c += kWaves) {
        const int t0 = cloth_tiles[c], t1 = cloth_tiles[c + 1];
        double r = 0.0, b = 0.0;
        const bool active = !cctl || cctl[c].state == CLH_CLOTH_RUNNING;
        if (active)
            for (int t = t0 + lane; t < t1; t += 64) {
                r += partials[2 * t + 0];
                b += partials[2 * t + 1];
            }
        for (int off = 32; off > 0; off >>= 1) {
            r += __shfl_xor(r, off, 64);
            b += __shfl_xor(b, off, 64);
        }
        if (cctl) {
            ClothCtl* q = cctl + c;
            int st;
            if (active) {
                // lane 0 checks the cloth; its new state reaches the other
                // lanes by a shuffle, not through memory
                st = CLH_CLOTH_RUNNING;
                if (lane == 0) {
                    cloth_check(q, ctl, par, kchk, r, b, r <= tol2 * b);
                    st = q->state;
                }
                st = __shfl(st, 0, 64);
            } else {
                r = q->r2;
                b = q->b2;
                st = q->state;
            }
            if (st == CLH_CLOTH_RUNNING) wrun = 1;
            if (st != CLH_CLOTH_CONVERGED) wall = 0;
            if (st == CLH_CLOTH_NONFINITE) {
                // a diverged cloth keeps x_n and stays out of the batch totals
                wnan = 1;
                if (lane == 0) {
                    cloth_res[2 * c + 0] = r;
                    cloth_res[2 * c + 1] = b;
                }
                continue;
            }
        }
        if (lane == 0) {
            cloth_res[2 * c + 0] = r;
            cloth_res[2 * c + 1] = b;
        }
        wr += r;
        wb += b;
        if (!(r <= tol2 * b)) wconv = 0;
    }
    __shared__ double sr[kWaves], sb[kWaves];
    __shared__ int sc[kWaves], srun[kWaves], sall[kWaves], snan[kWaves];
    if (lane == 0) {
        sr[wave] = wr;
        sb[wave] = wb;
        sc[wave] = wconv;
        srun[wave] = wrun;
        sall[wave] = wall;
        snan[wave] = wnan;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double tr = 0.0, tb = 0.0;
        int conv = 1, run = 0, all = 1, nan = 0;
        for (int i = 0; i < kWaves; ++i) {
            tr += sr[i];
            tb += sb[i];
            conv &= sc[i];
            run |= srun[i];
            all &= sall[i];
            nan |= snan[i];
        }
        if (ctl && cctl) {
            // per-cloth control: the cloths decided their own restarts; the
            // frame stops when none runs
            const int hs = ctl->checks;
            hist[2 * hs + 0] = tr;
            hist[2 * hs + 1] = tb;
            ctl->checks = hs + 1;
            (void)nan;
            if (!run) {
                ctl->converged = all;
                ctl->stalled = all ? 0 : 1;
                ctl->stop = 1;
                ctl->iters = kchk + 1;
            }
            if (ctl->mirror) ctl->mirror[hs & 1] = *ctl;
        } else if (ctl) {
            ctl_check(ctl, par, tr, tb, hist, conv != 0, kofs);
        } else {
            hist[2 * slot + 0] = tr;
            hist[2 * slot + 1] = tb;
        }
    }
}

// Starts the per-cloth control of a batch frame run to a tolerance.
__global__ void __launch_bounds__(256)
k_cloth_ctl_init(ClothCtl* __restrict__ q, int C, float rho)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= C) return;
    ClothCtl z;
    z.slot[0].omega = 1.0f;
    z.slot[0].blend = 0;
    z.slot[0].ck = 0;
    z.slot[0].k = 0;
    z.slot[1] = z.slot[0];
    z.rho = rho;
    z.state = CLH_CLOTH_RUNNING;
    z.restarts = 0;
    z.iters = 0;
    z.since_drop = 0;
    z.pad = 0;
    z.last_r2 = -1.0;
    z.best_rel2 = 1.0e300;
    z.stall_ref = 1.0e300;
    z.r2 = 0.0;
    z.b2 = 0.0;
    z.seq_min_r2 = -1.0;
    q[c] = z;
}

// Starts the control state of a frame run to a tolerance.
__global__ void k_ctl_frame_init(FrameCtl* __restrict__ c, float rho, int use_cheb, int delay,
                                 double tol2, FrameCtl* mirror, int stall_checks,
                                 int restart_rule, float restart_factor)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    FrameCtl f;
    f.slot[0].omega = 1.0f;
    f.slot[0].blend = 0;
    f.slot[0].ck = 0;
    f.slot[0].k = 0;
    f.slot[1] = f.slot[0];
    f.rho = rho;
    f.use_cheb = use_cheb;
    f.delay = delay;
    f.stop = 0;
    f.iters = 0;
    f.checks = 0;
    f.restarts = 0;
    f.converged = 0;
    f.nonfinite = 0;
    f.barrier_error = 0;
    f.last_r2 = -1.0;
    f.tol2 = tol2;
    f.stall_checks = stall_checks;
    f.since_drop = 0;
    f.best_rel2 = 1.0e300;
    f.stall_ref = 1.0e300;
    f.stalled = 0;
    f.restart_rule = restart_rule;
    f.restart_factor = restart_factor;
    f.seq_min_r2 = -1.0;
    f.mirror = mirror;
    *c = f;
}

// x_{n+1} and v = (x_{n+1} - x_n)(1 - d)/h; pinned vertices at target with
// zero velocity (src/pd_solver.cu:400-412, 602-618). In PD mode xf holds the
// displacement d, so x_{n+1} = x_n + d and v = d (1 - d)/h.
__global__ void __launch_bounds__(kBlock)
k_epilogue(const void* __restrict__ xf, const float* __restrict__ xprev,
           const int* __restrict__ pin_map, const float* __restrict__ target,
           float* __restrict__ pos, float* __restrict__ vel, int N, float scale, int disp)
{
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= N) return;
    const int slot = pin_map[v];
    if (slot >= 0) {
        st3(pos, v, ld3(target, slot));
        st3(vel, v, mk(0.0f, 0.0f, 0.0f));
        return;
    }
    const f3 xp = ld3(xprev, v);
    if (disp) {
        const d3 d = ld3d(static_cast<const double*>(xf), v);
        st3(pos, v, to_f(to_d(xp) + d));
        st3(vel, v, to_f(d * (double)scale));
    } else {
        const f3 x = ld3(static_cast<const float*>(xf), v);
        st3(pos, v, x);
        st3(vel, v, mk((x.x - xp.x) * scale, (x.y - xp.y) * scale, (x.z - xp.z) * scale));
    }
}

// Absolute positions of an iterate into a trace slot.
__global__ void __launch_bounds__(kBlock)
k_trace(const void* __restrict__ it, const float* __restrict__ xprev, float* __restrict__ out,
        int N, int disp)
{
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= N) return;
    st3(out, v,
        disp ? to_f(to_d(ld3(xprev, v)) + ld3d(static_cast<const double*>(it), v))
             : ld3(static_cast<const float*>(it), v));
}

// Gershgorin bound of the Jacobi-preconditioned global matrix D^-1 K, with
// K = M + h^2 (sum_t k_t A_t G_t^T G_t + sum_e k_e w_e w_e^T) (G_t the 2 x 3
// gradient rows of triangle t from Dm^-1): every eigenvalue of D^-1 K is at
// most max_v (sum_j |K_vj|) / D_v, and each element's block bounds its share
// of the row, |K_vj| <= sum over elements of |K^e_vj|. One thread per vertex
// sums its triangle corners and hinge slots in ascending order, so the row is
// the same on every run; the maximum of non-negative doubles is exact in any
// order, taken with an integer atomic max on their bits. With cloth vertex
// offsets seg[0..C] each cloth gets its own maximum in out[c] (the cloths
// share no element, so a cloth's rows are those of its solo mesh).
__global__ void __launch_bounds__(kBlock)
k_jacobi_bound(const int* __restrict__ vt_ptr, const int* __restrict__ vt_idx,
               const int* __restrict__ ve_ptr, const int* __restrict__ ve_idx,
               const float* __restrict__ mass, const float* __restrict__ diag,
               const float* __restrict__ Dm_inv, const float* __restrict__ rest_area,
               const float* __restrict__ stretch_k, const float* __restrict__ bend_w,
               const float* __restrict__ bend_k, int N, double h2, int use_stretch,
               int use_bend, const int* __restrict__ seg, int C,
               unsigned long long* __restrict__ out)
{
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= N) return;
    const double d = diag[v];
    if (!(d > 1e-10)) return;
    double row = fabs((double)mass[v]);
    if (use_stretch)
        for (int j = vt_ptr[v]; j < vt_ptr[v + 1]; ++j) {
            const int t = vt_idx[j] / 3, c = vt_idx[j] % 3;
            const double g00 = Dm_inv[4 * (size_t)t], g10 = Dm_inv[4 * (size_t)t + 1];
            const double g01 = Dm_inv[4 * (size_t)t + 2], g11 = Dm_inv[4 * (size_t)t + 3];
            const double D0[3] = {-g00 - g10, g00, g10}, D1[3] = {-g01 - g11, g01, g11};
            double sum = 0.0;
            for (int b = 0; b < 3; ++b) sum += fabs(D0[c] * D0[b] + D1[c] * D1[b]);
            row += fabs(h2 * (double)stretch_k[t] * (double)rest_area[t]) * sum;
        }
    if (use_bend)
        for (int j = ve_ptr[v]; j < ve_ptr[v + 1]; ++j) {
            const int e = ve_idx[j] >> 2, i = ve_idx[j] & 3;
            const float* w = bend_w + 4 * (size_t)e;
            const double sw = fabs((double)w[0]) + fabs((double)w[1]) + fabs((double)w[2]) +
                              fabs((double)w[3]);
            row += fabs(h2 * (double)bend_k[e]) * fabs((double)w[i]) * sw;
        }
    const double ratio = row / d;
    int lo = 0;
    if (seg) {  // the cloth c with seg[c] <= v < seg[c + 1]
        int hi = C;
        while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if (seg[mid] <= v) lo = mid;
            else hi = mid;
        }
    }
    atomicMax(out + lo, (unsigned long long)__double_as_longlong(ratio));
}

// ---------------------------------------------------------------------------
// host helpers
// ---------------------------------------------------------------------------

// CSR of vertex -> incident element slots (element * arity + corner), slots in
// ascending order.
void build_incidence(const std::vector<int>& elems, int arity, int N, std::vector<int>& ptr,
                     std::vector<int>& idx)
{
    const size_t ne = elems.size() / (size_t)arity;
    ptr.assign((size_t)N + 1, 0);
    for (size_t s = 0; s < ne * (size_t)arity; ++s) ptr[(size_t)elems[s] + 1]++;
    for (int v = 0; v < N; ++v) ptr[(size_t)v + 1] += ptr[(size_t)v];
    idx.assign(ne * (size_t)arity, 0);
    std::vector<int> fill(ptr.begin(), ptr.end() - 1);
    for (size_t s = 0; s < ne * (size_t)arity; ++s) idx[(size_t)fill[(size_t)elems[s]]++] = (int)s;
}

bool same_args(const clh_step_args& a, const clh_step_args& b)
{
    return std::memcmp(&a, &b, sizeof(a)) == 0;
}

}  // namespace

// ---------------------------------------------------------------------------
// solver object
// ---------------------------------------------------------------------------

struct clh_solver {
    clh_config cfg;
    int N = 0, T = 0, E = 0;
    hipStream_t stream = nullptr;      // stream work is enqueued on
    hipStream_t own_stream = nullptr;  // created by the solver
    bool user_stream = false;
    hipEvent_t ev_in = nullptr, ev_out = nullptr, ev_t0 = nullptr, ev_t1 = nullptr;

    // plan
    const int* plan_tris = nullptr;
    const int* plan_quads = nullptr;
    bool plan_ready = false;
    DevBuf<int> vt_ptr, vt_idx, ve_ptr, ve_idx;
    FusedPlan fused;
    std::vector<int> seg;        // cloth vertex offsets [C+1] (batched solver), else empty
    int num_cloths = 1;
    unsigned long long shape_key = 0;  // autotuner cache key of the mesh shape
    bool persist = false;        // tolerance frames in persistent cooperative launches
    int ipl = 1;                 // iterations per fused launch (2: two-iteration kernel)
    int regions = 1;             // element regions in LDS (2: separate triangle, hinge)
    bool pair_fits = false;      // the plan's two-iteration tables fit in LDS
    DevBuf<unsigned> bar;        // grid barrier words of the persistent kernel
    DevBuf<double> cloth_res;    // per-cloth (r^2, b^2) of the last residual evaluation
    DevBuf<ClothCtl> cctl;       // per-cloth control of batch frames run to a tolerance
    bool last_cloth_ctl = false; // the last frame used it
    int variant = CLH_KERNEL_UNFUSED;
    int tile_verts = 128;
    int block_threads = 256;
    DevBuf<int4> vinfo;  // tile-order vertex records (fused variant)

    // scratch
    DevBuf<float> y, xprev, fr, fd, bp, bd;
    DevBuf<double> it[kIterBufs];  // iterates: float triples (reference) or double (PD)
    DevBuf<int> pin_map;
    DevBuf<double> partials, hist;
    int nblocks_v = 0;

    // pinned host mirror of the residual history
    double* h_hist = nullptr;
    int h_hist_cap = 0;

    // graph of a fixed-length frame
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    clh_step_args graph_args = {};
    bool graph_valid = false;

    // frames run to a tolerance: device control block, its pinned host
    // snapshots (one per chunk in flight), and graphs of one chunk of
    // check_every iterations, by the chunk's first iteration modulo 4 (the
    // iterate buffers rotate by 4; the control slots alternate per launch,
    // every iteration or, with two-iteration launches, every second one)
    DevBuf<FrameCtl> ctl;
    FrameCtl* h_ctl = nullptr;       // [2], pinned and coherent
    FrameCtl* d_ctl_mirror = nullptr;  // device address of h_ctl
    hipEvent_t ev_chunk[2] = {nullptr, nullptr};
    hipGraphExec_t chunk_exec[kIterBufs] = {};
    clh_step_args chunk_args = {};
    int chunk_len = 0;

    // last pin list whose indices were range-checked
    const int* pins_checked_ptr = nullptr;
    int pins_checked_n = -1;

    // Jacobi step scale in use (clh_config.relax resolved; CLH_RELAX_AUTO:
    // from the Gershgorin bound of D^-1 K of the constants in relax_args)
    // (batches with CLH_RELAX_AUTO: each cloth's gamma in cloth_relax, from
    // its own bound; relax_eff and jacobi_bound are then the smallest gamma
    // and the largest bound)
    float relax_eff = 1.0f;
    double jacobi_bound = 0.0;
    bool relax_ready = false;
    clh_step_args relax_args = {};
    DevBuf<float> cloth_relax;
    bool cloth_relax_on = false;

    // statistics of the last frame
    int last_iters = 0;
    int last_checks = 0;
    int last_restarts = 0;
    float last_rho = 0.0f;
    int converged = 0;
    int frames = 0;
};

namespace {

void drop_chunk_graphs(clh_solver* s)
{
    for (auto& g : s->chunk_exec) {
        if (g) (void)hipGraphExecDestroy(g);
        g = nullptr;
    }
}

// FNV-1a over the counts that decide the fused kernel's cost: vertices,
// triangles, hinges, cloths, mode, whether two-iteration launches are
// candidates, and the histogram of (triangle, hinge) valence per vertex.
unsigned long long shape_key(const clh_solver& s, const std::vector<int>& tptr,
                             const std::vector<int>& eptr)
{
    std::vector<long long> words = {s.N, s.T, s.E, s.num_cloths, s.cfg.mode,
                                    s.cfg.iters_per_launch == 2 ? 1 : 0};
    std::vector<long long> hist(32 * 32, 0);
    for (int v = 0; v < s.N; ++v) {
        const int nt = std::min(31, tptr[(size_t)v + 1] - tptr[(size_t)v]);
        const int ne = std::min(31, eptr[(size_t)v + 1] - eptr[(size_t)v]);
        hist[(size_t)nt * 32 + ne]++;
    }
    words.insert(words.end(), hist.begin(), hist.end());
    unsigned long long h = 1469598103934665603ull;
    for (long long w : words)
        for (int b = 0; b < 8; ++b) {
            h ^= (unsigned long long)((w >> (8 * b)) & 0xFF);
            h *= 1099511628211ull;
        }
    return h;
}

int ensure_plan(clh_solver* s, const clh_step_args& a)
{
    if (s->plan_ready && s->plan_tris == a.d_tris && s->plan_quads == a.d_bend_quads)
        return CLH_OK;
    const int N = s->N, T = s->T;
    const int E = a.d_bend_quads ? s->E : 0;
    std::vector<int> tris((size_t)T * 3), quads((size_t)E * 4);
    if (T > 0) {
        if (!a.d_tris) return fail(CLH_ERR_INVALID, "clh_solver_step: null d_tris with num_tris > 0");
        CLH_HIP(hipMemcpy(tris.data(), a.d_tris, tris.size() * sizeof(int), hipMemcpyDeviceToHost));
    }
    if (E > 0)
        CLH_HIP(hipMemcpy(quads.data(), a.d_bend_quads, quads.size() * sizeof(int),
                          hipMemcpyDeviceToHost));
    for (int v : tris)
        if (v < 0 || v >= N) return fail(CLH_ERR_INVALID, "triangle vertex index out of range");
    for (int v : quads)
        if (v < 0 || v >= N) return fail(CLH_ERR_INVALID, "hinge vertex index out of range");

    std::vector<int> ptr, idx, tptr;
    build_incidence(tris, 3, N, ptr, idx);
    CLH_TRY(s->vt_ptr.upload(ptr));
    CLH_TRY(s->vt_idx.upload(idx));
    tptr.swap(ptr);
    build_incidence(quads, 4, N, ptr, idx);
    CLH_TRY(s->ve_ptr.upload(ptr));
    CLH_TRY(s->ve_idx.upload(idx));
    s->shape_key = shape_key(*s, tptr, ptr);

    s->variant = CLH_KERNEL_UNFUSED;
    s->pair_fits = false;
    if (s->cfg.kernel_variant != CLH_KERNEL_UNFUSED && N > 0 && a.d_pos) {
        std::vector<float> pos((size_t)N * 3);
        CLH_HIP(hipMemcpy(pos.data(), a.d_pos, pos.size() * sizeof(float), hipMemcpyDeviceToHost));
        // the tile order comes from these positions: a non-finite one would
        // give tiles of unrelated vertices
        for (float v : pos)
            if (!std::isfinite(v))
                return fail(CLH_ERR_NUMERIC, "clh_solver_step: non-finite positions at the first step");
        int dev = 0, lds_limit = 0, optin = 0;
        CLH_HIP(hipGetDevice(&dev));
        CLH_HIP(hipDeviceGetAttribute(&lds_limit, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
        // larger workgroup allocations by opt-in (gfx950: 160 KB of LDS per CU)
        if (hipDeviceGetAttribute(&optin, hipDeviceAttributeSharedMemPerBlockOptin, dev) == hipSuccess)
            lds_limit = std::max(lds_limit, optin);
        else
            (void)hipGetLastError();
        CLH_TRY(fused_enable_lds(lds_limit));
        int st = CLH_ERR_UNSUPPORTED;
        const bool want_pair = s->cfg.mode == CLH_MODE_PD && s->cfg.iters_per_launch == 2;
        for (int tv = s->tile_verts; tv >= 32; tv /= 2) {
            st = fused_plan_build(s->fused, N, tris, quads, pos, tv, s->seg, want_pair);
            if (st != CLH_OK) break;
            s->fused.threads = s->block_threads;
            if (fused_lds_bytes(s->fused, s->cfg.mode, true) <= (size_t)lds_limit) break;
            st = fail(CLH_ERR_UNSUPPORTED, "fused plan: tile does not fit in LDS");
        }
        // two-iteration launches only where their larger tables fit
        s->pair_fits = st == CLH_OK && s->fused.pair &&
                       fused_pair_lds_bytes(s->fused, true) <= (size_t)lds_limit;
        s->fused.lds_limit = lds_limit;
        s->fused.split = s->regions == 2 ? 1 : 0;
        if (st == CLH_OK) s->variant = CLH_KERNEL_FUSED;
        else if (s->cfg.kernel_variant == CLH_KERNEL_FUSED || s->num_cloths > 1) return st;
    }
    if (s->num_cloths > 1 && s->variant != CLH_KERNEL_FUSED)
        return fail(CLH_ERR_UNSUPPORTED, "batched cloths need the fused kernel variant");
    s->persist = s->cfg.persistent == CLH_PERSIST_ON && s->variant == CLH_KERNEL_FUSED &&
                 s->cfg.mode == CLH_MODE_PD && s->num_cloths == 1 && fused_persist_fits(s->fused);

    s->plan_tris = a.d_tris;
    s->plan_quads = a.d_bend_quads;
    s->plan_ready = true;
    s->graph_valid = false;
    drop_chunk_graphs(s);
    return CLH_OK;
}

int ensure_scratch(clh_solver* s)
{
    const size_t N = (size_t)s->N, T = (size_t)s->T, E = (size_t)s->E;
    // the fused variant keeps per-vertex state in tile order, padded
    const size_t Nt = s->variant == CLH_KERNEL_FUSED ? std::max(N, (size_t)s->fused.N) : N;
    CLH_TRY(s->y.resize(Nt * 3));
    CLH_TRY(s->xprev.resize(Nt * 3));
    // float iterates in reference mode need half of this
    for (auto& b : s->it) CLH_TRY(b.resize(Nt * 3));
    CLH_TRY(s->pin_map.resize(N));
    if (s->variant == CLH_KERNEL_FUSED) {
        CLH_TRY(s->vinfo.resize(Nt));
    }
    CLH_TRY(s->cloth_res.resize((size_t)s->num_cloths * 2));
    if (s->num_cloths > 1) CLH_TRY(s->cctl.resize((size_t)s->num_cloths));
    if (!s->bar.ptr) {
        CLH_TRY(s->bar.resize(kBarWords));
        CLH_HIP(hipMemset(s->bar.ptr, 0, kBarWords * sizeof(unsigned)));
    }
    CLH_TRY(s->fr.resize(T * 9));
    CLH_TRY(s->bp.resize(E * 3));
    if (s->cfg.mode == CLH_MODE_PD) {
        CLH_TRY(s->fd.resize(T * 9));
        CLH_TRY(s->bd.resize(E * 3));
    }
    s->nblocks_v = (int)grid_for(N);
    CLH_TRY(s->partials.resize((size_t)std::max(s->nblocks_v, fused_max_blocks(s->fused)) * 2));
    CLH_TRY(s->ctl.resize(1));
    if (!s->h_ctl) {
        CLH_HIP(hipHostMalloc((void**)&s->h_ctl, 2 * sizeof(FrameCtl),
                              hipHostMallocMapped | hipHostMallocCoherent));
        CLH_HIP(hipHostGetDevicePointer((void**)&s->d_ctl_mirror, s->h_ctl, 0));
    }
    for (auto& e : s->ev_chunk)
        if (!e) CLH_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    const int max_checks = s->cfg.max_iterations + 1;
    CLH_TRY(s->hist.resize((size_t)max_checks * 2));
    if (s->h_hist_cap < max_checks) {
        if (s->h_hist) (void)hipHostFree(s->h_hist);
        s->h_hist = nullptr;
        CLH_HIP(hipHostMalloc((void**)&s->h_hist, (size_t)max_checks * 2 * sizeof(double)));
        s->h_hist_cap = max_checks;
    }
    return CLH_OK;
}

// Chebyshev weight of the update producing x_{k+1} (Wang 2015 with delay S):
// 1 before S, 2/(2 - rho^2) at S, 4/(4 - rho^2 w) after.
float chebyshev_omega(const clh_config& c, int k, float prev)
{
    const int S = std::max(1, c.cheb_delay);
    const float r2 = c.rho * c.rho;
    if (k < S) return 1.0f;
    if (k == S) return 2.0f / (2.0f - r2);
    return 4.0f / (4.0f - r2 * prev);
}

bool is_check_iteration(const clh_config& c, int k)
{
    if (c.mode != CLH_MODE_PD) return false;
    if (k == c.max_iterations - 1) return true;
    return c.tolerance > 0.0f && c.check_every > 0 && ((k + 1) % c.check_every) == 0;
}

// Enqueues the prologue and pin scatter.
int enqueue_prologue(clh_solver* s, const clh_step_args& a)
{
    const clh_config& c = s->cfg;
    const float h2 = c.dt * c.dt;
    const float acc[3] = {c.ext_accel[0], c.gravity + c.ext_accel[1], c.ext_accel[2]};
    const float h2a0 = h2 * acc[0], h2a1 = h2 * acc[1], h2a2 = h2 * acc[2];
    const int disp = c.mode == CLH_MODE_PD ? 1 : 0;
    if (s->variant == CLH_KERNEL_FUSED) {
        const float h2a[3] = {h2a0, h2a1, h2a2};
        CLH_TRY(fused_pack_constants(s->fused, a, h2, s->stream));
        return fused_prologue(s->fused, a, s->xprev.ptr, s->y.ptr, s->it[0].ptr, s->vinfo.ptr, c.dt,
                              h2, acc, h2a, disp, s->stream);
    }
    if (s->N > 0)
        hipLaunchKernelGGL(k_prologue, dim3(grid_for(s->N)), dim3(kBlock), 0, s->stream, a.d_pos,
                           a.d_vel, a.d_init_guess, a.d_cloth_accel, s->xprev.ptr, s->y.ptr,
                           s->it[0].ptr, s->pin_map.ptr, s->N, c.dt, h2, acc[0], acc[1], acc[2],
                           h2a0, h2a1, h2a2, c.mode == CLH_MODE_PD ? 1 : 0);
    if (a.num_pinned > 0)
        hipLaunchKernelGGL(k_pin_scatter, dim3(grid_for(a.num_pinned)), dim3(kBlock), 0,
                           s->stream, a.d_pin_idx, a.d_pin_target, s->xprev.ptr, s->pin_map.ptr,
                           s->it[0].ptr, a.num_pinned, s->N, c.mode == CLH_MODE_PD ? 1 : 0);
    return hip_status(hipGetLastError(), "prologue launch");
}

// Residual reduction after a fused check launch: per-cloth sums for a batch,
// one sum otherwise; kofs = 1 when the check was a two-iteration launch's
// second iteration.
int enqueue_fused_reduce(clh_solver* s, int check_slot, FrameCtl* ctl, int par, int kofs)
{
    if (s->num_cloths > 1)
        hipLaunchKernelGGL(k_reduce_cloths, dim3(1), dim3(kReduceCloths), 0, s->stream,
                           s->partials.ptr, s->fused.cloth_tiles.ptr, s->num_cloths,
                           s->cloth_res.ptr, s->hist.ptr, check_slot, ctl, par, kofs,
                           ctl ? s->cctl.ptr : nullptr);
    else
        hipLaunchKernelGGL(k_reduce_partials, dim3(1), dim3(kBlock), 0, s->stream,
                           s->partials.ptr, s->fused.num_tiles, s->hist.ptr, check_slot, ctl, par,
                           kofs);
    return hip_status(hipGetLastError(), "residual reduction launch");
}

// Enqueues iteration k (x_k in it[k%4], x_{k-1} in it[(k+3)%4] -> x_{k+1} in
// it[(k+1)%4]); par is the control slot of the launch (default k & 1).
int enqueue_iteration(clh_solver* s, const clh_step_args& a, int k, float omega, bool blend_in,
                      bool check, int check_slot, bool device_ctl = false, int par = -1)
{
    if (par < 0) par = k & 1;
    const clh_config& c = s->cfg;
    FrameCtl* ctl = device_ctl ? s->ctl.ptr : nullptr;
    const float h2 = c.dt * c.dt;
    const int pd = (c.mode == CLH_MODE_PD);
    const void* xk = s->it[k % kIterBufs].ptr;
    const bool use_stretch = s->T > 0 && a.d_stretch_k != nullptr;
    const bool use_bend = s->E > 0 && a.d_bend_quads != nullptr;
    const bool blend = pd && c.use_chebyshev && blend_in && omega != 1.0f;

    if (s->variant == CLH_KERNEL_FUSED) {
        FusedIterArgs f;
        f.x = xk;
        f.xkm1 = s->it[(k + kIterBufs - 1) % kIterBufs].ptr;
        f.out = s->it[(k + 1) % kIterBufs].ptr;
        f.base = s->xprev.ptr;
        f.y = s->y.ptr;
        f.vinfo = s->vinfo.ptr;
        f.pin_target = a.d_pin_target;
        f.partials = check ? s->partials.ptr : nullptr;
        f.omega = omega;
        f.blend = blend ? 1 : 0;
        f.use_stretch = use_stretch ? 1 : 0;
        f.use_bend = use_bend ? 1 : 0;
        f.ctl = ctl;
        f.par = par;
        f.cctl = (ctl && s->num_cloths > 1) ? s->cctl.ptr : nullptr;
        f.relax = s->relax_eff;
        f.cloth_relax = s->cloth_relax_on ? s->cloth_relax.ptr : nullptr;
        CLH_TRY(fused_launch_iteration(s->fused, f, c.mode, s->stream));
        if (check) CLH_TRY(enqueue_fused_reduce(s, check_slot, ctl, par, 0));
        return hip_status(hipGetLastError(), "fused iteration launch");
    }

    if (use_stretch) {
        if (pd)
            hipLaunchKernelGGL(k_triangles<CLH_MODE_PD>, dim3(grid_for(s->T)), dim3(kBlock), 0,
                               s->stream, s->xprev.ptr, xk, a.d_tris, a.d_Dm_inv, a.d_rest_area, a.d_stretch_k,
                               s->fr.ptr, s->fd.ptr, s->T, s->N, h2, ctl);
        else
            hipLaunchKernelGGL(k_triangles<CLH_MODE_REF_CUDA>, dim3(grid_for(s->T)), dim3(kBlock),
                               0, s->stream, s->xprev.ptr, xk, a.d_tris, a.d_Dm_inv, a.d_rest_area,
                               a.d_stretch_k, s->fr.ptr, s->fd.ptr, s->T, s->N, h2, nullptr);
    }
    if (use_bend) {
        if (pd)
            hipLaunchKernelGGL(k_hinges<CLH_MODE_PD>, dim3(grid_for(s->E)), dim3(kBlock), 0,
                               s->stream, s->xprev.ptr, xk, a.d_bend_quads, a.d_bend_w, a.d_bend_n, s->bp.ptr,
                               s->bd.ptr, s->E, s->N, ctl);
        else
            hipLaunchKernelGGL(k_hinges<CLH_MODE_REF_CUDA>, dim3(grid_for(s->E)), dim3(kBlock), 0,
                               s->stream, s->xprev.ptr, xk, a.d_bend_quads, a.d_bend_w, a.d_bend_n, s->bp.ptr,
                               s->bd.ptr, s->E, s->N, nullptr);
    }
    VertexArgs va;
    va.base = s->xprev.ptr;
    va.x = xk;
    va.xkm1 = s->it[(k + kIterBufs - 1) % kIterBufs].ptr;
    va.out = s->it[(k + 1) % kIterBufs].ptr;
    va.y = s->y.ptr;
    va.mass = a.d_mass;
    va.diag = a.d_jacobi_diag;
    va.pin_map = s->pin_map.ptr;
    va.pin_target = a.d_pin_target;
    va.vt_ptr = s->vt_ptr.ptr;
    va.vt_idx = s->vt_idx.ptr;
    va.ve_ptr = s->ve_ptr.ptr;
    va.ve_idx = s->ve_idx.ptr;
    va.fr = s->fr.ptr;
    va.fd = s->fd.ptr;
    va.bp = s->bp.ptr;
    va.bd = s->bd.ptr;
    va.bk = a.d_bend_k;
    va.bw = a.d_bend_w;
    va.partials = check ? s->partials.ptr : nullptr;
    va.N = s->N;
    va.use_stretch = use_stretch ? 1 : 0;
    va.use_bend = use_bend ? 1 : 0;
    va.h2 = h2;
    va.omega = omega;
    va.blend = blend ? 1 : 0;
    va.ctl = ctl;
    va.par = k & 1;
    va.relax = s->relax_eff;
    if (s->N > 0) {
        if (pd)
            hipLaunchKernelGGL(k_vertex_pd, dim3(s->nblocks_v), dim3(kBlock), 0, s->stream, va);
        else
            hipLaunchKernelGGL(k_vertex_ref, dim3(s->nblocks_v), dim3(kBlock), 0, s->stream, va);
    }
    if (check && s->N > 0)
        hipLaunchKernelGGL(k_reduce_partials, dim3(1), dim3(kBlock), 0, s->stream, s->partials.ptr,
                           s->nblocks_v, s->hist.ptr, check_slot, ctl, par, 0);
    return hip_status(hipGetLastError(), "iteration launch");
}

bool pair_active(const clh_solver* s)
{
    return s->variant == CLH_KERNEL_FUSED && s->cfg.mode == CLH_MODE_PD && s->ipl == 2 &&
           !s->persist && s->pair_fits && fused_pair_usable(s->fused);
}

// Enqueues iterations k and k+1 in one launch (x_k in it[k%4], x_{k-1} in
// it[(k+3)%4] -> x_{k+1} in it[(k+1)%4], x_{k+2} in it[(k+2)%4]); a check
// measures iteration k+1.
int enqueue_pair(clh_solver* s, const clh_step_args& a, int k, float omega0, bool blend0_in,
                 float omega1, bool blend1_in, bool check, int check_slot, bool device_ctl, int par)
{
    const clh_config& c = s->cfg;
    FrameCtl* ctl = device_ctl ? s->ctl.ptr : nullptr;
    FusedIterArgs f;
    f.x = s->it[k % kIterBufs].ptr;
    f.xkm1 = s->it[(k + kIterBufs - 1) % kIterBufs].ptr;
    f.out = s->it[(k + 1) % kIterBufs].ptr;
    f.base = s->xprev.ptr;
    f.y = s->y.ptr;
    f.vinfo = s->vinfo.ptr;
    f.pin_target = a.d_pin_target;
    f.partials = check ? s->partials.ptr : nullptr;
    f.omega = omega0;
    f.blend = (c.use_chebyshev && blend0_in && omega0 != 1.0f) ? 1 : 0;
    f.use_stretch = (s->T > 0 && a.d_stretch_k != nullptr) ? 1 : 0;
    f.use_bend = (s->E > 0 && a.d_bend_quads != nullptr) ? 1 : 0;
    f.ctl = ctl;
    f.par = par;
    f.cctl = nullptr;  // batches to a tolerance run one iteration per launch
    f.relax = s->relax_eff;
    f.cloth_relax = s->cloth_relax_on ? s->cloth_relax.ptr : nullptr;
    const int blend1 = (c.use_chebyshev && blend1_in && omega1 != 1.0f) ? 1 : 0;
    CLH_TRY(fused_launch_pair(s->fused, f, s->it[(k + 2) % kIterBufs].ptr, omega1, blend1,
                              s->stream));
    if (check) CLH_TRY(enqueue_fused_reduce(s, check_slot, ctl, par, 1));
    return hip_status(hipGetLastError(), "two-iteration launch");
}

int enqueue_epilogue(clh_solver* s, const clh_step_args& a, int iters)
{
    const clh_config& c = s->cfg;
    const float scale = (1.0f - c.damping) / c.dt;
    if (s->variant == CLH_KERNEL_FUSED)
        return fused_epilogue(s->fused, a, s->it[iters % kIterBufs].ptr, s->xprev.ptr, s->vinfo.ptr,
                              scale, c.mode == CLH_MODE_PD ? 1 : 0, s->stream);
    if (s->N > 0)
        hipLaunchKernelGGL(k_epilogue, dim3(grid_for(s->N)), dim3(kBlock), 0, s->stream,
                           s->it[iters % kIterBufs].ptr, s->xprev.ptr, s->pin_map.ptr, a.d_pin_target,
                           a.d_pos, a.d_vel, s->N, scale, c.mode == CLH_MODE_PD ? 1 : 0);
    return hip_status(hipGetLastError(), "epilogue launch");
}

// Absolute positions of iterate k+1 into slot k of the trace.
int enqueue_trace(clh_solver* s, const clh_step_args& a, int k)
{
    if (s->variant == CLH_KERNEL_FUSED)
        return fused_trace(s->fused, s->it[(k + 1) % kIterBufs].ptr, s->xprev.ptr,
                           a.d_trace + (size_t)k * s->N * 3, s->cfg.mode == CLH_MODE_PD ? 1 : 0,
                           s->stream);
    if (s->N > 0)
        hipLaunchKernelGGL(k_trace, dim3(grid_for(s->N)), dim3(kBlock), 0, s->stream,
                           s->it[(k + 1) % kIterBufs].ptr, s->xprev.ptr,
                           a.d_trace + (size_t)k * s->N * 3, s->N,
                           s->cfg.mode == CLH_MODE_PD ? 1 : 0);
    return hip_status(hipGetLastError(), "trace launch");
}

// Fixed-length frame: every launch is known up front, so it can be captured.
int enqueue_fixed_frame(clh_solver* s, const clh_step_args& a)
{
    const int K = s->cfg.max_iterations;
    const bool pairs = pair_active(s) && !a.d_trace;
    CLH_TRY(enqueue_prologue(s, a));
    float omega = 1.0f;
    int checks = 0;
    for (int k = 0; k < K;) {
        const float w0 = chebyshev_omega(s->cfg, k, omega);
        if (pairs && k + 1 < K && !is_check_iteration(s->cfg, k)) {
            const float w1 = chebyshev_omega(s->cfg, k + 1, w0);
            const bool check = is_check_iteration(s->cfg, k + 1);
            CLH_TRY(enqueue_pair(s, a, k, w0, k > 0, w1, true, check, check ? checks : 0, false,
                                 (k >> 1) & 1));
            if (check) ++checks;
            omega = w1;
            k += 2;
            continue;
        }
        omega = w0;
        const bool check = is_check_iteration(s->cfg, k);
        CLH_TRY(enqueue_iteration(s, a, k, omega, k > 0, check, check ? checks : 0));
        if (check) ++checks;
        if (a.d_trace) CLH_TRY(enqueue_trace(s, a, k));
        ++k;
    }
    CLH_TRY(enqueue_epilogue(s, a, K));
    s->last_iters = K;
    s->last_checks = checks;
    s->converged = 0;
    return CLH_OK;
}

int validate_step_args(int N, const clh_step_args* a)
{
    if (!a) return fail(CLH_ERR_INVALID, "clh_solver_step: null args");
    if (N > 0 && (!a->d_pos || !a->d_vel || !a->d_mass || !a->d_jacobi_diag))
        return fail(CLH_ERR_INVALID, "clh_solver_step: null pos, vel, mass or jacobi_diag");
    if (a->d_stretch_k && (!a->d_tris || !a->d_Dm_inv || !a->d_rest_area))
        return fail(CLH_ERR_INVALID, "clh_solver_step: stretch enabled with null triangle data");
    if (a->d_bend_quads && (!a->d_bend_w || !a->d_bend_n || !a->d_bend_k))
        return fail(CLH_ERR_INVALID, "clh_solver_step: bending enabled with null hinge data");
    if (a->num_pinned < 0 || (a->num_pinned > 0 && (!a->d_pin_idx || !a->d_pin_target)))
        return fail(CLH_ERR_INVALID, "clh_solver_step: bad pin arguments");
    return CLH_OK;
}

int validate_args(const clh_solver* s, const clh_step_args* a) { return validate_step_args(s->N, a); }

// Resolves the Jacobi step scale of the solver's frames. With CLH_RELAX_AUTO
// in PD mode, gamma = min(1, (1 + rho) / mu) with mu the Gershgorin bound of
// D^-1 K (k_jacobi_bound): the damped Jacobi iteration matrix
// I - gamma D^-1 K then has every eigenvalue in [-rho, 1), inside the
// interval [-rho, rho] the Chebyshev weights assume at the bottom, where stiff
// bending otherwise puts eigenvalues below -1 that the weights amplify. dt
// and rho are fixed for the solver's lifetime, so the bound is recomputed
// only when the constants' or topology's pointers change.
int ensure_relax(clh_solver* s, const clh_step_args& a)
{
    const clh_config& c = s->cfg;
    if (!(c.relax < 0.0f && c.mode == CLH_MODE_PD)) {
        s->relax_eff = c.relax > 0.0f ? c.relax : 1.0f;
        s->cloth_relax_on = false;
        return CLH_OK;
    }
    const clh_step_args& r = s->relax_args;
    if (s->relax_ready && r.d_mass == a.d_mass && r.d_jacobi_diag == a.d_jacobi_diag &&
        r.d_Dm_inv == a.d_Dm_inv && r.d_rest_area == a.d_rest_area &&
        r.d_stretch_k == a.d_stretch_k && r.d_bend_w == a.d_bend_w && r.d_bend_k == a.d_bend_k &&
        r.d_bend_quads == a.d_bend_quads && r.d_tris == a.d_tris)
        return CLH_OK;
    const bool use_stretch = s->T > 0 && a.d_stretch_k != nullptr;
    const bool use_bend = s->E > 0 && a.d_bend_quads != nullptr;
    const int C = s->num_cloths;
    DevBuf<unsigned long long> mx;
    DevBuf<int> seg;
    CLH_TRY(mx.resize((size_t)C));
    if (C > 1) CLH_TRY(seg.upload(s->seg));
    CLH_HIP(hipMemsetAsync(mx.ptr, 0, sizeof(unsigned long long) * (size_t)C, s->stream));
    if (s->N > 0)
        hipLaunchKernelGGL(k_jacobi_bound, dim3(grid_for(s->N)), dim3(kBlock), 0, s->stream,
                           s->vt_ptr.ptr, s->vt_idx.ptr, s->ve_ptr.ptr, s->ve_idx.ptr, a.d_mass,
                           a.d_jacobi_diag, a.d_Dm_inv, a.d_rest_area, a.d_stretch_k, a.d_bend_w,
                           a.d_bend_k, s->N, (double)c.dt * (double)c.dt, use_stretch ? 1 : 0,
                           use_bend ? 1 : 0, C > 1 ? seg.ptr : nullptr, C, mx.ptr);
    CLH_HIP(hipGetLastError());
    std::vector<unsigned long long> bits((size_t)C, 0);
    CLH_HIP(hipMemcpyAsync(bits.data(), mx.ptr, sizeof(unsigned long long) * (size_t)C,
                           hipMemcpyDeviceToHost, s->stream));
    CLH_HIP(hipStreamSynchronize(s->stream));
    const double rho = std::min(std::max((double)c.rho, 0.0), 1.0);
    std::vector<float> gamma((size_t)C, 1.0f);
    double mu_max = 0.0;
    float g_min = 1.0f;
    for (int i = 0; i < C; ++i) {
        double mu = 0.0;
        std::memcpy(&mu, &bits[(size_t)i], sizeof(mu));
        if (!std::isfinite(mu))
            return fail(CLH_ERR_NUMERIC, "clh_solver_step: non-finite constants");
        gamma[(size_t)i] = (mu > 1.0 + rho) ? (float)((1.0 + rho) / mu) : 1.0f;
        mu_max = std::max(mu_max, mu);
        g_min = std::min(g_min, gamma[(size_t)i]);
    }
    if (C > 1) CLH_TRY(s->cloth_relax.upload(gamma));
    s->cloth_relax_on = C > 1;
    s->jacobi_bound = mu_max;
    s->relax_eff = g_min;
    s->relax_args = a;
    s->relax_ready = true;
    return CLH_OK;
}

int step_impl(clh_solver* s, const clh_step_args& a)
{
    CLH_TRY(ensure_plan(s, a));
    CLH_TRY(ensure_scratch(s));
    CLH_TRY(ensure_relax(s, a));
    if (a.num_pinned > 0 &&
        (s->pins_checked_ptr != a.d_pin_idx || s->pins_checked_n != a.num_pinned)) {
        // pin indices address the vertex arrays: check each new list once
        std::vector<int> idx((size_t)a.num_pinned);
        CLH_HIP(hipMemcpy(idx.data(), a.d_pin_idx, idx.size() * sizeof(int),
                          hipMemcpyDeviceToHost));
        for (int v : idx)
            if (v < 0 || v >= s->N) return fail(CLH_ERR_INVALID, "pinned index out of range");
        s->pins_checked_ptr = a.d_pin_idx;
        s->pins_checked_n = a.num_pinned;
    }
    const clh_config& c = s->cfg;
    const bool fixed = !(c.mode == CLH_MODE_PD && c.tolerance > 0.0f);
    if (fixed) s->last_cloth_ctl = false;

    if (fixed && c.use_graph && !a.d_trace) {
        if (!s->graph_valid || !same_args(s->graph_args, a)) {
            if (s->graph_exec) (void)hipGraphExecDestroy(s->graph_exec);
            if (s->graph) (void)hipGraphDestroy(s->graph);
            s->graph_exec = nullptr;
            s->graph = nullptr;
            CLH_HIP(hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal));
            const int st = enqueue_fixed_frame(s, a);
            hipGraph_t g = nullptr;
            const hipError_t e = hipStreamEndCapture(s->stream, &g);
            if (st != CLH_OK) {
                if (g) (void)hipGraphDestroy(g);
                return st;
            }
            CLH_HIP(e);
            s->graph = g;
            CLH_HIP(hipGraphInstantiate(&s->graph_exec, g, nullptr, nullptr, 0));
            s->graph_args = a;
            s->graph_valid = true;
        }
        CLH_HIP(hipGraphLaunch(s->graph_exec, s->stream));
        s->last_iters = c.max_iterations;
        s->last_checks = (c.mode == CLH_MODE_PD && c.max_iterations > 0) ? 1 : 0;
        s->converged = 0;
        return CLH_OK;
    }
    if (fixed) return enqueue_fixed_frame(s, a);

    // PD mode with a tolerance: a residual test every check_every iterations
    // (and at the last), evaluated and acted on by the device (FrameCtl). The
    // host enqueues the frame in chunks of check_every iterations, each a
    // replayed hipGraph, one chunk ahead of the control snapshot it reads;
    // once the device has stopped, the chunk in flight returns at once.
    const int K = c.max_iterations, P = std::max(1, c.check_every);
    const bool graphs = c.use_graph && !a.d_trace;
    if (!same_args(s->chunk_args, a) || s->chunk_len != P) {
        drop_chunk_graphs(s);
        s->chunk_args = a;
        s->chunk_len = P;
    }
    CLH_TRY(enqueue_prologue(s, a));
    // the persistent kernel's grid barrier reports its time-outs in the block
    // itself, so that path copies the block back; the others read the copies
    // the reductions write
    const bool mirrored = !s->persist || a.d_trace;
    hipLaunchKernelGGL(k_ctl_frame_init, dim3(1), dim3(64), 0, s->stream, s->ctl.ptr, c.rho,
                       c.use_chebyshev ? 1 : 0, c.cheb_delay,
                       (double)c.tolerance * (double)c.tolerance,
                       mirrored ? s->d_ctl_mirror : nullptr, c.stall_checks, c.restart_rule,
                       c.restart_factor);
    CLH_HIP(hipGetLastError());
    // a batch: every cloth has its own control (fused variant)
    s->last_cloth_ctl = s->num_cloths > 1 && s->variant == CLH_KERNEL_FUSED;
    if (s->last_cloth_ctl) {
        hipLaunchKernelGGL(k_cloth_ctl_init, dim3(grid_for(s->num_cloths)), dim3(kBlock), 0,
                           s->stream, s->cctl.ptr, s->num_cloths, c.rho);
        CLH_HIP(hipGetLastError());
    }
    // two-iteration launches when every check falls on a pair's second
    // iteration: chunks of even length start at even k, so pairs run at even
    // k, and a last odd iteration (K odd) runs alone; the control slot then
    // alternates per launch, (k >> 1) & 1
    const bool pairs = pair_active(s) && !a.d_trace && P % 2 == 0 && s->num_cloths == 1;
    auto enqueue_chunk = [&](int k0, int k1) -> int {
        for (int k = k0; k < k1;) {
            if (pairs && k + 1 < k1 && !is_check_iteration(c, k)) {
                CLH_TRY(enqueue_pair(s, a, k, 1.0f, false, 1.0f, false, is_check_iteration(c, k + 1),
                                     0, true, (k >> 1) & 1));
                k += 2;
                continue;
            }
            CLH_TRY(enqueue_iteration(s, a, k, 1.0f, false, is_check_iteration(c, k), 0, true,
                                      pairs ? (k >> 1) & 1 : k & 1));
            if (a.d_trace) CLH_TRY(enqueue_trace(s, a, k));
            ++k;
        }
        return CLH_OK;
    };
    auto issue = [&](int ci) -> int {
        const int k0 = ci * P, k1 = std::min(K, k0 + P);
        if (s->persist && !a.d_trace) {
            FusedIterArgs f = {};
            f.base = s->xprev.ptr;
            f.y = s->y.ptr;
            f.vinfo = s->vinfo.ptr;
            f.pin_target = a.d_pin_target;
            f.use_stretch = (s->T > 0 && a.d_stretch_k != nullptr) ? 1 : 0;
            f.use_bend = (s->E > 0 && a.d_bend_quads != nullptr) ? 1 : 0;
            f.relax = s->relax_eff;  // single cloth: cloth_relax stays null
            double* const its[kIterBufs] = {s->it[0].ptr, s->it[1].ptr, s->it[2].ptr, s->it[3].ptr};
            CLH_TRY(fused_launch_persist(s->fused, f, its, k0, k1, K, P, s->ctl.ptr,
                                         s->partials.ptr, s->hist.ptr, s->bar.ptr, s->stream));
        } else if (graphs && k1 - k0 == P) {
            hipGraphExec_t& ex = s->chunk_exec[k0 % kIterBufs];
            if (!ex) {
                CLH_HIP(hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal));
                const int st = enqueue_chunk(k0, k1);
                hipGraph_t g = nullptr;
                const hipError_t e = hipStreamEndCapture(s->stream, &g);
                if (st != CLH_OK) {
                    if (g) (void)hipGraphDestroy(g);
                    return st;
                }
                CLH_HIP(e);
                const hipError_t ei = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
                (void)hipGraphDestroy(g);
                CLH_HIP(ei);
            }
            CLH_HIP(hipGraphLaunch(ex, s->stream));
        } else {
            CLH_TRY(enqueue_chunk(k0, k1));
        }
        if (!mirrored)
            CLH_HIP(hipMemcpyAsync(&s->h_ctl[ci & 1], s->ctl.ptr, sizeof(FrameCtl),
                                   hipMemcpyDeviceToHost, s->stream));
        CLH_HIP(hipEventRecord(s->ev_chunk[ci & 1], s->stream));
        return CLH_OK;
    };
    const int nchunks = (K + P - 1) / P;
    int iters = K, checks = 0;
    s->converged = 0;
    s->last_restarts = 0;
    s->last_rho = c.rho;
    bool nonfinite = false;
    if (nchunks > 0) CLH_TRY(issue(0));
    if (nchunks > 1) CLH_TRY(issue(1));
    for (int ci = 0; ci < nchunks; ++ci) {
        CLH_HIP(hipEventSynchronize(s->ev_chunk[ci & 1]));
        const FrameCtl h = s->h_ctl[ci & 1];
        if (h.barrier_error) return fail(CLH_ERR_DEVICE, "persistent kernel: a grid barrier timed out");
        checks = h.checks;
        s->last_restarts = h.restarts;
        s->last_rho = h.rho;
        if (h.stop) {
            iters = h.iters;
            s->converged = h.converged;
            nonfinite = h.nonfinite != 0;
            break;
        }
        if (ci + 2 < nchunks) CLH_TRY(issue(ci + 2));
    }
    s->last_iters = iters;
    s->last_checks = checks;
    if (nonfinite) return fail(CLH_ERR_NUMERIC, "non-finite residual");
    return enqueue_epilogue(s, a, iters);
}

}  // namespace

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------

extern "C" int clh_config_default(clh_config* c)
{
    if (!c) return fail(CLH_ERR_INVALID, "clh_config_default: null pointer");
    std::memset(c, 0, sizeof(*c));
    c->mode = CLH_MODE_REF_CUDA;
    c->max_iterations = 50;
    c->use_chebyshev = 1;
    c->rho = 0.9f;
    c->gravity = -9.8f;
    c->dt = 0.01f;
    c->damping = 0.0f;
    c->tolerance = 0.0f;
    c->check_every = 10;
    c->cheb_delay = 0;
    c->kernel_variant = CLH_KERNEL_AUTO;
    c->use_graph = 1;
    return CLH_OK;
}

extern "C" int clh_solver_create(const clh_config* cfg, const clh_mesh_desc* mesh,
                                 clh_solver** out)
{
    if (!cfg || !mesh || !out) return fail(CLH_ERR_INVALID, "clh_solver_create: null pointer");
    *out = nullptr;
    if (mesh->num_verts < 0 || mesh->num_tris < 0 || mesh->num_bend < 0)
        return fail(CLH_ERR_INVALID, "clh_solver_create: negative count");
    if (cfg->mode != CLH_MODE_REF_CUDA && cfg->mode != CLH_MODE_PD)
        return fail(CLH_ERR_INVALID, "clh_solver_create: unknown mode");
    if (cfg->max_iterations < 0 || cfg->max_iterations > kMaxIterations)
        return fail(CLH_ERR_INVALID, "clh_solver_create: max_iterations out of range");
    if (!(cfg->dt > 0.0f)) return fail(CLH_ERR_INVALID, "clh_solver_create: dt must be > 0");
    if (cfg->kernel_variant < CLH_KERNEL_AUTO || cfg->kernel_variant > CLH_KERNEL_FUSED)
        return fail(CLH_ERR_INVALID, "clh_solver_create: unknown kernel variant");
    if (!((cfg->relax >= 0.0f && cfg->relax < 2.0f) || cfg->relax == CLH_RELAX_AUTO))
        return fail(CLH_ERR_INVALID, "clh_solver_create: relax must be in [0, 2) or CLH_RELAX_AUTO");
    if (cfg->restart_rule < CLH_RESTART_PREV || cfg->restart_rule > CLH_RESTART_NONE)
        return fail(CLH_ERR_INVALID, "clh_solver_create: unknown restart rule");
    if (cfg->restart_rule == CLH_RESTART_SEQ_MIN && !(cfg->restart_factor >= 1.0f &&
                                                      cfg->restart_factor < 1e30f))
        return fail(CLH_ERR_INVALID, "clh_solver_create: restart_factor must be finite and >= 1");
    clh_solver* s = new clh_solver();
    s->cfg = *cfg;
    if (s->cfg.check_every <= 0) s->cfg.check_every = 1;
    s->N = mesh->num_verts;
    s->T = mesh->num_tris;
    s->E = mesh->num_bend;
    if (cfg->tile_verts > 0) s->tile_verts = std::max(32, std::min(cfg->tile_verts, 1024));
    if (cfg->block_threads != 0 && cfg->block_threads != 256 && cfg->block_threads != 512 &&
        cfg->block_threads != 1024) {
        clh_solver_destroy(s);
        return fail(CLH_ERR_INVALID, "clh_solver_create: block_threads must be 0, 256, 512 or 1024");
    }
    if (cfg->block_threads > 0) s->block_threads = cfg->block_threads;
    if (cfg->iters_per_launch < 0 || cfg->iters_per_launch > 2) {
        clh_solver_destroy(s);
        return fail(CLH_ERR_INVALID, "clh_solver_create: iters_per_launch must be 0, 1 or 2");
    }
    s->ipl = cfg->iters_per_launch == 0 ? kDefaultItersPerLaunch : cfg->iters_per_launch;
    if (cfg->element_regions < 0 || cfg->element_regions > 2) {
        clh_solver_destroy(s);
        return fail(CLH_ERR_INVALID, "clh_solver_create: element_regions must be 0, 1 or 2");
    }
    s->regions = cfg->element_regions == 0 ? kDefaultElementRegions : cfg->element_regions;
    int st = hip_status(hipStreamCreateWithFlags(&s->own_stream, hipStreamNonBlocking),
                        "hipStreamCreate");
    if (st == CLH_OK) st = hip_status(hipEventCreateWithFlags(&s->ev_in, hipEventDisableTiming), "hipEventCreate");
    if (st == CLH_OK) st = hip_status(hipEventCreateWithFlags(&s->ev_out, hipEventDisableTiming), "hipEventCreate");
    if (st == CLH_OK) st = hip_status(hipEventCreate(&s->ev_t0), "hipEventCreate");
    if (st == CLH_OK) st = hip_status(hipEventCreate(&s->ev_t1), "hipEventCreate");
    if (st != CLH_OK) {
        clh_solver_destroy(s);
        return st;
    }
    s->stream = s->own_stream;
    // the plan is built at the first step, which also supplies the positions
    // the fused variant orders its tiles by
    *out = s;
    return CLH_OK;
}

extern "C" int clh_solver_destroy(clh_solver* s)
{
    if (!s) return CLH_OK;
    if (s->stream) (void)hipStreamSynchronize(s->stream);
    if (s->graph_exec) (void)hipGraphExecDestroy(s->graph_exec);
    if (s->graph) (void)hipGraphDestroy(s->graph);
    drop_chunk_graphs(s);
    for (hipEvent_t e : {s->ev_in, s->ev_out, s->ev_t0, s->ev_t1, s->ev_chunk[0], s->ev_chunk[1]})
        if (e) (void)hipEventDestroy(e);
    if (s->h_ctl) (void)hipHostFree(s->h_ctl);
    if (s->own_stream) (void)hipStreamDestroy(s->own_stream);
    if (s->h_hist) (void)hipHostFree(s->h_hist);
    fused_plan_release(s->fused);
    delete s;
    return CLH_OK;
}

extern "C" int clh_solver_reset(clh_solver* s)
{
    if (!s) return fail(CLH_ERR_INVALID, "clh_solver_reset: null solver");
    CLH_HIP(hipStreamSynchronize(s->stream));
    s->plan_ready = false;
    s->relax_ready = false;  // constants may have changed in place
    s->graph_valid = false;
    drop_chunk_graphs(s);
    s->pins_checked_ptr = nullptr;
    s->pins_checked_n = -1;
    s->frames = 0;
    return CLH_OK;
}

extern "C" int clh_solver_set_stream(clh_solver* s, void* stream)
{
    if (!s) return fail(CLH_ERR_INVALID, "clh_solver_set_stream: null solver");
    CLH_HIP(hipStreamSynchronize(s->stream));
    s->user_stream = (stream != nullptr);
    s->stream = stream ? static_cast<hipStream_t>(stream) : s->own_stream;
    s->graph_valid = false;
    return CLH_OK;
}

extern "C" int clh_solver_step(clh_solver* s, const clh_step_args* a)
{
    if (!s) return fail(CLH_ERR_INVALID, "clh_solver_step: null solver");
    CLH_TRY(validate_args(s, a));
    if (!s->user_stream) {
        // order after earlier work on the null stream, and make later
        // null-stream work (blocking copies) wait for this frame
        CLH_HIP(hipEventRecord(s->ev_in, 0));
        CLH_HIP(hipStreamWaitEvent(s->stream, s->ev_in, 0));
    }
    CLH_TRY(step_impl(s, *a));
    if (!s->user_stream) {
        CLH_HIP(hipEventRecord(s->ev_out, s->stream));
        CLH_HIP(hipStreamWaitEvent(0, s->ev_out, 0));
    }
    ++s->frames;
    return CLH_OK;
}

extern "C" int clh_solver_get_stats(clh_solver* s, clh_step_stats* out)
{
    if (!s || !out) return fail(CLH_ERR_INVALID, "clh_solver_get_stats: null pointer");
    CLH_HIP(hipStreamSynchronize(s->stream));
    std::memset(out, 0, sizeof(*out));
    out->iterations = s->last_iters;
    out->converged = s->converged;
    out->frames = s->frames;
    out->restarts = s->last_cloth_ctl ? 0 : s->last_restarts;
    out->rho = s->last_cloth_ctl ? s->cfg.rho : s->last_rho;
    out->relax = s->relax_eff;
    out->jacobi_bound = s->jacobi_bound;
    if (s->cfg.mode == CLH_MODE_PD && s->last_checks > 0 && s->hist.ptr) {
        std::vector<double> h((size_t)s->last_checks * 2);
        CLH_HIP(hipMemcpy(h.data(), s->hist.ptr, h.size() * sizeof(double), hipMemcpyDeviceToHost));
        out->residual0 = std::sqrt(h[0]);
        out->residual = std::sqrt(h[h.size() - 2]);
        out->rhs_norm = std::sqrt(h[h.size() - 1]);
    }
    return CLH_OK;
}

extern "C" int clh_solver_residual_history(clh_solver* s, double* out, int max_count, int* count)
{
    if (!s || !count) return fail(CLH_ERR_INVALID, "clh_solver_residual_history: null pointer");
    CLH_HIP(hipStreamSynchronize(s->stream));
    *count = (s->cfg.mode == CLH_MODE_PD) ? s->last_checks : 0;
    const int n = std::min(*count, std::max(max_count, 0));
    if (n > 0 && out) {
        std::vector<double> h((size_t)n * 2);
        CLH_HIP(hipMemcpy(h.data(), s->hist.ptr, h.size() * sizeof(double), hipMemcpyDeviceToHost));
        for (int i = 0; i < n; ++i)
            out[i] = h[2 * i + 1] > 0.0 ? std::sqrt(h[2 * i] / h[2 * i + 1]) : 0.0;
    }
    return CLH_OK;
}

extern "C" int clh_solver_run_frames(clh_solver* s, const clh_step_args* a, int num_frames,
                                     float* elapsed_ms)
{
    if (!s || !elapsed_ms || num_frames <= 0)
        return fail(CLH_ERR_INVALID, "clh_solver_run_frames: bad arguments");
    CLH_TRY(validate_args(s, a));
    CLH_HIP(hipDeviceSynchronize());
    CLH_HIP(hipEventRecord(s->ev_t0, s->stream));
    for (int f = 0; f < num_frames; ++f) {
        CLH_TRY(step_impl(s, *a));
        ++s->frames;
    }
    CLH_HIP(hipEventRecord(s->ev_t1, s->stream));
    CLH_HIP(hipEventSynchronize(s->ev_t1));
    CLH_HIP(hipEventElapsedTime(elapsed_ms, s->ev_t0, s->ev_t1));
    return CLH_OK;
}

namespace {
// Average duration (ms) of one launch of the iteration kernel in use, in the
// state of the frame loop's steady phase: a fresh prologue leaves x_0 in
// it[0], iteration 0 writes x_1 to it[1], and the timed launches continue
// with iterations 1, 2, ... (pairs 1-2, 3-4, ... with two-iteration
// launches) at a fixed Chebyshev weight with the blend on (in PD mode with
// Chebyshev). Each launch reads x_k and x_{k-1} and the iterate the launch
// before it wrote, as in a frame (re-running one iteration on unchanged
// inputs measured about 0.4 us less per launch at 256x256). Only solver
// scratch is written.
int time_iterations(clh_solver* s, const clh_step_args& a, int num_launches, float* avg_ms)
{
    const bool pairs = pair_active(s);
    const int step = pairs ? 2 : 1;
    int k = 1;
    auto launch = [&]() -> int {
        const int st = pairs ? enqueue_pair(s, a, k, 1.5f, true, 1.5f, true, false, 0, false, 0)
                             : enqueue_iteration(s, a, k, 1.5f, true, false, 0);
        k += step;
        return st;
    };
    CLH_TRY(enqueue_prologue(s, a));
    CLH_TRY(enqueue_iteration(s, a, 0, 1.0f, false, false, 0));
    CLH_TRY(launch());  // warm-up
    CLH_HIP(hipStreamSynchronize(s->stream));
    CLH_HIP(hipEventRecord(s->ev_t0, s->stream));
    for (int i = 0; i < num_launches; ++i) CLH_TRY(launch());
    CLH_HIP(hipEventRecord(s->ev_t1, s->stream));
    CLH_HIP(hipEventSynchronize(s->ev_t1));
    float ms = 0.0f;
    CLH_HIP(hipEventElapsedTime(&ms, s->ev_t0, s->ev_t1));
    *avg_ms = ms / (float)num_launches;
    return CLH_OK;
}
}  // namespace

extern "C" int clh_solver_time_iteration_kernel(clh_solver* s, const clh_step_args* a,
                                                int num_launches, float* avg_ms)
{
    if (!s || !avg_ms || num_launches <= 0)
        return fail(CLH_ERR_INVALID, "clh_solver_time_iteration_kernel: bad arguments");
    CLH_TRY(validate_args(s, a));
    CLH_TRY(ensure_plan(s, *a));
    CLH_TRY(ensure_scratch(s));
    CLH_TRY(ensure_relax(s, *a));  // the timed launches use the frames' step scale
    CLH_TRY(time_iterations(s, *a, num_launches, avg_ms));
    return CLH_OK;
}

extern "C" int clh_solver_plan_info(clh_solver* s, clh_plan_info* out)
{
    if (!s || !out) return fail(CLH_ERR_INVALID, "clh_solver_plan_info: null pointer");
    std::memset(out, 0, sizeof(*out));
    out->variant = s->plan_ready ? s->variant : CLH_KERNEL_AUTO;
    out->iters_per_launch = 1;
    if (s->plan_ready && s->variant == CLH_KERNEL_FUSED) {
        fused_plan_info(s->fused, out);
        out->lds_limit = s->fused.lds_limit;
    }
    out->persistent = s->plan_ready && s->persist ? 1 : 0;
    if (s->plan_ready && pair_active(s)) out->iters_per_launch = 2;
    if (!(s->plan_ready && s->variant == CLH_KERNEL_FUSED && s->cfg.mode == CLH_MODE_PD))
        out->element_regions = 1;
    return CLH_OK;
}

// ---------------------------------------------------------------------------
// batched cloths and the autotuner
// ---------------------------------------------------------------------------

extern "C" int clh_batch_create(const clh_config* cfg, const clh_mesh_desc* mesh, int num_cloths,
                                const int* offs, clh_solver** out)
{
    if (!cfg || !mesh || !out || !offs || num_cloths < 1)
        return fail(CLH_ERR_INVALID, "clh_batch_create: bad arguments");
    *out = nullptr;
    if (cfg->kernel_variant == CLH_KERNEL_UNFUSED)
        return fail(CLH_ERR_UNSUPPORTED, "clh_batch_create: batched cloths need the fused kernel");
    if (offs[0] != 0 || offs[num_cloths] != mesh->num_verts)
        return fail(CLH_ERR_INVALID, "clh_batch_create: offsets must run from 0 to num_verts");
    for (int c = 0; c < num_cloths; ++c)
        if (offs[c + 1] <= offs[c])
            return fail(CLH_ERR_INVALID, "clh_batch_create: every cloth needs at least one vertex");
    clh_solver* s = nullptr;
    CLH_TRY(clh_solver_create(cfg, mesh, &s));
    s->seg.assign(offs, offs + num_cloths + 1);
    s->num_cloths = num_cloths;
    *out = s;
    return CLH_OK;
}

extern "C" int clh_batch_cloth_residuals(clh_solver* s, double* out, int max_count, int* count)
{
    if (!s || !count) return fail(CLH_ERR_INVALID, "clh_batch_cloth_residuals: null pointer");
    CLH_HIP(hipStreamSynchronize(s->stream));
    *count = s->num_cloths;
    if (s->cfg.mode != CLH_MODE_PD || s->last_checks == 0 || !out) return CLH_OK;
    const int n = std::min(s->num_cloths, std::max(max_count, 0));
    std::vector<double> h((size_t)s->num_cloths * 2);
    if (s->num_cloths > 1) {
        CLH_HIP(hipMemcpy(h.data(), s->cloth_res.ptr, h.size() * sizeof(double),
                          hipMemcpyDeviceToHost));
    } else {
        // one cloth: the last entry of the history
        CLH_HIP(hipMemcpy(h.data(), s->hist.ptr + 2 * (size_t)(s->last_checks - 1),
                          2 * sizeof(double), hipMemcpyDeviceToHost));
    }
    for (int c = 0; c < n; ++c)
        out[c] = h[2 * (size_t)c + 1] > 0.0 ? std::sqrt(h[2 * (size_t)c] / h[2 * (size_t)c + 1]) : 0.0;
    return CLH_OK;
}

extern "C" int clh_batch_cloth_status(clh_solver* s, int* state, int* iterations, int max_count,
                                      int* count)
{
    if (!s || !count) return fail(CLH_ERR_INVALID, "clh_batch_cloth_status: null pointer");
    CLH_HIP(hipStreamSynchronize(s->stream));
    *count = s->num_cloths;
    const int n = std::min(s->num_cloths, std::max(max_count, 0));
    if (n <= 0 || (!state && !iterations)) return CLH_OK;
    std::vector<ClothCtl> h;
    if (s->last_cloth_ctl) {
        h.resize((size_t)s->num_cloths);
        CLH_HIP(hipMemcpy(h.data(), s->cctl.ptr, h.size() * sizeof(ClothCtl), hipMemcpyDeviceToHost));
    }
    for (int c = 0; c < n; ++c) {
        // without per-cloth control (fixed frames, one cloth) every cloth ran
        // the frame's iterations; a cloth still running at the end ran them all
        const int st = h.empty() ? (s->converged ? CLH_CLOTH_CONVERGED : CLH_CLOTH_RUNNING)
                                 : h[(size_t)c].state;
        if (state) state[c] = st;
        if (iterations)
            iterations[c] = (h.empty() || st == CLH_CLOTH_RUNNING) ? s->last_iters : h[(size_t)c].iters;
    }
    return CLH_OK;
}

namespace {
std::mutex g_tune_mu;
struct TuneChoice {
    int tile_verts, threads, ipl, regions;
};
std::map<unsigned long long, TuneChoice> g_tune_cache;

// Rebuilds the plan with the given launch shape; returns the tile size the
// plan actually got (smaller when the tile did not fit in LDS).
int replan(clh_solver* s, const clh_step_args& a, int tile_verts, int threads, int& got)
{
    s->tile_verts = tile_verts;
    s->block_threads = threads;
    s->plan_ready = false;
    CLH_TRY(ensure_plan(s, a));
    CLH_TRY(ensure_scratch(s));
    s->fused.threads = threads;
    got = s->variant == CLH_KERNEL_FUSED ? s->fused.tile_verts : 0;
    return CLH_OK;
}

// Sets the launch shape of an already built plan (workgroup size, iterations
// per launch, element regions); captured frames of the previous shape are
// dropped.
void set_shape(clh_solver* s, int threads, int ipl, int regions)
{
    s->block_threads = threads;
    s->fused.threads = threads;
    s->ipl = ipl;
    s->regions = regions;
    s->fused.split = regions == 2 ? 1 : 0;
    // the persistent kernel is built for some workgroup sizes only
    s->persist = s->cfg.persistent == CLH_PERSIST_ON && s->variant == CLH_KERNEL_FUSED &&
                 s->cfg.mode == CLH_MODE_PD && s->num_cloths == 1 && fused_persist_fits(s->fused);
    s->graph_valid = false;
    drop_chunk_graphs(s);
}
}  // namespace

extern "C" int clh_solver_autotune(clh_solver* s, const clh_step_args* a, int num_launches,
                                   clh_tune_report* out)
{
    if (!s || !out || num_launches <= 0)
        return fail(CLH_ERR_INVALID, "clh_solver_autotune: bad arguments");
    CLH_TRY(validate_args(s, a));
    if (s->cfg.kernel_variant == CLH_KERNEL_UNFUSED)
        return fail(CLH_ERR_UNSUPPORTED, "clh_solver_autotune: the unfused variant has no launch shape");
    std::memset(out, 0, sizeof(*out));
    CLH_TRY(ensure_plan(s, *a));
    if (s->variant != CLH_KERNEL_FUSED)
        return fail(CLH_ERR_UNSUPPORTED, "clh_solver_autotune: the mesh has no fused plan");
    CLH_TRY(ensure_relax(s, *a));  // candidates run at the frames' step scale
    out->key = s->shape_key;
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        auto it = g_tune_cache.find(s->shape_key);
        if (it != g_tune_cache.end()) {
            const TuneChoice c = it->second;
            int got = 0;
            if (s->fused.tile_verts != c.tile_verts) CLH_TRY(replan(s, *a, c.tile_verts, c.threads, got));
            set_shape(s, c.threads, c.ipl, c.regions);
            out->num_candidates = 1;
            out->tile_verts[0] = c.tile_verts;
            out->block_threads[0] = c.threads;
            out->iters_per_launch[0] = c.ipl;
            out->element_regions[0] = c.regions;
            out->best = 0;
            out->from_cache = 1;
            return CLH_OK;
        }
    }
    const int tiles[4] = {32, 64, 128, 256}, threads[3] = {256, 512, 1024};
    const bool pd = s->cfg.mode == CLH_MODE_PD;
    const bool try_pairs = pd && s->cfg.iters_per_launch == 2;
    int n = 0, best = -1;
    // The GPU's clocks ramp up over the first tens of milliseconds of work
    // that follows the host-side replan: each tile size starts with about
    // 20 ms of launches that are not kept, and each candidate takes the lower
    // of two timings.
    auto warm_up = [&]() -> int {
        float ms = 0.0f;
        CLH_TRY(time_iterations(s, *a, num_launches, &ms));
        const double n = std::ceil(20.0 / std::max((double)ms, 1e-4));
        CLH_TRY(time_iterations(s, *a, (int)std::min(n, 100000.0), &ms));
        return CLH_OK;
    };
    for (int tv : tiles) {
        int got = 0;
        CLH_TRY(replan(s, *a, tv, 256, got));
        if (got == tv) CLH_TRY(warm_up());
        for (int th : threads) {
            for (int ipl = 1; ipl <= (try_pairs ? 2 : 1); ++ipl) {
                for (int reg = 1; reg <= (pd && ipl == 1 ? 2 : 1); ++reg) {
                    out->tile_verts[n] = tv;
                    out->block_threads[n] = th;
                    out->iters_per_launch[n] = ipl;
                    out->element_regions[n] = reg;
                    out->us_per_iteration[n] = 0.0f;
                    if (got == tv) {
                        set_shape(s, th, ipl, reg);
                        clh_plan_info pi;
                        fused_plan_info(s->fused, &pi);
                        const bool runs = ipl == 2 ? pair_active(s) : (reg == 1 || pi.element_regions == 2);
                        if (runs) {
                            float ms0 = 0.0f, ms1 = 0.0f;
                            CLH_TRY(time_iterations(s, *a, num_launches, &ms0));
                            CLH_TRY(time_iterations(s, *a, num_launches, &ms1));
                            out->us_per_iteration[n] = 1e3f * std::min(ms0, ms1) / (float)ipl;
                            if (best < 0 || out->us_per_iteration[n] < out->us_per_iteration[best])
                                best = n;
                        }
                    }
                    ++n;
                }
            }
        }
    }
    out->num_candidates = n;
    if (best < 0) return fail(CLH_ERR_UNSUPPORTED, "clh_solver_autotune: no candidate fits");
    out->best = best;
    int got = 0;
    CLH_TRY(replan(s, *a, out->tile_verts[best], out->block_threads[best], got));
    set_shape(s, out->block_threads[best], out->iters_per_launch[best], out->element_regions[best]);
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tune_cache[s->shape_key] = {out->tile_verts[best], out->block_threads[best],
                                  out->iters_per_launch[best], out->element_regions[best]};
    return CLH_OK;
}

extern "C" int clh_solver_start_residual(clh_solver* s, const clh_step_args* a, double* rel)
{
    if (!s || !rel) return fail(CLH_ERR_INVALID, "clh_solver_start_residual: null pointer");
    CLH_TRY(validate_args(s, a));
    if (s->cfg.mode != CLH_MODE_PD)
        return fail(CLH_ERR_UNSUPPORTED, "clh_solver_start_residual: PD mode only");
    CLH_TRY(ensure_plan(s, *a));
    CLH_TRY(ensure_scratch(s));
    if (!s->user_stream) {
        CLH_HIP(hipEventRecord(s->ev_in, 0));
        CLH_HIP(hipStreamWaitEvent(s->stream, s->ev_in, 0));
    }
    // the prologue writes only solver scratch; iteration 0 with a residual
    // check measures its first iterate and writes x_1 to scratch as well
    CLH_TRY(enqueue_prologue(s, *a));
    CLH_TRY(enqueue_iteration(s, *a, 0, 1.0f, false, true, 0));
    double h[2] = {0.0, 0.0};
    CLH_HIP(hipMemcpyAsync(h, s->hist.ptr, sizeof(h), hipMemcpyDeviceToHost, s->stream));
    CLH_HIP(hipStreamSynchronize(s->stream));
    *rel = h[1] > 0.0 ? std::sqrt(h[0] / h[1]) : 0.0;
    return CLH_OK;
}

// ---------------------------------------------------------------------------
// fixed-length frames without allocation
// ---------------------------------------------------------------------------
//
// clh_step_noalloc runs the unfused kernels above (prologue, triangles,
// hinges, vertex update, epilogue) with every buffer they need carved out of
// one static device array, so a call allocates nothing. The per-vertex
// incidence lists that clh_solver_step builds on the host at its first step
// are built on the device at each call instead: count, scan, place, and a
// sort of each vertex's list into ascending slot order, which gives the
// host's lists exactly, and so the same gather order and the same bits.
//
// Scratch layout for N vertices, T triangles and E hinges, each region
// aligned to 256 bytes: y [3N] f32, x_n [3N] f32, three iterates [3N] f64,
// pin map [N], corner forces [9T] f32 twice, hinge vectors [3E] f32 twice,
// triangle and hinge incidence (offsets [N+1], slots [3T] and [4E]), and a
// counter per vertex: about 112 N + 84 T + 40 E bytes.

namespace {

constexpr size_t kNoAllocScratchBytes = CLH_NOALLOC_SCRATCH_BYTES;
constexpr int kNoAllocMaxDevices = 64;
constexpr int kNoAllocBufs = 3;  // x_{k-1}, x_k, x_{k+1}

__device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

// Device address of `scratch` per device, taken once with that device current.
std::mutex g_scratch_mu;
void* g_scratch_ptr[kNoAllocMaxDevices] = {};

int scratch_address(void** out)
{
    int dev = 0;
    CLH_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kNoAllocMaxDevices)
        return fail(CLH_ERR_UNSUPPORTED, "clh_step_noalloc: device ordinal beyond the table");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        CLH_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) return fail(CLH_ERR_DEVICE, "clh_step_noalloc: no address for the scratch array");
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return CLH_OK;
}

__global__ void __launch_bounds__(kBlock) k_fill_i32(int* __restrict__ p, int n, int value)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = value;
}

// count[v] = number of slots naming vertex v (indices outside [0, N) skipped)
__global__ void __launch_bounds__(kBlock)
k_count_slots(const int* __restrict__ elems, int nslots, int N, int* __restrict__ count)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nslots) return;
    const int v = elems[s];
    if ((unsigned)v < (unsigned)N) atomicAdd(&count[v], 1);
}

// ptr[0] = 0, ptr[v + 1] = ptr[v] + count[v], in one workgroup: each thread
// sums a contiguous run, the run sums are scanned in LDS, then each thread
// writes its run's offsets.
constexpr int kScanThreads = 1024;
__global__ void __launch_bounds__(kScanThreads)
k_scan_counts(const int* __restrict__ count, int N, int* __restrict__ ptr)
{
    __shared__ int part[kScanThreads];
    const int t = threadIdx.x;
    const int run = (N + kScanThreads - 1) / kScanThreads;
    const int b = min(N, t * run), e = min(N, b + run);
    int sum = 0;
    for (int v = b; v < e; ++v) sum += count[v];
    part[t] = sum;
    __syncthreads();
    for (int off = 1; off < kScanThreads; off <<= 1) {
        const int add = t >= off ? part[t - off] : 0;
        __syncthreads();
        part[t] += add;
        __syncthreads();
    }
    int acc = part[t] - sum;  // exclusive prefix of this run
    if (t == 0) ptr[0] = 0;
    for (int v = b; v < e; ++v) {
        acc += count[v];
        ptr[v + 1] = acc;
    }
}

// idx[ptr[v] + j] = slot for the j-th slot of v to arrive (order fixed later)
__global__ void __launch_bounds__(kBlock)
k_place_slots(const int* __restrict__ elems, int nslots, int N, const int* __restrict__ ptr,
              int* __restrict__ cursor, int* __restrict__ idx)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nslots) return;
    const int v = elems[s];
    if ((unsigned)v >= (unsigned)N) return;
    idx[ptr[v] + atomicAdd(&cursor[v], 1)] = s;
}

// Each vertex's slots into ascending order (insertion sort: lists are short).
__global__ void __launch_bounds__(kBlock)
k_sort_lists(const int* __restrict__ ptr, int* __restrict__ idx, int N)
{
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= N) return;
    const int b = ptr[v], e = ptr[v + 1];
    for (int i = b + 1; i < e; ++i) {
        const int key = idx[i];
        int j = i - 1;
        while (j >= b && idx[j] > key) {
            idx[j + 1] = idx[j];
            --j;
        }
        idx[j + 1] = key;
    }
}

struct NoAllocLayout {
    size_t y, xprev, it[kNoAllocBufs], pin_map, fr, fd, bp, bd, vt_ptr, vt_idx, ve_ptr, ve_idx,
        cursor, total;
};

NoAllocLayout noalloc_layout(size_t N, size_t T, size_t E)
{
    NoAllocLayout L;
    size_t off = 0;
    auto take = [&](size_t bytes) {
        const size_t at = off;
        off += (bytes + 255) & ~(size_t)255;
        return at;
    };
    L.y = take(N * 3 * sizeof(float));
    L.xprev = take(N * 3 * sizeof(float));
    for (auto& o : L.it) o = take(N * 3 * sizeof(double));
    L.pin_map = take(N * sizeof(int));
    L.fr = take(T * 9 * sizeof(float));
    L.fd = take(T * 9 * sizeof(float));
    L.bp = take(E * 3 * sizeof(float));
    L.bd = take(E * 3 * sizeof(float));
    L.vt_ptr = take((N + 1) * sizeof(int));
    L.vt_idx = take(T * 3 * sizeof(int));
    L.ve_ptr = take((N + 1) * sizeof(int));
    L.ve_idx = take(E * 4 * sizeof(int));
    L.cursor = take(N * sizeof(int));
    L.total = off;
    return L;
}

// CSR of vertex -> incident slots (element * arity + corner), ascending, on
// the device.
void enqueue_incidence(const int* elems, int nslots, int N, int* ptr, int* idx, int* cursor,
                       hipStream_t st)
{
    hipLaunchKernelGGL(k_fill_i32, dim3(grid_for(N)), dim3(kBlock), 0, st, cursor, N, 0);
    if (nslots > 0)
        hipLaunchKernelGGL(k_count_slots, dim3(grid_for(nslots)), dim3(kBlock), 0, st, elems,
                           nslots, N, cursor);
    hipLaunchKernelGGL(k_scan_counts, dim3(1), dim3(kScanThreads), 0, st, cursor, N, ptr);
    hipLaunchKernelGGL(k_fill_i32, dim3(grid_for(N)), dim3(kBlock), 0, st, cursor, N, 0);
    if (nslots > 0)
        hipLaunchKernelGGL(k_place_slots, dim3(grid_for(nslots)), dim3(kBlock), 0, st, elems,
                           nslots, N, ptr, cursor, idx);
    hipLaunchKernelGGL(k_sort_lists, dim3(grid_for(N)), dim3(kBlock), 0, st, ptr, idx, N);
}

}  // namespace

extern "C" int clh_step_noalloc(const clh_config* cfg, int num_verts, int num_tris, int num_bend,
                                const clh_step_args* a, void* stream)
{
    if (!cfg || !a) return fail(CLH_ERR_INVALID, "clh_step_noalloc: null pointer");
    if (num_verts < 0 || num_tris < 0 || num_bend < 0)
        return fail(CLH_ERR_INVALID, "clh_step_noalloc: negative count");
    if (cfg->mode != CLH_MODE_REF_CUDA && cfg->mode != CLH_MODE_PD)
        return fail(CLH_ERR_INVALID, "clh_step_noalloc: unknown mode");
    if (cfg->max_iterations < 0 || cfg->max_iterations > kMaxIterations)
        return fail(CLH_ERR_INVALID, "clh_step_noalloc: max_iterations out of range");
    if (!(cfg->dt > 0.0f)) return fail(CLH_ERR_INVALID, "clh_step_noalloc: dt must be > 0");
    if (!(cfg->relax >= 0.0f && cfg->relax < 2.0f))
        return fail(CLH_ERR_INVALID,
                    "clh_step_noalloc: relax must be in [0, 2) (no CLH_RELAX_AUTO: its bound "
                    "takes a reduction buffer)");
    if (cfg->mode == CLH_MODE_PD && cfg->tolerance > 0.0f)
        return fail(CLH_ERR_UNSUPPORTED,
                    "clh_step_noalloc: fixed-length frames only (tolerance must be 0)");
    if (a->d_trace) return fail(CLH_ERR_UNSUPPORTED, "clh_step_noalloc: no iterate trace");
    const NoAllocLayout L = noalloc_layout((size_t)num_verts, (size_t)num_tris, (size_t)num_bend);
    if (L.total > kNoAllocScratchBytes)
        return fail(CLH_ERR_UNSUPPORTED,
                    "clh_step_noalloc: the mesh needs " + std::to_string(L.total) +
                        " bytes of scratch, more than the " +
                        std::to_string(kNoAllocScratchBytes) + " bytes of the static array");
    // the same argument checks as clh_solver_step
    CLH_TRY(validate_step_args(num_verts, a));
    if (num_tris > 0 && a->d_stretch_k && !a->d_tris)
        return fail(CLH_ERR_INVALID, "clh_step_noalloc: null d_tris with num_tris > 0");
    if (num_verts == 0) return CLH_OK;

    void* base = nullptr;
    CLH_TRY(scratch_address(&base));
    unsigned char* sb = static_cast<unsigned char*>(base);
    float* y = reinterpret_cast<float*>(sb + L.y);
    float* xprev = reinterpret_cast<float*>(sb + L.xprev);
    void* it[kNoAllocBufs];
    for (int i = 0; i < kNoAllocBufs; ++i) it[i] = sb + L.it[i];
    int* pin_map = reinterpret_cast<int*>(sb + L.pin_map);
    float* fr = reinterpret_cast<float*>(sb + L.fr);
    float* fd = reinterpret_cast<float*>(sb + L.fd);
    float* bp = reinterpret_cast<float*>(sb + L.bp);
    float* bd = reinterpret_cast<float*>(sb + L.bd);
    int* vt_ptr = reinterpret_cast<int*>(sb + L.vt_ptr);
    int* vt_idx = reinterpret_cast<int*>(sb + L.vt_idx);
    int* ve_ptr = reinterpret_cast<int*>(sb + L.ve_ptr);
    int* ve_idx = reinterpret_cast<int*>(sb + L.ve_idx);
    int* cursor = reinterpret_cast<int*>(sb + L.cursor);

    const hipStream_t st = static_cast<hipStream_t>(stream);
    const clh_config& c = *cfg;
    const int N = num_verts, T = num_tris;
    const bool use_stretch = T > 0 && a->d_stretch_k != nullptr;
    const bool use_bend = num_bend > 0 && a->d_bend_quads != nullptr;
    const int E = use_bend ? num_bend : 0;
    const int pd = c.mode == CLH_MODE_PD ? 1 : 0;
    const float h2 = c.dt * c.dt;

    // incidence (an empty list where an element kind is off)
    enqueue_incidence(a->d_tris, use_stretch ? 3 * T : 0, N, vt_ptr, vt_idx, cursor, st);
    enqueue_incidence(a->d_bend_quads, 4 * E, N, ve_ptr, ve_idx, cursor, st);

    // prologue: as enqueue_prologue's unfused branch
    const float acc[3] = {c.ext_accel[0], c.gravity + c.ext_accel[1], c.ext_accel[2]};
    hipLaunchKernelGGL(k_prologue, dim3(grid_for(N)), dim3(kBlock), 0, st, a->d_pos, a->d_vel,
                       a->d_init_guess, a->d_cloth_accel, xprev, y, it[0], pin_map, N, c.dt, h2,
                       acc[0], acc[1], acc[2], h2 * acc[0], h2 * acc[1], h2 * acc[2], pd);
    if (a->num_pinned > 0)
        hipLaunchKernelGGL(k_pin_scatter, dim3(grid_for(a->num_pinned)), dim3(kBlock), 0, st,
                           a->d_pin_idx, a->d_pin_target, xprev, pin_map, it[0], a->num_pinned, N,
                           pd);

    // iterations: as enqueue_iteration's unfused branch, with the weights of
    // enqueue_fixed_frame
    float omega = 1.0f;
    for (int k = 0; k < c.max_iterations; ++k) {
        omega = chebyshev_omega(c, k, omega);
        const bool blend = pd && c.use_chebyshev && k > 0 && omega != 1.0f;
        const void* xk = it[k % kNoAllocBufs];
        if (use_stretch) {
            if (pd)
                hipLaunchKernelGGL(k_triangles<CLH_MODE_PD>, dim3(grid_for(T)), dim3(kBlock), 0, st,
                                   xprev, xk, a->d_tris, a->d_Dm_inv, a->d_rest_area,
                                   a->d_stretch_k, fr, fd, T, N, h2, nullptr);
            else
                hipLaunchKernelGGL(k_triangles<CLH_MODE_REF_CUDA>, dim3(grid_for(T)), dim3(kBlock),
                                   0, st, xprev, xk, a->d_tris, a->d_Dm_inv, a->d_rest_area,
                                   a->d_stretch_k, fr, fd, T, N, h2, nullptr);
        }
        if (use_bend) {
            if (pd)
                hipLaunchKernelGGL(k_hinges<CLH_MODE_PD>, dim3(grid_for(E)), dim3(kBlock), 0, st,
                                   xprev, xk, a->d_bend_quads, a->d_bend_w, a->d_bend_n, bp, bd, E,
                                   N, nullptr);
            else
                hipLaunchKernelGGL(k_hinges<CLH_MODE_REF_CUDA>, dim3(grid_for(E)), dim3(kBlock), 0,
                                   st, xprev, xk, a->d_bend_quads, a->d_bend_w, a->d_bend_n, bp, bd,
                                   E, N, nullptr);
        }
        VertexArgs va;
        va.base = xprev;
        va.x = xk;
        va.xkm1 = it[(k + kNoAllocBufs - 1) % kNoAllocBufs];
        va.out = it[(k + 1) % kNoAllocBufs];
        va.y = y;
        va.mass = a->d_mass;
        va.diag = a->d_jacobi_diag;
        va.pin_map = pin_map;
        va.pin_target = a->d_pin_target;
        va.vt_ptr = vt_ptr;
        va.vt_idx = vt_idx;
        va.ve_ptr = ve_ptr;
        va.ve_idx = ve_idx;
        va.fr = fr;
        va.fd = fd;
        va.bp = bp;
        va.bd = bd;
        va.bk = a->d_bend_k;
        va.bw = a->d_bend_w;
        va.partials = nullptr;
        va.N = N;
        va.use_stretch = use_stretch ? 1 : 0;
        va.use_bend = use_bend ? 1 : 0;
        va.h2 = h2;
        va.omega = omega;
        va.blend = blend ? 1 : 0;
        va.ctl = nullptr;
        va.par = k & 1;
        va.relax = c.relax;
        if (pd)
            hipLaunchKernelGGL(k_vertex_pd, dim3(grid_for(N)), dim3(kBlock), 0, st, va);
        else
            hipLaunchKernelGGL(k_vertex_ref, dim3(grid_for(N)), dim3(kBlock), 0, st, va);
    }
    hipLaunchKernelGGL(k_epilogue, dim3(grid_for(N)), dim3(kBlock), 0, st,
                       it[c.max_iterations % kNoAllocBufs], xprev, pin_map, a->d_pin_target,
                       a->d_pos, a->d_vel, N, (1.0f - c.damping) / c.dt, pd);
    return hip_status(hipGetLastError(), "clh_step_noalloc launch");
}
