// Learned warm start of the PD frame: per-vertex features, a BF16 MLP on
// MFMA (v_mfma_f32_16x16x32_bf16 for the 64-wide layers, 16x16x16 where K is
// 16; FP32 accumulation) for inference and training, and Adam on FP32 master
// weights.
//
// Prediction of the frame's converged positions x_{n+1} from the previous
// positions x_n, x_{n-1}, x_{n-2}:
//   x^ = 3 x_n - 3 x_{n-1} + x_{n-2} + L * s_o * MLP(f)
// i.e. the second-order extrapolation (constant acceleration) plus a learned
// change of acceleration. L = h^2 g0 (g0 = 9.81 m/s^2) is the length unit of
// the features, s_o the output scale (the spread of the training targets).
//
// Features (16 per vertex, in units of L per frame and of g0):
//   u = (x_n - x_{n-1}) / L, a = ((x_n - x_{n-1}) - (x_{n-1} - x_{n-2})) / L,
//   the external acceleration / g0, the pin flag, and the 1-ring means of u
//   and a minus the vertex's own (neighbours: the triangles' edges).
// Network: h1 = SiLU(W1 f + b1), h2 = SiLU(W2 h1 + b2), o = W3 h2 + b3, with
// W1 64x16, W2 64x64, W3 3x64 (stored as 16x64, rows 3.. zero).
//
// MFMA layout (16x16x16, 64 lanes, lane l = 16 q + r): the A operand holds
// A[r][4q..4q+3], the B operand B[4q..4q+3][r], the accumulator D[4q..4q+3][r].
// The layers are evaluated transposed, Z^T = W X^T, with weights as A and a
// group of 16 vertices as the columns: the accumulator of one layer (hidden
// units 4q..4q+3 of vertex r) is then exactly the B operand of the next, so
// the forward chain moves no data between lanes (the 16x16x32 steps join two
// such 4-wide blocks per lane, for A and B alike). Weight gradients reduce
// over vertices, which the operands hold along K; their inputs go through a
// per-wave LDS transpose.
#include "clh_internal.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

namespace clh {
namespace {

constexpr int kF = 16;   // features
constexpr int kH = 64;   // hidden width
constexpr int kO = 16;   // output rows of W3 (kOut used)
constexpr int kOut = 3;
// flat FP32 parameters
constexpr int oW1 = 0, oB1 = oW1 + kH * kF, oW2 = oB1 + kH, oB2 = oW2 + kH * kH, oW3 = oB2 + kH,
              oB3 = oW3 + kO * kH, kParams = oB3 + kO;
// BF16 weight copies: row-major W1, W2, W3 (forward A operands) and the
// transposes of W2, W3 (backward A operands)
constexpr int qW1 = 0, qW2 = qW1 + kH * kF, qW3 = qW2 + kH * kH, qW2T = qW3 + kO * kH,
              qW3T = qW2T + kH * kH, kQ = qW3T + kH * kO;
constexpr float kG0 = 9.81f;
constexpr int kWaves = 4;  // waves per workgroup of the MLP kernels
constexpr int kMaxGradWaves = 1024;

typedef short bf4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

typedef short s8v __attribute__((ext_vector_type(8)));
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));

// 16x16x16 product: lane l holds row (A) or column (B) l % 16 and the four K
// entries 4 (l / 16) + j. Used where K is 16 (the 16 features, the 16 samples
// of a weight-gradient tile).
__device__ __forceinline__ f4 mfma16(bf4 a, bf4 b, f4 c)
{
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

// gfx950's 16x16x32 product for the K = 64 layers, in two steps instead of
// four. Each lane joins two of its 4-wide blocks, cat8(x[2c], x[2c + 1]): its
// eight K entries are then units 32c + 4q + j and 32c + 16 + 4q + j (q = l / 16,
// j < 4). A and B are joined the same way, so the pairs multiplied are those
// of the 16x16x16 chain; only the order of the FP32 sum differs.
__device__ __forceinline__ bf8 cat8(bf4 a, bf4 b)
{
    const s8v s = __builtin_shufflevector(a, b, 0, 1, 2, 3, 4, 5, 6, 7);
    return __builtin_bit_cast(bf8, s);
}
__device__ __forceinline__ f4 mfma32(bf8 a, bf8 b, f4 c)
{
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// FP32 -> BF16, round to nearest even (finite inputs; NaN stays NaN).
__host__ __device__ __forceinline__ unsigned short f2bf(float f)
{
    unsigned u;
    __builtin_memcpy(&u, &f, sizeof(u));
    if ((u & 0x7F800000u) == 0x7F800000u) return (unsigned short)((u >> 16) | ((u & 0xFFFFu) ? 0x40u : 0u));
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}
__device__ __forceinline__ bf4 pack4(f4 v)
{
    bf4 o;
    o[0] = (short)f2bf(v[0]);
    o[1] = (short)f2bf(v[1]);
    o[2] = (short)f2bf(v[2]);
    o[3] = (short)f2bf(v[3]);
    return o;
}
// SiLU and its derivative from the hardware exponential and reciprocal
// (v_exp_f32, v_rcp_f32; about 1e-7 relative, far below the BF16 rounding of
// the activations, 4e-3). With the full-precision expf and division the
// forward over 4.2 M vertices took 516 us against 279. Large |z| saturates
// correctly: e = inf gives 1 / inf = 0.
__device__ __forceinline__ float sigmoid_fast(float z)
{
    return __builtin_amdgcn_rcpf(1.0f + __expf(-z));
}
__device__ __forceinline__ float silu(float z) { return z * sigmoid_fast(z); }
__device__ __forceinline__ float dsilu(float z)
{
    const float s = sigmoid_fast(z);
    return s * (1.0f + z * (1.0f - s));
}
__device__ __forceinline__ bf4 ld4(const unsigned short* p) { return *reinterpret_cast<const bf4*>(p); }
__device__ __forceinline__ f4 ldf4(const float* p)
{
    f4 o;
    o[0] = p[0];
    o[1] = p[1];
    o[2] = p[2];
    o[3] = p[3];
    return o;
}
// Orders this wave's LDS accesses (LDS operations of one wave complete in
// issue order; this keeps the compiler from moving them across).
__device__ __forceinline__ void wave_lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

struct v3 {
    float x, y, z;
};
__device__ __forceinline__ v3 ldv(const float* p, int i) { return v3{p[3 * i], p[3 * i + 1], p[3 * i + 2]}; }
__device__ __forceinline__ v3 operator-(v3 a, v3 b) { return v3{a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ v3 operator+(v3 a, v3 b) { return v3{a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ v3 operator*(v3 a, float s) { return v3{a.x * s, a.y * s, a.z * s}; }

// ---------------------------------------------------------------------------
// features
// ---------------------------------------------------------------------------

struct FeatArgs {
    const float* xm2;
    const float* xm1;
    const float* x;
    const float* xnext;        // converged x_{n+1}: targets (dataset), or null
    const float* cloth_accel;  // [C*3] or null
    const float* cloth_weight; // [C] sample weight of each cloth (targets), or null = 1
    const int* vcloth;         // [N] cloth of each vertex
    const unsigned char* pinned;
    const int* nb_ptr;
    const int* nb_idx;
    float acc0, acc1, acc2;    // uniform acceleration (gravity + wind), m/s^2
    int   use_ext;             // external acceleration in features 6..8 (else 0)
    float inv_L;               // 1 / (h^2 g0)
    float inv_so;              // 1 / output scale
    int N;
    unsigned short* F;         // [N*16]
    float4* T;                 // [N] (target / s_o, weight) or null
    float* extrap;             // [N*3] second-order extrapolation, or null
};

__global__ void __launch_bounds__(256) k_ws_features(FeatArgs a)
{
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= a.N) return;
    const v3 x = ldv(a.x, v), xm1 = ldv(a.xm1, v), xm2 = ldv(a.xm2, v);
    const v3 d1 = x - xm1, d0 = xm1 - xm2;
    const v3 u = d1 * a.inv_L;
    const v3 ac = (d1 - d0) * a.inv_L;
    v3 su = v3{0.0f, 0.0f, 0.0f}, sa = v3{0.0f, 0.0f, 0.0f};
    const int j0 = a.nb_ptr[v], j1 = a.nb_ptr[v + 1];
    for (int j = j0; j < j1; ++j) {
        const int w = a.nb_idx[j];
        const v3 y = ldv(a.x, w), ym1 = ldv(a.xm1, w), ym2 = ldv(a.xm2, w);
        const v3 e1 = y - ym1, e0 = ym1 - ym2;
        su = su + e1 * a.inv_L;
        sa = sa + (e1 - e0) * a.inv_L;
    }
    const int deg = j1 - j0;
    const float inv_deg = deg > 0 ? 1.0f / (float)deg : 0.0f;
    const v3 gu = deg > 0 ? su * inv_deg - u : v3{0.0f, 0.0f, 0.0f};
    const v3 ga = deg > 0 ? sa * inv_deg - ac : v3{0.0f, 0.0f, 0.0f};
    v3 ext = v3{a.acc0, a.acc1, a.acc2};
    if (a.cloth_accel) ext = ext + ldv(a.cloth_accel, a.vcloth[v]);
    if (!a.use_ext) ext = v3{0.0f, 0.0f, 0.0f};
    const float pin = a.pinned[v] ? 1.0f : 0.0f;
    const float f[kF] = {u.x,  u.y,  u.z,  ac.x, ac.y, ac.z, ext.x / kG0, ext.y / kG0,
                         ext.z / kG0, pin, gu.x, gu.y, gu.z, ga.x, ga.y, ga.z};
    unsigned short* o = a.F + (size_t)v * kF;
#pragma unroll
    for (int k = 0; k < kF; ++k) o[k] = f2bf(f[k]);
    if (a.T) {
        const v3 xn = ldv(a.xnext, v);
        const v3 d2 = xn - x;
        const v3 t = ((d2 - d1) * a.inv_L - ac) * a.inv_so;
        const float cw = a.cloth_weight ? a.cloth_weight[a.vcloth[v]] : 1.0f;
        a.T[v] = make_float4(t.x, t.y, t.z, a.pinned[v] ? 0.0f : cw);
    }
    if (a.extrap) {
        // x + d1 + (d1 - d0)
        const v3 e = x + d1 + (d1 - d0);
        a.extrap[3 * v + 0] = e.x;
        a.extrap[3 * v + 1] = e.y;
        a.extrap[3 * v + 2] = e.z;
    }
}

// ---------------------------------------------------------------------------
// forward (inference)
// ---------------------------------------------------------------------------

struct WaveWeights {
    bf4 w1[4];
    bf8 w2[4][2], w3[2];
    f4 b1[4], b2[4], b3;
};

__device__ __forceinline__ void load_forward_weights(const unsigned short* Wq, const float* P,
                                                     int r, int q, WaveWeights& w)
{
#pragma unroll
    for (int m = 0; m < 4; ++m) {
        w.w1[m] = ld4(Wq + qW1 + (m * 16 + r) * kF + 4 * q);
#pragma unroll
        for (int c = 0; c < 2; ++c) {
            const unsigned short* row = Wq + qW2 + (m * 16 + r) * kH + 32 * c + 4 * q;
            w.w2[m][c] = cat8(ld4(row), ld4(row + 16));
        }
        w.b1[m] = ldf4(P + oB1 + m * 16 + 4 * q);
        w.b2[m] = ldf4(P + oB2 + m * 16 + 4 * q);
    }
#pragma unroll
    for (int c = 0; c < 2; ++c) {
        const unsigned short* row = Wq + qW3 + r * kH + 32 * c + 4 * q;
        w.w3[c] = cat8(ld4(row), ld4(row + 16));
    }
    w.b3 = ldf4(P + oB3 + 4 * q);
}

// Forward of one group of 16 vertices: z1, z2 pre-activations (accumulator
// layout), h1, h2 activations (B-operand layout) and the output rows 4q.. of
// vertex r (bias included).
__device__ __forceinline__ f4 mlp_forward(const WaveWeights& w, bf4 fb, f4 z1[4], bf4 h1[4],
                                          f4 z2[4], bf4 h2[4])
{
    const f4 zero = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int m = 0; m < 4; ++m) {
        z1[m] = mfma16(w.w1[m], fb, zero) + w.b1[m];
        f4 h;
#pragma unroll
        for (int j = 0; j < 4; ++j) h[j] = silu(z1[m][j]);
        h1[m] = pack4(h);
    }
#pragma unroll
    for (int m = 0; m < 4; ++m) {
        f4 acc = zero;
#pragma unroll
        for (int c = 0; c < 2; ++c) acc = mfma32(w.w2[m][c], cat8(h1[2 * c], h1[2 * c + 1]), acc);
        z2[m] = acc + w.b2[m];
        f4 h;
#pragma unroll
        for (int j = 0; j < 4; ++j) h[j] = silu(z2[m][j]);
        h2[m] = pack4(h);
    }
    f4 o = zero;
#pragma unroll
    for (int c = 0; c < 2; ++c) o = mfma32(w.w3[c], cat8(h2[2 * c], h2[2 * c + 1]), o);
    return o + w.b3;
}

// Network correction L s_o o of every free vertex into corr (0 for pinned
// ones).
__global__ void __launch_bounds__(64 * kWaves)
k_ws_forward(const unsigned short* __restrict__ F, const unsigned short* __restrict__ Wq,
             const float* __restrict__ P, const unsigned char* __restrict__ pinned,
             float* __restrict__ corr, float out_len, int N)
{
    const int lane = threadIdx.x & 63, r = lane & 15, q = lane >> 4;
    const int wave = blockIdx.x * kWaves + (threadIdx.x >> 6);
    const int nwaves = gridDim.x * kWaves;
    WaveWeights w;
    load_forward_weights(Wq, P, r, q, w);
    const int ngroups = (N + 15) / 16;
    for (int g = wave; g < ngroups; g += nwaves) {
        const int v = g * 16 + r;
        bf4 fb = {0, 0, 0, 0};
        if (v < N) fb = ld4(F + (size_t)v * kF + 4 * q);
        f4 z1[4], z2[4];
        bf4 h1[4], h2[4];
        const f4 o = mlp_forward(w, fb, z1, h1, z2, h2);
        if (q == 0 && v < N) {
            const float s = pinned[v] ? 0.0f : out_len;
            corr[3 * v + 0] = s * o[0];
            corr[3 * v + 1] = s * o[1];
            corr[3 * v + 2] = s * o[2];
        }
    }
}

// One smoothing pass of the correction: the mean over the vertex and its
// 1-ring (pinned vertices keep 0), a low-pass filter that keeps the smooth
// part of the correction, which the iterations are slowest to find.
__global__ void __launch_bounds__(256)
k_ws_smooth(const float* __restrict__ in, float* __restrict__ out,
            const int* __restrict__ nb_ptr, const int* __restrict__ nb_idx,
            const unsigned char* __restrict__ pinned, int N)
{
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= N) return;
    if (pinned[v]) {
        out[3 * v + 0] = out[3 * v + 1] = out[3 * v + 2] = 0.0f;
        return;
    }
    v3 s = ldv(in, v);
    const int j0 = nb_ptr[v], j1 = nb_ptr[v + 1];
    for (int j = j0; j < j1; ++j) s = s + ldv(in, nb_idx[j]);
    const float inv = 1.0f / (float)(1 + j1 - j0);
    out[3 * v + 0] = s.x * inv;
    out[3 * v + 1] = s.y * inv;
    out[3 * v + 2] = s.z * inv;
}

// guess = extrap + corr for free vertices, x for pinned ones.
__global__ void __launch_bounds__(256)
k_ws_apply(const float* __restrict__ extrap, const float* __restrict__ corr,
           const float* __restrict__ x, const unsigned char* __restrict__ pinned,
           float* __restrict__ guess, int N)
{
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= N) return;
    for (int k = 0; k < 3; ++k)
        guess[3 * v + k] = pinned[v] ? x[3 * v + k] : extrap[3 * v + k] + corr[3 * v + k];
}

// guess = extrap (free) or x (pinned): the extrapolation without the network.
__global__ void __launch_bounds__(256)
k_ws_extrap_only(const float* __restrict__ extrap, const float* __restrict__ x,
                 const unsigned char* __restrict__ pinned, float* __restrict__ guess, int N)
{
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= N) return;
    const float* s = pinned[v] ? x : extrap;
    guess[3 * v + 0] = s[3 * v + 0];
    guess[3 * v + 1] = s[3 * v + 1];
    guess[3 * v + 2] = s[3 * v + 2];
}

// step = guess - x for free vertices, 0 for pinned ones.
__global__ void __launch_bounds__(256)
k_ws_step(const float* __restrict__ guess, const float* __restrict__ x,
          const unsigned char* __restrict__ pinned, float* __restrict__ step, int N)
{
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= N) return;
    for (int k = 0; k < 3; ++k) step[3 * v + k] = pinned[v] ? 0.0f : guess[3 * v + k] - x[3 * v + k];
}

// guess = x + gain * step for free vertices, x for pinned ones.
__global__ void __launch_bounds__(256)
k_ws_apply_step(const float* __restrict__ step, const float* __restrict__ x,
                const unsigned char* __restrict__ pinned, float gain, float* __restrict__ guess,
                int N)
{
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= N) return;
    for (int k = 0; k < 3; ++k)
        guess[3 * v + k] = pinned[v] ? x[3 * v + k] : fmaf(gain, step[3 * v + k], x[3 * v + k]);
}

// ---------------------------------------------------------------------------
// training: loss and gradient
// ---------------------------------------------------------------------------

__host__ __device__ __forceinline__ unsigned long long mix64(unsigned long long z)
{
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

struct GradArgs {
    const unsigned short* F;  // dataset features [S*16]
    const float4* T;          // dataset targets [S]
    const unsigned short* Wq;
    const float* P;
    long long first;          // contiguous mode: first sample (multiple of 16)
    long long nblocks;        // sampled mode: 16-sample blocks in the dataset
    unsigned long long seed;  // sampled mode when non-zero
    int ngroups;              // groups of 16 samples in this minibatch
    float inv_norm;           // 1 / (3 * samples)
    float* partial;           // [waves][kParams + 1]
};

// Loss sum_s w_s |o_s - t_s|^2 * inv_norm and its gradient with respect to
// every parameter, per wave, written to that wave's partial slot.
__global__ void __launch_bounds__(64 * kWaves) k_ws_grad(GradArgs a)
{
    // per-wave LDS: [hidden or feature][16 samples] tiles for the transposes
    __shared__ unsigned short lds[kWaves][(4 * kH + 2 * 16) * 16];
    const int lane = threadIdx.x & 63, r = lane & 15, q = lane >> 4;
    const int wv = threadIdx.x >> 6;
    const int wave = blockIdx.x * kWaves + wv;
    const int nwaves = gridDim.x * kWaves;
    unsigned short* Ft = lds[wv];            // [16 feat][16]
    unsigned short* dOt = Ft + 16 * 16;      // [16 out][16]
    unsigned short* H1t = dOt + 16 * 16;     // [64][16]
    unsigned short* H2t = H1t + kH * 16;
    unsigned short* dZ1t = H2t + kH * 16;
    unsigned short* dZ2t = dZ1t + kH * 16;

    WaveWeights w;
    load_forward_weights(a.Wq, a.P, r, q, w);
    bf8 w2t[4][2];
    bf4 w3t[4];
#pragma unroll
    for (int m = 0; m < 4; ++m) {
#pragma unroll
        for (int c = 0; c < 2; ++c) {
            const unsigned short* row = a.Wq + qW2T + (m * 16 + r) * kH + 32 * c + 4 * q;
            w2t[m][c] = cat8(ld4(row), ld4(row + 16));
        }
        w3t[m] = ld4(a.Wq + qW3T + (m * 16 + r) * kO + 4 * q);
    }
    const f4 zero = {0.0f, 0.0f, 0.0f, 0.0f};
    f4 gW1[4], gW2[4][4], gW3[4], gb1[4], gb2[4], gb3 = zero;
#pragma unroll
    for (int m = 0; m < 4; ++m) {
        gW1[m] = zero;
        gW3[m] = zero;
        gb1[m] = zero;
        gb2[m] = zero;
#pragma unroll
        for (int k = 0; k < 4; ++k) gW2[m][k] = zero;
    }
    float loss = 0.0f;

    for (int g = wave; g < a.ngroups; g += nwaves) {
        long long s0 = a.first + 16ll * g;
        if (a.seed) s0 = 16ll * (long long)(mix64(a.seed ^ ((unsigned long long)g << 20)) %
                                            (unsigned long long)a.nblocks);
        const bf4 fb = ld4(a.F + (size_t)(s0 + r) * kF + 4 * q);
        f4 z1[4], z2[4];
        bf4 h1[4], h2[4];
        const f4 o = mlp_forward(w, fb, z1, h1, z2, h2);

        // dL/do for the used output rows (lanes q == 0 hold rows 0..3)
        f4 d = zero;
        if (q == 0) {
            const float4 t = a.T[s0 + r];
            const float e0 = o[0] - t.x, e1 = o[1] - t.y, e2 = o[2] - t.z;
            loss += t.w * (e0 * e0 + e1 * e1 + e2 * e2);
            const float c = 2.0f * t.w * a.inv_norm;
            d[0] = c * e0;
            d[1] = c * e1;
            d[2] = c * e2;
        }
        gb3 += d;
        const bf4 db = pack4(d);
        f4 dz2[4], dz1[4];
        bf4 dz2b[4], dz1b[4];
#pragma unroll
        for (int m = 0; m < 4; ++m) {
            const f4 dh = mfma16(w3t[m], db, zero);
#pragma unroll
            for (int j = 0; j < 4; ++j) dz2[m][j] = dh[j] * dsilu(z2[m][j]);
            gb2[m] += dz2[m];
            dz2b[m] = pack4(dz2[m]);
        }
#pragma unroll
        for (int m = 0; m < 4; ++m) {
            f4 dh = zero;
#pragma unroll
            for (int c = 0; c < 2; ++c) dh = mfma32(w2t[m][c], cat8(dz2b[2 * c], dz2b[2 * c + 1]), dh);
#pragma unroll
            for (int j = 0; j < 4; ++j) dz1[m][j] = dh[j] * dsilu(z1[m][j]);
            gb1[m] += dz1[m];
            dz1b[m] = pack4(dz1[m]);
        }

        // transposes: element (row 4q+j, sample r) -> tile[row][r]
        wave_lds_sync();
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            Ft[(4 * q + j) * 16 + r] = (unsigned short)fb[j];
            dOt[(4 * q + j) * 16 + r] = (unsigned short)db[j];
        }
#pragma unroll
        for (int m = 0; m < 4; ++m)
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int row = (m * 16 + 4 * q + j) * 16 + r;
                H1t[row] = (unsigned short)h1[m][j];
                H2t[row] = (unsigned short)h2[m][j];
                dZ1t[row] = (unsigned short)dz1b[m][j];
                dZ2t[row] = (unsigned short)dz2b[m][j];
            }
        wave_lds_sync();

        // weight gradients: sum over the 16 samples (K)
        const bf4 aO = ld4(dOt + r * 16 + 4 * q);
#pragma unroll
        for (int n = 0; n < 4; ++n) gW3[n] = mfma16(aO, ld4(H2t + (n * 16 + r) * 16 + 4 * q), gW3[n]);
        const bf4 bF = ld4(Ft + r * 16 + 4 * q);
#pragma unroll
        for (int m = 0; m < 4; ++m) {
            gW1[m] = mfma16(ld4(dZ1t + (m * 16 + r) * 16 + 4 * q), bF, gW1[m]);
            const bf4 aZ = ld4(dZ2t + (m * 16 + r) * 16 + 4 * q);
#pragma unroll
            for (int n = 0; n < 4; ++n)
                gW2[m][n] = mfma16(aZ, ld4(H1t + (n * 16 + r) * 16 + 4 * q), gW2[m][n]);
        }
    }

    // bias gradients: sum the 16 samples (lanes r) of each row quad
#pragma unroll
    for (int off = 1; off < 16; off <<= 1) {
#pragma unroll
        for (int j = 0; j < 4; ++j) {
#pragma unroll
            for (int m = 0; m < 4; ++m) {
                gb1[m][j] += __shfl_xor(gb1[m][j], off, 64);
                gb2[m][j] += __shfl_xor(gb2[m][j], off, 64);
            }
            gb3[j] += __shfl_xor(gb3[j], off, 64);
        }
    }
    for (int off = 32; off > 0; off >>= 1) loss += __shfl_xor(loss, off, 64);

    float* out = a.partial + (size_t)wave * (kParams + 1);
#pragma unroll
    for (int m = 0; m < 4; ++m)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int row = m * 16 + 4 * q + j;
            out[oW1 + row * kF + r] = gW1[m][j];
#pragma unroll
            for (int n = 0; n < 4; ++n) out[oW2 + row * kH + n * 16 + r] = gW2[m][n][j];
            out[oW3 + (4 * q + j) * kH + m * 16 + r] = gW3[m][j];
            if (r == 0) {
                out[oB1 + row] = gb1[m][j];
                out[oB2 + row] = gb2[m][j];
            }
        }
    if (r == 0)
#pragma unroll
        for (int j = 0; j < 4; ++j) out[oB3 + 4 * q + j] = gb3[j];
    if (lane == 0) out[kParams] = loss * a.inv_norm;
}

// grad[p] = sum over waves of partial[w][p] in a fixed order: the waves are
// cut into kRedSlices contiguous runs, each summed in wave order by one wave
// of the workgroup (64 parameters per workgroup, one per lane, so every load
// of a wave is 256 contiguous bytes), and the run sums are added pairwise,
// (s_0 + s_8) + ..., in a fixed tree.
constexpr int kRedSlices = 16;
__global__ void __launch_bounds__(64 * kRedSlices)
k_ws_reduce(const float* __restrict__ partial, int nwaves, float* __restrict__ grad)
{
    __shared__ float run[kRedSlices][64];
    const int lane = threadIdx.x & 63, sl = threadIdx.x >> 6;
    const int p = blockIdx.x * 64 + lane;
    const int per = (nwaves + kRedSlices - 1) / kRedSlices;
    const int w0 = sl * per, w1 = min(nwaves, w0 + per);
    float s = 0.0f;
    if (p <= kParams) {
#pragma unroll 8
        for (int w = w0; w < w1; ++w) s += partial[(size_t)w * (kParams + 1) + p];
    }
    run[sl][lane] = s;
    __syncthreads();
    if (sl == 0 && p <= kParams) {
        float t[kRedSlices];
#pragma unroll
        for (int i = 0; i < kRedSlices; ++i) t[i] = run[i][lane];
#pragma unroll
        for (int h = kRedSlices / 2; h > 0; h >>= 1)
#pragma unroll
            for (int i = 0; i < h; ++i) t[i] = t[i] + t[i + h];
        grad[p] = t[0];
    }
}

// BF16 copies of one master weight (row-major and, for W2 and W3, transposed).
__device__ __forceinline__ void pack_param(int p, float val, unsigned short* Wq)
{
    const unsigned short b = f2bf(val);
    if (p < oB1) {
        Wq[qW1 + p] = b;
    } else if (p >= oW2 && p < oB2) {
        const int i = p - oW2, o = i / kH, c = i % kH;
        Wq[qW2 + i] = b;
        Wq[qW2T + c * kH + o] = b;
    } else if (p >= oW3 && p < oB3) {
        const int i = p - oW3, o = i / kH, c = i % kH;
        Wq[qW3 + i] = b;
        Wq[qW3T + c * kO + o] = b;
    }
}

__global__ void __launch_bounds__(256) k_ws_pack(const float* __restrict__ P, unsigned short* __restrict__ Wq)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < kParams) pack_param(p, P[p], Wq);
}

// Unused output rows (3..15 of W3 and b3) stay zero.
__device__ __forceinline__ bool frozen(int p)
{
    return (p >= oW3 + kOut * kH && p < oB3) || (p >= oB3 + kOut);
}

__global__ void __launch_bounds__(256)
k_ws_adam(float* __restrict__ P, float* __restrict__ m, float* __restrict__ v,
          const float* __restrict__ grad, unsigned short* __restrict__ Wq, float lr, float b1,
          float b2, float eps, float c1, float c2)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= kParams || frozen(p)) return;
    const float g = grad[p];
    const float mm = b1 * m[p] + (1.0f - b1) * g;
    const float vv = b2 * v[p] + (1.0f - b2) * g * g;
    m[p] = mm;
    v[p] = vv;
    const float val = P[p] - lr * (mm / c1) / (sqrtf(vv / c2) + eps);
    P[p] = val;
    pack_param(p, val, Wq);
}

// Records the loss of a training step.
__global__ void k_ws_store_loss(const float* __restrict__ grad, float* __restrict__ losses, int i)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) losses[i] = grad[kParams];
}

unsigned blocks256(long long n) { return (unsigned)((n + 255) / 256); }

}  // namespace
}  // namespace clh

using namespace clh;

struct clh_ws {
    int N = 0, C = 1;
    float dt = 0.0f;
    float acc[3] = {0.0f, 0.0f, 0.0f};
    float so = 1.0f;  // output scale
    int use_ext = 0;  // external acceleration among the features
    hipStream_t stream = nullptr;
    DevBuf<int> nb_ptr, nb_idx, vcloth;
    DevBuf<unsigned char> pinned;
    DevBuf<float> P, m, v, grad, partial, losses, extrap, corr[2];
    int smooth_passes = 0;
    int step_passes = 0;      // 1-ring smoothing passes of the whole predicted step
    float step_gain = 1.0f;   // factor on the (smoothed) predicted step
    DevBuf<unsigned short> Wq, Finf;
    DevBuf<unsigned short> F;  // dataset features
    DevBuf<float4> T;          // dataset targets
    long long size = 0, cap = 0;
    long long adam_t = 0;
    unsigned long long sel = 0;  // minibatch sampling stream
};

namespace {

float inv_L(const clh_ws* w) { return 1.0f / (w->dt * w->dt * kG0); }

int launch_features(clh_ws* w, const float* xm2, const float* xm1, const float* x,
                    const float* xnext, const float* cloth_accel, unsigned short* F, float4* T,
                    float* extrap, const float* cloth_weight = nullptr)
{
    // training targets are stored divided by the output scale
    if (T && !(w->so > 0.0f))
        return fail(CLH_ERR_INVALID, "warm start: training frames need an output scale > 0");
    FeatArgs a;
    a.cloth_weight = cloth_weight;
    a.xm2 = xm2;
    a.xm1 = xm1;
    a.x = x;
    a.xnext = xnext;
    a.cloth_accel = cloth_accel;
    a.vcloth = w->vcloth.ptr;
    a.pinned = w->pinned.ptr;
    a.nb_ptr = w->nb_ptr.ptr;
    a.nb_idx = w->nb_idx.ptr;
    a.acc0 = w->acc[0];
    a.acc1 = w->acc[1];
    a.acc2 = w->acc[2];
    a.use_ext = w->use_ext;
    a.inv_L = inv_L(w);
    a.inv_so = 1.0f / w->so;
    a.N = w->N;
    a.F = F;
    a.T = T;
    a.extrap = extrap;
    hipLaunchKernelGGL(k_ws_features, dim3(blocks256(w->N)), dim3(256), 0, w->stream, a);
    return hip_status(hipGetLastError(), "warm-start features launch");
}

int pack(clh_ws* w)
{
    hipLaunchKernelGGL(k_ws_pack, dim3(blocks256(kParams)), dim3(256), 0, w->stream, w->P.ptr,
                       w->Wq.ptr);
    return hip_status(hipGetLastError(), "warm-start pack launch");
}

// Gradient of the minibatch into w->grad (loss at grad[kParams]).
int launch_grad(clh_ws* w, long long first, int count, unsigned long long seed)
{
    if (count <= 0 || count % 16) return fail(CLH_ERR_INVALID, "warm start: minibatch must be a positive multiple of 16");
    const int ngroups = count / 16;
    const int nwaves = std::min(kMaxGradWaves, ngroups);
    const int blocks = (nwaves + kWaves - 1) / kWaves;
    GradArgs a;
    a.F = w->F.ptr;
    a.T = w->T.ptr;
    a.Wq = w->Wq.ptr;
    a.P = w->P.ptr;
    a.first = first;
    a.nblocks = w->size / 16;
    a.seed = seed;
    a.ngroups = ngroups;
    a.inv_norm = 1.0f / (3.0f * (float)count);
    a.partial = w->partial.ptr;
    hipLaunchKernelGGL(k_ws_grad, dim3(blocks), dim3(64 * kWaves), 0, w->stream, a);
    hipLaunchKernelGGL(k_ws_reduce, dim3((kParams + 1 + 63) / 64), dim3(64 * kRedSlices), 0,
                       w->stream, w->partial.ptr, blocks * kWaves, w->grad.ptr);
    return hip_status(hipGetLastError(), "warm-start gradient launch");
}

// Normal(0, 1) from SplitMix64 by Box-Muller.
struct HostRng {
    unsigned long long s;
    double uniform() { return (double)(mix64(s++) >> 11) * (1.0 / 9007199254740992.0); }
    double normal()
    {
        const double u1 = std::max(uniform(), 1e-300), u2 = uniform();
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
    }
};

}  // namespace

extern "C" int clh_ws_create(const clh_ws_desc* d, clh_ws** out)
{
    if (!d || !out) return fail(CLH_ERR_INVALID, "clh_ws_create: null pointer");
    *out = nullptr;
    if (d->num_verts <= 0 || d->num_tris < 0 || (d->num_tris > 0 && !d->d_tris) || !(d->dt > 0.0f) ||
        d->num_pinned < 0 || (d->num_pinned > 0 && !d->d_pin_idx) || d->num_cloths < 0 ||
        (d->num_cloths > 1 && !d->cloth_vert_offsets))
        return fail(CLH_ERR_INVALID, "clh_ws_create: bad description");
    const int N = d->num_verts, T = d->num_tris;
    std::vector<int> tris((size_t)T * 3), pins((size_t)d->num_pinned);
    if (T > 0) CLH_HIP(hipMemcpy(tris.data(), d->d_tris, tris.size() * sizeof(int), hipMemcpyDeviceToHost));
    if (d->num_pinned > 0)
        CLH_HIP(hipMemcpy(pins.data(), d->d_pin_idx, pins.size() * sizeof(int), hipMemcpyDeviceToHost));
    for (int v : tris) if (v < 0 || v >= N) return fail(CLH_ERR_INVALID, "clh_ws_create: triangle index out of range");
    for (int v : pins) if (v < 0 || v >= N) return fail(CLH_ERR_INVALID, "clh_ws_create: pin index out of range");
    const int C = std::max(1, d->num_cloths);
    std::vector<int> vcloth((size_t)N, 0);
    if (C > 1) {
        const int* o = d->cloth_vert_offsets;
        if (o[0] != 0 || o[C] != N) return fail(CLH_ERR_INVALID, "clh_ws_create: cloth offsets must run from 0 to num_verts");
        for (int c = 0; c < C; ++c) {
            if (o[c + 1] <= o[c]) return fail(CLH_ERR_INVALID, "clh_ws_create: empty cloth");
            for (int v = o[c]; v < o[c + 1]; ++v) vcloth[(size_t)v] = c;
        }
    }
    // neighbours: the other corners of the vertex's triangles, ascending
    std::vector<std::vector<int>> nb((size_t)N);
    for (int t = 0; t < T; ++t)
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j)
                if (i != j) nb[(size_t)tris[(size_t)t * 3 + i]].push_back(tris[(size_t)t * 3 + j]);
    std::vector<int> ptr((size_t)N + 1, 0), idx;
    for (int v = 0; v < N; ++v) {
        auto& l = nb[(size_t)v];
        std::sort(l.begin(), l.end());
        l.erase(std::unique(l.begin(), l.end()), l.end());
        l.erase(std::remove(l.begin(), l.end(), v), l.end());
        idx.insert(idx.end(), l.begin(), l.end());
        ptr[(size_t)v + 1] = (int)idx.size();
    }
    if (idx.empty()) idx.push_back(0);
    std::vector<unsigned char> pinned((size_t)N, 0);
    for (int v : pins) pinned[(size_t)v] = 1;

    clh_ws* w = new clh_ws();
    w->N = N;
    w->C = C;
    w->dt = d->dt;
    w->acc[0] = d->ext_accel[0];
    w->acc[1] = d->gravity + d->ext_accel[1];
    w->acc[2] = d->ext_accel[2];
    w->use_ext = d->ext_feature ? 1 : 0;
    int st = CLH_OK;
    auto up = [&](int s2) { if (st == CLH_OK) st = s2; };
    up(w->nb_ptr.upload(ptr));
    up(w->nb_idx.upload(idx));
    up(w->vcloth.upload(vcloth));
    up(w->pinned.upload(pinned));
    up(w->P.resize(kParams));
    up(w->m.resize(kParams));
    up(w->v.resize(kParams));
    up(w->grad.resize(kParams + 1));
    up(w->partial.resize((size_t)kMaxGradWaves * (kParams + 1)));
    up(w->Wq.resize(kQ));
    up(w->Finf.resize((size_t)N * kF));
    up(w->extrap.resize((size_t)N * 3));
    up(w->corr[0].resize((size_t)N * 3));
    up(w->corr[1].resize((size_t)N * 3));
    if (st != CLH_OK) {
        delete w;
        return st;
    }
    // initial weights: N(0, 1/fan_in) for W1 and W2, zero W3 and biases, so
    // the untrained network predicts the extrapolation itself
    std::vector<float> p((size_t)kParams, 0.0f);
    HostRng rng{d->seed ? d->seed : 1ull};
    for (int i = 0; i < kH * kF; ++i) p[(size_t)(oW1 + i)] = (float)(rng.normal() / std::sqrt((double)kF));
    for (int i = 0; i < kH * kH; ++i) p[(size_t)(oW2 + i)] = (float)(rng.normal() / std::sqrt((double)kH));
    w->sel = mix64(d->seed ^ 0x5E1EC7ull);
    st = hip_status(hipMemcpy(w->P.ptr, p.data(), p.size() * sizeof(float), hipMemcpyHostToDevice), "hipMemcpy");
    if (st == CLH_OK) st = hip_status(hipMemset(w->m.ptr, 0, kParams * sizeof(float)), "hipMemset");
    if (st == CLH_OK) st = hip_status(hipMemset(w->v.ptr, 0, kParams * sizeof(float)), "hipMemset");
    if (st == CLH_OK) st = hip_status(hipMemset(w->Wq.ptr, 0, kQ * sizeof(unsigned short)), "hipMemset");
    if (st == CLH_OK) st = pack(w);
    if (st == CLH_OK) st = hip_status(hipStreamSynchronize(w->stream), "hipStreamSynchronize");
    if (st != CLH_OK) {
        delete w;
        return st;
    }
    *out = w;
    return CLH_OK;
}

extern "C" int clh_ws_destroy(clh_ws* w)
{
    if (!w) return CLH_OK;
    (void)hipStreamSynchronize(w->stream);
    delete w;
    return CLH_OK;
}

extern "C" int clh_ws_set_stream(clh_ws* w, void* stream)
{
    if (!w) return fail(CLH_ERR_INVALID, "clh_ws_set_stream: null handle");
    CLH_HIP(hipStreamSynchronize(w->stream));
    w->stream = static_cast<hipStream_t>(stream);
    return CLH_OK;
}

extern "C" int clh_ws_param_count(void) { return kParams; }

extern "C" int clh_ws_get_params(clh_ws* w, float* out, int count, float* out_scale)
{
    if (!w || (count > 0 && !out)) return fail(CLH_ERR_INVALID, "clh_ws_get_params: null pointer");
    if (count != kParams && count != 0) return fail(CLH_ERR_INVALID, "clh_ws_get_params: count must equal clh_ws_param_count()");
    CLH_HIP(hipStreamSynchronize(w->stream));
    if (count) CLH_HIP(hipMemcpy(out, w->P.ptr, kParams * sizeof(float), hipMemcpyDeviceToHost));
    if (out_scale) *out_scale = w->so;
    return CLH_OK;
}

extern "C" int clh_ws_set_params(clh_ws* w, const float* in, int count, float out_scale)
{
    if (!w || (count > 0 && !in)) return fail(CLH_ERR_INVALID, "clh_ws_set_params: null pointer");
    if (count != kParams && count != 0) return fail(CLH_ERR_INVALID, "clh_ws_set_params: count must equal clh_ws_param_count()");
    // 0 turns the correction off (the prediction is then the extrapolation)
    if (!(out_scale >= 0.0f) || !std::isfinite(out_scale))
        return fail(CLH_ERR_INVALID, "clh_ws_set_params: output scale must be finite and >= 0");
    CLH_HIP(hipStreamSynchronize(w->stream));
    if (count) {
        std::vector<float> p(in, in + kParams);
        for (int i = oW3 + kOut * kH; i < oB3; ++i) p[(size_t)i] = 0.0f;  // unused rows
        for (int i = oB3 + kOut; i < kParams; ++i) p[(size_t)i] = 0.0f;
        CLH_HIP(hipMemcpy(w->P.ptr, p.data(), kParams * sizeof(float), hipMemcpyHostToDevice));
        CLH_TRY(pack(w));
    }
    w->so = out_scale;
    return CLH_OK;
}

extern "C" int clh_ws_predict(clh_ws* w, const float* xm2, const float* xm1, const float* x,
                              const float* cloth_accel, float* guess, int use_network)
{
    if (!w || !xm2 || !xm1 || !x || !guess) return fail(CLH_ERR_INVALID, "clh_ws_predict: null pointer");
    CLH_TRY(launch_features(w, xm2, xm1, x, nullptr, cloth_accel, w->Finf.ptr, nullptr, w->extrap.ptr));
    if (use_network) {
        const int ngroups = (w->N + 15) / 16;
        const int blocks = std::max(1, std::min(2048, (ngroups + kWaves - 1) / kWaves));
        hipLaunchKernelGGL(k_ws_forward, dim3(blocks), dim3(64 * kWaves), 0, w->stream,
                           w->Finf.ptr, w->Wq.ptr, w->P.ptr, w->pinned.ptr, w->corr[0].ptr,
                           w->so / inv_L(w), w->N);
        for (int i = 0; i < w->smooth_passes; ++i)
            hipLaunchKernelGGL(k_ws_smooth, dim3(blocks256(w->N)), dim3(256), 0, w->stream,
                               w->corr[i & 1].ptr, w->corr[(i + 1) & 1].ptr, w->nb_ptr.ptr,
                               w->nb_idx.ptr, w->pinned.ptr, w->N);
        hipLaunchKernelGGL(k_ws_apply, dim3(blocks256(w->N)), dim3(256), 0, w->stream,
                           w->extrap.ptr, w->corr[w->smooth_passes & 1].ptr, x, w->pinned.ptr,
                           guess, w->N);
    } else {
        hipLaunchKernelGGL(k_ws_extrap_only, dim3(blocks256(w->N)), dim3(256), 0, w->stream,
                           w->extrap.ptr, x, w->pinned.ptr, guess, w->N);
    }
    if (w->step_passes > 0 || w->step_gain != 1.0f) {
        // the whole step x_hat - x_n, smoothed and scaled (the correction
        // buffers are free once the guess is formed)
        hipLaunchKernelGGL(k_ws_step, dim3(blocks256(w->N)), dim3(256), 0, w->stream, guess, x,
                           w->pinned.ptr, w->corr[0].ptr, w->N);
        for (int i = 0; i < w->step_passes; ++i)
            hipLaunchKernelGGL(k_ws_smooth, dim3(blocks256(w->N)), dim3(256), 0, w->stream,
                               w->corr[i & 1].ptr, w->corr[(i + 1) & 1].ptr, w->nb_ptr.ptr,
                               w->nb_idx.ptr, w->pinned.ptr, w->N);
        hipLaunchKernelGGL(k_ws_apply_step, dim3(blocks256(w->N)), dim3(256), 0, w->stream,
                           w->corr[w->step_passes & 1].ptr, x, w->pinned.ptr, w->step_gain, guess,
                           w->N);
    }
    return hip_status(hipGetLastError(), "warm-start predict launch");
}

extern "C" int clh_ws_set_step_filter(clh_ws* w, int passes, float gain)
{
    if (!w || passes < 0 || passes > 1024 || !(gain >= 0.0f) || !std::isfinite(gain))
        return fail(CLH_ERR_INVALID,
                    "clh_ws_set_step_filter: passes must be in [0, 1024] and gain finite and >= 0");
    w->step_passes = passes;
    w->step_gain = gain;
    return CLH_OK;
}

extern "C" int clh_ws_set_smoothing(clh_ws* w, int passes)
{
    if (!w || passes < 0 || passes > 256) return fail(CLH_ERR_INVALID, "clh_ws_set_smoothing: passes must be in [0, 256]");
    w->smooth_passes = passes;
    return CLH_OK;
}

extern "C" int clh_ws_features(clh_ws* w, const float* xm2, const float* xm1, const float* x,
                               const float* cloth_accel, unsigned short* d_features)
{
    if (!w || !xm2 || !xm1 || !x || !d_features) return fail(CLH_ERR_INVALID, "clh_ws_features: null pointer");
    return launch_features(w, xm2, xm1, x, nullptr, cloth_accel, d_features, nullptr, nullptr);
}

extern "C" int clh_ws_dataset_reserve(clh_ws* w, long long samples)
{
    if (!w || samples < 0) return fail(CLH_ERR_INVALID, "clh_ws_dataset_reserve: bad arguments");
    const long long cap = (samples + 15) / 16 * 16;
    if (cap <= w->cap) return CLH_OK;
    CLH_HIP(hipStreamSynchronize(w->stream));
    DevBuf<unsigned short> F;
    DevBuf<float4> T;
    CLH_TRY(F.resize((size_t)cap * kF));
    CLH_TRY(T.resize((size_t)cap));
    if (w->size > 0) {
        CLH_HIP(hipMemcpy(F.ptr, w->F.ptr, (size_t)w->size * kF * sizeof(unsigned short), hipMemcpyDeviceToDevice));
        CLH_HIP(hipMemcpy(T.ptr, w->T.ptr, (size_t)w->size * sizeof(float4), hipMemcpyDeviceToDevice));
    }
    std::swap(w->F, F);
    std::swap(w->T, T);
    F.release();
    T.release();
    w->cap = cap;
    return CLH_OK;
}

extern "C" int clh_ws_dataset_add(clh_ws* w, const float* xm2, const float* xm1, const float* x,
                                  const float* xnext, const float* cloth_accel,
                                  const float* cloth_weight)
{
    if (!w || !xm2 || !xm1 || !x || !xnext) return fail(CLH_ERR_INVALID, "clh_ws_dataset_add: null pointer");
    const long long need = w->size + ((w->N + 15) / 16) * 16;
    if (need > w->cap) CLH_TRY(clh_ws_dataset_reserve(w, std::max(need, 2 * w->cap)));
    // every frame starts on a 16-sample block; the tail of the block is
    // padding with weight 0
    const long long pad = ((w->N + 15) / 16) * 16 - w->N;
    if (pad > 0) {
        CLH_HIP(hipMemsetAsync(w->F.ptr + (size_t)(w->size + w->N) * kF, 0, (size_t)pad * kF * sizeof(unsigned short), w->stream));
        CLH_HIP(hipMemsetAsync(w->T.ptr + (w->size + w->N), 0, (size_t)pad * sizeof(float4), w->stream));
    }
    CLH_TRY(launch_features(w, xm2, xm1, x, xnext, cloth_accel, w->F.ptr + (size_t)w->size * kF,
                            w->T.ptr + w->size, nullptr, cloth_weight));
    w->size = need;
    return CLH_OK;
}

extern "C" int clh_ws_dataset_append(clh_ws* w, const unsigned short* features,
                                     const float* targets, int count)
{
    if (!w || count < 0 || (count > 0 && (!features || !targets)))
        return fail(CLH_ERR_INVALID, "clh_ws_dataset_append: bad arguments");
    const long long padded = ((long long)count + 15) / 16 * 16;
    const long long need = w->size + padded;
    if (need > w->cap) CLH_TRY(clh_ws_dataset_reserve(w, std::max(need, 2 * w->cap)));
    CLH_HIP(hipStreamSynchronize(w->stream));
    std::vector<unsigned short> F((size_t)padded * kF, 0);
    std::vector<float4> T((size_t)padded, make_float4(0.0f, 0.0f, 0.0f, 0.0f));
    if (count > 0) {
        std::memcpy(F.data(), features, (size_t)count * kF * sizeof(unsigned short));
        std::memcpy(T.data(), targets, (size_t)count * sizeof(float4));
    }
    CLH_HIP(hipMemcpy(w->F.ptr + (size_t)w->size * kF, F.data(), F.size() * sizeof(unsigned short),
                      hipMemcpyHostToDevice));
    CLH_HIP(hipMemcpy(w->T.ptr + w->size, T.data(), T.size() * sizeof(float4), hipMemcpyHostToDevice));
    w->size = need;
    return CLH_OK;
}

extern "C" int clh_ws_dataset_size(clh_ws* w, long long* samples)
{
    if (!w || !samples) return fail(CLH_ERR_INVALID, "clh_ws_dataset_size: null pointer");
    *samples = w->size;
    return CLH_OK;
}

extern "C" int clh_ws_dataset_clear(clh_ws* w)
{
    if (!w) return fail(CLH_ERR_INVALID, "clh_ws_dataset_clear: null handle");
    w->size = 0;
    return CLH_OK;
}

extern "C" int clh_ws_dataset_read(clh_ws* w, long long first, int count, unsigned short* features,
                                   float* targets)
{
    if (!w || first < 0 || count < 0 || first + count > w->size)
        return fail(CLH_ERR_INVALID, "clh_ws_dataset_read: range outside the dataset");
    CLH_HIP(hipStreamSynchronize(w->stream));
    if (features && count)
        CLH_HIP(hipMemcpy(features, w->F.ptr + (size_t)first * kF, (size_t)count * kF * sizeof(unsigned short), hipMemcpyDeviceToHost));
    if (targets && count)
        CLH_HIP(hipMemcpy(targets, w->T.ptr + first, (size_t)count * sizeof(float4), hipMemcpyDeviceToHost));
    return CLH_OK;
}

extern "C" int clh_ws_fit_output_scale(clh_ws* w, float* out_scale)
{
    // root mean square of the weighted targets, measured at scale 1
    if (!w || w->size == 0) return fail(CLH_ERR_INVALID, "clh_ws_fit_output_scale: empty dataset");
    if (!(w->so > 0.0f)) return fail(CLH_ERR_INVALID, "clh_ws_fit_output_scale: output scale is 0");
    CLH_HIP(hipStreamSynchronize(w->stream));
    std::vector<float4> t((size_t)w->size);
    CLH_HIP(hipMemcpy(t.data(), w->T.ptr, t.size() * sizeof(float4), hipMemcpyDeviceToHost));
    double s2 = 0.0, n = 0.0;
    for (const float4& e : t) {
        s2 += (double)e.w * ((double)e.x * e.x + (double)e.y * e.y + (double)e.z * e.z);
        n += 3.0 * e.w;
    }
    const double rms = n > 0 ? std::sqrt(s2 / n) : 1.0;
    const float so = (float)(rms > 0 ? rms * (double)w->so : (double)w->so);
    // rescale stored targets to the new scale
    const float f = w->so / so;
    for (float4& e : t) {
        e.x *= f;
        e.y *= f;
        e.z *= f;
    }
    CLH_HIP(hipMemcpy(w->T.ptr, t.data(), t.size() * sizeof(float4), hipMemcpyHostToDevice));
    w->so = so;
    if (out_scale) *out_scale = so;
    return CLH_OK;
}

extern "C" int clh_ws_loss_grad(clh_ws* w, long long first, int count, float* loss, float* grad)
{
    if (!w || first < 0 || first % 16 || first + count > w->size)
        return fail(CLH_ERR_INVALID, "clh_ws_loss_grad: range outside the dataset or not 16-aligned");
    CLH_TRY(launch_grad(w, first, count, 0));
    CLH_HIP(hipStreamSynchronize(w->stream));
    std::vector<float> g((size_t)kParams + 1);
    CLH_HIP(hipMemcpy(g.data(), w->grad.ptr, g.size() * sizeof(float), hipMemcpyDeviceToHost));
    if (loss) *loss = g[(size_t)kParams];
    if (grad) std::memcpy(grad, g.data(), kParams * sizeof(float));
    return CLH_OK;
}

extern "C" int clh_ws_train(clh_ws* w, int steps, int batch, float lr, float* losses)
{
    if (!w || steps < 0 || batch <= 0 || batch % 16 || !(lr > 0.0f))
        return fail(CLH_ERR_INVALID, "clh_ws_train: bad arguments");
    if (w->size < 16) return fail(CLH_ERR_INVALID, "clh_ws_train: empty dataset");
    CLH_TRY(w->losses.resize((size_t)std::max(steps, 1)));
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    for (int s = 0; s < steps; ++s) {
        // cosine decay from lr to lr / 10
        const double c = 0.5 * (1.0 + std::cos(3.141592653589793 * (double)s / (double)steps));
        const float lrs = (float)((double)lr * (0.1 + 0.9 * c));
        w->sel = mix64(w->sel);
        CLH_TRY(launch_grad(w, 0, batch, w->sel | 1ull));
        ++w->adam_t;
        const float c1 = (float)(1.0 - std::pow((double)b1, (double)w->adam_t));
        const float c2 = (float)(1.0 - std::pow((double)b2, (double)w->adam_t));
        hipLaunchKernelGGL(k_ws_adam, dim3(blocks256(kParams)), dim3(256), 0, w->stream, w->P.ptr,
                           w->m.ptr, w->v.ptr, w->grad.ptr, w->Wq.ptr, lrs, b1, b2, eps, c1, c2);
        hipLaunchKernelGGL(k_ws_store_loss, dim3(1), dim3(64), 0, w->stream, w->grad.ptr,
                           w->losses.ptr, s);
        CLH_HIP(hipGetLastError());
    }
    if (losses && steps > 0) {
        CLH_HIP(hipStreamSynchronize(w->stream));
        CLH_HIP(hipMemcpy(losses, w->losses.ptr, (size_t)steps * sizeof(float), hipMemcpyDeviceToHost));
    }
    return CLH_OK;
}

extern "C" int clh_ws_adam_step(clh_ws* w, const float* grad, float lr)
{
    if (!w || !grad) return fail(CLH_ERR_INVALID, "clh_ws_adam_step: null pointer");
    CLH_HIP(hipStreamSynchronize(w->stream));
    CLH_HIP(hipMemcpy(w->grad.ptr, grad, kParams * sizeof(float), hipMemcpyHostToDevice));
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    ++w->adam_t;
    const float c1 = (float)(1.0 - std::pow((double)b1, (double)w->adam_t));
    const float c2 = (float)(1.0 - std::pow((double)b2, (double)w->adam_t));
    hipLaunchKernelGGL(k_ws_adam, dim3(blocks256(kParams)), dim3(256), 0, w->stream, w->P.ptr,
                       w->m.ptr, w->v.ptr, w->grad.ptr, w->Wq.ptr, lr, b1, b2, eps, c1, c2);
    return hip_status(hipGetLastError(), "warm-start adam launch");
}
