// Fused LDS-tiled iteration kernel: plan construction (host), constant
// packing, the per-iteration kernel, and the tile-order frame prologue /
// epilogue.
#include "pd_fused.h"
#include "pd_kernels.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <numeric>
#include <vector>

namespace clh {
namespace {

constexpr int kFB = 256;  // largest tile (one own vertex per thread)
constexpr int kTileInfo = 8;  // nhalo, tri_beg, ntri, hin_beg, nhin, gather word begin,
                               // gather entries, vbeg
constexpr int kPairInfo = 8;  // ring-2 count, tri_beg, ntri, hin_beg, nhin of the ring-1-only
                               // elements, ring-1 gather word begin, gather entries, 0

// ---------------------------------------------------------------------------
// host: plan
// ---------------------------------------------------------------------------

uint64_t spread21(uint64_t v)
{
    v &= 0x1FFFFF;
    v = (v | (v << 32)) & 0x1F00000000FFFFull;
    v = (v | (v << 16)) & 0x1F0000FF0000FFull;
    v = (v | (v << 8)) & 0x100F00F00F00F00Full;
    v = (v | (v << 4)) & 0x10C30C30C30C30C3ull;
    v = (v | (v << 2)) & 0x1249249249249249ull;
    return v;
}

// Vertex order of the original vertices [beg, end): 3-D Morton code of the
// positions (21 bits per axis over their bounding cube), ties by index.
std::vector<int> morton_order(const std::vector<float>& pos, int beg, int end)
{
    float lo[3] = {1e30f, 1e30f, 1e30f}, hi[3] = {-1e30f, -1e30f, -1e30f};
    for (int v = beg; v < end; ++v)
        for (int k = 0; k < 3; ++k) {
            lo[k] = std::min(lo[k], pos[(size_t)v * 3 + k]);
            hi[k] = std::max(hi[k], pos[(size_t)v * 3 + k]);
        }
    // one scale for all axes (the largest extent), so that a nearly flat
    // cloth orders by its two in-plane axes
    double ext = 0.0;
    for (int k = 0; k < 3; ++k) ext = std::max(ext, (double)hi[k] - (double)lo[k]);
    const int n = end - beg;
    std::vector<uint64_t> key((size_t)n);
    for (int v = beg; v < end; ++v) {
        uint64_t q[3];
        for (int k = 0; k < 3; ++k) {
            double f = ext > 0 ? ((double)pos[(size_t)v * 3 + k] - lo[k]) / ext : 0.0;
            f = std::min(std::max(f, 0.0), 1.0);
            q[k] = (uint64_t)(f * 2097151.0);
        }
        key[(size_t)(v - beg)] = spread21(q[0]) | (spread21(q[1]) << 1) | (spread21(q[2]) << 2);
    }
    std::vector<int> order((size_t)n);
    std::iota(order.begin(), order.end(), beg);
    std::stable_sort(order.begin(), order.end(),
                     [&](int a, int b) { return key[(size_t)(a - beg)] < key[(size_t)(b - beg)]; });
    return order;
}

void incidence(const std::vector<int>& elems, int arity, int N, std::vector<int>& ptr,
               std::vector<int>& ids)
{
    const size_t ne = elems.size() / (size_t)arity;
    ptr.assign((size_t)N + 1, 0);
    for (size_t s = 0; s < elems.size(); ++s) ptr[(size_t)elems[s] + 1]++;
    for (int v = 0; v < N; ++v) ptr[(size_t)v + 1] += ptr[(size_t)v];
    ids.assign(ptr[(size_t)N], 0);
    std::vector<int> fill(ptr.begin(), ptr.end() - 1);
    for (size_t e = 0; e < ne; ++e)
        for (int c = 0; c < arity; ++c) {
            const int v = elems[e * arity + c];
            // one entry per element even if the vertex repeats in it
            if (fill[(size_t)v] > ptr[(size_t)v] && ids[(size_t)fill[(size_t)v] - 1] == (int)e)
                continue;
            ids[(size_t)fill[(size_t)v]++] = (int)e;
        }
    // compact away the holes left by repeated vertices
    std::vector<int> nptr((size_t)N + 1, 0), nids;
    nids.reserve(ids.size());
    for (int v = 0; v < N; ++v) {
        for (int j = ptr[(size_t)v]; j < fill[(size_t)v]; ++j) nids.push_back(ids[(size_t)j]);
        nptr[(size_t)v + 1] = (int)nids.size();
    }
    ptr.swap(nptr);
    ids.swap(nids);
}

// vertex -> incident slots (element * arity + corner), ascending
void build_slots(const std::vector<int>& elems, int arity, int N, std::vector<int>& ptr,
                 std::vector<int>& ids)
{
    ptr.assign((size_t)N + 1, 0);
    for (size_t s = 0; s < elems.size(); ++s) ptr[(size_t)elems[s] + 1]++;
    for (int v = 0; v < N; ++v) ptr[(size_t)v + 1] += ptr[(size_t)v];
    ids.assign(elems.size(), 0);
    std::vector<int> fill(ptr.begin(), ptr.end() - 1);
    for (size_t s = 0; s < elems.size(); ++s) ids[(size_t)fill[(size_t)elems[s]]++] = (int)s;
    (void)arity;
}

struct HostPlan {
    FusedPlan meta;  // sizes only; device buffers unused
    std::vector<int> perm, inv, info, halo, tri_gid, hin_gid, tile_cloth, cloth_tiles;
    std::vector<int> nbr_ptr, nbr;  // neighbour tiles (owners of each tile's halo vertices)
    std::vector<uint2> tri_idx, hin_idx;
    std::vector<unsigned short> gl_idx;
    std::vector<unsigned> gl_meta;
    // two-iteration extension (meta.pair): ring-2 halo, the elements that
    // touch ring 1 but no own vertex, and ring 1's gather lists
    std::vector<int> pinfo, phalo, ptri_gid, phin_gid;
    std::vector<uint2> ptri_idx, phin_idx;
    std::vector<unsigned short> pgl_idx;
    std::vector<unsigned> pmeta;
};

// ---------------------------------------------------------------------------
// device
// ---------------------------------------------------------------------------

struct FusedKernelArgs {
    FusedIterArgs it;
    const int*      tile_info;
    const int*      halo;
    const uint2*    tri_idx;
    const float4*   tri_c0;
    const float*    tri_c1;
    const uint2*    hin_idx;
    const float4*   hin_w;
    const float2*   hin_c;
    const unsigned* gl_words;
    const int*      tile_cloth;   // [num_tiles] cloth of each tile
    const int*      cloth_tiles;  // [C + 1] first tile of each cloth
    int N, tile_verts, hstride;
    int lmax, tmax, hmax;
    int off_g, off_r;  // LDS byte offsets: gather list, element region
    int num_tiles;
    int xcd_remap;
};

__device__ __forceinline__ f3 lds3(const float* s, int i)
{
    return f3{s[3 * i], s[3 * i + 1], s[3 * i + 2]};
}
__device__ __forceinline__ void sts3(float* s, int i, f3 v)
{
    s[3 * i] = v.x;
    s[3 * i + 1] = v.y;
    s[3 * i + 2] = v.z;
}
// PD position table: three FP64 component arrays of stride n (a vertex's
// components sit in different banks from its neighbours')
__device__ __forceinline__ d3 ldsX(const double* s, int n, int i)
{
    return d3{s[i], s[n + i], s[2 * n + i]};
}
__device__ __forceinline__ void stsX(double* s, int n, int i, d3 v)
{
    s[i] = v.x;
    s[n + i] = v.y;
    s[2 * n + i] = v.z;
}

// Tile of workgroup b among n: workgroup b runs on XCD b % 8, and the
// workgroups of one XCD take a contiguous range of tiles (a spatial region,
// whose shared halo vertices then meet in that XCD's L2). Residue class x
// has n / 8 members, plus one for x < n % 8, so the ranges start at
// x (n / 8) + min(x, n % 8) and the map is a bijection for every n.
__device__ __forceinline__ int xcd_tile(int b, int n)
{
    const int x = b & 7;
    return x * (n >> 3) + min(x, n & 7) + (b >> 3);
}

// Per-vertex arithmetic shared by every kernel that updates a vertex, so that
// a vertex updated by two workgroups (as own vertex and as another tile's
// ring-1 vertex), or by two kernels, gets the same bits. Multiply-adds are
// explicit fused operations and the rest does not contract, so the rounding
// does not depend on what the compiler does with the surrounding code.

// PD update of a free vertex: residual r = f + m (dy - d) of the global
// system at the displacement d, the Jacobi step d + gamma r / D (gamma = 1:
// plain Jacobi; below 1 it pulls the modes with Jacobi eigenvalues below -1,
// which stiff bending produces, back inside the interval the Chebyshev
// weights assume), then the Chebyshev blend with d_{k-1} when `blend`.
__device__ __forceinline__ d3 pd_update(f3 ar, float om, float odg, f3 oy, d3 od, d3 oxp,
                                        float omega, int blend, float gamma, d3& r)
{
#pragma clang fp contract(off)
    const double md = om;
    r = d3{fma(md, (double)oy.x - od.x, (double)ar.x), fma(md, (double)oy.y - od.y, (double)ar.y),
           fma(md, (double)oy.z - od.z, (double)ar.z)};
    const double inv = (odg > 1e-10f) ? (double)gamma / (double)odg : 0.0;
    d3 xn = d3{fma(r.x, inv, od.x), fma(r.y, inv, od.y), fma(r.z, inv, od.z)};
    if (blend) {
        const double w = omega;
        xn = d3{fma(xn.x - oxp.x, w, oxp.x), fma(xn.y - oxp.y, w, oxp.y), fma(xn.z - oxp.z, w, oxp.z)};
    }
    return xn;
}

// The Jacobi step scale of a launch (0 in the arguments means 1).
__device__ __forceinline__ float relax_gamma(const FusedIterArgs& a)
{
    return a.relax > 0.0f ? a.relax : 1.0f;
}

// acc + kw * v, component-wise, fused.
__device__ __forceinline__ f3 fmadd3(f3 acc, float kw, f3 v)
{
    return f3{fmaf(kw, v.x, acc.x), fmaf(kw, v.y, acc.y), fmaf(kw, v.z, acc.z)};
}

// |r|^2 and |b|^2 of one vertex for the FP64 relative residual, with
// b = ab + m (x_n + dy) the right-hand side of the global system.
__device__ __forceinline__ void pd_residual_terms(d3 r, f3 ab, float om, f3 ob, f3 oy, double& rr,
                                                  double& bb)
{
#pragma clang fp contract(off)
    const f3 ya = ob + oy;
    const f3 b = fmadd3(ab, om, ya);
    rr = fma(r.z, r.z, fma(r.y, r.y, r.x * r.x));
    bb = fma((double)b.z, (double)b.z, fma((double)b.y, (double)b.y, (double)b.x * (double)b.x));
}

// Gather sums of one vertex, in four interleaved partial sums. Partial q
// takes the entries at offsets q, q + 4, q + 8, ... of each part of the
// vertex's list (triangle corners from j0, hinge slots from jt), in list
// order, and the vertex's sum is (P0 + P1) + (P2 + P3). kG adjacent lanes
// (1, 2 or 4) share a vertex: lane p of the group holds the partials
// q = p + kG m, m < 4 / kG, and the combination below uses only the
// commutativity of IEEE addition, so the sum has the same bits for every kG.
// Every entry is two dependent LDS reads (the slot, then its values); a lane
// issues U entries' reads before their sums (U = 4, or 2 with four lanes per
// vertex, whose lists then take one or two such steps per lane); reads past
// the end repeat the last entry and are not summed.
template <int kG>
struct GatherAcc {
    static constexpr int M = 4 / kG;
    static constexpr int U = kG == 4 ? 2 : 4;
    f3 r[M];   // force sums
    f3 b[M];   // right-hand-side parts (residual check)
    __device__ GatherAcc()
    {
#pragma unroll
        for (int m = 0; m < M; ++m) r[m] = b[m] = mk(0.0f, 0.0f, 0.0f);
    }
};

// Triangle corners j0 .. jt: corner forces st into acc.r and, with kCheck,
// the R parts st2 into acc.b. p is the lane's index in its group.
template <bool kCheck, int kG>
__device__ __forceinline__ void gather_corners(const unsigned short* gl, int j0, int jt, int p,
                                               const float* st, const float* st2,
                                               GatherAcc<kG>& acc)
{
    constexpr int M = GatherAcc<kG>::M;
    constexpr int U = GatherAcc<kG>::U;
    for (int j = j0 + p; j < jt; j += U * kG) {
        int s[U];
#pragma unroll
        for (int u = 0; u < U; ++u) s[u] = gl[min(j + kG * u, jt - 1)];
        f3 v[U];
#pragma unroll
        for (int u = 0; u < U; ++u) v[u] = lds3(st, s[u]);
#pragma unroll
        for (int u = 0; u < U; ++u)
            if (j + kG * u < jt) acc.r[u % M] = acc.r[u % M] + v[u];
        if (kCheck) {
#pragma unroll
            for (int u = 0; u < U; ++u) v[u] = lds3(st2, s[u]);
#pragma unroll
            for (int u = 0; u < U; ++u)
                if (j + kG * u < jt) acc.b[u % M] = acc.b[u % M] + v[u];
        }
    }
}

// Hinge slots jt .. j1: k h^2 w_i (arrays of stride hmax in skw) times the
// hinge vectors sq into acc.r (fused in PD mode, unfused in reference mode)
// and, with kCheck, times the projections sp into acc.b.
template <bool kPD, bool kCheck, int kG>
__device__ __forceinline__ void gather_hinges(const unsigned short* gl, int jt, int j1, int p,
                                              const float* skw, int hmax, const float* sq,
                                              const float* sp, GatherAcc<kG>& acc)
{
    constexpr int M = GatherAcc<kG>::M;
    constexpr int U = GatherAcc<kG>::U;
    for (int j = jt + p; j < j1; j += U * kG) {
        int s[U];
#pragma unroll
        for (int u = 0; u < U; ++u) s[u] = gl[min(j + kG * u, j1 - 1)];
        float kw[U];
        f3 d[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            kw[u] = skw[(s[u] & 3) * hmax + (s[u] >> 2)];
            d[u] = lds3(sq, s[u] >> 2);
        }
#pragma unroll
        for (int u = 0; u < U; ++u)
            if (j + kG * u < j1)
                acc.r[u % M] = kPD ? fmadd3(acc.r[u % M], kw[u], d[u]) : madd_nc(acc.r[u % M], kw[u], d[u]);
        if (kPD && kCheck) {
#pragma unroll
            for (int u = 0; u < U; ++u) d[u] = lds3(sp, s[u] >> 2);
#pragma unroll
            for (int u = 0; u < U; ++u)
                if (j + kG * u < j1) acc.b[u % M] = fmadd3(acc.b[u % M], kw[u], d[u]);
        }
    }
}

__device__ __forceinline__ f3 shfl_xor3(f3 v, int mask)
{
    return f3{__shfl_xor(v.x, mask, 64), __shfl_xor(v.y, mask, 64), __shfl_xor(v.z, mask, 64)};
}

// (P0 + P1) + (P2 + P3) from the partials of the lane's group. With kG > 1
// every lane of the wave must execute it (cross-lane reads).
template <int kG>
__device__ __forceinline__ f3 gather_total(const f3* a)
{
    if constexpr (kG == 1) {
        return (a[0] + a[1]) + (a[2] + a[3]);
    } else if constexpr (kG == 2) {
        const f3 s0 = a[0] + shfl_xor3(a[0], 1);
        const f3 s1 = a[1] + shfl_xor3(a[1], 1);
        return s0 + s1;
    } else {
        const f3 s = a[0] + shfl_xor3(a[0], 1);
        return s + shfl_xor3(s, 2);
    }
}

template <int kT>
__device__ void fused_block_sum2(double r, double b, double* out, int slot)
{
    __shared__ double sr[kT / 64], sb[kT / 64];
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
        for (int i = 0; i < kT / 64; ++i) {
            tr += sr[i];
            tb += sb[i];
        }
        out[2 * slot + 0] = tr;
        out[2 * slot + 1] = tb;
    }
}

// One local-global iteration of one tile. kMode selects the reference
// update (absolute FP32 positions, rhs / diag) or PD (FP64 displacements,
// Jacobi + Chebyshev); kCheck adds the right-hand-side norm for the FP64
// residual.
//
// Memory trips: (1) the tile descriptor, the own vertices' data (their range
// is tile * tile_verts) and the halo ids (fixed stride per tile) are all
// independent; (2) halo positions, the first element records and the gather
// list follow the descriptor. Further element records are fetched one pass
// ahead of their use.
//
// LDS: position table (PD: FP64 x_n + d as three component arrays;
// reference: FP32 triples), the tile's gather list, and one element region
// that holds the triangle corner forces and, once they are summed, the hinge
// weights (four component arrays) and vectors. With kSplit the hinges have a
// region of their own: their first records then come with trip 2, triangles
// and hinges are evaluated in one phase, and the own vertices sum both after
// one barrier (two barriers fewer; same sums in the same order).
//
// kG adjacent lanes share an own vertex in the gathers (four partial sums,
// same bits for every kG); the group's first lane loads the vertex's data and
// updates it. The host picks kG = min(4, kT / tile_verts).
template <int kMode, bool kCheck, int kT, bool kSplit, int kG>
__global__ void __launch_bounds__(kT) k_fused_iter(FusedKernelArgs A)
{
    constexpr bool PD = (kMode == CLH_MODE_PD);
    extern __shared__ double smem[];
    FusedIterArgs& a = A.it;
    const int tid = threadIdx.x;
    // control state of a frame run to a tolerance: loaded with the first
    // trip, acted on before anything is written
    int stopped = 0;
    CtlSlot cs = {a.omega, a.blend, 0, 0};
    if (a.ctl) {
        stopped = a.ctl->stop;
        cs = a.ctl->slot[a.par];
    }

    const int L = A.lmax;
    const int tile = A.xcd_remap ? xcd_tile(blockIdx.x, A.num_tiles) : (int)blockIdx.x;
    const int vbeg = tile * A.tile_verts;
    const int nown = min(A.tile_verts, A.N - vbeg);

    char* lds = reinterpret_cast<char*>(smem);
    double* sX = smem;                                   // PD positions
    float* sx = reinterpret_cast<float*>(smem);          // reference positions
    unsigned* sgw = reinterpret_cast<unsigned*>(lds + A.off_g);
    const unsigned short* sg = reinterpret_cast<const unsigned short*>(sgw);
    float* st = reinterpret_cast<float*>(lds + A.off_r);     // tri corner forces [tmax*9]
    float* st2 = st + A.tmax * 9;                            // tri R part [tmax*9] (PD, check)
    float* skw = kSplit ? st + A.tmax * 9 * ((PD && kCheck) ? 2 : 1)
                        : st;                                // hinge k h^2 w_i [4][hmax]
    float* sq = skw + A.hmax * 4;                            // hinge p-e (PD) or p [hmax*3]
    float* sp = sq + A.hmax * 3;                             // hinge p [hmax*3] (PD, check)

    const float* xf = static_cast<const float*>(a.x);
    const double* xd = static_cast<const double*>(a.x);

    // Each trip's loads are issued, without branches, before any of their
    // values is used: lanes with nothing to load read an in-range element
    // (index 0, or the tile's first own vertex), so no consumer placed inside a
    // branch makes the lane wait for one load before issuing the next, and a
    // trip is one round trip to memory.

    // ---- trip 1: descriptor, own vertices, halo ids --------------------------
    const int* ti = A.tile_info + (size_t)tile * kTileInfo;
    const int vl = tid / kG, lane_p = tid % kG;  // own vertex of the lane's group
    const bool own = vl < nown;
    const bool lead = own && lane_p == 0;        // loads and updates the vertex
    const size_t g = (size_t)vbeg + (own ? vl : 0);
    const int4 vld = a.vinfo[g];
    const f3 oy = ld3(a.y, g);
    d3 od = d3{0.0, 0.0, 0.0}, oxp = d3{0.0, 0.0, 0.0};
    f3 ob;
    if (PD) {
        od = ld3d(xd, g);
        ob = ld3(a.base, g);
        // under device control the blend flag arrives with this trip
        if (a.blend || a.ctl) oxp = ld3d(static_cast<const double*>(a.xkm1), g);
    } else {
        ob = ld3(xf, g);
    }
    const int* hids = A.halo + (size_t)tile * A.hstride;
    const int hid0 = hids[min(tid, A.hstride - 1)];
    const int4 vi = own ? vld : make_int4(0, 0, -1, 0);

    const int cloth = (a.cctl || a.cloth_relax) ? A.tile_cloth[tile] : 0;

    const int nhalo = ti[0], tbeg = ti[1], ntri = ti[2], ebeg = ti[3], nhin = ti[4];
    const int gbeg = ti[5], glen = ti[6];
    if (a.ctl) {
        if (stopped) return;
        a.omega = cs.omega;
        a.blend = cs.blend;
        ctl_write_next(a.ctl, a.par, cs);
    }
    // per-cloth control: the cloth's state arrives with trip 2
    ClothCtl* cc = a.cctl ? a.cctl + cloth : nullptr;
    CtlSlot ccs = cs;
    int cstate = 0, cfirst = -1;
    float crho = 0.0f;
    if (cc) {
        ccs = cc->slot[a.par];
        cstate = cc->state;
        crho = cc->rho;
        cfirst = A.cloth_tiles[cloth];
    }

    // ---- trip 2: halo positions, first records, gather list -------------------
    const int hsel = tid < nhalo ? hid0 : vbeg;
    f3 hb = mk(0.0f, 0.0f, 0.0f), hf = mk(0.0f, 0.0f, 0.0f);
    d3 hx = d3{0.0, 0.0, 0.0};
    if (PD) {
        hb = ld3(a.base, hsel);
        hx = ld3d(xd, hsel);
    } else {
        hf = ld3(xf, hsel);
    }
    const size_t tj = tid < ntri ? (size_t)tbeg + tid : 0;
    uint2 rt = A.tri_idx[tj];
    float4 rG = A.tri_c0[tj];
    float rwA = A.tri_c1[tj];
    uint2 rh = make_uint2(0, 0);
    float4 rw = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    float2 rc = make_float2(0.0f, 0.0f);
    if (kSplit) {
        const size_t hj = tid < nhin ? (size_t)ebeg + tid : 0;
        rh = A.hin_idx[hj];
        rw = A.hin_w[hj];
        rc = A.hin_c[hj];
    }
    const int nwords = (glen + 1) >> 1;
    const unsigned gw0 = A.gl_words[tid < nwords ? gbeg + tid : 0];
    // consumers
    if (tid < nhalo) {
        if (PD) stsX(sX, L, nown + tid, to_d(hb) + hx);
        else    sts3(sx, nown + tid, hf);
    }
    for (int i = tid + kT; i < nhalo; i += kT) {
        const int hg = hids[i];
        if (PD) stsX(sX, L, nown + i, to_d(ld3(a.base, hg)) + ld3d(xd, hg));
        else    sts3(sx, nown + i, ld3(xf, hg));
    }
    if (tid < nwords) sgw[tid] = gw0;
    for (int j = tid + kT; j < nwords; j += kT) sgw[j] = A.gl_words[gbeg + j];
    if (lead) {
        if (PD) stsX(sX, L, vl, to_d(ob) + od);
        else    sts3(sx, vl, ob);
    }
    if (cc) {
        // a cloth that has stopped keeps x_{k+1} = x_k, and one that diverged
        // goes back to x_n (pinned vertices at their targets) for the rest of
        // the frame; the whole workgroup returns (the condition is the tile's)
        if (cstate != CLH_CLOTH_RUNNING) {
            if (lead && PD) {
                d3 keep = od;
                if (cstate == CLH_CLOTH_NONFINITE)
                    keep = vi.z >= 0 ? to_d(ld3(a.pin_target, vi.z)) - to_d(ob) : d3{0.0, 0.0, 0.0};
                st3d(static_cast<double*>(a.out), g, keep);
            }
            return;
        }
        a.omega = ccs.omega;
        a.blend = ccs.blend;
        if (tile == cfirst && tid == 0)
            cc->slot[a.par ^ 1] = ctl_next_rho(crho, a.ctl->delay, a.ctl->use_cheb, ccs);
    }
    __syncthreads();

    // ---- triangles -> corner forces (src/pd_solver.cu:48-298) -----------------
    if (a.use_stretch) {
        for (int j = tid; j < ntri; j += kT) {
            const uint2 id = rt;
            const float4 G = rG;
            const float wA = rwA;
            {  // next pass's record (index 0 after the last pass; no branch)
                const size_t jn = j + kT < ntri ? (size_t)tbeg + j + kT : 0;
                rt = A.tri_idx[jn];
                rG = A.tri_c0[jn];
                rwA = A.tri_c1[jn];
            }
            const int l0 = id.x & 0xFFFF, l1 = id.x >> 16, l2 = id.y & 0xFFFF;
            f3 c[3];
            if (PD) {
                f3 cr[3];
                pd_triangle(ldsX(sX, L, l0), ldsX(sX, L, l1), ldsX(sX, L, l2), G, wA, cr, c, kCheck);
                if (kCheck) {
                    float* q = st2 + (size_t)j * 9;
                    sts3(q, 0, cr[0]);
                    sts3(q, 1, cr[1]);
                    sts3(q, 2, cr[2]);
                }
            } else {
                const f3 x0 = lds3(sx, l0);
                f3 F0, F1, R0, R1;
                deformation_nc(lds3(sx, l1) - x0, lds3(sx, l2) - x0, G.x, G.y, G.z, G.w, F0, F1);
                ref_cuda_rotation(F0, F1, R0, R1);
                corner_forces_nc(G.x, G.y, G.z, G.w, wA, R0, R1, c);
            }
            float* o = st + (size_t)j * 9;
            sts3(o, 0, c[0]);
            sts3(o, 1, c[1]);
            sts3(o, 2, c[2]);
        }
    }
    if (!kSplit) {
        // first hinge record (its latency overlaps the triangle sums; no branch)
        const size_t hj = tid < nhin ? (size_t)ebeg + tid : 0;
        rh = A.hin_idx[hj];
        rw = A.hin_w[hj];
        rc = A.hin_c[hj];
        __syncthreads();
    }

    // ---- own vertices: triangle corners (ascending element id) ---------------
    const int opin = vi.z;
    const unsigned meta = (unsigned)vi.w;
    const int j0 = (int)(meta & 0xFFFFu), jt = j0 + (int)((meta >> 16) & 0xFFu),
              j1 = j0 + (int)(meta >> 24);
    const bool gather = own && opin < 0;
    GatherAcc<kG> acc;
    auto sum_triangles = [&]() {
        if (gather && a.use_stretch) gather_corners<PD && kCheck, kG>(sg, j0, jt, lane_p, st, st2, acc);
    };
    if (!kSplit) sum_triangles();
    if (kSplit && !a.use_bend) {
        __syncthreads();
        sum_triangles();
    }
    if (a.use_bend) {
        if (!kSplit) __syncthreads();  // the hinge vectors reuse the triangle region
        // ---- hinges -> bend vectors (src/pd_solver.cu:304-358) ----------------
        for (int j = tid; j < nhin; j += kT) {
            const uint2 id = rh;
            const float4 w = rw;
            const float2 c = rc;  // (n_rest, k h^2)
            {  // next pass's record (index 0 after the last pass; no branch)
                const size_t jn = j + kT < nhin ? (size_t)ebeg + j + kT : 0;
                rh = A.hin_idx[jn];
                rw = A.hin_w[jn];
                rc = A.hin_c[jn];
            }
            const int q0 = id.x & 0xFFFF, q1 = id.x >> 16, q2 = id.y & 0xFFFF, q3 = id.y >> 16;
            f3 p, pe;
            if (PD) {
                const d3 X[4] = {ldsX(sX, L, q0), ldsX(sX, L, q1), ldsX(sX, L, q2), ldsX(sX, L, q3)};
                pd_hinge(X, w, c.x, p, pe);
            } else {
                f3 acc = mk(0.0f, 0.0f, 0.0f);
                acc = madd_nc(acc, w.x, lds3(sx, q0));
                acc = madd_nc(acc, w.y, lds3(sx, q1));
                acc = madd_nc(acc, w.z, lds3(sx, q2));
                acc = madd_nc(acc, w.w, lds3(sx, q3));
                p = hinge_projection_nc(acc, c.x);
            }
            skw[j] = c.y * w.x;
            skw[A.hmax + j] = c.y * w.y;
            skw[2 * A.hmax + j] = c.y * w.z;
            skw[3 * A.hmax + j] = c.y * w.w;
            sts3(sq, j, PD ? pe : p);
            if (PD && kCheck) sts3(sp, j, p);
        }
        __syncthreads();
        if (kSplit) sum_triangles();
        if (gather) gather_hinges<PD, PD && kCheck, kG>(sg, jt, j1, lane_p, skw, A.hmax, sq, sp, acc);
    }
    const f3 ar = gather_total<kG>(acc.r);
    const f3 ab = (PD && kCheck) ? gather_total<kG>(acc.b) : mk(0.0f, 0.0f, 0.0f);

    // ---- own vertices: update, pins, residual ----------------------------------
    double rr = 0.0, bb = 0.0;
    if (lead) {
        const float om = __int_as_float(vi.x), odg = __int_as_float(vi.y);
        if (opin >= 0) {
            const f3 tgt = ld3(a.pin_target, opin);
            if (PD) st3d(static_cast<double*>(a.out), g, to_d(tgt) - to_d(ob));
            else    st3(static_cast<float*>(a.out), g, tgt);
        } else if (!PD) {
            const float inv = (odg > 1e-10f) ? (1.0f / odg) : 0.0f;
            st3(static_cast<float*>(a.out), g, madd_nc(ar, om, oy) * inv);
        } else {
            d3 r;
            const float gamma = a.cloth_relax ? a.cloth_relax[cloth] : relax_gamma(a);
            const d3 xn = pd_update(ar, om, odg, oy, od, oxp, a.omega, a.blend, gamma, r);
            st3d(static_cast<double*>(a.out), g, xn);
            if (kCheck) {
                pd_residual_terms(r, ab, om, ob, oy, rr, bb);
            }
        }
    }
    if (kCheck) {
        if (kG > 1) {
            // the block sum takes vertex i's terms in thread i, as with one
            // lane per vertex; the position table is free (its last reads
            // came before the barrier ahead of the gathers)
            if (lead) {
                sX[vl] = rr;
                sX[L + vl] = bb;
            }
            __syncthreads();
            rr = tid < nown ? sX[tid] : 0.0;
            bb = tid < nown ? sX[L + tid] : 0.0;
        }
        fused_block_sum2<kT>(rr, bb, a.partials, tile);
    }
}

// ---- two PD iterations per launch ---------------------------------------------
//
// Iteration k updates the tile's own vertices and its ring-1 halo (the
// vertices sharing an element with an own vertex) from the positions of own,
// ring 1 and ring 2; iteration k+1 then updates the own vertices from the
// first iteration's own and ring-1 results, all in LDS. A ring-1 vertex is
// updated exactly as its own tile updates it: same element records, same
// gather order (ascending element id), same update function, so x_{k+1} of
// every vertex is the same in every tile that computes it, and the two
// iterations give the bits of two single-iteration launches. Memory trips are
// those of k_fused_iter: (1) descriptors, own vertex data, pair halo ids and
// ring-1 gather metas; (2) halo positions, ring-1 vertex data, first element
// records and both gather lists. The first element records of the second
// iteration are the first iteration's own-list records, kept in registers.
// x_{k+1} of own vertices goes to a.out (the next launch's x_{k-1}), x_{k+2}
// to out2. The residual check, when requested, belongs to iteration k+1.
//
// LDS: position table (FP64 x_n + d of own, ring 1, ring 2 as three
// component arrays), own gather list, ring-1 gather list, element region
// (triangle corner forces of both element lists, then hinge weights and
// vectors).

struct PairKernelArgs {
    FusedKernelArgs A;      // plan, frame arrays; A.it.out receives x_{k+1}
    void* out2;             // x_{k+2}
    float omega1;           // second iteration (host control)
    int   blend1;
    const int*      pinfo;
    const int*      phalo;
    const unsigned* pmeta;
    const uint2*    ptri_idx;
    const float4*   ptri_c0;
    const float*    ptri_c1;
    const uint2*    phin_idx;
    const float4*   phin_w;
    const float2*   phin_c;
    const unsigned* pgl_words;
    int pstride, plmax, ptmax, phmax;
    int off_g2;             // LDS byte offset of the ring-1 gather list
};

template <bool kCheck, int kT>
__global__ void __launch_bounds__(kT) k_fused_pair(PairKernelArgs P)
{
    extern __shared__ double smem[];
    const FusedKernelArgs& A = P.A;
    const FusedIterArgs& a = A.it;
    const int tid = threadIdx.x;
    int stopped = 0;
    CtlSlot c0 = {a.omega, a.blend, 0, 0};
    if (a.ctl) {
        stopped = a.ctl->stop;
        c0 = a.ctl->slot[a.par];
    }

    const int L = P.plmax;
    const int tile = A.xcd_remap ? xcd_tile(blockIdx.x, A.num_tiles) : (int)blockIdx.x;
    const float gamma = a.cloth_relax ? a.cloth_relax[A.tile_cloth[tile]] : relax_gamma(a);
    const int vbeg = tile * A.tile_verts;
    const int nown = min(A.tile_verts, A.N - vbeg);

    char* lds = reinterpret_cast<char*>(smem);
    double* sX = smem;
    unsigned* sgw = reinterpret_cast<unsigned*>(lds + A.off_g);
    unsigned* sgw2 = reinterpret_cast<unsigned*>(lds + P.off_g2);
    const unsigned short* sg = reinterpret_cast<const unsigned short*>(sgw);
    const unsigned short* sg2 = reinterpret_cast<const unsigned short*>(sgw2);
    float* st = reinterpret_cast<float*>(lds + A.off_r);     // tri corner forces [ptmax*9]
    float* st2 = st + P.ptmax * 9;                           // tri R part [tmax*9] (check)
    float* skw = st;                                         // hinge k h^2 w_i [4][phmax]
    float* sq = skw + P.phmax * 4;                           // hinge p-e [phmax*3]
    float* sp = sq + P.phmax * 3;                            // hinge p [hmax*3] (check)

    const double* xd = static_cast<const double*>(a.x);
    const double* xm = static_cast<const double*>(a.xkm1);

    // (as in k_fused_iter, each trip's loads are issued without branches
    // before any of their values is used; lanes with nothing to load read an
    // in-range element)

    // ---- trip 1: descriptors, own vertices, pair halo ids, ring-1 metas -----
    const int* ti = A.tile_info + (size_t)tile * kTileInfo;
    const int* pi = P.pinfo + (size_t)tile * kPairInfo;
    const bool own = tid < nown;
    const int r1 = tid - nown;  // ring-1 index of this thread's vertex
    const size_t g = (size_t)vbeg + (own ? tid : 0);
    const int4 vo = a.vinfo[g];
    const f3 oyo = ld3(a.y, g);
    const d3 odo = ld3d(xd, g);
    const f3 obo = ld3(a.base, g);
    d3 oxpo = d3{0.0, 0.0, 0.0};
    if (a.blend || a.ctl) oxpo = ld3d(xm, g);
    const int* ph = P.phalo + (size_t)tile * P.pstride;
    const int hid0 = ph[min(tid, P.pstride - 1)];
    const int r1c = min(max(r1, 0), A.hstride - 1);
    const int rid = ph[r1c];
    const unsigned rmeta = P.pmeta[(size_t)tile * A.hstride + r1c];
    const int nh1 = ti[0], tbeg = ti[1], ntri = ti[2], ebeg = ti[3], nhin = ti[4];
    const int gbeg = ti[5], glen = ti[6];
    const int nh2 = pi[0], tbeg2 = pi[1], ntri2 = pi[2], ebeg2 = pi[3], nhin2 = pi[4];
    const int pgbeg = pi[5], pglen = pi[6];
    CtlSlot c1 = {P.omega1, P.blend1, 0, 0};
    if (a.ctl) {
        if (stopped) return;
        c1 = ctl_next(*a.ctl, c0);
        // the launch after this one starts at iteration k + 2
        ctl_write_next(a.ctl, a.par, c1);
    }
    const bool ring1 = r1 >= 0 && r1 < nh1;
    const bool cv = own || ring1;  // updated in the first iteration

    // ---- trip 2: halo positions, ring-1 data, first records, gather lists ---
    const int nhh = nh1 + nh2;
    const int hsel = tid < nhh ? hid0 : vbeg;
    const f3 hb = ld3(a.base, hsel);
    const d3 hx = ld3d(xd, hsel);
    const size_t rsel = ring1 ? (size_t)rid : g;
    const int4 vr = a.vinfo[rsel];
    const f3 oyr = ld3(a.y, rsel);
    const d3 odr = ld3d(xd, rsel);
    const f3 obr = ld3(a.base, rsel);
    d3 oxpr = d3{0.0, 0.0, 0.0};
    if (c0.blend) oxpr = ld3d(xm, rsel);
    const int ntA = ntri + ntri2, nhA = nhin + nhin2;
    // first records of both iterations: slot tid of the first iteration's
    // list, which for tid < ntri is also slot tid of the second's
    const size_t ta = tid < ntri ? (size_t)tbeg + tid : 0;
    const size_t tp = (tid >= ntri && tid < ntA) ? (size_t)tbeg2 + tid - ntri : 0;
    const uint2 rta = A.tri_idx[ta];
    const float4 rGa = A.tri_c0[ta];
    const float rwAa = A.tri_c1[ta];
    const uint2 rtp = P.ptri_idx[tp];
    const float4 rGp = P.ptri_c0[tp];
    const float rwAp = P.ptri_c1[tp];
    const int nw = (glen + 1) >> 1, pnw = (pglen + 1) >> 1;
    const unsigned gw0 = A.gl_words[tid < nw ? gbeg + tid : 0];
    const unsigned pgw0 = P.pgl_words[tid < pnw ? pgbeg + tid : 0];
    // consumers
    if (tid < nhh) stsX(sX, L, nown + tid, to_d(hb) + hx);
    for (int i = tid + kT; i < nhh; i += kT) {
        const int hg = ph[i];
        stsX(sX, L, nown + i, to_d(ld3(a.base, hg)) + ld3d(xd, hg));
    }
    int4 vi = own ? vo : make_int4(0, 0, -1, 0);
    f3 oy = oyo, ob = obo;
    d3 od = odo, oxp = oxpo;
    if (ring1) {
        vi = vr;
        oy = oyr;
        od = odr;
        ob = obr;
        oxp = oxpr;
    }
    const bool t_own = tid < ntri;
    const uint2 rt0 = t_own ? rta : rtp;
    const float4 rG0 = t_own ? rGa : rGp;
    const float rwA0 = t_own ? rwAa : rwAp;
    if (tid < nw) sgw[tid] = gw0;
    for (int j = tid + kT; j < nw; j += kT) sgw[j] = A.gl_words[gbeg + j];
    if (tid < pnw) sgw2[tid] = pgw0;
    for (int j = tid + kT; j < pnw; j += kT) sgw2[j] = P.pgl_words[pgbeg + j];
    if (own) stsX(sX, L, tid, to_d(ob) + od);
    __syncthreads();

    const int opin = vi.z;
    const float om = __int_as_float(vi.x), odg = __int_as_float(vi.y);
    const bool gather = cv && opin < 0;

    // ---- iteration k: elements of own and ring-1 vertices --------------------
    if (a.use_stretch) {
        uint2 rt = rt0;
        float4 rG = rG0;
        float rwA = rwA0;
        for (int j = tid; j < ntA; j += kT) {
            const uint2 id = rt;
            const float4 G = rG;
            const float wA = rwA;
            {  // next pass's record from either list (index 0 when unused; no branch)
                const int jn = j + kT;
                const size_t qa = jn < ntri ? (size_t)tbeg + jn : 0;
                const size_t qp = (jn >= ntri && jn < ntA) ? (size_t)tbeg2 + jn - ntri : 0;
                const uint2 ta_ = A.tri_idx[qa], tp_ = P.ptri_idx[qp];
                const float4 ga_ = A.tri_c0[qa], gp_ = P.ptri_c0[qp];
                const float wa_ = A.tri_c1[qa], wp_ = P.ptri_c1[qp];
                rt = jn < ntri ? ta_ : tp_;
                rG = jn < ntri ? ga_ : gp_;
                rwA = jn < ntri ? wa_ : wp_;
            }
            const int l0 = id.x & 0xFFFF, l1 = id.x >> 16, l2 = id.y & 0xFFFF;
            f3 cr[3], c[3];
            pd_triangle(ldsX(sX, L, l0), ldsX(sX, L, l1), ldsX(sX, L, l2), G, wA, cr, c, false);
            float* o = st + (size_t)j * 9;
            sts3(o, 0, c[0]);
            sts3(o, 1, c[1]);
            sts3(o, 2, c[2]);
        }
    }
    // first hinge records of both iterations (no branch)
    const size_t ha = tid < nhin ? (size_t)ebeg + tid : 0;
    const size_t hp = (tid >= nhin && tid < nhA) ? (size_t)ebeg2 + tid - nhin : 0;
    const uint2 rha = A.hin_idx[ha], rhp = P.phin_idx[hp];
    const float4 rwa = A.hin_w[ha], rwp = P.phin_w[hp];
    const float2 rca = A.hin_c[ha], rcp = P.phin_c[hp];
    const uint2 rh0 = tid < nhin ? rha : rhp;
    const float4 rw0 = tid < nhin ? rwa : rwp;
    const float2 rc0 = tid < nhin ? rca : rcp;
    __syncthreads();

    const unsigned meta = own ? (unsigned)vi.w : rmeta;
    const unsigned short* gl = own ? sg : sg2;
    const int j0 = (int)(meta & 0xFFFFu), jt = j0 + (int)((meta >> 16) & 0xFFu),
              j1 = j0 + (int)(meta >> 24);
    GatherAcc<1> acc1;
    if (gather && a.use_stretch) gather_corners<false, 1>(gl, j0, jt, 0, st, nullptr, acc1);
    if (a.use_bend) {
        __syncthreads();
        uint2 rh = rh0;
        float4 rw = rw0;
        float2 rc = rc0;
        for (int j = tid; j < nhA; j += kT) {
            const uint2 id = rh;
            const float4 w = rw;
            const float2 c = rc;
            {  // next pass's record from either list (index 0 when unused; no branch)
                const int jn = j + kT;
                const size_t qa = jn < nhin ? (size_t)ebeg + jn : 0;
                const size_t qp = (jn >= nhin && jn < nhA) ? (size_t)ebeg2 + jn - nhin : 0;
                const uint2 ha_ = A.hin_idx[qa], hp_ = P.phin_idx[qp];
                const float4 wa_ = A.hin_w[qa], wp_ = P.phin_w[qp];
                const float2 ca_ = A.hin_c[qa], cp_ = P.phin_c[qp];
                rh = jn < nhin ? ha_ : hp_;
                rw = jn < nhin ? wa_ : wp_;
                rc = jn < nhin ? ca_ : cp_;
            }
            const int q0 = id.x & 0xFFFF, q1 = id.x >> 16, q2 = id.y & 0xFFFF, q3 = id.y >> 16;
            f3 p, pe;
            const d3 X[4] = {ldsX(sX, L, q0), ldsX(sX, L, q1), ldsX(sX, L, q2), ldsX(sX, L, q3)};
            pd_hinge(X, w, c.x, p, pe);
            skw[j] = c.y * w.x;
            skw[P.phmax + j] = c.y * w.y;
            skw[2 * P.phmax + j] = c.y * w.z;
            skw[3 * P.phmax + j] = c.y * w.w;
            sts3(sq, j, pe);
        }
        __syncthreads();
        if (gather) gather_hinges<true, false, 1>(gl, jt, j1, 0, skw, P.phmax, sq, nullptr, acc1);
    }
    const f3 ar1 = gather_total<1>(acc1.r);
    // x_{k+1} of own and ring-1 vertices; the position table is free (its
    // last reads were in the element loops, before the barrier above)
    d3 x1 = d3{0.0, 0.0, 0.0};
    if (cv) {
        if (opin >= 0) {
            x1 = to_d(ld3(a.pin_target, opin)) - to_d(ob);
        } else {
            d3 r;
            x1 = pd_update(ar1, om, odg, oy, od, oxp, c0.omega, c0.blend, gamma, r);
        }
        if (own) st3d(static_cast<double*>(a.out), g, x1);
        stsX(sX, L, tid, to_d(ob) + x1);
    }
    __syncthreads();

    // ---- iteration k+1: elements of own vertices -------------------------------
    GatherAcc<1> acc2;
    const bool gather2 = own && opin < 0;
    if (a.use_stretch) {
        uint2 rt = rt0;
        float4 rG = rG0;
        float rwA = rwA0;
        for (int j = tid; j < ntri; j += kT) {
            const uint2 id = rt;
            const float4 G = rG;
            const float wA = rwA;
            {  // next pass's record (index 0 after the last pass; no branch)
                const size_t jn = j + kT < ntri ? (size_t)tbeg + j + kT : 0;
                rt = A.tri_idx[jn];
                rG = A.tri_c0[jn];
                rwA = A.tri_c1[jn];
            }
            const int l0 = id.x & 0xFFFF, l1 = id.x >> 16, l2 = id.y & 0xFFFF;
            f3 cr[3], c[3];
            pd_triangle(ldsX(sX, L, l0), ldsX(sX, L, l1), ldsX(sX, L, l2), G, wA, cr, c, kCheck);
            if (kCheck) {
                float* q = st2 + (size_t)j * 9;
                sts3(q, 0, cr[0]);
                sts3(q, 1, cr[1]);
                sts3(q, 2, cr[2]);
            }
            float* o = st + (size_t)j * 9;
            sts3(o, 0, c[0]);
            sts3(o, 1, c[1]);
            sts3(o, 2, c[2]);
        }
    }
    __syncthreads();
    const int k0 = (int)((unsigned)vi.w & 0xFFFFu), kt = k0 + (int)(((unsigned)vi.w >> 16) & 0xFFu),
              k1 = k0 + (int)((unsigned)vi.w >> 24);
    if (gather2 && a.use_stretch) gather_corners<kCheck, 1>(sg, k0, kt, 0, st, st2, acc2);
    if (a.use_bend) {
        __syncthreads();
        uint2 rh = rh0;
        float4 rw = rw0;
        float2 rc = rc0;
        for (int j = tid; j < nhin; j += kT) {
            const uint2 id = rh;
            const float4 w = rw;
            const float2 c = rc;
            {  // next pass's record (index 0 after the last pass; no branch)
                const size_t jn = j + kT < nhin ? (size_t)ebeg + j + kT : 0;
                rh = A.hin_idx[jn];
                rw = A.hin_w[jn];
                rc = A.hin_c[jn];
            }
            const int q0 = id.x & 0xFFFF, q1 = id.x >> 16, q2 = id.y & 0xFFFF, q3 = id.y >> 16;
            f3 p, pe;
            const d3 X[4] = {ldsX(sX, L, q0), ldsX(sX, L, q1), ldsX(sX, L, q2), ldsX(sX, L, q3)};
            pd_hinge(X, w, c.x, p, pe);
            skw[j] = c.y * w.x;
            skw[P.phmax + j] = c.y * w.y;
            skw[2 * P.phmax + j] = c.y * w.z;
            skw[3 * P.phmax + j] = c.y * w.w;
            sts3(sq, j, pe);
            if (kCheck) sts3(sp, j, p);
        }
        __syncthreads();
        if (gather2) gather_hinges<true, kCheck, 1>(sg, kt, k1, 0, skw, P.phmax, sq, sp, acc2);
    }
    const f3 ar = gather_total<1>(acc2.r);
    const f3 ab = kCheck ? gather_total<1>(acc2.b) : mk(0.0f, 0.0f, 0.0f);

    // ---- own vertices: x_{k+2}, pins, residual ---------------------------------
    double rr = 0.0, bb = 0.0;
    if (own) {
        d3 x2;
        if (opin >= 0) {
            x2 = to_d(ld3(a.pin_target, opin)) - to_d(ob);
        } else {
            d3 r;
            x2 = pd_update(ar, om, odg, oy, x1, od, c1.omega, c1.blend, gamma, r);
            if (kCheck) {
                pd_residual_terms(r, ab, om, ob, oy, rr, bb);
            }
        }
        st3d(static_cast<double*>(P.out2), g, x2);
    }
    if (kCheck) fused_block_sum2<kT>(rr, bb, a.partials, tile);
}

__global__ void __launch_bounds__(256) k_fill_u32(unsigned* __restrict__ p, int n, unsigned v)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// ---- persistent PD iteration kernel ------------------------------------------
//
// One cooperative launch runs iterations [k0, k1) of a frame under device
// control (FrameCtl): every tile's workgroup stays resident, keeps its element
// records, gather list, halo ids and halo base positions in LDS and its own
// vertices' state (x_k, x_{k-1}, constants) in registers. Only the iterate
// goes through memory: each iteration writes x_{k+1} of the own vertices and
// reads x_k of the halo. The arithmetic and every summation order are those
// of k_fused_iter and k_reduce_partials, so both paths give the same iterates
// bit for bit.
//
// Synchronization. A tile reads the halo of its neighbour tiles (those owning
// a vertex of its halo; the relation is symmetric, as elements are shared),
// so iteration k of a tile waits only until every neighbour has finished
// iteration k-1, as counted by the neighbour's progress word. A tile is then
// never more than one iteration ahead of a neighbour, so the four rotating
// iterate buffers are never overwritten while read. Iterates and progress
// words are written and read as relaxed device-scope atomics, which are
// coherent across the XCDs without cache write-backs: a workgroup waits for
// its iterate stores to complete before it publishes its progress. (A
// device-scope release or acquire fence writes back or invalidates the
// XCD's L2 and serializes over the XCD's workgroups: measured 21.8 us per
// grid barrier with one per workgroup at 256 workgroups, 2.06 us without.)
// The Chebyshev state advances in every workgroup by the same recurrence as
// the device control's; a check iteration adds the fixed-order FP64 residual
// sum and the control decision (workgroup 0, whose writes it releases once)
// between two grid barriers, after which every workgroup reloads the state.

struct PersistArgs {
    FusedKernelArgs A;   // plan and frame arrays (A.it.x / xkm1 / out unused)
    double* it[kIterBufs];  // the rotating iterate buffers (tile order)
    int k0, k1, K, check_every;
    FrameCtl* ctl;
    double* partials;    // [num_tiles * 2]
    double* hist;        // residual history
    unsigned* bar;       // barrier words (kBarWords unsigned, zero at rest)
    unsigned* prog;      // [num_tiles * kBarLine] progress words (iterations done)
    const int* nbr_ptr;  // [num_tiles + 1] neighbour tiles of each tile
    const int* nbr;
    int off_rt, off_rh, off_hb;  // LDS byte offsets: tri records, hinge records, halo ids + base
};

constexpr long kSpinLimit = 1l << 24;

__device__ __forceinline__ double ld_coh(const double* p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ d3 ld3d_coh(const double* p, size_t i)
{
    return d3{ld_coh(p + 3 * i), ld_coh(p + 3 * i + 1), ld_coh(p + 3 * i + 2)};
}
__device__ __forceinline__ void st_coh(double* p, double v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st3d_coh(double* p, size_t i, d3 v)
{
    st_coh(p + 3 * i, v.x);
    st_coh(p + 3 * i + 1, v.y);
    st_coh(p + 3 * i + 2, v.z);
}
__device__ __forceinline__ int ctl_load_int(const int* p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ float ctl_load_float(const float* p)
{
    return __int_as_float(ctl_load_int(reinterpret_cast<const int*>(p)));
}

// Grid-wide barrier of a cooperative launch, at the check iterations: a
// two-level arrival (workgroup b counts on the word of its group b % 8, the
// group's last arrival on the global word, whose last arrival resets it and
// advances the generation word that every workgroup waits on), all as relaxed
// device-scope atomics, each word on a 128-byte line of its own. The data a
// barrier orders is exchanged through coherent atomics (or released once by
// its writer), so the barrier itself writes back and invalidates nothing. A
// wait longer than kSpinLimit sleeps raises the error flag `err`, which every
// later wait honours, so a stuck launch still ends. Returns false once the
// flag is up.
__device__ __forceinline__ bool grid_barrier(unsigned* bar, int* err, unsigned nblocks, int* sflag)
{
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned* gen_w = bar + kBarGen * kBarLine;
        unsigned* all_w = bar + kBarAll * kBarLine;
        const unsigned grp = blockIdx.x % kBarGroups;
        const unsigned ngrp = nblocks < kBarGroups ? nblocks : kBarGroups;
        const unsigned in_grp = nblocks / kBarGroups + (grp < nblocks % kBarGroups ? 1u : 0u);
        unsigned* grp_w = bar + grp * kBarLine;
        const unsigned gen = __hip_atomic_load(gen_w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        // this workgroup's earlier stores (coherent atomics) have completed
        __builtin_amdgcn_s_waitcnt(0);
        bool last = false;
        const unsigned a1 =
            __hip_atomic_fetch_add(grp_w, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) + 1u;
        if (a1 == in_grp) {
            __hip_atomic_store(grp_w, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            const unsigned a2 =
                __hip_atomic_fetch_add(all_w, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) + 1u;
            last = a2 == ngrp;
        }
        if (last) {
            __hip_atomic_store(all_w, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __builtin_amdgcn_s_waitcnt(0);
            __hip_atomic_store(gen_w, gen + 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        } else {
            long spins = 0;
            while (__hip_atomic_load(gen_w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == gen) {
                if (__hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) break;
                if (++spins > kSpinLimit) {
                    __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    break;
                }
                __builtin_amdgcn_s_sleep(1);
            }
        }
        *sflag = __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) ? 1 : 0;
    }
    __syncthreads();
    return *sflag == 0;
}

// Waits until every neighbour tile of `tile` has finished `k` iterations
// (thread i polls neighbour i; at most kT neighbours, checked on the host).
// The same time-out as the grid barrier. Returns false once the flag is up.
__device__ __forceinline__ bool wait_neighbours(const unsigned* prog, const int* nbr, int nb,
                                                unsigned k, int* err, int* sflag)
{
    if (threadIdx.x == 0) *sflag = 0;
    __syncthreads();
    if ((int)threadIdx.x < nb) {
        const unsigned* w = prog + (size_t)nbr[threadIdx.x] * kBarLine;
        long spins = 0;
        while (__hip_atomic_load(w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < k) {
            if (__hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
                *sflag = 1;
                break;
            }
            if (++spins > kSpinLimit) {
                __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                *sflag = 1;
                break;
            }
            __builtin_amdgcn_s_sleep(1);
        }
    }
    __syncthreads();
    return *sflag == 0;
}

// The persistent kernel's residual partials: the block sum of
// fused_block_sum2, stored as coherent atomics.
template <int kT>
__device__ void persist_block_sum2(double r, double b, double* out, int slot)
{
    __shared__ double sr[kT / 64], sb[kT / 64];
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
        for (int i = 0; i < kT / 64; ++i) {
            tr += sr[i];
            tb += sb[i];
        }
        st_coh(out + 2 * slot + 0, tr);
        st_coh(out + 2 * slot + 1, tb);
    }
}

// kG adjacent lanes share an own vertex in the gathers, as in k_fused_iter
// (same bits for every kG); the group's first lane holds the vertex's state.
template <int kT, int kG>
__global__ void __launch_bounds__(kT, 4) k_fused_persist(PersistArgs P)
{
    extern __shared__ double smem[];
    __shared__ double red_r[kT], red_b[kT];
    __shared__ int s_flag;
    const FusedKernelArgs& A = P.A;
    const FusedIterArgs& a = A.it;
    const int tid = threadIdx.x;
    const int L = A.lmax;
    const int tile = A.xcd_remap ? xcd_tile(blockIdx.x, A.num_tiles) : (int)blockIdx.x;
    const int vbeg = tile * A.tile_verts;
    const int nown = min(A.tile_verts, A.N - vbeg);
    int* const err = &P.ctl->barrier_error;

    char* lds = reinterpret_cast<char*>(smem);
    double* sX = smem;
    unsigned* sgw = reinterpret_cast<unsigned*>(lds + A.off_g);
    const unsigned short* sg = reinterpret_cast<const unsigned short*>(sgw);
    float* st = reinterpret_cast<float*>(lds + A.off_r);
    float* st2 = st + A.tmax * 9;
    float* skw = st;
    float* sq = skw + A.hmax * 4;
    float* sp = sq + A.hmax * 3;
    uint2* rti = reinterpret_cast<uint2*>(lds + P.off_rt);
    float4* rtg = reinterpret_cast<float4*>(rti + A.tmax + (A.tmax & 1));
    float* rtw = reinterpret_cast<float*>(rtg + A.tmax);
    uint2* rhi = reinterpret_cast<uint2*>(lds + P.off_rh);
    float4* rhw = reinterpret_cast<float4*>(rhi + A.hmax + (A.hmax & 1));
    float2* rhc = reinterpret_cast<float2*>(rhw + A.hmax);
    int* shid = reinterpret_cast<int*>(lds + P.off_hb);
    float* shb = reinterpret_cast<float*>(shid + A.hstride);

    // ---- setup: tile data into LDS, own vertices into registers --------------
    const int* ti = A.tile_info + (size_t)tile * kTileInfo;
    const int nhalo = ti[0], tbeg = ti[1], ntri = ti[2], ebeg = ti[3], nhin = ti[4];
    const int gbeg = ti[5], glen = ti[6];
    const int nb0 = P.nbr_ptr[tile], nb = P.nbr_ptr[tile + 1] - nb0;
    const int* nbr = P.nbr + nb0;
    for (int j = tid; j < ntri; j += kT) {
        rti[j] = A.tri_idx[tbeg + j];
        rtg[j] = A.tri_c0[tbeg + j];
        rtw[j] = A.tri_c1[tbeg + j];
    }
    for (int j = tid; j < nhin; j += kT) {
        rhi[j] = A.hin_idx[ebeg + j];
        rhw[j] = A.hin_w[ebeg + j];
        rhc[j] = A.hin_c[ebeg + j];
    }
    for (int j = tid; j < (glen + 1) >> 1; j += kT) sgw[j] = A.gl_words[gbeg + j];
    const int* hids = A.halo + (size_t)tile * A.hstride;
    for (int i = tid; i < nhalo; i += kT) {
        const int hg = hids[i];
        shid[i] = hg;
        const f3 hb = ld3(a.base, hg);
        shb[3 * i + 0] = hb.x;
        shb[3 * i + 1] = hb.y;
        shb[3 * i + 2] = hb.z;
    }
    // own vertices: constants in LDS, the two FP64 iterates in registers
    int4* sov = reinterpret_cast<int4*>(shb + 3 * A.hstride);  // 16 * hstride bytes in
    float* sob = reinterpret_cast<float*>(sov + A.tile_verts);
    float* soy = sob + 3 * A.tile_verts;
    const int vl = tid / kG, lane_p = tid % kG;  // own vertex of the lane's group
    const bool own = vl < nown;
    const bool lead = own && lane_p == 0;
    const size_t g = (size_t)vbeg + (own ? vl : 0);
    d3 od = d3{0.0, 0.0, 0.0}, oxp = d3{0.0, 0.0, 0.0};
    if (lead) {
        sov[vl] = a.vinfo[g];
        sts3(sob, vl, ld3(a.base, g));
        sts3(soy, vl, ld3(a.y, g));
        od = ld3d(P.it[P.k0 % kIterBufs], g);
        oxp = ld3d(P.it[(P.k0 + kIterBufs - 1) % kIterBufs], g);
    }
    // control state of iteration k0 (the launch boundary orders it)
    if (ctl_load_int(&P.ctl->stop)) return;
    CtlSlot cs;
    cs.omega = ctl_load_float(&P.ctl->slot[P.k0 & 1].omega);
    cs.blend = ctl_load_int(&P.ctl->slot[P.k0 & 1].blend);
    cs.ck = ctl_load_int(&P.ctl->slot[P.k0 & 1].ck);
    cs.k = ctl_load_int(&P.ctl->slot[P.k0 & 1].k);
    float rho = ctl_load_float(&P.ctl->rho);
    const int delay = ctl_load_int(&P.ctl->delay), use_cheb = ctl_load_int(&P.ctl->use_cheb);
    __syncthreads();

    for (int k = P.k0; k < P.k1; ++k) {
        const int par = k & 1;
        // block 0 keeps the device control's slots as the launches do
        const CtlSlot cn = ctl_next_rho(rho, delay, use_cheb, cs);
        if (blockIdx.x == 0 && tid == 0) P.ctl->slot[par ^ 1] = cn;
        const bool check = ((k + 1) % P.check_every == 0) || (k == P.K - 1);
        const double* xk = P.it[k % kIterBufs];
        double* out = P.it[(k + 1) % kIterBufs];

        // ---- the neighbours have written x_k
        if (k > P.k0 && !wait_neighbours(P.prog, nbr, nb, (unsigned)k, err, &s_flag)) break;

        // ---- positions of the tile: own from registers, halo from memory
        for (int i = tid; i < nhalo; i += kT) {
            const d3 hx = ld3d_coh(xk, shid[i]);
            const d3 hb = d3{(double)shb[3 * i], (double)shb[3 * i + 1], (double)shb[3 * i + 2]};
            stsX(sX, L, nown + i, hb + hx);
        }
        if (lead) stsX(sX, L, vl, to_d(lds3(sob, vl)) + od);
        __syncthreads();

        // ---- triangles -> corner forces
        if (a.use_stretch) {
            for (int j = tid; j < ntri; j += kT) {
                const uint2 id = rti[j];
                const int l0 = id.x & 0xFFFF, l1 = id.x >> 16, l2 = id.y & 0xFFFF;
                f3 c[3], cr[3];
                pd_triangle(ldsX(sX, L, l0), ldsX(sX, L, l1), ldsX(sX, L, l2), rtg[j], rtw[j], cr,
                            c, check);
                if (check) {
                    float* q = st2 + (size_t)j * 9;
                    sts3(q, 0, cr[0]);
                    sts3(q, 1, cr[1]);
                    sts3(q, 2, cr[2]);
                }
                float* o = st + (size_t)j * 9;
                sts3(o, 0, c[0]);
                sts3(o, 1, c[1]);
                sts3(o, 2, c[2]);
            }
        }
        __syncthreads();
        const int4 vi = own ? sov[vl] : make_int4(0, 0, -1, 0);
        const unsigned meta = (unsigned)vi.w;
        const int j0 = (int)(meta & 0xFFFFu), jt = j0 + (int)((meta >> 16) & 0xFFu),
                  j1 = j0 + (int)(meta >> 24);
        const bool gather = own && vi.z < 0;
        GatherAcc<kG> acc;
        if (gather && a.use_stretch) {
            if (check) gather_corners<true, kG>(sg, j0, jt, lane_p, st, st2, acc);
            else       gather_corners<false, kG>(sg, j0, jt, lane_p, st, st2, acc);
        }
        if (a.use_bend) {
            __syncthreads();
            for (int j = tid; j < nhin; j += kT) {
                const uint2 id = rhi[j];
                const float4 w = rhw[j];
                const float2 c = rhc[j];
                const int q0 = id.x & 0xFFFF, q1 = id.x >> 16, q2 = id.y & 0xFFFF, q3 = id.y >> 16;
                const d3 X[4] = {ldsX(sX, L, q0), ldsX(sX, L, q1), ldsX(sX, L, q2), ldsX(sX, L, q3)};
                f3 pp, pe;
                pd_hinge(X, w, c.x, pp, pe);
                skw[j] = c.y * w.x;
                skw[A.hmax + j] = c.y * w.y;
                skw[2 * A.hmax + j] = c.y * w.z;
                skw[3 * A.hmax + j] = c.y * w.w;
                sts3(sq, j, pe);
                if (check) sts3(sp, j, pp);
            }
            __syncthreads();
            if (gather) {
                if (check) gather_hinges<true, true, kG>(sg, jt, j1, lane_p, skw, A.hmax, sq, sp, acc);
                else       gather_hinges<true, false, kG>(sg, jt, j1, lane_p, skw, A.hmax, sq, sp, acc);
            }
        }
        const f3 ar = gather_total<kG>(acc.r);
        const f3 ab = gather_total<kG>(acc.b);

        // ---- own vertices: update, pins, residual
        double rr = 0.0, bb = 0.0;
        if (lead) {
            const f3 ob = lds3(sob, vl);
            const float om = __int_as_float(vi.x), odg = __int_as_float(vi.y);
            d3 xn;
            if (vi.z >= 0) {
                xn = to_d(ld3(a.pin_target, vi.z)) - to_d(ob);
            } else {
                const f3 oy = lds3(soy, vl);
                d3 r;
                xn = pd_update(ar, om, odg, oy, od, oxp, cs.omega, cs.blend, relax_gamma(a), r);
                if (check) {
                    pd_residual_terms(r, ab, om, ob, oy, rr, bb);
                }
            }
            st3d_coh(out, g, xn);
            oxp = od;
            od = xn;
        }
        cs = cn;
        // every lane's iterate stores have completed (a workgroup barrier
        // orders execution, not the completion of other waves' stores)
        __builtin_amdgcn_s_waitcnt(0);
        if (check) {
            if (kG > 1) {
                // vertex i's terms in thread i, as with one lane per vertex
                // (the position table is free: its last reads came before
                // the barrier ahead of the hinge sums)
                __syncthreads();
                if (lead) {
                    sX[vl] = rr;
                    sX[L + vl] = bb;
                }
                __syncthreads();
                rr = tid < nown ? sX[tid] : 0.0;
                bb = tid < nown ? sX[L + tid] : 0.0;
            }
            persist_block_sum2<kT>(rr, bb, P.partials, tile);
            if (!grid_barrier(P.bar, err, gridDim.x, &s_flag)) break;
            if (blockIdx.x == 0) {
                // fixed-order sum of the tile partials (k_reduce_partials)
                double r = 0.0, b = 0.0;
                for (int i = tid; i < A.num_tiles; i += kT) {
                    r += ld_coh(P.partials + 2 * i + 0);
                    b += ld_coh(P.partials + 2 * i + 1);
                }
                red_r[tid] = r;
                red_b[tid] = b;
                __syncthreads();
                for (int s2 = kT / 2; s2 > 0; s2 >>= 1) {
                    if (tid < s2) {
                        red_r[tid] += red_r[tid + s2];
                        red_b[tid] += red_b[tid + s2];
                    }
                    __syncthreads();
                }
                if (tid == 0) {
                    ctl_check(P.ctl, par, red_r[0], red_b[0], P.hist,
                              red_r[0] <= P.ctl->tol2 * red_b[0]);
                    // the control block and history, once, for every XCD
                    __threadfence();
                }
            }
            if (!grid_barrier(P.bar, err, gridDim.x, &s_flag)) break;
            if (ctl_load_int(&P.ctl->stop)) break;
            // a restart may have replaced the next state and lowered rho
            cs.omega = ctl_load_float(&P.ctl->slot[par ^ 1].omega);
            cs.blend = ctl_load_int(&P.ctl->slot[par ^ 1].blend);
            cs.ck = ctl_load_int(&P.ctl->slot[par ^ 1].ck);
            cs.k = ctl_load_int(&P.ctl->slot[par ^ 1].k);
            rho = ctl_load_float(&P.ctl->rho);
        } else {
            __syncthreads();
        }
        // publish x_{k+1} of this tile
        if (tid == 0)
            __hip_atomic_store(P.prog + (size_t)tile * kBarLine, (unsigned)(k + 1),
                               __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

// ---- tile-order frame kernels ------------------------------------------------

// Tile-order padding (perm < 0) gets a zero state and an inert record: no
// gather list, zero mass and diagonal, so its iterate stays 0.
__global__ void __launch_bounds__(256)
k_fused_prologue(const int* __restrict__ perm, const float* __restrict__ pos,
                 const float* __restrict__ vel, const float* __restrict__ guess,
                 const float* __restrict__ mass_in, const float* __restrict__ diag_in,
                 const unsigned* __restrict__ gl_meta, const int* __restrict__ tile_cloth,
                 const float* __restrict__ cloth_accel, float* __restrict__ xprev,
                 float* __restrict__ y, void* __restrict__ x0, int4* __restrict__ vinfo, int N,
                 int tile_verts, float dt, float h2, float acc0, float acc1, float acc2,
                 float h2a0, float h2a1, float h2a2, int disp)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const int o = perm[i];
    if (o < 0) {
        st3(xprev, i, mk(0.0f, 0.0f, 0.0f));
        st3(y, i, mk(0.0f, 0.0f, 0.0f));
        if (disp) st3d(static_cast<double*>(x0), i, d3{0.0, 0.0, 0.0});
        else      st3(static_cast<float*>(x0), i, mk(0.0f, 0.0f, 0.0f));
        vinfo[i] = make_int4(0, 0, -1, 0);
        return;
    }
    if (cloth_accel) {
        // per-cloth acceleration on top of the frame's uniform one
        const f3 ca = ld3(cloth_accel, tile_cloth[i / tile_verts]);
        h2a0 = h2 * (acc0 + ca.x);
        h2a1 = h2 * (acc1 + ca.y);
        h2a2 = h2 * (acc2 + ca.z);
    }
    const f3 x = ld3(pos, o), u = ld3(vel, o);
    st3(xprev, i, x);
    const f3 dy = mk(dt * u.x + h2a0, dt * u.y + h2a1, dt * u.z + h2a2);
    if (disp) {
        st3(y, i, dy);
        st3d(static_cast<double*>(x0), i,
             guess ? to_d(ld3(guess, o)) - to_d(x) : d3{0.0, 0.0, 0.0});
    } else {
        st3(y, i, predict_nc(x, u, dt, h2a0, h2a1, h2a2));
        st3(static_cast<float*>(x0), i, guess ? ld3(guess, o) : x);
    }
    vinfo[i] = make_int4(__float_as_int(mass_in[o]), __float_as_int(diag_in[o]), -1,
                         (int)gl_meta[i]);
}

__global__ void __launch_bounds__(256)
k_fused_pins(const int* __restrict__ inv_perm, const int* __restrict__ idx,
             const float* __restrict__ target, const float* __restrict__ xprev,
             int4* __restrict__ vinfo, void* __restrict__ x0, int P, int disp)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= P) return;
    const int i = inv_perm[idx[k]];
    vinfo[i].z = k;
    if (disp)
        st3d(static_cast<double*>(x0), i, to_d(ld3(target, k)) - to_d(ld3(xprev, i)));
}

__global__ void __launch_bounds__(256)
k_fused_epilogue(const int* __restrict__ perm, const void* __restrict__ xf,
                 const float* __restrict__ xprev, const int4* __restrict__ vinfo,
                 const float* __restrict__ target, float* __restrict__ pos,
                 float* __restrict__ vel, int N, float scale, int disp)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const int o = perm[i];
    if (o < 0) return;  // padding
    const int slot = vinfo[i].z;
    if (slot >= 0) {
        st3(pos, o, ld3(target, slot));
        st3(vel, o, mk(0.0f, 0.0f, 0.0f));
        return;
    }
    const f3 xp = ld3(xprev, i);
    if (disp) {
        const d3 d = ld3d(static_cast<const double*>(xf), i);
        st3(pos, o, to_f(to_d(xp) + d));
        st3(vel, o, to_f(d * (double)scale));
    } else {
        const f3 x = ld3(static_cast<const float*>(xf), i);
        st3(pos, o, x);
        st3(vel, o, mk((x.x - xp.x) * scale, (x.y - xp.y) * scale, (x.z - xp.z) * scale));
    }
}

__global__ void __launch_bounds__(256)
k_fused_trace(const int* __restrict__ perm, const void* __restrict__ it,
              const float* __restrict__ xprev, float* __restrict__ out, int N, int disp)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N || perm[i] < 0) return;
    st3(out, perm[i],
        disp ? to_f(to_d(ld3(xprev, i)) + ld3d(static_cast<const double*>(it), i))
             : ld3(static_cast<const float*>(it), i));
}

__global__ void __launch_bounds__(256)
k_pack_tris(const int* __restrict__ gid, const float* __restrict__ dminv,
            const float* __restrict__ area, const float* __restrict__ kst,
            float4* __restrict__ c0, float* __restrict__ c1, int n, float h2)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const size_t t = (size_t)gid[j];
    c0[j] = make_float4(dminv[t * 4 + 0], dminv[t * 4 + 1], dminv[t * 4 + 2], dminv[t * 4 + 3]);
    c1[j] = kst ? kst[t] * area[t] * h2 : 0.0f;
}

__global__ void __launch_bounds__(256)
k_pack_hinges(const int* __restrict__ gid, const float* __restrict__ w,
              const float* __restrict__ nrest, const float* __restrict__ bk,
              float4* __restrict__ hw, float2* __restrict__ hc, int n, float h2)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const size_t e = (size_t)gid[j];
    hw[j] = make_float4(w[e * 4 + 0], w[e * 4 + 1], w[e * 4 + 2], w[e * 4 + 3]);
    hc[j] = make_float2(nrest[e], bk[e] * h2);
}

size_t align16(size_t b) { return (b + 15) & ~(size_t)15; }

// LDS byte offsets of the fused kernel: position table, gather list, element
// region (triangle corner forces, then hinge vectors).
struct LdsLayout {
    size_t off_g, off_r, total;
};

LdsLayout lds_layout(const FusedPlan& p, int mode, bool check, bool split = false)
{
    const bool pd = (mode == CLH_MODE_PD), two = pd && check;
    LdsLayout L;
    L.off_g = align16((size_t)p.lmax * (pd ? 24 : 12));
    L.off_r = L.off_g + align16((size_t)(p.gmax + 1) / 2 * 4);
    const size_t ra = (size_t)p.tmax * 9 * (two ? 2 : 1) * sizeof(float);
    const size_t rb = (size_t)p.hmax * (4 + 3 + (two ? 3 : 0)) * sizeof(float);
    // split: the hinge region follows the triangle region (ra is a multiple
    // of 4 bytes, as the hinge arrays need)
    L.total = L.off_r + (split ? align16(ra + rb) : align16(std::max(ra, rb)));
    return L;
}

size_t lds_bytes(const FusedPlan& p, int mode, bool check) { return lds_layout(p, mode, check).total; }

// Separate triangle and hinge regions for this launch: PD mode, asked for by
// the plan, and within the workgroup's LDS.
bool use_split(const FusedPlan& p, int mode, bool check)
{
    return p.split && mode == CLH_MODE_PD &&
           lds_layout(p, mode, check, true).total <= (size_t)p.lds_limit;
}

// LDS byte offsets of the two-iteration kernel: position table (own, ring 1,
// ring 2), own gather list, ring-1 gather list, element region (triangles of
// both lists plus the second iteration's R parts on a check; or hinges of
// both lists plus the second iteration's projections).
struct PairLdsLayout {
    size_t off_g, off_g2, off_r, total;
};

PairLdsLayout pair_lds_layout(const FusedPlan& p, bool check)
{
    PairLdsLayout L;
    L.off_g = align16((size_t)p.plmax * 24);
    L.off_g2 = L.off_g + align16((size_t)(p.gmax + 1) / 2 * 4);
    L.off_r = L.off_g2 + align16((size_t)(p.pgmax + 1) / 2 * 4);
    const size_t ra = ((size_t)p.ptmax * 9 + (check ? (size_t)p.tmax * 9 : 0)) * sizeof(float);
    const size_t rb = ((size_t)p.phmax * 7 + (check ? (size_t)p.hmax * 3 : 0)) * sizeof(float);
    L.total = L.off_r + align16(std::max(ra, rb));
    return L;
}

void fill_kernel_args(const FusedPlan& p, const FusedIterArgs& a, FusedKernelArgs& A)
{
    A.it = a;
    A.tile_info = p.tile_info.ptr;
    A.halo = p.halo.ptr;
    A.tri_idx = p.tri_idx.ptr;
    A.tri_c0 = p.tri_c0.ptr;
    A.tri_c1 = p.tri_c1.ptr;
    A.hin_idx = p.hin_idx.ptr;
    A.hin_w = p.hin_w.ptr;
    A.hin_c = p.hin_c.ptr;
    A.gl_words = p.gl_words.ptr;
    A.tile_cloth = p.tile_cloth.ptr;
    A.cloth_tiles = p.cloth_tiles.ptr;
    A.N = p.N;
    A.tile_verts = p.tile_verts;
    A.hstride = p.hstride;
    A.lmax = p.lmax;
    A.tmax = p.tmax;
    A.hmax = p.hmax;
    A.num_tiles = p.num_tiles;
    A.xcd_remap = p.xcd_remap;
}

// LDS of the persistent kernel: the iteration kernel's layout (PD, residual
// variant), then the tile's triangle and hinge records, halo ids and halo
// base positions.
struct PersistLayout {
    LdsLayout it;
    size_t off_rt, off_rh, off_hb, total;
};

PersistLayout persist_layout(const FusedPlan& p)
{
    PersistLayout o;
    o.it = lds_layout(p, CLH_MODE_PD, true);
    const size_t t = (size_t)p.tmax, h = (size_t)p.hmax;
    o.off_rt = align16(o.it.total);
    o.off_rh = align16(o.off_rt + (t + (t & 1)) * 8 + t * 16 + t * 4);
    o.off_hb = align16(o.off_rh + (h + (h & 1)) * 8 + h * 16 + h * 8);
    // halo ids and base positions, then the own vertices' records, base and y
    o.total = align16(o.off_hb + (size_t)p.hstride * 16 + (size_t)p.tile_verts * (16 + 12 + 12));
    return o;
}

unsigned blocks_for(int n) { return (unsigned)((n + 255) / 256); }

}  // namespace

int host_plan_build(HostPlan& h, int N, const std::vector<int>& tris,
                    const std::vector<int>& quads, const std::vector<float>& pos, int tile_verts,
                    const std::vector<int>& seg_in, bool want_pair)
{
    if (N <= 0) return fail(CLH_ERR_UNSUPPORTED, "fused plan: empty mesh");
    for (float v : pos)
        if (!std::isfinite(v)) return fail(CLH_ERR_NUMERIC, "fused plan: non-finite vertex positions");
    const int T = (int)(tris.size() / 3), E = (int)(quads.size() / 4);
    const int V = std::max(32, std::min(tile_verts, kFB));  // one own vertex per thread

    // cloths: contiguous ranges of original vertices; elements stay inside one
    const std::vector<int> seg = seg_in.empty() ? std::vector<int>{0, N} : seg_in;
    const int C = (int)seg.size() - 1;
    if (C < 1 || seg[0] != 0 || seg[(size_t)C] != N)
        return fail(CLH_ERR_INVALID, "fused plan: cloth offsets must run from 0 to num_verts");
    std::vector<int> vcloth((size_t)N);
    for (int c = 0; c < C; ++c) {
        if (seg[(size_t)c + 1] <= seg[(size_t)c])
            return fail(CLH_ERR_INVALID, "fused plan: every cloth needs at least one vertex");
        for (int v = seg[(size_t)c]; v < seg[(size_t)c + 1]; ++v) vcloth[(size_t)v] = c;
    }
    for (int e = 0; e < T; ++e)
        for (int k = 1; k < 3; ++k)
            if (vcloth[(size_t)tris[(size_t)e * 3 + k]] != vcloth[(size_t)tris[(size_t)e * 3]])
                return fail(CLH_ERR_INVALID, "fused plan: a triangle spans two cloths");
    for (int e = 0; e < E; ++e)
        for (int k = 1; k < 4; ++k)
            if (vcloth[(size_t)quads[(size_t)e * 4 + k]] != vcloth[(size_t)quads[(size_t)e * 4]])
                return fail(CLH_ERR_INVALID, "fused plan: a hinge spans two cloths");

    // tile order: each cloth's vertices in Morton order, padded with inert
    // vertices (perm = -1) to a whole number of tiles
    std::vector<int> perm, tile_cloth, cloth_tiles((size_t)C + 1, 0);
    for (int c = 0; c < C; ++c) {
        const std::vector<int> o = morton_order(pos, seg[(size_t)c], seg[(size_t)c + 1]);
        perm.insert(perm.end(), o.begin(), o.end());
        while (perm.size() % (size_t)V) perm.push_back(-1);
        cloth_tiles[(size_t)c + 1] = (int)(perm.size() / (size_t)V);
        while ((int)tile_cloth.size() < cloth_tiles[(size_t)c + 1]) tile_cloth.push_back(c);
    }
    const int Np = (int)perm.size();
    std::vector<int> inv((size_t)N);
    for (int i = 0; i < Np; ++i)
        if (perm[(size_t)i] >= 0) inv[(size_t)perm[(size_t)i]] = i;

    std::vector<int> vt_ptr, vt_ids, ve_ptr, ve_ids;
    incidence(tris, 3, N, vt_ptr, vt_ids);
    incidence(quads, 4, N, ve_ptr, ve_ids);

    const int ntiles = Np / V;
    std::vector<int> info((size_t)ntiles * kTileInfo), halo_list, halo_beg((size_t)ntiles + 1, 0);
    std::vector<int> tri_gid, hin_gid;
    std::vector<uint2> tri_idx, hin_idx;
    std::vector<unsigned short> gl_idx;
    std::vector<unsigned> gl_meta((size_t)Np, 0);
    std::vector<int> loc((size_t)Np, -1);  // tile-order id -> local id in the current tile
    int lmax = 0, tmax = 0, hmax = 0, gmax = 0, hstride = 1;
    // two-iteration extension
    bool pair_ok = want_pair;
    std::vector<int> pinfo, phalo_list, phalo_beg((size_t)ntiles + 1, 0), ptri_gid, phin_gid;
    std::vector<uint2> ptri_idx, phin_idx;
    std::vector<unsigned short> pgl_idx;
    std::vector<unsigned> pmeta_list;
    int pstride = 1, plmax = 0, ptmax = 0, phmax = 0, pgmax = 0, pcmax = 0;
    if (pair_ok) pinfo.assign((size_t)ntiles * kPairInfo, 0);

    for (int t = 0; t < ntiles; ++t) {
        const int vbeg = t * V, nown = V;
        std::vector<int> tl, hl, touched;
        for (int i = 0; i < nown; ++i) {
            const int o = perm[(size_t)vbeg + i];
            if (o < 0) continue;  // padding
            for (int j = vt_ptr[(size_t)o]; j < vt_ptr[(size_t)o + 1]; ++j) tl.push_back(vt_ids[(size_t)j]);
            for (int j = ve_ptr[(size_t)o]; j < ve_ptr[(size_t)o + 1]; ++j) hl.push_back(ve_ids[(size_t)j]);
        }
        std::sort(tl.begin(), tl.end());
        tl.erase(std::unique(tl.begin(), tl.end()), tl.end());
        std::sort(hl.begin(), hl.end());
        hl.erase(std::unique(hl.begin(), hl.end()), hl.end());

        for (int i = 0; i < nown; ++i) {
            loc[(size_t)vbeg + i] = i;
            touched.push_back(vbeg + i);
        }
        int nl = nown;
        auto local_of = [&](int orig) {
            const int g = inv[(size_t)orig];
            if (loc[(size_t)g] < 0) {
                loc[(size_t)g] = nl++;
                halo_list.push_back(g);
                touched.push_back(g);
            }
            return loc[(size_t)g];
        };
        const int tbeg = (int)tri_gid.size();
        for (int e : tl) {
            const int a = local_of(tris[(size_t)e * 3]), b = local_of(tris[(size_t)e * 3 + 1]),
                      c = local_of(tris[(size_t)e * 3 + 2]);
            tri_gid.push_back(e);
            tri_idx.push_back(make_uint2((unsigned)a | ((unsigned)b << 16), (unsigned)c | 0xFFFF0000u));
        }
        const int ebeg = (int)hin_gid.size();
        for (int e : hl) {
            const int a = local_of(quads[(size_t)e * 4]), b = local_of(quads[(size_t)e * 4 + 1]),
                      c = local_of(quads[(size_t)e * 4 + 2]), d = local_of(quads[(size_t)e * 4 + 3]);
            hin_gid.push_back(e);
            hin_idx.push_back(make_uint2((unsigned)a | ((unsigned)b << 16), (unsigned)c | ((unsigned)d << 16)));
        }
        if (nl > 65535 || (int)tl.size() * 3 > 65535 || (int)hl.size() * 4 > 65535)
            return fail(CLH_ERR_UNSUPPORTED, "fused plan: tile exceeds 16-bit local indexing");
        halo_beg[(size_t)t + 1] = (int)halo_list.size();

        // gather lists: triangle corners then hinge slots, ascending element
        // id; each tile's list starts on a 32-bit word
        if (gl_idx.size() & 1) gl_idx.push_back(0);
        const size_t gbeg = gl_idx.size();
        for (int i = 0; i < nown; ++i) {
            const int g = vbeg + i, o = perm[(size_t)g];
            if (o < 0) continue;  // padding: no gather (meta 0)
            const size_t start = gl_idx.size() - gbeg;
            int nt = 0;
            for (int j = vt_ptr[(size_t)o]; j < vt_ptr[(size_t)o + 1]; ++j) {
                const int e = vt_ids[(size_t)j];
                const int slot = (int)(std::lower_bound(tl.begin(), tl.end(), e) - tl.begin());
                for (int c = 0; c < 3; ++c)
                    if (tris[(size_t)e * 3 + c] == o) {
                        gl_idx.push_back((unsigned short)(slot * 3 + c));
                        ++nt;
                    }
            }
            for (int j = ve_ptr[(size_t)o]; j < ve_ptr[(size_t)o + 1]; ++j) {
                const int e = ve_ids[(size_t)j];
                const int slot = (int)(std::lower_bound(hl.begin(), hl.end(), e) - hl.begin());
                for (int c = 0; c < 4; ++c)
                    if (quads[(size_t)e * 4 + c] == o) gl_idx.push_back((unsigned short)(slot * 4 + c));
            }
            const size_t nall = gl_idx.size() - gbeg - start;
            if (nall > 255) return fail(CLH_ERR_UNSUPPORTED, "fused plan: vertex valence above 255");
            if (start > 0xFFFF) return fail(CLH_ERR_UNSUPPORTED, "fused plan: tile gather list too long");
            gl_meta[(size_t)g] = (unsigned)start | ((unsigned)nt << 16) | ((unsigned)nall << 24);
        }
        const size_t glen = gl_idx.size() - gbeg;
        if (glen > 0xFFFF + 255) return fail(CLH_ERR_UNSUPPORTED, "fused plan: tile gather list too long");
        int* ti = &info[(size_t)t * kTileInfo];
        ti[0] = nl - nown; ti[1] = tbeg; ti[2] = (int)tl.size(); ti[3] = ebeg; ti[4] = (int)hl.size();
        ti[5] = (int)(gbeg / 2); ti[6] = (int)glen; ti[7] = vbeg;
        gmax = std::max(gmax, (int)glen);
        lmax = std::max(lmax, nl);
        tmax = std::max(tmax, (int)tl.size());
        hmax = std::max(hmax, (int)hl.size());
        hstride = std::max(hstride, nl - nown);

        if (pair_ok) {
            // ring 1 = this tile's halo (local ids nown .. nl-1); the elements
            // touching ring 1 but no own vertex, ascending; their vertices
            // outside own and ring 1 form ring 2 (local ids nl ..)
            const int nh1 = nl - nown;
            const int* r1 = halo_list.data() + halo_beg[(size_t)t];
            std::vector<int> tl2, hl2, r2;
            for (int i = 0; i < nh1; ++i) {
                const int o = perm[(size_t)r1[i]];
                for (int j = vt_ptr[(size_t)o]; j < vt_ptr[(size_t)o + 1]; ++j) tl2.push_back(vt_ids[(size_t)j]);
                for (int j = ve_ptr[(size_t)o]; j < ve_ptr[(size_t)o + 1]; ++j) hl2.push_back(ve_ids[(size_t)j]);
            }
            std::sort(tl2.begin(), tl2.end());
            tl2.erase(std::unique(tl2.begin(), tl2.end()), tl2.end());
            tl2.erase(std::remove_if(tl2.begin(), tl2.end(),
                                     [&](int e) { return std::binary_search(tl.begin(), tl.end(), e); }),
                      tl2.end());
            std::sort(hl2.begin(), hl2.end());
            hl2.erase(std::unique(hl2.begin(), hl2.end()), hl2.end());
            hl2.erase(std::remove_if(hl2.begin(), hl2.end(),
                                     [&](int e) { return std::binary_search(hl.begin(), hl.end(), e); }),
                      hl2.end());
            int nl2 = nl;
            auto local2 = [&](int orig) {
                const int g = inv[(size_t)orig];
                if (loc[(size_t)g] < 0) {
                    loc[(size_t)g] = nl2++;
                    r2.push_back(g);
                    touched.push_back(g);
                }
                return loc[(size_t)g];
            };
            const int tbeg2 = (int)ptri_gid.size();
            for (int e : tl2) {
                const int a = local2(tris[(size_t)e * 3]), b = local2(tris[(size_t)e * 3 + 1]),
                          c = local2(tris[(size_t)e * 3 + 2]);
                ptri_gid.push_back(e);
                ptri_idx.push_back(make_uint2((unsigned)a | ((unsigned)b << 16), (unsigned)c | 0xFFFF0000u));
            }
            const int ebeg2 = (int)phin_gid.size();
            for (int e : hl2) {
                const int a = local2(quads[(size_t)e * 4]), b = local2(quads[(size_t)e * 4 + 1]),
                          c = local2(quads[(size_t)e * 4 + 2]), d = local2(quads[(size_t)e * 4 + 3]);
                phin_gid.push_back(e);
                phin_idx.push_back(make_uint2((unsigned)a | ((unsigned)b << 16), (unsigned)c | ((unsigned)d << 16)));
            }
            // ring-1 gather lists over the slots of both element lists (the
            // own vertices' elements first, then the ring-1-only ones), in
            // ascending element id as for own vertices
            const int nt1 = (int)tl.size(), ne1 = (int)hl.size();
            if (pgl_idx.size() & 1) pgl_idx.push_back(0);
            const size_t pgbeg = pgl_idx.size();
            for (int i = 0; i < nh1 && pair_ok; ++i) {
                const int o = perm[(size_t)r1[i]];
                const size_t start = pgl_idx.size() - pgbeg;
                int nt = 0;
                for (int j = vt_ptr[(size_t)o]; j < vt_ptr[(size_t)o + 1]; ++j) {
                    const int e = vt_ids[(size_t)j];
                    auto it1 = std::lower_bound(tl.begin(), tl.end(), e);
                    const int slot = (it1 != tl.end() && *it1 == e)
                                         ? (int)(it1 - tl.begin())
                                         : nt1 + (int)(std::lower_bound(tl2.begin(), tl2.end(), e) - tl2.begin());
                    for (int c = 0; c < 3; ++c)
                        if (tris[(size_t)e * 3 + c] == o) {
                            pgl_idx.push_back((unsigned short)(slot * 3 + c));
                            ++nt;
                        }
                }
                for (int j = ve_ptr[(size_t)o]; j < ve_ptr[(size_t)o + 1]; ++j) {
                    const int e = ve_ids[(size_t)j];
                    auto it1 = std::lower_bound(hl.begin(), hl.end(), e);
                    const int slot = (it1 != hl.end() && *it1 == e)
                                         ? (int)(it1 - hl.begin())
                                         : ne1 + (int)(std::lower_bound(hl2.begin(), hl2.end(), e) - hl2.begin());
                    for (int c = 0; c < 4; ++c)
                        if (quads[(size_t)e * 4 + c] == o) pgl_idx.push_back((unsigned short)(slot * 4 + c));
                }
                const size_t nall = pgl_idx.size() - pgbeg - start;
                if (nall > 255 || start > 0xFFFF) pair_ok = false;
                pmeta_list.push_back((unsigned)start | ((unsigned)nt << 16) | ((unsigned)nall << 24));
            }
            const size_t pglen = pgl_idx.size() - pgbeg;
            if (nl2 > 65535 || (nt1 + (int)tl2.size()) * 3 > 65535 ||
                (ne1 + (int)hl2.size()) * 4 > 65535 || pglen > 0xFFFF + 255 || nown + nh1 > 512)
                pair_ok = false;
            int* pi = &pinfo[(size_t)t * kPairInfo];
            pi[0] = (int)r2.size(); pi[1] = tbeg2; pi[2] = (int)tl2.size(); pi[3] = ebeg2;
            pi[4] = (int)hl2.size(); pi[5] = (int)(pgbeg / 2); pi[6] = (int)pglen; pi[7] = 0;
            phalo_list.insert(phalo_list.end(), r1, r1 + nh1);
            phalo_list.insert(phalo_list.end(), r2.begin(), r2.end());
            phalo_beg[(size_t)t + 1] = (int)phalo_list.size();
            pstride = std::max(pstride, nl2 - nown);
            plmax = std::max(plmax, nl2);
            ptmax = std::max(ptmax, nt1 + (int)tl2.size());
            phmax = std::max(phmax, ne1 + (int)hl2.size());
            pgmax = std::max(pgmax, (int)pglen);
            pcmax = std::max(pcmax, nl);
        }
        for (int g : touched) loc[(size_t)g] = -1;
    }
    if (gl_idx.size() & 1) gl_idx.push_back(0);
    if (gl_idx.empty()) gl_idx.assign(2, 0);

    // halo ids at a fixed stride per tile, padded with the tile's first own
    // vertex (an in-range id that is never read as a halo vertex)
    std::vector<int> halo((size_t)ntiles * hstride);
    for (int t = 0; t < ntiles; ++t) {
        const int n = halo_beg[(size_t)t + 1] - halo_beg[(size_t)t];
        for (int i = 0; i < hstride; ++i)
            halo[(size_t)t * hstride + i] = i < n ? halo_list[(size_t)halo_beg[(size_t)t] + i] : t * V;
    }

    if (pair_ok) {
        // fixed strides, padded as the halo ids; ring-1 metas at the halo stride
        h.phalo.assign((size_t)ntiles * pstride, 0);
        h.pmeta.assign((size_t)ntiles * hstride, 0u);
        size_t mo = 0;
        for (int t = 0; t < ntiles; ++t) {
            const int n = phalo_beg[(size_t)t + 1] - phalo_beg[(size_t)t];
            for (int i = 0; i < pstride; ++i)
                h.phalo[(size_t)t * pstride + i] = i < n ? phalo_list[(size_t)phalo_beg[(size_t)t] + i] : t * V;
            const int nh1 = info[(size_t)t * kTileInfo];
            for (int i = 0; i < nh1; ++i) h.pmeta[(size_t)t * hstride + i] = pmeta_list[mo++];
        }
        if (pgl_idx.size() & 1) pgl_idx.push_back(0);
        if (pgl_idx.empty()) pgl_idx.assign(2, 0);
        if (ptri_gid.empty()) { ptri_gid.push_back(0); ptri_idx.push_back(make_uint2(0, 0)); }
        if (phin_gid.empty()) { phin_gid.push_back(0); phin_idx.push_back(make_uint2(0, 0)); }
        h.pinfo.swap(pinfo);
        h.ptri_gid.swap(ptri_gid);
        h.phin_gid.swap(phin_gid);
        h.ptri_idx.swap(ptri_idx);
        h.phin_idx.swap(phin_idx);
        h.pgl_idx.swap(pgl_idx);
    }

    FusedPlan& p = h.meta;
    p.pair = pair_ok ? 1 : 0;
    if (pair_ok) {
        p.pstride = pstride;
        p.plmax = plmax;
        p.ptmax = std::max(ptmax, 1);
        p.phmax = std::max(phmax, 1);
        p.pgmax = std::max(pgmax, 2);
        p.pcmax = pcmax;
        // per launch: own x_k, x_{k-1} read and x_{k+1}, x_{k+2} written
        // (24 B each), base, dy, record; per ring-1 entry x_k, x_{k-1}, base,
        // dy, record, id and gather meta; per ring-2 entry x_k, base and id;
        // both element lists; both gather lists; descriptors
        size_t n2 = 0;
        for (int t = 0; t < ntiles; ++t) n2 += (size_t)h.pinfo[(size_t)t * kPairInfo];
        p.pbytes_per_launch = (double)Np * (4 * 24 + 2 * 12 + 16) + (double)halo_list.size() * 96 +
                              (double)n2 * 40 +
                              (double)ntiles * (kTileInfo + kPairInfo + hstride) * 4 +
                              (double)(tri_gid.size() + h.ptri_gid.size()) * 28 +
                              (double)(hin_gid.size() + h.phin_gid.size()) * 32 +
                              (double)(gl_idx.size() + h.pgl_idx.size()) * 2;
    }
    p.N = Np;
    p.num_verts = N;
    p.num_cloths = C;
    p.num_tiles = ntiles;
    p.tile_verts = V;
    p.hstride = hstride;
    p.lmax = lmax;
    p.tmax = std::max(tmax, 1);
    p.hmax = std::max(hmax, 1);
    p.gmax = std::max(gmax, 2);
    p.vmax = V;
    p.tri_factor = T > 0 ? (double)tri_gid.size() / T : 0.0;
    p.bend_factor = E > 0 ? (double)hin_gid.size() / E : 0.0;
    p.xcd_remap = 1;
    // bytes the PD iteration streams: per vertex the double iterates x_k,
    // x_{k-1}, x_{k+1} (24 B each), base and dy (12 B each) and the 16-byte
    // vertex record; per tile its descriptor and padded halo id block; per
    // halo vertex its iterate and base; packed triangle (28 B) and hinge
    // (32 B) records; gather entries (2 B)
    p.bytes_per_iter = (double)Np * (3 * 24 + 2 * 12 + 16) + (double)halo_list.size() * 36 +
                       (double)ntiles * (kTileInfo + hstride) * 4 +
                       (double)tri_gid.size() * 28 + (double)hin_gid.size() * 32 +
                       (double)gl_idx.size() * 2;

    if (tri_gid.empty()) { tri_gid.push_back(0); tri_idx.push_back(make_uint2(0, 0)); }
    if (hin_gid.empty()) { hin_gid.push_back(0); hin_idx.push_back(make_uint2(0, 0)); }
    h.perm.swap(perm);
    h.inv.swap(inv);
    h.info.swap(info);
    h.halo.swap(halo);
    h.tri_gid.swap(tri_gid);
    h.hin_gid.swap(hin_gid);
    h.tri_idx.swap(tri_idx);
    h.hin_idx.swap(hin_idx);
    h.gl_idx.swap(gl_idx);
    h.gl_meta.swap(gl_meta);
    h.tile_cloth.swap(tile_cloth);
    h.cloth_tiles.swap(cloth_tiles);
    // neighbour tiles: owners of the halo vertices (tile t owns the
    // tile-order range [t V, t V + V)), ascending, without t itself
    h.nbr_ptr.assign(1, 0);
    h.nbr.clear();
    int nbr_max = 0;
    std::vector<int> seen;
    for (int t = 0; t < ntiles; ++t) {
        const int nh = h.info[(size_t)t * kTileInfo + 0];
        seen.clear();
        for (int i = 0; i < nh; ++i) {
            const int o = h.halo[(size_t)t * hstride + i] / V;
            if (o != t) seen.push_back(o);
        }
        std::sort(seen.begin(), seen.end());
        seen.erase(std::unique(seen.begin(), seen.end()), seen.end());
        h.nbr.insert(h.nbr.end(), seen.begin(), seen.end());
        h.nbr_ptr.push_back((int)h.nbr.size());
        nbr_max = std::max(nbr_max, (int)seen.size());
    }
    h.meta.nbr_max = nbr_max;
    return CLH_OK;
}

// Checks every index of a host plan against the mesh: the tile order is a
// permutation of the vertices plus padding, each cloth's vertices fill its
// own whole tiles, local ids lie inside the tile's table, halo ids lie
// outside the own range and in the tile's cloth, each local slot names the
// element's vertex, padding has no gather list, and each own vertex's gather
// list equals its incident (element, corner) slots in ascending order (the
// order of the unfused pipeline and the oracle) inside the tile's segment.
// Two-iteration extension of tile t: ring 1 in the pair halo is the tile's
// halo; ring 2 lies outside own and ring 1, without repeats, in the tile's
// cloth; every ring-1-only element names its vertices through local ids of
// the table, touches ring 1 and no own vertex; each ring-1 vertex's gather
// list equals its incident (element, corner) slots in ascending order over
// both element lists.
template <typename ClothOf>
int pair_tile_check(const HostPlan& h, int t, const std::vector<int>& tris,
                    const std::vector<int>& quads, const std::vector<int>& vt_ptr,
                    const std::vector<int>& vt_ids, const std::vector<int>& ve_ptr,
                    const std::vector<int>& ve_ids, ClothOf cloth_of)
{
    const FusedPlan& p = h.meta;
    const int V = p.tile_verts, vbeg = t * V, nown = V;
    const int* ti = &h.info[(size_t)t * kTileInfo];
    const int* pi = &h.pinfo[(size_t)t * kPairInfo];
    const int nh1 = ti[0], tbeg = ti[1], ntri = ti[2], ebeg = ti[3], nhin = ti[4];
    const int nh2 = pi[0], tbeg2 = pi[1], ntri2 = pi[2], ebeg2 = pi[3], nhin2 = pi[4];
    const size_t pgbeg = (size_t)pi[5] * 2;
    const int pglen = pi[6];
    const int nl = nown + nh1, nl2 = nl + nh2;
    if (nh2 < 0 || nh1 + nh2 > p.pstride || nl2 > p.plmax || ntri + ntri2 > p.ptmax ||
        nhin + nhin2 > p.phmax || nl > p.pcmax || pglen < 0 || pglen > p.pgmax)
        return fail(CLH_ERR_INVALID, "plan: pair maxima");
    if (tbeg2 < 0 || (size_t)tbeg2 + ntri2 > h.ptri_idx.size() || ebeg2 < 0 ||
        (size_t)ebeg2 + nhin2 > h.phin_idx.size() ||
        pgbeg + (size_t)((pglen + 1) & ~1) > h.pgl_idx.size())
        return fail(CLH_ERR_INVALID, "plan: pair ranges");
    const int* ph = &h.phalo[(size_t)t * p.pstride];
    auto pid = [&](int l) { return l < nown ? vbeg + l : ph[l - nown]; };
    for (int l = nown; l < nl; ++l)
        if (pid(l) != h.halo[(size_t)t * p.hstride + l - nown])
            return fail(CLH_ERR_INVALID, "plan: pair ring 1 differs from the halo");
    std::vector<int> ids;
    for (int l = nown; l < nl2; ++l) {
        const int g = pid(l);
        if (g < 0 || g >= p.N || (g >= vbeg && g < vbeg + nown))
            return fail(CLH_ERR_INVALID, "plan: pair halo id");
        if (h.perm[(size_t)g] < 0 || cloth_of(h.perm[(size_t)g]) != h.tile_cloth[(size_t)t])
            return fail(CLH_ERR_INVALID, "plan: pair halo vertex outside the tile's cloth");
        ids.push_back(g);
    }
    std::sort(ids.begin(), ids.end());
    if (std::adjacent_find(ids.begin(), ids.end()) != ids.end())
        return fail(CLH_ERR_INVALID, "plan: pair halo repeats a vertex");
    const int T = (int)(tris.size() / 3), E = (int)(quads.size() / 4);
    for (int j = 0; j < ntri2; ++j) {
        const uint2 id = h.ptri_idx[(size_t)tbeg2 + j];
        const int e = h.ptri_gid[(size_t)tbeg2 + j];
        const int l[3] = {(int)(id.x & 0xFFFF), (int)(id.x >> 16), (int)(id.y & 0xFFFF)};
        if (e < 0 || e >= T) return fail(CLH_ERR_INVALID, "plan: pair triangle id");
        bool ring1 = false;
        for (int c = 0; c < 3; ++c) {
            if (l[c] >= nl2 || h.perm[(size_t)pid(l[c])] != tris[(size_t)e * 3 + c])
                return fail(CLH_ERR_INVALID, "plan: pair triangle local index");
            if (l[c] < nown) return fail(CLH_ERR_INVALID, "plan: pair triangle touches an own vertex");
            ring1 = ring1 || l[c] < nl;
        }
        if (!ring1) return fail(CLH_ERR_INVALID, "plan: pair triangle misses ring 1");
    }
    for (int j = 0; j < nhin2; ++j) {
        const uint2 id = h.phin_idx[(size_t)ebeg2 + j];
        const int e = h.phin_gid[(size_t)ebeg2 + j];
        const int l[4] = {(int)(id.x & 0xFFFF), (int)(id.x >> 16), (int)(id.y & 0xFFFF), (int)(id.y >> 16)};
        if (e < 0 || e >= E) return fail(CLH_ERR_INVALID, "plan: pair hinge id");
        bool ring1 = false;
        for (int c = 0; c < 4; ++c) {
            if (l[c] >= nl2 || h.perm[(size_t)pid(l[c])] != quads[(size_t)e * 4 + c])
                return fail(CLH_ERR_INVALID, "plan: pair hinge local index");
            if (l[c] < nown) return fail(CLH_ERR_INVALID, "plan: pair hinge touches an own vertex");
            ring1 = ring1 || l[c] < nl;
        }
        if (!ring1) return fail(CLH_ERR_INVALID, "plan: pair hinge misses ring 1");
    }
    for (int i = 0; i < nh1; ++i) {
        const int o = h.perm[(size_t)pid(nown + i)];
        const unsigned m = h.pmeta[(size_t)t * p.hstride + i];
        const int j0 = (int)(m & 0xFFFFu), nt = (int)((m >> 16) & 0xFFu), na = (int)(m >> 24);
        if (nt > na || j0 + na > pglen) return fail(CLH_ERR_INVALID, "plan: pair gather meta");
        std::vector<int> got;
        for (int j = j0; j < j0 + na; ++j) {
            const int sl = h.pgl_idx[pgbeg + (size_t)j];
            if (j < j0 + nt) {
                const int k = sl / 3;
                if (k >= ntri + ntri2) return fail(CLH_ERR_INVALID, "plan: pair gather triangle slot");
                const int e = k < ntri ? h.tri_gid[(size_t)tbeg + k] : h.ptri_gid[(size_t)tbeg2 + k - ntri];
                got.push_back(e * 3 + sl % 3);
            } else {
                const int k = sl / 4;
                if (k >= nhin + nhin2) return fail(CLH_ERR_INVALID, "plan: pair gather hinge slot");
                const int e = k < nhin ? h.hin_gid[(size_t)ebeg + k] : h.phin_gid[(size_t)ebeg2 + k - nhin];
                got.push_back(-1 - (e * 4 + sl % 4));
            }
        }
        std::vector<int> want;
        for (int j = vt_ptr[(size_t)o]; j < vt_ptr[(size_t)o + 1]; ++j) want.push_back(vt_ids[(size_t)j]);
        for (int j = ve_ptr[(size_t)o]; j < ve_ptr[(size_t)o + 1]; ++j) want.push_back(-1 - ve_ids[(size_t)j]);
        if (got != want) return fail(CLH_ERR_INVALID, "plan: pair gather list differs from element order");
    }
    return CLH_OK;
}

int host_plan_check(const HostPlan& h, const std::vector<int>& tris, const std::vector<int>& quads,
                    const std::vector<int>& seg_in)
{
    const FusedPlan& p = h.meta;
    const int Np = p.N, N = p.num_verts, V = p.tile_verts, C = p.num_cloths;
    const std::vector<int> seg = seg_in.empty() ? std::vector<int>{0, N} : seg_in;
    if ((int)seg.size() != C + 1 || (int)h.cloth_tiles.size() != C + 1 ||
        (int)h.tile_cloth.size() != p.num_tiles)
        return fail(CLH_ERR_INVALID, "plan: cloth tables");
    if ((int)h.perm.size() != Np || (int)h.inv.size() != N) return fail(CLH_ERR_INVALID, "plan: perm size");
    if ((long long)p.num_tiles * V != Np) return fail(CLH_ERR_INVALID, "plan: tile count");
    std::vector<int> seen((size_t)N, 0);
    for (int i = 0; i < Np; ++i) {
        const int o = h.perm[(size_t)i];
        if (o < 0) continue;
        if (o >= N || h.inv[(size_t)o] != i || seen[(size_t)o]++)
            return fail(CLH_ERR_INVALID, "plan: perm is not a permutation");
    }
    for (int v = 0; v < N; ++v)
        if (!seen[(size_t)v]) return fail(CLH_ERR_INVALID, "plan: vertex missing from the tile order");
    auto cloth_of = [&](int o) {
        return (int)(std::upper_bound(seg.begin(), seg.end(), o) - seg.begin()) - 1;
    };
    if (h.cloth_tiles[0] != 0 || h.cloth_tiles[(size_t)C] != p.num_tiles)
        return fail(CLH_ERR_INVALID, "plan: cloth tile ranges");
    for (int c = 0; c < C; ++c)
        for (int t = h.cloth_tiles[(size_t)c]; t < h.cloth_tiles[(size_t)c + 1]; ++t)
            if (h.tile_cloth[(size_t)t] != c) return fail(CLH_ERR_INVALID, "plan: tile cloth");
    for (int i = 0; i < Np; ++i) {
        const int o = h.perm[(size_t)i];
        if (o >= 0 && cloth_of(o) != h.tile_cloth[(size_t)(i / V)])
            return fail(CLH_ERR_INVALID, "plan: vertex in a tile of another cloth");
        if (o < 0 && h.gl_meta[(size_t)i] != 0) return fail(CLH_ERR_INVALID, "plan: padding gathers");
    }
    if (h.halo.size() != (size_t)p.num_tiles * p.hstride) return fail(CLH_ERR_INVALID, "plan: halo size");
    for (int v : h.halo)
        if (v < 0 || v >= Np) return fail(CLH_ERR_INVALID, "plan: halo id range");
    std::vector<int> vt_ptr, vt_ids, ve_ptr, ve_ids;
    build_slots(tris, 3, N, vt_ptr, vt_ids);
    build_slots(quads, 4, N, ve_ptr, ve_ids);
    const int T = (int)(tris.size() / 3), E = (int)(quads.size() / 4);
    for (int t = 0; t < p.num_tiles; ++t) {
        const int* ti = &h.info[(size_t)t * kTileInfo];
        const int vbeg = t * V, nown = V;
        const int nhalo = ti[0], tbeg = ti[1], ntri = ti[2], ebeg = ti[3], nhin = ti[4];
        const size_t gbeg = (size_t)ti[5] * 2;
        const int glen = ti[6];
        const int nl = nown + nhalo;
        if (ti[7] != vbeg) return fail(CLH_ERR_INVALID, "plan: own range");
        if (nhalo < 0 || nhalo > p.hstride || nl > p.lmax || ntri > p.tmax || nhin > p.hmax)
            return fail(CLH_ERR_INVALID, "plan: tile maxima");
        if (tbeg < 0 || (size_t)tbeg + ntri > h.tri_idx.size() || ebeg < 0 ||
            (size_t)ebeg + nhin > h.hin_idx.size())
            return fail(CLH_ERR_INVALID, "plan: element range");
        if (glen < 0 || glen > p.gmax || gbeg + (size_t)((glen + 1) & ~1) > h.gl_idx.size())
            return fail(CLH_ERR_INVALID, "plan: gather range");
        auto tile_id = [&](int l) { return l < nown ? vbeg + l : h.halo[(size_t)t * p.hstride + l - nown]; };
        for (int l = nown; l < nl; ++l) {
            const int g = tile_id(l);
            if (g >= vbeg && g < vbeg + nown) return fail(CLH_ERR_INVALID, "plan: halo id");
            if (h.perm[(size_t)g] < 0 || cloth_of(h.perm[(size_t)g]) != h.tile_cloth[(size_t)t])
                return fail(CLH_ERR_INVALID, "plan: halo vertex outside the tile's cloth");
        }
        for (int j = 0; j < ntri; ++j) {
            const uint2 id = h.tri_idx[(size_t)tbeg + j];
            const int e = h.tri_gid[(size_t)tbeg + j];
            const int l[3] = {(int)(id.x & 0xFFFF), (int)(id.x >> 16), (int)(id.y & 0xFFFF)};
            if (e < 0 || e >= T) return fail(CLH_ERR_INVALID, "plan: triangle id");
            for (int c = 0; c < 3; ++c)
                if (l[c] >= nl || h.perm[(size_t)tile_id(l[c])] != tris[(size_t)e * 3 + c])
                    return fail(CLH_ERR_INVALID, "plan: triangle local index");
        }
        for (int j = 0; j < nhin; ++j) {
            const uint2 id = h.hin_idx[(size_t)ebeg + j];
            const int e = h.hin_gid[(size_t)ebeg + j];
            const int l[4] = {(int)(id.x & 0xFFFF), (int)(id.x >> 16), (int)(id.y & 0xFFFF), (int)(id.y >> 16)};
            if (e < 0 || e >= E) return fail(CLH_ERR_INVALID, "plan: hinge id");
            for (int c = 0; c < 4; ++c)
                if (l[c] >= nl || h.perm[(size_t)tile_id(l[c])] != quads[(size_t)e * 4 + c])
                    return fail(CLH_ERR_INVALID, "plan: hinge local index");
        }
        for (int i = 0; i < nown; ++i) {
            const int g = vbeg + i, o = h.perm[(size_t)g];
            if (o < 0) continue;  // padding (checked above)
            const unsigned m = h.gl_meta[(size_t)g];
            const int j0 = (int)(m & 0xFFFFu), nt = (int)((m >> 16) & 0xFFu), na = (int)(m >> 24);
            if (nt > na || j0 + na > glen) return fail(CLH_ERR_INVALID, "plan: gather meta");
            std::vector<int> got;
            for (int j = j0; j < j0 + na; ++j) {
                const int sl = h.gl_idx[gbeg + (size_t)j];
                if (j < j0 + nt) {
                    if (sl / 3 >= ntri) return fail(CLH_ERR_INVALID, "plan: gather triangle slot");
                    got.push_back(h.tri_gid[(size_t)tbeg + sl / 3] * 3 + sl % 3);
                } else {
                    if (sl / 4 >= nhin) return fail(CLH_ERR_INVALID, "plan: gather hinge slot");
                    got.push_back(-1 - (h.hin_gid[(size_t)ebeg + sl / 4] * 4 + sl % 4));
                }
            }
            std::vector<int> want;
            for (int j = vt_ptr[(size_t)o]; j < vt_ptr[(size_t)o + 1]; ++j) want.push_back(vt_ids[(size_t)j]);
            for (int j = ve_ptr[(size_t)o]; j < ve_ptr[(size_t)o + 1]; ++j) want.push_back(-1 - ve_ids[(size_t)j]);
            if (got != want) return fail(CLH_ERR_INVALID, "plan: gather list differs from element order");
        }
        if (p.pair) CLH_TRY(pair_tile_check(h, t, tris, quads, vt_ptr, vt_ids, ve_ptr, ve_ids, cloth_of));
    }
    // neighbour tiles (the persistent kernel's waits): ascending, unique,
    // without the tile itself, every halo owner among them, no tile beyond
    // them, and the relation symmetric, so a tile never runs more than one
    // iteration ahead of a tile that reads its vertices
    if ((int)h.nbr_ptr.size() != p.num_tiles + 1 || h.nbr_ptr[0] != 0 ||
        h.nbr_ptr.back() != (int)h.nbr.size())
        return fail(CLH_ERR_INVALID, "plan: neighbour table size");
    std::vector<std::vector<int>> adj((size_t)p.num_tiles);
    for (int t = 0; t < p.num_tiles; ++t) {
        const int b = h.nbr_ptr[(size_t)t], e = h.nbr_ptr[(size_t)t + 1];
        if (e < b || e - b > p.nbr_max) return fail(CLH_ERR_INVALID, "plan: neighbour range");
        std::vector<int> want;
        const int nh = h.info[(size_t)t * kTileInfo + 0];
        for (int i = 0; i < nh; ++i) {
            const int o = h.halo[(size_t)t * p.hstride + i] / V;
            if (o != t) want.push_back(o);
        }
        std::sort(want.begin(), want.end());
        want.erase(std::unique(want.begin(), want.end()), want.end());
        adj[(size_t)t].assign(h.nbr.begin() + b, h.nbr.begin() + e);
        if (adj[(size_t)t] != want) return fail(CLH_ERR_INVALID, "plan: neighbour tiles differ from halo owners");
    }
    for (int t = 0; t < p.num_tiles; ++t)
        for (int o : adj[(size_t)t])
            if (!std::binary_search(adj[(size_t)o].begin(), adj[(size_t)o].end(), t))
                return fail(CLH_ERR_INVALID, "plan: neighbour relation not symmetric");
    return CLH_OK;
}

int fused_plan_build(FusedPlan& p, int N, const std::vector<int>& tris,
                     const std::vector<int>& quads, const std::vector<float>& pos, int tile_verts,
                     const std::vector<int>& seg, bool pair)
{
    fused_plan_release(p);
    HostPlan h;
    CLH_TRY(host_plan_build(h, N, tris, quads, pos, tile_verts, seg, pair));
    const FusedPlan& m = h.meta;
    p.N = m.N; p.num_verts = m.num_verts; p.num_cloths = m.num_cloths; p.num_tiles = m.num_tiles; p.tile_verts = m.tile_verts; p.hstride = m.hstride;
    p.lmax = m.lmax; p.tmax = m.tmax; p.hmax = m.hmax; p.gmax = m.gmax; p.vmax = m.vmax;
    p.tri_factor = m.tri_factor; p.bend_factor = m.bend_factor; p.xcd_remap = m.xcd_remap;
    p.bytes_per_iter = m.bytes_per_iter;
    CLH_TRY(p.perm.upload(h.perm));
    CLH_TRY(p.inv_perm.upload(h.inv));
    CLH_TRY(p.tile_info.upload(h.info));
    CLH_TRY(p.halo.upload(h.halo));
    CLH_TRY(p.tri_gid.upload(h.tri_gid));
    CLH_TRY(p.tri_idx.upload(h.tri_idx));
    CLH_TRY(p.hin_gid.upload(h.hin_gid));
    CLH_TRY(p.hin_idx.upload(h.hin_idx));
    CLH_TRY(p.tri_c0.resize(h.tri_gid.size()));
    CLH_TRY(p.tri_c1.resize(h.tri_gid.size()));
    CLH_TRY(p.hin_w.resize(h.hin_gid.size()));
    CLH_TRY(p.hin_c.resize(h.hin_gid.size()));
    std::vector<unsigned> words(h.gl_idx.size() / 2);
    std::memcpy(words.data(), h.gl_idx.data(), words.size() * sizeof(unsigned));
    CLH_TRY(p.gl_words.upload(words));
    CLH_TRY(p.gl_meta.upload(h.gl_meta));
    CLH_TRY(p.tile_cloth.upload(h.tile_cloth));
    CLH_TRY(p.cloth_tiles.upload(h.cloth_tiles));
    p.nbr_max = m.nbr_max;
    CLH_TRY(p.nbr_ptr.upload(h.nbr_ptr));
    {
        std::vector<int> nb = h.nbr;
        if (nb.empty()) nb.push_back(0);
        CLH_TRY(p.nbr.upload(nb));
    }
    CLH_TRY(p.prog.resize((size_t)std::max(p.num_tiles, 1) * kBarLine));
    p.pair = m.pair;
    if (p.pair) {
        p.pstride = m.pstride; p.plmax = m.plmax; p.ptmax = m.ptmax; p.phmax = m.phmax;
        p.pgmax = m.pgmax; p.pcmax = m.pcmax; p.pbytes_per_launch = m.pbytes_per_launch;
        CLH_TRY(p.pinfo.upload(h.pinfo));
        CLH_TRY(p.phalo.upload(h.phalo));
        CLH_TRY(p.pmeta.upload(h.pmeta));
        CLH_TRY(p.ptri_gid.upload(h.ptri_gid));
        CLH_TRY(p.ptri_idx.upload(h.ptri_idx));
        CLH_TRY(p.phin_gid.upload(h.phin_gid));
        CLH_TRY(p.phin_idx.upload(h.phin_idx));
        CLH_TRY(p.ptri_c0.resize(h.ptri_gid.size()));
        CLH_TRY(p.ptri_c1.resize(h.ptri_gid.size()));
        CLH_TRY(p.phin_w.resize(h.phin_gid.size()));
        CLH_TRY(p.phin_c.resize(h.phin_gid.size()));
        std::vector<unsigned> pw(h.pgl_idx.size() / 2);
        std::memcpy(pw.data(), h.pgl_idx.data(), pw.size() * sizeof(unsigned));
        CLH_TRY(p.pgl_words.upload(pw));
    }
    std::memset(p.packed_src, 0, sizeof(p.packed_src));
    p.packed_h2 = -1.0f;
    return CLH_OK;
}

int fused_pack_constants(FusedPlan& p, const clh_step_args& a, float h2, hipStream_t stream)
{
    const void* src[8] = {a.d_Dm_inv, a.d_rest_area, a.d_stretch_k, a.d_bend_w,
                          a.d_bend_n,  a.d_bend_k,    a.d_tris,      a.d_bend_quads};
    if (p.packed_h2 == h2 && std::memcmp(src, p.packed_src, sizeof(src)) == 0) return CLH_OK;
    const int nt = (int)p.tri_gid.count, nh = (int)p.hin_gid.count;
    if (a.d_Dm_inv && a.d_rest_area)
        hipLaunchKernelGGL(k_pack_tris, dim3(blocks_for(nt)), dim3(256), 0, stream, p.tri_gid.ptr,
                           a.d_Dm_inv, a.d_rest_area, a.d_stretch_k, p.tri_c0.ptr, p.tri_c1.ptr, nt,
                           h2);
    if (a.d_bend_w && a.d_bend_n && a.d_bend_k)
        hipLaunchKernelGGL(k_pack_hinges, dim3(blocks_for(nh)), dim3(256), 0, stream,
                           p.hin_gid.ptr, a.d_bend_w, a.d_bend_n, a.d_bend_k, p.hin_w.ptr,
                           p.hin_c.ptr, nh, h2);
    if (p.pair) {
        const int pt = (int)p.ptri_gid.count, ph = (int)p.phin_gid.count;
        if (a.d_Dm_inv && a.d_rest_area)
            hipLaunchKernelGGL(k_pack_tris, dim3(blocks_for(pt)), dim3(256), 0, stream,
                               p.ptri_gid.ptr, a.d_Dm_inv, a.d_rest_area, a.d_stretch_k,
                               p.ptri_c0.ptr, p.ptri_c1.ptr, pt, h2);
        if (a.d_bend_w && a.d_bend_n && a.d_bend_k)
            hipLaunchKernelGGL(k_pack_hinges, dim3(blocks_for(ph)), dim3(256), 0, stream,
                               p.phin_gid.ptr, a.d_bend_w, a.d_bend_n, a.d_bend_k, p.phin_w.ptr,
                               p.phin_c.ptr, ph, h2);
    }
    CLH_HIP(hipGetLastError());
    std::memcpy(p.packed_src, src, sizeof(src));
    p.packed_h2 = h2;
    return CLH_OK;
}

int fused_max_blocks(const FusedPlan& p) { return p.num_tiles; }

size_t fused_persist_lds_bytes(const FusedPlan& p) { return persist_layout(p).total; }

int gather_lanes(const FusedPlan& p);

// The persistent kernel of the plan's workgroup size and gather lanes.
const void* persist_kernel(const FusedPlan& p)
{
    if (p.threads == 1024) return reinterpret_cast<const void*>(&k_fused_persist<1024, 4>);
    switch (gather_lanes(p)) {
    case 4: return reinterpret_cast<const void*>(&k_fused_persist<256, 4>);
    case 2: return reinterpret_cast<const void*>(&k_fused_persist<256, 2>);
    default: return reinterpret_cast<const void*>(&k_fused_persist<256, 1>);
    }
}

bool fused_persist_fits(const FusedPlan& p)
{
    if ((p.threads != 256 && p.threads != 1024) || p.num_tiles <= 0 || p.nbr_max > p.threads)
        return false;
    int dev = 0, cus = 0, coop = 0, per = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, dev) != hipSuccess) {
        (void)hipGetLastError();
        return false;
    }
    if (!coop) return false;
    const hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(
        &per, reinterpret_cast<const void*>(persist_kernel(p)), p.threads, fused_persist_lds_bytes(p));
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return false;
    }
    return (long long)per * cus >= p.num_tiles;
}

int fused_launch_persist(const FusedPlan& p, const FusedIterArgs& a, double* const it[kIterBufs], int k0,
                         int k1, int K, int check_every, FrameCtl* ctl, double* partials,
                         double* hist, unsigned* bar, hipStream_t stream)
{
    PersistArgs P;
    FusedKernelArgs& A = P.A;
    fill_kernel_args(p, a, A);
    const PersistLayout L = persist_layout(p);
    A.off_g = (int)L.it.off_g;
    A.off_r = (int)L.it.off_r;
    for (int i = 0; i < kIterBufs; ++i) P.it[i] = it[i];
    P.k0 = k0;
    P.k1 = k1;
    P.K = K;
    P.check_every = check_every;
    P.ctl = ctl;
    P.partials = partials;
    P.hist = hist;
    P.bar = bar;
    P.prog = p.prog.ptr;
    P.nbr_ptr = p.nbr_ptr.ptr;
    P.nbr = p.nbr.ptr;
    P.off_rt = (int)L.off_rt;
    P.off_rh = (int)L.off_rh;
    P.off_hb = (int)L.off_hb;
    if (k0 == 0) {
        // a frame's progress counts start at 0 (a fill kernel: no memset
        // node in captured work)
        const int n = p.num_tiles * kBarLine;
        hipLaunchKernelGGL(k_fill_u32, dim3((n + 255) / 256), dim3(256), 0, stream, p.prog.ptr, n, 0u);
    }
    void* args[] = {&P};
    const void* fn = persist_kernel(p);
    return hip_status(hipLaunchCooperativeKernel(fn, dim3(p.num_tiles), dim3(p.threads), args,
                                                 L.total, stream),
                      "persistent iteration launch");
}

size_t fused_pair_lds_bytes(const FusedPlan& p, bool check) { return pair_lds_layout(p, check).total; }

// the two-iteration kernel is built for 256-, 512- and 1024-thread
// workgroups, which must hold a tile's own and ring-1 vertices
bool fused_pair_usable(const FusedPlan& p)
{
    return p.pair && p.threads <= 1024 && p.pcmax <= p.threads;
}

int fused_launch_pair(const FusedPlan& p, const FusedIterArgs& a, void* out2, float omega1,
                      int blend1, hipStream_t stream)
{
    if (!fused_pair_usable(p))
        return fail(CLH_ERR_UNSUPPORTED, "two-iteration launch: no pair plan for this workgroup size");
    PairKernelArgs P;
    fill_kernel_args(p, a, P.A);
    const bool check = a.partials != nullptr;
    const PairLdsLayout L = pair_lds_layout(p, check);
    P.A.off_g = (int)L.off_g;
    P.A.off_r = (int)L.off_r;
    P.off_g2 = (int)L.off_g2;
    P.out2 = out2;
    P.omega1 = omega1;
    P.blend1 = blend1;
    P.pinfo = p.pinfo.ptr;
    P.phalo = p.phalo.ptr;
    P.pmeta = p.pmeta.ptr;
    P.ptri_idx = p.ptri_idx.ptr;
    P.ptri_c0 = p.ptri_c0.ptr;
    P.ptri_c1 = p.ptri_c1.ptr;
    P.phin_idx = p.phin_idx.ptr;
    P.phin_w = p.phin_w.ptr;
    P.phin_c = p.phin_c.ptr;
    P.pgl_words = p.pgl_words.ptr;
    P.pstride = p.pstride;
    P.plmax = p.plmax;
    P.ptmax = p.ptmax;
    P.phmax = p.phmax;
    const dim3 grid(p.num_tiles);
    if (p.threads == 1024) {
        if (check) hipLaunchKernelGGL((k_fused_pair<true, 1024>), grid, dim3(1024), L.total, stream, P);
        else       hipLaunchKernelGGL((k_fused_pair<false, 1024>), grid, dim3(1024), L.total, stream, P);
    } else if (p.threads == 512) {
        if (check) hipLaunchKernelGGL((k_fused_pair<true, 512>), grid, dim3(512), L.total, stream, P);
        else       hipLaunchKernelGGL((k_fused_pair<false, 512>), grid, dim3(512), L.total, stream, P);
    } else {
        if (check) hipLaunchKernelGGL((k_fused_pair<true, 256>), grid, dim3(256), L.total, stream, P);
        else       hipLaunchKernelGGL((k_fused_pair<false, 256>), grid, dim3(256), L.total, stream, P);
    }
    return hip_status(hipGetLastError(), "two-iteration launch");
}

// Lanes per own vertex in the gathers of k_fused_iter: as many as the
// workgroup has threads per tile vertex, at most 4 (a vertex's four partial
// sums).
int gather_lanes(const FusedPlan& p)
{
    const int r = p.threads / p.tile_verts;
    return r >= 4 ? 4 : (r >= 2 ? 2 : 1);
}

template <int kMode, bool kCheck, int kT, bool kSplit>
void launch_iter(int G, dim3 grid, size_t lds, hipStream_t stream, const FusedKernelArgs& A)
{
    if (G == 4)      hipLaunchKernelGGL((k_fused_iter<kMode, kCheck, kT, kSplit, 4>), grid, dim3(kT), lds, stream, A);
    else if (G == 2) hipLaunchKernelGGL((k_fused_iter<kMode, kCheck, kT, kSplit, 2>), grid, dim3(kT), lds, stream, A);
    else             hipLaunchKernelGGL((k_fused_iter<kMode, kCheck, kT, kSplit, 1>), grid, dim3(kT), lds, stream, A);
}

template <int kMode, bool kCheck, int kT, bool kSplit>
void push_iter_kernels(std::vector<const void*>& v)
{
    v.push_back(reinterpret_cast<const void*>(&k_fused_iter<kMode, kCheck, kT, kSplit, 1>));
    v.push_back(reinterpret_cast<const void*>(&k_fused_iter<kMode, kCheck, kT, kSplit, 2>));
    v.push_back(reinterpret_cast<const void*>(&k_fused_iter<kMode, kCheck, kT, kSplit, 4>));
}

int fused_launch_iteration(const FusedPlan& p, const FusedIterArgs& a, int mode,
                           hipStream_t stream)
{
    FusedKernelArgs A;
    fill_kernel_args(p, a, A);
    const bool check = a.partials != nullptr;
    const bool split = use_split(p, mode, check);
    const LdsLayout L = lds_layout(p, mode, check, split);
    A.off_g = (int)L.off_g;
    A.off_r = (int)L.off_r;
    const size_t lds = L.total;
    const dim3 grid(p.num_tiles);
    const int G = gather_lanes(p);
    if (p.threads == 1024) {
        if (mode == CLH_MODE_PD && split) {
            if (check) launch_iter<CLH_MODE_PD, true, 1024, true>(G, grid, lds, stream, A);
            else       launch_iter<CLH_MODE_PD, false, 1024, true>(G, grid, lds, stream, A);
        } else if (mode == CLH_MODE_PD) {
            if (check) launch_iter<CLH_MODE_PD, true, 1024, false>(G, grid, lds, stream, A);
            else       launch_iter<CLH_MODE_PD, false, 1024, false>(G, grid, lds, stream, A);
        } else {
            launch_iter<CLH_MODE_REF_CUDA, false, 1024, false>(G, grid, lds, stream, A);
        }
    } else if (p.threads == 512) {
        if (mode == CLH_MODE_PD && split) {
            if (check) launch_iter<CLH_MODE_PD, true, 512, true>(G, grid, lds, stream, A);
            else       launch_iter<CLH_MODE_PD, false, 512, true>(G, grid, lds, stream, A);
        } else if (mode == CLH_MODE_PD) {
            if (check) launch_iter<CLH_MODE_PD, true, 512, false>(G, grid, lds, stream, A);
            else       launch_iter<CLH_MODE_PD, false, 512, false>(G, grid, lds, stream, A);
        } else {
            launch_iter<CLH_MODE_REF_CUDA, false, 512, false>(G, grid, lds, stream, A);
        }
    } else {
        if (mode == CLH_MODE_PD && split) {
            if (check) launch_iter<CLH_MODE_PD, true, 256, true>(G, grid, lds, stream, A);
            else       launch_iter<CLH_MODE_PD, false, 256, true>(G, grid, lds, stream, A);
        } else if (mode == CLH_MODE_PD) {
            if (check) launch_iter<CLH_MODE_PD, true, 256, false>(G, grid, lds, stream, A);
            else       launch_iter<CLH_MODE_PD, false, 256, false>(G, grid, lds, stream, A);
        } else {
            launch_iter<CLH_MODE_REF_CUDA, false, 256, false>(G, grid, lds, stream, A);
        }
    }
    return hip_status(hipGetLastError(), "fused iteration launch");
}

size_t fused_lds_bytes(const FusedPlan& p, int mode, bool check) { return lds_bytes(p, mode, check); }

int fused_enable_lds(int& bytes)
{
    // above 64 KB the kernels must opt in, each for the device's limit less
    // its own static LDS (the block-sum arrays); `bytes` becomes the dynamic
    // LDS every kernel was granted. When the runtime refuses, stay at 64 KB.
    if (bytes <= 64 * 1024) return CLH_OK;
    std::vector<const void*> fns;
    push_iter_kernels<CLH_MODE_PD, true, 256, false>(fns);
    push_iter_kernels<CLH_MODE_PD, false, 256, false>(fns);
    push_iter_kernels<CLH_MODE_REF_CUDA, false, 256, false>(fns);
    push_iter_kernels<CLH_MODE_PD, true, 512, false>(fns);
    push_iter_kernels<CLH_MODE_PD, false, 512, false>(fns);
    push_iter_kernels<CLH_MODE_REF_CUDA, false, 512, false>(fns);
    push_iter_kernels<CLH_MODE_PD, true, 256, true>(fns);
    push_iter_kernels<CLH_MODE_PD, false, 256, true>(fns);
    push_iter_kernels<CLH_MODE_PD, true, 512, true>(fns);
    push_iter_kernels<CLH_MODE_PD, false, 512, true>(fns);
    push_iter_kernels<CLH_MODE_PD, true, 1024, false>(fns);
    push_iter_kernels<CLH_MODE_PD, false, 1024, false>(fns);
    push_iter_kernels<CLH_MODE_REF_CUDA, false, 1024, false>(fns);
    push_iter_kernels<CLH_MODE_PD, true, 1024, true>(fns);
    push_iter_kernels<CLH_MODE_PD, false, 1024, true>(fns);
    fns.push_back(reinterpret_cast<const void*>(&k_fused_persist<256, 1>));
    fns.push_back(reinterpret_cast<const void*>(&k_fused_persist<256, 2>));
    fns.push_back(reinterpret_cast<const void*>(&k_fused_persist<256, 4>));
    fns.push_back(reinterpret_cast<const void*>(&k_fused_persist<1024, 4>));
    fns.push_back(reinterpret_cast<const void*>(&k_fused_pair<true, 256>));
    fns.push_back(reinterpret_cast<const void*>(&k_fused_pair<false, 256>));
    fns.push_back(reinterpret_cast<const void*>(&k_fused_pair<true, 512>));
    fns.push_back(reinterpret_cast<const void*>(&k_fused_pair<false, 512>));
    fns.push_back(reinterpret_cast<const void*>(&k_fused_pair<true, 1024>));
    fns.push_back(reinterpret_cast<const void*>(&k_fused_pair<false, 1024>));
    int granted = bytes;
    for (const void* f : fns) {
        hipFuncAttributes fa;
        if (hipFuncGetAttributes(&fa, f) != hipSuccess) {
            (void)hipGetLastError();
            bytes = 64 * 1024;
            return CLH_OK;
        }
        const int dyn = bytes - (int)fa.sharedSizeBytes;
        if (dyn <= 64 * 1024 ||
            hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, dyn) != hipSuccess) {
            (void)hipGetLastError();
            bytes = 64 * 1024;
            return CLH_OK;
        }
        granted = std::min(granted, dyn);
    }
    bytes = granted;
    return CLH_OK;
}

int fused_prologue(const FusedPlan& p, const clh_step_args& a, float* xprev, float* y, void* x0,
                   int4* vinfo, float dt, float h2, const float acc[3], const float h2a[3],
                   int disp, hipStream_t stream)
{
    hipLaunchKernelGGL(k_fused_prologue, dim3(blocks_for(p.N)), dim3(256), 0, stream, p.perm.ptr,
                       a.d_pos, a.d_vel, a.d_init_guess, a.d_mass, a.d_jacobi_diag, p.gl_meta.ptr,
                       p.tile_cloth.ptr, a.d_cloth_accel, xprev, y, x0, vinfo, p.N, p.tile_verts,
                       dt, h2, acc[0], acc[1], acc[2], h2a[0], h2a[1], h2a[2], disp);
    if (a.num_pinned > 0)
        hipLaunchKernelGGL(k_fused_pins, dim3(blocks_for(a.num_pinned)), dim3(256), 0, stream,
                           p.inv_perm.ptr, a.d_pin_idx, a.d_pin_target, xprev, vinfo, x0,
                           a.num_pinned, disp);
    return hip_status(hipGetLastError(), "fused prologue launch");
}

int fused_epilogue(const FusedPlan& p, const clh_step_args& a, const void* xf, const float* xprev,
                   const int4* vinfo, float scale, int disp, hipStream_t stream)
{
    hipLaunchKernelGGL(k_fused_epilogue, dim3(blocks_for(p.N)), dim3(256), 0, stream, p.perm.ptr,
                       xf, xprev, vinfo, a.d_pin_target, a.d_pos, a.d_vel, p.N, scale, disp);
    return hip_status(hipGetLastError(), "fused epilogue launch");
}

int fused_trace(const FusedPlan& p, const void* it, const float* xprev, float* out, int disp,
                hipStream_t stream)
{
    hipLaunchKernelGGL(k_fused_trace, dim3(blocks_for(p.N)), dim3(256), 0, stream, p.perm.ptr, it,
                       xprev, out, p.N, disp);
    return hip_status(hipGetLastError(), "fused trace launch");
}

}  // namespace clh

extern "C" int clh_fused_plan_check(int N, int T, int E, const int* tris, const int* quads,
                                    const float* pos, int tile_verts, int mode, int num_cloths,
                                    const int* cloth_vert_offsets, clh_plan_info* out)
{
    using namespace clh;
    if (N <= 0 || T < 0 || E < 0 || !pos || (T > 0 && !tris) || (E > 0 && !quads) ||
        num_cloths < 0 || (num_cloths > 0 && !cloth_vert_offsets))
        return fail(CLH_ERR_INVALID, "clh_fused_plan_check: bad arguments");
    std::vector<int> t(tris, tris + (size_t)T * 3), q;
    if (E > 0) q.assign(quads, quads + (size_t)E * 4);
    for (int v : t) if (v < 0 || v >= N) return fail(CLH_ERR_INVALID, "triangle vertex index out of range");
    for (int v : q) if (v < 0 || v >= N) return fail(CLH_ERR_INVALID, "hinge vertex index out of range");
    std::vector<int> seg;
    if (num_cloths > 0) seg.assign(cloth_vert_offsets, cloth_vert_offsets + num_cloths + 1);
    std::vector<float> x(pos, pos + (size_t)N * 3);
    HostPlan h;
    CLH_TRY(host_plan_build(h, N, t, q, x, tile_verts, seg, mode == CLH_MODE_PD));
    CLH_TRY(host_plan_check(h, t, q, seg));
    if (out) {
        std::memset(out, 0, sizeof(*out));
        out->variant = CLH_KERNEL_FUSED;
        fused_plan_info(h.meta, out);
        out->lds_bytes = (double)fused_lds_bytes(h.meta, mode, true);
    }
    return CLH_OK;
}

namespace clh {

void fused_plan_info(const FusedPlan& p, clh_plan_info* out)
{
    out->num_cloths = p.num_cloths;
    out->padded_verts = p.N;
    out->tri_max = p.tmax;
    out->hinge_max = p.hmax;
    out->halo_max = p.hstride;
    out->num_tiles = p.num_tiles;
    out->tile_verts = p.tile_verts;
    out->halo_tri_factor = p.tri_factor;
    out->halo_bend_factor = p.bend_factor;
    out->bytes_per_iter = p.bytes_per_iter;
    out->block_threads = p.threads;
    out->pair = p.pair;
    out->iters_per_launch = 1;
    if (p.pair) {
        out->pair_halo_max = p.pstride;
        out->pair_tri_max = p.ptmax;
        out->pair_hinge_max = p.phmax;
        out->pair_bytes_per_launch = p.pbytes_per_launch;
        out->pair_compute_max = p.pcmax;
        out->pair_lds_bytes = (double)pair_lds_layout(p, true).total;
    }
    out->split_lds_bytes = (double)lds_layout(p, CLH_MODE_PD, true, true).total;
    out->element_regions = use_split(p, CLH_MODE_PD, false) ? 2 : 1;
    out->gather_lanes = gather_lanes(p);
}

void fused_plan_release(FusedPlan& p)
{
    p.perm.release();
    p.inv_perm.release();
    p.tile_info.release();
    p.halo.release();
    p.tri_gid.release();
    p.hin_gid.release();
    p.tri_idx.release();
    p.hin_idx.release();
    p.tri_c0.release();
    p.hin_w.release();
    p.tri_c1.release();
    p.hin_c.release();
    p.gl_words.release();
    p.gl_meta.release();
    p.tile_cloth.release();
    p.cloth_tiles.release();
    p.nbr_ptr.release();
    p.nbr.release();
    p.prog.release();
    p.nbr_max = 0;
    p.pinfo.release();
    p.phalo.release();
    p.pmeta.release();
    p.ptri_gid.release();
    p.phin_gid.release();
    p.ptri_idx.release();
    p.phin_idx.release();
    p.ptri_c0.release();
    p.phin_w.release();
    p.ptri_c1.release();
    p.phin_c.release();
    p.pgl_words.release();
    p.pair = 0;
    p.num_tiles = 0;
}

}  // namespace clh
