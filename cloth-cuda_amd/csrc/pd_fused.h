#pragma once
// Fused LDS-tiled iteration kernel: plan and launch interface.
//
// Vertices are renumbered into tiles (Morton order of the positions at plan
// time, cut into runs of tile_verts); every per-vertex array of the
// iteration lives in that tile order. A tile owns its vertices and carries
// a private copy of every triangle and hinge touching them, with local
// (16-bit) vertex indices into its LDS position table (own vertices first,
// then the halo). One workgroup per tile per iteration:
//   1. stage own + halo positions in LDS,
//   2. triangles -> corner forces in LDS; each own vertex sums its corners,
//   3. hinges -> bend vectors in the same LDS region; each own vertex sums
//      its hinge slots (ascending element id after the triangles, the order
//      of the unfused pipeline), then Jacobi / Chebyshev / pins and FP64
//      residual partials; x_{k+1} goes to HBM.
#include "clh_internal.h"

#include <vector>

namespace clh {

struct FusedIterArgs {
    const void*  x;        // iterate k (tile order): float (reference) or double (PD)
    const void*  xkm1;     // iterate k-1
    void*        out;      // iterate k+1
    const float* base;     // x_n (tile order; PD iterates are displacements)
    const float* y;        // y (reference mode) or y - x_n (PD), tile order
    const int4*  vinfo;    // tile order: (mass, diag as float bits, pin slot or -1, gather meta)
    const float* pin_target;
    double*      partials; // per-tile |r|^2, |b|^2 when non-null
    float omega;
    int   blend;
    int   use_stretch;
    int   use_bend;
    FrameCtl* ctl;         // when non-null, omega and blend come from ctl->slot[par]
    int   par;
};

struct FusedPlan {
    int N = 0;             // tile-order length: every cloth's run padded to whole tiles
    int num_verts = 0;     // the caller's vertex count (original order)
    int num_cloths = 1;
    int num_tiles = 0;
    int tile_verts = 0;
    int threads = 256;                                     // workgroup size (256 or 512)
    int hstride = 0;                                       // halo ids per tile (padded)
    int lmax = 0, tmax = 0, hmax = 0, gmax = 0, vmax = 0;  // per-tile maxima (LDS sizing)
    double tri_factor = 0.0, bend_factor = 0.0, bytes_per_iter = 0.0;
    int xcd_remap = 1;

    DevBuf<int>   perm, inv_perm;             // tile order <-> original
    DevBuf<int>   tile_info;                  // [num_tiles * kTileInfo]
    DevBuf<int>   halo;                       // [num_tiles * hstride] tile-order halo ids
    DevBuf<int>   tri_gid, hin_gid;           // original element ids
    DevBuf<uint2> tri_idx, hin_idx;           // packed u16 local vertex ids
    DevBuf<float4> tri_c0, hin_w;             // Dm_inv | bend weights
    DevBuf<float> tri_c1;                     // h^2 k A
    DevBuf<float2> hin_c;                     // (n_rest, h^2 k)
    DevBuf<unsigned> gl_words;                // gather entries, two u16 per word
    DevBuf<unsigned> gl_meta;                 // [N] start in tile | n_tri << 16 | n_all << 24
    DevBuf<int>   tile_cloth;                 // [num_tiles] cloth owning each tile
    DevBuf<int>   cloth_tiles;                // [num_cloths + 1] first tile of each cloth

    // packed-constant provenance (repack when any pointer changes)
    const void* packed_src[8] = {};
    float packed_h2 = -1.0f;
};

// seg: cloth vertex offsets [C+1] (cloth c owns original vertices
// [seg[c], seg[c+1])), or empty for one cloth. Every element must lie in one
// cloth; each cloth's vertices are Morton-ordered on their own and padded to
// whole tiles with inert vertices (no elements, zero mass and diagonal).
int  fused_plan_build(FusedPlan& p, int N, const std::vector<int>& tris,
                      const std::vector<int>& quads, const std::vector<float>& pos,
                      int tile_verts, const std::vector<int>& seg);
int  fused_pack_constants(FusedPlan& p, const clh_step_args& a, float h2, hipStream_t stream);
int  fused_max_blocks(const FusedPlan& p);
size_t fused_lds_bytes(const FusedPlan& p, int mode, bool check);
// Lets the fused kernels use up to `bytes` of dynamic LDS per workgroup;
// lowers `bytes` to 64 KB when the runtime does not allow more.
int  fused_enable_lds(int& bytes);
int  fused_launch_iteration(const FusedPlan& p, const FusedIterArgs& a, int mode,
                            hipStream_t stream);

// Persistent PD iterations [k0, k1) in one cooperative launch (device control
// through ctl; iterates in it[k % 3]; grid barrier words bar[2], zero at
// rest). fused_persist_fits tells whether every tile can be resident at once.
size_t fused_persist_lds_bytes(const FusedPlan& p);
bool fused_persist_fits(const FusedPlan& p);
int  fused_launch_persist(const FusedPlan& p, const FusedIterArgs& a, double* const it[3], int k0,
                          int k1, int K, int check_every, FrameCtl* ctl, double* partials,
                          double* hist, unsigned* bar, hipStream_t stream);
void fused_plan_info(const FusedPlan& p, clh_plan_info* out);
void fused_plan_release(FusedPlan& p);

// Frame prologue / epilogue in tile order (gather from and scatter to the
// caller's original-order arrays).
// h2a = h^2 a for the frame's uniform acceleration a; with a.d_cloth_accel
// the prediction uses h^2 (a + a_c) of each vertex's cloth c instead.
int fused_prologue(const FusedPlan& p, const clh_step_args& a, float* xprev, float* y, void* x0,
                   int4* vinfo, float dt, float h2, const float acc[3], const float h2a[3],
                   int disp, hipStream_t stream);
int fused_epilogue(const FusedPlan& p, const clh_step_args& a, const void* xf, const float* xprev,
                   const int4* vinfo, float scale, int disp, hipStream_t stream);
int fused_trace(const FusedPlan& p, const void* it, const float* xprev, float* out, int disp,
                hipStream_t stream);

}  // namespace clh
