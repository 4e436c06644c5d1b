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

// Rotating iterate buffers of a frame: iteration k reads x_k from it[k % 4]
// and x_{k-1} from it[(k + 3) % 4] and writes x_{k+1} to it[(k + 1) % 4]; a
// two-iteration launch also writes x_{k+2} to it[(k + 2) % 4], which no tile
// reads during that launch.
constexpr int kIterBufs = 4;

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
    ClothCtl* cctl;        // batches run to a tolerance: per-cloth control (with ctl), or null
    float relax;           // PD: scale gamma of the Jacobi step (0 means 1)
    const float* cloth_relax;  // batches: each cloth's gamma (overrides relax), or null
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
    // launch choices (not part of the plan's data): separate triangle and
    // hinge LDS regions in PD launches whose layout fits lds_limit
    int split = 0;
    int lds_limit = 65536;

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
    // persistent launches: the tiles owning each tile's halo vertices, and
    // one progress word per tile (its own 128-byte line)
    DevBuf<int>   nbr_ptr, nbr;               // [num_tiles + 1], neighbour tile ids
    DevBuf<unsigned> prog;                    // [num_tiles * kBarLine]
    int nbr_max = 0;                          // most neighbours of a tile

    // Two-iteration launches (PD): each tile also updates its ring-1 halo
    // (the vertices sharing an element with an own vertex) in the first
    // iteration, from the positions of ring 1 and ring 2, so the second
    // iteration of its own vertices needs no other tile's result.
    int pair = 0;          // 1 when the extension below was built
    int pstride = 0;       // ring-1 + ring-2 ids per tile (padded)
    int plmax = 0;         // local vertex table (own + ring 1 + ring 2)
    int ptmax = 0, phmax = 0;  // triangles / hinges touching own or ring-1 vertices
    int pgmax = 0;         // ring-1 gather entries
    int pcmax = 0;         // vertices updated in the first iteration (own + ring 1)
    double pbytes_per_launch = 0.0;
    DevBuf<int>   pinfo;                      // [num_tiles * kPairInfo]
    DevBuf<int>   phalo;                      // [num_tiles * pstride] ring 1, then ring 2
    DevBuf<unsigned> pmeta;                   // [num_tiles * hstride] ring-1 gather meta
    DevBuf<int>   ptri_gid, phin_gid;         // ring-1-only elements (original ids)
    DevBuf<uint2> ptri_idx, phin_idx;
    DevBuf<float4> ptri_c0, phin_w;
    DevBuf<float> ptri_c1;
    DevBuf<float2> phin_c;
    DevBuf<unsigned> pgl_words;               // ring-1 gather entries, two u16 per word

    // packed-constant provenance (repack when any pointer changes)
    const void* packed_src[8] = {};
    float packed_h2 = -1.0f;
};

// seg: cloth vertex offsets [C+1] (cloth c owns original vertices
// [seg[c], seg[c+1])), or empty for one cloth. Every element must lie in one
// cloth; each cloth's vertices are Morton-ordered on their own and padded to
// whole tiles with inert vertices (no elements, zero mass and diagonal).
// pair: also build the two-iteration extension (kept off, p.pair = 0, when a
// tile exceeds its limits).
int  fused_plan_build(FusedPlan& p, int N, const std::vector<int>& tris,
                      const std::vector<int>& quads, const std::vector<float>& pos,
                      int tile_verts, const std::vector<int>& seg, bool pair);
int  fused_pack_constants(FusedPlan& p, const clh_step_args& a, float h2, hipStream_t stream);
int  fused_max_blocks(const FusedPlan& p);
size_t fused_lds_bytes(const FusedPlan& p, int mode, bool check);
// Lets the fused kernels use up to `bytes` of dynamic LDS per workgroup;
// lowers `bytes` to 64 KB when the runtime does not allow more.
int  fused_enable_lds(int& bytes);
int  fused_launch_iteration(const FusedPlan& p, const FusedIterArgs& a, int mode,
                            hipStream_t stream);
// Two PD iterations k, k+1 in one launch: a.x = x_k, a.xkm1 = x_{k-1},
// a.out receives x_{k+1} and out2 x_{k+2}; a.omega / a.blend weight
// iteration k and omega1 / blend1 iteration k+1 (or both come from a.ctl:
// slot a.par and the state after it; block 0 writes the state of k+2 to the
// other slot). A residual check (a.partials) measures iteration k+1.
// Needs the pair extension with every tile's own + ring-1 count within the
// workgroup (fused_pair_usable).
bool fused_pair_usable(const FusedPlan& p);
size_t fused_pair_lds_bytes(const FusedPlan& p, bool check);
int  fused_launch_pair(const FusedPlan& p, const FusedIterArgs& a, void* out2, float omega1,
                       int blend1, hipStream_t stream);

// Grid barrier words of the persistent kernel: one arrival word per group of
// workgroups (blockIdx % kBarGroups), a global arrival word and the
// generation, each on a 128-byte line of its own.
constexpr int kBarGroups = 8;
constexpr int kBarLine = 32;                      // unsigned per 128-byte line
constexpr int kBarAll = kBarGroups;               // line of the global arrival word
constexpr int kBarGen = kBarGroups + 1;           // line of the generation word
constexpr int kBarWords = (kBarGroups + 2) * kBarLine;

// Persistent PD iterations [k0, k1) in one cooperative launch (device control
// through ctl; iterates in it[k % kIterBufs]; grid barrier words bar[kBarWords],
// zero at rest). fused_persist_fits tells whether every tile can be resident at once.
size_t fused_persist_lds_bytes(const FusedPlan& p);
bool fused_persist_fits(const FusedPlan& p);
int  fused_launch_persist(const FusedPlan& p, const FusedIterArgs& a, double* const it[kIterBufs], int k0,
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
