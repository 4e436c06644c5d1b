#pragma once
// Fused LDS-tiled iteration kernel: plan and launch interface.
#include "clh_internal.h"

#include <vector>

namespace clh {

struct FusedIterArgs {
    const float* x;        // x_k, original vertex order [N*3]
    const float* xkm1;     // x_{k-1}
    float*       out;      // x_{k+1}
    const float* y;
    const float* mass;
    const float* diag;
    const int*   pin_map;
    const float* pin_target;
    double*      partials; // per-tile |r|^2, |b|^2 when non-null
    float h2;
    float omega;
    int   blend;
    int   use_stretch;
    int   use_bend;
};

struct FusedPlan {
    int num_tiles = 0;
    int tile_verts = 0;
    double tri_factor = 0.0, bend_factor = 0.0, bytes_per_iter = 0.0;
};

int  fused_plan_build(FusedPlan& p, int N, const std::vector<int>& tris,
                      const std::vector<int>& quads, int mode);
int  fused_max_blocks(const FusedPlan& p);
int  fused_launch_iteration(const FusedPlan& p, const FusedIterArgs& a, int mode,
                            hipStream_t stream);
void fused_plan_info(const FusedPlan& p, clh_plan_info* out);
void fused_plan_release(FusedPlan& p);

}  // namespace clh
