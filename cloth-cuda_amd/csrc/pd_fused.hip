// Fused LDS-tiled iteration kernel (not yet available: the unfused
// pipeline is used).
#include "pd_fused.h"

namespace clh {

int fused_plan_build(FusedPlan&, int, const std::vector<int>&, const std::vector<int>&, int)
{
    return fail(CLH_ERR_UNSUPPORTED, "fused kernel variant not available");
}
int fused_max_blocks(const FusedPlan& p) { return p.num_tiles; }
int fused_launch_iteration(const FusedPlan&, const FusedIterArgs&, int, hipStream_t)
{
    return fail(CLH_ERR_UNSUPPORTED, "fused kernel variant not available");
}
void fused_plan_info(const FusedPlan& p, clh_plan_info* out)
{
    out->num_tiles = p.num_tiles;
    out->tile_verts = p.tile_verts;
    out->halo_tri_factor = p.tri_factor;
    out->halo_bend_factor = p.bend_factor;
    out->bytes_per_iter = p.bytes_per_iter;
}
void fused_plan_release(FusedPlan&) {}

}  // namespace clh
