"""The fused kernel's tile plan, built and checked on the host: every local
index inside its tile's LDS table, halo ids outside the tile, each local
slot naming its element's vertex, and each vertex's gather list equal to its
incident element slots in ascending order (clh_fused_plan_check)."""
import numpy as np
import pytest

import clothhip as ch
from helpers import CFG1, CFG2


CASES = {
    "cfg1": dict(CFG1),
    "cross": dict(CFG1, nrows=12, ncols=12, size=1 / 11, type=3),
    "checker": dict(CFG1, nrows=20, ncols=17, size=0.05, type=1),
    "irregular": dict(CFG1, nrows=48, ncols=48, size=1 / 47, type=ch.MESH_IRREGULAR,
                      rest_jitter=0.3, seed=3),
    "no_bend": dict(CFG1, bend_k=0.0),
    "tiny": dict(CFG1, nrows=2, ncols=2, size=1.0),
}


@pytest.mark.parametrize("case", list(CASES))
@pytest.mark.parametrize("tile", [32, 128, 256])
def test_plan_is_consistent(case, tile):
    s = ch.Scene(**CASES[case])
    info = ch.fused_plan_check(s.N, s.get("tris"), s.get("bend_quads") if s.E else None,
                               s.get("pos0"), tile)
    assert info.num_tiles == -(-s.N // info.tile_verts)
    assert 1.0 <= info.halo_tri_factor < 3.0
    # PD plans carry the two-iteration extension (ring 2, ring-1 gather lists),
    # checked by the same call
    assert info.pair == 1
    assert info.pair_halo_max >= info.halo_max and info.pair_tri_max >= info.tri_max
    assert info.pair_hinge_max >= info.hinge_max


def test_reference_mode_plan_has_no_pair_extension():
    s = ch.Scene(**CFG1)
    info = ch.fused_plan_check(s.N, s.get("tris"), s.get("bend_quads"), s.get("pos0"), 128,
                               mode=ch.MODE_REF_CUDA)
    assert info.pair == 0


def test_plan_quality_on_grid():
    # Morton tiles of a flat grid are near-square: little duplication
    s = ch.Scene(**CFG2)
    info = ch.fused_plan_check(s.N, s.get("tris"), s.get("bend_quads"), s.get("pos0"), 128)
    assert info.halo_tri_factor < 1.25 and info.halo_bend_factor < 1.3


def test_plan_rejects_bad_indices():
    s = ch.Scene(**CFG1)
    tris = s.get("tris").copy()
    tris[3, 2] = s.N
    with pytest.raises(ch.ClothHipError, match="out of range"):
        ch.fused_plan_check(s.N, tris, s.get("bend_quads"), s.get("pos0"), 128)


def test_plan_rejects_non_finite_positions():
    # the tile order is built from the positions; NaN would order at random
    s = ch.Scene(**CFG1)
    pos = s.get("pos0").copy()
    pos[7, 1] = np.nan
    with pytest.raises(ch.ClothHipError, match="non-finite"):
        ch.fused_plan_check(s.N, s.get("tris"), s.get("bend_quads"), pos, 128)
