"""Batched cloths on the host: the concatenated scene equals its cloths built
alone (indices shifted by the cloth offsets), and the fused kernel's tile plan
of a batch keeps every tile inside one cloth (clh_fused_plan_check with cloth
offsets). No device calls."""
import numpy as np
import pytest

import clothhip as ch
from helpers import CFG1

SPECS = [
    dict(CFG1, nrows=12, ncols=9, size=1 / 11, seed=21, stretch_k=3e2, bend_k=1e-3,
         pin_mode=ch.PIN_RANDOM_BOUNDARY, pin_count=3),
    dict(CFG1, nrows=20, ncols=20, size=1 / 19, seed=22, stretch_k=5e3, bend_k=5e-2,
         pin_mode=ch.PIN_TOP_ROW),
    dict(CFG1, nrows=5, ncols=7, size=0.2, type=3, seed=23, stretch_k=1e2, bend_k=0.0,
         pin_mode=ch.PIN_CORNERS),
    dict(CFG1, nrows=16, ncols=16, size=1 / 15, type=ch.MESH_IRREGULAR, rest_jitter=0.3,
         seed=24, stretch_k=5e2, stretch_k_max=5e3, pin_mode=ch.PIN_RANDOM_BOUNDARY, pin_count=1),
]

SHIFTED = {"tris": "vert", "bend_quads": "vert", "pin_idx": "vert"}
PLAIN = ["pos0", "rest_pos", "Dm_inv", "rest_area", "mass", "stretch_k", "jacobi_diag", "bend_w",
         "bend_n", "bend_k", "pin_target"]


def test_batch_scene_is_the_concatenation():
    b = ch.Scene(specs=SPECS)
    assert b.C == len(SPECS)
    offs = {"vert": b.vert_off, "tri": b.tri_off, "bend": b.bend_off, "pin": b.pin_off}
    cnt = {"pos0": "vert", "rest_pos": "vert", "mass": "vert", "jacobi_diag": "vert",
           "tris": "tri", "Dm_inv": "tri", "rest_area": "tri", "stretch_k": "tri",
           "bend_quads": "bend", "bend_w": "bend", "bend_n": "bend", "bend_k": "bend",
           "pin_idx": "pin", "pin_target": "pin"}
    whole = {k: b.get(k) for k in cnt}
    for c, spec in enumerate(SPECS):
        s = ch.Scene(**spec)
        assert (b.vert_off[c + 1] - b.vert_off[c], b.tri_off[c + 1] - b.tri_off[c],
                b.bend_off[c + 1] - b.bend_off[c], b.pin_off[c + 1] - b.pin_off[c]) == \
            (s.N, s.T, s.E, s.P)
        for k, which in cnt.items():
            part = whole[k][offs[which][c]:offs[which][c + 1]]
            mine = s.get(k)
            if k in SHIFTED:
                mine = mine + b.vert_off[c]
            np.testing.assert_array_equal(part, mine, err_msg=f"{k} of cloth {c}")


@pytest.mark.parametrize("tile", [32, 64, 128, 256])
def test_batch_plan_keeps_tiles_inside_cloths(tile):
    b = ch.Scene(specs=SPECS)
    info = ch.fused_plan_check(b.N, b.get("tris"), b.get("bend_quads"), b.get("pos0"), tile,
                               cloth_offsets=b.vert_off)
    V = info.tile_verts
    sizes = np.diff(b.vert_off)
    assert info.num_cloths == len(SPECS)
    assert info.padded_verts == int(np.sum(-(-sizes // V)) * V)
    assert info.num_tiles * V == info.padded_verts


def test_batch_plan_rejects_elements_across_cloths():
    b = ch.Scene(specs=SPECS)
    tris = b.get("tris").copy()
    tris[0, 2] = b.vert_off[1]  # first vertex of the second cloth
    with pytest.raises(ch.ClothHipError, match="spans two cloths"):
        ch.fused_plan_check(b.N, tris, b.get("bend_quads"), b.get("pos0"), 128,
                            cloth_offsets=b.vert_off)


def test_batch_plan_rejects_bad_offsets():
    b = ch.Scene(specs=SPECS)
    bad = b.vert_off.copy()
    bad[-1] -= 1
    with pytest.raises(ch.ClothHipError, match="offsets"):
        ch.fused_plan_check(b.N, b.get("tris"), b.get("bend_quads"), b.get("pos0"), 128,
                            cloth_offsets=bad)
    bad = b.vert_off.copy()
    bad[2] = bad[1]
    with pytest.raises(ch.ClothHipError, match="at least one vertex"):
        ch.fused_plan_check(b.N, b.get("tris"), b.get("bend_quads"), b.get("pos0"), 128,
                            cloth_offsets=bad)


def test_single_cloth_offsets_equal_default_plan():
    s = ch.Scene(**SPECS[1])
    a = ch.fused_plan_check(s.N, s.get("tris"), s.get("bend_quads"), s.get("pos0"), 64)
    b = ch.fused_plan_check(s.N, s.get("tris"), s.get("bend_quads"), s.get("pos0"), 64,
                            cloth_offsets=[0, s.N])
    assert (a.num_tiles, a.padded_verts, a.bytes_per_iter) == \
        (b.num_tiles, b.padded_verts, b.bytes_per_iter)
