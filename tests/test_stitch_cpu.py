"""Tile / window plan of the stitched (batching-form) codec: the tiles cover the frame,
the clamped last tile ends at the edge, and the raised-cosine blend weights sum to one at
every pixel (the property the reference's Hann ramps rely on, vtem/vtdm.py:252-268)."""

import numpy as np
import pytest

from vtem_amd.stitch import StitchPlan, blend_weights, coverage, tile_origins


@pytest.mark.parametrize("P,B,O", [(512, 64, 16), (128, 64, 16), (400, 100, 25), (64, 64, 16),
                                   (100, 64, 16), (300, 96, 20), (115, 64, 16), (1080, 64, 16),
                                   (410, 100, 25)])
def test_tiles_cover_and_blend_to_one(P, B, O):
    org = tile_origins(P, B, O)
    assert org[0] == 0 and org[-1] + B == max(P, B)
    assert all(b > a for a, b in zip(org, org[1:]))
    assert all(b - a <= B - O for a, b in zip(org, org[1:]))      # overlaps at least O
    acc = np.zeros(max(P, B))
    for a, w in zip(org, blend_weights(org, B)):
        assert np.all(w > 0) and np.all(w <= 1)
        acc[a:a + B] += w
    np.testing.assert_allclose(acc[:P], 1.0, atol=1e-15)
    # interleaved passes: tiles i and i + k (k = coverage) never overlap
    k = coverage(org, B)
    for i in range(len(org) - k):
        assert org[i + k] >= org[i] + B


@pytest.mark.parametrize("Px,Py,F", [(115, 1080, 410), (512, 512, 400), (100, 300, 250)])
def test_plan_passes_are_disjoint_and_complete(Px, Py, F):
    """Every block is in exactly one pass, and no two blocks of a pass overlap in all
    three axes (so the plain += of the accumulation kernel has one writer per element)."""
    p = StitchPlan(Px, Py, F, 64, 16, 100, 25)
    seen = []
    for grp in p.passes():
        seen += grp
        for a in range(len(grp)):
            for b in range(a + 1, len(grp)):
                (y1, x1, t1), (y2, x2, t2) = grp[a], grp[b]
                ov = (abs(p.y0[y1] - p.y0[y2]) < 64 and abs(p.x0[x1] - p.x0[x2]) < 64
                      and abs(p.f0[t1] - p.f0[t2]) < 100)
                assert not ov, (grp[a], grp[b])
    assert sorted(seen) == sorted(p.blocks)


def test_config3_plan():
    """512^2 px, 4 s at 100 fps: 64 px tiles with 16 px overlap and 1 s windows with
    0.25 s overlap give 11 x 11 tiles x 5 windows = 605 blocks."""
    p = StitchPlan(512, 512, 400, 64, 16, 100, 25)
    assert p.x0 == [0, 48, 96, 144, 192, 240, 288, 336, 384, 432, 448]
    assert p.f0 == [0, 75, 150, 225, 300]
    assert len(p.blocks) == 605
