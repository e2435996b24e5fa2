"""Host side of the block-sparse group fit (astrophot_amd/group_fit.py): the
per-tile source lists and overlap pairs against brute force, which groups
may take the route, and the joint-group scene generator.  No GPU."""

import math

import numpy as np
import pytest

import astrophot_amd as ap
from astrophot_amd.group_fit import TILE_H, TILE_W, SparseGroup, plan_tiles, scene_group
from astrophot_amd.population import group_scene


def brute(wins, nx, ny):
    ntx, nty = math.ceil(nx / TILE_W), math.ceil(ny / TILE_H)
    lists = []
    for ty in range(nty):
        for tx in range(ntx):
            t = []
            for s, (x0, y0, w, h) in enumerate(wins):
                if x0 < (tx + 1) * TILE_W and x0 + w > tx * TILE_W and y0 < (ty + 1) * TILE_H and y0 + h > ty * TILE_H:
                    t.append(s)
            lists.append(t)
    pairs = []
    for a in range(len(wins)):
        for b in range(a, len(wins)):
            xa, ya, wa, ha = wins[a]
            xb, yb, wb, hb = wins[b]
            if min(xa + wa, xb + wb) > max(xa, xb) and min(ya + ha, yb + hb) > max(ya, yb):
                pairs.append((a, b))
    return lists, pairs


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_plan_tiles_matches_brute_force(seed):
    rng = np.random.default_rng(seed)
    nx, ny = 300, 170
    wins = []
    for _ in range(25):
        x0, y0 = rng.integers(0, nx - 1), rng.integers(0, ny - 1)
        wins.append((x0, y0, rng.integers(1, nx - x0 + 1), rng.integers(1, ny - y0 + 1)))
    wins = np.array(wins)
    ntx, nty, tp, ts, pairs = plan_tiles(wins, nx, ny)
    lists, bp = brute(wins, nx, ny)
    assert (ntx, nty) == (math.ceil(nx / TILE_W), math.ceil(ny / TILE_H))
    assert [list(ts[tp[k]:tp[k + 1]]) for k in range(ntx * nty)] == lists
    assert [tuple(p) for p in pairs.reshape(-1, 2)] == bp


def test_group_scene_priors():
    s = group_scene(400, 7, side=1024)
    t = s.truth
    assert t.shape == (400, 7) and s.psf.shape == (65, 65) and abs(s.psf.sum() - 1) < 1e-12
    assert np.all((t[:, 0] >= 0) & (t[:, 0] <= 1024)) and np.all((t[:, 4] >= 0.5) & (t[:, 4] <= 6))
    assert np.all((s.windows[:, 0] >= 0) & (s.windows[:, 1] <= 1024) & (s.windows[:, 1] > s.windows[:, 0]))
    half = np.maximum(5 * t[:, 5], 33)
    inner = (t[:, 0] - half > 0) & (t[:, 0] + half < 1024) & (t[:, 1] - half > 0) & (t[:, 1] + half < 1024)
    assert np.all(s.windows[inner, 1] - s.windows[inner, 0] >= 2 * half[inner])
    assert 2.5 <= s.beta <= 4.5 and abs(s.Rd - 3 / (2 * math.sqrt(2 ** (1 / s.beta) - 1))) < 1e-12


def small_scene():
    s = group_scene(6, 11, side=120, psf_side=11, min_half=16, re_range=(1.5, 5.0))
    tgt = ap.image.Target_Image(data=np.zeros((120, 120)), pixelscale=1.0, variance=np.ones((120, 120)),
                                psf=ap.image.PSF_Image(data=s.psf, pixelscale=1.0))
    return s, tgt


def test_scene_group_takes_the_sparse_route():
    s, tgt = small_scene()
    g = scene_group(s, tgt)
    assert g.parameters.size == 6 * 7 + 1
    assert SparseGroup.unsupported(g, g.target[g.window].window) is None
    # the last source's parameters sit just before the sky's
    assert list(g.parameters.vector_names())[-8:-1] == ["center:0", "center:1", "q", "PA", "n", "Re", "Ie"]


def test_unsupported_groups_say_why():
    s, tgt = small_scene()
    g = scene_group(s, tgt)
    g.models["src2"]["n"].locked = True
    assert "free" in SparseGroup.unsupported(g, g.target[g.window].window)
    g.models["src2"]["n"].locked = False
    g.models["src3"].psf_mode = "none"
    assert "psf_mode" in SparseGroup.unsupported(g, g.target[g.window].window)
    g.models["src3"].psf_mode = "full"
    g.add_model(ap.models.AstroPhot_Model(name="m", model_type="moffat galaxy model", target=tgt,
                                          parameters={"center": [50, 50], "q": 0.5, "PA": 1.0, "n": 2.0,
                                                      "Rd": 3.0, "I0": 0.0}))
    assert "only Sersic" in SparseGroup.unsupported(g, g.target[g.window].window)
    single = ap.models.AstroPhot_Model(name="x", model_type="sersic galaxy model", target=tgt,
                                       parameters={"center": [50, 50], "q": 0.5, "PA": 1.0, "n": 2.0,
                                                   "Re": 3.0, "Ie": 0.0})
    assert SparseGroup.unsupported(single, single.window) == "not a group model"


def test_group_window_union_follows_reference_rounding():
    """window_object.py:483-495 takes the union's end as max(pixel_shape,
    other's far corner in pixel coordinates) while corners sit at n - 0.5, and
    rounds the shape half to even: an odd-width window united with one that
    ends inside it grows by a pixel.  The block-sparse fit runs on the
    target's part of such a group window."""
    tgt = ap.image.Target_Image(data=np.zeros((40, 80)), pixelscale=1.0)
    a = tgt.window.copy().crop_to_pixel([[0, 67], [0, 40]])
    b = tgt.window.copy().crop_to_pixel([[10, 30], [5, 20]])
    u = a | b
    # origin stays at -0.5, end = max(67, 29.5) = 67 -> shape 67.5 -> 68 (half to even)
    assert (u.nx, u.ny) == (68, 40)
    a2 = tgt.window.copy().crop_to_pixel([[0, 66], [0, 40]])
    assert ((a2 | b).nx, (a2 | b).ny) == (66, 40)
    s = group_scene(30, 20261205, side=300, psf_side=11, min_half=16, re_range=(1.5, 5.0))
    t = ap.image.Target_Image(data=np.zeros((300, 300)), pixelscale=1.0, variance=np.ones((300, 300)),
                              psf=ap.image.PSF_Image(data=s.psf, pixelscale=1.0))
    g = scene_group(s, t)
    grid = g.target[g.window].window
    assert (grid.nx, grid.ny) == (300, 300)
    assert SparseGroup.unsupported(g, grid) is None
