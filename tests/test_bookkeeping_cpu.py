"""The scheduler's tile bookkeeping, on the CPU: row-major tile ids (src/eikonal_hcm.cpp:209),
partial edge tiles (data/eikonal_ghcm_fim.cl:253), neighbours across the four sides and the
border bit each activation sets (src/eikonal_hcm.cpp:18, 442-462).  ghcm_tile_info is the
code the kernels and the host scheduler both use (tile_geom.h)."""
import numpy as np
import pytest

import ghcm_amd as g

NO_TILE = 0xFFFFFFFF
SHAPES = [(1, 1, 8, 8), (8, 8, 8, 8), (100, 70, 32, 32), (257, 129, 64, 16), (64, 64, 64, 64),
          (65, 63, 64, 64), (4096, 3, 32, 8), (3, 4096, 16, 16)]


@pytest.mark.parametrize("w,h,tw,th", SHAPES)
def test_tiles_cover_the_grid_once_in_row_major_order(w, h, tw, th):
    base = 17
    tx_n, ty_n = -(-w // tw), -(-h // th)
    owner = np.full((h, w), -1, np.int64)
    for t in range(base, base + tx_n * ty_n):
        d = g.tile_info(w, h, tw, th, t, tile_base=base)
        assert (d["tiles_x"], d["tiles_y"]) == (tx_n, ty_n)
        # row-major: id = base + tx + tiles_x * ty
        assert t == base + d["tx"] + tx_n * d["ty"]
        assert (d["x0"], d["y0"]) == (d["tx"] * tw, d["ty"] * th)
        # partial tiles only on the far edges, full size elsewhere
        assert d["w"] == min(tw, w - d["x0"]) and d["h"] == min(th, h - d["y0"])
        assert d["w"] >= 1 and d["h"] >= 1
        blk = owner[d["y0"]:d["y0"] + d["h"], d["x0"]:d["x0"] + d["w"]]
        assert np.all(blk == -1), "tiles overlap"
        blk[...] = t
        # the cell at the tile's far corner maps back to the tile
        assert d["seed_tile"] == t
    assert np.all(owner >= 0), "tiles leave cells uncovered"


@pytest.mark.parametrize("w,h,tw,th", SHAPES)
def test_neighbours_and_border_bits(w, h, tw, th):
    tx_n, ty_n = -(-w // tw), -(-h // th)
    for t in range(tx_n * ty_n):
        d = g.tile_info(w, h, tw, th, t)
        steps = {0: (-1, 0), 1: (0, -1), 2: (1, 0), 3: (0, 1)}     # 0 x-, 1 y-, 2 x+, 3 y+
        for s, (dx, dy) in steps.items():
            nx, ny = d["tx"] + dx, d["ty"] + dy
            if 0 <= nx < tx_n and 0 <= ny < ty_n:
                assert d["neighbour"][s] == nx + tx_n * ny
                # the neighbour's neighbour on the opposite side is this tile
                back = g.tile_info(w, h, tw, th, d["neighbour"][s])
                assert back["neighbour"][s ^ 2] == t
            else:
                assert d["neighbour"][s] == NO_TILE
            # the activated neighbour learns which of its own sides changed: the one facing
            # this tile, s ^ 2 (the reference sets byte s ^ 2 of a byte mask; here bit s ^ 2)
            assert d["border_bit"][s] == 1 << (s ^ 2)
    # opposite sides pair up 0 <-> 2, 1 <-> 3 and the four bits are distinct
    assert sorted(g.tile_info(w, h, tw, th, 0)["border_bit"]) == [1, 2, 4, 8]


def test_tile_of_a_random_cell(rng):
    w, h, tw, th = 1000, 777, 64, 32
    tx_n = -(-w // tw)
    for _ in range(200):
        x, y = int(rng.integers(0, w)), int(rng.integers(0, h))
        t = (y // th) * tx_n + x // tw
        d = g.tile_info(w, h, tw, th, t)
        assert d["x0"] <= x < d["x0"] + d["w"] and d["y0"] <= y < d["y0"] + d["h"]


def test_tile_info_rejects_bad_arguments():
    with pytest.raises(g.GhcmError):
        g.tile_info(100, 70, 32, 32, 12)          # 4 x 3 tiles: ids 0..11
    with pytest.raises(g.GhcmError):
        g.tile_info(100, 70, 32, 32, 3, tile_base=4)
    with pytest.raises(g.GhcmError):
        g.tile_info(0, 70, 32, 32, 0)
