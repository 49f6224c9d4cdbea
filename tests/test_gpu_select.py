"""The device top-n selection of a wave boundary, pinned on known lists -- needs an MI355X.

A wave pops at most n tiles of lowest key (src/eikonal_hcm.cpp:344-358, a std::priority_queue
pop).  The device boundary (k_wave_end) resolves keys to 2^7-ulp bins: with lo the bin of the
n-th smallest key, every key below lo is selected, none at or above lo + 128, and inside the
bin the first ones in list order fill the quota.  That is a complete specification, so each
case is checked exactly against it (a sort on the host), together with what happens to the
tiles that were not selected: each is carried once into the next wave's list with its key and
border bits, or, after a later boundary of a long list, handed to the tile launch in list
order.  Every call goes through the C-ABI (ghcm_debug_select)."""
import numpy as np
import pytest

import ghcm_amd as g

pytestmark = pytest.mark.gpu
SHARDS = 32


def list_order(tiles):
    """Order of the boundary's list: shard g % 32, then insertion order within a shard."""
    idx = np.arange(len(tiles))
    return idx[np.lexsort((idx, tiles % SHARDS))]


def expected(tiles, keys, n):
    order = list_order(tiles)
    c = len(tiles)
    if c <= n:
        return tiles[order], np.array([], np.uint32)
    bits = keys.view(np.uint32).astype(np.int64)
    kth = np.sort(bits)[n - 1]
    lo = kth & ~np.int64(127)
    below = bits < lo
    tie = (bits >= lo) & (bits < lo + 128)
    quota = n - int(below.sum())
    assert quota >= 1
    sel = below.copy()
    taken = 0
    for i in order:
        if tie[i] and taken < quota:
            sel[i] = True
            taken += 1
    jobs = tiles[[i for i in order if sel[i]]]
    carried = tiles[[i for i in order if not sel[i]]]
    return jobs, carried


def make_keys(rng, c, kind):
    if kind == "uniform":
        return rng.uniform(0.0, 1000.0, c).astype(np.float32)
    if kind == "duplicates":
        return rng.choice(np.array([0.0, 1.5, 1.5, 7.25, 300.0], np.float32), c)
    if kind == "one_bin":
        # every key inside one 2^7-ulp bin: the quota decides in list order
        b = np.float32(123.0).view(np.uint32) & ~np.uint32(127)
        return (b + rng.integers(0, 128, c).astype(np.uint32)).view(np.float32)
    if kind == "with_inf":
        k = rng.uniform(0.0, 50.0, c).astype(np.float32)
        k[rng.random(c) < 0.3] = np.inf
        return k
    raise ValueError(kind)


CASES = [(5, 1), (300, 64), (300, 299), (300, 300), (300, 512), (2047, 512), (2049, 512),
         (5000, 512), (5000, 4999), (16384, 4096), (20000, 512), (20000, 19999)]


@pytest.fixture(scope="module")
def ctx():
    with g.Context(0) as c:
        yield c


@pytest.mark.parametrize("handoff", [False, True])
@pytest.mark.parametrize("kind", ["uniform", "duplicates", "one_bin", "with_inf"])
@pytest.mark.parametrize("c,n", CASES)
def test_selection_matches_its_specification(ctx, c, n, kind, handoff):
    rng = np.random.default_rng(c * 7919 + n + (1 if handoff else 0))
    total = 3 * c + 11
    tiles = rng.choice(total, c, replace=False).astype(np.uint32)
    keys = make_keys(rng, c, kind)
    masks = rng.integers(0, 32, c).astype(np.uint32)
    print(f"case select c={c} n={n} keys={kind} handoff={handoff}", flush=True)
    jobs, handed, nxt, nkeys, nmasks = ctx.debug_select(total, tiles, keys, masks, n, handoff)
    exp_jobs, exp_carried = expected(tiles, keys, n)
    # exactly min(n, c) jobs, no duplicates, the right ones in list order
    assert len(jobs) == min(n, c)
    assert len(np.unique(jobs)) == len(jobs)
    assert np.array_equal(jobs, exp_jobs)
    if c > n:
        bits = keys.view(np.uint32).astype(np.int64)
        pos = {int(t): i for i, t in enumerate(tiles)}
        sel_bits = bits[[pos[int(t)] for t in jobs]]
        car_bits = bits[[pos[int(t)] for t in exp_carried]]
        # every selected key is at most every carried key plus the 2^7-ulp bin width
        assert sel_bits.max() < car_bits.min() + 128
    # the tiles not selected: handed to the tile launch (a later boundary of a long staged
    # list) or carried by the boundary into the next list, once each, key and bits kept
    long_staged = 2 * 1024 < c <= 16384
    if handoff and long_staged and c > n:
        assert np.array_equal(handed, exp_carried)
        assert len(nxt) == 0
    else:
        assert len(handed) == 0
        assert len(nxt) == len(exp_carried)
        assert np.array_equal(np.sort(nxt), np.sort(exp_carried))
        key_of = dict(zip(tiles.tolist(), keys.tolist()))
        mask_of = dict(zip(tiles.tolist(), masks.tolist()))
        for t, k, m in zip(nxt.tolist(), nkeys.tolist(), nmasks.tolist()):
            assert (k == key_of[t]) or (np.isinf(k) and np.isinf(key_of[t]))
            assert m == mask_of[t]
