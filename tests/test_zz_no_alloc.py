"""GPU check of the allocation-free point lookups
(g500_msbfs_gather_levels_noalloc) against g500_msbfs_gather_levels.

Both read the same batch; the allocation-free form stages its pairs and levels
in the library's fixed device array.  Two calls in a row, the second smaller
than the first, must each equal the allocating entry point's answer byte for
byte (the second call must not see the first call's leftovers).  Counts above
the fixed maximum are refused with G500_ERR_ARG, and the refusal leaves the
engine usable.
"""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def G():
    import g500bfs
    g500bfs.load()
    return g500bfs


@pytest.fixture(scope="module")
def batch(G):
    g = G.Graph.generate(14).construct()
    roots = g.find_roots(96, r1=3)
    mb = G.MultiBfs(g, max_roots=96)
    mb.run(roots)
    return g, mb, roots


def _pairs(rng, n, nroots, count):
    ri = rng.integers(0, nroots, count)
    vs = rng.integers(-2, n + 2, count)  # a few out-of-range vertices too
    return ri, vs


def test_noalloc_lookups_equal_the_allocating_ones(G, batch):
    print("\n[noalloc] two calls, the second smaller", flush=True)
    g, mb, roots = batch
    rng = np.random.default_rng(17)
    n = g.info.n
    for count in (G.NOALLOC_MAX_PAIRS, 1000):
        ri, vs = _pairs(rng, n, len(roots), count)
        want = mb.gather(ri, vs)
        got = mb.gather(ri, vs, noalloc=True)
        assert got.dtype == np.uint8 and len(got) == count
        assert np.array_equal(got, want), count
        # spot check against the level arrays themselves
        for i in range(0, count, max(1, count // 50)):
            v = int(vs[i])
            exp = 0xFF if v < 0 or v >= n else int(mb.levels(int(ri[i]))[v])
            assert int(got[i]) == exp


def test_noalloc_rejects_counts_over_the_maximum(G, batch):
    print("\n[noalloc] over the maximum", flush=True)
    g, mb, roots = batch
    ri = np.zeros(G.NOALLOC_MAX_PAIRS + 1, np.int64)
    with pytest.raises(G.G500Error, match="at most"):
        mb.gather(ri, ri, noalloc=True)
    with pytest.raises(G.G500Error):
        mb.gather([len(roots)], [0], noalloc=True)  # root index out of range
    # still usable, and an empty call is a no-op
    assert len(mb.gather([], [], noalloc=True)) == 0
    ri, vs = _pairs(np.random.default_rng(3), g.info.n, len(roots), 77)
    assert np.array_equal(mb.gather(ri, vs, noalloc=True), mb.gather(ri, vs))
