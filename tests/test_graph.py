"""CPU tests of the BFS inputs and of the BFS oracle.

The synthetic graphs must have the CSR form bfs_bench.cpp:181-200 builds
(symmetric, sorted, deduplicated, no self loops).  The serial BFS oracle is
pinned against closed forms (full lattice: depth = Manhattan distance from
the corner; path; star; disconnected pieces) and against scipy's
breadth-first search, a library routine, on random graphs."""
import numpy as np
import pytest
import scipy.sparse as sp
from scipy.sparse.csgraph import shortest_path


def _check_csr(off, idx, nv):
    assert off.dtype == np.int32 and idx.dtype == np.int32
    assert len(off) == nv + 1 and off[0] == 0 and off[-1] == len(idx)
    assert np.all(np.diff(off) >= 0)
    if len(idx) == 0:
        return
    assert idx.min() >= 0 and idx.max() < nv
    src = np.repeat(np.arange(nv, dtype=np.int64), np.diff(off))
    assert not np.any(src == idx), "self loop"
    # Sorted and unique within each row: keys strictly increase along the array.
    key = src * nv + idx
    assert np.all(np.diff(key) > 0), "rows not sorted/deduplicated"
    # Symmetric: the set of (v, u) equals the set of (u, v).
    rkey = np.sort(idx.astype(np.int64) * nv + src)
    assert np.array_equal(rkey, key), "not symmetric"


def test_full_lattice_matches_grid_and_manhattan(gcq, orc):
    from gcq_amd import bfs
    R, C = 37, 53
    off, idx = bfs.lattice(R, C, keep=1.0, diag=0.0, seed=7)
    _check_csr(off, idx, R * C)
    # bfs_bench.cpp:245-256 full 4-neighbour grid: 2 * (R(C-1) + (R-1)C) directed edges.
    assert len(idx) == 2 * (R * (C - 1) + (R - 1) * C)
    d, reached, es = orc.bfs_serial(off, idx, 0)
    r, c = np.divmod(np.arange(R * C), C)
    assert np.array_equal(d, (r + c).astype(np.int32))
    assert reached == R * C and es == len(idx)
    # From an interior source the closed form is |dr| + |dc|.
    s = 17 * C + 29
    d, _, _ = orc.bfs_serial(off, idx, s)
    assert np.array_equal(d, (np.abs(r - 17) + np.abs(c - 29)).astype(np.int32))


@pytest.mark.parametrize("keep,diag,seed", [(0.75, 0.01, 7), (0.5, 0.2, 3), (0.9, 1.0, 11)])
def test_lattice_form_and_density(gcq, keep, diag, seed):
    from gcq_amd import bfs
    R, C = 120, 90
    off, idx = bfs.lattice(R, C, keep=keep, diag=diag, seed=seed)
    _check_csr(off, idx, R * C)
    grid = R * (C - 1) + (R - 1) * C
    src = np.repeat(np.arange(R * C), np.diff(off))
    dr = np.abs(idx // C - src // C)
    dc = np.abs(idx % C - src % C)
    assert np.all(dr <= 1) and np.all(dc <= 1), "only lattice and short diagonal edges"
    straight = int(np.sum(dr + dc == 1)) // 2
    diagonal = int(np.sum(dr + dc == 2)) // 2
    assert abs(straight - keep * grid) < 5 * np.sqrt(grid * keep * (1 - keep)) + 1
    # Each vertex not on the last row draws one diagonal; about a quarter of
    # the last column's draws point outside the grid.
    exp_diag = diag * (R - 1) * (C - 0.5)
    assert abs(diagonal - exp_diag) < 5 * np.sqrt(exp_diag + 1) + 2
    off2, idx2 = bfs.lattice(R, C, keep=keep, diag=diag, seed=seed)
    assert np.array_equal(off, off2) and np.array_equal(idx, idx2), "deterministic in the seed"
    off3, idx3 = bfs.lattice(R, C, keep=keep, diag=diag, seed=seed + 1)
    assert not (len(idx3) == len(idx) and np.array_equal(idx3, idx))


def test_config3_lattice_size(gcq):
    # The 4900 x 4900 lattice, p = 0.75, 1% diagonals, seed 7: the graph the
    # BFS benchmark runs (a fixed property of this generator).
    from gcq_amd import bfs
    with bfs.Graph.lattice(4900, 4900, 0.75, 0.01, 7) as g:
        nv, ne = g.info()
        assert nv == 24_010_000
        assert ne == 72_502_886
        assert g.pick_source(7, 0) >= 0


@pytest.mark.parametrize("scale", [6, 10, 14])
def test_rmat_form(gcq, scale):
    from gcq_amd import bfs
    with bfs.Graph.rmat(scale, 16, 0.57, 0.19, 0.19, 11) as g:
        off, idx = g.csr()
        _check_csr(off, idx, 1 << scale)
        assert len(idx) <= 32 * (1 << scale)
        deg = np.diff(off)
        # R-MAT skew: the maximum degree is far above the mean.
        assert deg.max() > 4 * deg.mean()
        for k in range(8):
            s = g.pick_source(11, k)
            assert 0 <= s < (1 << scale) and deg[s] > 0


def test_from_csr_validation(gcq):
    from gcq_amd import bfs
    with pytest.raises(RuntimeError):
        bfs.Graph.from_csr(np.array([0, 2, 1], np.int32), np.array([1], np.int32))  # bad end
    with pytest.raises(RuntimeError):
        bfs.Graph.from_csr(np.array([0, 2, 1, 3], np.int32), np.array([1, 2, 0], np.int32))  # decreasing
    with pytest.raises(RuntimeError):
        bfs.Graph.from_csr(np.array([0, 1, 2], np.int32), np.array([1, 5], np.int32))  # index out of range
    with bfs.Graph.from_csr(np.array([0, 0, 0], np.int32), np.array([], np.int32)) as g:
        assert g.info() == (2, 0)


def _scipy_depths(off, idx, src):
    nv = len(off) - 1
    a = sp.csr_matrix((np.ones(len(idx)), idx, off), shape=(nv, nv))
    d = shortest_path(a, unweighted=True, directed=True, indices=src)
    return np.where(np.isinf(d), -1, d).astype(np.int32)


def _random_csr(rng, nv, ne):
    u = rng.integers(0, nv, ne)
    v = rng.integers(0, nv, ne)
    keep = u != v
    return _csr_from_edges(nv, u[keep], v[keep])


@pytest.mark.parametrize("seed", range(6))
def test_bfs_oracle_matches_scipy(orc, seed):
    rng = np.random.default_rng(seed)
    nv = int(rng.integers(2, 400))
    off, idx = _random_csr(rng, nv, int(rng.integers(0, 3 * nv)))
    src = int(rng.integers(0, nv))
    d, reached, es = orc.bfs_serial(off, idx, src)
    ref = _scipy_depths(off, idx, src)
    assert np.array_equal(d, ref)
    assert reached == int(np.sum(ref >= 0))
    deg = np.diff(off)
    assert es == int(deg[ref >= 0].sum())
    dp, rp = orc.bfs_parallel(off, idx, src, 4)
    assert np.array_equal(dp, ref) and rp == reached


def test_bfs_oracle_special_graphs(orc):
    # Path 0-1-...-9: depth i.  Star centred at 0: depth 1.  Two components.
    n = 10
    a = np.arange(n - 1)
    off, idx = _csr_from_edges(n, a, a + 1)
    assert np.array_equal(orc.bfs_serial(off, idx, 0)[0], np.arange(n, dtype=np.int32))
    assert np.array_equal(orc.bfs_serial(off, idx, n - 1)[0], np.arange(n - 1, -1, -1, dtype=np.int32))
    off, idx = _csr_from_edges(n, np.zeros(n - 1, int), np.arange(1, n))
    d = orc.bfs_serial(off, idx, 0)[0]
    assert d[0] == 0 and np.all(d[1:] == 1)
    off, idx = _csr_from_edges(6, np.array([0, 1, 3, 4]), np.array([1, 2, 4, 5]))
    d, reached, _ = orc.bfs_serial(off, idx, 0)
    assert list(d) == [0, 1, 2, -1, -1, -1] and reached == 3
    # Out-of-range source falls back to vertex 0 (bfs_bench.cpp:555).
    assert np.array_equal(orc.bfs_serial(off, idx, 99)[0], d)


def _csr_from_edges(nv, u, v):
    key = np.unique(np.concatenate([u * nv + v, v * nv + u]))
    s, t = np.divmod(key, nv)
    off = np.zeros(nv + 1, np.int32)
    np.add.at(off, s + 1, 1)
    return np.cumsum(off).astype(np.int32), t.astype(np.int32)
