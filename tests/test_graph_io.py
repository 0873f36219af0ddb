"""Graph-file loading (Matrix Market and SNAP edge lists) through the C ABI,
host side only.

Pinned by hand-worked known answers (tests/golden/graphs), by agreement
with the oracle's plain-Python restatement of the reference loaders
(bfs_bench.cpp:202-243) on generated files, and by round trips of
generated graphs through both formats."""
import json
import os

import numpy as np
import pytest

from conftest import GOLDEN

GDIR = os.path.join(GOLDEN, "graphs")


def _load(path, fmt="auto", undirected=True):
    from gcq_amd import bfs
    with bfs.Graph.load(path, fmt, undirected) as g:
        return g.csr()


def test_known_answers():
    from oracle import graph_io
    exp = json.load(open(os.path.join(GDIR, "expected.json")))
    for key, want in exp.items():
        name = key.split("#")[0]
        path = os.path.join(GDIR, name)
        off, idx = _load(path, undirected=want["undirected"])
        assert off.tolist() == want["offsets"], key
        assert idx.tolist() == want["indices"], key
        text = open(path).read()
        ro, ri = (graph_io.load_mtx if name.endswith(".mtx") else graph_io.load_snap)(text, want["undirected"])
        assert ro.tolist() == want["offsets"] and ri.tolist() == want["indices"], key


def _random_text(rng, fmt):
    n = int(rng.integers(2, 40))
    m = int(rng.integers(1, 120))
    base = 1 if (fmt == "mtx" or rng.random() < 0.5) else 0
    lines = []
    if fmt == "mtx":
        lines.append("%%MatrixMarket matrix coordinate pattern general")
        lines.append("% generated")
        nr, nc = n, int(rng.integers(1, n + 1))
        lines.append(f"{nr} {nc} {m}")
    for _ in range(m):
        r = rng.random()
        if r < 0.05:
            lines.append("% comment" if fmt == "mtx" else "# comment")
            continue
        if r < 0.08:
            lines.append("")
            continue
        u = int(rng.integers(0, n + 2)) + base - (1 if rng.random() < 0.03 else 0)
        v = u if rng.random() < 0.05 else int(rng.integers(0, n + 2)) + base
        tail = [" 1.5", "\t7", "", " x"][int(rng.integers(0, 4))]
        sep = [" ", "\t", "  "][int(rng.integers(0, 3))]
        lines.append(f"{u}{sep}{v}{tail}")
    return "\n".join(lines) + ("\n" if rng.random() < 0.7 else "")


@pytest.mark.parametrize("fmt", ["mtx", "snap"])
def test_matches_oracle_on_generated_files(tmp_path, fmt):
    from oracle import graph_io
    rng = np.random.default_rng(5)
    ref = graph_io.load_mtx if fmt == "mtx" else graph_io.load_snap
    checked = 0
    for k in range(150):
        text = _random_text(rng, fmt)
        path = tmp_path / f"g{k}.{'mtx' if fmt == 'mtx' else 'txt'}"
        path.write_text(text)
        for und in (True, False):
            try:
                ro, ri = ref(text, und)
            except ValueError:
                with pytest.raises(Exception):
                    _load(str(path), undirected=und)
                continue
            off, idx = _load(str(path), undirected=und)
            assert np.array_equal(off, ro) and np.array_equal(idx, ri), (k, und, text)
            checked += 1
    assert checked > 200


def test_round_trip_lattice_through_both_formats(tmp_path):
    from gcq_amd import bfs
    off, idx = bfs.lattice(30, 41, keep=0.75, diag=0.01, seed=7)
    nv = len(off) - 1
    src = np.repeat(np.arange(nv), np.diff(off))
    upper = src < idx  # one direction of each undirected edge
    mtx = tmp_path / "lat.mtx"
    with open(mtx, "w") as f:
        f.write("%%MatrixMarket matrix coordinate pattern symmetric\n")
        f.write(f"{nv} {nv} {int(upper.sum())}\n")
        for u, v in zip(src[upper], idx[upper]):
            f.write(f"{v + 1} {u + 1}\n")  # lower triangle, as symmetric files store it
    snap = tmp_path / "lat.txt"
    with open(snap, "w") as f:
        f.write("# lattice 30x41\n")
        for u, v in zip(src, idx):
            f.write(f"{u}\t{v}\n")
    for path, fmt in ((mtx, "auto"), (snap, "auto"), (snap, "snap")):
        o2, i2 = _load(str(path), fmt)
        assert np.array_equal(o2, off) and np.array_equal(i2, idx), path
    # written directed, read directed: the same CSR again
    o3, i3 = _load(str(snap), "snap", undirected=False)
    assert np.array_equal(o3, off) and np.array_equal(i3, idx)


def test_errors(tmp_path):
    from gcq_amd import GcqError
    bad = tmp_path / "nobanner.mtx"
    bad.write_text("3 3 1\n1 2\n")
    empty = tmp_path / "empty.txt"
    empty.write_text("# nothing\n5 5\n")
    for p, fmt in ((bad, "auto"), (empty, "snap"), (tmp_path / "missing.txt", "snap")):
        with pytest.raises(GcqError):
            _load(str(p), fmt)
