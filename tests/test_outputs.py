"""Byte-level checks of the library's file outputs, through the C ABI
(host-only functions, so they run without a GPU):
 - gcq_hist_write_jsonl: the reference's JSONL history schema
   (history.cpp:249-283).  Re-writing the reference's own golden histories
   must reproduce their bytes, and a written history must read back through
   the oracle's JSONL reader field for field;
 - gcq_rt_write_ppm: save_ppm's header and truncating quantisation
   (rt_final.cpp:689-700);
 - gcq_rt_append_csv: the header and column order of rt_queue_perf.csv
   (rt_final.cpp:972-1001) and rt_performance.csv
   (baselines/rt_compaction.cpp:849-864)."""
import os
import re

import numpy as np
import pytest

from conftest import GOLDEN

KEYS = re.compile(r'^\{"proc":\d+,"op":[01],"arg":\d+,"ret":\d+,"call":\d+,"end":\d+\}$')


def _as_hist(gcq, ops):
    h = np.zeros(len(ops), dtype=gcq.HIST_DTYPE)
    for f in ("proc", "op", "arg", "ret", "call", "end"):
        h[f] = ops[f]
    return h


@pytest.mark.parametrize("name", ["history_glfq_m0_t64_o16.jsonl", "history_gwfq_m0_t32_o8.jsonl"])
def test_jsonl_rewrite_of_reference_histories_is_byte_identical(gcq, orc, tmp_path, name):
    src = os.path.join(GOLDEN, name)
    ops = orc.read_jsonl(src)
    out = tmp_path / "sub" / "dir" / name  # parent directories are created, as in the reference
    gcq.write_jsonl(_as_hist(gcq, ops), str(out))
    assert out.read_bytes() == open(src, "rb").read()


def test_jsonl_round_trip_and_key_order(gcq, orc, tmp_path):
    rng = np.random.default_rng(5)
    n = 500
    h = np.zeros(n, dtype=gcq.HIST_DTYPE)
    h["proc"] = rng.integers(0, 65536, n)
    h["op"] = rng.integers(0, 2, n)
    h["arg"] = np.where(h["op"] == 0, (h["proc"].astype(np.uint64) << np.uint64(32)) | np.uint64(7), 0)
    h["ret"] = rng.integers(0, 2**63, n, dtype=np.uint64)
    h["call"] = np.arange(1, 2 * n, 2)
    h["end"] = h["call"] + rng.integers(1, 1000, n).astype(np.uint64)
    path = tmp_path / "h.jsonl"
    gcq.write_jsonl(h, str(path))
    lines = path.read_text().splitlines()
    assert len(lines) == n and all(KEYS.match(x) for x in lines)
    back = orc.read_jsonl(str(path))
    for f in ("proc", "op", "arg", "ret", "call", "end"):
        assert np.array_equal(back[f], h[f]), f


def test_ppm_bytes(gcq, tmp_path):
    from gcq_amd import rt
    vals = np.array([0.0, 0.5, 1.0, 1.7, -0.2, 0.99999, np.float32(1 / 255), 0.25, np.nan, 3e38, 1e-9, 0.7],
                    np.float32)
    img = np.resize(vals, (3, 5, 3)).astype(np.float32)  # 5 wide, 3 high
    path = tmp_path / "x.ppm"
    rt.write_ppm(img, path)
    data = path.read_bytes()
    head = b"P6\n5 3\n255\n"
    assert data[: len(head)] == head and len(data) == len(head) + 45
    with np.errstate(over="ignore"):  # 3e38 * 255 is inf, as in the C float product
        x = img.reshape(-1) * np.float32(255)
    x = np.where(np.isnan(x), np.float32(0), x)  # fmaxf(0, NaN) is 0
    expect = np.minimum(np.float32(255), np.maximum(np.float32(0), x)).astype(np.uint8)
    assert np.array_equal(np.frombuffer(data[len(head):], np.uint8), expect)
    # truncation, not rounding: 0.5 -> 127, 0.99999 -> 254
    assert expect[1] == 127 and expect[5] == 254


def test_csv_rows(gcq, tmp_path):
    from gcq_amd import rt
    res = {"primary_rays": 921600, "total_rays": 1712002, "kernel_ms": 1.25, "total_ms": 3.5, "tiles": 64}
    q = tmp_path / "rt_queue_perf.csv"
    rt.append_csv(q, rt.CSV_QUEUE, res, "AMD Instinct MI355X", "GLFQ", "complex", 1024, 4)
    rt.append_csv(q, rt.CSV_QUEUE, res, "AMD Instinct MI355X", "GWFQ", "cornell", 512, 2)
    lines = q.read_text().splitlines()
    assert lines[0] == "GPU,QUEUE,SCENE,THREADS,BOUNCES,PRIMARY,TOTAL,GEN_MS,TRACE_MS,TOTAL_MS,MRAYS"
    assert len(lines) == 3
    cols = lines[1].split(",")
    assert len(cols) == 11
    assert cols[:7] == ['"AMD Instinct MI355X"', "GLFQ", "complex", "1024", "4", "921600", "1712002"]
    assert float(cols[7]) == 0 and float(cols[8]) == 1.25 and float(cols[9]) == 3.5
    assert abs(float(cols[10]) - 1712002 / 1e6 / 1.25e-3) < 1e-3 * float(cols[10])
    assert lines[2].split(",")[1:3] == ["GWFQ", "cornell"]
    c = tmp_path / "rt_performance.csv"
    rt.append_csv(c, rt.CSV_COMPACT, res, "MI355X", "COMPACT", threads=256)
    lines = c.read_text().splitlines()
    assert lines[0] == "DEVICE,QUEUE,THREADS,TOTAL_RAYS,TOTAL_TIME_MS,TOTAL_MRAYS"
    cols = lines[1].split(",")
    assert cols[:5] == ["MI355X", "COMPACT", "256", "1712002", "1.25"]
