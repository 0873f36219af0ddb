"""The renderer against the reference's own output: the PPMs that the
reference's ray tracing harness and G-LFQ queue wrote for its two scenes
(tests/golden/rt_glfq_*.ppm.gz, compiled without FMA contraction, the
arithmetic of this repository's renderer) must equal, byte for byte, the
image of the CPU restatement written by this repository's PPM writer
(rt_final.cpp:689-700) and, on the GPU, the images of the queue renderer
(every queue kind) and of the stream-compaction renderer."""
import gzip
import os

import numpy as np
import pytest

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
SCENES = [(0, "complex"), (1, "cornell")]


def _golden(name):
    with gzip.open(os.path.join(GOLDEN, f"rt_glfq_{name}.ppm.gz")) as f:
        return f.read()


def _ppm_bytes(img, tmp_path):
    from gcq_amd import rt
    p = str(tmp_path / "img.ppm")
    rt.write_ppm(img, p)
    with open(p, "rb") as f:
        return f.read()


def _first_diff(a, b):
    n = min(len(a), len(b))
    d = np.flatnonzero(np.frombuffer(a[:n], np.uint8) != np.frombuffer(b[:n], np.uint8))
    return (int(d[0]) if len(d) else n), len(d), len(a), len(b)


@pytest.mark.parametrize("preset,name", SCENES, ids=[n for _, n in SCENES])
def test_oracle_image_equals_reference_ppm(gcq, orc, preset, name, tmp_path):
    from gcq_amd import rt
    with rt.Scene.preset(preset) as s:
        img, _ = orc.rt_render(*s.arrays(), 1280, 720, max_depth=4)
    mine, ref = _ppm_bytes(img, tmp_path), _golden(name)
    assert mine == ref, f"{name}: first difference at byte {_first_diff(mine, ref)}"


@pytest.mark.gpu
@pytest.mark.parametrize("kind", [0, 1, 2, 3], ids=["gwfq", "glfq", "wfq", "sfq"])
@pytest.mark.parametrize("preset,name", SCENES, ids=[n for _, n in SCENES])
def test_gpu_image_equals_reference_ppm(gcq, preset, name, kind, tmp_path):
    from gcq_amd import rt
    print(f"[rt ref ppm] {name} kind={kind}", flush=True)
    with rt.Scene.preset(preset) as s:
        img, _ = s.render(kind=kind)
    mine, ref = _ppm_bytes(img, tmp_path), _golden(name)
    assert mine == ref, f"{name} kind {kind}: first difference at byte {_first_diff(mine, ref)}"


@pytest.mark.gpu
@pytest.mark.parametrize("preset,name", SCENES, ids=[n for _, n in SCENES])
def test_gpu_compaction_image_equals_reference_ppm(gcq, preset, name, tmp_path):
    from gcq_amd import rt
    with rt.Scene.preset(preset) as s:
        img, _ = s.render_compact(width=1280, height=720, block=1024)
    mine, ref = _ppm_bytes(img, tmp_path), _golden(name)
    assert mine == ref, f"{name} compaction: first difference at byte {_first_diff(mine, ref)}"
