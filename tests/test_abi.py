"""The C ABI library loads and exports every function include/*.h declares
(no compute calls: runs without a GPU)."""
import ctypes
import glob
import os
import re

from conftest import ROOT


def _declared():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"\b(gcq_[a-z0-9_]+)\s*\(", text):
            names.add(m.group(1))
    return names


def test_header_declares_entry_points():
    names = _declared()
    for must in ["gcq_queue_create", "gcq_queue_destroy", "gcq_tb_run", "gcq_hist_run", "gcq_faa_peak",
                 "gcq_fifo_check"]:
        assert must in names


def test_library_exports_every_declared_symbol(gcq):
    lib = ctypes.CDLL(gcq.LIB_PATH)
    missing = [n for n in sorted(_declared()) if not hasattr(lib, n)]
    assert not missing, missing


def test_errors_cross_the_abi_as_codes(gcq):
    lib = gcq.load()
    assert lib.gcq_strerror(-1) == b"invalid argument"
    cfg = gcq.default_config(gcq.GWFQ)
    cfg.capacity = 1000  # not a power of two
    h = ctypes.c_void_p()
    rc = lib.gcq_queue_create(ctypes.byref(cfg), ctypes.byref(h))
    assert rc < 0  # GCQ_EINVAL here, GCQ_ENODEV on a machine without a GPU
    assert h.value is None


def test_reference_defaults(gcq):
    cfg = gcq.default_config(gcq.GWFQ)
    assert (cfg.help_delay, cfg.enq_patience, cfg.deq_patience) == (64, 128, 256)  # gwf_queue.hpp:296-306
    assert gcq.default_config(gcq.WFQ).enq_patience == 16  # WF_PATIENCE wfqueue_hip_opt.hpp:29-31
