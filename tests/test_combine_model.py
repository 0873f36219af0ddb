"""CPU model of the workgroup claim-combining protocol (combine_claim in
gcq_amd/csrc/gcq/prim.hpp): tickets unique and dense, every group member's
call before the leader's Tail read, no deadlock, across 2-16 waves; and the
same for the unbounded group FAA of producers' Tail tickets (group_faa); and a
scripted wrap of the 16-bit group id.  The same model with the Tail read moved
before the group's close, or without the result-word clear before a join
(planted mistakes), must be caught.  CPU only."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "combine_model.cpp")


def _build(tmp_path, name, *defs):
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    exe = str(tmp_path / name)
    subprocess.run(["g++", "-O2", "-std=c++17", "-pthread", *defs, SRC, "-o", exe], check=True)
    return exe


def test_combine_protocol_model(tmp_path):
    exe = _build(tmp_path, "combine_model")
    r = subprocess.run([exe, "45", "3000"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "failures 0" in r.stdout


def test_combine_model_catches_late_close(tmp_path):
    exe = _build(tmp_path, "combine_model_bad", "-DPLANT_LATE_CLOSE")
    r = subprocess.run([exe, "30", "2000"], capture_output=True, text=True, timeout=120)
    assert r.returncode != 0
    assert "a member started after the leader read Tail" in r.stderr


def test_group_faa_model(tmp_path):
    exe = _build(tmp_path, "combine_model_faa")
    r = subprocess.run([exe, "45", "3000", "faa"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "failures 0" in r.stdout


def test_combine_group_id_wrap(tmp_path):
    """A wave that joins group X, leads 65,535 groups alone, then joins group
    X + 65,536 (the same 16-bit id) must wait for its new result."""
    exe = _build(tmp_path, "combine_model_wrap")
    r = subprocess.run([exe, "1", "1", "wrap"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "failures 0" in r.stdout


def test_combine_model_catches_stale_group(tmp_path):
    exe = _build(tmp_path, "combine_model_stale", "-DPLANT_STALE_GROUP")
    r = subprocess.run([exe, "1", "1", "wrap"], capture_output=True, text=True, timeout=120)
    assert r.returncode != 0
    assert "ticket handed out twice" in r.stderr
