"""CPU check of the persistent ragged launch's work queue.

One launch of the persistent ragged kernel takes units in the order
``csrc/fftlib_ragged_queue.h`` builds: single groups of plain tasks, and
whole long transforms (column pass, then row pass, in one workgroup).
``ragged_queue_check.cpp`` compiles that header with g++ and checks coverage,
unit shapes and order; no GPU is involved.
"""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "..", "c7x_fftlib_amd", "csrc")


def test_ragged_queue_units(tmp_path):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    exe = str(tmp_path / "ragged_queue_check")
    src = os.path.join(HERE, "ragged_queue_check.cpp")
    build = subprocess.run([gxx, "-std=c++17", "-O1", "-Wall", "-Werror", "-I" + CSRC, src, "-o", exe],
                           capture_output=True, text=True)
    assert build.returncode == 0, build.stderr
    run = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert run.returncode == 0, run.stdout + run.stderr
    assert run.stdout.strip().endswith("ok")
