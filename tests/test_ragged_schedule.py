"""CPU check of the ragged plan's capture schedule and event pool.

The library captures a ragged call by walking the steps of
``csrc/fftlib_ragged_schedule.h`` (fork, class launches, join per phase and
direction) and sizes its event pool with ``events_needed`` of the
two-direction schedule. ``ragged_schedule_check.cpp`` compiles that header
with g++ and checks the event accounting and the launch ordering it implies
(happens-before over event record / stream wait) for many plans; no GPU is
involved.
"""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "..", "c7x_fftlib_amd", "csrc")


def test_ragged_schedule_events_and_order(tmp_path):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    exe = str(tmp_path / "ragged_schedule_check")
    src = os.path.join(HERE, "ragged_schedule_check.cpp")
    build = subprocess.run([gxx, "-std=c++17", "-O1", "-Wall", "-Werror", "-I" + CSRC, src, "-o", exe],
                           capture_output=True, text=True)
    assert build.returncode == 0, build.stderr
    run = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert run.returncode == 0, run.stdout + run.stderr
    assert run.stdout.strip().endswith("ok")
