"""Host-side checks of the ragged-batch API that need no GPU: argument
validation before any device work, and the tuning-cache file format."""
import ctypes

import pytest

from tests import fftlib_calls as G


def create(lengths):
    lib = G.lib()
    h = ctypes.c_void_p()
    arr = (ctypes.c_uint32 * max(1, len(lengths)))(*lengths)
    return lib.FFTLIB_GPU_raggedCreate(ctypes.byref(h), arr, len(lengths)), h


@pytest.mark.parametrize("lengths", [[], [960, 17], [8], [960, 1 << 21], [0]])
def test_ragged_create_rejects(lengths):
    s, h = create(lengths)
    assert s == G.F.FFTLIB_ERR_INVALID_DIMENSION
    assert not h.value


def test_ragged_null_arguments():
    lib = G.lib()
    assert lib.FFTLIB_GPU_raggedCreate(None, None, 1) == G.F.FFTLIB_ERR_NULL_POINTER
    assert lib.FFTLIB_GPU_raggedExec(None, None, None, None, 0) == G.F.FFTLIB_ERR_NULL_POINTER
    assert lib.FFTLIB_GPU_raggedDestroy(None) == G.F.FFTLIB_ERR_NULL_POINTER
    assert lib.FFTLIB_GPU_raggedPoints(None) == 0


def test_tune_file_round_trip(tmp_path):
    lib = G.lib()
    src = tmp_path / "in.txt"
    src.write_text("fftlib-amd-tune 1\n960 64 512 12.500\n4096 8 0 3.250\n")
    assert lib.FFTLIB_GPU_tuneLoad(str(src).encode()) == 0
    out = tmp_path / "out.txt"
    assert lib.FFTLIB_GPU_tuneSave(str(out).encode()) == 0
    lines = out.read_text().splitlines()
    assert lines[0] == "fftlib-amd-tune 1"
    assert "960 64 512 12.500" in lines and "4096 8 0 3.250" in lines


@pytest.mark.parametrize("text", ["not-a-tune-file 1\n", "fftlib-amd-tune 2\n", "fftlib-amd-tune 1\n960 x\n"])
def test_tune_file_rejects_bad_input(tmp_path, text):
    lib = G.lib()
    p = tmp_path / "bad.txt"
    p.write_text(text)
    assert lib.FFTLIB_GPU_tuneLoad(str(p).encode()) == G.F.FFTLIB_ERR_INVALID_VALUE


def test_tune_file_missing():
    lib = G.lib()
    assert lib.FFTLIB_GPU_tuneLoad(b"/nonexistent/dir/tune.txt") == G.F.FFTLIB_ERR_FAILURE
