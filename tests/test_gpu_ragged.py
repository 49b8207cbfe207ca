"""Ragged batches (FFTLIB_GPU_raggedCreate / Exec / Tune) on the GPU through
the C-ABI: every transform against the oracle's batched C2C / IFFT of its
own length, graph replays against the eager result, forward-then-inverse
round trips, and the tuned plans."""
import ctypes

import numpy as np
import pytest

import oracle
from tests import fftlib_calls as G

pytestmark = pytest.mark.gpu

REL_TOL = 1e-6


def rng(seed):
    return np.random.Generator(np.random.Philox(seed))


class Ragged:
    def __init__(self, lengths):
        import torch
        self.torch = torch
        self.lib = G.lib()
        self.lengths = [int(n) for n in lengths]
        arr = (ctypes.c_uint32 * len(lengths))(*self.lengths)
        self.h = ctypes.c_void_p()
        s = self.lib.FFTLIB_GPU_raggedCreate(ctypes.byref(self.h), arr, len(lengths))
        assert s == 0, (s, G.F.last_error())
        self.total = sum(self.lengths)
        assert self.lib.FFTLIB_GPU_raggedPoints(self.h) == self.total

    def buffers(self, seed):
        r = rng(seed)
        host = (r.uniform(-1, 1, self.total) + 1j * r.uniform(-1, 1, self.total)).astype(np.complex64)
        dev = lambda a: self.torch.from_numpy(a.view(np.float32).copy()).to("cuda:0")
        x = dev(host)
        y = self.torch.zeros_like(x)
        z = self.torch.zeros_like(x)
        return host, x, y, z

    def run(self, x, y, z, direction):
        self.lib.FFTLIB_setFunctionStyle(G.F.FFTLIB_FUNCTION_GPU)
        s = self.lib.FFTLIB_GPU_raggedExec(self.h, x.data_ptr(), y.data_ptr(), z.data_ptr(), direction)
        assert s == 0, (s, G.F.last_error())
        self.torch.cuda.synchronize()

    def split(self, t):
        a = t.cpu().numpy().view(np.complex64)
        out, o = [], 0
        for n in self.lengths:
            out.append(a[o:o + n])
            o += n
        return out

    def close(self):
        assert self.lib.FFTLIB_GPU_raggedDestroy(self.h) == 0


def check_against_oracle(rg, host, got, inverse):
    o = 0
    for n, y in zip(rg.lengths, got):
        ref = oracle.c2c(host[o:o + n][None, :], n, 1, inverse=inverse)[0]
        assert G.rel_l2(y, ref) <= REL_TOL, n
        o += n


MIXED = [16, 60, 960, 1000, 4096, 16384, 18, 20000, 960, 75, 65536, 3125, 243, 8192, 30720, 96, 960]


@pytest.mark.parametrize("direction", [0, 1])
def test_ragged_one_direction(direction):
    rg = Ragged(MIXED)
    host, x, y, z = rg.buffers(1 + direction)
    rg.run(x, y, z, direction)
    check_against_oracle(rg, host, rg.split(y), inverse=direction == 1)
    before = x.cpu().numpy().view(np.complex64)
    assert np.array_equal(before, host), "input was modified"
    rg.close()


def test_ragged_forward_then_inverse_round_trip():
    rg = Ragged(MIXED)
    host, x, y, z = rg.buffers(7)
    rg.run(x, y, z, 2)
    check_against_oracle(rg, host, rg.split(y), inverse=False)
    back = np.concatenate(rg.split(z))
    assert G.rel_l2(back, host) <= 1e-6
    rg.close()


def test_ragged_graph_replays_match_eager():
    """Replays of the captured graph, with other GPU work and synchronisation
    in between, give the same bits each time."""
    rg = Ragged(MIXED * 3)
    host, x, y, z = rg.buffers(11)
    rg.run(x, y, z, 2)
    first_y, first_z = y.clone(), z.clone()
    for i in range(3):
        y.zero_()
        z.zero_()
        junk = rg.torch.randn(1 << 20, device="cuda:0").sum()  # unrelated eager work
        rg.torch.cuda.synchronize()
        rg.run(x, y, z, 2)
        assert rg.torch.equal(y, first_y), i
        assert rg.torch.equal(z, first_z), i
        del junk
    rg.close()


def test_ragged_buffers_change_recaptures():
    rg = Ragged([960, 1024, 20000])
    host, x, y, z = rg.buffers(3)
    rg.run(x, y, z, 0)
    host2, x2, y2, z2 = rg.buffers(4)
    rg.run(x2, y2, z2, 0)
    check_against_oracle(rg, host2, rg.split(y2), inverse=False)
    rg.run(x, y, z, 1)
    check_against_oracle(rg, host, rg.split(y), inverse=True)
    rg.close()


def test_ragged_tuning_keeps_results(tmp_path):
    rg = Ragged([960] * 40 + [4096] * 8 + [250] * 100 + [40000] * 2)
    host, x, y, z = rg.buffers(5)
    assert rg.lib.FFTLIB_GPU_raggedTune(rg.h, x.data_ptr(), y.data_ptr()) == 0, G.F.last_error()
    rg.run(x, y, z, 0)
    check_against_oracle(rg, host, rg.split(y), inverse=False)
    path = str(tmp_path / "tune.txt").encode()
    assert rg.lib.FFTLIB_GPU_tuneSave(path) == 0
    text = open(path, "rb").read().decode()
    assert text.startswith("fftlib-amd-tune 1\n")
    keys = {int(line.split()[0]) for line in text.strip().splitlines()[1:]}
    # phase 0: octave kernels of 250 and of 40000's columns (200; both in
    # (128, 256]), 960 (512, 1024], 4096 (2048, 4096]; phase 1: 40000's rows
    # (200), whose column tasks are in octave 4 (key + 1000 * (4 + 1))
    assert {100000 + 4, 100000 + 6, 100000 + 8, 100000 + 100 + 4 + 5000} <= keys
    rg.close()
