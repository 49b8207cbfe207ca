"""Ragged batches (FFTLIB_GPU_raggedCreate / Exec / Tune) on the GPU through
the C-ABI, in both plan forms (one persistent launch per direction, and the
graph of per-octave launches): every transform against the oracle's batched
C2C / IFFT of its own length, graph replays against the eager result,
forward-then-inverse round trips, the two forms against each other, and the
tuned plans."""
import ctypes

import numpy as np
import pytest

import oracle
from tests import fftlib_calls as G

pytestmark = pytest.mark.gpu

REL_TOL = 1e-6
FORMS = [1, 0]   # FFTLIB_RAGGED_PERSISTENT, FFTLIB_RAGGED_GRAPH


def rng(seed):
    return np.random.Generator(np.random.Philox(seed))


class Ragged:
    def __init__(self, lengths, form=None):
        import torch
        self.torch = torch
        self.lib = G.lib()
        self.lengths = [int(n) for n in lengths]
        arr = (ctypes.c_uint32 * len(lengths))(*self.lengths)
        self.h = ctypes.c_void_p()
        print(f"[ragged] {len(lengths)} transforms, form {form}", flush=True)
        s = self.lib.FFTLIB_GPU_raggedCreate(ctypes.byref(self.h), arr, len(lengths))
        assert s == 0, (s, G.F.last_error())
        if form is not None:
            assert self.lib.FFTLIB_GPU_raggedSetForm(self.h, form) == 0
            assert self.lib.FFTLIB_GPU_raggedGetForm(self.h) == form
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


@pytest.mark.parametrize("form", FORMS)
@pytest.mark.parametrize("direction", [0, 1])
def test_ragged_one_direction(direction, form):
    rg = Ragged(MIXED, form)
    host, x, y, z = rg.buffers(1 + direction)
    rg.run(x, y, z, direction)
    check_against_oracle(rg, host, rg.split(y), inverse=direction == 1)
    before = x.cpu().numpy().view(np.complex64)
    assert np.array_equal(before, host), "input was modified"
    rg.close()


@pytest.mark.parametrize("form", FORMS)
def test_ragged_forward_then_inverse_round_trip(form):
    rg = Ragged(MIXED, form)
    host, x, y, z = rg.buffers(7)
    rg.run(x, y, z, 2)
    check_against_oracle(rg, host, rg.split(y), inverse=False)
    back = np.concatenate(rg.split(z))
    assert G.rel_l2(back, host) <= 1e-6
    rg.close()


@pytest.mark.parametrize("form", FORMS)
def test_ragged_graph_replays_match_eager(form):
    """Replays of the captured graph, with other GPU work and synchronisation
    in between, give the same bits each time (the persistent form's queue and
    counters are cleared by each launch's last workgroup)."""
    rg = Ragged(MIXED * 3, form)
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


@pytest.mark.parametrize("form", FORMS)
def test_ragged_buffers_change_recaptures(form):
    rg = Ragged([960, 1024, 20000], form)
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
    # the persistent launch's grid and the chosen form
    assert {200000, 200001} <= keys
    assert rg.lib.FFTLIB_GPU_raggedGetForm(rg.h) in (0, 1)
    rg.close()


def test_ragged_forms_agree():
    """The two forms differ only in how the same class-kernel groups are
    scheduled, except that the persistent form also runs the lengths 4097 ..
    16384 as two passes: both within the oracle tolerance of each other on a
    mix with many lengths of every octave."""
    r = rng(23)
    lengths = sorted({int(n) for n in r.choice(
        [16, 48, 125, 243, 256, 960, 1000, 2187, 4096, 4374, 5000, 8192, 9375, 12288, 16384, 16875, 20000, 32768,
         65536], 40)})
    outs = []
    for form in FORMS:
        rg = Ragged(lengths * 2, form)
        host, x, y, z = rg.buffers(29)
        rg.run(x, y, z, 2)
        check_against_oracle(rg, host, rg.split(y), inverse=False)
        outs.append((y.cpu().numpy().view(np.complex64), z.cpu().numpy().view(np.complex64)))
        rg.close()
    assert G.rel_l2(outs[0][0], outs[1][0]) <= REL_TOL
    assert G.rel_l2(outs[0][1], outs[1][1]) <= REL_TOL


def test_ragged_set_form_rejects_unknown():
    rg = Ragged([960, 1024])
    assert rg.lib.FFTLIB_GPU_raggedSetForm(rg.h, 2) == G.F.FFTLIB_ERR_INVALID_VALUE
    assert rg.lib.FFTLIB_GPU_raggedSetForm(rg.h, -1) == G.F.FFTLIB_ERR_INVALID_VALUE
    rg.close()
