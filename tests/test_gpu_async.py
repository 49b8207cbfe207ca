"""Asynchronous entry points on the GPU through the C-ABI.

FFTLIB_GPU_kernelAsync runs a plan on a caller's stream; the host pipeline
runs host-memory calls through pinned staging on its own streams. Both use
the same kernels as the synchronous _kernel, so their outputs must equal the
synchronous outputs bit for bit; every case is also checked against the
oracle (relative L2 <= 1e-6 for fp32, bit-exact for int16), including those
no other test covers (the single-channel C2C of 300000 points, C2R inMode 1
at N = 480).
"""
import ctypes

import numpy as np
import pytest

import c7x_fftlib_amd as F
from tests import fftlib_calls as G

pytestmark = pytest.mark.gpu

F32, I16, U32 = F.FFTLIB_FLOAT32, F.FFTLIB_INT16, F.FFTLIB_UINT32


def _bp(dt, n):
    return ctypes.byref(F.BufParams1D(dt, n))


def plan(kind, n, b, mode=0, nshift=0):
    """pBlock planned by the _init of `kind` (buffers are not read by _init)."""
    L = G.lib()
    assert L.FFTLIB_setFunctionStyle(F.FFTLIB_FUNCTION_GPU) == 0
    z = 0
    if kind in ("c2c", "ic2c"):
        name = {"c2c": "FFTLIB_fft1dBatched_i32fc_c32fc_o32fc", "ic2c": "FFTLIB_ifft1dBatched_i32fc_c32fc_o32fc"}[kind]
        pb = G.pblock(name)
        d = _bp(F32, 2 * n * b)
        st = getattr(L, name + "_init")(z, d, z, _bp(F32, 2 * n), z, d, n, b, pb)
    elif kind == "c2c1":
        pb = G.pblock("FFTLIB_fft1d_i32fc_c32fc_o32fc")
        d = _bp(F32, 2 * n)
        st = L.FFTLIB_fft1d_i32fc_c32fc_o32fc_init(z, d, z, d, z, d, pb)
    elif kind == "i16":
        pb = G.pblock("FFTLIB_fft1dBatched_i16sc_c16sc_o16sc")
        d = _bp(I16, 2 * n * b)
        st = L.FFTLIB_fft1dBatched_i16sc_c16sc_o16sc_init(z, d, z, _bp(I16, 2 * n), z, d, z, _bp(U32, nshift), n, b,
                                                          pb)
    elif kind == "r2c":
        pb = G.pblock("FFTLIB_fft1dBatched_i32f_c32fc_o32fc")
        d = _bp(F32, n * b)
        st = L.FFTLIB_fft1dBatched_i32f_c32fc_o32fc_init(z, d, z, d, z, d, z, d, z, d, n, b, mode, pb)
    elif kind == "c2r":
        pb = G.pblock("FFTLIB_ifft1dBatched_i32fc_c32fc_o32f")
        d = _bp(F32, n * b)
        st = L.FFTLIB_ifft1dBatched_i32fc_c32fc_o32f_init(z, d, z, d, z, d, z, d, z, d, n, b, mode, pb)
    else:
        raise ValueError(kind)
    assert st == 0, (kind, n, st)
    return pb


import oracle

REL_TOL = 1e-6


def assert_oracle(kind, got, ref):
    """got / ref: the product's and the oracle's outputs of one case."""
    if kind == "i16":
        assert np.array_equal(np.asarray(got).reshape(-1), np.asarray(ref).reshape(-1))
    else:
        assert G.rel_l2(np.asarray(got).reshape(-1), np.asarray(ref).reshape(-1)) <= REL_TOL


def oracle_ref(kind, x, n, b, mode, aux):
    """The oracle's output for the inputs `case` builds (same flattening)."""
    if kind in ("c2c", "ic2c", "c2c1"):
        xc = x.view(np.complex64).reshape(b, n)
        return oracle.c2c(xc, n, b, inverse=kind == "ic2c", single=kind == "c2c1").view(np.float32).reshape(-1)
    if kind == "i16":
        return oracle.i16(x.reshape(b, n, 2), n, b, aux).reshape(-1)
    if kind == "r2c":
        return oracle.r2c(x.reshape(b, n), n, b, outmode=mode).view(np.float32).reshape(-1)
    xc = x.view(np.complex64).reshape(b, -1)
    return oracle.c2r(xc, n, b, inmode=mode).reshape(-1)


def case(kind, n, b, mode=0, seed=0):
    """Host inputs (x, w, aux), output element count/dtype, and the
    synchronous GPU result of the same call (checked against the oracle)."""
    got = _case(kind, n, b, mode, seed)
    x, _, aux, _, want, s = got
    if s == 0:
        assert_oracle(kind, np.asarray(want).view(np.int16 if kind == "i16" else np.float32),
                      oracle_ref(kind, x, n, b, mode, aux))
    return got


def _case(kind, n, b, mode=0, seed=0):
    r = np.random.Generator(np.random.Philox(seed))
    if kind in ("c2c", "ic2c", "c2c1"):
        x = (r.uniform(-1, 1, (b, n)) + 1j * r.uniform(-1, 1, (b, n))).astype(np.complex64)
        s, want = G.c2c(x, n, b, inverse=kind == "ic2c", single=kind == "c2c1")
        return x.view(np.float32).reshape(-1), G.tw_c2c(n), None, (2 * n * b, np.float32), want, s
    if kind == "i16":
        x = r.integers(-2048, 2048, (b, n, 2)).astype(np.int16)
        sh = np.asarray([0, 1, 1, 1, 1, 1, 1, 1, 1, 1][:max(1, int(np.log2(n)) // 2)], np.uint32)
        s, want = G.i16(x, n, b, sh)
        return x.reshape(-1), G.tw_i16(n), sh, (2 * n * b, np.int16), want, s
    if kind == "r2c":
        x = r.standard_normal((b, n)).astype(np.float32)
        s, want = G.r2c(x, n, b, outmode=mode)
        nout = n if mode else n // 2 + 1
        return x.reshape(-1), G.tw_r2c(n), G.sf_r2c(n), (2 * nout * b, np.float32), want, s
    if kind == "c2r":
        nin = n if mode else n // 2 + 1
        spec = np.fft.rfft(r.standard_normal((b, n)), axis=1)
        if mode:
            full = np.fft.fft(r.standard_normal((b, n)), axis=1)
            x = full.astype(np.complex64)
        else:
            x = spec.astype(np.complex64)
        assert x.shape == (b, nin)
        s, want = G.c2r(x, n, b, inmode=mode)
        return x.view(np.float32).reshape(-1), G.tw_c2r(n), G.sf_c2r(n), (n * b, np.float32), want, s
    raise ValueError(kind)


CASES = [("c2c", 960, 64, 0), ("ic2c", 1000, 33, 0), ("c2c1", 1024, 1, 0), ("c2c", 1 << 20, 2, 0),
         ("c2c", 300000, 1, 0), ("i16", 256, 40, 0), ("i16", 2048, 3, 0), ("r2c", 4096, 17, 0),
         ("r2c", 960, 5, 1), ("c2r", 4096, 9, 0), ("c2r", 480, 4, 1)]


@pytest.mark.parametrize("kind,n,b,mode", CASES)
def test_kernel_async_on_caller_stream_equals_synchronous(kind, n, b, mode):
    import torch
    x, w, aux, (ycount, ydt), want, s = case(kind, n, b, mode, seed=n + b)
    assert s == 0, G.F.last_error()
    pb = plan(kind, n, b, mode, nshift=0 if aux is None or kind != "i16" else aux.size)
    stream = torch.cuda.Stream()
    dx = torch.from_numpy(x.copy()).cuda()
    dw = torch.from_numpy(w.copy()).cuda()
    da = torch.from_numpy(aux.view(np.int32) if aux is not None and aux.dtype == np.uint32 else aux).cuda() \
        if aux is not None else None
    tdt = torch.int16 if ydt == np.int16 else torch.float32
    dy = torch.full((ycount,), -7, dtype=tdt, device="cuda")
    torch.cuda.synchronize()
    st = G.lib().FFTLIB_GPU_kernelAsync(pb, dx.data_ptr(), dw.data_ptr(), da.data_ptr() if da is not None else None,
                                        dy.data_ptr(), ctypes.c_void_p(stream.cuda_stream))
    assert st == 0, G.F.last_error()
    stream.synchronize()
    got = dy.cpu().numpy()
    assert np.array_equal(got.view(np.uint16 if ydt == np.int16 else np.uint32),
                          np.ascontiguousarray(want).reshape(-1).view(np.uint16 if ydt == np.int16 else np.uint32))
    assert np.array_equal(dx.cpu().numpy(), x)                          # input untouched


def test_kernel_async_long_transforms_on_two_streams_overlap():
    """Two long (two-pass, scratch-using) transforms enqueued on two streams
    without synchronisation: each stream has its own scratch."""
    import torch
    n, b = 1 << 20, 2
    xa, w, _, _, want_a, sa = case("c2c", n, b, seed=11)
    xb, _, _, _, want_b, sb = case("c2c", n, b, seed=12)
    assert sa == 0 and sb == 0
    pb = plan("c2c", n, b)
    s1, s2 = torch.cuda.Stream(), torch.cuda.Stream()
    dw = torch.from_numpy(w).cuda()
    dxa, dxb = torch.from_numpy(xa).cuda(), torch.from_numpy(xb).cuda()
    dya = torch.empty(2 * n * b, dtype=torch.float32, device="cuda")
    dyb = torch.empty_like(dya)
    torch.cuda.synchronize()
    L = G.lib()
    for _ in range(3):
        assert L.FFTLIB_GPU_kernelAsync(pb, dxa.data_ptr(), dw.data_ptr(), None, dya.data_ptr(),
                                        ctypes.c_void_p(s1.cuda_stream)) == 0
        assert L.FFTLIB_GPU_kernelAsync(pb, dxb.data_ptr(), dw.data_ptr(), None, dyb.data_ptr(),
                                        ctypes.c_void_p(s2.cuda_stream)) == 0
    torch.cuda.synchronize()
    assert np.array_equal(dya.cpu().numpy().view(np.uint32), want_a.reshape(-1).view(np.complex64).view(np.uint32))
    assert np.array_equal(dyb.cpu().numpy().view(np.uint32), want_b.reshape(-1).view(np.complex64).view(np.uint32))


@pytest.mark.parametrize("depth", [1, 2, 3])
def test_host_pipeline_equals_synchronous(depth):
    """Host-memory calls of several kernels and sizes, more calls than slots
    (slots are reused), outputs delivered at reuse or wait."""
    L = G.lib()
    pipe = ctypes.c_void_p()
    assert L.FFTLIB_GPU_pipelineCreate(ctypes.byref(pipe), depth) == 0, G.F.last_error()
    try:
        jobs = []
        for i, (kind, n, b, mode) in enumerate(CASES + CASES[:3]):
            x, w, aux, (ycount, ydt), want, s = case(kind, n, b, mode, seed=100 + i)
            assert s == 0
            pb = plan(kind, n, b, mode, nshift=aux.size if kind == "i16" else 0)
            y = np.full(ycount, -7, ydt)
            keep = (x, w, aux, pb)
            st = L.FFTLIB_GPU_pipelineSubmit(pipe, pb, x.ctypes.data, w.ctypes.data,
                                             aux.ctypes.data if aux is not None else None, y.ctypes.data)
            assert st == 0, (kind, n, G.F.last_error())
            jobs.append((kind, n, y, want, keep))
        assert L.FFTLIB_GPU_pipelineWait(pipe) == 0
        for kind, n, y, want, _ in jobs:
            v = np.uint16 if y.dtype == np.int16 else np.uint32
            assert np.array_equal(y.view(v), np.ascontiguousarray(want).reshape(-1).view(y.dtype).view(v)), (kind, n)
    finally:
        assert L.FFTLIB_GPU_pipelineDestroy(pipe) == 0


def test_host_pipeline_same_output_twice_keeps_the_later_call():
    L = G.lib()
    pipe = ctypes.c_void_p()
    assert L.FFTLIB_GPU_pipelineCreate(ctypes.byref(pipe), 2) == 0
    try:
        n, b = 960, 8
        xa, w, _, (ycount, _), _, _ = case("c2c", n, b, seed=1)
        xb, _, _, _, want_b, _ = case("c2c", n, b, seed=2)
        pb = plan("c2c", n, b)
        y = np.zeros(ycount, np.float32)
        assert L.FFTLIB_GPU_pipelineSubmit(pipe, pb, xa.ctypes.data, w.ctypes.data, None, y.ctypes.data) == 0
        assert L.FFTLIB_GPU_pipelineSubmit(pipe, pb, xb.ctypes.data, w.ctypes.data, None, y.ctypes.data) == 0
        assert L.FFTLIB_GPU_pipelineWait(pipe) == 0
        assert np.array_equal(y.view(np.uint32), want_b.reshape(-1).view(np.complex64).view(np.uint32))
        assert L.FFTLIB_GPU_pipelineWait(pipe) == 0                      # nothing pending: no-op
    finally:
        assert L.FFTLIB_GPU_pipelineDestroy(pipe) == 0


def _pinned(a):
    """Copy of numpy array `a` in pinned host memory (torch pin_memory)."""
    import torch
    t = torch.from_numpy(np.ascontiguousarray(a).view(np.int32) if a.dtype == np.uint32 else
                         np.ascontiguousarray(a)).pin_memory()
    return t, t.numpy()


@pytest.mark.parametrize("kind,n,b,mode", [("c2c", 960, 64, 0), ("c2c", 1 << 20, 2, 0), ("i16", 256, 40, 0),
                                           ("r2c", 4096, 17, 0), ("c2r", 62500, 2, 1)])
def test_host_pipeline_pinned_buffers(kind, n, b, mode):
    """Pinned inputs and outputs are copied by DMA straight from and to the
    caller's memory; the results equal the synchronous ones."""
    import torch
    L = G.lib()
    pipe = ctypes.c_void_p()
    assert L.FFTLIB_GPU_pipelineCreate(ctypes.byref(pipe), 2) == 0
    try:
        keep, jobs = [], []
        for i in range(3):
            x, w, aux, (ycount, ydt), want, s = case(kind, n, b, mode, seed=300 + i)
            assert s == 0
            pb = plan(kind, n, b, mode, nshift=aux.size if kind == "i16" else 0)
            tx, px = _pinned(x)
            tw, pw = _pinned(w)
            ta, pa = _pinned(aux) if aux is not None else (None, None)
            ty = torch.full((ycount,), -7, dtype=torch.int16 if ydt == np.int16 else torch.float32).pin_memory()
            keep.append((tx, tw, ta, ty, pb))
            st = L.FFTLIB_GPU_pipelineSubmit(pipe, pb, tx.data_ptr(), tw.data_ptr(),
                                             ta.data_ptr() if ta is not None else None, ty.data_ptr())
            assert st == 0, G.F.last_error()
            jobs.append((ty, want))
        assert L.FFTLIB_GPU_pipelineWait(pipe) == 0
        for ty, want in jobs:
            y = ty.numpy()
            v = np.uint16 if y.dtype == np.int16 else np.uint32
            assert np.array_equal(y.view(v), np.ascontiguousarray(want).reshape(-1).view(y.dtype).view(v))
    finally:
        assert L.FFTLIB_GPU_pipelineDestroy(pipe) == 0


def test_host_pipeline_pinned_same_output_keeps_the_later_call():
    """Outputs copied by DMA on different slot streams still land in
    submission order."""
    import torch
    L = G.lib()
    pipe = ctypes.c_void_p()
    assert L.FFTLIB_GPU_pipelineCreate(ctypes.byref(pipe), 3) == 0
    try:
        n, b = 1 << 20, 2                   # the long first call finishes last without ordering
        xa, w, _, (ycount, _), _, _ = case("c2c", n, b, seed=7)
        pb_long = plan("c2c", n, b)
        ty = torch.zeros(ycount, dtype=torch.float32).pin_memory()
        tx, _ = _pinned(xa)
        tw, _ = _pinned(w)
        assert L.FFTLIB_GPU_pipelineSubmit(pipe, pb_long, tx.data_ptr(), tw.data_ptr(), None, ty.data_ptr()) == 0
        n2, b2 = 1024, 2048                 # same output bytes, much less work
        xb, w2, _, (ycount2, _), want_b, _ = case("c2c", n2, b2, seed=8)
        assert ycount2 == ycount
        pb_short = plan("c2c", n2, b2)
        tx2, _ = _pinned(xb)
        tw2, _ = _pinned(w2)
        assert L.FFTLIB_GPU_pipelineSubmit(pipe, pb_short, tx2.data_ptr(), tw2.data_ptr(), None, ty.data_ptr()) == 0
        assert L.FFTLIB_GPU_pipelineWait(pipe) == 0
        assert np.array_equal(ty.numpy().view(np.uint32), want_b.reshape(-1).view(np.complex64).view(np.uint32))
    finally:
        assert L.FFTLIB_GPU_pipelineDestroy(pipe) == 0
