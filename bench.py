"""Benchmark of the FFTLIB batched FFT hot path on one MI355X.

Default workload (the configuration the headline metric is quoted on):
  FFTLIB_fft1dBatched_i32fc_c32fc_o32fc, N = 960 = 2^6*3*5, 65536 channels,
  interleaved complex64 input i.i.d. uniform[-1, 1) from a Philox seed,
  twiddles from the library's generator, buffers resident in HBM.

A step is one call of the C-ABI _kernel entry point (FFTLIB_FUNCTION_GPU,
device pointers, asynchronous mode) over the whole batch. W warm-up steps,
then untimed steps until --min-warmup-s (0.3 s) of wall time has passed (the
GPU clocks ramp up over tens of ms of sustained work), are followed by
exactly K timed steps bracketed by device synchronisation; HIP
events recorded on the stream the library launches on give the average launch
duration. One JSON line is printed:
  value     = nominal GFLOP/s = 5 N log2(N) B / t
  roofline  = algorithmic HBM bytes (one read + one write of 8 B per point)
              / t against the 8 TB/s HBM3E peak, plus measured traffic when a
              rocprofv3 counter summary for the workload is committed
  cpu_baseline = the oracle (C restatement of the reference _cn kernel) timed
              on this host on a bounded sample of the same workload.

Other workloads: --workload i16_256 (int16, N=256, 8192 channels),
r2c_4096 (R2C then C2R, N=4096, 16384 channels), ragged_10000 (10000 C2C
transforms of lengths drawn log-uniformly from [64, 65536] and snapped to
2/3/5-smooth lengths, forward then inverse in one graph launch),
copy_503m (plain device copy of the N=960 workload's bytes, for the HBM
counter calibration), c2c_960_host (the N=960 workload with input and output
in pinned host memory, streamed through the library's host pipeline with two
slots: the PCIe-inclusive rate; the synchronous host-pointer style is timed
beside it).
"""
import argparse
import ctypes
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

RAGGED_FORMS = {"graph": 0, "persistent": 1}   # FFTLIB_RAGGED_GRAPH, FFTLIB_RAGGED_PERSISTENT
RAGGED_KERNELS = {
    "persistent": "ragged_class_kernel (one persistent launch per direction: lengths <= 4096 whole, each "
                  "longer transform a column then a row pass in one workgroup)",
    "graph": "ragged graph (class kernels: lengths <= 16384 in one pass, longer lengths in a column and a row pass)",
}
METRIC = "Batched C2C FFT GFLOP/s & %HBM peak; split-BF16 MFMA vs FP32 at ≤2e-6 err"
HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="c2c_960", choices=["c2c_960", "i16_256", "r2c_4096", "ragged_10000", "copy_503m",
                             "c2c_960_host"])
    ap.add_argument("--no-tune", action="store_true", help="ragged_10000: skip the autotuning pass")
    ap.add_argument("--ragged-form", default="auto", choices=["auto", "persistent", "graph"],
                    help="ragged_10000: plan form (auto: the autotuner's choice; without tuning the graph form)")
    ap.add_argument("--variant", default="fp32", choices=["fp32", "bf16x3", "fp16x2", "bf16x2"],
                    help="C2C arithmetic: fp32 mixed radix, or MFMA DFT-as-GEMM with split bf16/fp16 operands")
    ap.add_argument("--generic", action="store_true",
                    help="run the generic (runtime stage table) kernels instead of the length-specialised ones")
    ap.add_argument("--min-warmup-s", type=float, default=0.3,
                    help="after the W warmup steps, keep stepping (untimed) until this much wall time has passed: "
                         "the GPU clocks ramp up over tens of ms of sustained work after the idle host-side setup")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-variants", action="store_true",
                    help="c2c_960 fp32: skip timing the split-precision MFMA variants after the headline")
    ap.add_argument("--cpu-seconds", type=float, default=2.0, help="target seconds per CPU baseline run")
    ap.add_argument("--sweep-mfma", action="store_true",
                    help="instead of the workload: fp32 against the three MFMA variants at every MFMA length, "
                         "2^25 points per call, one JSON line")
    return ap.parse_args()


def philox(seed):
    return np.random.Generator(np.random.Philox(seed))


class Workload:
    """Holds device buffers and a launch closure for one workload."""

    def __init__(self, args, F, lib, torch):
        self.args, self.F, self.lib, self.torch = args, F, lib, torch
        getattr(self, "_setup_" + args.workload)()

    # -- batched fp32 C2C, N = 960, B = 65536 --------------------------------
    def _setup_c2c_960(self):
        F, lib, torch = self.F, self.lib, self.torch
        self.n, self.b = 960, 65536
        n, b = self.n, self.b
        r = philox(960)
        host = np.empty((b, n), np.complex64)
        chunk = 8192
        for i in range(0, b, chunk):
            host[i:i + chunk].real = r.uniform(-1, 1, (chunk, n))
            host[i:i + chunk].imag = r.uniform(-1, 1, (chunk, n))
        self.host_x = host
        w = np.zeros(2 * n, np.float32)
        lib.FFTLIB_tw_gen_c2c(w.ctypes.data, n)
        self.host_w = w
        self.x = torch.from_numpy(host.view(np.float32).reshape(-1)).to("cuda:0")
        self.w = torch.from_numpy(w).to("cuda:0")
        self.y = torch.empty_like(self.x)
        self.bp_x = F.BufParams1D(F.FFTLIB_FLOAT32, 2 * n * b)
        self.bp_w = F.BufParams1D(F.FFTLIB_FLOAT32, 2 * n)
        self.pblock = ctypes.create_string_buffer(F.PBLOCK_SIZE["FFTLIB_fft1dBatched_i32fc_c32fc_o32fc"])
        st = lib.FFTLIB_fft1dBatched_i32fc_c32fc_o32fc_init(self.x.data_ptr(), ctypes.byref(self.bp_x),
                                                             self.w.data_ptr(), ctypes.byref(self.bp_w),
                                                             self.y.data_ptr(), ctypes.byref(self.bp_x), n, b,
                                                             self.pblock)
        assert st == 0, st
        self.flops = 5.0 * n * math.log2(n) * b
        self.bytes = 16.0 * n * b
        self.dtype = "f32"
        self.unit = "GFLOP/s"
        self.config = {"workload": "FFTLIB_fft1dBatched_i32fc_c32fc_o32fc N=960 B=65536 (fp32 C2C)",
                       "numPoints": n, "numChannels": b, "variant": self.args.variant,
                       "input": "complex64 uniform[-1,1) Philox seed 960"}
        if self.args.variant != "fp32":
            self.kernel_name = "c2c_mfma_kernel<%s, 15, 4>" % self.args.variant
        elif self.args.generic:
            self.kernel_name = "c2c_lds_kernel<0, true>"
        else:
            self.kernel_name = "fixed_kernel<960, C2C>"
        xp, wp, yp = self.x.data_ptr(), self.w.data_ptr(), self.y.data_ptr()
        bx, bw, pb = ctypes.byref(self.bp_x), ctypes.byref(self.bp_w), self.pblock
        fn = lib.FFTLIB_fft1dBatched_i32fc_c32fc_o32fc_kernel

        def step():
            s = fn(xp, bx, wp, bw, yp, bx, n, b, pb)
            if s != 0:
                raise RuntimeError(f"kernel status {s}: {F.last_error()}")
        self.step = step

    def check_c2c_960(self):
        """Relative L2 of 256 channels against the FP64 DFT and the oracle."""
        import oracle
        idx = np.arange(0, self.b, self.b // 256)
        y = self.y.view(-1, 2 * self.n)[idx].cpu().numpy().view(np.complex64)
        x = self.host_x[idx]
        ref64 = np.fft.fft(x.astype(np.complex128), axis=1)
        err64 = float(np.linalg.norm(y - ref64) / np.linalg.norm(ref64))
        refo = oracle.c2c(x, self.n, len(idx))
        erro = float(np.linalg.norm(y - refo) / np.linalg.norm(refo))
        return {"rel_l2_vs_fp64": err64, "rel_l2_vs_oracle": erro, "checked_channels": int(len(idx))}

    def cpu_c2c_960(self, seconds):
        import oracle
        n = self.n
        w = oracle.tw_c2c(n)
        probe = 64
        xs = self.host_x[:probe].view(np.float32).reshape(-1).copy()
        ys = np.zeros_like(xs)
        t0 = time.perf_counter()
        oracle.raw_c2c_batched(xs.copy(), w, ys, n, probe)
        per = (time.perf_counter() - t0) / probe
        cnt = int(max(64, min(self.b, seconds / max(per, 1e-9))))
        src = self.host_x[:cnt].view(np.float32).reshape(-1)
        times = []
        for _ in range(5):
            xin = src.copy()  # the natural-C kernel clobbers its input
            y = np.zeros_like(xin)
            t0 = time.perf_counter()
            oracle.raw_c2c_batched(xin, w, y, n, cnt)
            times.append(time.perf_counter() - t0)
        t = float(np.median(times))
        flops = 5.0 * n * math.log2(n) * cnt
        res = {"value": flops / t / 1e9, "unit": "GFLOP/s", "cores": 1, "kind": "port",
               "sample": f"oracle FFTLIB_fft1dBatched_i32fc_c32fc_o32fc_cn, N=960, {cnt} of 65536 channels, "
                         f"median of 5 runs ({t:.3f} s each), input re-copied before each run"}
        threads, cores_note = usable_cores()
        res["cpu_model"] = cpu_model()
        if threads > 1:
            tt = []
            for _ in range(3):
                xin = src.copy()
                y = np.zeros_like(xin)
                t0 = time.perf_counter()
                oracle.raw_threads_c2c(xin, w, y, n, cnt, False, threads)
                tt.append(time.perf_counter() - t0)
            res["all_cores"] = {"value": flops / float(np.median(tt)) / 1e9, "cores": threads,
                                "cores_note": cores_note,
                                "sample": "same sample, channels split over POSIX threads, median of 3"}
        return res

    # -- the N=960 workload with host-resident buffers (PCIe-inclusive) --------
    def _setup_c2c_960_host(self):
        F, lib, torch = self.F, self.lib, self.torch
        self._setup_c2c_960()
        n, b = self.n, self.b
        # pinned host copies of the input and output; the device buffers of
        # the device workload are released
        self.hx = torch.from_numpy(self.host_x.view(np.float32).reshape(-1)).pin_memory()
        self.hw = torch.from_numpy(self.host_w).pin_memory()
        self.hy = torch.empty(2 * n * b, dtype=torch.float32).pin_memory()
        self.x = self.w = None
        self.y = None
        torch.cuda.empty_cache()
        self.pipe = ctypes.c_void_p()
        assert lib.FFTLIB_GPU_pipelineCreate(ctypes.byref(self.pipe), 2) == 0, F.last_error()
        self.kernel_name = "fixed_kernel<960, C2C>"
        self.wall_timing = True
        self.config = dict(self.config, workload="FFTLIB_fft1dBatched_i32fc_c32fc_o32fc N=960 B=65536 (fp32 C2C), "
                                                 "input and output in pinned host memory, host pipeline depth 2",
                           pcie="PCIe-inclusive: 8 B per point each way")
        xp, wp, yp, pb = self.hx.data_ptr(), self.hw.data_ptr(), self.hy.data_ptr(), self.pblock

        def step():
            st = lib.FFTLIB_GPU_pipelineSubmit(self.pipe, pb, xp, wp, None, yp)
            if st != 0:
                raise RuntimeError(f"pipeline status {st}: {F.last_error()}")
        self.step = step

        def finish():
            st = lib.FFTLIB_GPU_pipelineWait(self.pipe)
            if st != 0:
                raise RuntimeError(f"pipeline wait status {st}: {F.last_error()}")
        self.finish = finish

    def extra_c2c_960_host(self, steps):
        """The same transform through the synchronous host-pointer style
        (FFTLIB_FUNCTION_OPTIMIZED: copy in, transform, copy out, wait)."""
        F, lib = self.F, self.lib
        n, b = self.n, self.b
        pb = ctypes.create_string_buffer(F.PBLOCK_SIZE["FFTLIB_fft1dBatched_i32fc_c32fc_o32fc"])
        assert lib.FFTLIB_setFunctionStyle(F.FFTLIB_FUNCTION_OPTIMIZED) == 0
        bx, bw = ctypes.byref(self.bp_x), ctypes.byref(self.bp_w)
        xp, wp, yp = self.hx.data_ptr(), self.hw.data_ptr(), self.hy.data_ptr()
        assert lib.FFTLIB_fft1dBatched_i32fc_c32fc_o32fc_init(xp, bx, wp, bw, yp, bx, n, b, pb) == 0
        fn = lib.FFTLIB_fft1dBatched_i32fc_c32fc_o32fc_kernel
        assert fn(xp, bx, wp, bw, yp, bx, n, b, pb) == 0, F.last_error()
        t0 = time.perf_counter()
        for _ in range(steps):
            assert fn(xp, bx, wp, bw, yp, bx, n, b, pb) == 0, F.last_error()
        t = (time.perf_counter() - t0) / steps
        assert lib.FFTLIB_setFunctionStyle(F.FFTLIB_FUNCTION_GPU) == 0
        return {"host_sync_style": {"ms_per_step": round(t * 1e3, 3),
                                    "gflops": round(self.flops / t / 1e9, 1),
                                    "host_gb_s_each_way": round(8.0 * n * b / t / 1e9, 2)}}

    def check_c2c_960_host(self):
        """Relative L2 of 256 channels of the pinned host output against the
        FP64 DFT and the oracle."""
        import oracle
        idx = np.arange(0, self.b, self.b // 256)
        y = self.hy.numpy().reshape(self.b, 2 * self.n)[idx].copy().view(np.complex64)
        x = self.host_x[idx]
        ref64 = np.fft.fft(x.astype(np.complex128), axis=1)
        refo = oracle.c2c(x, self.n, len(idx))
        return {"rel_l2_vs_fp64": float(np.linalg.norm(y - ref64) / np.linalg.norm(ref64)),
                "rel_l2_vs_oracle": float(np.linalg.norm(y - refo) / np.linalg.norm(refo)),
                "checked_channels": int(len(idx))}

    def cpu_c2c_960_host(self, seconds):
        return self.cpu_c2c_960(seconds)

    # -- plain device copy of the C2C workload's input size (counter calibration)
    # -- ragged batch: 10000 C2C transforms, forward then inverse -----------
    @staticmethod
    def ragged_lengths(count, seed):
        """Lengths log-uniform on [64, 65536], each snapped to the 2/3/5-smooth
        length nearest in log distance (ties to the smaller)."""
        smooth = []
        for a in range(17):
            for b in range(11):
                for c in range(7):
                    v = (2 ** a) * (3 ** b) * (5 ** c)
                    if 16 <= v <= 65536:
                        smooth.append(v)
        smooth = np.array(sorted(smooth), np.float64)
        logs = np.log(smooth)
        r = philox(seed)
        want = r.uniform(np.log(64.0), np.log(65536.0), count)
        idx = np.abs(logs[None, :] - want[:, None]).argmin(axis=1)
        return smooth[idx].astype(np.int64)

    def _setup_ragged_10000(self):
        F, lib, torch = self.F, self.lib, self.torch
        lengths = self.ragged_lengths(10000, 10000)
        self.lengths = lengths
        total = int(lengths.sum())
        self.n, self.b = 0, len(lengths)
        arr = (ctypes.c_uint32 * len(lengths))(*[int(v) for v in lengths])
        self.h = ctypes.c_void_p()
        st = lib.FFTLIB_GPU_raggedCreate(ctypes.byref(self.h), arr, len(lengths))
        if st != 0:
            raise RuntimeError(f"raggedCreate status {st}: {F.last_error()}")
        if self.args.ragged_form != "auto":
            assert lib.FFTLIB_GPU_raggedSetForm(self.h, RAGGED_FORMS[self.args.ragged_form]) == 0
        r = philox(10001)
        host = np.empty(total, np.complex64)
        chunk = 1 << 24
        for i in range(0, total, chunk):
            m = min(chunk, total - i)
            host[i:i + m].real = r.uniform(-1, 1, m)
            host[i:i + m].imag = r.uniform(-1, 1, m)
        self.host_x = host
        self.x = torch.from_numpy(host.view(np.float32)).to("cuda:0")
        self.y = torch.empty_like(self.x)
        self.z = torch.empty_like(self.x)
        if not self.args.no_tune:
            st = lib.FFTLIB_GPU_raggedTune(self.h, self.x.data_ptr(), self.y.data_ptr())
            if st != 0:
                raise RuntimeError(f"raggedTune status {st}: {F.last_error()}")
        self.form = {v: k for k, v in RAGGED_FORMS.items()}[lib.FFTLIB_GPU_raggedGetForm(self.h)]
        lg = np.log2(lengths.astype(np.float64))
        self.flops = float(2 * (5.0 * lengths * lg).sum())
        self.bytes = float(2 * 16.0 * total)
        self.dtype = "f32"
        self.unit = "GFLOP/s"
        self.config = {"workload": "ragged batch of 10000 C2C FFTs, lengths log-uniform on [64, 65536] snapped to "
                                   "2/3/5-smooth (seed 10000), forward then inverse in one graph launch",
                       "transforms": int(len(lengths)), "distinct_lengths": int(len(np.unique(lengths))),
                       "points": total, "tuned": not self.args.no_tune, "form": self.form,
                       "input": "complex64 uniform[-1,1) Philox seed 10001"}
        self.kernel_name = RAGGED_KERNELS[self.form]
        xp, yp, zp, h = self.x.data_ptr(), self.y.data_ptr(), self.z.data_ptr(), self.h
        fn = lib.FFTLIB_GPU_raggedExec

        def step():
            s = fn(h, xp, yp, zp, 2)
            if s != 0:
                raise RuntimeError(f"raggedExec status {s}: {F.last_error()}")
        self.step = step

    def extra_ragged_10000(self, steps):
        """Both plan forms timed on the same buffers after the headline steps
        (ms per forward + inverse step, HIP events, best of 3 x `steps`)."""
        lib, torch = self.lib, self.torch
        stream = torch.cuda.current_stream()
        forms = {}
        for name, code in RAGGED_FORMS.items():
            assert lib.FFTLIB_GPU_raggedSetForm(self.h, code) == 0
            for _ in range(3):
                self.step()
            best = float("inf")
            for _ in range(3):
                e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                e0.record(stream)
                for _ in range(steps):
                    self.step()
                e1.record(stream)
                torch.cuda.synchronize()
                best = min(best, e0.elapsed_time(e1) / steps)
            forms[name] = {"ms_per_step": round(best, 5), "kernel": RAGGED_KERNELS[name]}
        assert lib.FFTLIB_GPU_raggedSetForm(self.h, RAGGED_FORMS[self.form]) == 0
        return {"forms": forms}

    def check_ragged_10000(self):
        """Forward output of 64 transforms against the FP64 FFT; round-trip error
        of the whole batch."""
        idx = np.linspace(0, len(self.lengths) - 1, 64).astype(np.int64)
        offs = np.concatenate([[0], np.cumsum(self.lengths)[:-1]])
        y = self.y.cpu().numpy().view(np.complex64)
        num = den = 0.0
        for i in idx:
            o, n = int(offs[i]), int(self.lengths[i])
            ref = np.fft.fft(self.host_x[o:o + n].astype(np.complex128))
            num += float(np.sum(np.abs(y[o:o + n] - ref) ** 2))
            den += float(np.sum(np.abs(ref) ** 2))
        z = self.z.cpu().numpy().view(np.complex64)
        rt = float(np.linalg.norm(z - self.host_x) / np.linalg.norm(self.host_x))
        return {"fwd_rel_l2_vs_fp64": math.sqrt(num / den), "roundtrip_rel_l2": rt, "checked_transforms": 64}

    def cpu_ragged_10000(self, seconds):
        import oracle
        offs = np.concatenate([[0], np.cumsum(self.lengths)[:-1]])
        order = philox(7).permutation(len(self.lengths))
        done_flops, t_total, count = 0.0, 0.0, 0
        for i in order:
            o, n = int(offs[i]), int(self.lengths[i])
            x = self.host_x[o:o + n][None, :]
            t0 = time.perf_counter()
            y = oracle.c2c(x, n, 1)
            oracle.c2c(y, n, 1, inverse=True)
            t_total += time.perf_counter() - t0
            done_flops += 2 * 5.0 * n * math.log2(n)
            count += 1
            if t_total > seconds:
                break
        return {"value": done_flops / t_total / 1e9, "unit": "GFLOP/s", "cores": 1, "kind": "port",
                "sample": f"oracle C2C forward + inverse (natural-C restatement), {count} of the 10000 transforms "
                          f"in random order, {t_total:.2f} s"}

    def _setup_copy_503m(self):
        torch = self.torch
        self.n, self.b = 960, 65536
        count = 2 * self.n * self.b
        self.x = torch.ones(count, dtype=torch.float32, device="cuda:0")
        self.y = torch.empty_like(self.x)
        self.flops = 0.0
        self.bytes = 2.0 * 4 * count
        self.dtype = "f32"
        self.unit = "GB/s"
        self.config = {"workload": "torch copy_ of 503,316,480 bytes (HBM counter calibration)"}
        self.kernel_name = "copy"

        def step():
            self.y.copy_(self.x)
        self.step = step

    def check_copy_503m(self):
        return {"copied_equal": bool(torch_equal(self.x, self.y))}

    def cpu_copy_503m(self, seconds):
        return None

    # -- int16, N = 256, B = 8192 ---------------------------------------------
    def _setup_i16_256(self):
        F, lib, torch = self.F, self.lib, self.torch
        self.n, self.b = 256, 8192
        n, b = self.n, self.b
        r = philox(256)
        host = r.integers(-2048, 2048, size=(b, n, 2)).astype(np.int16)
        self.host_x = host
        w = np.zeros(2 * n, np.int16)
        lib.FFTLIB_tw_gen_i16sc(w.ctypes.data, n)
        self.shifts = np.array([0, 1, 1, 1], np.uint32)
        self.x = torch.from_numpy(host.reshape(-1)).to("cuda:0")
        self.w = torch.from_numpy(w).to("cuda:0")
        self.sh = torch.from_numpy(self.shifts.view(np.int32)).to("cuda:0")
        self.y = torch.empty_like(self.x)
        self.bp_x = F.BufParams1D(F.FFTLIB_INT16, 2 * n * b)
        self.bp_w = F.BufParams1D(F.FFTLIB_INT16, 2 * n)
        self.bp_s = F.BufParams1D(F.FFTLIB_UINT32, 4)
        self.pblock = ctypes.create_string_buffer(F.PBLOCK_SIZE["FFTLIB_fft1dBatched_i16sc_c16sc_o16sc"])
        args = (self.x.data_ptr(), ctypes.byref(self.bp_x), self.w.data_ptr(), ctypes.byref(self.bp_w),
                self.y.data_ptr(), ctypes.byref(self.bp_x), self.sh.data_ptr(), ctypes.byref(self.bp_s), n, b,
                self.pblock)
        assert lib.FFTLIB_fft1dBatched_i16sc_c16sc_o16sc_init(*args) == 0
        self.flops = 5.0 * n * math.log2(n) * b
        self.bytes = 8.0 * n * b
        self.dtype = "int16"
        self.unit = "GOP/s"
        self.config = {"workload": "FFTLIB_fft1dBatched_i16sc_c16sc_o16sc N=256 B=8192 shifts [0,1,1,1]",
                       "numPoints": n, "numChannels": b}
        self.kernel_name = "i16"
        fn = lib.FFTLIB_fft1dBatched_i16sc_c16sc_o16sc_kernel

        def step():
            s = fn(*args)
            if s != 0:
                raise RuntimeError(f"kernel status {s}: {F.last_error()}")
        self.step = step

    def check_i16_256(self):
        import oracle
        y = self.y.cpu().numpy().reshape(self.b, self.n, 2)
        ref = oracle.i16(self.host_x, self.n, self.b, self.shifts)
        return {"bit_exact_vs_oracle": bool(np.array_equal(y, ref)), "checked_channels": self.b}

    def cpu_i16_256(self, seconds):
        import oracle
        t = []
        for _ in range(5):
            t0 = time.perf_counter()
            oracle.i16(self.host_x, self.n, self.b, self.shifts)
            t.append(time.perf_counter() - t0)
        tm = float(np.median(t))
        return {"value": self.flops / tm / 1e9, "unit": "GOP/s", "cores": 1, "kind": "port",
                "sample": f"oracle int16 _cn, full workload N=256 B=8192, median of 5 ({tm:.3f} s)"}

    # -- R2C + C2R, N = 4096, B = 16384 ------------------------------------------
    def _setup_r2c_4096(self):
        F, lib, torch = self.F, self.lib, self.torch
        self.n, self.b = 4096, 16384
        n, b = self.n, self.b
        r = philox(4096)
        self.host_x = r.standard_normal((b, n), dtype=np.float32)
        w = np.zeros(n, np.float32)
        sf = np.zeros(2 * n, np.float32)
        lib.tw_gen_r2c(w.ctypes.data, n)
        lib.split_factor_gen_r2c(sf.ctypes.data, n)
        wi = np.zeros(n, np.float32)
        sfi = np.zeros(2 * n, np.float32)
        lib.ifft_c2r_tw_gen(wi.ctypes.data, n)
        lib.ifft_c2r_split_factor_gen(sfi.ctypes.data, n)
        dev = lambda a: torch.from_numpy(a).to("cuda:0")
        self.x, self.w, self.sf, self.wi, self.sfi = map(dev, (self.host_x.reshape(-1), w, sf, wi, sfi))
        nout = n // 2 + 1
        self.spec = torch.empty(2 * nout * b, dtype=torch.float32, device="cuda:0")
        self.back = torch.empty(n * b, dtype=torch.float32, device="cuda:0")
        self.scratch = torch.empty(1, dtype=torch.float32, device="cuda:0")
        bp = F.BufParams1D(F.FFTLIB_FLOAT32, n * b)
        self.bp = bp
        self.pb_f = ctypes.create_string_buffer(F.PBLOCK_SIZE["FFTLIB_fft1dBatched_i32f_c32fc_o32fc"])
        self.pb_i = ctypes.create_string_buffer(F.PBLOCK_SIZE["FFTLIB_ifft1dBatched_i32fc_c32fc_o32f"])
        B = ctypes.byref(bp)
        fa = (self.x.data_ptr(), B, self.w.data_ptr(), B, self.scratch.data_ptr(), B, self.sf.data_ptr(), B,
              self.spec.data_ptr(), B, n, b, 0, self.pb_f)
        ia = (self.spec.data_ptr(), B, self.wi.data_ptr(), B, self.scratch.data_ptr(), B, self.sfi.data_ptr(), B,
              self.back.data_ptr(), B, n, b, 0, self.pb_i)
        assert lib.FFTLIB_fft1dBatched_i32f_c32fc_o32fc_init(*fa) == 0
        assert lib.FFTLIB_ifft1dBatched_i32fc_c32fc_o32f_init(*ia) == 0
        fwd, inv = lib.FFTLIB_fft1dBatched_i32f_c32fc_o32fc_kernel, lib.FFTLIB_ifft1dBatched_i32fc_c32fc_o32f_kernel
        self.flops = 2 * 2.5 * n * math.log2(n) * b
        self.bytes = 2.0 * (4.0 * n * b + 8.0 * nout * b)
        self.dtype = "f32"
        self.unit = "GFLOP/s"
        self.config = {"workload": "R2C then C2R (FFTLIB_fft1dBatched_i32f_c32fc_o32fc + "
                                   "FFTLIB_ifft1dBatched_i32fc_c32fc_o32f) N=4096 B=16384, outmode/inMode 0",
                       "numPoints": n, "numChannels": b}
        self.kernel_name = ("r2c_lds_kernel + c2r_lds_kernel" if self.args.generic
                            else "fixed_kernel<2048, R2C> + fixed_kernel<2048, C2R>")

        def step():
            s1 = fwd(*fa)
            s2 = inv(*ia)
            if s1 or s2:
                raise RuntimeError(f"status {s1},{s2}: {F.last_error()}")
        self.step = step

    def check_r2c_4096(self):
        idx = np.arange(0, self.b, self.b // 64)
        nout = self.n // 2 + 1
        spec = self.spec.view(self.b, 2 * nout)[idx].cpu().numpy().view(np.complex64)
        back = self.back.view(self.b, self.n)[idx].cpu().numpy()
        x = self.host_x[idx].astype(np.float64)
        ref = np.fft.rfft(x, axis=1)
        return {"fwd_rel_l2_vs_fp64": float(np.linalg.norm(spec - ref) / np.linalg.norm(ref)),
                "roundtrip_rel_l2": float(np.linalg.norm(back - x) / np.linalg.norm(x)),
                "checked_channels": int(len(idx))}

    def cpu_r2c_4096(self, seconds):
        import oracle
        cnt = 256
        t = []
        for _ in range(5):
            t0 = time.perf_counter()
            spec = oracle.r2c(self.host_x[:cnt], self.n, cnt, 0)
            oracle.c2r(spec, self.n, cnt, 0)
            t.append(time.perf_counter() - t0)
        tm = float(np.median(t))
        flops = 2 * 2.5 * self.n * math.log2(self.n) * cnt
        return {"value": flops / tm / 1e9, "unit": "GFLOP/s", "cores": 1, "kind": "port",
                "sample": f"oracle R2C+C2R _cn, N=4096, {cnt} of 16384 channels, median of 5 ({tm:.3f} s)"}


def cpu_model():
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def usable_cores():
    """Threads for the all-cores CPU baseline: the CPUs this process may run
    on (affinity mask), bounded by the cgroup CPU quota when one is set."""
    n = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    note = f"affinity mask {n} CPUs, os.cpu_count() {os.cpu_count()}"
    try:
        with open("/sys/fs/cgroup/cpu.max") as fh:
            quota, period = fh.read().split()[:2]
        if quota != "max":
            q = max(1, int(math.ceil(int(quota) / int(period))))
            note += f", cgroup quota {q} CPUs"
            n = min(n, q)
    except (OSError, ValueError):
        pass
    return max(1, n), note


def torch_equal(a, b):
    return bool((a == b).all().item())


def load_traffic(workload, kernel):
    """Measured HBM bytes per launch from the committed rocprofv3 summary of
    the same kernel (None when no summary covers it)."""
    path = os.path.join(ROOT, "profiles", f"{workload}_traffic.json")
    if os.path.exists(path):
        with open(path) as fh:
            t = json.load(fh)
        if t.get("kernel") == kernel:
            return t
    return None


def time_steps(wl, args, stream, min_warmup_s):
    """W warmup steps, untimed steps until min_warmup_s of wall time, then K
    timed steps between HIP events on the launch stream. Returns (ms per
    step, wall seconds of the timed steps, ramp steps, ramp seconds)."""
    import torch

    def settle():
        if hasattr(wl, "finish"):
            wl.finish()
        torch.cuda.synchronize()

    for _ in range(args.warmup):
        wl.step()
    settle()
    # Clock ramp: after the host-side setup the device idles at low clocks and
    # needs tens of ms of sustained work to reach steady state (measured on
    # MI355X for the N=960 call: 385 us per launch for the first 10, 176 us
    # after about 100). Steps run in chunks that double until one takes 20 ms.
    ramp_steps, chunk, t_ramp = 0, 1, time.perf_counter()
    while time.perf_counter() - t_ramp < min_warmup_s:
        t_chunk = time.perf_counter()
        for _ in range(chunk):
            wl.step()
        settle()
        ramp_steps += chunk
        if time.perf_counter() - t_chunk < 0.02:
            chunk *= 2
    ramp_s = time.perf_counter() - t_ramp
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    ev0.record(stream)
    for _ in range(args.steps):
        wl.step()
    if hasattr(wl, "finish"):
        wl.finish()
    ev1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms = ev0.elapsed_time(ev1) / args.steps
    if getattr(wl, "wall_timing", False):
        ms = wall * 1e3 / args.steps       # host-to-host: the pipeline's own streams
    return ms, wall, ramp_steps, ramp_s


# Error bars of the C2C variants against the FP64 DFT: the split-precision
# goal is 2e-6 for every variant. BF16x2 (three bf16 products, about 4.6e-6)
# does not meet it: it is timed and reported with within_bar false, and only
# the variants built to meet the bar (BF16x3, FP16x2) gate the run; its own
# check bound (--variant bf16x2) is VARIANT_CHECK.
VARIANT_BARS = {"fp32": 1e-6, "bf16x3": 2e-6, "fp16x2": 2e-6, "bf16x2": 2e-6}
VARIANT_CHECK = {"fp32": 1e-6, "bf16x3": 2e-6, "fp16x2": 2e-6, "bf16x2": 1e-5}
VARIANT_GATED = ("bf16x3", "fp16x2")
VARIANT_CODES = {"fp32": 0, "bf16x3": 1, "fp16x2": 2, "bf16x2": 3}


def check_passes(args, chk):
    if args.workload in ("c2c_960", "c2c_960_host"):
        if args.variant == "fp32":
            return chk["rel_l2_vs_oracle"] <= 1e-6 and chk["rel_l2_vs_fp64"] <= 1e-6
        return chk["rel_l2_vs_fp64"] <= VARIANT_CHECK[args.variant]
    if args.workload == "ragged_10000":
        return chk["fwd_rel_l2_vs_fp64"] <= 1e-6 and chk["roundtrip_rel_l2"] <= 1e-6
    if args.workload == "r2c_4096":
        return chk["fwd_rel_l2_vs_fp64"] <= 1e-6 and chk["roundtrip_rel_l2"] <= 1e-6
    if args.workload == "i16_256":
        return bool(chk["bit_exact_vs_oracle"])
    if args.workload == "copy_503m":
        return bool(chk["copied_equal"])
    return True


def time_variants(wl, args, lib, F, torch, stream, fp32_ms):
    """The split-precision MFMA variants of the same call, timed in the same
    process right after the fp32 headline: ms per step, roofline fraction,
    relative L2 against the FP64 DFT on 256 channels, and the ratio to the
    fp32 time (below 1 = faster than fp32)."""
    res = {}
    try:
        for name in ("bf16x3", "fp16x2", "bf16x2"):
            assert lib.FFTLIB_GPU_setC2CVariant(VARIANT_CODES[name]) == 0, F.last_error()
            lib.FFTLIB_GPU_setSynchronous(0)
            ms, _, _, _ = time_steps(wl, args, stream, min(0.15, args.min_warmup_s))
            lib.FFTLIB_GPU_setSynchronous(1)
            chk = wl.check_c2c_960()
            frac = wl.bytes / (ms / 1e3) / 1e9 / HBM_PEAK_GBS
            res[name] = {"ms_per_step": round(ms, 5), "gflops": round(wl.flops / (ms / 1e3) / 1e9, 1),
                         "roofline_frac": round(frac, 4), "rel_l2_vs_fp64": chk["rel_l2_vs_fp64"],
                         "bar": VARIANT_BARS[name], "within_bar": chk["rel_l2_vs_fp64"] <= VARIANT_BARS[name],
                         "time_vs_fp32": round(ms / fp32_ms, 4),
                         "kernel": "c2c_mfma_kernel<%s, 15, 4>" % name}
    finally:
        assert lib.FFTLIB_GPU_setC2CVariant(0) == 0
    return res


MFMA_LENGTHS = [32, 48, 64, 80, 96, 128, 144, 160, 192, 240, 256, 288, 320, 384, 480, 512, 576, 640, 768, 960, 1024]


def sweep_mfma(lib, F, torch, stream):
    """Per MFMA length: fp32 and the three split variants on 2^25 points per
    call (best of 3 x 20 launches between HIP events on the launch stream) and
    each one's rel L2 error on 64 transforms against numpy's FP64 FFT."""
    points = 1 << 25
    name = "FFTLIB_fft1dBatched_i32fc_c32fc_o32fc"
    fn = lib.FFTLIB_fft1dBatched_i32fc_c32fc_o32fc_kernel
    r = philox(25)
    rows = []
    for n in MFMA_LENGTHS:
        b = points // n
        host = np.empty((b, n), np.complex64)
        host.real = r.uniform(-1, 1, (b, n))
        host.imag = r.uniform(-1, 1, (b, n))
        x = torch.from_numpy(host.view(np.float32).reshape(-1)).to("cuda:0")
        y = torch.empty_like(x)
        w = np.zeros(2 * n, np.float32)
        lib.FFTLIB_tw_gen_c2c(w.ctypes.data, n)
        wd = torch.from_numpy(w).to("cuda:0")
        pb = ctypes.create_string_buffer(F.PBLOCK_SIZE[name])
        bx, bw = F.BufParams1D(F.FFTLIB_FLOAT32, 2 * n * b), F.BufParams1D(F.FFTLIB_FLOAT32, 2 * n)
        args = (x.data_ptr(), ctypes.byref(bx), wd.data_ptr(), ctypes.byref(bw), y.data_ptr(), ctypes.byref(bx), n, b,
                pb)
        assert lib.FFTLIB_fft1dBatched_i32fc_c32fc_o32fc_init(*args) == 0, F.last_error()
        ref = np.fft.fft(host[:64].astype(np.complex128), axis=1)
        row = {"n": n, "channels": b}
        for vname, code in VARIANT_CODES.items():
            assert lib.FFTLIB_GPU_setC2CVariant(code) == 0, F.last_error()
            for _ in range(5):
                assert fn(*args) == 0, F.last_error()
            torch.cuda.synchronize()
            best = float("inf")
            for _ in range(3):
                e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                e0.record(stream)
                for _ in range(20):
                    assert fn(*args) == 0, F.last_error()
                e1.record(stream)
                torch.cuda.synchronize()
                best = min(best, e0.elapsed_time(e1) / 20 * 1e3)
            got = y[: 2 * n * 64].cpu().numpy().view(np.complex64).reshape(64, n).astype(np.complex128)
            row[vname] = {"us": round(best, 1), "rel_l2_vs_fp64": float(np.linalg.norm(got - ref) / np.linalg.norm(ref))}
        assert lib.FFTLIB_GPU_setC2CVariant(0) == 0
        rows.append(row)
        print(f"sweep n={n}: " + ", ".join(f"{k} {row[k]['us']} us" for k in VARIANT_CODES), file=sys.stderr,
              flush=True)
        del x, y
    return {"sweep": "C2C fp32 vs split-precision MFMA variants per length", "points_per_call": points,
            "timing": "best of 3 x 20 launches, HIP events on the launch stream", "rows": rows}


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this benchmark runs on one GPU only: the batched FFT has no multi-GPU path "
                                   "(a batch would be split by channels, one share per GPU)", "n_gpus": args.gpus}))
        sys.exit(2)
    import torch

    import c7x_fftlib_amd as F
    lib = F.load()
    assert lib.FFTLIB_setFunctionStyle(F.FFTLIB_FUNCTION_GPU) == 0
    torch.cuda.init()
    stream = torch.cuda.current_stream()
    lib.FFTLIB_GPU_setStream(ctypes.c_void_p(stream.cuda_stream))
    lib.FFTLIB_GPU_setSpecializedKernels(0 if args.generic else 1)
    if args.variant != "fp32":
        sel = getattr(lib, "FFTLIB_GPU_setC2CVariant", None)
        if sel is None:
            raise SystemExit("variant selection not available in this library build")
        code = {"fp32": 0, "bf16x3": 1, "fp16x2": 2, "bf16x2": 3}[args.variant]
        assert sel(code) == 0, F.last_error()

    if args.sweep_mfma:
        lib.FFTLIB_GPU_setSynchronous(0)
        print(json.dumps(sweep_mfma(lib, F, torch, stream)), flush=True)
        lib.FFTLIB_GPU_setSynchronous(1)
        return
    wl = Workload(args, F, lib, torch)
    lib.FFTLIB_GPU_setSynchronous(0)
    ms, wall, ramp_steps, ramp_s = time_steps(wl, args, stream, args.min_warmup_s)
    lib.FFTLIB_GPU_setSynchronous(1)
    t = ms / 1e3
    achieved = wl.bytes / t / 1e9
    value = wl.flops / t / 1e9 if wl.flops else achieved
    traffic = load_traffic(args.workload, wl.kernel_name)
    out = {
        "metric": METRIC, "value": round(value, 2), "unit": wl.unit, "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup,
        "clock_ramp": {"untimed_steps": ramp_steps, "seconds": round(ramp_s, 3),
                       "why": "extra untimed steps after the W warmup steps, until --min-warmup-s of wall time"},
        "ms_per_step": round(ms, 5), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": wl.dtype, "data": "synthetic",
        "config": dict(wl.config, kernels="generic" if args.generic else "specialised"),
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": traffic.get("bytes_per_launch") if traffic else None,
                     "algorithmic_bytes_per_launch": wl.bytes, "kernel": wl.kernel_name,
                     "timing": "HIP events on the launch stream, average over the timed steps"},
        "wall_s_timed": round(wall, 4),
    }
    if getattr(wl, "wall_timing", False):
        out["roofline"] = {"bound": "pcie", "achieved": round(wl.bytes / 2 / t / 1e9, 2), "peak": 64.0,
                           "unit": "GB/s", "frac": round(wl.bytes / 2 / t / 1e9 / 64.0, 4), "traffic": None,
                           "kernel": wl.kernel_name,
                           "timing": "wall clock over the timed submits and the final pipeline wait; bytes "
                                     "per direction against PCIe Gen5 x16 (64 GB/s each way)"}
    if hasattr(wl, "extra_" + args.workload):
        out.update(getattr(wl, "extra_" + args.workload)(max(3, args.steps // 4)))
    out["check"] = getattr(wl, "check_" + args.workload)()
    ok = check_passes(args, out["check"])
    out["check"]["passed"] = ok
    if args.workload == "c2c_960" and args.variant == "fp32" and not args.generic and not args.no_variants:
        out["variants"] = time_variants(wl, args, lib, F, torch, stream, ms)
        ok = ok and all(out["variants"][k]["within_bar"] for k in VARIANT_GATED)
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = getattr(wl, "cpu_" + args.workload)(args.cpu_seconds)
    print(json.dumps(out), flush=True)
    if not ok:
        print("bench: result check failed (see the check / variants fields)", file=sys.stderr, flush=True)
        sys.exit(1)


if __name__ == "__main__":
    main()
