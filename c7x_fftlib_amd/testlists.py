"""The reference's test lists and performance files on the GPU route.

The reference drives each fft_c7x kernel from a test list,
``test/fft_c7x/FFTLIB_idat_gen/<kernel>/test_cases_list.csv`` (one row per
case: ID, numChannels, numPoints, outmode / inMode, numShifts, shiftVector,
testType, testCategory), with the STATIC cases' input and expected output in
``staticRefCase<ID>.h`` C headers; its drivers (``test/fft_c7x/<kernel>/
<kernel>_d.c``) fill the buffers, call _init and _kernel, compare, and write
one performance CSV per kernel whose published copies live in
``docs/sphinx/perf_data/<device>/<kernel>.csv``.

This module runs such a list unchanged on the GPU through the C-ABI:

* :func:`read_test_cases` and :func:`read_static_ref_case` read the two
  reference formats;
* :func:`make_case_data` builds a case's buffers as the kernel's driver does
  (STATIC cases from their header when one is given, else from the input
  distribution of the reference's generator for that list; RANDOM and
  RANDOM_SIGNED cases from the driver's distribution, seeded);
* :func:`run_case` plans and runs the case on device buffers and times the
  warm calls with HIP events;
* :func:`check_expected` compares a STATIC case with its header's expected
  output under the driver's tolerance;
* :func:`write_perf_csv` writes results in the perf_data schema, with the
  time in nanoseconds per call where the reference has EVM cycles.

Usage::

    python -m c7x_fftlib_amd.testlists KERNEL test_cases_list.csv \\
        [--static-dir DIR] [--perf-csv OUT.csv] [--warm 4] [--ids 1,5,9]

Nothing here computes a reference result: the STATIC expected outputs are the
reference's own; comparisons with the natural-C restatement live in the
test suite.
"""
import argparse
import csv
import ctypes
import os
import re
import sys
from dataclasses import dataclass, field

import numpy as np

from . import (FFTLIB_FLOAT32, FFTLIB_FUNCTION_GPU, FFTLIB_INT16, FFTLIB_UINT32, PBLOCK_SIZE, BufParams1D,
               last_error, load)


@dataclass
class TestCase:
    """One row of a test_cases_list.csv (numChannels is 1 for the
    single-channel lists, which have no such column)."""
    id: int
    numPoints: int
    numChannels: int = 1
    mode: int = 0              # outmode (R2C) / inMode (C2R); 0 where the list has none
    shifts: list = field(default_factory=list)
    testType: str = "STATIC"
    testCategory: int = 200


@dataclass
class KernelInfo:
    kind: str                  # c2c, ic2c, i16, r2c, c2r
    batched: bool
    perf_columns: list         # columns of the perf_data CSV before the time
    random_range: float        # RANDOM / RANDOM_SIGNED fp32 inputs ~ U[-r, r); 0: int16 full range
    static_gen: str            # input distribution of the reference generator for STATIC cases
    tol_expected: float        # driver tolerance against staticRefOut (0: bit-exact)
    driver: str                # where the driver fills and compares


_C = "test/fft_c7x/"
KERNELS = {
    "FFTLIB_fft1dBatched_i32fc_c32fc_o32fc": KernelInfo(
        "c2c", True, ["Test ID", "Num Channels", "Num Points"], 256.0, "complex_uniform", 2.0 ** -5,
        _C + "FFTLIB_fft1dBatched_i32fc_c32fc_o32fc/FFTLIB_fft1dBatched_i32fc_c32fc_o32fc_d.c:183-311"),
    "FFTLIB_ifft1dBatched_i32fc_c32fc_o32fc": KernelInfo(
        "ic2c", True, ["Test ID", "Num Channels", "Num Points"], 256.0, "complex_0_100", 2.0 ** -5,
        _C + "FFTLIB_ifft1dBatched_i32fc_c32fc_o32fc/FFTLIB_ifft1dBatched_i32fc_c32fc_o32fc_d.c:184-312"),
    "FFTLIB_fft1d_i32fc_c32fc_o32fc": KernelInfo(
        "c2c", False, ["Test ID", "Num Points"], 1.0, "complex_uniform", 2.0 ** -10,
        _C + "FFTLIB_fft1d_i32fc_c32fc_o32fc/FFTLIB_fft1d_i32fc_c32fc_o32fc_d.c:149-227"),
    "FFTLIB_ifft1d_i32fc_c32fc_o32fc": KernelInfo(
        "ic2c", False, ["Test ID", "Num Points"], 128.0, "complex_uniform", 1e-3,
        _C + "FFTLIB_ifft1d_i32fc_c32fc_o32fc/FFTLIB_ifft1d_i32fc_c32fc_o32fc_d.c:148-240"),
    "FFTLIB_fft1dBatched_i16sc_c16sc_o16sc": KernelInfo(
        "i16", True, ["Test ID", "Num Channels", "Num Points"], 0.0, "int_1000", 0.0,
        _C + "FFTLIB_fft1dBatched_i16sc_c16sc_o16sc/FFTLIB_fft1dBatched_i16sc_c16sc_o16sc_d.c:142-236"),
    "FFTLIB_fft1d_i16sc_c16sc_o16sc": KernelInfo(
        "i16", False, ["Test ID", "Num Points"], 0.0, "int_1000", 0.0,
        _C + "FFTLIB_fft1d_i16sc_c16sc_o16sc/FFTLIB_fft1d_i16sc_c16sc_o16sc_d.c:123-199"),
    "FFTLIB_fft1dBatched_i32f_c32fc_o32fc": KernelInfo(
        "r2c", True, ["Test ID", "Num Channels", "Num Points", "Outmode"], 128.0, "real_uniform", 2.0 ** -3,
        _C + "FFTLIB_fft1dBatched_i32f_c32fc_o32fc/FFTLIB_fft1dBatched_i32f_c32fc_o32fc_d.c:108-235"),
    "FFTLIB_fft1d_i32f_c32fc_o32fc": KernelInfo(
        "r2c", False, ["Test ID", "Num Points"], 128.0, "real_uniform", 2.0 ** -3,
        _C + "FFTLIB_fft1d_i32f_c32fc_o32fc/FFTLIB_fft1d_i32f_c32fc_o32fc_d.c:96-196"),
    "FFTLIB_ifft1dBatched_i32fc_c32fc_o32f": KernelInfo(
        "c2r", True, ["Test ID", "Num Channels", "Num Points", "Inmode"], 128.0, "complex_uniform", 1e-3,
        _C + "FFTLIB_ifft1dBatched_i32fc_c32fc_o32f/FFTLIB_ifft1dBatched_i32fc_c32fc_o32f_d.c:103-203"),
    "FFTLIB_ifft1d_i32fc_c32fc_o32f": KernelInfo(
        "c2r", False, ["Test ID", "Num Points"], 128.0, "complex_uniform", 1e-1,
        _C + "FFTLIB_ifft1d_i32fc_c32fc_o32f/FFTLIB_ifft1d_i32fc_c32fc_o32f_d.c:96-194"),
}

# Deviation: the single-channel complex driver fills RANDOM cases with rand()
# integers written as raw float bit patterns (TI_fillBuffer -> fillRand,
# test/common/TI_memory.c:1288-1295); the runner draws U[-1, 1) instead.

TIME_COLUMN = "MI355X ns"


# ---- reading the reference formats ------------------------------------------------
def read_test_cases(path):
    """Rows of a test_cases_list.csv as TestCase objects, in file order."""
    cases = []
    with open(path, newline="") as fh:
        for row in csv.DictReader(fh):
            row = {k.strip(): (v or "").strip() for k, v in row.items() if k}
            shifts = []
            if row.get("shiftVector"):
                shifts = [int(v) for v in row["shiftVector"].strip("[]").split(",") if v.strip()]
                shifts = shifts[:int(row.get("numShifts", len(shifts)) or len(shifts))]
            mode = row.get("outmode", row.get("inMode", "0")) or "0"
            cases.append(TestCase(id=int(row["ID"]), numPoints=int(row["numPoints"]),
                                  numChannels=int(row.get("numChannels", "1") or 1), mode=int(mode),
                                  shifts=shifts, testType=row.get("testType", "STATIC"),
                                  testCategory=int(row.get("testCategory", "200") or 200)))
    return cases


_CTYPES = {"float": np.float32, "FFTLIB_F32": np.float32, "double": np.float64, "int16_t": np.int16,
           "int32_t": np.int32, "uint32_t": np.uint32}
_ARRAY = re.compile(r"(?:static\s+)?(?:const\s+)?(\w+)\s+(staticRef(?:In|Out)\d+)\s*\[\s*\]\s*=\s*\{(.*?)\}\s*;",
                    re.S)


def read_static_ref_case(path):
    """The C array literals of a staticRefCase<ID>.h header: returns
    {"in": array, "out": array} typed by the arrays' declared C type (the
    values are parsed as text; nothing in the file is executed)."""
    with open(path) as fh:
        text = re.sub(r"//[^\n]*", "", fh.read())
    found = {}
    for ctype, name, body in _ARRAY.findall(text):
        dt = _CTYPES.get(ctype)
        if dt is None:
            raise ValueError(f"{path}: array {name} of unsupported type {ctype}")
        vals = [v.strip() for v in body.replace("\n", " ").split(",") if v.strip()]
        arr = np.asarray([float(v) for v in vals], dtype=np.float64)
        found["in" if name.startswith("staticRefIn") else "out"] = arr.astype(dt)
    if "in" not in found or "out" not in found:
        raise ValueError(f"{path}: no staticRefIn/staticRefOut arrays")
    return found


def shape_of(kernel, case):
    """(input elements, output elements) of one case, in scalars."""
    k = KERNELS[kernel]
    n, b = case.numPoints, (case.numChannels if k.batched else 1)
    if k.kind in ("c2c", "ic2c", "i16"):
        return 2 * n * b, 2 * n * b
    if k.kind == "r2c":
        nout = n if (case.mode or not k.batched) else n // 2 + 1
        return n * b, 2 * nout * b
    nin = n if (case.mode or not k.batched) else n // 2 + 1
    return 2 * nin * b, n * b


def find_static(static_dir, kernel, case):
    """(input, expected output) of a STATIC case from static_dir, or None.
    The header is staticRefCase<ID>.h; the inverse real-output batched list's
    headers are numbered one above their rows, so when the arrays do not have
    the case's lengths the next number is tried."""
    want = shape_of(kernel, case)
    for cid in (case.id, case.id + 1):
        path = os.path.join(static_dir, f"staticRefCase{cid}.h")
        if not os.path.exists(path):
            continue
        arrs = read_static_ref_case(path)
        if (arrs["in"].size, arrs["out"].size) == want:
            return arrs["in"], arrs["out"]
    return None


# ---- building and running a case ----------------------------------------------------
@dataclass
class CaseData:
    x: np.ndarray              # input scalars (float32 or int16)
    w: np.ndarray              # twiddle table
    aux: object                # split factors (float32), shift vector (uint32) or None
    out_count: int
    out_dtype: type
    expected: object = None    # STATIC expected output (reference header), when known


def _tables(kind, n):
    L = load()
    if kind in ("c2c", "ic2c"):
        w = np.zeros(max(2 * n, 2), np.float32)
        L.FFTLIB_tw_gen_c2c(w.ctypes.data, n)
        return w, None
    if kind == "i16":
        w = np.zeros(max(2 * n, 2), np.int16)
        L.FFTLIB_tw_gen_i16sc(w.ctypes.data, n)
        return w, None
    w = np.zeros(max(n, 2), np.float32)
    sf = np.zeros(max(2 * n, 2), np.float32)
    if kind == "r2c":
        L.tw_gen_r2c(w.ctypes.data, n)
        L.split_factor_gen_r2c(sf.ctypes.data, n)
    else:
        L.ifft_c2r_tw_gen(w.ctypes.data, n)
        L.ifft_c2r_split_factor_gen(sf.ctypes.data, n)
    return w, sf


def _generated_input(info, count, test_type, r):
    """Input scalars of a case without a header: RANDOM / RANDOM_SIGNED as
    the driver fills them, STATIC as the reference's generator draws them."""
    if info.kind == "i16":
        if test_type == "STATIC":          # generator: randint(-1000, 1001) per re/im
            return r.integers(-1000, 1001, count).astype(np.int16)
        return r.integers(-32768, 32768, count).astype(np.int16)   # all 16-bit patterns
    if test_type != "STATIC":
        return r.uniform(-info.random_range, info.random_range, count).astype(np.float32)
    if info.static_gen == "complex_0_100":
        return np.round(100.0 * r.random(count), 6).astype(np.float32)
    return r.uniform(-1.0, 1.0, count).astype(np.float32)


def make_case_data(kernel, case, static=None, static_dir=None, seed=None):
    """Buffers of one case. `static` = (input, expected) overrides the
    header lookup in `static_dir`."""
    info = KERNELS[kernel]
    nin, nout = shape_of(kernel, case)
    w, sf = _tables(info.kind, case.numPoints)
    aux = sf
    if info.kind == "i16":
        aux = np.asarray(case.shifts if case.shifts else [0], np.uint32)
    if static is None and static_dir and case.testType == "STATIC":
        static = find_static(static_dir, kernel, case)
    r = np.random.Generator(np.random.Philox(case.id if seed is None else seed))
    dt = np.int16 if info.kind == "i16" else np.float32
    if static is not None:
        x = np.asarray(static[0]).astype(dt)
        expected = np.asarray(static[1])
    else:
        x = _generated_input(info, nin, case.testType, r)
        expected = None
    if x.size != nin:
        raise ValueError(f"case {case.id}: input has {x.size} elements, the list's shape needs {nin}")
    return CaseData(x=x, w=w, aux=aux, out_count=nout, out_dtype=dt, expected=expected)


@dataclass
class CaseResult:
    id: int
    status: int
    y: object = None
    cold_ns: float = 0.0
    warm_ns: float = 0.0
    error: str = ""


def _bp(dt, n):
    return ctypes.byref(BufParams1D(dt, n))


def _calls(kernel, case, ptrs, info):
    """(init, kernel) thunks with the kernel's C signature."""
    L = load()
    n, b = case.numPoints, (case.numChannels if info.batched else 1)
    x, w, aux, y, pb = ptrs
    f32, i16 = FFTLIB_FLOAT32, FFTLIB_INT16
    init, kern = getattr(L, kernel + "_init"), getattr(L, kernel + "_kernel")
    if info.kind in ("c2c", "ic2c"):
        if info.batched:
            d, dw = _bp(f32, max(2 * n * b, 128 if n * b < 64 else 0)), _bp(f32, 2 * n)
            args = (x, d, w, dw, y, d, n, b, pb)
        else:
            d = _bp(f32, 2 * n)
            args = (x, d, w, d, y, d, pb)
    elif info.kind == "i16":
        bs = _bp(FFTLIB_UINT32, max(len(case.shifts), 1))
        if info.batched:
            d, dw = _bp(i16, max(2 * n * b, 256 if n * b < 128 else 0)), _bp(i16, 2 * n)
            args = (x, d, w, dw, y, d, aux, bs, n, b, pb)
        else:
            d = _bp(i16, 2 * n)
            args = (x, d, w, d, y, d, aux, bs, pb)
    else:
        d = _bp(f32, n * b)
        xfft = None            # the GPU route never writes pXFFT
        if info.batched:
            args = (x, d, w, d, xfft, d, aux, d, y, d, n, b, case.mode, pb)
        else:
            args = (x, d, w, d, xfft, d, aux, d, y, d, pb)
    return (lambda: init(*args)), (lambda: kern(*args))


def run_case(kernel, case, data, warm=4):
    """Plans and runs one case on the GPU (FFTLIB_FUNCTION_GPU, device
    buffers, synchronous _kernel): one cold call, then `warm` timed calls
    (median reported). Returns a CaseResult with the output of the cold
    call."""
    import torch
    info = KERNELS[kernel]
    L = load()
    if L.FFTLIB_setFunctionStyle(FFTLIB_FUNCTION_GPU) != 0:
        raise RuntimeError("cannot select the GPU function style")
    L.FFTLIB_GPU_setSynchronous(1)
    stream = torch.cuda.current_stream()
    old_stream = L.FFTLIB_GPU_getStream()
    L.FFTLIB_GPU_setStream(ctypes.c_void_p(stream.cuda_stream))
    try:
        dev = torch.device("cuda", torch.cuda.current_device())
        tx = torch.from_numpy(np.ascontiguousarray(data.x)).to(dev)
        tw = torch.from_numpy(np.ascontiguousarray(data.w)).to(dev)
        ta = None
        if data.aux is not None:
            a = np.ascontiguousarray(data.aux)
            ta = torch.from_numpy(a.view(np.int32) if a.dtype == np.uint32 else a).to(dev)
        tdt = torch.int16 if data.out_dtype == np.int16 else torch.float32
        ty = torch.zeros(max(data.out_count, 1), dtype=tdt, device=dev)
        pb = ctypes.create_string_buffer(PBLOCK_SIZE[kernel])
        ptrs = (tx.data_ptr(), tw.data_ptr(), ta.data_ptr() if ta is not None else None, ty.data_ptr(), pb)
        init, kern = _calls(kernel, case, ptrs, info)
        st = init()
        if st != 0:
            return CaseResult(case.id, st, error=last_error())
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        st = kern()
        e1.record(stream)
        if st != 0:
            return CaseResult(case.id, st, error=last_error())
        torch.cuda.synchronize()
        y = ty[:data.out_count].cpu().numpy()
        res = CaseResult(case.id, 0, y=y, cold_ns=e0.elapsed_time(e1) * 1e6)
        times = []
        for _ in range(warm):
            e0.record(stream)
            st = kern()
            e1.record(stream)
            if st != 0:
                return CaseResult(case.id, st, error=last_error())
            e1.synchronize()
            times.append(e0.elapsed_time(e1) * 1e6)
        res.warm_ns = float(np.median(times)) if times else res.cold_ns
        return res
    finally:
        L.FFTLIB_GPU_setStream(ctypes.c_void_p(old_stream))


def check_expected(kernel, data, y):
    """(passed, max abs difference) of an output against the case's STATIC
    expected output under the driver's tolerance (exact for int16); None
    when the case has no expected output."""
    if data.expected is None:
        return None
    info = KERNELS[kernel]
    exp = np.asarray(data.expected)
    if info.tol_expected == 0.0:
        ok = np.array_equal(np.asarray(y, np.int64), exp.astype(np.int64))
        return ok, float(np.max(np.abs(np.asarray(y, np.int64) - exp.astype(np.int64)))) if exp.size else 0.0
    diff = float(np.max(np.abs(np.asarray(y, np.float64) - exp.astype(np.float64)))) if exp.size else 0.0
    return diff <= info.tol_expected, diff


def write_perf_csv(path, kernel, cases, results):
    """Performance CSV in the docs/sphinx/perf_data schema of `kernel`
    ("Test ID,Num Channels,Num Points[,Outmode|Inmode],<time>" with ", "
    between values), one row per case that ran, time = warm ns per call."""
    info = KERNELS[kernel]
    by_id = {c.id: c for c in cases}
    with open(path, "w") as fh:
        fh.write(",".join(info.perf_columns + [TIME_COLUMN]) + "\n")
        for r in results:
            if r.status != 0:
                continue
            c = by_id[r.id]
            vals = {"Test ID": c.id, "Num Channels": c.numChannels, "Num Points": c.numPoints,
                    "Outmode": c.mode, "Inmode": c.mode}
            row = [str(vals[k]) for k in info.perf_columns] + [str(int(round(r.warm_ns)))]
            fh.write(", ".join(row) + "\n")


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("kernel", choices=sorted(KERNELS))
    ap.add_argument("cases", help="test_cases_list.csv")
    ap.add_argument("--static-dir", help="directory of the staticRefCase<ID>.h headers")
    ap.add_argument("--perf-csv", help="write the perf_data CSV here")
    ap.add_argument("--warm", type=int, default=4)
    ap.add_argument("--ids", help="comma-separated test IDs to run (default: all)")
    args = ap.parse_args(argv)
    cases = read_test_cases(args.cases)
    if args.ids:
        keep = {int(v) for v in args.ids.split(",")}
        cases = [c for c in cases if c.id in keep]
    results, failed = [], 0
    for c in cases:
        data = make_case_data(args.kernel, c, static_dir=args.static_dir)
        r = run_case(args.kernel, c, data, warm=args.warm)
        chk = check_expected(args.kernel, data, r.y) if r.status == 0 else None
        verdict = "ok" if r.status == 0 and (chk is None or chk[0]) else "FAIL"
        failed += verdict != "ok"
        note = f" vs staticRefOut max|diff| {chk[1]:.3g}" if chk else ""
        print(f"case {c.id:4d} {c.testType:13s} N={c.numPoints:7d} B={c.numChannels:5d} status {r.status} "
              f"{r.warm_ns:10.0f} ns{note} {verdict}{' ' + r.error if r.error else ''}", flush=True)
        results.append(r)
    if args.perf_csv:
        write_perf_csv(args.perf_csv, args.kernel, cases, results)
    return 1 if failed else 0


if __name__ == "__main__":
    sys.exit(main())
