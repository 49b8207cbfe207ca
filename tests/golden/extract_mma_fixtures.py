"""Convert the reference's committed fft_c7xmma test vectors into .npz test data.

The reference commits, for each of its four MMA FFT kernels, input /
expected-output pairs and shift vectors as C array literals in
test/fft_c7xmma/<kernel>/<kernel>_idat.c, with the shape of every case in the
same file's testParams table ({STATIC, in, out, shifts, data type, batchSize,
fftSize, interleave, reps, testID}). This script reads those literals (no
code is executed), checks that every array has the length its case implies
(2 * fftSize * batchSize values, NUMSHIFTS shifts) and writes
tests/golden/mma_<alias>.npz with, per case ID: in_<id>, out_<id>, shift_<id>
and a params table (id, dtype, batchSize, fftSize, interleave), dtype 1 =
FFTLIB_INT16, 2 = FFTLIB_INT32. The reference driver compares with a
tolerance of 50 (test/fft_c7xmma/<kernel>/<kernel>_d.c, TI_compare_mem_2D).

It runs only where the reference tree is present; the tests read the .npz.
Usage: python tests/golden/extract_mma_fixtures.py [reference_root]
"""
import os
import re
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
MAX_VALUES = 1 << 16   # in + out values per kept case

KERNELS = {
    "fft": "FFTLIB_FFT_fft_ixX_cxX_oxX",
    "dftsmall": "FFTLIB_FFT_dftSmall_ixX_cxX_oxX",
    "dftlarge": "FFTLIB_FFT_dftLarge_ixX_cxX_oxX",
    "highradix": "FFTLIB_FFT_highRadixDecompositions_ixX_cxX_oxX",
}
NUMSHIFTS = {"fft": 5, "dftsmall": 1, "dftlarge": 1, "highradix": 5}

ARRAY_RE = re.compile(r"static\s+(int16_t|int32_t|uint32_t)\s+(\w+)\s*\[[^\]]*\]\s*=\s*\{(.*?)\};", re.S)
CASE_RE = re.compile(r"\{\s*STATIC\s*,\s*(\w+)\s*,\s*(\w+)\s*,\s*(\w+)\s*,\s*FFTLIB_(INT16|INT32)\s*,"
                     r"\s*(\d+)\s*,\s*(\d+)\s*,\s*(\d+)\s*,\s*(\d+)\s*,\s*(\d+)\s*\}")


def parse(path):
    text = open(path).read()
    arrays = {}
    for ctype, name, body in ARRAY_RE.findall(text):
        vals = [int(v) for v in re.findall(r"-?\d+", body)]
        arrays[name] = (ctype, vals)
    cases = []
    for m in CASE_RE.finditer(text):
        cin, cout, csh, dt, batch, fft, inter, _reps, tid = m.groups()
        cases.append(dict(inp=cin, out=cout, shift=csh, dtype=1 if dt == "INT16" else 2, batch=int(batch),
                          fft=int(fft), interleave=int(inter), id=int(tid)))
    return arrays, cases


def main(ref_root):
    for alias, kernel in KERNELS.items():
        path = os.path.join(ref_root, "test", "fft_c7xmma", kernel, kernel + "_idat.c")
        arrays, cases = parse(path)
        data, params, seen = {}, [], set()
        for c in cases:
            if c["id"] in seen:
                continue
            n = 2 * c["fft"] * c["batch"]
            if 2 * n > MAX_VALUES:
                continue
            ti, vi = arrays[c["inp"]]
            to, vo = arrays[c["out"]]
            _, vs = arrays[c["shift"]]
            want = "int16_t" if c["dtype"] == 1 else "int32_t"
            assert ti == want and to == want, (alias, c["id"], ti, to)
            assert len(vi) == n and len(vo) == n, (alias, c["id"], len(vi), len(vo), n)
            assert len(vs) == NUMSHIFTS[alias], (alias, c["id"], len(vs))
            np_t = np.int16 if c["dtype"] == 1 else np.int32
            data[f"in_{c['id']}"] = np.array(vi, np_t)
            data[f"out_{c['id']}"] = np.array(vo, np_t)
            data[f"shift_{c['id']}"] = np.array(vs, np.uint32)
            params.append([c["id"], c["dtype"], c["batch"], c["fft"], c["interleave"]])
            seen.add(c["id"])
        data["params"] = np.array(params, np.int64)
        out = os.path.join(HERE, f"mma_{alias}.npz")
        np.savez_compressed(out, **data)
        print(f"{alias}: {len(params)} cases of {len(cases)} -> {out} ({os.path.getsize(out)} bytes)")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
