"""Convert the reference's committed FFT fixtures into compact .npz test data.

The reference commits input/expected-output pairs for its real-input,
real-output and single-channel inverse transforms under
test/fft_c7x/<kernel>/test_data/extended_test_cases/staticRefCase<ID>.h
(numpy FFT results rounded to 3-6 decimals), with the shape of case <ID> in
test/fft_c7x/FFTLIB_idat_gen/<kernel>/test_cases_list.csv. This script reads
those two data files (C array literals and CSV, no code is executed), checks
that every array has the length the listed shape implies, and keeps the cases
small enough for the CPU test suite in tests/golden/ref_<alias>.npz.

It also copies the ten kernels' test lists (data files of the reference's
test generator) to tests/golden/testlists/<kernel>.csv, so that the lists run
unchanged against the GPU route without the reference tree.

It runs only where the reference tree is present; the tests read the copies.
Usage: python tests/golden/extract_reference_fixtures.py [reference_root]
"""
import csv
import os
import re
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
MAX_FLOATS = 24000  # in + out floats per kept case

# alias -> (kernel directory, function giving (in_len, out_len, params) from a CSV row)
def _r2cb(row):
    n, b, om = int(row["numPoints"]), int(row["numChannels"]), int(row["outmode"])
    nout = n if om == 1 else n // 2 + 1
    return n * b, 2 * nout * b, dict(numPoints=n, numChannels=b, mode=om)


def _c2rb(row):
    n, b, im = int(row["numPoints"]), int(row["numChannels"]), int(row["inMode"])
    nin = n // 2 + 1 if im == 0 else n
    return 2 * nin * b, n * b, dict(numPoints=n, numChannels=b, mode=im)


def _r2c1(row):
    n = int(row["numPoints"])
    return n, 2 * n, dict(numPoints=n, numChannels=1, mode=1)


def _c2r1(row):
    n = int(row["numPoints"])
    return 2 * n, n, dict(numPoints=n, numChannels=1, mode=1)


def _ic2c1(row):
    n = int(row["numPoints"])
    return 2 * n, 2 * n, dict(numPoints=n, numChannels=1, mode=0)


KERNELS = {
    "r2cb": ("FFTLIB_fft1dBatched_i32f_c32fc_o32fc", _r2cb),
    "c2rb": ("FFTLIB_ifft1dBatched_i32fc_c32fc_o32f", _c2rb),
    "r2c1": ("FFTLIB_fft1d_i32f_c32fc_o32fc", _r2c1),
    "c2r1": ("FFTLIB_ifft1d_i32fc_c32fc_o32f", _c2r1),
    "ic2c1": ("FFTLIB_ifft1d_i32fc_c32fc_o32fc", _ic2c1),
}

_ARRAY = re.compile(r"static\s+float\s+(staticRef(?:In|Out)\d+)\[\]\s*=\s*\{(.*?)\};", re.S)


def parse_arrays(path):
    with open(path) as fh:
        text = fh.read()
    arrays = {}
    for name, body in _ARRAY.findall(text):
        vals = [v for v in body.replace("\n", " ").split(",") if v.strip()]
        arrays[name] = np.asarray([float(v) for v in vals], dtype=np.float64)
    return arrays


LISTED = ["FFTLIB_fft1dBatched_i32fc_c32fc_o32fc", "FFTLIB_ifft1dBatched_i32fc_c32fc_o32fc",
          "FFTLIB_fft1d_i32fc_c32fc_o32fc", "FFTLIB_ifft1d_i32fc_c32fc_o32fc",
          "FFTLIB_fft1dBatched_i16sc_c16sc_o16sc", "FFTLIB_fft1d_i16sc_c16sc_o16sc",
          "FFTLIB_fft1dBatched_i32f_c32fc_o32fc", "FFTLIB_fft1d_i32f_c32fc_o32fc",
          "FFTLIB_ifft1dBatched_i32fc_c32fc_o32f", "FFTLIB_ifft1d_i32fc_c32fc_o32f"]


def copy_test_lists(ref_root):
    dst_dir = os.path.join(HERE, "testlists")
    os.makedirs(dst_dir, exist_ok=True)
    for kdir in LISTED:
        src = os.path.join(ref_root, "test/fft_c7x/FFTLIB_idat_gen", kdir, "test_cases_list.csv")
        with open(src, "rb") as fh:
            data = fh.read()
        with open(os.path.join(dst_dir, kdir + ".csv"), "wb") as fh:
            fh.write(data)
    print(f"copied {len(LISTED)} test lists -> {dst_dir}")


def main():
    ref_root = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    copy_test_lists(ref_root)
    for alias, (kdir, shape_of) in KERNELS.items():
        csv_path = os.path.join(ref_root, "test/fft_c7x/FFTLIB_idat_gen", kdir, "test_cases_list.csv")
        data_dir = os.path.join(ref_root, "test/fft_c7x", kdir, "test_data/extended_test_cases")
        with open(csv_path) as fh:
            rows = {int(r["ID"]): r for r in csv.DictReader(fh)}
        out, kept = {}, []
        for fname in sorted(os.listdir(data_dir)):
            m = re.match(r"staticRefCase(\d+)\.h$", fname)
            if not m:
                continue
            cid = int(m.group(1))
            arrs = parse_arrays(os.path.join(data_dir, fname))
            x, y = arrs.get(f"staticRefIn{cid}"), arrs.get(f"staticRefOut{cid}")
            if x is None or y is None:
                continue
            # The C2R batched files are numbered one above their CSV row
            # (staticRefCase2.h holds row 1: N=128, B=1, inMode 0); take the
            # row whose implied lengths match, preferring the same number.
            match = None
            for rid in (cid, cid - 1):
                if rid in rows:
                    in_len, out_len, params = shape_of(rows[rid])
                    if x.size == in_len and y.size == out_len:
                        match = params
                        break
            if match is None:
                print(f"{alias}: case {cid} skipped (array lengths match no listed shape)")
                continue
            params = match
            if x.size + y.size > MAX_FLOATS:
                continue
            out[f"c{cid}_x"] = x.astype(np.float32)
            out[f"c{cid}_y"] = y.astype(np.float32)
            out[f"c{cid}_meta"] = np.asarray([params["numPoints"], params["numChannels"], params["mode"]],
                                             dtype=np.int64)
            kept.append(cid)
        out["case_ids"] = np.asarray(sorted(kept), dtype=np.int64)
        dst = os.path.join(HERE, f"ref_{alias}.npz")
        np.savez_compressed(dst, **out)
        print(f"{alias}: kept cases {sorted(kept)} -> {os.path.basename(dst)}")


if __name__ == "__main__":
    main()
