"""Write int16 FFT golden vectors from the reference's own Python model.

The reference keeps a Python model of its int16 batched FFT
(test/fft_c7x/FFTLIB_idat_gen/FFTLIB_fft1dBatched_i16sc_c16sc_o16sc/
fft1dBatched_i16sc_c16sc_o16sc.py:8-350) that restates the natural-C kernel
with unbounded Python integers. Where no 16-bit add wraps, its output is the
natural-C kernel's output bit for bit, so it pins the exact fixed-point bits
(rounding of the Q15 products, the per-stage rounding shift, the shared shift
of the 8-point and final 2-point stages, the bit-reversed order) that no
closed form does.

This script runs only where the reference tree is present (it is not needed
to run the tests): it imports that one module from the reference tree, runs it
on seeded inputs and saves inputs, twiddles, shift vectors and outputs to
tests/golden/i16_model_vectors.npz. Inputs are kept small enough in amplitude
that no intermediate 16-bit sum can wrap (the model does not wrap), and the
shift vectors are the ones of the reference's test list
(test/fft_c7x/FFTLIB_idat_gen/FFTLIB_fft1dBatched_i16sc_c16sc_o16sc/
test_cases_list.csv) plus a few of this file's own.

Usage: python tests/golden/gen_i16_vectors.py [reference_root]
"""
import importlib.util
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
MODEL_REL = ("test/fft_c7x/FFTLIB_idat_gen/FFTLIB_fft1dBatched_i16sc_c16sc_o16sc/"
             "fft1dBatched_i16sc_c16sc_o16sc.py")

# (numPoints, numChannels, shift vector, input amplitude) -- shift vectors of
# the reference test list IDs 1..40 plus extra shapes (N = 4, 8, 16, 32).
CASES = [
    (4, 3, [1], 4096),
    (4, 2, [0], 4096),
    (8, 3, [1, 1], 4096),
    (8, 2, [0, 2], 2048),
    (16, 2, [1, 1], 2048),
    (32, 2, [1, 1, 1], 2048),
    (64, 4, [0, 1, 1], 2048),
    (64, 2, [1, 0, 1], 2048),
    (64, 2, [1, 1, 0], 1024),
    (128, 2, [0, 1, 1], 2048),
    (128, 2, [1, 0, 1], 2048),
    (128, 2, [0, 0, 1], 512),
    (256, 4, [0, 1, 1, 1], 2048),
    (256, 2, [1, 0, 1, 1], 2048),
    (256, 2, [0, 0, 1, 1], 512),
    (256, 2, [1, 1, 0, 1], 2048),
    (512, 2, [0, 1, 1, 1], 2048),
    (512, 2, [1, 0, 1, 1], 2048),
    (1024, 2, [0, 1, 1, 1, 1], 2048),
    (1024, 1, [1, 0, 1, 0, 1], 1024),
    (2048, 1, [0, 1, 1, 1, 1], 2048),
    (4096, 1, [0, 1, 1, 0, 0, 1], 256),
]


def load_model(ref_root):
    path = os.path.join(ref_root, MODEL_REL)
    spec = importlib.util.spec_from_file_location("ref_i16_model", path)
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod.FFTLIB_fft1dBatched_i16sc_c16sc_o16sc


def main():
    ref_root = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    Model = load_model(ref_root)
    rng = np.random.default_rng(20261203)
    out = {}
    for idx, (n, b, shifts, amp) in enumerate(CASES):
        re = rng.integers(-amp, amp, size=(b, n))
        im = rng.integers(-amp, amp, size=(b, n))
        g = re + 1j * im
        model = Model(b, n, list(shifts))
        y, _ = model.exec(g)
        y_int = np.stack([np.real(y), np.imag(y)], axis=-1)
        assert np.all(y_int == np.round(y_int)), "model output not integral"
        assert np.all(np.abs(y_int) <= 32768), "model output outside int16"
        out[f"c{idx}_x"] = np.stack([re, im], axis=-1).astype(np.int16)
        out[f"c{idx}_w"] = np.asarray(model.pW, dtype=np.int16)
        out[f"c{idx}_shift"] = np.asarray(shifts, dtype=np.uint32)
        out[f"c{idx}_y"] = y_int.astype(np.int16)
        out[f"c{idx}_meta"] = np.asarray([n, b], dtype=np.int64)
    out["num_cases"] = np.asarray(len(CASES))
    dst = os.path.join(HERE, "i16_model_vectors.npz")
    np.savez_compressed(dst, **out)
    print(f"wrote {len(CASES)} cases to {dst}")


if __name__ == "__main__":
    main()
