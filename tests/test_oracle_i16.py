"""Pins the int16 oracle bit for bit.

  * vectors written by the reference's own Python model of the int16 kernel
    (tests/golden/gen_i16_vectors.py; inputs that never wrap, where that model
    equals the natural-C kernel), including its twiddle tables;
  * hand-worked 4-point cases: the twiddle-free radix-4 stage is an exact
    integer DFT-4, so the output is round-half-up(DFT4(x) / 2^s) computed
    directly, plus cases where a 16-bit sum wraps or -(-32768) saturates;
  * the FP64 DFT, scaled by 2^-(sum of shifts), within the fixed-point error
    bound (pins ordering, signs and twiddle placement).
"""
import os

import numpy as np
import pytest

import oracle

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def _model_cases():
    d = np.load(os.path.join(GOLDEN, "i16_model_vectors.npz"))
    return [(c, d[f"c{c}_x"], d[f"c{c}_w"], d[f"c{c}_shift"], d[f"c{c}_y"], d[f"c{c}_meta"])
            for c in range(int(d["num_cases"]))]


@pytest.mark.parametrize("case", _model_cases(), ids=lambda c: f"n{c[5][0]}_b{c[5][1]}_c{c[0]}")
def test_matches_reference_python_model(case):
    _, x, w, shift, y, (n, b) = case
    assert np.array_equal(oracle.tw_i16(int(n)), w)
    assert np.array_equal(oracle.i16(x, int(n), int(b), shift, w=w), y)
    if b == 1:
        assert np.array_equal(oracle.i16(x, int(n), 1, shift, w=w, single=True), y)


def _round_shift(v, s):
    """Round-half-up division by 2^s on exact integers."""
    if s == 0:
        return v
    return (v + (1 << (s - 1))) >> s


def _dft4(x):
    re = [int(v[0]) for v in x]
    im = [int(v[1]) for v in x]
    out = []
    for k in range(4):
        ar = ai = 0
        for nn in range(4):
            # exp(-2 pi i k n / 4) = (-i)^(k n)
            e = (k * nn) % 4
            cr, ci = [(1, 0), (0, -1), (-1, 0), (0, 1)][e]
            ar += re[nn] * cr - im[nn] * ci
            ai += re[nn] * ci + im[nn] * cr
        out.append((ar, ai))
    return out


@pytest.mark.parametrize("shift", [0, 1, 2, 3])
def test_four_point_exact(shift):
    r = np.random.Generator(np.random.Philox(shift))
    for _ in range(50):
        x = r.integers(-4000, 4000, size=(1, 4, 2)).astype(np.int16)
        got = oracle.i16(x, 4, 1, [shift])[0]
        want = [(_round_shift(a, shift), _round_shift(b, shift)) for a, b in _dft4(x[0])]
        assert [tuple(v) for v in got.tolist()] == want


def test_four_point_wraps_and_saturates():
    # all four points 16384: x0+x2 = 32768 wraps to -32768, the DC sum wraps to 0
    x = np.array([[[16384, 0]] * 4], np.int16)
    got = oracle.i16(x, 4, 1, [0])[0]
    assert got[0].tolist() == [0, 0]
    # x1 - x3 = (0, -32768): the rotation of -32768 saturates to 32767, so
    # X1.re = (x0 - x2).re - 32767 instead of - 32768
    x = np.array([[[100, 0], [0, -32768], [0, 0], [0, 0]]], np.int16)
    got = oracle.i16(x, 4, 1, [0])[0]
    # X1 = (x0 - x2) - j (x1 - x3): re = 100 + (-32768) with saturation -> 100 - 32767
    assert got[1].tolist() == [100 - 32767, 0]
    # X3 = (x0 - x2) + j (x1 - x3): re = 100 + 32767 -> wraps to 32867 - 65536
    assert got[3].tolist() == [100 + 32767 - 65536, 0]


@pytest.mark.parametrize("n,shifts", [(16, [1, 1]), (64, [0, 1, 1]), (256, [0, 1, 1, 1]), (512, [1, 1, 1, 1]),
                                      (1024, [1, 1, 1, 1, 1])])
def test_close_to_scaled_dft(n, shifts):
    r = np.random.Generator(np.random.Philox(n))
    x = r.integers(-2048, 2048, size=(2, n, 2)).astype(np.int16)
    y = oracle.i16(x, n, 2, shifts)
    xc = x[..., 0].astype(np.float64) + 1j * x[..., 1]
    ref = np.fft.fft(xc, axis=1) / (2.0 ** sum(shifts))
    yc = y[..., 0].astype(np.float64) + 1j * y[..., 1]
    err = np.max(np.abs(yc - ref))
    # One rounding (<= 1 unit) per stage, grown by at most 2 per later stage
    # (4 summed inputs, shift >= 1 or not), plus Q15 twiddle quantisation
    # (relative 2^-15 per stage). A wrong sign, twiddle or order errs by
    # O(max |ref|) instead.
    stages = len(shifts) + (1 if (n.bit_length() - 1) % 2 else 0)
    assert err <= 2.0 ** stages + 1e-4 * np.max(np.abs(ref)) * stages, err


def test_shift_zero_is_identity_and_modular_counts():
    r = np.random.Generator(np.random.Philox(3))
    x = r.integers(-32768, 32768, size=(3, 64, 2)).astype(np.int16)
    a = oracle.i16(x, 64, 3, [0, 0, 0])
    b = oracle.i16(x, 64, 3, [32, 32, 32])  # counts taken modulo 32
    assert np.array_equal(a, b)
