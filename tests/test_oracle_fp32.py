"""Pins the fp32 oracle (C restatement of the reference natural-C kernels).

Checks that do not restate the oracle's formulas:
  * the FP64 DFT as a dense matrix product (closed form) for every radix mix;
  * textbook invariants: impulse -> flat spectrum, single tone -> delta,
    linearity, Parseval, inverse(forward(x)) = x, Hermitian symmetry of a real
    input's spectrum;
  * the reference's committed input/output fixtures for R2C, C2R and the
    single-channel IFFT, at the reference drivers' own tolerances;
  * twiddle and split-factor tables against cos/sin of their closed-form
    angles;
  * a planted mistake (a conjugated twiddle table) is caught by the DFT check.
"""
import os

import numpy as np
import pytest

import oracle

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def rng(seed):
    return np.random.Generator(np.random.Philox(seed))


def cplx(r, b, n, s=1.0):
    return (r.uniform(-s, s, (b, n)) + 1j * r.uniform(-s, s, (b, n))).astype(np.complex64)


def dft_matrix(n, sign=-1):
    k = np.arange(n)
    return np.exp(sign * 2j * np.pi * np.outer(k, k) / n)


def rel(a, b):
    return np.linalg.norm(np.asarray(a, np.complex128) - b) / np.linalg.norm(b)


SMALL = [1, 2, 3, 4, 5, 6, 8, 9, 10, 12, 15, 16, 20, 25, 27, 30, 32, 45, 48, 60, 64, 75, 81, 96, 100, 120, 125, 128,
         243, 250, 256, 384, 500, 512]


@pytest.mark.parametrize("n", SMALL)
def test_c2c_matches_dense_dft(n):
    x = cplx(rng(n), 3, n)
    y = oracle.c2c(x, n, 3)
    ref = x.astype(np.complex128) @ dft_matrix(n).T
    assert rel(y, ref) < 4e-7
    yi = oracle.c2c(x, n, 3, inverse=True)
    refi = x.astype(np.complex128) @ dft_matrix(n, +1).T / n
    assert rel(yi, refi) < 4e-7
    ys = oracle.c2c(x[:1], n, 1, single=True)
    assert rel(ys, ref[:1]) < 4e-7
    yis = oracle.c2c(x[:1], n, 1, inverse=True, single=True)
    assert rel(yis, refi[:1]) < 4e-7


@pytest.mark.parametrize("n", [960, 1024, 1536, 3125, 4096, 6000, 8192, 15625])
def test_c2c_large_vs_fp64_fft(n):
    x = cplx(rng(n + 1), 2, n)
    assert rel(oracle.c2c(x, n, 2), np.fft.fft(x.astype(np.complex128), axis=1)) < 5e-7
    assert rel(oracle.c2c(x, n, 2, inverse=True), np.fft.ifft(x.astype(np.complex128), axis=1)) < 5e-7


@pytest.mark.parametrize("n", [8, 60, 64, 960])
def test_c2c_invariants(n):
    b = 2
    imp = np.zeros((b, n), np.complex64)
    imp[:, 0] = 1.0
    assert np.allclose(oracle.c2c(imp, n, b), 1.0, atol=1e-6)          # impulse -> flat
    k0 = 3 % n
    tone = np.exp(2j * np.pi * k0 * np.arange(n) / n).astype(np.complex64)[None, :].repeat(b, 0)
    spec = oracle.c2c(tone, n, b)
    expect = np.zeros(n)
    expect[k0] = n
    assert np.allclose(np.abs(spec), expect[None, :], atol=2e-4 * n)   # tone -> delta at k0
    r = rng(n)
    x1, x2 = cplx(r, b, n), cplx(r, b, n)
    lin = oracle.c2c(2.0 * x1 + x2, n, b)
    assert rel(lin, 2.0 * oracle.c2c(x1, n, b).astype(np.complex128) + oracle.c2c(x2, n, b)) < 1e-6
    y = oracle.c2c(x1, n, b)
    assert abs(np.sum(np.abs(y) ** 2) / n - np.sum(np.abs(x1) ** 2)) / np.sum(np.abs(x1) ** 2) < 1e-6  # Parseval
    back = oracle.c2c(oracle.c2c(x1, n, b), n, b, inverse=True)
    assert rel(back, x1) < 1e-6


def test_two_inverse_scalings_agree_to_rounding():
    # batched IFFT multiplies by (1.0f/N) first, single-channel IFFT divides by
    # N at the end: same value up to float rounding
    n = 960
    x = cplx(rng(5), 1, n)
    a = oracle.c2c(x, n, 1, inverse=True)
    b = oracle.c2c(x, n, 1, inverse=True, single=True)
    assert rel(a, b) < 3e-7


@pytest.mark.parametrize("n", [2, 4, 10, 50, 64, 250, 1024, 1250, 4096])
def test_real_transforms_vs_fp64(n):
    b = 3
    x = rng(n).standard_normal((b, n)).astype(np.float32)
    ref = np.fft.rfft(x.astype(np.float64), axis=1)
    y0 = oracle.r2c(x, n, b, outmode=0)
    assert rel(y0, ref) < 5e-7
    y1 = oracle.r2c(x, n, b, outmode=1)
    full = np.fft.fft(x.astype(np.float64), axis=1)
    assert rel(y1, full) < 5e-7
    ys = oracle.r2c(x[:1], n, 1, single=True)
    assert rel(ys, full[:1]) < 5e-7
    back = oracle.c2r(y0, n, b, inmode=0)
    assert rel(back, x.astype(np.float64)) < 1e-6
    back1 = oracle.c2r(y1, n, b, inmode=1)
    assert rel(back1, x.astype(np.float64)) < 1e-6
    backs = oracle.c2r(ys, n, 1, single=True)
    assert rel(backs, x[:1].astype(np.float64)) < 1e-6


def test_full_spectrum_is_hermitian():
    n = 128
    y = oracle.r2c(rng(1).standard_normal((1, n)).astype(np.float32), n, 1, outmode=1)[0]
    assert np.allclose(y[1:][::-1], np.conj(y[1:]), atol=1e-5)


def _fixture_cases(alias):
    d = np.load(os.path.join(GOLDEN, f"ref_{alias}.npz"))
    return [(int(c), d[f"c{c}_x"], d[f"c{c}_y"], d[f"c{c}_meta"]) for c in d["case_ids"]]


# tolerances of the reference drivers' comparisons against their static
# references (test/fft_c7x/<kernel>/<kernel>_d.c)
@pytest.mark.parametrize("case", _fixture_cases("r2cb"), ids=lambda c: f"case{c[0]}")
def test_fixture_r2c_batched(case):
    _, x, y, (n, b, om) = case
    out = oracle.r2c(x.reshape(b, n), int(n), int(b), outmode=int(om))
    assert np.max(np.abs(out.view(np.float32).reshape(-1) - y)) <= 2.0 ** -3


@pytest.mark.parametrize("case", _fixture_cases("c2rb"), ids=lambda c: f"case{c[0]}")
def test_fixture_c2r_batched(case):
    _, x, y, (n, b, im) = case
    out = oracle.c2r(x.view(np.complex64).reshape(int(b), -1), int(n), int(b), inmode=int(im))
    assert np.max(np.abs(out.reshape(-1) - y)) <= 1e-3


@pytest.mark.parametrize("case", _fixture_cases("r2c1"), ids=lambda c: f"case{c[0]}")
def test_fixture_r2c_single(case):
    _, x, y, (n, _b, _m) = case
    out = oracle.r2c(x.reshape(1, n), int(n), 1, single=True)
    assert np.max(np.abs(out.view(np.float32).reshape(-1) - y)) <= 2.0 ** -3


@pytest.mark.parametrize("case", _fixture_cases("c2r1"), ids=lambda c: f"case{c[0]}")
def test_fixture_c2r_single(case):
    _, x, y, (n, _b, _m) = case
    out = oracle.c2r(x.view(np.complex64).reshape(1, -1), int(n), 1, single=True)
    assert np.max(np.abs(out.reshape(-1) - y)) <= 1e-1


@pytest.mark.parametrize("case", _fixture_cases("ic2c1"), ids=lambda c: f"case{c[0]}")
def test_fixture_ifft_single(case):
    _, x, y, (n, _b, _m) = case
    out = oracle.c2c(x.view(np.complex64).reshape(1, -1), int(n), 1, inverse=True, single=True)
    assert np.max(np.abs(out.view(np.float32).reshape(-1) - y)) <= 1e-3


def _stage_blocks(n):
    """(radix, M, complex offset) of every twiddle block: radix-5 stages, then
    radix-3, then radix-4; a block holds (r-1)*(M/r) values."""
    out, off, m = [], 0, n
    for r in (5, 3, 4):
        while m % r == 0 and (r != 4 or m >= 4):
            out.append((r, m, off))
            off += (r - 1) * (m // r)
            m //= r
    return out


@pytest.mark.parametrize("n", [4, 8, 15, 60, 64, 960, 1000, 1024, 6250])
def test_c2c_twiddle_table_closed_form(n):
    w = oracle.tw_c2c(n).view(np.complex64)
    for r, m, off in _stage_blocks(n):
        q = m // r
        for mm in range(1, r):
            p = np.arange(q)
            expect = np.exp(-2j * np.pi * mm * p / m)
            got = w[off + (mm - 1) * q: off + mm * q]
            assert np.max(np.abs(got - expect)) < 2e-7, (r, m, mm)


@pytest.mark.parametrize("n", [4, 8, 16, 64, 256, 1024, 2048])
def test_i16_twiddle_table_closed_form(n):
    w = oracle.tw_i16(n).reshape(-1, 2).astype(np.int64)
    off, m = 0, n
    while m >= 4:
        q = m // 4
        for mm in range(1, 4):
            p = np.arange(q)
            ang = 2 * np.pi * mm * p / m
            exp_re = np.clip(np.floor(0.5 + 32767.5 * np.cos(ang)), -32768, 32767)
            exp_im = np.clip(np.floor(0.5 - 32767.5 * np.sin(ang)), -32768, 32767)
            got = w[off + (mm - 1) * q: off + mm * q]
            assert np.max(np.abs(got[:, 0] - exp_re)) <= 1 and np.max(np.abs(got[:, 1] - exp_im)) <= 1
        off += 3 * q
        m //= 4


@pytest.mark.parametrize("n", [8, 64, 1000, 4096])
def test_split_factors_closed_form(n):
    i = np.arange(n // 2)
    ang = 2 * np.pi * i / n
    s = oracle.sf_r2c(n)
    assert np.allclose(s[2 * i], 0.5 * (1 - np.sin(ang)), atol=1e-7)
    assert np.allclose(s[2 * i + 1], -0.5 * np.cos(ang), atol=1e-7)
    assert np.allclose(s[2 * n - 2 * i - 2], 0.5 * (1 + np.sin(ang)), atol=1e-7)
    assert np.allclose(s[2 * n - 2 * i - 1], 0.5 * np.cos(ang), atol=1e-7)
    c = oracle.sf_c2r(n)
    assert np.allclose(c[2 * i + 1], 0.5 * np.cos(ang), atol=1e-7)
    assert np.allclose(c[n + 2 * i], 0.5 * (1 + np.sin(ang)), atol=1e-7)


def test_planted_twiddle_mistake_is_caught():
    # a conjugated twiddle table (a plausible sign slip) must fail the DFT check
    n = 60
    x = cplx(rng(9), 1, n)
    bad = oracle.tw_c2c(n).copy()
    bad[1::2] *= -1
    y = oracle.c2c(x, n, 1, w=bad)
    assert rel(y, np.fft.fft(x.astype(np.complex128), axis=1)) > 1e-2
