"""Every case of the reference's ten fft_c7x test lists (committed copies of
test/fft_c7x/FFTLIB_idat_gen/<kernel>/test_cases_list.csv) run unchanged on
the GPU through the test-list runner, and each output compared with the
oracle on the same input:

  * int16 bit-exact;
  * fp32 relative L2 <= 1e-6, and the maximum absolute difference within
    the kernel's driver tolerance or 2e-6 of the output's peak magnitude,
    whichever is larger (the drivers' absolute tolerances assume their own
    input ranges);
  * STATIC cases with a committed reference header (tests/golden/ref_*.npz)
    also against the header's expected output with the driver's tolerance.
"""
import os

import numpy as np
import pytest

import oracle
from c7x_fftlib_amd import testlists as T

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURES = {"FFTLIB_fft1dBatched_i32f_c32fc_o32fc": "r2cb", "FFTLIB_ifft1dBatched_i32fc_c32fc_o32f": "c2rb",
            "FFTLIB_fft1d_i32f_c32fc_o32fc": "r2c1", "FFTLIB_ifft1d_i32fc_c32fc_o32f": "c2r1",
            "FFTLIB_ifft1d_i32fc_c32fc_o32fc": "ic2c1"}


def static_cases(kernel):
    """{row id: (input, expected)} from the committed reference headers."""
    alias = FIXTURES.get(kernel)
    if alias is None:
        return {}
    d = np.load(os.path.join(HERE, "golden", f"ref_{alias}.npz"))
    off = 1 if alias == "c2rb" else 0          # headers numbered one above their rows
    return {int(c) - off: (d[f"c{c}_x"], d[f"c{c}_y"]) for c in d["case_ids"]}


def oracle_out(kernel, case, data):
    info = T.KERNELS[kernel]
    n, b = case.numPoints, (case.numChannels if info.batched else 1)
    single = not info.batched
    if info.kind in ("c2c", "ic2c"):
        x = data.x.view(np.complex64).reshape(b, n)
        return oracle.c2c(x, n, b, inverse=info.kind == "ic2c", single=single).view(np.float32).reshape(-1)
    if info.kind == "i16":
        return oracle.i16(data.x.reshape(b, n, 2), n, b, data.aux, single=single).reshape(-1)
    if info.kind == "r2c":
        y = oracle.r2c(data.x.reshape(b, n), n, b, outmode=case.mode, single=single)
        return np.ascontiguousarray(y).view(np.float32).reshape(-1)
    nin = n if (case.mode or single) else n // 2 + 1
    x = data.x.view(np.complex64).reshape(b, nin)
    return oracle.c2r(x, n, b, inmode=case.mode, single=single).reshape(-1)


@pytest.mark.parametrize("kernel", sorted(T.KERNELS))
def test_reference_test_list(kernel):
    info = T.KERNELS[kernel]
    cases = T.read_test_cases(os.path.join(HERE, "golden", "testlists", kernel + ".csv"))
    static = static_cases(kernel)
    checked_static = 0
    for case in cases:
        data = T.make_case_data(kernel, case, static=static.get(case.id))
        res = T.run_case(kernel, case, data, warm=0)
        assert res.status == 0, (case, res.error)
        want = oracle_out(kernel, case, data)
        y = res.y
        assert y.shape == want.shape, case
        if info.kind == "i16":
            assert np.array_equal(y, want), case
        else:
            diff = np.abs(y.astype(np.float64) - want.astype(np.float64))
            den = np.linalg.norm(want.astype(np.float64))
            assert np.linalg.norm(diff) <= 1e-6 * (den if den > 0 else 1.0), case
            peak = float(np.max(np.abs(want))) if want.size else 0.0
            assert float(np.max(diff)) <= max(info.tol_expected, 2e-6 * peak), (case, float(np.max(diff)))
        chk = T.check_expected(kernel, data, y)
        if chk is not None:
            assert chk[0], (case, chk[1])
            checked_static += 1
    assert checked_static == len(static)
