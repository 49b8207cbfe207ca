"""Host-side checks of the test-list runner (c7x_fftlib_amd.testlists): the
reference's list and header formats, case shapes, and the perf_data CSV
schema. Where the reference tree is present (not on GPU machines) its own
headers and perf files are read as well."""
import os

import numpy as np
import pytest

import c7x_fftlib_amd as F
from c7x_fftlib_amd import testlists as T

HERE = os.path.dirname(os.path.abspath(__file__))
LISTS = os.path.join(HERE, "golden", "testlists")
REF = "/root/reference"
ROWS = {"FFTLIB_fft1dBatched_i16sc_c16sc_o16sc": 40, "FFTLIB_fft1dBatched_i32f_c32fc_o32fc": 236,
        "FFTLIB_fft1dBatched_i32fc_c32fc_o32fc": 182, "FFTLIB_fft1d_i16sc_c16sc_o16sc": 14,
        "FFTLIB_fft1d_i32f_c32fc_o32fc": 101, "FFTLIB_fft1d_i32fc_c32fc_o32fc": 109,
        "FFTLIB_ifft1dBatched_i32fc_c32fc_o32f": 236, "FFTLIB_ifft1dBatched_i32fc_c32fc_o32fc": 182,
        "FFTLIB_ifft1d_i32fc_c32fc_o32f": 104, "FFTLIB_ifft1d_i32fc_c32fc_o32fc": 109}


@pytest.mark.parametrize("kernel", sorted(T.KERNELS))
def test_read_committed_test_lists(kernel):
    cases = T.read_test_cases(os.path.join(LISTS, kernel + ".csv"))
    assert len(cases) == ROWS[kernel]
    assert len({c.id for c in cases}) == len(cases)
    for c in cases:
        assert c.numPoints > 0 and c.numChannels > 0
        assert c.testType in ("STATIC", "RANDOM", "RANDOM_SIGNED")
        if T.KERNELS[kernel].kind == "i16":
            assert 1 <= len(c.shifts) <= 8, c
        if not T.KERNELS[kernel].batched:
            assert c.numChannels == 1


def test_read_test_cases_fields(tmp_path):
    p = tmp_path / "list.csv"
    p.write_text('ID,numChannels,numPoints,numShifts,shiftVector,testType,testCategory\n'
                 '7,4,256,4,"[0,1,1,1]",STATIC,200\n9,2,512,3,"[1,0,1,1,1]",RANDOM_SIGNED,201\n')
    a, b = T.read_test_cases(str(p))
    assert (a.id, a.numChannels, a.numPoints, a.shifts, a.testType, a.testCategory) == \
        (7, 4, 256, [0, 1, 1, 1], "STATIC", 200)
    assert b.shifts == [1, 0, 1] and b.testCategory == 201       # numShifts entries are used
    p.write_text("ID,numChannels,numPoints,inMode,testType,testCategory\n3,2,128,1,RANDOM,200\n")
    (c,) = T.read_test_cases(str(p))
    assert (c.mode, c.numChannels) == (1, 2)


def test_read_static_ref_case_format(tmp_path):
    """The header layout the reference's generator writes (attribute,
    static C array, values over several lines, trailing comma)."""
    p = tmp_path / "staticRefCase5.h"
    p.write_text("// header comment, 1.0, 2.0\n\n"
                 '__attribute__((section(".staticData"))) static float staticRefIn5[] = {\n'
                 "    0.756, -0.792, 3e-3,\n    1.5,\n\n};\n\n"
                 '__attribute__((section(".staticData"))) static float staticRefOut5[] = {\n'
                 "    32.419, 0.0, -1.0, 2.0,};\n")
    d = T.read_static_ref_case(str(p))
    assert d["in"].dtype == np.float32 and d["out"].dtype == np.float32
    assert np.array_equal(d["in"], np.asarray([0.756, -0.792, 3e-3, 1.5], np.float32))
    assert np.array_equal(d["out"], np.asarray([32.419, 0.0, -1.0, 2.0], np.float32))
    p.write_text("static int16_t staticRefIn1[] = {1, -2, 3, 4};\nstatic int16_t staticRefOut1[] = {-32768, 5};\n")
    d = T.read_static_ref_case(str(p))
    assert d["in"].dtype == np.int16 and list(d["out"]) == [-32768, 5]


@pytest.mark.skipif(not os.path.isdir(REF), reason="reference tree not present")
@pytest.mark.parametrize("alias,kernel", [("r2cb", "FFTLIB_fft1dBatched_i32f_c32fc_o32fc"),
                                          ("c2rb", "FFTLIB_ifft1dBatched_i32fc_c32fc_o32f"),
                                          ("r2c1", "FFTLIB_fft1d_i32f_c32fc_o32fc"),
                                          ("c2r1", "FFTLIB_ifft1d_i32fc_c32fc_o32f"),
                                          ("ic2c1", "FFTLIB_ifft1d_i32fc_c32fc_o32fc")])
def test_reference_headers_match_committed_fixtures(alias, kernel):
    """The reader, pointed at the reference's own headers, finds every case
    of the committed fixtures (including the inverse real batched list whose
    headers are numbered one above their rows) with identical values."""
    d = np.load(os.path.join(HERE, "golden", f"ref_{alias}.npz"))
    cases = {c.id: c for c in T.read_test_cases(os.path.join(LISTS, kernel + ".csv"))}
    static_dir = os.path.join(REF, "test/fft_c7x", kernel, "test_data/extended_test_cases")
    found = 0
    for cid in d["case_ids"]:
        x, y = d[f"c{cid}_x"], d[f"c{cid}_y"]
        row = cid if alias != "c2rb" else cid - 1
        got = T.find_static(static_dir, kernel, cases[int(row)])
        assert got is not None, cid
        assert np.array_equal(got[0], x) and np.array_equal(got[1], y), cid
        found += 1
    assert found == len(d["case_ids"])


@pytest.mark.parametrize("kernel", sorted(T.KERNELS))
def test_case_data_shapes(kernel):
    """Buffers of every listed case have the sizes the kernel's C signature
    implies (checked on the first cases of each list; the library's table
    generators run on the CPU)."""
    cases = T.read_test_cases(os.path.join(LISTS, kernel + ".csv"))[:6]
    info = T.KERNELS[kernel]
    for c in cases:
        d = T.make_case_data(kernel, c)
        nin, nout = T.shape_of(kernel, c)
        assert d.x.size == nin and d.out_count == nout
        assert d.x.dtype == (np.int16 if info.kind == "i16" else np.float32)
        if info.kind == "i16":
            assert list(d.aux) == c.shifts
        elif info.kind in ("r2c", "c2r"):
            assert d.aux.size >= 2 * c.numPoints
        if c.testType != "STATIC" and info.kind != "i16":
            assert np.max(np.abs(d.x)) <= info.random_range


@pytest.mark.parametrize("kernel", sorted(T.KERNELS))
def test_perf_csv_schema(kernel, tmp_path):
    cases = T.read_test_cases(os.path.join(LISTS, kernel + ".csv"))[:3]
    res = [T.CaseResult(c.id, 0, warm_ns=1234.4 + i) for i, c in enumerate(cases)]
    res.append(T.CaseResult(999, 5))                                 # failed cases are not written
    out = tmp_path / "perf.csv"
    T.write_perf_csv(str(out), kernel, cases, res)
    lines = out.read_text().splitlines()
    head = lines[0].split(",")
    assert head[:-1] == T.KERNELS[kernel].perf_columns and head[-1] == T.TIME_COLUMN
    assert len(lines) == 4
    first = [v.strip() for v in lines[1].split(",")]
    assert first[0] == str(cases[0].id) and first[-1] == "1234"
    ref_csv = os.path.join(REF, "docs/sphinx/perf_data/c7504", kernel + ".csv")
    if os.path.exists(ref_csv):
        with open(ref_csv) as fh:
            ref_head = fh.readline().strip().split(",")
        assert ref_head[:-1] == head[:-1]                            # same columns; time replaces EVM Cycles


def test_unknown_kernel_is_refused():
    with pytest.raises(KeyError):
        T.make_case_data("FFTLIB_nope", T.TestCase(1, 64))
    assert F.load() is not None
