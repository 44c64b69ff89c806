"""The C ABI library: loads, exports every declared symbol, and its host-only
steps (tables, read preprocessing, post-processing) match the oracle. No
device work here."""
import ctypes
import os
import re

import numpy as np
import pytest

import oracle
import rovaca_amd as rv
from conftest import ROOT


def declared_functions():
    names = []
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if fn.endswith(".h"):
            text = open(os.path.join(ROOT, "include", fn)).read()
            names += re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(rvc_\w+)\s*\(", text, re.M)
    return sorted(set(names))


def test_library_exports_every_declared_symbol():
    L = rv.lib()
    names = declared_functions()
    assert len(names) >= 17
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing
    assert sorted(rv.EXPORTS) == names


def test_tables_bit_exact():
    pf, mf, pd, md = rv.tables()
    of, om = oracle.tables_f32()
    od, odm = oracle.tables_f64()
    assert np.array_equal(pf.view(np.uint32), of.view(np.uint32))
    assert np.array_equal(mf.view(np.uint32), om.view(np.uint32))
    assert np.array_equal(pd.view(np.uint64), od.view(np.uint64))
    assert np.array_equal(md.view(np.uint64), odm.view(np.uint64))


def test_tables_closed_form():
    pf, mf, pd, md = rv.tables()
    for q in (0, 10, 30, 45, 127):
        assert abs(pd[q] - 10 ** (-q / 10)) <= 1e-15 * max(1.0, 10 ** (-q / 10)) + 1e-300
    # match-to-match for GOP 45/45 is 1 - 2*10^-4.5 up to the 1e-4-step log table
    assert abs(md[45 * 46 // 2 + 45] - (1 - 2 * 10 ** -4.5)) < 1e-8


@pytest.mark.parametrize("pcr", [0, 1, 2, 3])
def test_preprocess_matches_oracle(pcr):
    rng = np.random.default_rng(100 + pcr)
    for trial in range(40):
        n = int(rng.integers(1, 160))
        # repeat-rich reads exercise the PCR model
        unit = np.frombuffer(b"ACGT", np.uint8)[rng.integers(0, 4, int(rng.integers(1, 5)))]
        bases = np.resize(unit, n).copy()
        flip = rng.random(n) < 0.2
        bases[flip] = np.frombuffer(b"ACGTN", np.uint8)[rng.integers(0, 5, int(flip.sum()))]
        qual = rng.integers(0, 60, n).astype(np.uint8)
        mq = int(rng.integers(0, 70))
        min_q = int(rng.integers(0, 30))
        q, g, c = rv.preprocess_read(bases, qual, mq, min_q, pcr)
        oq, oi, od, oc = oracle.preprocess_read(bases, qual, mq, min_q, pcr)
        assert np.array_equal(q, oq) and np.array_equal(g, oi) and np.array_equal(g, od) and np.array_equal(c, oc)


def test_preprocess_known_values():
    # no PCR model: GOP 45, GCP 10; quals capped at MQ then floored to 6 below the threshold
    q, g, c = rv.preprocess_read(b"ACGTACGT", np.array([10, 20, 30, 40, 50, 60, 5, 17], np.uint8), 40, 18, 0)
    assert list(q) == [6, 20, 30, 40, 40, 40, 6, 6]
    assert list(g) == [45] * 8 and list(c) == [10] * 8
    # a 12-base poly-A run under CONSERVATIVE: repeat length 12 caps GOP at 38, last base keeps 45
    q, g, c = rv.preprocess_read(b"A" * 12, np.full(12, 30, np.uint8), 60, 0, 3)
    assert g[-1] == 45 and all(x <= 40 for x in g[:-1])


def test_postprocess_matches_oracle():
    rng = np.random.default_rng(9)
    for trial in range(30):
        n, h = int(rng.integers(1, 30)), int(rng.integers(1, 10))
        rm = rng.uniform(-20, -0.1, (n, h))
        rm[rng.random((n, h)) < 0.05] = np.nan
        rl = rng.integers(0, 300, n).astype(np.int32)
        a, ka = rv.postprocess(rm, rl)
        b, kb = oracle.postprocess(rm, rl)
        assert np.array_equal(ka, kb)
        assert np.array_equal(np.isnan(a), np.isnan(b))
        assert np.array_equal(a[~np.isnan(a)], b[~np.isnan(b)])


def test_postprocess_closed_form():
    # read 0 is kept (best -1); read 1 (best -9 < -8 for 100 bp) is dropped; floor at best - 4.5
    rm = np.array([[-1.0, -10.0], [-9.0, -12.0]])
    out, kept = rv.postprocess(rm, np.array([100, 100], np.int32))
    assert list(kept) == [0]
    assert out[0, 0] == -1.0 and out[1, 0] == -5.5
    assert np.isnan(out[:, 1]).all()


def test_postprocess_drops_a_middle_read_and_packs_the_rest():
    # The reference erases a dropped read's row from the read-major matrix
    # (collect_poorly_modelled_indices, ph/pairhmm_internal.cpp:489-497) before
    # transpose_likelihood_matrix (:499-508) copies the remaining rows into the
    # first columns of the NaN-initialised [H][N] result (:581, :613): kept
    # reads are packed to the front, the last columns stay NaN.
    rm = np.array([[-1.0, -2.0], [-20.0, -30.0], [-3.0, -1.5]])
    for impl in (rv.postprocess, oracle.postprocess):
        out, kept = impl(rm, np.array([100, 100, 100], np.int32))
        assert list(kept) == [0, 2]
        assert out.shape == (2, 3)
        assert list(out[:, 0]) == [-1.0, -2.0]      # read 0 (nothing below best - 4.5 = -5.5)
        assert list(out[:, 1]) == [-3.0, -1.5]      # read 2, moved into column 1
        assert np.isnan(out[:, 2]).all()            # the dropped read's column is the last one


def test_postprocess_threshold_for_short_reads():
    # min(2, ceil(0.02 R)) * -4.0 (ph/pairhmm_internal.cpp:38-41, ph/adapter.h:20-22):
    # -4.0 up to R = 50, -8.0 from R = 51; an empty read (all NaN) has best -inf
    # below its threshold 0 and is dropped
    cases = [(50, -4.01, False), (50, -3.99, True), (1, -4.01, False), (51, -7.99, True), (51, -8.01, False),
             (300, -7.99, True)]
    for R, best, keep in cases:
        rm = np.array([[best, best - 1.0]])
        for impl in (rv.postprocess, oracle.postprocess):
            out, kept = impl(rm, np.array([R], np.int32))
            assert (list(kept) == [0]) == keep, (R, best)
    rm = np.array([[np.nan, np.nan], [-1.0, -2.0]])
    for impl in (rv.postprocess, oracle.postprocess):
        out, kept = impl(rm, np.array([0, 10], np.int32))
        assert list(kept) == [1]


def test_postprocess_nan_first_row():
    # normalize takes the best with std::max_element (ph/pairhmm_internal.cpp:248-258):
    # a NaN in front compares false against everything and stays the "best",
    # so nothing in that row is floored; a NaN further on is skipped. The
    # drop test ignores NaNs either way (:33-37).
    rm = np.array([[np.nan, -10.0, -1.0], [-10.0, np.nan, -1.0]])
    for impl in (rv.postprocess, oracle.postprocess):
        out, kept = impl(rm, np.array([100, 100], np.int32))
        assert list(kept) == [0, 1]
        assert np.isnan(out[0, 0]) and out[1, 0] == -10.0 and out[2, 0] == -1.0
        assert out[0, 1] == -5.5 and np.isnan(out[1, 1]) and out[2, 1] == -1.0


def test_calls_fail_cleanly_without_init():
    reg = {"n_haps": 1, "n_reads": 1, "hap_off": np.array([0, 4], np.int32),
           "hap_bases": np.frombuffer(b"ACGT", np.uint8).copy(), "read_off": np.array([0, 4], np.int32),
           "read_bases": np.frombuffer(b"ACGT", np.uint8).copy()}
    for k in "qidc":
        reg[k] = np.full(4, 30, np.uint8)
    with pytest.raises(rv.EngineError) as e:
        rv.region_sync(reg)
    assert e.value.code == rv.RVC_E_STATE


def test_struct_layouts_match_the_headers():
    # the ctypes mirrors of the ABI's structs against the C compiler's view of
    # the headers: same field names, offsets and sizes
    import shutil
    import subprocess
    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        pytest.skip("no C compiler")
    structs = {"rvc_region": rv.RvcRegion, "rvc_raw_region": rv.RvcRawRegion, "rvc_config": rv.RvcConfig,
               "rvc_batch_opts": rv.RvcBatchOpts, "rvc_batch_info": rv.RvcBatchInfo,
               "rvc_kernel_time": rv.RvcKernelTime, "rvc_server_stats": rv.RvcServerStats,
               "rvc_loadgen_result": rv.RvcLoadgenResult, "rvc_sw_stats": rv.RvcSwStats,
               "rvc_sw_region": rv.RvcSwRegion}
    lines = ["#include <stddef.h>", "#include <stdio.h>", '#include "rovaca_phmm_loadgen.h"', "int main(void) {"]
    for name, py in structs.items():
        lines.append(f'  printf("{name} %zu\\n", sizeof({name}));')
        for field, _ in py._fields_:
            lines.append(f'  printf("{name}.{field} %zu\\n", offsetof({name}, {field}));')
    lines += ["  return 0;", "}"]
    out_dir = os.path.join(ROOT, "rovaca_amd", "_build")
    os.makedirs(out_dir, exist_ok=True)
    src, exe = os.path.join(out_dir, "abi_layout.c"), os.path.join(out_dir, "abi_layout")
    with open(src, "w") as f:
        f.write("\n".join(lines) + "\n")
    subprocess.run([cc, "-std=c11", "-I", os.path.join(ROOT, "include"), "-o", exe, src], check=True)
    c = dict(line.rsplit(" ", 1) for line in subprocess.run([exe], check=True, capture_output=True,
                                                              text=True).stdout.split("\n") if line)
    for name, py in structs.items():
        assert int(c[name]) == ctypes.sizeof(py), name
        for field, _ in py._fields_:
            assert int(c[f"{name}.{field}"]) == getattr(py, field).offset, (name, field)
