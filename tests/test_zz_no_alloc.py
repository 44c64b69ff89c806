"""The one-region entry point that allocates no device memory
(rvc_phmm_region_sync_noalloc) against rvc_phmm_region_sync: the same
matrix and flags, bit for bit, on calls in a row over the shared device
scratch array (the later ones smaller, so stale contents would show), and
the size limit refused with RVC_E_LIMIT."""
import sys

import numpy as np
import pytest

import rovaca_amd as rv
from rovaca_amd import workload

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[gpu-test] {msg}", flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def engine():
    say("init device 0")
    rv.init(0)
    yield
    say("shutdown")
    rv.shutdown()


def same(a, b):
    assert np.array_equal(np.isnan(a), np.isnan(b))
    assert np.array_equal(a[~np.isnan(a)], b[~np.isnan(b)])


def test_noalloc_equals_region_sync_on_calls_in_a_row():
    big = workload.single_region(0)                       # 8 x 32, reads of 150 bases
    small = workload.batched_regions(6, seed=7)
    small = min(small, key=lambda r: r["n_haps"] * r["n_reads"])
    stress = [r for r in workload.stress_regions(8, seed=12) if r["n_haps"] * r["n_reads"] < 3000][0]
    for name, reg in (("8 x 32 region", big), ("smaller region", small), ("precision-stress region", stress),
                      ("the 8 x 32 region again", big)):
        say(f"no-allocation call: {name} ({reg['n_haps']} x {reg['n_reads']})")
        got, gf = rv.region_sync_noalloc(reg)
        want, wf = rv.region_sync(reg)
        same(got, want)
        assert np.array_equal(gf, wf)
    assert (rv.region_sync_noalloc(stress)[1] & rv.PAIR_FP64).any()


def test_noalloc_refuses_a_region_beyond_its_scratch():
    say("no-allocation call: a region beyond the 32 MiB scratch is refused")
    rng = np.random.default_rng(3)
    haps = [workload.ACGT[rng.integers(0, 4, 8)] for _ in range(4096)]
    reads = [workload.ACGT[rng.integers(0, 4, 8)] for _ in range(600)]
    reg = workload.make_region(haps, reads, [np.full(8, 30, np.uint8)] * len(reads))
    with pytest.raises(rv.EngineError) as e:
        rv.region_sync_noalloc(reg)
    assert e.value.code == rv.RVC_E_LIMIT
    # a normal call still works after the refusal
    reg2 = workload.single_region(1)
    same(rv.region_sync_noalloc(reg2)[0], rv.region_sync(reg2)[0])
