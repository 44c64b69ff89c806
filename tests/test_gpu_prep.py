"""Device-side read preprocessing (MQ cap, quality floor, GOP 45 lowered by
the PCR indel model, GCP 10) behind rvc_phmm_call_pairhmm: bit for bit the
same likelihoods and kept reads as the host preprocessing followed by the
plain region call. This module runs the engine without the serving loop (the
one-region path); test_gpu_parity.py covers the served path."""
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
    say("init device 0 (no serving loop, reads preprocessed on the device)")
    rv.init(0, host_prep=2)
    yield
    rv.shutdown()


def repeat_rich_region(seed):
    """A region whose reads carry homopolymers and short tandem repeats, so
    the PCR model caps many gap-open values."""
    rng = np.random.default_rng(seed)
    units = [b"A", b"CA", b"GTT", b"ACGT", b"TTAGG", b"C"]
    hap0 = bytearray()
    while len(hap0) < 300:
        if rng.random() < 0.3:
            hap0 += units[int(rng.integers(0, len(units)))] * int(rng.integers(2, 9))
        else:
            hap0 += bytes(workload.ACGT[rng.integers(0, 4, int(rng.integers(3, 12)))])
    hap0 = np.frombuffer(bytes(hap0[:300]), np.uint8)
    haps = [hap0] + [workload._mutate(rng, hap0) for _ in range(5)]
    reads, quals = [], []
    for x in range(24):
        h = haps[x % len(haps)]
        L = int(rng.integers(30, 160))
        o = int(rng.integers(0, len(h) - L))
        rd = h[o:o + L].copy()
        rd[rng.random(L) < 0.01] = ord("N")
        reads.append(rd)
        quals.append(rng.integers(2, 45, L).astype(np.uint8))
    mqs = rng.integers(0, 61, len(reads)).astype(np.uint8)
    return haps, reads, quals, mqs


def host_path(haps, reads, quals, mqs, min_q, pcr):
    """Host preprocessing of every read, the plain region call, host post-processing."""
    pre = [rv.preprocess_read(reads[r], quals[r], int(mqs[r]), min_q, pcr) for r in range(len(reads))]
    reg = workload.make_region(haps, reads, [p[0] for p in pre])
    reg["i"] = np.concatenate([p[1][:len(reads[r])] for r, p in enumerate(pre)] + [np.zeros(1, np.uint8)])
    reg["d"] = reg["i"].copy()
    reg["c"] = np.concatenate([p[2][:len(reads[r])] for r, p in enumerate(pre)] + [np.zeros(1, np.uint8)])
    rm, _ = rv.region_sync(reg)
    return rv.postprocess(rm, np.array([len(r) for r in reads], np.int32))


@pytest.mark.parametrize("pcr", [rv.PCR_NONE, rv.PCR_HOSTILE, rv.PCR_AGGRESSIVE, rv.PCR_CONSERVATIVE])
@pytest.mark.parametrize("min_q", [0, 18])
def test_device_prep_equals_host_prep(pcr, min_q):
    say(f"device preprocessing == host preprocessing, pcr {pcr}, min_q {min_q}")
    for seed in (1, 2):
        haps, reads, quals, mqs = repeat_rich_region(100 + seed)
        rc, mat, kept = rv.call_pairhmm(haps, reads, quals, mqs, min_quality_threshold=min_q, pcr=pcr)
        assert rc == 0
        want, wkept = host_path(haps, reads, quals, mqs, min_q, pcr)
        assert np.array_equal(kept, wkept)
        assert np.array_equal(np.isnan(mat), np.isnan(want))
        m = ~np.isnan(want)
        assert np.array_equal(mat[m], want[m])
