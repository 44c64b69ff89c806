"""GPU parity of the batched Smith-Waterman (read against haplotype) with the
oracle, which the CPU tests pin to Rovaca's own aligner: CIGAR, offset and
score identical."""
import sys

import numpy as np
import pytest

import oracle
import rovaca_amd as rv
from test_oracle_sw import ACGT, _mutate, pairs_with_other_bytes, random_pairs, reference_vectors

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[gpu-test] {msg}", flush=True)
    sys.stdout.flush()


# the module runs twice: on the base sweep shapes only (16, 32 or 64 lanes
# by read length) and on the wide ones only (64 lanes, 2-6 rows per lane);
# on its own the engine picks the wide ones for passes too small to fill the
# device, which every batch here but one is
WIDE = {"base": 1, "wide": 2}
_mode = {"wide": 1}


@pytest.fixture(scope="module", autouse=True, params=sorted(WIDE))
def engine(request):
    # every sweep shape launches on its own, however few its pairs, so that
    # small batches reach every kernel
    _mode["wide"] = WIDE[request.param]
    say(f"init device 0, {request.param} shapes")
    rv.init(0, sw_min_launch_pairs=1, sw_wide_shapes=_mode["wide"])
    yield
    say("shutdown")
    rv.shutdown()


def reinit(**kw):
    rv.shutdown()
    rv.init(0, sw_min_launch_pairs=1, sw_wide_shapes=_mode["wide"], **kw)


def check(haps, reads, read_hap, assemble=False):
    got = rv.sw_align_batch(haps, reads, read_hap, assemble)
    for r, (g, read) in enumerate(zip(got, reads)):
        want = oracle.sw_align(haps[read_hap[r]], read, assemble)
        assert g == want, (r, len(haps[read_hap[r]]), len(read), g, want)
    return got


@pytest.mark.parametrize("assemble", [False, True])
def test_random_pairs(assemble):
    say(f"random pairs, assemble={assemble}")
    pairs = random_pairs(21 + assemble, 600)
    haps = [h for h, _ in pairs]
    reads = [r for _, r in pairs]
    check(haps, reads, list(range(len(pairs))), assemble)


@pytest.mark.parametrize("assemble", [False, True])
def test_committed_reference_vectors(assemble):
    # the device against the outputs of Rovaca's own aligner stored with
    # their inputs (no oracle involved)
    say(f"committed reference vectors, assemble={assemble}")
    haps, reads, want = reference_vectors(assemble)
    got = rv.sw_align_batch(haps, reads, list(range(len(reads))), assemble)
    for r, (g, w) in enumerate(zip(got, want)):
        assert g == w, (r, len(haps[r]), len(reads[r]), g, w)


def test_every_shape_and_its_edges():
    say("read lengths at the sweep shapes' edges")
    rng = np.random.default_rng(3)
    haps, reads, read_hap = [], [], []
    # the sweep shapes hold 64, 112, 128, 160, 192, 224, 256, 320, 384, 512,
    # 768 and 1024 rows
    edges = (64, 112, 128, 160, 192, 224, 256, 320, 384, 512, 768)
    for L in (1, 2, 15, 16, 17) + tuple(x + d for x in edges for d in (-1, 0, 1)) + (1000, 1023, 1024):
        hap = ACGT[rng.integers(0, 4, L + 200)].tobytes()
        haps.append(hap)
        for _ in range(2):
            a = int(rng.integers(0, 150))
            reads.append((_mutate(rng, hap[a:a + L], int(rng.integers(0, 4))) or b"C")[:L])  # insertions stay within L
            read_hap.append(len(haps) - 1)
        reads.append(ACGT[rng.integers(0, 4, L)].tobytes())
        read_hap.append(len(haps) - 1)
    check(haps, reads, read_hap)


def test_haplotype_lengths_and_refusals():
    say("haplotypes up to 2047 bases; 2048 refused")
    rng = np.random.default_rng(4)
    haps = [ACGT[rng.integers(0, 4, n)].tobytes() for n in (1, 2, 500, 2047, 2048)]
    reads = [haps[2][10:160], haps[3][1900:2040], b"ACGTACGT", haps[3][:100], haps[4][:100], b"A", b"G" * 30]
    read_hap = [2, 3, 0, 3, 4, 1, 3]
    got = check(haps, reads, read_hap)
    assert got[4] == (None, 1, None)


def test_empty_haplotypes_and_reads():
    say("empty haplotypes and reads: not refused, as the reference library")
    haps = [b"", b"ACGTACGTAC", b"A" * 2047]
    reads = [b"ACGT", b"", b"", b"ACGTACGTACGTACGTACGTAA", b"ACGTAC", b"", b"GT" * 40]
    read_hap = [0, 1, 0, 0, 1, 2, 0]
    for served in (False, True):
        got = rv.sw_align_batch(haps, reads, read_hap, served=served)
        for r, g in enumerate(got):
            assert g == oracle.sw_align(haps[read_hap[r]], reads[r]), (r, g)
    assert got[0] == ([("S", 4)], 0, 0) and got[1] == ([], 10, 0)


def test_shared_haplotypes_and_batch_independence():
    say("many reads per haplotype; results do not depend on the batch")
    rng = np.random.default_rng(5)
    haps = [ACGT[rng.integers(0, 4, int(rng.integers(300, 600)))].tobytes() for _ in range(8)]
    reads, read_hap = [], []
    for _ in range(400):
        h = int(rng.integers(0, 8))
        L = int(rng.choice([100, 150, 250]))
        a = int(rng.integers(0, len(haps[h]) - L))
        reads.append(_mutate(rng, haps[h][a:a + L], int(rng.integers(0, 5))) or b"T")
        read_hap.append(h)
    whole = check(haps, reads, read_hap)
    part = rv.sw_align_batch(haps, reads[123:170], read_hap[123:170])
    assert part == whole[123:170]


@pytest.mark.parametrize("assemble", [False, True])
def test_long_reads_1025_to_2047(assemble):
    # the reference aligns every length below 2048 (sw_core_template.h:356-361):
    # reads past the 1024 rows of the regular shapes take the 2048-row one
    say(f"reads of 1025-2047 bases, assemble={assemble}")
    rng = np.random.default_rng(1025 + assemble)
    haps = [ACGT[rng.integers(0, 4, n)].tobytes() for n in (2047, 1800, 1100, 600)]
    reads, read_hap = [], []
    for n in (1024, 1025, 1026, 1300, 1536, 2000, 2047):
        h = int(rng.integers(0, 2))
        src = haps[h]
        o = int(rng.integers(0, len(src) - min(n, len(src)) + 1))
        rd = bytearray(src[o:o + n])
        rd += ACGT[rng.integers(0, 4, n - len(rd))].tobytes()
        for x in rng.integers(0, n, n // 50):
            rd[x] = ACGT[rng.integers(0, 4)]
        reads.append(bytes(rd))
        read_hap.append(h)
    reads.append(haps[3][100:500] + ACGT[rng.integers(0, 4, 900)].tobytes())  # longer than its haplotype
    read_hap.append(3)
    reads.append(b"A" * 1100)                                                   # poly-A against a random one
    read_hap.append(2)
    got = check(haps, reads, read_hap, assemble)
    assert got[-1][0] is not None
    served = rv.sw_align_batch(haps, reads, read_hap, assemble, served=True)
    assert served == got


def test_limits():
    say("limit errors")
    with pytest.raises(rv.EngineError) as e:
        rv.sw_align_batch([b"ACGT"], [b"AC"], [1])
    assert e.value.code == rv.RVC_E_INVALID


@pytest.mark.parametrize("assemble", [False, True])
def test_packed_and_int32_sweeps_mixed(assemble):
    # pairs two to a lane (the packed int16 sweep) next to pairs swept alone:
    # haplotype lengths far apart within a shape, odd counts per shape, reads
    # of 15 / 16 bases (the 16-bit core's edge) and, with the assembly scores,
    # short pairs (packed) beside long ones (int32)
    say(f"packed and int32 sweeps mixed, assemble={assemble}")
    rng = np.random.default_rng(11 + assemble)
    haps, reads, read_hap = [], [], []
    for L in (15, 16, 17, 40, 59, 60, 61, 100, 130, 250, 400):
        for t in range(7 if L % 2 else 4):
            H = int(rng.choice([L + 1, L + 30, 3 * L, 1500, 2000]))
            hap = ACGT[rng.integers(0, 4, H)].tobytes()
            haps.append(hap)
            a = int(rng.integers(0, max(1, H - L)))
            piece = hap[a:a + L]
            reads.append(_mutate(rng, piece, int(rng.integers(0, 3))) if t % 3 else ACGT[rng.integers(0, 4, L)].tobytes())
            read_hap.append(len(haps) - 1)
    for h in rng.permutation(len(reads))[:9]:  # extra reads on haplotypes already used
        reads.append(haps[read_hap[h]][:len(reads[h])] or b"A")
        read_hap.append(read_hap[h])
    check(haps, reads, read_hap, assemble)


def test_passes_under_a_small_backtrack_budget():
    # a 1 MiB backtrack budget splits the batch into many device passes (one
    # pair alone where its own area exceeds it): the same results as the
    # oracle, read by read
    say("many device passes under a 1 MiB backtrack budget")
    reinit(sw_bt_budget_mib=1)
    try:
        rng = np.random.default_rng(41)
        pairs = random_pairs(41, 300)
        haps = [h for h, _ in pairs]
        reads = [r for _, r in pairs]
        hap = ACGT[rng.integers(0, 4, 2000)].tobytes()
        haps.append(hap)
        for L in (700, 1024):  # areas of 1-2 MiB each
            reads.append(hap[100:100 + L])
        read_hap = list(range(len(pairs))) + [len(haps) - 1] * 2
        b = rv.SwBatch(haps, reads, read_hap)
        st = b.run()
        assert st["chunks"] > 3, st
        got = b.results()
        for r, read in enumerate(reads):
            assert got[r] == oracle.sw_align(haps[read_hap[r]], read), r
    finally:
        reinit()


def test_cigar_longer_than_its_slot():
    say("a CIGAR longer than max_cigar is an error")
    rng = np.random.default_rng(42)
    hap = ACGT[rng.integers(0, 4, 300)].tobytes()
    read = hap[10:60] + b"T" + hap[60:110]  # M I M: three elements
    ok = rv.SwBatch([hap], [read], [0], max_cigar=3)
    ok.run()
    assert ok.results()[0] == oracle.sw_align(hap, read)
    with pytest.raises(rv.EngineError) as e:
        rv.SwBatch([hap], [read], [0], max_cigar=2).run()
    assert e.value.code == rv.RVC_E_LIMIT


def test_concurrent_callers_equal_alone():
    # workers align from their own threads (each gets its own workspace) while
    # others run call_pairhmm on the same device
    say("aligner from 6 threads and call_pairhmm from 2 at once equal each call alone")
    import threading

    from rovaca_amd import workload
    batches = []
    for t in range(6):
        pairs = random_pairs(60 + t, 150)
        batches.append(([h for h, _ in pairs], [r for _, r in pairs], list(range(len(pairs))), bool(t % 2)))
    calls = []
    rng = np.random.default_rng(74)
    for reg in [r for r in workload.batched_regions(12, seed=73) if r["n_haps"] * r["n_reads"] < 4000][:2]:
        H, N = reg["n_haps"], reg["n_reads"]
        haps = [reg["hap_bases"][reg["hap_off"][h]:reg["hap_off"][h + 1]] for h in range(H)]
        reads = [reg["read_bases"][reg["read_off"][r]:reg["read_off"][r + 1]] for r in range(N)]
        quals = [reg["q"][reg["read_off"][r]:reg["read_off"][r + 1]] for r in range(N)]
        calls.append((haps, reads, quals, rng.integers(20, 61, N).astype(np.uint8)))
    alone_sw = [rv.sw_align_batch(*b) for b in batches]
    alone_ph = [rv.call_pairhmm(*c) for c in calls]
    got_sw, got_ph = [None] * len(batches), [None] * len(calls)

    def sw(x):
        got_sw[x] = rv.sw_align_batch(*batches[x])

    def ph(x):
        got_ph[x] = rv.call_pairhmm(*calls[x])

    th = [threading.Thread(target=sw, args=(x,)) for x in range(len(batches))]
    th += [threading.Thread(target=ph, args=(x,)) for x in range(len(calls))]
    for t in th:
        t.start()
    for t in th:
        t.join()
    assert got_sw == alone_sw
    for (rc0, m0, k0), (rc1, m1, k1) in zip(alone_ph, got_ph):
        assert rc0 == 0 and rc1 == 0 and np.array_equal(k0, k1)
        assert np.array_equal(np.isnan(m0), np.isnan(m1))
        assert np.array_equal(m0[~np.isnan(m0)], m1[~np.isnan(m1)])


def test_served_calls_equal_batch():
    # one region per call from many threads, merged by the serving threads
    say("served calls from 12 threads equal the batched calls")
    import threading
    batches = []
    for t in range(12):
        pairs = random_pairs(80 + t, 40 + 13 * t)
        batches.append(([h for h, _ in pairs], [r for _, r in pairs], list(range(len(pairs))), t % 3 == 0))
    want = [rv.sw_align_batch(*b) for b in batches]
    got = [None] * len(batches)

    def worker(x):
        got[x] = rv.sw_align_batch(*batches[x], served=True)

    th = [threading.Thread(target=worker, args=(x,)) for x in range(len(batches))]
    for t in th:
        t.start()
    for t in th:
        t.join()
    assert got == want
    for (haps, reads, rh, asm), g in zip(batches[:3], got[:3]):
        for r in range(len(reads)):
            assert g[r] == oracle.sw_align(haps[rh[r]], reads[r], asm)


def test_served_errors_stay_with_their_call():
    say("a served call that fails does not fail the calls merged with it")
    import threading
    rng = np.random.default_rng(43)
    hap = ACGT[rng.integers(0, 4, 300)].tobytes()
    gapped = hap[10:60] + b"T" + hap[60:110]  # three CIGAR elements
    good = [random_pairs(90 + t, 60) for t in range(6)]
    out, errs = [None] * 8, [None] * 8

    def ok(x):
        p = good[x]
        out[x] = rv.sw_align_batch([h for h, _ in p], [r for _, r in p], list(range(len(p))), served=True)

    def too_long_cigar(x):
        try:
            rv.SwBatch([hap], [gapped], [0], max_cigar=2).run(served=True)
        except rv.EngineError as e:
            errs[x] = e.code

    def bad_index(x):
        try:
            rv.sw_align_batch([b"A" * 1200], [b"A" * 1030], [1], served=True)
        except rv.EngineError as e:
            errs[x] = e.code

    th = [threading.Thread(target=ok, args=(x,)) for x in range(6)]
    th += [threading.Thread(target=too_long_cigar, args=(6,)), threading.Thread(target=bad_index, args=(7,))]
    for t in th:
        t.start()
    for t in th:
        t.join()
    assert errs[6] == rv.RVC_E_LIMIT and errs[7] == rv.RVC_E_INVALID
    for x in range(6):
        p = good[x]
        assert out[x] == [oracle.sw_align(h, r) for h, r in p]


def test_native_workers_served_and_batched():
    # Rovaca's worker pattern from C++ threads: one call per region, served
    # and batched; every region's results equal its own batched call
    say("64 region calls from 16 native workers, served and batched")
    batches, want = [], []
    for t in range(64):
        pairs = random_pairs(200 + t, 20 + t)
        haps, reads = [h for h, _ in pairs], [r for _, r in pairs]
        batches.append(rv.SwBatch(haps, reads, list(range(len(pairs)))))
        want.append(rv.sw_align_batch(haps, reads, list(range(len(pairs)))))
    for served in (True, False):
        res, lat = rv.loadgen_sw_align(batches, served=served, n_workers=16)
        assert res["completed"] == 64 and res["failed"] == 0 and (lat > 0).all()
        assert [b.results() for b in batches] == want


def test_engine_restart_and_folded_shapes():
    # shutdown releases the aligner's per-thread workspace; the next call
    # after init builds a new one. With the default setting the sparse shapes
    # of a small batch run folded into larger ones: the same results
    say("aligner across an engine restart; sparse shapes folded")
    pairs = random_pairs(31, 400)
    haps = [h for h, _ in pairs]
    reads = [r for _, r in pairs]
    first = check(haps, reads, list(range(len(pairs))))
    rv.sw_align_batch(haps, reads, list(range(len(pairs))), served=True)  # the serving threads started
    rv.shutdown()
    rv.init(0)
    assert rv.sw_align_batch(haps, reads, list(range(len(pairs)))) == first
    assert rv.sw_align_batch(haps, reads, list(range(len(pairs))), served=True) == first
    for assemble in (False, True):
        check(haps, reads, list(range(len(pairs))), assemble)
    reinit()


def test_shapes_chosen_by_pass_size():
    # with the default setting a pass whose pairs take at most three waves
    # per SIMD on the wide shapes (300 pairs) runs on them, a larger one (8000
    # pairs, ~4000 packed duos: more than three waves for each of the
    # device's 1024 SIMDs) on the base shapes; both equal the oracle read by
    # read, served or not
    say("wide shapes for a small pass, base shapes for a large one")
    rv.shutdown()
    rv.init(0)
    try:
        for seed, n in ((51, 300), (52, 8000)):
            pairs = random_pairs(seed, n)
            haps = [h for h, _ in pairs]
            reads = [r for _, r in pairs]
            got = check(haps, reads, list(range(n)))
            assert rv.sw_align_batch(haps, reads, list(range(n)), served=True) == got
    finally:
        reinit()


@pytest.mark.parametrize("assemble", [False, True])
def test_reads_with_n_bases(assemble):
    say(f"reads and haplotypes with N bases, assemble={assemble}")
    pairs = pairs_with_other_bytes(41 + assemble, 400)
    check([h for h, _ in pairs], [r for _, r in pairs], list(range(len(pairs))), assemble)
