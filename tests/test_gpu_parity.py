"""GPU parity: the HIP path (through the C ABI) against the oracle, Rovaca's
known answers and the committed reference fixtures.

Tolerance: |delta log10 L| <= 1e-4 against the oracle's reference routing
(the target bar is 1e-3; Rovaca's own tests use 1e-5 relative, which the
known-answer test below also meets)."""
import os
import sys
import threading

import numpy as np
import pytest

import oracle
import rovaca_amd as rv
from conftest import GOLDEN, load_debug_regions
from rovaca_amd import workload

pytestmark = pytest.mark.gpu
TOL = 1e-4


def say(msg):
    print(f"[gpu-test] {msg}", flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def engine():
    say("init device 0")
    rv.init(0, server=True)
    yield
    say("shutdown")
    rv.shutdown()



def _region1(hap, rd, q, i, d, c):
    return {"n_haps": 1, "n_reads": 1, "hap_off": np.array([0, len(hap)], np.int32),
            "hap_bases": np.append(np.asarray(hap, np.uint8), 0).astype(np.uint8),
            "read_off": np.array([0, len(rd)], np.int32),
            "read_bases": np.append(np.asarray(rd, np.uint8), 0).astype(np.uint8),
            "q": np.append(q, 0).astype(np.uint8), "i": np.append(i, 0).astype(np.uint8),
            "d": np.append(d, 0).astype(np.uint8), "c": np.append(c, 0).astype(np.uint8)}


def check_region(reg, tol=TOL):
    got, flags = rv.region_sync(reg)
    want, route, ok = oracle.region_reference(reg)
    assert ok
    nan_w, nan_g = np.isnan(want), np.isnan(got)
    assert np.array_equal(nan_w, nan_g)
    d = np.abs(got[~nan_g] - want[~nan_w]) if (~nan_g).any() else np.zeros(1)
    assert d.max() <= tol, f"max |dlog10| {d.max()}"
    assert np.array_equal((flags & 1) == 1, ~nan_g)
    return got, flags, d.max()


def test_kats(kats):
    say("known-answer tests, one pair per region")
    worst = 0.0
    for hap, rd, q, i, d, c, exp in kats:
        got, flags = rv.region_sync(_region1(hap, rd, q, i, d, c))
        v = got[0, 0]
        assert abs(v - exp) / abs(exp) < 1e-5
        worst = max(worst, abs(v - oracle.intra_log10(hap, rd, q, i, d, c)))
    assert worst < 1e-5
    say(f"KAT worst |d| vs oracle {worst:.2e}")


def test_single_region_against_reference_fixture():
    say("single region (8 x 32) against the reference fixture")
    g = np.load(os.path.join(GOLDEN, "single_region_seed0.npz"))
    reg = {k: g[k] for k in ("hap_off", "hap_bases", "read_off", "read_bases", "q", "i", "d", "c")}
    reg["n_haps"], reg["n_reads"] = len(reg["hap_off"]) - 1, len(reg["read_off"]) - 1
    got, flags, dmax = check_region(reg)
    assert np.abs(got - g["ref_intra"]).max() < TOL
    assert np.abs(got - g["fp64"]).max() < TOL
    say(f"single region max |d| {dmax:.2e}")


@pytest.mark.parametrize("name", sorted(f for f in os.listdir(GOLDEN) if f.startswith(("batched_", "stress_"))))
def test_fixture_regions(name):
    say(f"fixture {name}")
    g = np.load(os.path.join(GOLDEN, name))
    reg = {k: g[k] for k in ("hap_off", "hap_bases", "read_off", "read_bases", "q", "i", "d", "c")}
    reg["n_haps"], reg["n_reads"] = len(reg["hap_off"]) - 1, len(reg["read_off"]) - 1
    got, flags = rv.region_sync(reg)
    assert np.abs(got - g["ref_intra"]).max() < TOL
    assert np.abs(got - g["fp64"]).max() < TOL


def test_rescue_path():
    say("stress regions: FP64 rescue")
    regs = [r for r in workload.stress_regions(12, seed=5) if r["n_haps"] * r["n_reads"] < 4000][:4]
    n_resc = 0
    for reg in regs:
        got, flags, _ = check_region(reg)
        n_resc += int(((flags & 2) != 0).sum())
        # rescued pairs are exactly those below the FP32 threshold: log10 L < -64.12 (+/- float rounding)
        resc = (flags & 2) != 0
        assert (got[resc] < -63.9).all()
        assert (got[(flags == 1)] > -64.3).all()
    assert n_resc > 0


def test_rescue_both_fp64_kernels():
    say("FP64 rescue of pairs with and without N in the haplotype (compare and prior-table kernels)")
    regs = [r for r in workload.stress_regions(12, seed=6) if r["n_haps"] * r["n_reads"] < 4000][:3]
    r0 = dict(regs[0])
    hb = r0["hap_bases"].copy()
    for h in range(0, r0["n_haps"], 2):
        hb[r0["hap_off"][h] + 7] = ord("N")
    r0["hap_bases"] = hb
    regs[0] = r0
    b = rv.Batch(regs)
    b.run()
    out, flags = b.results()
    names = {k["name"] for k in b.kernels() if k["precision"] == 64}
    assert any(n.startswith("k_fp64pt<") for n in names) and any(n.startswith("k_fp64<") for n in names)
    for x, reg in enumerate(regs):
        got = b.region_result(out, x)
        fl = b.region_result(flags, x)
        want, route, ok = oracle.region_reference(reg)
        m = ~np.isnan(want)
        assert np.array_equal(m, ~np.isnan(got))
        assert np.abs(got[m] - want[m]).max() < TOL
        if x == 0:
            hn = np.zeros(reg["n_haps"], bool)
            hn[::2] = True
            assert ((fl & 2) != 0)[:, hn].any() and ((fl & 2) != 0)[:, ~hn].any()
    b.close()


def test_long_rescue_lists_equal_alone():
    # enough rescued pairs in one FP64 list that each segment of the FP64
    # prior-table kernel takes several consecutive entries (a read set up once
    # for several of its pairs, entries appended a read at a time by the FP32
    # pass); every result equals the region computed alone, whose short lists
    # give one entry per segment
    say("long FP64 rescue lists: batch equals each region alone bit for bit; a sample against the FP64 oracle")
    regs = workload.stress_regions(64, seed=81)
    b = rv.Batch(regs)
    b.run()
    out, flags = b.results()
    assert b.info()["n_rescued"] > 40000
    rng = np.random.default_rng(82)
    worst, checked = 0.0, 0
    for x, reg in enumerate(regs):
        a, af = rv.region_sync(reg)
        m = b.region_result(out, x)
        mf = b.region_result(flags, x)
        assert np.array_equal(np.isnan(a), np.isnan(m)), x
        assert np.array_equal(a[~np.isnan(a)], m[~np.isnan(m)]), x
        assert np.array_equal(af, mf), x
        resc = np.argwhere((mf & 2) != 0)
        for r, h in resc[rng.permutation(len(resc))[:3]]:
            o0, o1 = reg["read_off"][r], reg["read_off"][r + 1]
            hb = reg["hap_bases"][reg["hap_off"][h]:reg["hap_off"][h + 1]]
            want = oracle.pair_log10_f64(hb, reg["read_bases"][o0:o1], *[reg[k][o0:o1] for k in "qidc"])
            worst = max(worst, abs(m[r, h] - want))
            checked += 1
    b.close()
    assert checked > 100 and worst < TOL
    say(f"{checked} rescued pairs vs FP64 oracle, worst |d| {worst:.2e}")


def _gap_region(seed, gop_lo, gop_hi, gcp, n_haps=5, n_reads=6):
    """Reads whose gap-open qualities alternate between gop_lo and gop_hi row by
    row (the largest step between rows), gap continuation gcp."""
    rng = np.random.default_rng(seed)
    haps = [workload.ACGT[rng.integers(0, 4, int(rng.integers(1, 160)))] for _ in range(n_haps)]
    haps[0] = haps[0][:1]
    reads, quals, gi, gd, gc = [], [], [], [], []
    for x in range(n_reads):
        h = haps[1 + x % (n_haps - 1)]
        L = int(rng.integers(1, min(len(h), 120) + 1))
        r = h[:L].copy()
        r[rng.random(L) < 0.05] = workload.ACGT[rng.integers(0, 4)]
        reads.append(r)
        quals.append(rng.integers(10, 41, L).astype(np.uint8))
        alt = np.where(np.arange(L) % 2 == 0, gop_lo, gop_hi).astype(np.uint8)
        gi.append(alt)
        gd.append(alt[::-1].copy())
        gc.append(np.full(L, gcp, np.uint8))
    reg = workload.make_region(haps, reads, quals)
    reg["i"] = np.append(np.concatenate(gi), 0).astype(np.uint8)
    reg["d"] = np.append(np.concatenate(gd), 0).astype(np.uint8)
    reg["c"] = np.append(np.concatenate(gc), 0).astype(np.uint8)
    return reg


# (gap open low, high, gap continuation, form the FP32 pass must use): the
# scaled form takes gap qualities >= 6 with a gap-open spread up to 48
@pytest.mark.parametrize("lo,hi,gcp,form", [(45, 45, 10, "scaled"), (10, 45, 10, "scaled"), (6, 54, 6, "scaled"),
                                            (6, 55, 6, "plain"), (5, 45, 10, "plain"), (10, 45, 5, "plain"),
                                            (10, 45, 0, "plain"), (30, 78, 20, "scaled")])
def test_gap_quality_forms(lo, hi, gcp, form):
    say(f"gap qualities {lo}..{hi}, continuation {gcp}: {form} form")
    reg = _gap_region(lo * 131 + hi * 7 + gcp, lo, hi, gcp)
    b = rv.Batch([reg])
    b.run()
    out, flags = b.results()
    names = {k["name"] for k in b.kernels() if k["precision"] == 32}
    assert names and all(f",{form}," in n for n in names), names
    got = b.region_result(out, 0)
    b.close()
    want, route, ok = oracle.region_reference(reg)
    assert ok
    m = ~np.isnan(want)
    assert np.array_equal(m, ~np.isnan(got))
    # gap continuation 0 leaves no way back from a gap: equal infinities
    same = got[m] == want[m]
    assert same.any() or np.isfinite(want[m]).all()
    with np.errstate(invalid="ignore"):  # inf - inf where both are the same infinity
        close = np.abs(got[m] - want[m]) <= TOL
    assert np.all(same | close)
    # alone through the synchronous path: the same form, bit for bit
    alone, _ = rv.region_sync(reg)
    assert np.array_equal(alone[m], got[m])


def test_mixed_forms_equal_alone():
    say("plain-form regions among scaled ones: each equals the region alone, batched and served")
    base = [r for r in workload.batched_regions(10, seed=61) if r["n_haps"] * r["n_reads"] < 6000]
    # two regions the scaled form cannot take (gap open 5; spread 49) and two
    # it takes at other exponents than the batch's default regions (spreads
    # 48, the lowest exponent, and 48 from gap open 30 up)
    odd = [_gap_region(901, 5, 45, 10), _gap_region(902, 6, 55, 6), _gap_region(903, 6, 54, 6),
           _gap_region(904, 30, 78, 20)]
    regs = base[:4] + odd[:2] + base[4:7] + odd[2:] + base[7:]
    alone = [rv.region_sync(r) for r in regs]
    b = rv.Batch(regs)
    b.run()
    out, flags = b.results()
    names = {k["name"] for k in b.kernels() if k["precision"] == 32}
    assert any(",plain," in n for n in names) and any(",scaled," in n for n in names), names
    for x, reg in enumerate(regs):
        a, af = alone[x]
        m = b.region_result(out, x)
        assert np.array_equal(np.isnan(a), np.isnan(m)), x
        assert np.array_equal(a[~np.isnan(a)], m[~np.isnan(m)]), x
        assert np.array_equal(af, b.region_result(flags, x)), x
    b.close()
    # the serving loop batches concurrent submissions together
    results = [None] * len(regs)
    errors = []

    def worker(idx):
        try:
            for x in idx:
                t = rv.submit(regs[x])
                results[x] = rv.wait(t, regs[x])[0]
        except Exception as e:  # noqa: BLE001
            errors.append(e)

    th = [threading.Thread(target=worker, args=(list(range(k, len(regs), 8)),)) for k in range(8)]
    for t in th:
        t.start()
    for t in th:
        t.join(timeout=120)
    assert not errors
    for (a, _), m in zip(alone, results):
        assert np.array_equal(np.isnan(a), np.isnan(m))
        assert np.array_equal(a[~np.isnan(a)], m[~np.isnan(m)])


# the FP64 bins hold reads of fewer than 64, 128, 160 and 256 bases (their
# last row is a padding row; one at 63, 127, 159, 255) and up to 768
@pytest.mark.parametrize("R", [63, 64, 65, 127, 128, 129, 159, 160, 161, 255, 256, 257, 512, 767, 768])
def test_rescue_bin_edges(R):
    say(f"FP64 rescue at read length {R} (edges of the rescue bins' row counts)")
    rng = np.random.default_rng(1000 + R)
    hap = np.full(800, ord("A"), np.uint8)
    haps = [hap, hap.copy()]
    haps[1][5] = ord("N")  # one haplotype per FP64 kernel
    # 30 C's in every other base of a poly-A read at quality 40 with gap
    # continuation 20: a mismatch costs 4.5 in log10, an insertion 4.5 to open
    # and 2 a base, so log10 L is near -120, below the FP32 range (about -64)
    # and inside the FP64 one
    reads, quals = [], []
    for x in range(3):
        r = np.full(R, ord("A"), np.uint8)
        first = int(rng.integers(0, R - 60 + 1))
        r[first + 1 + x % 2:first + 60:2] = ord("C")
        reads.append(r)
        quals.append(np.full(R, 40, np.uint8))
    reg = workload.make_region(haps, reads, quals)
    reg["c"][:] = 20
    got, flags, _ = check_region(reg)
    assert ((flags & 2) != 0).all() and np.isfinite(got).all()


@pytest.mark.parametrize("R,H", [(1, 1), (1, 40), (40, 1), (16, 16), (17, 300), (100, 350), (102, 300), (103, 300),
                                 (112, 300), (127, 200), (128, 200), (158, 300), (159, 300), (254, 200), (255, 200),
                                 (256, 256), (257, 100), (300, 700), (510, 512), (511, 512), (512, 512), (513, 90),
                                 (700, 1536), (767, 900), (768, 1536)])
def test_shapes(R, H):
    say(f"shape R={R} H={H}")
    rng = np.random.default_rng(R * 7919 + H)
    hap = workload.ACGT[rng.integers(0, 4, H)]
    start = int(rng.integers(0, max(1, H - R)))
    rd = np.resize(hap[start:start + R], R).copy()
    rd[rng.random(R) < 0.05] = workload.ACGT[rng.integers(0, 4)]
    q = rng.integers(10, 41, R).astype(np.uint8)
    reg = _region1(hap, rd, q, np.full(R, 45, np.uint8), np.full(R, 45, np.uint8), np.full(R, 10, np.uint8))
    check_region(reg)


def test_n_and_empty_sequences():
    say("N bases and empty sequences")
    rng = np.random.default_rng(21)
    haps = [workload.ACGT[rng.integers(0, 4, 120)] for _ in range(3)] + [np.zeros(0, np.uint8)]
    haps[1] = haps[1].copy()
    haps[1][::17] = ord("N")
    reads, quals = [], []
    for x in range(7):
        L = 0 if x == 3 else int(rng.integers(20, 110))
        s = haps[x % 3][:L].copy()
        s[rng.random(L) < 0.05] = ord("N")
        reads.append(s)
        quals.append(rng.integers(5, 41, L).astype(np.uint8))
    reg = workload.make_region(haps, reads, quals)
    got, flags, _ = check_region(reg)
    assert np.isnan(got[3]).all() and np.isnan(got[:, 3]).all()


def test_limits():
    say("limit errors")
    R = 10
    for H, Rr, code in [(1537, R, rv.RVC_E_LIMIT), (100, 769, rv.RVC_E_LIMIT)]:
        reg = _region1(np.full(H, 65, np.uint8), np.full(Rr, 65, np.uint8), np.full(Rr, 30, np.uint8),
                       np.full(Rr, 45, np.uint8), np.full(Rr, 45, np.uint8), np.full(Rr, 10, np.uint8))
        with pytest.raises(rv.EngineError) as e:
            rv.region_sync(reg)
        assert e.value.code == code


@pytest.mark.parametrize("byte", [b"R", b"Y", b"=", b"a", b"n", b"."])
def test_other_bytes_are_refused(byte):
    # The reference's result for a byte other than A, C, G, T, N depends on
    # its routing (raw-byte compare in the inter kernel, 'A' or an index past
    # its table in the intra kernel): every entry point refuses the region
    # with RVC_E_BASES and computes nothing; the caller keeps its CPU path.
    say(f"byte {byte!r} in a read and in a haplotype: RVC_E_BASES")
    base = workload.single_region(3)
    for where in ("read_bases", "hap_bases"):
        reg = dict(base)
        arr = reg[where].copy()
        arr[7] = byte[0]
        reg[where] = arr
        with pytest.raises(rv.EngineError) as e:
            rv.region_sync(reg)
        assert e.value.code == rv.RVC_E_BASES
        with pytest.raises(rv.EngineError) as e:
            rv.submit(reg)
        assert e.value.code == rv.RVC_E_BASES
        with pytest.raises(rv.EngineError) as e:
            rv.Batch([base, reg])
        assert e.value.code == rv.RVC_E_BASES
        H, N = reg["n_haps"], reg["n_reads"]
        haps = [reg["hap_bases"][reg["hap_off"][h]:reg["hap_off"][h + 1]] for h in range(H)]
        reads = [reg["read_bases"][reg["read_off"][r]:reg["read_off"][r + 1]] for r in range(N)]
        quals = [reg["q"][reg["read_off"][r]:reg["read_off"][r + 1]] for r in range(N)]
        rc, mat, kept = rv.call_pairhmm(haps, reads, quals, np.full(N, 60, np.uint8))
        assert rc == rv.RVC_E_BASES and np.isnan(mat).all() and list(kept) == list(range(N))
    # the same region with only A, C, G, T, N still runs
    check_region(base)


def test_call_pairhmm_drops_a_middle_read():
    say("call_pairhmm: a poorly modelled read in the middle is dropped, the others packed to the front")
    reg = workload.single_region(5)
    H, N = reg["n_haps"], reg["n_reads"]
    haps = [reg["hap_bases"][reg["hap_off"][h]:reg["hap_off"][h + 1]] for h in range(H)]
    reads = [reg["read_bases"][reg["read_off"][r]:reg["read_off"][r + 1]].copy() for r in range(N)]
    quals = [reg["q"][reg["read_off"][r]:reg["read_off"][r + 1]].copy() for r in range(N)]
    rng = np.random.default_rng(12345)  # not the region's own seed (its first draws are haplotype 0)
    bad = N // 2
    reads[bad] = workload.ACGT[rng.integers(0, 4, len(reads[bad]))]  # unrelated to every haplotype
    quals[bad][:] = 40
    rc, mat, kept = rv.call_pairhmm(haps, reads, quals, np.full(N, 60, np.uint8))
    assert rc == 0
    assert bad not in set(kept.tolist()) and len(kept) < N
    assert np.isnan(mat[:, len(kept):]).all() and not np.isnan(mat[:, :len(kept)]).any()
    # the kept columns are the reads' normalised rows, in order
    oreg = workload.make_region(haps, reads, quals)
    rm, _, ok = oracle.region_reference(oreg)
    want, wkept = oracle.postprocess(rm, np.array([len(r) for r in reads], np.int32))
    assert np.array_equal(kept, wkept)
    assert np.abs(mat[:, :len(kept)] - want[:, :len(kept)]).max() < TOL


def test_many_haplotypes():
    say("4096 haplotypes x 2 reads")
    rng = np.random.default_rng(4)
    haps = [workload.ACGT[rng.integers(0, 4, int(rng.integers(1, 60)))] for _ in range(4096)]
    reads = [haps[5][:], haps[77][:]]
    quals = [np.full(len(r), 30, np.uint8) for r in reads]
    check_region(workload.make_region(haps, reads, quals))


def test_batch_equals_alone():
    say("batch of mixed regions equals each region alone (bit for bit)")
    regs = workload.batched_regions(24, seed=31) + workload.stress_regions(4, seed=32)
    regs = [r for r in regs if r["n_haps"] * r["n_reads"] < 6000]
    b = rv.Batch(regs)
    b.run()
    out, flags = b.results()
    for x, reg in enumerate(regs):
        alone, af = rv.region_sync(reg)
        mixed = b.region_result(out, x)
        assert np.array_equal(np.isnan(alone), np.isnan(mixed))
        assert np.array_equal(alone[~np.isnan(alone)], mixed[~np.isnan(mixed)])
    info = b.info()
    assert info["n_pairs"] == sum(r["n_haps"] * r["n_reads"] for r in regs)
    b.close()


@pytest.mark.parametrize("run", [3, 64])
def test_hap_runs_equal_alone(run):
    say(f"tiles sweeping runs of {run} haplotypes equal each region alone (bit for bit)")
    regs = [r for r in workload.batched_regions(16, seed=61) if r["n_haps"] * r["n_reads"] < 6000][:8]
    # N in some haplotypes of the first region breaks its runs: those take the
    # compare-and-select kernel, the others the prior-table kernel
    r0 = dict(regs[0])
    hb = r0["hap_bases"].copy()
    for h in range(1, r0["n_haps"], 3):
        hb[r0["hap_off"][h] + 5] = ord("N")
    r0["hap_bases"] = hb
    regs[0] = r0
    b = rv.Batch(regs, hap_run=run)
    b.run()
    out, flags = b.results()
    for x, reg in enumerate(regs):
        alone, af = rv.region_sync(reg)
        mixed = b.region_result(out, x)
        assert np.array_equal(np.isnan(alone), np.isnan(mixed))
        assert np.array_equal(alone[~np.isnan(alone)], mixed[~np.isnan(mixed)])
        assert np.array_equal(af, b.region_result(flags, x))
    names = {k["name"] for k in b.kernels() if k["precision"] == 32}
    assert any(n.startswith("k_fp32pt") for n in names) and any(n.startswith("k_fp32<") for n in names)
    b.close()


def test_small_bins_fold_into_larger_shapes():
    say("tiles of small shape bins swept by the next larger shape: fewer launches, the same bits")
    rng = np.random.default_rng(71)
    regs = []
    for x in range(8):
        base = workload.batched_regions(1, seed=700 + x)[0]
        H = base["n_haps"]
        haps = [base["hap_bases"][base["hap_off"][h]:base["hap_off"][h + 1]] for h in range(H)]
        haps = [h if len(h) >= 260 else np.concatenate([h, h, h])[:300] for h in haps]
        reads, quals = [], []
        for r in range(30):
            n = int(rng.choice([100, 120, 130, 140, 150, 155, 180, 200, 250]))  # many shape bins
            src = haps[int(rng.integers(0, H))]
            o = int(rng.integers(0, len(src) - n + 1))
            reads.append(src[o:o + n].copy())
            quals.append(rng.integers(10, 41, n).astype(np.uint8))
        regs.append(workload.make_region(haps, reads, quals))
    runs = {}
    for fold in (True, False):
        b = rv.Batch(regs, shape_fold=fold)
        b.run()
        out, flags = b.results()
        runs[fold] = ([b.region_result(out, x) for x in range(len(regs))],
                      [b.region_result(flags, x) for x in range(len(regs))],
                      [k for k in b.kernels() if k["precision"] == 32])
        b.close()
    (o1, f1, k1), (o0, f0, k0) = runs[True], runs[False]
    assert len(k1) < len(k0), ([k["name"] for k in k1], [k["name"] for k in k0])
    assert sum(k["cells"] for k in k1) == sum(k["cells"] for k in k0)
    for a, b_, fa, fb in zip(o1, o0, f1, f0):
        assert np.array_equal(np.isnan(a), np.isnan(b_)) and np.array_equal(a[~np.isnan(a)], b_[~np.isnan(b_)])
        assert np.array_equal(fa, fb)
    alone, _ = rv.region_sync(regs[0])
    assert np.array_equal(np.isnan(alone), np.isnan(o1[0])) and np.array_equal(alone[~np.isnan(alone)],
                                                                               o1[0][~np.isnan(o1[0])])


@pytest.mark.parametrize("width", [8, 32])
def test_segment_width_equals_default(width):
    say(f"{width}-lane segments give the same bits as the default shapes")
    rng = np.random.default_rng(81)
    regs = []
    for x in range(6):
        reg = workload.batched_regions(1, seed=82 + x)[0]
        # reads of 20..120 bases (the 8-lane shapes hold up to 127)
        H, N = reg["n_haps"], min(reg["n_reads"], 40)
        haps = [reg["hap_bases"][reg["hap_off"][h]:reg["hap_off"][h + 1]] for h in range(H)]
        reads, quals = [], []
        for r in range(N):
            o = reg["read_off"][r]
            n = min(int(rng.integers(20, 121)), int(reg["read_off"][-1] - o))  # bases of the region only
            reads.append(reg["read_bases"][o:o + n])
            quals.append(reg["q"][o:o + n])
        if x == 0:  # N in a haplotype: compare-and-select kernel
            haps[1] = haps[1].copy()
            haps[1][3] = ord("N")
        regs.append(workload.make_region(haps, reads, quals))
    b = rv.Batch(regs, segment_width=width, shape_fold=False)
    b.run()
    out, flags = b.results()
    names = {k["name"] for k in b.kernels() if k["precision"] == 32}
    assert all(f"<{width}," in n for n in names)
    for x, reg in enumerate(regs):
        alone, af = rv.region_sync(reg)
        mixed = b.region_result(out, x)
        assert np.array_equal(np.isnan(alone), np.isnan(mixed))
        assert np.array_equal(alone[~np.isnan(alone)], mixed[~np.isnan(mixed)])
        assert np.array_equal(af, b.region_result(flags, x))
    b.close()


def test_graph_replays_match_eager():
    say("hipGraph replays equal the eager run")
    regs = [r for r in workload.stress_regions(8, seed=41) if r["n_haps"] * r["n_reads"] < 5000]
    eager = rv.Batch(regs)
    eager.run()
    ref_out, ref_flags = eager.results()
    g = rv.Batch(regs, use_graph=True)
    for rep in range(4):
        g.run()
        out, flags = g.results()
        assert np.array_equal(np.isnan(out), np.isnan(ref_out))
        m = ~np.isnan(out)
        assert np.array_equal(out[m], ref_out[m]) and np.array_equal(flags, ref_flags)
        rv.region_sync(regs[0])  # unrelated eager work between replays
    eager.close()
    g.close()


def test_served_results_equal_alone():
    say("serving: 16 threads submit concurrently; each result equals the region alone")
    regs = [r for r in workload.batched_regions(48, seed=51) if r["n_haps"] * r["n_reads"] < 8000]
    alone = [rv.region_sync(r)[0] for r in regs]
    results = [None] * len(regs)
    errors = []

    def worker(idx):
        try:
            for x in idx:
                t = rv.submit(regs[x])
                results[x] = rv.wait(t, regs[x])[0]
        except Exception as e:  # noqa: BLE001
            errors.append(e)

    th = [threading.Thread(target=worker, args=(list(range(k, len(regs), 16)),)) for k in range(16)]
    for t in th:
        t.start()
    for t in th:
        t.join(timeout=120)
    assert not errors
    for a, m in zip(alone, results):
        assert np.array_equal(np.isnan(a), np.isnan(m))
        assert np.array_equal(a[~np.isnan(a)], m[~np.isnan(m)])
    st = rv.server_stats()
    assert st["completed"] >= len(regs)


def test_served_rescue_equals_alone():
    # stress regions among ordinary ones from 8 threads: a served launch whose
    # FP32 pass flags pairs runs the FP64 lists in its second phase, one
    # without flagged pairs skips them; each result equals the region alone
    say("serving with FP64 rescue: stress and ordinary regions from 8 threads, each equals the region alone")
    stress = [r for r in workload.stress_regions(16, seed=71) if r["n_haps"] * r["n_reads"] < 5000][:6]
    plain = [r for r in workload.batched_regions(12, seed=72) if r["n_haps"] * r["n_reads"] < 5000]
    regs = [r for pair in zip(stress, plain) for r in pair] + plain[len(stress):]
    alone = [rv.region_sync(r) for r in regs]
    assert sum(int(((f & 2) != 0).sum()) for _, f in alone) > 0
    results = [None] * len(regs)
    errors = []

    def worker(idx):
        try:
            for x in idx:
                t = rv.submit(regs[x])
                results[x] = rv.wait(t, regs[x])[0]
        except Exception as e:  # noqa: BLE001
            errors.append(e)

    for rep in range(2):
        th = [threading.Thread(target=worker, args=(list(range(k, len(regs), 8)),)) for k in range(8)]
        for t in th:
            t.start()
        for t in th:
            t.join(timeout=120)
        assert not errors
        for (a, _), m in zip(alone, results):
            assert np.array_equal(np.isnan(a), np.isnan(m)), rep
            assert np.array_equal(a[~np.isnan(a)], m[~np.isnan(m)]), rep


def test_call_pairhmm_region_adapter():
    say("call_pairhmm adapter: preprocessing + likelihoods + normalize/filter/transpose")
    rng = np.random.default_rng(61)
    reg = workload.batched_regions(1, seed=62)[0]
    H, N = reg["n_haps"], reg["n_reads"]
    haps = [reg["hap_bases"][reg["hap_off"][h]:reg["hap_off"][h + 1]] for h in range(H)]
    reads = [reg["read_bases"][reg["read_off"][r]:reg["read_off"][r + 1]] for r in range(N)]
    # the region's own base qualities with some noise; MQ 20-60 caps some of them
    quals = [np.clip(reg["q"][reg["read_off"][r]:reg["read_off"][r + 1]].astype(int)
                     + rng.integers(-3, 4, reads[r].size), 0, 60).astype(np.uint8) for r in range(N)]
    mqs = rng.integers(20, 61, N).astype(np.uint8)
    for pcr in (rv.PCR_NONE, rv.PCR_CONSERVATIVE):
        rc, mat, kept = rv.call_pairhmm(haps, reads, quals, mqs, min_quality_threshold=18, pcr=pcr)
        assert rc == 0
        # oracle: same preprocessing, reference routing, post-processing
        pre = [oracle.preprocess_read(reads[r], quals[r], int(mqs[r]), 18, pcr) for r in range(N)]
        oreg = workload.make_region(haps, reads, [p[0] for p in pre])
        oreg["i"] = np.concatenate([p[1] for p in pre] + [np.zeros(1, np.uint8)])
        oreg["d"] = np.concatenate([p[2] for p in pre] + [np.zeros(1, np.uint8)])
        oreg["c"] = np.concatenate([p[3] for p in pre] + [np.zeros(1, np.uint8)])
        rm, _, ok = oracle.region_reference(oreg)
        want, wkept = oracle.postprocess(rm, np.diff(reg["read_off"]))
        assert np.array_equal(kept, wkept)
        m = ~np.isnan(want)
        assert np.array_equal(m, ~np.isnan(mat))
        assert np.abs(mat[m] - want[m]).max() < TOL


def test_call_pairhmm_concurrent_equals_alone():
    say("call_pairhmm from 12 threads at once (served, batched together) equals each call alone")
    rng = np.random.default_rng(71)
    regs = [r for r in workload.batched_regions(24, seed=72) if r["n_haps"] * r["n_reads"] < 6000][:12]
    calls = []
    for reg in regs:
        H, N = reg["n_haps"], reg["n_reads"]
        haps = [reg["hap_bases"][reg["hap_off"][h]:reg["hap_off"][h + 1]] for h in range(H)]
        reads = [reg["read_bases"][reg["read_off"][r]:reg["read_off"][r + 1]] for r in range(N)]
        quals = [reg["q"][reg["read_off"][r]:reg["read_off"][r + 1]] for r in range(N)]
        mqs = rng.integers(20, 61, N).astype(np.uint8)
        calls.append((haps, reads, quals, mqs))
    alone = [rv.call_pairhmm(*c, min_quality_threshold=10, pcr=rv.PCR_CONSERVATIVE) for c in calls]
    mixed = [None] * len(calls)

    def worker(x):
        mixed[x] = rv.call_pairhmm(*calls[x], min_quality_threshold=10, pcr=rv.PCR_CONSERVATIVE)

    th = [threading.Thread(target=worker, args=(x,)) for x in range(len(calls))]
    for t in th:
        t.start()
    for t in th:
        t.join()
    for (rc0, m0, k0), (rc1, m1, k1) in zip(alone, mixed):
        assert rc0 == 0 and rc1 == 0
        assert np.array_equal(k0, k1)
        assert np.array_equal(np.isnan(m0), np.isnan(m1))
        assert np.array_equal(m0[~np.isnan(m0)], m1[~np.isnan(m1)])


def test_call_pairhmm_served_device_prep_equals_host_prep():
    say("served call_pairhmm with device preprocessing == host preprocessing + plain call (bit for bit)")
    from test_gpu_prep import host_path, repeat_rich_region
    for pcr in (rv.PCR_NONE, rv.PCR_CONSERVATIVE, rv.PCR_HOSTILE):
        haps, reads, quals, mqs = repeat_rich_region(200 + pcr)
        rc, mat, kept = rv.call_pairhmm(haps, reads, quals, mqs, min_quality_threshold=10, pcr=pcr)
        assert rc == 0
        want, wkept = host_path(haps, reads, quals, mqs, 10, pcr)
        assert np.array_equal(kept, wkept)
        m = ~np.isnan(want)
        assert np.array_equal(m, ~np.isnan(mat)) and np.array_equal(mat[m], want[m])


def test_loadgen_serves_poisson_stream():
    say("load generator: 300 regions at 3000/s from 8 producers")
    regs = [r for r in workload.batched_regions(40, seed=71) if r["n_haps"] * r["n_reads"] < 6000]
    rng = np.random.default_rng(1)
    arr = np.cumsum(rng.exponential(1 / 3000.0, 300))
    res, lat = rv.loadgen_run(regs, arr, np.arange(300) % len(regs), 8)
    assert res["completed"] == 300 and res["failed"] == 0
    assert (lat > 0).all() and res["lat_p99_ms"] >= res["lat_p50_ms"] > 0
    assert res["cells"] == sum(workload.cells([regs[i % len(regs)]]) for i in range(300))


def test_loadgen_region_calls():
    say("worker threads calling call_pairhmm (C++ load generator)")
    rng = np.random.default_rng(91)
    pool = [r for r in workload.batched_regions(12, seed=92) if r["n_haps"] * r["n_reads"] < 6000]
    raws = [rv.raw_region(r, rng.integers(20, 61, int(r["n_reads"]))) for r in pool]
    res, lat = rv.loadgen_call_pairhmm(raws, 48, None, 10, rv.PCR_CONSERVATIVE, 8)
    assert res["completed"] == 48 and res["failed"] == 0 and (lat > 0).all()


def test_budget_admission():
    say("admission by budget: a region larger than the whole budget is refused")
    rv.shutdown()
    rv.init(0, server=True, hbm_budget_bytes=1 << 20)
    try:
        big = workload.deep_region(seed=81)  # 256k pairs: several MB on the device
        with pytest.raises(rv.EngineError) as e:
            rv.submit(big)
        assert e.value.code == rv.RVC_E_BUDGET
        small = [r for r in workload.batched_regions(30, seed=82) if r["n_haps"] * r["n_reads"] < 2000][:6]
        # several small ones still go through one at a time under the tight budget
        tickets = [rv.submit(r) for r in small]
        for t, r in zip(tickets, small):
            got, _ = rv.wait(t, r)
            alone, _ = rv.region_sync(r)
            assert np.array_equal(np.isnan(got), np.isnan(alone))
            assert np.array_equal(got[~np.isnan(got)], alone[~np.isnan(alone)])
    finally:
        rv.shutdown()
        rv.init(0, server=True)


def test_real_regions_against_reference_fixture():
    # Rovaca's own schedule test data: 8 real chrY regions, expected values from
    # Rovaca's intra kernel, tolerance 1e-5 as ph/test/pairhmm_unitest_from_assemble.cpp:158
    say("8 real regions (ph/resource/pairhmm_debug.txt) against Rovaca's intra kernel, 1e-5")
    regs = load_debug_regions()
    ref = np.load(os.path.join(GOLDEN, "pairhmm_debug_ref.npz"))["ref_intra"]
    b = rv.Batch(regs)
    b.run()
    bout, _ = b.results()
    o, worst = 0, 0.0
    for x, reg in enumerate(regs):
        n = reg["n_haps"] * reg["n_reads"]
        want = ref[o:o + n].reshape(reg["n_reads"], reg["n_haps"])
        o += n
        got, flags = rv.region_sync(reg)
        assert np.isfinite(got).all() and (flags & 1).all()
        worst = max(worst, float(np.abs(got - want).max()))
        assert np.array_equal(got, b.region_result(bout, x))
    b.close()
    assert o == len(ref)
    assert worst <= 1e-5, worst
    say(f"real regions: worst |d| {worst:.2e}")


def _rescue_check(got, flags, route):
    # the FP64 flag against the reference routing's FP64 routes (2: inter
    # FP64, 4/5: intra FP64); a disagreement is allowed only within 1e-3 of
    # the threshold log10(1e-28f) - 36.1236000061f = -64.1236
    resc = (flags & rv.PAIR_FP64) != 0
    ref_resc = np.isin(route, (2, 4, 5))
    diff = resc != ref_resc
    assert (np.abs(got[diff] + 64.1236000061) < 1e-3).all(), got[diff]
    return int(resc.sum()), int(diff.sum())


def test_rescue_membership_matches_reference_routing():
    say("FP64 rescue membership against the reference routing on stress regions")
    regs = [r for r in workload.stress_regions(24, seed=91) if r["n_haps"] * r["n_reads"] < 6000][:8]
    for name in sorted(f for f in os.listdir(GOLDEN) if f.startswith("stress_")):
        g = np.load(os.path.join(GOLDEN, name))
        reg = {k: g[k] for k in ("hap_off", "hap_bases", "read_off", "read_bases", "q", "i", "d", "c")}
        reg["n_haps"], reg["n_reads"] = len(reg["hap_off"]) - 1, len(reg["read_off"]) - 1
        regs.append(reg)
    n_resc = n_diff = 0
    for reg in regs:
        got, flags = rv.region_sync(reg)
        want, route, ok = oracle.region_reference(reg)
        a, d = _rescue_check(got, flags, route)
        n_resc += a
        n_diff += d
    assert n_resc > 100
    say(f"rescued {n_resc} pairs, {n_diff} within 1e-3 of the threshold on the other side")


def test_deep_region_config4():
    # 2000 reads of 250 bases x 128 haplotypes of 1000 bases (6.4e10 cells)
    say("deep region 2000 x 128 (250 / 1000 bp): batch, plain and served mixed with small regions")
    deep = workload.deep_region(seed=4)
    b = rv.Batch([deep])
    b.run()
    out, flags = b.results()
    got = b.region_result(out, 0)
    fl = b.region_result(flags, 0)
    b.close()
    # every pair by property: finite, <= 0, FP64 exactly below the threshold
    assert np.isfinite(got).all() and (got <= 0).all() and (fl & 1).all()
    resc = (fl & rv.PAIR_FP64) != 0
    assert (got[resc] < -64.1236 + 0.2).all() and (got[~resc] > -64.1236 - 0.2).all()
    # 512 seeded pairs against the FP64 oracle
    rng = np.random.default_rng(44)
    rr = rng.integers(0, deep["n_reads"], 512)
    hh = rng.integers(0, deep["n_haps"], 512)
    worst = 0.0
    for r, h in zip(rr, hh):
        o0, o1 = deep["read_off"][r], deep["read_off"][r + 1]
        hb = deep["hap_bases"][deep["hap_off"][h]:deep["hap_off"][h + 1]]
        want = oracle.pair_log10_f64(hb, deep["read_bases"][o0:o1], *[deep[k][o0:o1] for k in "qidc"])
        worst = max(worst, abs(got[r, h] - want))
    assert worst < 1e-4, worst
    say(f"deep region: 512 pairs vs FP64 oracle, worst |d| {worst:.2e}; {int(resc.sum())} rescued")
    # the plain one-region path gives the same bits
    alone, af = rv.region_sync(deep)
    assert np.array_equal(alone, got) and np.array_equal(af, fl)
    # served, mixed into a stream of small regions from 8 threads: same bits
    small = [r for r in workload.batched_regions(24, seed=45) if r["n_haps"] * r["n_reads"] < 6000][:16]
    small_alone = [rv.region_sync(r)[0] for r in small]
    jobs = [("deep", deep)] + [(x, r) for x, r in enumerate(small)]
    results, errors = {}, []

    def worker(items):
        try:
            for key, reg in items:
                t = rv.submit(reg)
                results[key] = rv.wait(t, reg)[0]
        except Exception as e:  # noqa: BLE001
            errors.append(e)

    th = [threading.Thread(target=worker, args=(jobs[k::8],)) for k in range(8)]
    for t in th:
        t.start()
    for t in th:
        t.join(timeout=120)
    assert not errors
    assert np.array_equal(results["deep"], got)
    for x, a in enumerate(small_alone):
        m = results[x]
        assert np.array_equal(np.isnan(a), np.isnan(m)) and np.array_equal(a[~np.isnan(a)], m[~np.isnan(m)])


def test_served_poisson_stream_32_producers():
    # the served stream at its stated shape: Poisson arrivals at 8k regions/s
    # for 0.5 s from 32 producer threads, regions of every size of the batched
    # workload (up to 64 haplotypes x 500 reads); every returned matrix equal
    # to the region alone, bit for bit
    say("served stream: 4000 Poisson arrivals at 8000/s from 32 producers, each result == alone")
    pool = workload.batched_regions(48, seed=77)
    assert max(r["n_haps"] * r["n_reads"] for r in pool) > 6000
    alone = [rv.region_sync(r)[0] for r in pool]
    rng = np.random.default_rng(78)
    n = 4000
    arr = np.cumsum(rng.exponential(1 / 8000.0, n))
    order = rng.integers(0, len(pool), n).astype(np.uint32)
    res, lat, outs = rv.loadgen_run(pool, arr, order, 32, outputs=True)
    assert res["completed"] == n and res["failed"] == 0
    for i in range(n):
        a, m = alone[order[i]], outs[i]
        assert np.array_equal(np.isnan(a), np.isnan(m)), i
        assert np.array_equal(a[~np.isnan(a)], m[~np.isnan(m)]), i
    # a sample of the regions against the oracle's reference routing
    for x in [x for x, r in enumerate(pool) if r["n_haps"] * r["n_reads"] < 1500][:3]:
        want, route, ok = oracle.region_reference(pool[x])
        m = ~np.isnan(want)
        assert np.abs(alone[x][m] - want[m]).max() < TOL
    say(f"stream: p50 {res['lat_p50_ms']:.2f} ms, p99 {res['lat_p99_ms']:.2f} ms")


def test_call_pairhmm_submit_wait_equals_call_and_overlaps_cpu_work():
    # Rovaca's worker blocks on call_pairhmm per region
    # (hc/haplotypecaller_engine.cpp:149-289); the two-half call lets it go on
    # with CPU work between submit and wait. Results equal the one-call form.
    say("call_pairhmm_submit / _wait: equal to call_pairhmm; the worker works while the region is on the device")
    rng = np.random.default_rng(101)
    regs = [r for r in workload.batched_regions(16, seed=102) if r["n_haps"] * r["n_reads"] < 6000][:6]
    calls = []
    for reg in regs:
        H, N = reg["n_haps"], reg["n_reads"]
        haps = [reg["hap_bases"][reg["hap_off"][h]:reg["hap_off"][h + 1]] for h in range(H)]
        reads = [reg["read_bases"][reg["read_off"][r]:reg["read_off"][r + 1]] for r in range(N)]
        quals = [reg["q"][reg["read_off"][r]:reg["read_off"][r + 1]] for r in range(N)]
        calls.append((haps, reads, quals, rng.integers(20, 61, N).astype(np.uint8)))
    for pcr in (rv.PCR_NONE, rv.PCR_CONSERVATIVE):
        want = [rv.call_pairhmm(*c, min_quality_threshold=10, pcr=pcr) for c in calls]
        tickets = []
        for c in calls:
            rc, t = rv.call_pairhmm_submit(*c, min_quality_threshold=10, pcr=pcr)
            assert rc == 0
            tickets.append(t)
        for (rc0, m0, k0), t, c in zip(want, tickets, calls):
            rc1, m1, k1 = rv.call_pairhmm_wait(t, len(c[0]), len(c[1]))
            assert rc0 == 0 and rc1 == 0 and np.array_equal(k0, k1)
            assert np.array_equal(np.isnan(m0), np.isnan(m1)) and np.array_equal(m0[~np.isnan(m0)], m1[~np.isnan(m1)])
    # a deep region: submit returns before the device is done, the worker
    # does CPU work of its own, then waits
    deep = workload.deep_region(seed=103)
    H, N = deep["n_haps"], deep["n_reads"]
    haps = [deep["hap_bases"][deep["hap_off"][h]:deep["hap_off"][h + 1]] for h in range(H)]
    reads = [deep["read_bases"][deep["read_off"][r]:deep["read_off"][r + 1]] for r in range(N)]
    quals = [deep["q"][deep["read_off"][r]:deep["read_off"][r + 1]] for r in range(N)]
    mqs = np.full(N, 60, np.uint8)
    rc, t = rv.call_pairhmm_submit(haps, reads, quals, mqs)
    assert rc == 0
    assert rv.poll(t) == 0          # 6.4e10 cells: not done the moment submit returns
    work = sum(int(x) for x in range(200000))  # the worker's own CPU work meanwhile
    assert work > 0
    rc1, m1, k1 = rv.call_pairhmm_wait(t, H, N)
    rc0, m0, k0 = rv.call_pairhmm(haps, reads, quals, mqs)
    assert rc0 == 0 and rc1 == 0 and np.array_equal(k0, k1)
    assert np.array_equal(np.isnan(m0), np.isnan(m1)) and np.array_equal(m0[~np.isnan(m0)], m1[~np.isnan(m1)])
    # a refused region is refused at submit, with nothing queued
    bad = [h.copy() for h in calls[0][1]]
    bad[0][0] = ord("R")
    rc, t = rv.call_pairhmm_submit(calls[0][0], bad, calls[0][2], calls[0][3])
    assert rc == rv.RVC_E_BASES


def test_call_pairhmm_marginalises_sites_on_the_device():
    # the region call with its variant sites: the device marginalises the
    # normalised matrix per site (allele_likelihoods.hpp:701-732); the kept
    # reads' columns equal the oracle's marginalisation of the call's own
    # matrix, exactly (a max is exact)
    say("call_pairhmm with sites: haplotype -> allele marginals from the device")
    rng = np.random.default_rng(111)
    regs = [r for r in workload.batched_regions(20, seed=112) if r["n_haps"] * r["n_reads"] < 8000][:8]
    regs.append(workload.single_region(113))
    pending = []
    for reg in regs:
        H, N = reg["n_haps"], reg["n_reads"]
        haps = [reg["hap_bases"][reg["hap_off"][h]:reg["hap_off"][h + 1]] for h in range(H)]
        reads = [reg["read_bases"][reg["read_off"][r]:reg["read_off"][r + 1]] for r in range(N)]
        quals = [reg["q"][reg["read_off"][r]:reg["read_off"][r + 1]] for r in range(N)]
        mqs = rng.integers(20, 61, N).astype(np.uint8)
        n_sites = int(rng.integers(1, 6))
        n_alleles = rng.integers(1, 5, n_sites).astype(np.uint32)
        hap_allele = np.stack([rng.integers(-1, int(a), H) for a in n_alleles]).astype(np.int32)
        rc, t = rv.call_pairhmm_submit(haps, reads, quals, mqs, min_quality_threshold=10, sites=(n_alleles, hap_allele))
        assert rc == 0
        pending.append((t, H, N, n_alleles, hap_allele, (haps, reads, quals, mqs)))
    for t, H, N, n_alleles, hap_allele, call in pending:
        rc, mat, kept, marg = rv.call_pairhmm_wait(t, H, N, n_alleles=n_alleles)
        assert rc == 0
        rc0, mat0, kept0 = rv.call_pairhmm(*call, min_quality_threshold=10)
        assert np.array_equal(kept, kept0)
        assert np.array_equal(np.isnan(mat), np.isnan(mat0)) and np.array_equal(mat[~np.isnan(mat)], mat0[~np.isnan(mat0)])
        want = oracle.marginalize(mat, len(kept), n_alleles, hap_allele)
        for g, w in zip(marg, want):
            assert np.array_equal(np.isnan(g), np.isnan(w))
            assert np.array_equal(g[~np.isnan(g)], w[~np.isnan(w)])


@pytest.mark.parametrize("lo,hi", [(128, 151), (100, 160)])
def test_eight_lane_long_shapes_equal_default(lo, hi):
    # 8-lane segments of 17-19 rows per lane (reads of 128-151 bases), on the
    # prior-table kernel and, for a haplotype with N, the compare kernel; reads
    # past 151 bases fall back to the 16-lane shape: all the same bits as the
    # default shapes
    say(f"8-lane shapes for reads of {lo}-{hi} bases give the default bits")
    rng = np.random.default_rng(lo + hi)
    regs = []
    for x in range(4):
        base = workload.batched_regions(1, seed=300 + x)[0]
        H = base["n_haps"]
        haps = [base["hap_bases"][base["hap_off"][h]:base["hap_off"][h + 1]] for h in range(H)]
        haps = [h if len(h) >= hi else np.concatenate([h, h, h])[:hi + 40] for h in haps]
        if x == 0:
            haps[1] = haps[1].copy()
            haps[1][3] = ord("N")
        reads, quals = [], []
        for r in range(24):
            n = int(rng.integers(lo, hi + 1))
            src = haps[int(rng.integers(0, H))]
            o = int(rng.integers(0, len(src) - n + 1))
            rd = src[o:o + n].copy()
            rd[rng.random(n) < 0.02] = workload.ACGT[rng.integers(0, 4)]
            reads.append(rd)
            quals.append(rng.integers(10, 41, n).astype(np.uint8))
        regs.append(workload.make_region(haps, reads, quals))
    b = rv.Batch(regs, segment_width=8, shape_fold=False)  # each tile in its own shape
    b.run()
    out, flags = b.results()
    names = {k["name"] for k in b.kernels() if k["precision"] == 32}
    assert any(f"<8,{r}," in n for n in names for r in (17, 18, 19)), names
    for x, reg in enumerate(regs):
        alone, af = rv.region_sync(reg)
        mixed = b.region_result(out, x)
        assert np.array_equal(np.isnan(alone), np.isnan(mixed))
        assert np.array_equal(alone[~np.isnan(alone)], mixed[~np.isnan(mixed)])
        assert np.array_equal(af, b.region_result(flags, x))
    b.close()
    want, route, ok = oracle.region_reference(regs[1])
    got = rv.region_sync(regs[1])[0]
    m = ~np.isnan(want)
    assert np.abs(got[m] - want[m]).max() < TOL
