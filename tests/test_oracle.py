"""The oracle checked against what pins it: Rovaca's known-answer tests and
closed forms, brute-force path enumeration, an unscaled long-double
evaluation, and (where it was built) Rovaca's own intra kernel."""
import os

import numpy as np
import pytest

import oracle
from conftest import GOLDEN
from rovaca_amd import workload


def test_kats_fp32_route(kats):
    # ph/test/pairhmm_unittest.cpp:389-402: relative error < 1e-5
    for hap, rd, q, i, d, c, exp in kats:
        v = oracle.intra_log10(hap, rd, q, i, d, c)
        assert abs(v - exp) / abs(exp) < 1e-5


def test_kats_fp64(kats):
    for hap, rd, q, i, d, c, exp in kats:
        v = oracle.intra_log10(hap, rd, q, i, d, c, use_double=True)
        assert abs(v - exp) / abs(exp) < 1e-5
        assert abs(oracle.pair_log10_f64(hap, rd, q, i, d, c) - v) < 1e-12


def test_closed_forms():
    # ph/test/pairhmm_unittest.cpp:108-141
    q4 = np.full(4, 43, np.uint8)
    assert abs(oracle.intra_log10(b"ACGT", b"ACGT", q4, q4, q4, q4) - (-0.6022797)) < 1e-5 * 0.6022797
    q5 = np.full(5, 43, np.uint8)
    v = oracle.intra_log10(b"ACGTA", b"ACGTA", q5, q5, q5, q5)
    assert abs(v - (-0.69925308227539062)) < 1e-6


def _rand_pair(rng, H, R, alphabet=b"ACGT", nprob=0.0):
    al = np.frombuffer(alphabet, np.uint8)
    hap = al[rng.integers(0, len(al), H)].copy()
    rd = al[rng.integers(0, len(al), R)].copy()
    rd[rng.random(R) < nprob] = ord("N")
    hap[rng.random(H) < nprob] = ord("N")
    q = rng.integers(2, 45, R).astype(np.uint8)
    i = rng.integers(10, 50, R).astype(np.uint8)
    d = rng.integers(10, 50, R).astype(np.uint8)
    c = rng.integers(5, 20, R).astype(np.uint8)
    return hap, rd, q, i, d, c


def test_bruteforce_matches_dp():
    """Explicit enumeration of every state path (independent of the DP) equals
    the long-double DP and the FP64 oracle on tiny pairs, with and without N."""
    rng = np.random.default_rng(7)
    for trial in range(60):
        H, R = int(rng.integers(1, 6)), int(rng.integers(1, 6))
        args = _rand_pair(rng, H, R, nprob=0.2 if trial % 3 == 0 else 0.0)
        bf = oracle.bruteforce_log10(*args)
        ld = oracle.pair_log10_ld(*args)
        f64 = oracle.pair_log10_f64(*args)
        assert abs(bf - ld) < 1e-12, (H, R)
        assert abs(f64 - ld) < 1e-10, (H, R)


def test_planted_mistake_is_caught():
    """A transposed operand (insertion and deletion qualities swapped) moves the
    oracle away from brute force on asymmetric inputs."""
    rng = np.random.default_rng(3)
    hap, rd, q, i, d, c = _rand_pair(rng, 4, 4)
    i[:] = 20
    d[:] = 40
    bf = oracle.bruteforce_log10(hap, rd, q, i, d, c)
    wrong = oracle.pair_log10_ld(hap, rd, q, d, i, c)
    assert abs(bf - wrong) > 1e-6


def test_scaling_constant_independence():
    """FP64 with C = 2^1020 equals the unscaled long-double evaluation, so the
    constant cancels exactly as the log10 offset says."""
    rng = np.random.default_rng(11)
    for _ in range(10):
        args = _rand_pair(rng, int(rng.integers(20, 120)), int(rng.integers(10, 90)), nprob=0.01)
        assert abs(oracle.pair_log10_f64(*args) - oracle.pair_log10_ld(*args)) < 1e-9


def test_n_is_wildcard_symmetric():
    rng = np.random.default_rng(5)
    hap, rd, q, i, d, c = _rand_pair(rng, 30, 20)
    rd2 = rd.copy()
    rd2[5] = ord("N")
    hap2 = hap.copy()
    # N in the read matches whatever the haplotype holds: equal to a read whose base 5 equals the aligned hap base
    v_n = oracle.pair_log10_f64(hap, rd2, q, i, d, c)
    v_a = [oracle.pair_log10_f64(hap2, np.where(np.arange(20) == 5, b, rd), q, i, d, c) for b in b"ACGT"]
    assert v_n >= max(v_a) - 1e-12


@pytest.mark.skipif(oracle.ref() is None, reason="oracle/_ref (Rovaca's intra kernel) not built")
def test_against_reference_intra_kernel():
    """The restatement equals Rovaca's compiled intra kernel: bit for bit on
    nearly every pair; the FP32 route may differ by a float ulp of log10f where
    the reference's compiler fused a multiply-add."""
    rng = np.random.default_rng(13)
    exact = 0
    for trial in range(200):
        H, R = int(rng.integers(1, 300)), int(rng.integers(1, 200))
        args = _rand_pair(rng, H, R, nprob=0.01 if trial % 2 else 0.0)
        a, b = oracle.intra_log10(*args), oracle.ref_intra_log10(*args)
        assert abs(a - b) <= 8e-6 * max(1.0, abs(b) / 64)
        exact += a == b
        assert abs(oracle.intra_log10(*args, use_double=True) - oracle.ref_intra_log10(*args, use_double=True)) < 1e-9
    assert exact >= 190


def test_region_routing_and_golden():
    """Region routing restatement on the single-region workload: every pair is
    within 1e-5 of the FP64 oracle and of the committed reference fixture."""
    import os
    from conftest import GOLDEN
    g = np.load(os.path.join(GOLDEN, "single_region_seed0.npz"))
    reg = {k: g[k] for k in ("hap_off", "hap_bases", "read_off", "read_bases", "q", "i", "d", "c")}
    reg["n_haps"], reg["n_reads"] = len(reg["hap_off"]) - 1, len(reg["read_off"]) - 1
    out, route, ok = oracle.region_reference(reg)
    assert ok
    assert np.abs(out - g["ref_intra"]).max() < 1e-5
    assert np.abs(out - g["fp64"]).max() < 1e-5
    assert set(np.unique(route)) <= {1, 2, 3, 4, 5}


# Known answers of Rovaca's own repeat-count test
# (genotype/test/lush_variant_context_utils_unit_test.cpp:349-370); the
# counter there and the one the PCR indel model uses
# (pairhmm/pairhmm_internal.cpp:510-542) walk the test string the same way.
REPETITION_KATS = [
    ("AT", "GATAT", False, 2), ("AT", "GATAT", True, 0), ("A", "ATATG", True, 1), ("AT", "ATATG", True, 2),
    ("CCC", "CCCCCCCC", True, 2), ("CCC", "CCCCCCCC", False, 2),
    ("ATG", "ATGATGATGATG", True, 4), ("G", "ATGATGATGATG", True, 0), ("T", "T", True, 1),
    ("AT", "ATGATGATCATG", True, 1), ("CCCCCCCC", "CCC", True, 0), ("AT", "AT", True, 1), ("AT", "", True, 0),
    ("ATG", "ATGATGATGATG", False, 4), ("G", "ATGATGATGATG", False, 1), ("T", "T", False, 1),
    ("AT", "ATGATGATCATG", False, 0), ("CCCCCCCC", "CCC", False, 0), ("AT", "AT", False, 1), ("AT", "", False, 0),
]


@pytest.mark.parametrize("unit,test,leading,expected", REPETITION_KATS)
def test_count_repetitions_kats(unit, test, leading, expected):
    assert oracle.count_repetitions(unit.encode(), test.encode(), leading) == expected


def test_repeat_length_closed_forms():
    # A read that is one homopolymer of length L has repeat length min(L, 20)
    # at every base: backward and forward runs of the one-base unit add up to L.
    for L in (2, 5, 12, 20, 21, 33):
        assert [oracle.repeat_length(b"A" * L, x) for x in range(L)] == [min(L, 20)] * L
    # Dinucleotide ACACACAC. Base 0: no backward unit repeats, forward unit CA
    # occurs 3 times and once-shifted AC has no copy in "A": 0 + 3.  Base 1:
    # forward unit AC occurs 3 times, its backward count over "AC" is 1: 1 + 3.
    assert oracle.repeat_length(b"ACACACAC", 0) == 3
    assert oracle.repeat_length(b"ACACACAC", 1) == 4
    # A run inside unrelated bases: GCTTTTTCA, base 3 (second T) sees two
    # T's behind it (including itself) and three ahead.
    assert oracle.repeat_length(b"GCTTTTTCA", 3) == 5
    # No repeat at all: the forward count of the whole tail is 1.
    assert oracle.repeat_length(b"ACGTTGCA", 0) == 1


def test_pcr_gop_follows_repeat_length():
    # HOSTILE caps a repeat of length 12 at GOP 10, AGGRESSIVE at 34,
    # CONSERVATIVE at 37 (pairhmm/adapter.h:13-15); the last base keeps 45.
    for pcr, cap in ((1, 10), (2, 34), (3, 37)):
        _, gi, gd, _ = oracle.preprocess_read(b"A" * 12, np.full(12, 30, np.uint8), 60, 0, pcr)
        assert list(gi) == [cap] * 11 + [45]
        assert list(gd) == list(gi)
    # Without a repeat (length 1 at both inner bases of ACG) HOSTILE caps at 40.
    assert oracle.repeat_length(b"ACG", 0) == oracle.repeat_length(b"ACG", 1) == 1
    _, gi, _, _ = oracle.preprocess_read(b"ACG", np.full(3, 30, np.uint8), 60, 0, 1)
    assert list(gi) == [40, 40, 45]


def test_oracle_on_real_regions_against_reference_fixture():
    # the 8 real chrY regions of Rovaca's schedule test (ph/resource/pairhmm_debug.txt):
    # the oracle's reference routing against Rovaca's intra kernel at 1e-5
    # (ph/test/pairhmm_unitest_from_assemble.cpp:158)
    from conftest import load_debug_regions
    regs = load_debug_regions()
    ref = np.load(os.path.join(GOLDEN, "pairhmm_debug_ref.npz"))
    o = 0
    for reg in regs:
        n = reg["n_haps"] * reg["n_reads"]
        want = ref["ref_intra"][o:o + n].reshape(reg["n_reads"], reg["n_haps"])
        f64 = ref["fp64"][o:o + n].reshape(reg["n_reads"], reg["n_haps"])
        o += n
        got, route, ok = oracle.region_reference(reg)
        assert ok and np.isfinite(got).all()
        assert np.abs(got - want).max() <= 1e-5
        assert np.abs(f64 - want).max() <= 1e-5
    assert o == len(ref["ref_intra"])


def test_marginalize_closed_form():
    # allele_likelihoods.hpp:701-732: per read, the largest likelihood of the
    # haplotypes mapped to the allele, -inf for an allele no haplotype maps
    # to, a NaN never the largest; unmapped haplotypes skipped; columns past
    # the evidence count (dropped reads) NaN
    hm = np.array([[-1.0, -5.0, np.nan], [-2.0, np.nan, -3.0], [-0.5, -7.0, -1.0]])
    s0, s1 = oracle.marginalize(hm, 2, [2, 3], [[0, 1, 0], [-1, 0, 0]])
    assert s0[0, :2].tolist() == [-0.5, -5.0] and s0[1, :2].tolist() == [-2.0, -np.inf]
    assert s1[0, :2].tolist() == [-0.5, -7.0] and np.isneginf(s1[1:, :2]).all()
    assert np.isnan(s0[:, 2]).all() and np.isnan(s1[:, 2]).all()
    with pytest.raises(ValueError):
        oracle.marginalize(hm, 2, [2], [[0, 2, 0]])


def test_marginalize_matches_the_reference_tests_property():
    # the check of AlleleLikelihoodsUnitTest.testMarginalization
    # (genotype/test/allele_likelihoods_unit_test.cpp:210-244): each new value
    # equals the max over its old alleles, on random likelihoods
    rng = np.random.default_rng(7)
    for _ in range(50):
        H, N = int(rng.integers(1, 12)), int(rng.integers(1, 20))
        hm = rng.uniform(-30, 0, (H, N))
        A = int(rng.integers(1, 5))
        m = rng.integers(-1, A, H)
        (got,) = oracle.marginalize(hm, N, [A], [m])
        for a in range(A):
            old = [h for h in range(H) if m[h] == a]
            want = hm[old].max(axis=0) if old else np.full(N, -np.inf)
            assert np.array_equal(got[a], want)
