"""The Smith-Waterman oracle (read against its best haplotype) pinned against
Rovaca's own aligner compiled from its sources (oracle/_ref), closed forms and
brute force on tiny inputs."""
import itertools
import os
import sys

import numpy as np
import pytest

import oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

ACGT = np.frombuffer(b"ACGT", np.uint8)


def _mutate(rng, s, n_edits):
    s = bytearray(s)
    for _ in range(n_edits):
        k, p = int(rng.integers(0, 3)), int(rng.integers(0, len(s)))
        if k == 0:
            s[p] = int(ACGT[rng.integers(0, 4)])
        elif k == 1:
            s.insert(p, int(ACGT[rng.integers(0, 4)]))
        elif len(s) > 1:
            del s[p]
    return bytes(s)


def random_pairs(seed, n, max_hap=400, max_read=300):
    """Reads that are edited pieces of their haplotype (as after assembly),
    pieces with foreign ends, and unrelated reads."""
    rng = np.random.default_rng(seed)
    out = []
    for t in range(n):
        H = int(rng.integers(1, max_hap))
        hap = ACGT[rng.integers(0, 4, H)].tobytes()
        if t % 4 == 0:
            read = ACGT[rng.integers(0, 4, int(rng.integers(1, max_read)))].tobytes()
        else:
            a = int(rng.integers(0, H))
            b = int(rng.integers(a + 1, H + 1))
            read = _mutate(rng, hap[a:b], int(rng.integers(0, 6))) or b"A"
            if t % 4 == 3:
                read = ACGT[rng.integers(0, 4, 5)].tobytes() + read + ACGT[rng.integers(0, 4, 7)].tobytes()
        out.append((hap, read))
    return out


@pytest.mark.skipif(oracle.ref_sw() is None, reason="oracle/_ref/libref_sw.so is not built")
@pytest.mark.parametrize("assemble", [False, True])
def test_matches_reference_aligner(assemble):
    for hap, read in random_pairs(11 + assemble, 1500):
        assert oracle.sw_align(hap, read, assemble) == oracle.ref_sw_align(hap, read, assemble), (hap, read)


@pytest.mark.skipif(oracle.ref_sw() is None, reason="oracle/_ref/libref_sw.so is not built")
def test_matches_reference_aligner_edges():
    # short reads (the 32-bit core below 16 bases), single bases, reads longer
    # than the haplotype, repeats (ties), the 2048 limit
    rng = np.random.default_rng(5)
    cases = [(b"A", b"A"), (b"A", b"C"), (b"AC", b"A"), (b"A", b"ACGT"), (b"ACGTACGT", b"ACGTACGTACGTACGTACGT"),
             (b"AAAAAAAAAAAAAAAAAAAA", b"AAAA"), (b"AAAAAAAAAA", b"AAAAAAAAAAAAAAAAAAAAAAAAA"),
             (b"ACACACACACACACAC", b"CACACA"), (b"A" * 2047, b"A" * 10), (b"A" * 2048, b"A" * 10)]
    for L in (1, 2, 7, 8, 15, 16, 17, 31, 32, 33):
        hap = ACGT[rng.integers(0, 4, 3 * L + 5)].tobytes()
        cases.append((hap, hap[L:2 * L]))
        cases.append((hap, _mutate(rng, hap[L:2 * L], 1) or b"G"))
        cases.append((hap[:L], ACGT[rng.integers(0, 4, L)].tobytes()))
    for hap, read in cases:
        for assemble in (False, True):
            assert oracle.sw_align(hap, read, assemble) == oracle.ref_sw_align(hap, read, assemble), (hap, read)


@pytest.mark.skipif(oracle.ref_sw() is None, reason="oracle/_ref/libref_sw.so is not built")
def test_long_reads_match_reference_aligner():
    # reads of 1025-2047 bases (the device's 2048-row shape) against the
    # reference library, both score sets
    rng = np.random.default_rng(2047)
    for n1, n2 in ((2047, 1025), (1500, 2047), (2047, 2047), (900, 1400)):
        hap = ACGT[rng.integers(0, 4, n1)].tobytes()
        o = int(rng.integers(0, max(1, n1 - n2)))
        read = bytearray(hap[o:o + n2]) + ACGT[rng.integers(0, 4, max(0, n2 - (n1 - o)))].tobytes()
        for x in rng.integers(0, n2, n2 // 40):
            read[x] = ACGT[rng.integers(0, 4)]
        for assemble in (False, True):
            assert oracle.sw_align(hap, bytes(read), assemble) == oracle.ref_sw_align(hap, bytes(read), assemble)


@pytest.mark.skipif(oracle.ref_sw() is None, reason="oracle/_ref/libref_sw.so is not built")
def test_empty_sequences_match_a_fresh_reference_aligner():
    # An empty haplotype or read is not refused by Rovaca's aligner (only a
    # length >= 2048 is, sw_core_template.h:356-361). Its instance keeps the
    # previous call's CIGAR count when the read is empty, so each case runs in
    # a process of its own, on a fresh instance.
    import subprocess
    cases = [(b"", b"ACGT"), (b"", b"ACGTACGTACGTACGTACGTAA"), (b"", b"A" * 1500), (b"ACGT", b""),
             (b"A" * 2047, b""), (b"", b""), (b"A" * 2048, b"")]
    for hap, read in cases:
        for assemble in (False, True):
            code = (f"import oracle; print(repr(oracle.ref_sw_align({hap!r}, {read!r}, {assemble})))")
            out = subprocess.run([sys.executable, "-c", code], cwd=ROOT, check=True, capture_output=True,
                                 text=True).stdout.strip()
            assert repr(oracle.sw_align(hap, read, assemble)) == out, (len(hap), len(read), assemble)
    # closed forms: every read base soft-clipped against an empty haplotype
    assert oracle.sw_align(b"", b"ACGTA") == ([("S", 5)], 0, 0)
    assert oracle.sw_align(b"ACG", b"") == ([], 3, 0)


def test_closed_forms():
    hap = b"GATTACACCGTAGCTTAGGCATCGATCGGATCCATGCA"
    # a read equal to a piece of the haplotype: one match run at its offset
    for a, b in [(0, len(hap)), (5, 25), (13, 21), (10, 30)]:  # pieces that occur once
        assert oracle.sw_align(hap, hap[a:b]) == ([("M", b - a)], a, 10 * (b - a))
    # one substitution inside long flanks stays a match run (-15 < clipping a flank)
    read = bytearray(hap[4:34])
    read[15] = ord("A") if read[15] != ord("A") else ord("C")
    assert oracle.sw_align(hap, bytes(read)) == ([("M", 30)], 4, 10 * 29 - 15)
    # one inserted and one deleted base inside 20-base flanks: gap open -30
    ins = hap[2:22] + b"T" + hap[22:37]
    cig, off, sc = oracle.sw_align(hap, ins)
    assert off == 2 and sc == 10 * 35 - 30 and [op for op, _ in cig] == ["M", "I", "M"] and cig[1] == ("I", 1)
    dele = hap[1:20] + hap[21:38]
    cig, off, sc = oracle.sw_align(hap, dele)
    assert off == 1 and sc == 10 * 36 - 30 and [op for op, _ in cig] == ["M", "D", "M"] and cig[1] == ("D", 1)
    # read bases hanging over either end of the haplotype become soft clips
    # (elsewhere the read has to be aligned: the alignment starts on row 0 or
    # column 0 and ends on the last row or column)
    assert oracle.sw_align(hap, b"TTTTT" + hap[:20]) == ([("S", 5), ("M", 20)], 0, 200)
    assert oracle.sw_align(hap, hap[18:] + b"GGGG") == ([("M", len(hap) - 18), ("S", 4)], 18, 10 * (len(hap) - 18))
    # the refused lengths
    assert oracle.sw_align(b"A" * 2048, b"AAAA") == (None, 1, None)


def _brute_best(hap, read, p):
    """Best score over all alignments from a first-row/first-column cell
    (score 0) to a last-row/last-column cell, by exhaustive path search."""
    match, mismatch, gap_open, gap_ext = p
    n1, n2 = len(hap), len(read)
    memo = {}

    def best_from(i, j, state):
        # best score of a path that starts at (i, j) and ends on the last row
        # or column; state: 0 after a match/start, 1 inside an insertion, 2 inside a deletion
        key = (i, j, state)
        if key in memo:
            return memo[key]
        res = 0 if (i == n1 or j == n2) else None
        cands = []
        if i < n1 and j < n2:
            cands.append(best_from(i + 1, j + 1, 0) + (match if hap[i] == read[j] else mismatch))
        if j < n2 and i > 0:
            cands.append(best_from(i, j + 1, 1) + (gap_ext if state == 1 else gap_open))
        if i < n1 and j > 0:
            cands.append(best_from(i + 1, j, 2) + (gap_ext if state == 2 else gap_open))
        cands = [c for c in cands if c is not None]
        if cands:
            res = max(cands) if res is None else max(res, max(cands))
        memo[key] = res
        return res

    starts = [(i, 0) for i in range(n1 + 1)] + [(0, j) for j in range(1, n2 + 1)]
    # an alignment must take at least one cell step
    best = None
    for i, j in starts:
        if i < n1 and j < n2:
            v = best_from(i + 1, j + 1, 0) + (match if hap[i] == read[j] else mismatch)
            best = v if best is None else max(best, v)
        if j < n2 and i > 0:
            v = best_from(i, j + 1, 1) + gap_open
            best = max(best, v) if best is not None else v
        if i < n1 and j > 0:
            v = best_from(i + 1, j, 2) + gap_open
            best = max(best, v) if best is not None else v
    return best


def test_score_is_the_best_alignment_by_brute_force():
    rng = np.random.default_rng(8)
    for n1, n2 in itertools.product(range(1, 6), range(1, 6)):
        for _ in range(3):
            hap = ACGT[rng.integers(0, 2, n1)].tobytes()  # two letters: more ties and gaps
            read = ACGT[rng.integers(0, 2, n2)].tobytes()
            for assemble in (False, True):
                p = oracle.SW_ASSEMBLE if assemble else oracle.SW_GENOTYPE
                _, _, sc = oracle.sw_align(hap, read, assemble)
                assert sc == _brute_best(hap, read, p), (hap, read, assemble)


def reference_vectors(assemble):
    """The committed fixture (tests/golden/make_golden_sw.py): pairs and the
    outputs of Rovaca's own aligner, as (haps, reads, expected results)."""
    import os
    z = np.load(os.path.join(os.path.dirname(__file__), "golden", "sw_reference_pairs.npz"))
    hb, ho, rb, ro = z["hap_bases"].tobytes(), z["hap_off"], z["read_bases"].tobytes(), z["read_off"]
    haps = [hb[ho[k]:ho[k + 1]] for k in range(len(ho) - 1)]
    reads = [rb[ro[k]:ro[k + 1]] for k in range(len(ro) - 1)]
    a = int(bool(assemble))
    cig, cnt, off, sc = z[f"cigar_{a}"], z[f"count_{a}"], z[f"offset_{a}"], z[f"score_{a}"]
    want, pos = [], 0
    for k in range(len(reads)):
        if cnt[k] < 0:
            want.append((None, int(off[k]), None))
            continue
        c = cig[pos:pos + cnt[k]]
        pos += int(cnt[k])
        want.append(([("MIDNSHP=X"[int(x) & 15], int(x) >> 4) for x in c], int(off[k]), int(sc[k])))
    return haps, reads, want


@pytest.mark.parametrize("assemble", [False, True])
def test_matches_committed_reference_vectors(assemble):
    haps, reads, want = reference_vectors(assemble)
    assert len(want) > 200
    for h, r, w in zip(haps, reads, want):
        assert oracle.sw_align(h, r, assemble) == w, (h, r)


def pairs_with_other_bytes(seed, n):
    """Pairs whose reads and haplotypes carry N (and a few other bytes) as
    real reads do: the aligner compares raw bytes, so N matches N only."""
    rng = np.random.default_rng(seed)
    out = []
    for hap, read in random_pairs(seed, n):
        h = np.frombuffer(hap, np.uint8).copy()
        r = np.frombuffer(read, np.uint8).copy()
        h[rng.random(len(h)) < 0.02] = ord("N")
        r[rng.random(len(r)) < 0.05] = ord("N")
        if len(r) > 3 and rng.random() < 0.3:
            r[int(rng.integers(0, len(r)))] = ord("a")  # a lowercase base: no match to 'A'
        out.append((h.tobytes(), r.tobytes()))
    return out


@pytest.mark.skipif(oracle.ref_sw() is None, reason="oracle/_ref/libref_sw.so is not built")
@pytest.mark.parametrize("assemble", [False, True])
def test_matches_reference_aligner_with_n_bases(assemble):
    for hap, read in pairs_with_other_bytes(31 + assemble, 400):
        assert oracle.sw_align(hap, read, assemble) == oracle.ref_sw_align(hap, read, assemble), (hap, read)
