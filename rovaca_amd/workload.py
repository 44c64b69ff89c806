"""Synthetic active regions for the PairHMM workloads.

A region is a dict of flat numpy arrays in the layout of ``rvc_region``
(include/rovaca_phmm.h): ``hap_off``/``hap_bases`` and
``read_off``/``read_bases`` plus per read base the Phred arrays ``q``, ``i``,
``d``, ``c``. Gap-open quality is 45 and gap continuation 10, the values
Rovaca's scheduler uses without a PCR indel model
(hc/genotype/read_record.cpp:184-210, ph/roDV_schedule.cpp:509-519).

Workloads:
  single_region(seed)       8 haplotypes x 32 reads: haplotype 0 uniform ACGT
                            of 300 bp, haplotypes 1-7 carry 1-3 SNPs and 0-1
                            indels of 1-10 bp; reads are 150 bp substrings of a
                            random haplotype; quals 90 % U[30,40], 10 % U[2,20];
                            each base substituted with probability 10^(-Q/10);
                            0.1 % of read bases are N.
  batched_regions(n, seed)  haplotypes U[2,64], reads log-uniform[10,500],
                            haplotype length U[100,600], read length one of
                            {100, 150, 250} per region.
  stress_regions(n, seed)   as batched_regions, but 30 % of reads are 250 bp
                            with a 40 % substitution rate and quals U[2,10].
  deep_region(seed)         2000 reads of 250 bp x 128 haplotypes of 1000 bp.
When a read is longer than its source haplotype it is cut to that length.
"""
import numpy as np

ACGT = np.frombuffer(b"ACGT", np.uint8)
GOP = 45
GCP = 10


def _mutate(rng, hap0):
    """hap0 with 1-3 SNPs and 0-1 indels of 1-10 bp."""
    h = hap0.copy()
    for _ in range(rng.integers(1, 4)):
        p = rng.integers(0, len(h))
        choices = ACGT[ACGT != h[p]]
        h[p] = choices[rng.integers(0, 3)]
    if rng.integers(0, 2):
        L = int(rng.integers(1, 11))
        p = int(rng.integers(0, max(1, len(h) - L)))
        if rng.integers(0, 2):
            h = np.concatenate([h[:p], ACGT[rng.integers(0, 4, L)], h[p:]])
        else:
            h = np.concatenate([h[:p], h[p + L:]])
    return h


def _reads(rng, haps, n, length, low=False, sub_rate=None):
    """n reads of up to `length` bases, each a substring of a random haplotype
    at a uniform offset, with quality-driven substitutions and 0.1 % N."""
    if n == 0:
        return [], []
    lens = np.array([len(h) for h in haps])
    src = rng.integers(0, len(haps), n)
    L = np.minimum(length, lens[src])
    off = (rng.random(n) * (lens[src] - L + 1)).astype(np.int64)
    width = int(L.max())
    pad = np.zeros((len(haps), int(lens.max()) + width), np.uint8)
    for x, h in enumerate(haps):
        pad[x, :len(h)] = h
    seq = pad[src[:, None], off[:, None] + np.arange(width)[None, :]]
    if low:
        q = rng.integers(2, 11, (n, width)).astype(np.uint8)
    else:
        q = np.where(rng.random((n, width)) < 0.9, rng.integers(30, 41, (n, width)),
                     rng.integers(2, 21, (n, width))).astype(np.uint8)
    p = 10.0 ** (-q.astype(np.float64) / 10.0) if sub_rate is None else np.full((n, width), sub_rate)
    sub = rng.random((n, width)) < p
    code = np.searchsorted(ACGT, seq)
    # a substituted base becomes one of the three other bases
    seq = np.where(sub, ACGT[(code + rng.integers(1, 4, (n, width))) % 4], seq).astype(np.uint8)
    seq[rng.random((n, width)) < 0.001] = ord("N")
    return [seq[x, :L[x]] for x in range(n)], [q[x, :L[x]] for x in range(n)]


def make_region(haps, reads, quals):
    hap_off = np.zeros(len(haps) + 1, np.int32)
    hap_off[1:] = np.cumsum([len(h) for h in haps])
    read_off = np.zeros(len(reads) + 1, np.int32)
    read_off[1:] = np.cumsum([len(r) for r in reads])
    total = int(read_off[-1])
    return {
        "n_haps": len(haps), "n_reads": len(reads),
        "hap_off": hap_off,
        "hap_bases": np.concatenate(list(haps) + [np.zeros(1, np.uint8)]).astype(np.uint8),
        "read_off": read_off,
        "read_bases": np.concatenate(list(reads) + [np.zeros(1, np.uint8)]).astype(np.uint8),
        "q": np.concatenate(list(quals) + [np.zeros(1, np.uint8)]).astype(np.uint8),
        "i": np.full(total + 1, GOP, np.uint8),
        "d": np.full(total + 1, GOP, np.uint8),
        "c": np.full(total + 1, GCP, np.uint8),
    }


def _region(rng, n_haps, n_reads, hap_len, read_len, stress=False):
    hap0 = ACGT[rng.integers(0, 4, hap_len)]
    haps = [hap0] + [_mutate(rng, hap0) for _ in range(n_haps - 1)]
    n_bad = int((rng.random(n_reads) < 0.3).sum()) if stress else 0
    reads, quals = _reads(rng, haps, n_reads - n_bad, read_len)
    r2, q2 = _reads(rng, haps, n_bad, 250, low=True, sub_rate=0.4)
    return make_region(haps, reads + r2, quals + q2)


def single_region(seed=0):
    rng = np.random.default_rng(seed)
    return _region(rng, 8, 32, 300, 150)


def _batched(n, seed, stress):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        n_haps = int(rng.integers(2, 65))
        n_reads = int(np.exp(rng.uniform(np.log(10), np.log(500))))
        hap_len = int(rng.integers(100, 601))
        read_len = int(rng.choice([100, 150, 250]))
        out.append(_region(rng, n_haps, n_reads, hap_len, read_len, stress))
    return out


def batched_regions(n=4096, seed=1):
    return _batched(n, seed, False)


def stress_regions(n=2048, seed=2):
    return _batched(n, seed, True)


def deep_region(seed=4):
    rng = np.random.default_rng(seed)
    return _region(rng, 128, 2000, 1000, 250)


def cells(regions):
    """Sum over pairs of read_len * hap_len (the GCUPS numerator)."""
    tot = 0
    for r in regions:
        hl = np.diff(r["hap_off"]).astype(np.int64)
        rl = np.diff(r["read_off"]).astype(np.int64)
        tot += int(hl.sum()) * int(rl.sum())
    return tot
