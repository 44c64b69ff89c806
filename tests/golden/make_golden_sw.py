"""Writes the committed golden fixture of the aligner tests: seeded pairs and
the outputs of Rovaca's own aligner (sw_avx_init + avx_function,
hc/common/base/lib/sw_avx/sw_avx.c:402-451 over the ro_sw cores), built from
the reference sources into oracle/_ref by oracle/Makefile.

Inputs are stored with the outputs, so the fixture does not depend on the
generator staying unchanged. Per pair and score set: the CIGAR in BAM
encoding (length << 4 | op), the alignment offset and the score; a refused
pair has count -1, offset 1, score 0.
Run from the repository root: python tests/golden/make_golden_sw.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))
import oracle  # noqa: E402
from test_oracle_sw import ACGT, _mutate, random_pairs  # noqa: E402

OPS = {op: k for k, op in enumerate("MIDNSHP=X")}


def pairs():
    out = random_pairs(1001, 240, max_hap=500, max_read=300)
    rng = np.random.default_rng(1002)
    # short reads (the 32-bit core below 16 bases), repeats (ties), long reads
    for L in (1, 5, 15, 16, 40, 150, 250, 600):
        hap = ACGT[rng.integers(0, 4, L + 300)].tobytes()
        out.append((hap, _mutate(rng, hap[100:100 + L], 2) or b"A"))
    out += [(b"ACACACACACACACACACAC", b"CACACA"), (b"A" * 60, b"A" * 20), (b"A" * 2048, b"A" * 10)]
    return out


def main():
    ps = pairs()
    haps = [h for h, _ in ps]
    reads = [r for _, r in ps]
    res = {}
    for asm in (0, 1):
        cig, cnt, off, sc = [], [], [], []
        for h, r in ps:
            c, o, s = oracle.ref_sw_align(h, r, bool(asm))
            if c is None:
                cnt.append(-1)
                off.append(o)
                sc.append(0)
                continue
            cnt.append(len(c))
            cig += [(n << 4) | OPS[op] for op, n in c]
            off.append(o)
            sc.append(s)
        res[f"cigar_{asm}"] = np.array(cig, np.uint32)
        res[f"count_{asm}"] = np.array(cnt, np.int32)
        res[f"offset_{asm}"] = np.array(off, np.int32)
        res[f"score_{asm}"] = np.array(sc, np.int32)
    np.savez_compressed(
        os.path.join(ROOT, "tests", "golden", "sw_reference_pairs.npz"),
        hap_bases=np.frombuffer(b"".join(haps), np.uint8), hap_off=np.cumsum([0] + [len(h) for h in haps]),
        read_bases=np.frombuffer(b"".join(reads), np.uint8), read_off=np.cumsum([0] + [len(r) for r in reads]),
        **res)
    print(len(ps), "pairs")


if __name__ == "__main__":
    main()
