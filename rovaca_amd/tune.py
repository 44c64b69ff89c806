"""Tile-shape tuner for the FP32 kernel.

For every read length class of the 16-lane kernel (RPL = 1..16 rows per
lane, reads of 16*RPL bases) it times each candidate shape on the device --
16 lanes per read at 2, 3 and 4 waves per SIMD, and 32 or 64 lanes per read --
on a fixed synthetic batch, and writes the fastest as a table that the engine
loads at init (tuned/mi355x.txt beside the library):

    occ16   <waves per SIMD of the 16-lane kernel, index = RPL 0..16>
    shape16 <0 = 16 lanes, else W*256 + RPL, index = RPL 0..16>

Usage: python -m rovaca_amd.tune [--out PATH] [--regions N]
"""
import argparse
import datetime
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def base_pool(n_regions, seed=11):
    """Regions of 64 haplotypes (~350 bp) x 256 reads of 256 bases."""
    from . import workload
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n_regions):
        hap0 = workload.ACGT[rng.integers(0, 4, 350 + int(rng.integers(0, 40)))]
        haps = [hap0] + [workload._mutate(rng, hap0) for _ in range(63)]
        reads, quals = workload._reads(rng, [np.concatenate([h, h]) for h in haps], 256, 256)
        out.append((haps, reads, quals))
    return out


def truncated(pool, R):
    from . import workload
    return [workload.make_region(h, [r[:R] for r in rd], [q[:R] for q in qs]) for h, rd, qs in pool]


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=os.path.join(HERE, "tuned", "mi355x.txt"))
    ap.add_argument("--regions", type=int, default=96)
    ap.add_argument("--repeats", type=int, default=3)
    args = ap.parse_args()

    import rovaca_amd as rv
    from . import workload
    t0 = time.time()
    pool = base_pool(args.regions)
    print(f"[tune] pool of {len(pool)} regions in {time.time() - t0:.1f} s", flush=True)
    rv.init(0, tune_table=os.devnull)  # start from the built-in defaults
    occ16 = [4] * 17
    shape16 = [0] * 17
    report = []
    for rpl in range(1, 17):
        R = 16 * rpl
        regs = truncated(pool, R)
        cells = workload.cells(regs)
        cands = [(16, o) for o in (2, 3, 4)]
        for w in (32, 64):
            if -(-R // w) <= (16 if w == 32 else 12):
                cands.append((w, 2))
        best = None
        best16 = None
        for w, occ in cands:
            b = rv.Batch(regs, rpl_round=1, occupancy=occ, segment_width=w)
            b.run()
            ms = min((b.run(), b.info()["fp32_ms"])[1] for _ in range(args.repeats))
            b.close()
            gcups = cells / (ms * 1e-3) / 1e9
            report.append((rpl, w, occ, ms, gcups))
            print(f"[tune] R={R:3d} W={w:2d} occ={occ}: {ms:8.3f} ms {gcups:8.1f} GCUPS", flush=True)
            if best is None or ms < best[2]:
                best = (w, occ, ms)
            if w == 16 and (best16 is None or ms < best16[1]):
                best16 = (occ, ms)
        # the 16-lane occupancy is kept even where a wider shape wins, because
        # reads of other lengths can round up to this row count
        occ16[rpl] = best16[0]
        w, occ, _ = best
        if w != 16:
            shape16[rpl] = w * 256 + (-(-R // w))
    occ16[0] = occ16[1]
    rv.shutdown()
    os.makedirs(os.path.dirname(args.out), exist_ok=True)
    with open(args.out, "w") as f:
        f.write(f"# FP32 kernel shapes tuned by rovaca_amd/tune.py on {datetime.date.today()}\n")
        f.write(f"# batch: {args.regions} regions x 64 haplotypes (~350 bp) x 256 reads of 16*RPL bases\n")
        for rpl, w, occ, ms, g in report:
            f.write(f"# RPL {rpl:2d} W {w:2d} occ {occ}: {ms:.3f} ms, {g:.1f} GCUPS\n")
        f.write("occ16 " + " ".join(str(x) for x in occ16) + "\n")
        f.write("shape16 " + " ".join(str(x) for x in shape16) + "\n")
        f.write("steps3 1\nfold 1\n")
    print(f"[tune] wrote {args.out}", flush=True)


if __name__ == "__main__":
    sys.exit(main())
