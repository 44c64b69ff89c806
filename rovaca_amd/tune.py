"""Tile-shape tuner for the FP32 kernel.

For every read length class of the 16-lane kernel (RPL = 1..16 rows per
lane, reads of 16*RPL - 1 bases: a segment always keeps one padding row) it times, on a fixed synthetic batch, the
compare-and-select kernel at 2, 3 and 4 waves per SIMD and the prior-table
kernel with 8, 16, 32 and 64 lanes per read at each occupancy, and writes the
fastest as a table that the engine loads at init (tuned/mi355x.txt beside the
library):

    occ16   <waves per SIMD, compare-and-select 16-lane kernel, index = RPL 0..16>
    occpt16 <the same for the prior-table 16-lane kernel>
    occpt32 <the same for the prior-table 32-lane kernel, index = its own RPL>
    occpt8  <the same for the prior-table 8-lane kernel, index = its own RPL>
    shape16 <0 = 16 lanes, else W*256 + RPL (RPL 0: fewest rows of W lanes), index = RPL 0..16>

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
    occ16, occpt16, occpt32, occpt8 = [4] * 17, [4] * 17, [4] * 17, [4] * 17  # rows past 16 run at 2
    shape16 = [0] * 17
    report = []

    def time_batch(regs, cells, **kw):
        b = rv.Batch(regs, rpl_round=1, **kw)
        b.run()
        ms = min((b.run(), b.info()["fp32_ms"])[1] for _ in range(args.repeats))
        b.close()
        return ms, cells / (ms * 1e-3) / 1e9

    for rpl in range(1, 17):
        R = 16 * rpl - 1  # the longest read of this row count (one padding row stays)
        regs = truncated(pool, R)
        cells = workload.cells(regs)
        # compare-and-select kernels (used for haplotypes that contain N): occupancy of the 16-lane shape
        best16 = None
        for occ in (2, 3, 4):
            ms, g = time_batch(regs, cells, occupancy=occ, segment_width=16, prior_table=2)
            report.append((rpl, "cmp", 16, occ, ms, g))
            print(f"[tune] R={R:3d} cmp W=16 occ={occ}: {ms:8.3f} ms {g:8.1f} GCUPS", flush=True)
            if best16 is None or ms < best16[1]:
                best16 = (occ, ms)
        occ16[rpl] = best16[0]
        # prior-table kernels (haplotypes without N): shape and occupancy
        best, bpt = None, {16: None, 32: None, 8: None}
        cands = [(16, o) for o in (2, 3, 4)]
        if -(-(R + 1) // 8) <= 16:
            cands += [(8, o) for o in (2, 3, 4)]
        if -(-(R + 1) // 32) <= 16:
            cands += [(32, o) for o in (2, 3, 4)]
        if -(-(R + 1) // 64) <= 13:
            cands.append((64, 2))
        for w, occ in cands:
            ms, g = time_batch(regs, cells, occupancy=occ, segment_width=w, prior_table=1)
            report.append((rpl, "table", w, occ, ms, g))
            print(f"[tune] R={R:3d} table W={w:2d} occ={occ}: {ms:8.3f} ms {g:8.1f} GCUPS", flush=True)
            if best is None or ms < best[2]:
                best = (w, occ, ms)
            if w in bpt and (bpt[w] is None or ms < bpt[w][1]):
                bpt[w] = (occ, ms)
        occpt16[rpl] = bpt[16][0]
        # the 32-lane table kernel of this read length has ceil((R+1)/32) rows per lane
        if bpt[32] is not None:
            occpt32[-(-(R + 1) // 32)] = bpt[32][0]
        if bpt[8] is not None:
            occpt8[-(-(R + 1) // 8)] = bpt[8][0]
        w, occ, _ = best
        if w == 8:
            shape16[rpl] = 8 * 256  # rows per lane follow the read length
        elif w != 16:
            shape16[rpl] = w * 256 + (-(-(R + 1) // w))
        # 8-lane shapes of 17-19 rows (prior-table kernel, 2 waves per SIMD)
        # hold reads up to 151 bases: where the class has such reads, time
        # them against the chosen shape on the class's reads that fit
        R8 = min(R, 8 * 19 - 1)
        if w == 16 and 16 * (rpl - 1) < R8 + 1 and -(-(R8 + 1) // 8) > 16:
            regs8 = truncated(pool, R8)
            cells8 = workload.cells(regs8)
            ms8, g8 = time_batch(regs8, cells8, occupancy=2, segment_width=8, prior_table=1)
            ms16, g16 = time_batch(regs8, cells8, occupancy=occpt16[rpl], segment_width=16, prior_table=1)
            rows = -(-(R8 + 1) // 8)
            report.append((rpl, "table", 8, 2, ms8, g8))
            print(f"[tune] R={R8:3d} table W= 8 ({rows} rows) occ=2: {ms8:8.3f} ms {g8:8.1f} GCUPS "
                  f"(W=16: {ms16:8.3f} ms)", flush=True)
            if ms8 < ms16:
                shape16[rpl] = 8 * 256
    occ16[0], occpt16[0], occpt32[0] = occ16[1], occpt16[1], occpt32[1]
    # 8-lane shapes were timed at even row counts; odd ones take the next one's
    for r in range(15, 0, -2):
        occpt8[r] = occpt8[r + 1]
    occpt8[0] = occpt8[1]
    rv.shutdown()
    os.makedirs(os.path.dirname(args.out), exist_ok=True)
    with open(args.out, "w") as f:
        f.write(f"# FP32 kernel shapes tuned by rovaca_amd/tune.py on {datetime.date.today()}\n")
        f.write(f"# batch: {args.regions} regions x 64 haplotypes (~350 bp) x 256 reads of 16*RPL bases\n")
        for rpl, kind, w, occ, ms, g in report:
            f.write(f"# RPL {rpl:2d} {kind:5s} W {w:2d} occ {occ}: {ms:.3f} ms, {g:.1f} GCUPS\n")
        f.write("occ16 " + " ".join(str(x) for x in occ16) + "\n")
        f.write("occpt16 " + " ".join(str(x) for x in occpt16) + "\n")
        f.write("occpt32 " + " ".join(str(x) for x in occpt32) + "\n")
        f.write("occpt8 " + " ".join(str(x) for x in occpt8) + "\n")
        f.write("shape16 " + " ".join(str(x) for x in shape16) + "\n")
        f.write("steps3 1\nfold 1\ntable 1\n")
    print(f"[tune] wrote {args.out}", flush=True)


if __name__ == "__main__":
    sys.exit(main())
