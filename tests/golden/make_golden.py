"""Writes the committed golden fixtures of the PairHMM tests.

Inputs come from rovaca_amd.workload (seeded) and are stored with the
outputs, so the fixtures do not depend on the generator staying unchanged.
Outputs per pair:
  ref_intra  Rovaca's own intra-pair kernel (ph/roDV_intra/pairhmm_api.cpp:14-23,
             FP32 then FP64 below 1e-28), built from the reference sources into
             oracle/_ref by oracle/Makefile
  fp64       the oracle's FP64 evaluation (C = 2^1020, N wildcard)
For the 8 real regions of Rovaca's own schedule test data
(pairhmm_debug.txt, ph/resource/) it writes pairhmm_debug_ref.npz: per region
the ref_intra and fp64 matrices, flattened region after region (read-major),
as the reference test computes its expected values
(ph/test/pairhmm_unitest_from_assemble.cpp:69-107).
Run from the repository root: python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))
import oracle  # noqa: E402
from conftest import load_debug_regions  # noqa: E402
from rovaca_amd import workload  # noqa: E402

KEYS = ("hap_off", "hap_bases", "read_off", "read_bases", "q", "i", "d", "c")


def outputs(reg):
    H, N = reg["n_haps"], reg["n_reads"]
    ref = np.full((N, H), np.nan)
    f64 = np.full((N, H), np.nan)
    for r in range(N):
        o0, o1 = reg["read_off"][r], reg["read_off"][r + 1]
        args = [reg[k][o0:o1] for k in "qidc"]
        rd = reg["read_bases"][o0:o1]
        for h in range(H):
            hb = reg["hap_bases"][reg["hap_off"][h]:reg["hap_off"][h + 1]]
            if len(hb) == 0 or len(rd) == 0:
                continue
            ref[r, h] = oracle.ref_intra_log10(hb, rd, *args)
            f64[r, h] = oracle.pair_log10_f64(hb, rd, *args)
    return ref, f64


def save(name, reg):
    ref, f64 = outputs(reg)
    np.savez_compressed(os.path.join(ROOT, "tests", "golden", name), ref_intra=ref, fp64=f64,
                        **{k: reg[k] for k in KEYS})


def main():
    if oracle.ref() is None:
        sys.exit("oracle/_ref is not built: run `make -C oracle ref`")
    save("single_region_seed0.npz", workload.single_region(0))
    small = [r for r in workload.batched_regions(60, seed=101) if r["n_haps"] * r["n_reads"] <= 1500][:4]
    for x, r in enumerate(small):
        save(f"batched_seed101_{x}.npz", r)
    stress = [r for r in workload.stress_regions(60, seed=202) if r["n_haps"] * r["n_reads"] <= 1500][:3]
    for x, r in enumerate(stress):
        save(f"stress_seed202_{x}.npz", r)
    refs, f64s = [], []
    for reg in load_debug_regions():
        ref, f64 = outputs(reg)
        refs.append(ref.ravel())
        f64s.append(f64.ravel())
    np.savez_compressed(os.path.join(ROOT, "tests", "golden", "pairhmm_debug_ref.npz"),
                        ref_intra=np.concatenate(refs), fp64=np.concatenate(f64s))


if __name__ == "__main__":
    main()
