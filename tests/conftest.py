import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")


GOLDEN = os.path.join(ROOT, "tests", "golden")


def load_kats():
    """The 104 known-answer tests Rovaca's unit tests read (ph/resource/pairhmm-testdata.txt;
    quality characters are Phred+33, ph/test/pairhmm_unittest.cpp:86-92)."""
    import numpy as np
    out = []
    with open(os.path.join(GOLDEN, "pairhmm-testdata.txt")) as f:
        for ln in f:
            if not ln.strip() or ln.startswith("#"):
                continue
            hap, rd, q, i, d, c, exp = ln.split()
            cv = lambda s: (np.frombuffer(s.encode(), np.uint8) - 33).astype(np.uint8)
            out.append((np.frombuffer(hap.encode(), np.uint8).copy(), np.frombuffer(rd.encode(), np.uint8).copy(),
                        cv(q), cv(i), cv(d), cv(c), float(exp)))
    return out


@pytest.fixture(scope="session")
def kats():
    return load_kats()


def load_debug_regions():
    """The 8 real active regions (chrY) of Rovaca's schedule test data
    (ph/resource/pairhmm_debug.txt, read by ph/test/pairhmm_unitest_from_assemble.cpp:19-170):
    region dicts in the rvc_region layout, qualities as given (Phred, not +33).
    A read whose quality arrays do not match its length is skipped by the
    reference test (expected 0.0, :84-90); it is marked in 'valid'."""
    import numpy as np
    regions, cur, read = [], None, None
    with open(os.path.join(GOLDEN, "pairhmm_debug.txt")) as f:
        for ln in f:
            s = ln.strip()
            if s.startswith("=== Region:"):
                cur = {"name": s[len("=== Region:"):].strip(" ="), "haps": [], "reads": []}
                regions.append(cur)
            elif s[:1] == "H" and ":" in s and s.split(":")[0][1:].isdigit():
                cur["haps"].append(s.split(":", 1)[1].strip().encode())
            elif s[:1] == "R" and ":" in s and s.split(":")[0][1:].isdigit():
                read = {"seq": s.split(":", 1)[1].strip().encode()}
                cur["reads"].append(read)
            elif s.startswith(("read-qual:", "read-ins-qual:", "read-del-qual:", "gcp:")):
                k, v = s.split(":", 1)
                read[k] = [int(x) for x in v.split()]
    out = []
    for g in regions:
        valid = [all(len(r.get(k, [])) == len(r["seq"]) for k in ("read-qual", "read-ins-qual", "read-del-qual",
                                                                   "gcp")) for r in g["reads"]]
        cols = {"q": "read-qual", "i": "read-ins-qual", "d": "read-del-qual", "c": "gcp"}
        reg = {"name": g["name"], "n_haps": len(g["haps"]), "n_reads": len(g["reads"]), "valid": np.array(valid)}
        reg["hap_off"] = np.zeros(len(g["haps"]) + 1, np.int32)
        reg["hap_off"][1:] = np.cumsum([len(h) for h in g["haps"]])
        reg["hap_bases"] = np.frombuffer(b"".join(g["haps"]) + b"\0", np.uint8).copy()
        reg["read_off"] = np.zeros(len(g["reads"]) + 1, np.int32)
        reg["read_off"][1:] = np.cumsum([len(r["seq"]) for r in g["reads"]])
        reg["read_bases"] = np.frombuffer(b"".join(r["seq"] for r in g["reads"]) + b"\0", np.uint8).copy()
        for k, name in cols.items():
            vals = []
            for r, ok in zip(g["reads"], valid):
                v = r.get(name, []) if ok else [30] * len(r["seq"])
                vals.append(np.array(v[:len(r["seq"])], np.uint8))
            reg[k] = np.concatenate(vals + [np.zeros(1, np.uint8)])
        out.append(reg)
    return out
