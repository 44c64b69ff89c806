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
