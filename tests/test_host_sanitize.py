"""The library's host code under AddressSanitizer and UBSan, on the CPU.

tests/sanitize/host_pack_harness.cpp is built together with the host sources
of the PairHMM library (region checks, packing, tile runs, staging, merging
for a serving launch, result unpermuting, read preprocessing and
post-processing) with the sanitizers on the host side only, and run over
random and edge-case regions: empty haplotypes and reads, N bases, refused
bytes, reads up to 768 bases, up to 300 haplotypes, low gap qualities (the
plain form), device-preprocessing records, several regions per pack and
merged serving launches. Every input array is allocated at exactly its size
(the C ABI promises nothing past it). Any sanitizer report fails the test.
"""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
CSRC = os.path.join(ROOT, "rovaca_amd", "csrc")


def test_host_code_under_sanitizers(tmp_path):
    if not os.path.exists(HIPCC):
        pytest.skip("hipcc not found")
    exe = str(tmp_path / "host_pack_harness")
    srcs = [os.path.join(ROOT, "tests", "sanitize", "host_pack_harness.cpp")] + [
        os.path.join(CSRC, f) for f in ("phmm_engine.cpp", "phmm_tables.cpp", "phmm_adapter.cpp")]
    san = ["-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined",
           "-Xarch_host", "-fno-sanitize-recover=all"]
    cmd = [HIPCC, "--offload-arch=gfx950", "-std=c++17", "-O1", "-g", "-fno-omit-frame-pointer"] + san + [
        "-I" + CSRC, "-I" + os.path.join(ROOT, "include")] + srcs + [
        "-Wl,--wrap=hipHostMalloc", "-Wl,--wrap=hipHostFree", "-o", exe]
    b = subprocess.run(cmd, capture_output=True, text=True, timeout=600)
    assert b.returncode == 0, b.stderr[-4000:]
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:halt_on_error=1", UBSAN_OPTIONS="print_stacktrace=1")
    for seed in (1, 2):
        r = subprocess.run([exe, "200", str(seed)], capture_output=True, text=True, timeout=600, env=env)
        out = r.stdout + r.stderr
        assert r.returncode == 0 and "ERROR: AddressSanitizer" not in out and "runtime error" not in out, out[-4000:]
        assert "packs" in r.stdout
