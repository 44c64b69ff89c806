"""Builds librovaca_phmm.so (HIP for gfx950 + host C++) in this folder.

Run as ``python -m rovaca_amd.build`` or through ``__graft_entry__.build()``.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB = os.path.join(HERE, "librovaca_phmm.so")
SOURCES = ["phmm_kernels.hip", "phmm_engine.cpp", "phmm_tables.cpp", "phmm_capi.cpp", "phmm_adapter.cpp",
           "phmm_server.cpp", "phmm_loadgen.cpp", "phmm_noalloc.cpp", "sw_kernels.hip", "sw_capi.cpp",
           "sw_server.cpp"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
# -packed-fp32-ops: keep the compiler from forming v_pk_{add,mul,fma}_f32.
FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-value", "-Wno-unused-result",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def build(verbose=False):
    csrc = os.path.join(HERE, "csrc")
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    objs = []
    procs = []
    for src in SOURCES:
        obj = os.path.join(objdir, os.path.splitext(src)[0] + ".o")
        lang = [] if src.endswith(".hip") else ["-x", "hip"]
        cmd = [HIPCC] + FLAGS + lang + ["-c", os.path.join(csrc, src), "-o", obj]
        procs.append((src, cmd, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)))
        objs.append(obj)
    failed = False
    for src, cmd, p in procs:
        out, _ = p.communicate()
        text = "\n".join(l for l in out.decode(errors="replace").splitlines() if "packed-fp32-ops" not in l)
        if p.returncode != 0:
            failed = True
            sys.stderr.write(f"[build] {src} failed:\n{text}\n")
        elif verbose and text.strip():
            sys.stderr.write(text + "\n")
    if failed:
        raise RuntimeError("HIP build failed")
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs + ["-lpthread"]
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


def build_oracle():
    """Builds oracle/liboracle.so and, where the Rovaca tree exists, oracle/_ref,
    from scratch every time (make -B)."""
    odir = os.path.join(ROOT, "oracle")
    subprocess.run(["make", "-B", "-s", "-j4", "-C", odir, "all"], check=True)
    ref = os.environ.get("ROVACA_REF", "/root/reference/src/haplotypecaller/pairhmm")
    if os.path.isdir(ref):
        subprocess.run(["make", "-B", "-s", "-j4", "-C", odir, "ref", f"REF={ref}"], check=True)


if __name__ == "__main__":
    print(build(verbose=True))
    build_oracle()
