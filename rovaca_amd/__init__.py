"""MI355X PairHMM engine for Rovaca's HaplotypeCaller (Python bindings).

The engine itself is ``librovaca_phmm.so`` (HIP kernels for gfx950 plus a C++
host side) with the C ABI declared in ``include/rovaca_phmm.h``. This module
binds that ABI with ctypes for tests, the smoke check and the benchmark. It
never computes likelihoods itself: if the library is missing, every call
raises.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "librovaca_phmm.so")

RVC_OK = 0
RVC_E_INVALID = 1
RVC_E_LIMIT = 2
RVC_E_DEVICE = 3
RVC_E_STATE = 4
RVC_E_BUDGET = 5
RVC_E_TIMEOUT = 6
RVC_E_SHUTDOWN = 7
RVC_E_BASES = 8
PAIR_COMPUTED = 1
PAIR_FP64 = 2
PCR_NONE, PCR_HOSTILE, PCR_AGGRESSIVE, PCR_CONSERVATIVE = 0, 1, 2, 3

# Every symbol include/rovaca_phmm.h declares.
EXPORTS = [
    "rvc_phmm_init", "rvc_phmm_shutdown", "rvc_phmm_last_error", "rvc_phmm_region_sync",
    "rvc_phmm_region_sync_noalloc",
    "rvc_phmm_call_pairhmm", "rvc_phmm_call_pairhmm_submit", "rvc_phmm_call_pairhmm_wait", "rvc_phmm_batch_create", "rvc_phmm_batch_run", "rvc_phmm_batch_results",
    "rvc_phmm_batch_info", "rvc_phmm_batch_destroy", "rvc_phmm_submit", "rvc_phmm_wait", "rvc_phmm_poll",
    "rvc_phmm_server_stats", "rvc_phmm_preprocess_read", "rvc_phmm_postprocess", "rvc_phmm_tables",
    "rvc_loadgen_run", "rvc_loadgen_call_pairhmm", "rvc_loadgen_region_latency", "rvc_phmm_batch_kernels", "rvc_sw_align_batch", "rvc_sw_align_served", "rvc_loadgen_sw_align",
]

_u8p = ctypes.POINTER(ctypes.c_uint8)
_i32p = ctypes.POINTER(ctypes.c_int32)
_u32p = ctypes.POINTER(ctypes.c_uint32)
_f64p = ctypes.POINTER(ctypes.c_double)


class RvcRegion(ctypes.Structure):
    _fields_ = [("n_haps", ctypes.c_uint32), ("n_reads", ctypes.c_uint32), ("hap_off", _i32p),
                ("hap_bases", _u8p), ("read_off", _i32p), ("read_bases", _u8p), ("q", _u8p), ("i", _u8p),
                ("d", _u8p), ("c", _u8p)]


class RvcRawRegion(ctypes.Structure):
    _fields_ = [("n_haps", ctypes.c_uint32), ("n_reads", ctypes.c_uint32), ("hap_off", _i32p),
                ("hap_bases", _u8p), ("read_off", _i32p), ("read_bases", _u8p), ("read_qual", _u8p),
                ("read_mq", _u8p)]


class RvcSites(ctypes.Structure):
    _fields_ = [("n_sites", ctypes.c_uint32), ("n_alleles", _u32p), ("hap_allele", _i32p)]


class RvcConfig(ctypes.Structure):
    _fields_ = [("device", ctypes.c_int32), ("server_threads", ctypes.c_uint32),
                ("hbm_budget_bytes", ctypes.c_uint64), ("max_batch_regions", ctypes.c_uint32),
                ("host_prep", ctypes.c_uint32), ("tune_table", ctypes.c_char_p),
                ("sw_min_launch_pairs", ctypes.c_uint32), ("sw_bt_budget_mib", ctypes.c_uint32),
                ("sw_wide_shapes", ctypes.c_uint32), ("reserved", ctypes.c_uint32 * 3)]


class RvcBatchOpts(ctypes.Structure):
    _fields_ = [("persistent_waves_per_cu", ctypes.c_uint32), ("use_graph", ctypes.c_uint32),
                ("fp32_fold", ctypes.c_uint32), ("rpl_round", ctypes.c_uint32), ("occupancy", ctypes.c_uint32),
                ("segment_width", ctypes.c_uint32), ("prior_table", ctypes.c_uint32),
                ("hap_run", ctypes.c_uint32), ("untimed", ctypes.c_uint32), ("shape_fold", ctypes.c_uint32)]


class RvcBatchInfo(ctypes.Structure):
    _fields_ = [("n_pairs", ctypes.c_uint64), ("n_cells", ctypes.c_uint64), ("n_rescued", ctypes.c_uint64),
                ("rescued_cells", ctypes.c_uint64), ("device_bytes", ctypes.c_uint64),
                ("fp32_ms", ctypes.c_double), ("fp64_ms", ctypes.c_double), ("total_ms", ctypes.c_double),
                ("n_fp32_launches", ctypes.c_uint32), ("n_fp64_launches", ctypes.c_uint32)]


class RvcKernelTime(ctypes.Structure):
    _fields_ = [("name", ctypes.c_char * 48), ("precision", ctypes.c_uint32), ("width", ctypes.c_uint32),
                ("rows_per_lane", ctypes.c_uint32), ("occupancy", ctypes.c_uint32), ("tiles", ctypes.c_uint64),
                ("cells", ctypes.c_uint64), ("ms", ctypes.c_double)]


class RvcServerStats(ctypes.Structure):
    _fields_ = [("submitted", ctypes.c_uint64), ("completed", ctypes.c_uint64), ("launches", ctypes.c_uint64),
                ("cells", ctypes.c_uint64), ("rescued_pairs", ctypes.c_uint64),
                ("resident_bytes", ctypes.c_uint64), ("peak_resident_bytes", ctypes.c_uint64),
                ("busy_ms", ctypes.c_double), ("stage_ms", ctypes.c_double), ("buffer_grows", ctypes.c_uint64)]


class RvcLoadgenResult(ctypes.Structure):
    _fields_ = [("submitted", ctypes.c_uint64), ("completed", ctypes.c_uint64), ("failed", ctypes.c_uint64),
                ("cells", ctypes.c_uint64), ("wall_s", ctypes.c_double), ("lat_p50_ms", ctypes.c_double),
                ("lat_p99_ms", ctypes.c_double), ("lat_max_ms", ctypes.c_double), ("lat_mean_ms", ctypes.c_double),
                ("queue_p50_ms", ctypes.c_double), ("queue_p99_ms", ctypes.c_double),
                ("late_p99_ms", ctypes.c_double)]


class EngineError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"rvc error {code}: {msg}")
        self.code = code


_lib = None


def lib():
    """The loaded engine library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m rovaca_amd.build`")
        L = ctypes.CDLL(LIB_PATH)
        L.rvc_phmm_init.argtypes = [ctypes.POINTER(RvcConfig)]
        L.rvc_phmm_last_error.restype = ctypes.c_char_p
        L.rvc_phmm_region_sync.argtypes = [ctypes.POINTER(RvcRegion), _f64p, _u8p]
        L.rvc_phmm_region_sync_noalloc.argtypes = [ctypes.POINTER(RvcRegion), _f64p, _u8p]
        L.rvc_phmm_call_pairhmm.argtypes = [ctypes.POINTER(RvcRawRegion), ctypes.c_int32, ctypes.c_int32, _f64p,
                                            _u32p, _u32p]
        L.rvc_phmm_call_pairhmm_submit.argtypes = [ctypes.POINTER(RvcRawRegion), ctypes.c_int32, ctypes.c_int32,
                                                   ctypes.POINTER(RvcSites), ctypes.POINTER(ctypes.c_uint64)]
        L.rvc_phmm_call_pairhmm_wait.argtypes = [ctypes.c_uint64, _f64p, _u32p, _u32p, _f64p, ctypes.c_int64]
        L.rvc_phmm_batch_create.argtypes = [ctypes.POINTER(RvcRegion), ctypes.c_uint32,
                                            ctypes.POINTER(RvcBatchOpts), ctypes.POINTER(ctypes.c_void_p)]
        L.rvc_phmm_batch_run.argtypes = [ctypes.c_void_p]
        L.rvc_phmm_batch_results.argtypes = [ctypes.c_void_p, _f64p, _u8p]
        L.rvc_phmm_batch_info.argtypes = [ctypes.c_void_p, ctypes.POINTER(RvcBatchInfo)]
        L.rvc_phmm_batch_destroy.argtypes = [ctypes.c_void_p]
        L.rvc_phmm_batch_kernels.argtypes = [ctypes.c_void_p, ctypes.POINTER(RvcKernelTime), ctypes.c_uint32, _u32p]
        L.rvc_phmm_submit.argtypes = [ctypes.POINTER(RvcRegion), ctypes.POINTER(ctypes.c_uint64)]
        L.rvc_phmm_wait.argtypes = [ctypes.c_uint64, _f64p, _u8p, ctypes.c_int64, ctypes.c_void_p]
        L.rvc_phmm_poll.argtypes = [ctypes.c_uint64]
        L.rvc_phmm_server_stats.argtypes = [ctypes.POINTER(RvcServerStats)]
        L.rvc_phmm_preprocess_read.argtypes = [_u8p, _u8p, ctypes.c_int32, ctypes.c_uint8, ctypes.c_int32,
                                               ctypes.c_int32, _u8p, _u8p, _u8p]
        L.rvc_phmm_postprocess.argtypes = [ctypes.c_uint32, ctypes.c_uint32, _f64p, _i32p, _f64p, _u32p]
        L.rvc_loadgen_run.argtypes = [ctypes.POINTER(RvcRegion), ctypes.c_uint32, _u32p, _f64p, ctypes.c_uint32,
                                      ctypes.c_uint32, _f64p, ctypes.POINTER(_f64p),
                                      ctypes.POINTER(RvcLoadgenResult)]
        L.rvc_loadgen_call_pairhmm.argtypes = [ctypes.POINTER(RvcRawRegion), ctypes.c_uint32, _u32p, ctypes.c_uint32,
                                               ctypes.c_int32, ctypes.c_int32, ctypes.c_uint32, _f64p,
                                               ctypes.POINTER(RvcLoadgenResult)]
        L.rvc_loadgen_region_latency.argtypes = [ctypes.POINTER(RvcRegion), ctypes.c_uint32, ctypes.c_int32, _f64p]
        L.rvc_phmm_tables.argtypes = [ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float), _f64p,
                                      _f64p]
        _lib = L
    return _lib


def _check(rc):
    if rc != RVC_OK:
        raise EngineError(rc, lib().rvc_phmm_last_error().decode(errors="replace"))


def _ptr(a, t=_u8p):
    return a.ctypes.data_as(t)


def make_c_region(region):
    """ctypes rvc_region viewing the arrays of a region dict (keep the dict alive)."""
    cr = RvcRegion()
    cr.n_haps = int(region["n_haps"])
    cr.n_reads = int(region["n_reads"])
    cr.hap_off = _ptr(region["hap_off"], _i32p)
    cr.hap_bases = _ptr(region["hap_bases"])
    cr.read_off = _ptr(region["read_off"], _i32p)
    cr.read_bases = _ptr(region["read_bases"])
    for k in "qidc":
        setattr(cr, k, _ptr(region[k]))
    return cr


def init(device=0, server=False, hbm_budget_bytes=0, max_batch_regions=0, tune_table=None, host_prep=0,
         sw_min_launch_pairs=0, sw_bt_budget_mib=0, sw_wide_shapes=0):
    """host_prep: where call_pairhmm preprocesses reads (0 = device when the
    PCR model is on, else host; 1 = host; 2 = device). sw_min_launch_pairs:
    the aligner runs a sweep shape with fewer pairs together with the next
    larger one (0 = enough pairs to fill the device once; 1 = never).
    sw_bt_budget_mib: the aligner's backtrack bytes per device pass (0 = 12 GiB).
    sw_wide_shapes: the aligner sweeps a pass on 64 lanes per pair (0 = when
    the pass is too small to fill the device; 1 = never; 2 = always)."""
    cfg = RvcConfig()
    cfg.host_prep = int(host_prep)
    cfg.sw_min_launch_pairs = int(sw_min_launch_pairs)
    cfg.sw_bt_budget_mib = int(sw_bt_budget_mib)
    cfg.sw_wide_shapes = int(sw_wide_shapes)
    cfg.tune_table = tune_table.encode() if tune_table else None
    cfg.device = device
    cfg.server_threads = 1 if server else 0
    cfg.hbm_budget_bytes = hbm_budget_bytes
    cfg.max_batch_regions = max_batch_regions
    _check(lib().rvc_phmm_init(ctypes.byref(cfg)))


def shutdown():
    _check(lib().rvc_phmm_shutdown())


def region_sync(region):
    """Read-major [n_reads, n_haps] log10 likelihoods and per-pair flags of one region."""
    nr, nh = int(region["n_reads"]), int(region["n_haps"])
    out = np.empty(nr * nh, np.float64)
    flags = np.zeros(nr * nh, np.uint8)
    cr = make_c_region(region)
    _check(lib().rvc_phmm_region_sync(ctypes.byref(cr), _ptr(out, _f64p), _ptr(flags)))
    return out.reshape(nr, nh), flags.reshape(nr, nh)


def region_sync_noalloc(region):
    """region_sync through the entry point that allocates no device memory
    (rvc_phmm_region_sync_noalloc): the same matrix and flags."""
    nr, nh = int(region["n_reads"]), int(region["n_haps"])
    out = np.empty(nr * nh, np.float64)
    flags = np.zeros(nr * nh, np.uint8)
    cr = make_c_region(region)
    _check(lib().rvc_phmm_region_sync_noalloc(ctypes.byref(cr), _ptr(out, _f64p), _ptr(flags)))
    return out.reshape(nr, nh), flags.reshape(nr, nh)


def _raw_arrays(haps, reads, quals, mqs):
    hap_off = np.zeros(len(haps) + 1, np.int32)
    hap_off[1:] = np.cumsum([len(h) for h in haps])
    read_off = np.zeros(len(reads) + 1, np.int32)
    read_off[1:] = np.cumsum([len(r) for r in reads])
    hb = np.frombuffer(b"".join(bytes(h) for h in haps) or b"\0", np.uint8).copy()
    rb = np.frombuffer(b"".join(bytes(r) for r in reads) or b"\0", np.uint8).copy()
    rq = np.concatenate([np.asarray(q, np.uint8) for q in quals] + [np.zeros(1, np.uint8)])
    mq = np.asarray(mqs, np.uint8)
    raw = RvcRawRegion()
    raw.n_haps, raw.n_reads = len(haps), len(reads)
    raw.hap_off, raw.hap_bases = _ptr(hap_off, _i32p), _ptr(hb)
    raw.read_off, raw.read_bases = _ptr(read_off, _i32p), _ptr(rb)
    raw.read_qual, raw.read_mq = _ptr(rq), _ptr(mq)
    raw._keep = (hap_off, read_off, hb, rb, rq, mq)
    return raw


def call_pairhmm(haps, reads, quals, mqs, min_quality_threshold=0, pcr=PCR_NONE):
    """The region-level call of the reference (raw reads in, [H][N] matrix and kept reads out)."""
    raw = _raw_arrays(haps, reads, quals, mqs)
    H, N = len(haps), len(reads)
    out = np.empty(max(H * N, 1), np.float64)
    kept = np.zeros(max(N, 1), np.uint32)
    nk = ctypes.c_uint32(0)
    rc = lib().rvc_phmm_call_pairhmm(ctypes.byref(raw), min_quality_threshold, pcr, _ptr(out, _f64p),
                                     _ptr(kept, _u32p), ctypes.byref(nk))
    return rc, out[:H * N].reshape(H, N), kept[:nk.value].copy()


def call_pairhmm_submit(haps, reads, quals, mqs, min_quality_threshold=0, pcr=PCR_NONE, sites=None):
    """First half of the region-level call: returns (rc, ticket) at once.
    sites = (n_alleles per site, hap_allele [n_sites][n_haps] with -1 for none)
    asks the device to marginalise the region's matrix per site."""
    raw = _raw_arrays(haps, reads, quals, mqs)
    t = ctypes.c_uint64(0)
    sp = None
    if sites is not None:
        na = np.ascontiguousarray(sites[0], np.uint32)
        ha = np.ascontiguousarray(sites[1], np.int32).reshape(-1)
        st = RvcSites(len(na), _ptr(na, _u32p), _ptr(ha, _i32p))
        st._keep = (na, ha)
        sp = ctypes.byref(st)
    rc = lib().rvc_phmm_call_pairhmm_submit(ctypes.byref(raw), min_quality_threshold, pcr, sp, ctypes.byref(t))
    return rc, t.value


def call_pairhmm_wait(ticket, n_haps, n_reads, timeout_ms=-1, n_alleles=None):
    """Second half: (rc, [H][N] matrix, kept reads), as call_pairhmm returns
    them; with n_alleles (the sites given at submit) also the per-site
    [n_alleles[s]][N] marginals (kept reads' columns first, the rest NaN)."""
    out = np.empty(max(n_haps * n_reads, 1), np.float64)
    kept = np.zeros(max(n_reads, 1), np.uint32)
    nk = ctypes.c_uint32(0)
    marg = None
    if n_alleles is not None:
        marg = np.empty(max(int(np.sum(n_alleles)) * n_reads, 1), np.float64)
    rc = lib().rvc_phmm_call_pairhmm_wait(ticket, _ptr(out, _f64p), _ptr(kept, _u32p), ctypes.byref(nk),
                                          _ptr(marg, _f64p) if marg is not None else None, timeout_ms)
    res = (rc, out[:n_haps * n_reads].reshape(n_haps, n_reads), kept[:nk.value].copy())
    if n_alleles is None:
        return res
    per, o = [], 0
    for A in n_alleles:
        per.append(marg[o:o + int(A) * n_reads].reshape(int(A), n_reads))
        o += int(A) * n_reads
    return res + (per,)


def poll(ticket):
    """1 if the ticket is done, 0 if not yet, negative for an unknown ticket."""
    return lib().rvc_phmm_poll(ticket)


class Batch:
    """Many regions packed, uploaded and launched together."""

    def __init__(self, regions, use_graph=False, fp32_fold=0, rpl_round=0, waves_per_cu=0, occupancy=0,
                 segment_width=0, prior_table=0, hap_run=0, untimed=False, shape_fold=True):
        self._regions = regions
        arr = (RvcRegion * len(regions))()
        for x, r in enumerate(regions):
            arr[x] = make_c_region(r)
        self._arr = arr
        opts = RvcBatchOpts()
        opts.use_graph = 1 if use_graph else 0
        opts.fp32_fold = fp32_fold
        opts.rpl_round = rpl_round
        opts.persistent_waves_per_cu = waves_per_cu
        opts.occupancy = occupancy
        opts.segment_width = segment_width
        opts.prior_table = prior_table
        opts.hap_run = hap_run
        opts.untimed = 1 if untimed else 0
        opts.shape_fold = 0 if shape_fold else 1
        h = ctypes.c_void_p()
        _check(lib().rvc_phmm_batch_create(arr, len(regions), ctypes.byref(opts), ctypes.byref(h)))
        self.h = h
        self.offsets = np.zeros(len(regions) + 1, np.int64)
        self.offsets[1:] = np.cumsum([int(r["n_reads"]) * int(r["n_haps"]) for r in regions])

    def run(self):
        _check(lib().rvc_phmm_batch_run(self.h))

    def info(self):
        inf = RvcBatchInfo()
        _check(lib().rvc_phmm_batch_info(self.h, ctypes.byref(inf)))
        return {k: getattr(inf, k) for k, _ in RvcBatchInfo._fields_}

    def kernels(self):
        """Per-launch device times of the last non-graph run."""
        n = ctypes.c_uint32(0)
        arr = (RvcKernelTime * 64)()
        _check(lib().rvc_phmm_batch_kernels(self.h, arr, 64, ctypes.byref(n)))
        out = []
        for x in range(min(n.value, 64)):
            k = arr[x]
            out.append({"name": k.name.decode(), "precision": k.precision, "width": k.width,
                        "rows_per_lane": k.rows_per_lane, "occupancy": k.occupancy, "tiles": k.tiles,
                        "cells": k.cells, "ms": k.ms})
        return out

    def results(self):
        n = int(self.offsets[-1])
        out = np.empty(n, np.float64)
        flags = np.zeros(n, np.uint8)
        _check(lib().rvc_phmm_batch_results(self.h, _ptr(out, _f64p), _ptr(flags)))
        return out, flags

    def region_result(self, out, r):
        reg = self._regions[r]
        return out[self.offsets[r]:self.offsets[r + 1]].reshape(int(reg["n_reads"]), int(reg["n_haps"]))

    def close(self):
        if self.h:
            lib().rvc_phmm_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def submit(region):
    cr = make_c_region(region)
    t = ctypes.c_uint64()
    _check(lib().rvc_phmm_submit(ctypes.byref(cr), ctypes.byref(t)))
    return t.value


def wait(ticket, region, timeout_ms=-1):
    nr, nh = int(region["n_reads"]), int(region["n_haps"])
    out = np.empty(nr * nh, np.float64)
    flags = np.zeros(nr * nh, np.uint8)
    _check(lib().rvc_phmm_wait(ticket, _ptr(out, _f64p), _ptr(flags), timeout_ms, None))
    return out.reshape(nr, nh), flags.reshape(nr, nh)


def preprocess_read(bases, qual, mq, min_quality_threshold=0, pcr=PCR_NONE):
    """Host-side read preprocessing: returns (q, gop, gcp) uint8 arrays."""
    b = np.ascontiguousarray(np.frombuffer(bytes(bases), np.uint8) if isinstance(bases, (bytes, bytearray))
                             else bases, dtype=np.uint8)
    ql = np.ascontiguousarray(qual, dtype=np.uint8)
    n = len(b)
    q, g, c = (np.zeros(max(n, 1), np.uint8) for _ in range(3))
    _check(lib().rvc_phmm_preprocess_read(_ptr(b), _ptr(ql), n, mq, min_quality_threshold, pcr, _ptr(q), _ptr(g),
                                          _ptr(c)))
    return q[:n], g[:n], c[:n]


def postprocess(read_major, read_len):
    """Host-side normalize / filter / transpose: returns ([H][N] matrix, kept read indices)."""
    rm = np.ascontiguousarray(read_major, dtype=np.float64).copy()
    n, h = rm.shape
    rl = np.ascontiguousarray(read_len, dtype=np.int32)
    out = np.empty((h, n), np.float64)
    kept = np.zeros(n, np.uint32)
    nk = lib().rvc_phmm_postprocess(n, h, _ptr(rm, _f64p), _ptr(rl, _i32p), _ptr(out, _f64p), _ptr(kept, _u32p))
    if nk < 0:
        _check(-nk)
    return out, kept[:nk].copy()


def tables():
    """(ph2pr_f32[128], mm_f32[8256], ph2pr_f64[128], mm_f64[8256]) as the engine builds them."""
    pf = np.zeros(128, np.float32)
    mf = np.zeros(128 * 129 // 2, np.float32)
    pd = np.zeros(128, np.float64)
    md = np.zeros(128 * 129 // 2, np.float64)
    fp = ctypes.POINTER(ctypes.c_float)
    _check(lib().rvc_phmm_tables(_ptr(pf, fp), _ptr(mf, fp), _ptr(pd, _f64p), _ptr(md, _f64p)))
    return pf, mf, pd, md


def loadgen_run(regions, arrival_s, order=None, n_producers=32, outputs=False):
    """Submits regions at the given arrival times from n_producers threads (C++)
    and waits for all of them. Returns (summary dict, per-request latency ms),
    and with outputs=True also each arrival's read-major matrix."""
    arr = (RvcRegion * len(regions))()
    for x, r in enumerate(regions):
        arr[x] = make_c_region(r)
    arrival = np.ascontiguousarray(arrival_s, np.float64)
    n = len(arrival)
    o = None if order is None else np.ascontiguousarray(order, np.uint32)
    lat = np.zeros(n, np.float64)
    res = RvcLoadgenResult()
    outs, optr = None, None
    if outputs:
        idx = o if o is not None else np.arange(n) % len(regions)
        outs = [np.empty((int(regions[i]["n_reads"]), int(regions[i]["n_haps"])), np.float64) for i in idx]
        optr = (_f64p * n)(*[_ptr(x, _f64p) for x in outs])
    rc = lib().rvc_loadgen_run(arr, len(regions), _ptr(o, _u32p) if o is not None else None, _ptr(arrival, _f64p),
                               n, n_producers, _ptr(lat, _f64p), optr, ctypes.byref(res))
    _check(rc)
    summary = {k: getattr(res, k) for k, _ in RvcLoadgenResult._fields_}
    return (summary, lat, outs) if outputs else (summary, lat)


def region_latency(region, n_calls, noalloc=False):
    """Per-call latency (ms) of n_calls one-region calls made from C++
    (rvc_loadgen_region_latency), without Python in the timed loop."""
    cr = make_c_region(region)
    lat = np.zeros(n_calls, np.float64)
    _check(lib().rvc_loadgen_region_latency(ctypes.byref(cr), n_calls, 1 if noalloc else 0, _ptr(lat, _f64p)))
    return lat


def raw_region(region, mqs):
    """The rvc_raw_region view of a region dict: its q array as the raw base
    qualities and the given per-read mapping qualities (arrays kept alive on
    the returned object)."""
    raw = RvcRawRegion()
    keep = [np.ascontiguousarray(region["hap_off"], np.int32), np.ascontiguousarray(region["hap_bases"], np.uint8),
            np.ascontiguousarray(region["read_off"], np.int32), np.ascontiguousarray(region["read_bases"], np.uint8),
            np.ascontiguousarray(region["q"], np.uint8), np.ascontiguousarray(mqs, np.uint8)]
    raw.n_haps, raw.n_reads = int(region["n_haps"]), int(region["n_reads"])
    raw.hap_off, raw.hap_bases = _ptr(keep[0], _i32p), _ptr(keep[1])
    raw.read_off, raw.read_bases = _ptr(keep[2], _i32p), _ptr(keep[3])
    raw.read_qual, raw.read_mq = _ptr(keep[4]), _ptr(keep[5])
    raw._keep = keep
    return raw


def loadgen_call_pairhmm(raw_regions, n_calls, order=None, min_quality_threshold=0, pcr=PCR_NONE, n_workers=32):
    """n_calls region calls (rvc_phmm_call_pairhmm) from n_workers threads
    (C++), one region at a time per thread. Returns (summary, latency ms)."""
    arr = (RvcRawRegion * len(raw_regions))()
    for x, r in enumerate(raw_regions):
        arr[x] = r
    o = None if order is None else np.ascontiguousarray(order, np.uint32)
    lat = np.zeros(n_calls, np.float64)
    res = RvcLoadgenResult()
    _check(lib().rvc_loadgen_call_pairhmm(arr, len(raw_regions), _ptr(o, _u32p) if o is not None else None, n_calls,
                                          min_quality_threshold, pcr, n_workers, _ptr(lat, _f64p),
                                          ctypes.byref(res)))
    return {k: getattr(res, k) for k, _ in RvcLoadgenResult._fields_}, lat


def server_stats():
    st = RvcServerStats()
    _check(lib().rvc_phmm_server_stats(ctypes.byref(st)))
    return {k: getattr(st, k) for k, _ in RvcServerStats._fields_}


# ---- batched Smith-Waterman (include/rovaca_sw.h) ---------------------------
class RvcSwStats(ctypes.Structure):
    _fields_ = [("fill_ms", ctypes.c_double), ("trace_ms", ctypes.c_double), ("total_ms", ctypes.c_double),
                ("cells", ctypes.c_uint64), ("launches", ctypes.c_uint32), ("chunks", ctypes.c_uint32)]


_SW_OPS = "MIDNSHP=X"


class SwBatch:
    """A batch of reads to align against their haplotypes on the device,
    packed once, with output arrays kept for repeated runs (a caller that
    aligns batch after batch reuses its result buffers the same way)."""

    def __init__(self, haps, reads, read_hap, assemble=False, max_cigar=None):
        self.hap_off = np.zeros(len(haps) + 1, np.int32)
        self.hap_off[1:] = np.cumsum([len(h) for h in haps])
        self.read_off = np.zeros(len(reads) + 1, np.int32)
        self.read_off[1:] = np.cumsum([len(r) for r in reads])
        self.hb = np.frombuffer(b"".join(bytes(h) for h in haps) + b"\0", np.uint8)
        self.rb = np.frombuffer(b"".join(bytes(r) for r in reads) + b"\0", np.uint8)
        self.rh = np.ascontiguousarray(read_hap, np.uint32)
        self.n_haps, self.n = len(haps), len(reads)
        self.assemble = 1 if assemble else 0
        # 2 * longest read + 2 elements always suffice
        self.max_cigar = int(max_cigar) if max_cigar else 2 * max([len(r) for r in reads] + [1]) + 2
        self.cig = np.zeros(self.n * self.max_cigar, np.uint32)
        self.cnt = np.zeros(self.n, np.int32)
        self.off = np.zeros(self.n, np.int32)
        self.sc = np.zeros(self.n, np.int32)

    def run(self, served=False):
        """Aligns every read; returns the call's statistics as a dict (the
        batched call), or None (served: rvc_sw_align_served, merged with
        other threads' concurrent calls)."""
        args = [self.hb.ctypes.data_as(_u8p), self.hap_off.ctypes.data_as(_i32p), self.n_haps,
                self.rb.ctypes.data_as(_u8p), self.read_off.ctypes.data_as(_i32p), self.n,
                self.rh.ctypes.data_as(_u32p), self.assemble, self.max_cigar, self.cig.ctypes.data_as(_u32p),
                self.cnt.ctypes.data_as(_i32p), self.off.ctypes.data_as(_i32p), self.sc.ctypes.data_as(_i32p)]
        types = [_u8p, _i32p, ctypes.c_uint32, _u8p, _i32p, ctypes.c_uint32, _u32p, ctypes.c_int32,
                 ctypes.c_uint32, _u32p, _i32p, _i32p, _i32p]
        if served:
            fn = lib().rvc_sw_align_served
            fn.argtypes = types
            _check(fn(*args))
            return None
        fn = lib().rvc_sw_align_batch
        fn.argtypes = types + [ctypes.POINTER(RvcSwStats)]
        st = RvcSwStats()
        _check(fn(*args, ctypes.byref(st)))
        return {f: getattr(st, f) for f, _ in RvcSwStats._fields_}

    def results(self):
        """[(cigar [(op, length)], offset, score)] per read of the last run,
        (None, 1, None) where the aligner refuses the pair."""
        out = []
        for r in range(self.n):
            k = int(self.cnt[r])
            if k < 0:
                out.append((None, int(self.off[r]), None))
                continue
            c = self.cig[r * self.max_cigar:r * self.max_cigar + k]
            out.append(([(_SW_OPS[int(x) & 15], int(x) >> 4) for x in c], int(self.off[r]), int(self.sc[r])))
        return out


class RvcSwRegion(ctypes.Structure):
    _fields_ = [("hap_bases", _u8p), ("hap_off", _i32p), ("n_haps", ctypes.c_uint32), ("read_bases", _u8p),
                ("read_off", _i32p), ("n_reads", ctypes.c_uint32), ("read_hap", _u32p), ("max_cigar", ctypes.c_uint32),
                ("cigar", _u32p), ("count", _i32p), ("offset", _i32p), ("score", _i32p)]


def loadgen_sw_align(batches, served=True, n_workers=32, assemble=False):
    """One aligner call per SwBatch (a region's reads) from n_workers C++
    threads, served (rvc_sw_align_served) or batched per call. Results land
    in each SwBatch. Returns (summary, latency ms per call)."""
    arr = (RvcSwRegion * len(batches))()
    for x, b in enumerate(batches):
        arr[x] = RvcSwRegion(b.hb.ctypes.data_as(_u8p), b.hap_off.ctypes.data_as(_i32p), b.n_haps,
                             b.rb.ctypes.data_as(_u8p), b.read_off.ctypes.data_as(_i32p), b.n,
                             b.rh.ctypes.data_as(_u32p), b.max_cigar, b.cig.ctypes.data_as(_u32p),
                             b.cnt.ctypes.data_as(_i32p), b.off.ctypes.data_as(_i32p), b.sc.ctypes.data_as(_i32p))
    lat = np.zeros(len(batches), np.float64)
    res = RvcLoadgenResult()
    _check(lib().rvc_loadgen_sw_align(arr, len(batches), 1 if assemble else 0, 1 if served else 0, n_workers,
                                      _ptr(lat, _f64p), ctypes.byref(res)))
    return {k: getattr(res, k) for k, _ in RvcLoadgenResult._fields_}, lat


def sw_align_batch(haps, reads, read_hap, assemble=False, with_stats=False, served=False):
    """Aligns reads[r] against haps[read_hap[r]] on the device (the aligner of
    Rovaca's read realignment). Returns [(cigar [(op, length)], offset, score)]
    per read, (None, 1, None) where the aligner refuses the pair; with
    with_stats, also the call's statistics as a dict (None when served)."""
    b = SwBatch(haps, reads, read_hap, assemble)
    st = b.run(served)
    out = b.results()
    return (out, st) if with_stats else out
