"""Benchmark of the batched transform path on one MI355X.

Default workload (the metric's configuration): 8192 independent 1D complex
DFTs of uneven power-of-two lengths N_i = 2^k_i, k_i = 6 + floor(11 u_i)
(log-uniform over 2^6..2^16, numpy PCG64 seed 3), data iid uniform[-1,1) in
re/im, FP32 butterflies with twiddles generated in FP64, one graph launch per
step (minfft_mkaux_batch / minfft_batch_run).  Inputs and outputs stay in
HBM; a step = one forward transform of every entry.

Printed: one JSON line with the throughput in GFLOP/s (5 N log2 N per entry),
the whole-step HBM roofline figure, the FP64 GPU path on the same workload,
the error vs the FP64 oracle on a sample of entries and the reference minfft
(compiled from its sources, single precision, one host core) on a bounded
sample of the same workload.
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

SEED = 3
COUNT = 8192
HBM_PEAK_GBS = 8000.0
METRIC = "batched C2C GFLOP/s + HBM GB/s, FP32/split-FP16 vs FP64 plain, rel-L2<=1e-6"


def workload(count=COUNT, seed=SEED):
    rng = np.random.default_rng(seed)
    ks = 6 + np.floor(11 * rng.random(count)).astype(np.int64)
    return (1 << ks).astype(np.int64)


def log(msg):
    print(msg, file=sys.stderr, flush=True)


# Untimed replays before the warm-up steps of every timed workload, for at
# least SETTLE_MS of wall time.  After the GPU has been idle for ~10 ms or
# more, the first 25-50 replays of a memory-bound workload run 5-10 % slower
# (default batch: ~455 us per step for the first 25 replays after 1 s idle,
# 415-417 us once settled).  A compute-bound warm-up (150 ms of matmuls) does
# not remove it and 50 ms of device copies removes half of it; the workload's
# own replays do: it is device state (memory-side clocks, translation and
# cache contents) that the workload warms itself.  The steady state is what a
# caller running the transforms in a loop sees; the timed steps themselves
# are unchanged.  --settle-ms 0 turns this off.
SETTLE_MS = 60.0
SETTLED = []


def settle(torch, fn):
    """Replay fn untimed until SETTLE_MS of wall time has passed; returns
    (replays, ms)."""
    if SETTLE_MS <= 0:
        return (0, 0.0)
    n = 0
    t0 = time.perf_counter()
    while (time.perf_counter() - t0) * 1e3 < SETTLE_MS:
        for _ in range(8):
            fn()
        n += 8
        torch.cuda.synchronize()
    return (n, round((time.perf_counter() - t0) * 1e3, 1))


def gpu_batch(sizes, prec, steps, warmup, torch, mf, seed=SEED):
    """Time `steps` graph launches of the batch; returns (ms/step, in, out, offs)."""
    cdt = torch.complex64 if prec == mf.PREC_FP32 else torch.complex128
    rdt = torch.float32 if prec == mf.PREC_FP32 else torch.float64
    esz = 8 if prec == mf.PREC_FP32 else 16
    offs = np.concatenate([[0], np.cumsum(sizes)[:-1]])
    total = int(sizes.sum())
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    xin = (torch.rand(total, 2, generator=g, device="cuda", dtype=torch.float32) * 2 - 1).to(rdt)
    xin = torch.view_as_complex(xin.contiguous()).to(cdt)
    xout = torch.empty(total, dtype=cdt, device="cuda")
    entries = [("dft", (int(n),), int(o) * esz, int(o) * esz) for n, o in zip(sizes, offs)]
    plan = mf.Batch(entries, prec)
    # a stream of our own (default priority): its handle is what the graph
    # launches on and what the events below are recorded on
    st = torch.cuda.Stream()
    sh = st.cuda_stream
    assert sh, "need a non-null stream handle"
    torch.cuda.synchronize()
    plan.run(xin.data_ptr(), xout.data_ptr(), sh)  # first run captures the graph
    SETTLED.append(settle(torch, lambda: plan.run(xin.data_ptr(), xout.data_ptr(), sh)))
    for _ in range(warmup):
        plan.run(xin.data_ptr(), xout.data_ptr(), sh)
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record(st)
    for _ in range(steps):
        plan.run(xin.data_ptr(), xout.data_ptr(), sh)
    e1.record(st)
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / steps
    return ms, xin, xout, offs, plan


def dominant_group(sizes, xin, xout, offs, prec, steps, torch, mf):
    """Time the largest-bytes length group alone (one kernel launch per step
    for the long lengths), with HIP events on the stream it runs on."""
    esz = 8 if prec == mf.PREC_FP32 else 16
    vals, counts = np.unique(sizes, return_counts=True)
    n = int(vals[np.argmax(vals * counts)])
    idx = np.nonzero(sizes == n)[0]
    plan = mf.Batch([("dft", (n,), int(offs[i]) * esz, int(offs[i]) * esz) for i in idx], prec)
    st = torch.cuda.Stream()
    sh = st.cuda_stream
    plan.run(xin.data_ptr(), xout.data_ptr(), sh)
    SETTLED.append(settle(torch, lambda: plan.run(xin.data_ptr(), xout.data_ptr(), sh)))
    for _ in range(2):
        plan.run(xin.data_ptr(), xout.data_ptr(), sh)
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record(st)
    for _ in range(steps):
        plan.run(xin.data_ptr(), xout.data_ptr(), sh)
    e1.record(st)
    torch.cuda.synchronize()
    plan.close()
    ms = e0.elapsed_time(e1) / steps
    return n, len(idx), ms, float(len(idx) * n * esz * 2)


def read_traffic(kernel_tag, config="c3"):
    """HBM bytes per launch (or per transform, for the other configurations)
    measured by rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE and scaled by copies
    of known size (profiles/traffic.txt)."""
    path = os.path.join(ROOT, "profiles", "traffic.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic ") and f"config={config} " in line and kernel_tag in line:
            kv = dict(t.split("=", 1) for t in line.split()[1:] if "=" in t)
            try:
                return float(kv["bytes_per_launch"])
            except (KeyError, ValueError):
                return None
    return None


def read_step_traffic(alg_bytes):
    """HBM bytes of one whole step of the default workload: the sum of the
    per-launch traffic of every kernel of the step (profiles/traffic.txt);
    None unless the profiled launches' algorithmic bytes add up to the
    step's (every launch of the step is covered)."""
    path = os.path.join(ROOT, "profiles", "traffic.txt")
    if not os.path.exists(path):
        return None
    tot, alg = 0.0, 0.0
    for line in open(path):
        if line.startswith("traffic ") and "config=c3 " in line:
            kv = dict(t.split("=", 1) for t in line.split()[1:] if "=" in t)
            tot += float(kv["bytes_per_launch"])
            alg += float(kv["algorithmic_bytes"])
    return tot if abs(alg - alg_bytes) <= 0.01 * alg_bytes else None


def config_copy(a, torch, mf):
    """Calibration copies for the HBM byte counters: 512 MiB contiguous, then
    the same source read as 64-byte and as 32-byte segments at a 128-byte
    stride (whole 128-byte lines fetched, half / a quarter of them used)."""
    src = torch.zeros(128 << 20, dtype=torch.float32, device="cuda")  # 512 MiB
    dst = torch.empty_like(src)
    v = src.view(-1, 32)  # rows of 128 bytes
    d64 = torch.empty((v.shape[0], 16), dtype=torch.float32, device="cuda")
    d32 = torch.empty((v.shape[0], 8), dtype=torch.float32, device="cuda")
    for _ in range(3):
        dst.copy_(src)
        d64.copy_(v[:, :16])
        d32.copy_(v[:, :8])
    torch.cuda.synchronize()
    return {"copies": "512 MiB contiguous; 64 B of every 128 B (256 MiB); 32 B of every 128 B (128 MiB)",
            "bytes_read": [512 << 20, 256 << 20, 128 << 20], "bytes_written": [512 << 20, 256 << 20, 128 << 20]}


def config_groups(a, torch, mf):
    """Each length group of the default workload alone (its own plan, tuned
    as in the batch, settled), then the whole batch: where the step's time
    goes.  The short lengths (2^6..2^12) form one unit, as in the batch."""
    sizes = workload()
    offs = np.concatenate([[0], np.cumsum(sizes)[:-1]])
    total = int(sizes.sum())
    g = torch.Generator(device="cuda")
    g.manual_seed(SEED)
    xin = torch.view_as_complex((torch.rand(total, 2, generator=g, device="cuda") * 2 - 1).contiguous())
    xout = torch.empty_like(xin)
    st = torch.cuda.Stream()
    sh = st.cuda_stream
    groups = [[65536], [32768], [16384], [8192], [64, 128, 256, 512, 1024, 2048, 4096]]
    rows = []
    for grp in groups:
        idx = np.nonzero(np.isin(sizes, grp))[0]
        plan = mf.Batch([("dft", (int(sizes[i]),), int(offs[i]) * 8, int(offs[i]) * 8) for i in idx], mf.PREC_FP32)
        ms = _time_enqueue(torch, lambda: plan.run(xin.data_ptr(), xout.data_ptr(), sh), a.steps, a.warmup, st)
        nb = float(sizes[idx].sum()) * 16
        rows.append({"lengths": grp if len(grp) > 1 else grp[0], "entries": int(len(idx)), "ms": round(ms, 5),
                     "alg_bytes": nb, "frac": round(nb / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                     "schedule": [ln.split(": ", 1)[1] for ln in plan.describe().strip().split("\n")][0]})
        log(f"[bench] group {rows[-1]['lengths']} x{len(idx)}: {ms * 1e3:.1f} us, frac {rows[-1]['frac']}")
        plan.close()
    plan = mf.Batch([("dft", (int(n),), int(o) * 8, int(o) * 8) for n, o in zip(sizes, offs)], mf.PREC_FP32)
    ms = _time_enqueue(torch, lambda: plan.run(xin.data_ptr(), xout.data_ptr(), sh), a.steps, a.warmup, st)
    plan.close()
    alg = float(total) * 16
    flops = float(np.sum(5.0 * sizes * np.log2(sizes)))
    return _line(METRIC, round(flops / (ms * 1e-3) / 1e9, 3), "GFLOP/s", ms, a.steps, a.warmup, "f32",
                 "default batch (8192 x 1D C2C FP32) and each of its length groups alone",
                 {"groups": rows, "sum_of_groups_ms": round(sum(r["ms"] for r in rows), 5),
                  "step_frac": round(alg / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)})


def check_sample(xin, xout, sizes, offs, nsample=48):
    """rel-L2 and max-abs error vs the FP64 oracle on a spread of entries."""
    from oracle.oracle import Oracle
    o = Oracle(64)
    idx = np.unique(np.linspace(0, len(sizes) - 1, nsample).astype(int))
    # always include one entry of each length
    for n in np.unique(sizes):
        idx = np.append(idx, np.nonzero(sizes == n)[0][0])
    idx = np.unique(idx)
    worst_rel, worst_abs, num, den = 0.0, 0.0, 0.0, 0.0
    xi = xin.cpu().numpy()
    xo = xout.cpu().numpy()
    for i in idx:
        n, o0 = int(sizes[i]), int(offs[i])
        x = xi[o0:o0 + n].astype(np.complex128)
        ref = o.run("dft", x, (n,))
        y = xo[o0:o0 + n].astype(np.complex128)
        err = np.linalg.norm(y - ref) / np.linalg.norm(ref)
        worst_rel = max(worst_rel, err)
        worst_abs = max(worst_abs, float(np.max(np.abs(y - ref))))
        num += float(np.linalg.norm(y - ref) ** 2)
        den += float(np.linalg.norm(ref) ** 2)
    return {"entries_checked": int(len(idx)), "rel_l2_max": worst_rel,
            "rel_l2_pooled": math.sqrt(num / den), "max_abs": worst_abs}


def _cpu_harness(ref):
    """oracle/libcpubatch.so (a C timing loop over the reference's
    functions on POSIX threads) bound to `ref`'s FP32 dft, or None if it is
    not built."""
    import ctypes
    p = os.path.join(ROOT, "oracle", "libcpubatch.so")
    if not os.path.exists(p):
        return None
    h = ctypes.CDLL(p)
    f = h.cpu_batch_seconds
    f.restype = ctypes.c_double
    f.argtypes = [ctypes.POINTER(ctypes.c_int), ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                  ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
    addr = lambda name: ctypes.cast(ref.fn(name), ctypes.c_void_p).value  # noqa: E731

    def run(sizes, threads, reps):
        arr = (ctypes.c_int * len(sizes))(*[int(n) for n in sizes])
        return f(arr, len(sizes), threads, reps, 8, addr("mkaux_dft"), addr("dft"), addr("free_aux"))
    return run


def _ref32():
    from oracle.oracle import Reference, ref_available
    flags = open("/proc/cpuinfo").read()
    fast = ref_available(32, fast=True) and " avx2 " in flags and " fma " in flags
    if not ref_available(32, fast=fast):
        return None, fast
    return Reference(32, fast=fast), fast


def cpu_baseline(sizes, budget_s=12.0):
    """Reference minfft (oracle/_ref, FP32, -O3) on one host core over the
    workload's entries: the C harness (oracle/cpu_batch.c) runs the whole
    workload on one thread, repeated to fill a few seconds; without it,
    entries in order through ctypes until the time budget is spent."""
    ref, fast = _ref32()
    if ref is None:
        return None
    opt = "-O3 -march=x86-64-v3" if fast else "-O2"
    harness = _cpu_harness(ref)
    if harness is not None:
        flops = float(np.sum(5.0 * sizes * np.log2(sizes)))
        t = harness(sizes, 1, 1)  # one pass (~0.5 s) sizes the timed run
        if t > 0:
            reps = max(1, int(math.ceil(3.0 / t)))
            t = harness(sizes, 1, reps)
        if t > 0:
            return {"value": reps * flops / t / 1e9, "unit": "GFLOP/s", "cores": 1, "kind": "reference",
                    "sample": f"the whole workload ({len(sizes)} entries) {reps} times on one thread, reference "
                              f"minfft FP32 {opt} called from a C loop (oracle/cpu_batch.c), one aux per size"}
    rng = np.random.default_rng(SEED + 1)
    auxes = {}
    flops = 0.0
    spent = 0.0
    done = 0
    pts = 0
    f = ref.fn("dft")
    for n in sizes:
        n = int(n)
        if n not in auxes:
            auxes[n] = ref.mkaux("dft", (n,))
        x = (rng.uniform(-1, 1, n) + 1j * rng.uniform(-1, 1, n)).astype(np.complex64)
        y = np.empty_like(x)
        xp, yp, ap = x.ctypes.data, y.ctypes.data, auxes[n]
        t0 = time.perf_counter()
        f(xp, yp, ap)
        spent += time.perf_counter() - t0
        flops += 5.0 * n * math.log2(n)
        done += 1
        pts += n
        if spent >= budget_s:
            break
    for a in auxes.values():
        ref.free(a)
    return {"value": flops / spent / 1e9, "unit": "GFLOP/s", "cores": 1, "kind": "reference",
            "sample": f"first {done} of {len(sizes)} entries ({pts} points), reference minfft "
                      f"FP32 {opt}, one transform at a time through ctypes"}


def cpu_baseline_all_cores(sizes, single, budget_s=4.0, max_threads=16):
    """Reference minfft (oracle/_ref, FP32) on the host's cores: one thread
    per core (at most 16, the host share of one GPU), each with its own
    auxes (a reference aux owns scratch), thread k taking entries k, k+T, ...
    of the workload, repeated to fill about budget_s of wall time -- in the
    C harness (oracle/cpu_batch.c) when it is built, else in Python threads
    (ctypes releases the GIL inside the calls)."""
    if not single:
        return None
    ref, fast = _ref32()
    if ref is None:
        return None
    opt = "-O3 -march=x86-64-v3" if fast else "-O2"
    threads = max(1, min(max_threads, os.cpu_count() or 1))
    step_flops = float(sum(5.0 * int(n) * math.log2(int(n)) for n in sizes))
    reps = max(1, int(math.ceil(budget_s * threads * single["value"] * 1e9 / step_flops)))
    harness = _cpu_harness(ref)
    if harness is not None:
        wall = harness(sizes, threads, reps)
        if wall > 0:
            return {"value": reps * step_flops / wall / 1e9, "unit": "GFLOP/s", "cores": threads,
                    "kind": "reference",
                    "sample": f"the whole workload {reps} times, entries split over {threads} threads, reference "
                              f"minfft FP32 {opt} called from C (oracle/cpu_batch.c), one aux per size per thread"}
    from concurrent.futures import ThreadPoolExecutor
    f = ref.fn("dft")
    nmax = int(max(sizes))

    def worker(k):
        auxes = {}
        rng = np.random.default_rng(SEED + 2 + k)
        x = (rng.uniform(-1, 1, nmax) + 1j * rng.uniform(-1, 1, nmax)).astype(np.complex64)
        y = np.empty_like(x)
        xp, yp = x.ctypes.data, y.ctypes.data
        mine = [int(n) for n in sizes[k::threads]]
        for n in set(mine):
            auxes[n] = ref.mkaux("dft", (n,))
        barrier.wait()
        for _ in range(reps):
            for n in mine:
                f(xp, yp, auxes[n])
        for a in auxes.values():
            ref.free(a)

    import threading
    barrier = threading.Barrier(threads + 1)
    with ThreadPoolExecutor(max_workers=threads) as ex:
        futs = [ex.submit(worker, k) for k in range(threads)]
        barrier.wait()
        t0 = time.perf_counter()
        for fu in futs:
            fu.result()
        wall = time.perf_counter() - t0
    return {"value": reps * step_flops / wall / 1e9, "unit": "GFLOP/s", "cores": threads, "kind": "reference",
            "sample": f"the whole workload {reps} times, entries split over {threads} threads, reference minfft "
                      f"FP32 {opt} through ctypes in Python threads, one aux per size per thread"}


# --------------------------------------------------------------------------
# other configurations (printed with --config; the default line is C3)
# --------------------------------------------------------------------------
def _events(torch, st):
    return torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)


def _time_enqueue(torch, fn, steps, warmup, st):
    fn()
    SETTLED.append(settle(torch, fn))
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    e0, e1 = _events(torch, st)
    e0.record(st)
    for _ in range(steps):
        fn()
    e1.record(st)
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / steps


def settle_info():
    """What the untimed settling replays of this run were (see SETTLE_MS)."""
    return {"min_ms_per_workload": SETTLE_MS, "untimed_replays": [n for n, _ in SETTLED],
            "what": "untimed replays of each timed workload before its warm-up steps (device clock and "
                    "cache state after idle); the timed steps are unchanged"}


def _line(metric, value, unit, ms, steps, warmup, dtype, workload, extra, higher=True):
    out = {"metric": metric, "value": value, "unit": unit, "n_gpus": 1, "steps": steps, "warmup": warmup,
           "ms_per_step": round(ms, 5), "higher_is_better": higher, "scaling": "weak", "vs_baseline": None,
           "dtype": dtype, "data": "synthetic", "config": {"workload": workload, "parallelism": "single GPU"}}
    out.update(extra)
    out["settle"] = settle_info()
    return out


def config_c1(a, torch, mf):
    """1D C2C N=1024 FP64, dft then invdft, single transform."""
    from oracle.oracle import Oracle, Reference, ref_available
    N = 1024
    rng = np.random.default_rng(1)
    x = (rng.uniform(-1, 1, N) + 1j * rng.uniform(-1, 1, N)).astype(np.complex128)
    lib = mf.Minfft(64)
    aux = mf.Aux(lib, "dft", (N,))
    st = torch.cuda.Stream()
    xd = torch.from_numpy(x).cuda()
    yd = torch.empty_like(xd)
    bd = torch.empty_like(xd)
    sh = st.cuda_stream

    def step():
        aux.enqueue("dft", xd, yd, sh)
        aux.enqueue("invdft", yd, bd, sh)
    ms = _time_enqueue(torch, step, a.steps * 50, a.warmup, st)
    # host-pointer (reference-style) latency: the 27-symbol functions called
    # through ctypes with the pointers taken once, as the CPU baseline below
    # is (the error channel is read once, after the loop)
    y = np.empty_like(x)
    b = np.empty_like(x)
    fd, fi = lib.lib.minfft_dft, lib.lib.minfft_invdft
    xp, yp, bp, ah = x.ctypes.data, y.ctypes.data, b.ctypes.data, aux.h
    for _ in range(20):
        fd(xp, yp, ah)
        fi(yp, bp, ah)
    reps = 1000
    t0 = time.perf_counter()
    for _ in range(reps):
        fd(xp, yp, ah)
        fi(yp, bp, ah)
    host_us = (time.perf_counter() - t0) / reps * 1e6
    err = mf.last_error()
    assert not err, err
    assert np.linalg.norm(b / N - x) <= 1e-13 * np.linalg.norm(x)
    o = Oracle(64)
    ref = o.run("dft", x, (N,))
    yg = yd.cpu().numpy()
    rel = float(np.linalg.norm(yg - ref) / np.linalg.norm(ref))
    rt = float(np.linalg.norm(bd.cpu().numpy() / N - x) / np.linalg.norm(x))
    relref = None
    if ref_available(64):
        r = Reference(64).run("dft", x, (N,))
        relref = float(np.linalg.norm(yg - r) / np.linalg.norm(r))
    cpu = None
    if ref_available(64, fast=True):
        R = Reference(64, fast=True)
        ra = R.mkaux("dft", (N,))
        f = R.fn("dft")
        fi = R.fn("invdft")
        yy = np.empty_like(x)
        bb = np.empty_like(x)
        xq, yq, bq = x.ctypes.data, yy.ctypes.data, bb.ctypes.data
        reps = 20000
        t0 = time.perf_counter()
        for _ in range(reps):
            f(xq, yq, ra)
            fi(yq, bq, ra)
        dt = (time.perf_counter() - t0) / reps
        R.free(ra)
        cpu = {"value": round(dt * 1e6, 3), "unit": "us per dft+invdft pair", "cores": 1, "kind": "reference",
               "sample": f"{reps} round trips of N=1024 through ctypes (pointers taken once)"}
    aux.close()
    flops = 2 * 5.0 * N * 10
    return _line("1D C2C FP64 dft+invdft latency", round(ms * 1000, 3), "us per dft+invdft pair (device)", ms,
                 a.steps * 50, a.warmup, "f64", "C1: N=1024 C2C FP64 round trip",
                 {"gflops": round(flops / (ms * 1e-3) / 1e9, 3), "host_pointer_us": round(host_us, 2),
                  "rel_l2_vs_oracle": rel, "rel_l2_vs_reference": relref, "round_trip_rel_l2": rt,
                  "tolerance": 1e-14 * 10, "cpu_baseline": cpu}, higher=False)


def config_c2(a, torch, mf):
    """2D real DFT and DCT-2/DCT-3 on 2048x2048 FP64."""
    from oracle.oracle import Oracle, Reference, ref_available
    n = 2048
    shape = (n, n)
    g = torch.Generator(device="cuda")
    g.manual_seed(2)
    xd = torch.randn(shape, generator=g, device="cuda", dtype=torch.float64)
    zd = torch.empty((n, n // 2 + 1), dtype=torch.complex128, device="cuda")
    yd = torch.empty_like(xd)
    cd = torch.empty_like(xd)
    dd = torch.empty_like(xd)
    lib = mf.Minfft(64)
    ar = mf.Aux(lib, "realdft", shape)
    at = mf.Aux(lib, "t2t3", shape)
    st = torch.cuda.Stream()
    sh = st.cuda_stream
    res = {}
    for name, fn in (("realdft", lambda: ar.enqueue("realdft", xd, zd, sh)),
                     ("invrealdft", lambda: ar.enqueue("invrealdft", zd, yd, sh)),
                     ("dct2", lambda: at.enqueue("dct2", xd, cd, sh)),
                     ("dct3", lambda: at.enqueue("dct3", cd, dd, sh))):
        res[name] = _time_enqueue(torch, fn, a.steps, a.warmup, st)
    o = Oracle(64)
    x = xd.cpu().numpy()
    t0 = time.perf_counter()
    zr = o.run("realdft", x, shape)
    cr = o.run("dct2", x, shape)
    oracle_s = time.perf_counter() - t0
    N = n * n
    tol = 1e-14 * 22
    err = {"realdft": float(np.linalg.norm(zd.cpu().numpy() - zr) / np.linalg.norm(zr)),
           "dct2": float(np.linalg.norm(cd.cpu().numpy() - cr) / np.linalg.norm(cr)),
           "invrealdft_round_trip": float(np.linalg.norm(yd.cpu().numpy() / N - x) / np.linalg.norm(x)),
           "dct3_of_dct2": float(np.linalg.norm(dd.cpu().numpy() / (4 * N) - x) / np.linalg.norm(x))}
    # host arrays through the synchronous API (PCIe-inclusive): the aux's
    # pinned bounce buffer in 4 MiB pieces, host copies overlapping the DMA,
    # against one piece (MINFFT_HOST_PIPE=0)
    host = {}
    zh = np.empty((n, n // 2 + 1), dtype=np.complex128)
    for pipe in ("1", "0"):
        os.environ["MINFFT_HOST_PIPE"] = pipe
        ar("realdft", x, zh)
        t0 = time.perf_counter()
        for _ in range(10):
            ar("realdft", x, zh)
        host["pipelined" if pipe == "1" else "one_piece"] = round((time.perf_counter() - t0) / 10 * 1e3, 3)
    os.environ.pop("MINFFT_HOST_PIPE", None)
    host["unit"] = "ms per realdft of 2048x2048 FP64 host arrays (copies included)"
    cpu = None
    if ref_available(64, fast=True):
        R = Reference(64, fast=True)
        ra = R.mkaux("realdft", shape)
        zz = np.empty((n, n // 2 + 1), dtype=np.complex128)
        t0 = time.perf_counter()
        R.fn("realdft")(x.ctypes.data, zz.ctypes.data, ra)
        dt = time.perf_counter() - t0
        R.free(ra)
        cpu = {"value": round(2.5 * N * 22 / dt / 1e9, 3), "unit": "GFLOP/s (2.5 N log2 N)", "cores": 1,
               "kind": "reference", "sample": f"one 2048x2048 realdft, {dt*1e3:.1f} ms"}
    alg = {"realdft": N * 8 + n * (n // 2 + 1) * 16, "invrealdft": N * 8 + n * (n // 2 + 1) * 16,
           "dct2": 2 * N * 8, "dct3": 2 * N * 8}
    flops = 2.5 * N * 22
    per = {k: {"ms": round(v, 5), "gflops": round(flops / (v * 1e-3) / 1e9, 2),
               "alg_hbm_gbs": round(alg[k] / (v * 1e-3) / 1e9, 1)} for k, v in res.items()}
    ms = res["realdft"]
    return _line("2D real DFT 2048x2048 FP64 GFLOP/s", round(flops / (ms * 1e-3) / 1e9, 3), "GFLOP/s", ms,
                 a.steps, a.warmup, "f64", "C2: 2048x2048 realdft/invrealdft + DCT-2/DCT-3, FP64",
                 {"per_transform": per, "rel_l2": err, "tolerance": tol, "oracle_seconds": round(oracle_s, 2),
                  "host_arrays": host,
                  "roofline": {"bound": "hbm", "achieved": per["realdft"]["alg_hbm_gbs"], "peak": HBM_PEAK_GBS,
                               "unit": "GB/s", "frac": round(per["realdft"]["alg_hbm_gbs"] / HBM_PEAK_GBS, 4),
                               "traffic": read_traffic("kind=realdft", "c2"), "kernel": "realdft (2 axis passes)"},
                  "cpu_baseline": cpu})


def config_rlong(a, torch, mf):
    """Long 1D real transforms whose half-length FFT exceeds one tile: 32
    entries of 2^20 points (FP32 real DFT, DCT-2, DCT-3, DCT-4; FP64 real
    DFT, 16 entries) as batches, two HBM passes (pre/post-processing fused
    into the two-step FFT) against the four-pass form (schedule rlong=0)."""
    from oracle.oracle import Oracle, Reference, ref_available
    n = 1 << 20
    st = torch.cuda.Stream()
    sh = st.cuda_stream
    o = Oracle(64)
    res = {}
    for prec, kind, cnt in ((32, "realdft", 32), (32, "dct2", 32), (32, "dct3", 32), (32, "dct4", 32),
                            (64, "realdft", 16)):
        rs = 4 if prec == 32 else 8
        ient = n * rs
        oent = (n // 2 + 1) * 2 * rs if kind == "realdft" else n * rs
        stride = (max(ient, oent) + 4095) // 4096 * 4096
        g = torch.Generator(device="cuda")
        g.manual_seed(9)
        dt = torch.float32 if prec == 32 else torch.float64
        xin = torch.rand(cnt * stride // rs, generator=g, device="cuda", dtype=dt) * 2 - 1
        xout = torch.empty_like(xin)
        b = mf.Batch([(kind, (n,), i * stride, i * stride) for i in range(cnt)],
                     mf.PREC_FP32 if prec == 32 else mf.PREC_FP64)
        ng = len(b.describe().strip().split("\n"))
        ms = {}
        for spec in ("rlong=0", "rlong=1"):
            for gi in range(ng):
                b.schedule(gi, spec)
            ms[spec] = _time_enqueue(torch, lambda: b.run(xin.data_ptr(), xout.data_ptr(), sh), a.steps, a.warmup, st)
        b.close()
        # the first entry against the FP64 oracle
        x0 = xin[: n].cpu().numpy().astype(np.float64)
        ref = o.run(kind, x0, (n,))
        odt = (torch.complex64 if prec == 32 else torch.complex128) if kind == "realdft" else dt
        nout = n // 2 + 1 if kind == "realdft" else n
        y0 = xout.view(torch.uint8)[: nout * (2 if kind == "realdft" else 1) * rs].view(odt).cpu().numpy()
        err = float(np.linalg.norm(y0 - ref) / np.linalg.norm(ref))
        alg = cnt * (ient + oent)
        res[f"fp{prec}_{kind}"] = {
            "entries": cnt, "two_pass_ms": round(ms["rlong=1"], 5), "four_pass_ms": round(ms["rlong=0"], 5),
            "speedup": round(ms["rlong=0"] / ms["rlong=1"], 3),
            "alg_hbm_gbs": round(alg / (ms["rlong=1"] * 1e-3) / 1e9, 1),
            "frac": round(alg / (ms["rlong=1"] * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
            "traffic": read_traffic(f"kind={kind} prec={prec}", "rlong"), "alg_bytes": alg,
            "rel_l2_vs_fp64_oracle": err}
        del xin, xout
        torch.cuda.empty_cache()
    cpu = None
    if not a.no_cpu and ref_available(32, fast=True):
        R = Reference(32, fast=True)
        ra = R.mkaux("realdft", (n,))
        xh = np.random.default_rng(9).uniform(-1, 1, n).astype(np.float32)
        zh = np.empty(n // 2 + 1, dtype=np.complex64)
        t0 = time.perf_counter()
        for _ in range(8):
            R.fn("realdft")(xh.ctypes.data, zh.ctypes.data, ra)
        dt = (time.perf_counter() - t0) / 8
        R.free(ra)
        cpu = {"value": round(2.5 * n * 20 / dt / 1e9, 3), "unit": "GFLOP/s (2.5 N log2 N)", "cores": 1,
               "kind": "reference", "sample": f"8 FP32 realdft of 2^20 points, {dt*1e3:.2f} ms each"}
    main = res["fp32_realdft"]
    flops = 2.5 * n * 20 * main["entries"]
    return _line("long 1D real DFT 2^20 FP32 GFLOP/s", round(flops / (main["two_pass_ms"] * 1e-3) / 1e9, 3),
                 "GFLOP/s", main["two_pass_ms"], a.steps, a.warmup, "f32",
                 "32 x 1D realdft of 2^20 points (and DCT-2/3/4; FP64 realdft), two HBM passes vs four",
                 {"per_kind": res,
                  "roofline": {"bound": "hbm", "achieved": main["alg_hbm_gbs"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                               "frac": main["frac"], "traffic": main["traffic"],
                               "kernel": "rstep1_kernel + rstep2_kernel (2 passes)"},
                  "cpu_baseline": cpu})


def config_c4(a, torch, mf):
    """3D C2C 512^3 complex64 (FP32 path) vs the FP64 path on the same input."""
    n = 512
    shape = (n, n, n)
    rng = np.random.default_rng(4)
    ks = set()
    while len(ks) < 64:
        ks.add(tuple(int(v) for v in rng.integers(0, n, 3)))
    ks = sorted(ks)
    phis = rng.uniform(0, 2 * np.pi, 64)
    # the plane-wave field = unnormalised inverse DFT of its 64-spike spectrum
    spec = torch.zeros(shape, dtype=torch.complex128, device="cuda")
    for (kx, ky, kz), p in zip(ks, phis):
        spec[kx, ky, kz] = complex(np.cos(p), np.sin(p))
    lib64 = mf.Minfft(64)
    lib32 = mf.Minfft(32)
    a64 = mf.Aux(lib64, "dft", shape)
    st = torch.cuda.Stream()
    sh = st.cuda_stream
    field = torch.empty_like(spec)
    a64.enqueue("invdft", spec, field, sh)
    torch.cuda.synchronize()
    del spec
    g = torch.Generator(device="cuda")
    g.manual_seed(4)
    field += torch.complex(torch.randn(shape, generator=g, device="cuda", dtype=torch.float64),
                           torch.randn(shape, generator=g, device="cuda", dtype=torch.float64)) * 1e-3
    x32 = field.to(torch.complex64)
    del field
    torch.cuda.empty_cache()
    y32 = torch.empty_like(x32)
    # split-FP16 MFMA form: one batch entry of 512^3 (radix-16 stages on MFMA)
    ps = mf.Batch([("dft", shape, 0, 0)], mf.PREC_SPLIT_FP16)
    ms_s = _time_enqueue(torch, lambda: ps.run(x32.data_ptr(), y32.data_ptr(), sh), a.steps, a.warmup, st)
    ps.close()
    ys = y32.clone()
    # FP32 form (all passes on the vector ALUs)
    a32 = mf.Aux(lib32, "dft", shape)
    ms32 = _time_enqueue(torch, lambda: a32.enqueue("dft", x32, y32, sh), a.steps, a.warmup, st)
    x64 = x32.to(torch.complex128)
    y64 = torch.empty_like(x64)
    ms64 = _time_enqueue(torch, lambda: a64.enqueue("dft", x64, y64, sh), a.steps, a.warmup, st)
    nrm = torch.linalg.vector_norm(y64)
    d = (ys.to(torch.complex128) - y64)
    rel = float(torch.linalg.vector_norm(d) / nrm)
    mx = float(d.abs().max())
    d32 = (y32.to(torch.complex128) - y64)
    rel32 = float(torch.linalg.vector_norm(d32) / nrm)
    del d, d32
    # closed form at the spikes: 512^3 e^{i phi}
    spikes = torch.stack([y64[kx, ky, kz] for (kx, ky, kz) in ks]).cpu().numpy()
    want = (n ** 3) * np.exp(1j * phis)
    spike_rel = float(np.max(np.abs(spikes - want)) / n ** 3)
    N = n ** 3
    alg = 2 * N * 8
    flops = 5.0 * N * 27
    a32.close()
    a64.close()
    ms = ms_s
    # CPU baseline: the reference's own 3D transform (FP32: its aux of
    # 2^27 points still fits the reference's int sizes) on this input, once
    cpu = None
    from oracle.oracle import Reference, ref_available
    if not a.no_cpu and ref_available(32, fast=True):
        R = Reference(32, fast=True)
        xh = x32.cpu().numpy()
        yh = np.empty_like(xh)
        ra = R.mkaux("dft", shape)
        t0 = time.perf_counter()
        R.fn("dft")(xh.ctypes.data, yh.ctypes.data, ra)
        dt = time.perf_counter() - t0
        R.free(ra)
        del xh, yh
        cpu = {"value": round(flops / dt / 1e9, 3), "unit": "GFLOP/s", "cores": 1, "kind": "reference",
               "sample": f"one 512^3 transform of this input, reference minfft FP32 -O3, {dt:.2f} s"}
    return _line("3D C2C 512^3 complex64 GFLOP/s", round(flops / (ms * 1e-3) / 1e9, 3), "GFLOP/s", ms,
                 a.steps, a.warmup, "f32+split-f16",
                 "C4: 512^3 C2C, 64 plane waves + N(0,1e-3) noise; radix-16 stages as split-FP16 MFMA",
                 {"fp32_ms": round(ms32, 4), "fp64_gpu_ms": round(ms64, 4),
                  "speedup_vs_fp64_gpu": round(ms64 / ms, 3),
                  "rel_l2_vs_fp64": rel, "max_abs_vs_fp64": mx, "fp32_rel_l2_vs_fp64": rel32,
                  "spike_closed_form_rel": spike_rel,
                  "roofline": {"bound": "hbm", "achieved": round(alg / (ms * 1e-3) / 1e9, 1),
                               "peak": HBM_PEAK_GBS, "unit": "GB/s",
                               "frac": round(alg / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                               "traffic": read_traffic("kind=split-fp16", "c4"), "kernel": "3 axis passes"},
                  "cpu_baseline": cpu})


def config_c5(a, torch, mf):
    """256 images 1024x1024 FP32, DCT-2 / DST-2 / DCT-4 by image index mod 3, one batch."""
    from oracle.oracle import Oracle
    n, count = 1024, 256
    kinds = ("dct2", "dst2", "dct4")
    g = torch.Generator(device="cuda")
    g.manual_seed(5)
    m = torch.rand((count, n // 8, n // 8), generator=g, device="cuda", dtype=torch.float32)
    x = m.repeat_interleave(8, 1).repeat_interleave(8, 2)
    x += (torch.rand((count, n, n), generator=g, device="cuda", dtype=torch.float32) - 0.5) * 0.1
    x = x.contiguous()
    y = torch.empty_like(x)
    img = n * n * 4
    plan = mf.Batch([(kinds[i % 3], (n, n), i * img, i * img) for i in range(count)], mf.PREC_FP32)
    st = torch.cuda.Stream()
    sh = st.cuda_stream
    ms = _time_enqueue(torch, lambda: plan.run(x.data_ptr(), y.data_ptr(), sh), a.steps, a.warmup, st)
    plan.close()
    o = Oracle(64)
    errs = {}
    for i in range(3):
        xi = x[i].cpu().numpy().astype(np.float64)
        r = o.run(kinds[i], xi, (n, n))
        errs[kinds[i]] = float(np.linalg.norm(y[i].cpu().numpy() - r) / np.linalg.norm(r))
    # FP64 batch on the same workload
    x64 = x.to(torch.float64)
    y64 = torch.empty_like(x64)
    p64 = mf.Batch([(kinds[i % 3], (n, n), i * img * 2, i * img * 2) for i in range(count)], mf.PREC_FP64)
    ms64 = _time_enqueue(torch, lambda: p64.run(x64.data_ptr(), y64.data_ptr(), sh), a.steps, a.warmup, st)
    p64.close()
    N = n * n * count
    alg = 2 * N * 4
    flops = 2.5 * N * 20
    # CPU baseline: the reference's own 2D transforms (FP32) on 6 images,
    # two of each type
    cpu = None
    from oracle.oracle import Reference, ref_available
    if not a.no_cpu and ref_available(32, fast=True):
        R = Reference(32, fast=True)
        dt = 0.0
        for i in range(6):
            xh = x[i].cpu().numpy()
            yh = np.empty_like(xh)
            ra = R.mkaux(kinds[i % 3], (n, n))
            t0 = time.perf_counter()
            R.fn(kinds[i % 3])(xh.ctypes.data, yh.ctypes.data, ra)
            dt += time.perf_counter() - t0
            R.free(ra)
        cpu = {"value": round(6 * 2.5 * n * n * 20 / dt / 1e9, 3), "unit": "GFLOP/s", "cores": 1,
               "kind": "reference", "sample": f"6 of the 256 images (2 per type), reference minfft FP32 -O3, "
                                               f"{dt * 1e3:.0f} ms"}
    return _line("batched 2D DCT-2/DST-2/DCT-4 1024x1024 FP32 GFLOP/s", round(flops / (ms * 1e-3) / 1e9, 3),
                 "GFLOP/s", ms, a.steps, a.warmup, "f32",
                 "C5: 256 x 1024^2 FP32, type = (DCT-2, DST-2, DCT-4)[i mod 3], one batch",
                 {"rel_l2_vs_fp64_oracle": errs, "fp64_gpu_ms": round(ms64, 4),
                  "speedup_vs_fp64_gpu": round(ms64 / ms, 3),
                  "roofline": {"bound": "hbm", "achieved": round(alg / (ms * 1e-3) / 1e9, 1), "peak": HBM_PEAK_GBS,
                               "unit": "GB/s", "frac": round(alg / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                               "traffic": read_traffic("kind=fp32", "c5"), "kernel": "2 axis passes per image"},
                  "cpu_baseline": cpu})


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-fp64", action="store_true", help="skip the FP64 comparison leg")
    ap.add_argument("--copy-probe", action="store_true",
                    help="also run a 512 MiB device copy (calibrates the HBM byte counters)")
    ap.add_argument("--config", default="c3", choices=["c1", "c2", "c3", "c4", "c5", "rlong", "copy", "groups"],
                    help="c3 (default) is the metric's workload; the others print their own line")
    ap.add_argument("--settle-ms", type=float, default=SETTLE_MS,
                    help="untimed replays of each timed workload before its warm-up steps, in ms (0 = none)")
    a = ap.parse_args()
    globals()["SETTLE_MS"] = a.settle_ms
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: the batched transforms are independent, so "
              "N GPUs would each run their own share; multi-GPU runs are not part of this build.",
              file=sys.stderr)
        sys.exit(2)
    import torch
    import minfft_amd as mf

    if a.config != "c3":
        fn = {"c1": config_c1, "c2": config_c2, "c4": config_c4, "c5": config_c5, "rlong": config_rlong,
              "copy": config_copy, "groups": config_groups}[a.config]
        print(json.dumps(fn(a, torch, mf)), flush=True)
        return

    sizes = workload()
    total_pts = int(sizes.sum())
    flops = float(np.sum(5.0 * sizes * np.log2(sizes)))
    alg_bytes = float(total_pts) * 8 * 2  # complex64 read + write
    log(f"[bench] workload: {len(sizes)} entries, {total_pts} points, {alg_bytes / 1e9:.3f} GB")

    ms32, xin, xout, offs, plan = gpu_batch(sizes, mf.PREC_FP32, a.steps, a.warmup, torch, mf)
    log(f"[bench] fp32 batch: {ms32:.4f} ms/step")
    schedule = [ln.split(": ", 1)[1] for ln in plan.describe().strip().split("\n") if ": " in ln]
    log("[bench] schedule:\n  " + "\n  ".join(schedule))
    acc = check_sample(xin, xout, sizes, offs)
    log(f"[bench] accuracy: {acc}")
    plan.close()
    dom_n, dom_cnt, dom_ms, dom_bytes = dominant_group(sizes, xin, xout, offs, mf.PREC_FP32, a.steps, torch, mf)
    log(f"[bench] dominant group N={dom_n} x{dom_cnt}: {dom_ms:.4f} ms")
    del xin, xout
    torch.cuda.empty_cache()

    fp64 = None
    if not a.no_fp64:
        ms64, xi64, xo64, _, p64 = gpu_batch(sizes, mf.PREC_FP64, a.steps, a.warmup, torch, mf)
        fp64 = {"ms_per_step": ms64, "gflops": flops / (ms64 * 1e-3) / 1e9,
                "hbm_gbs_algorithmic": 2 * alg_bytes / (ms64 * 1e-3) / 1e9}
        log(f"[bench] fp64 batch: {ms64:.4f} ms/step")
        p64.close()
        del xi64, xo64
        torch.cuda.empty_cache()

    cpu = None if a.no_cpu else cpu_baseline(sizes)
    if cpu:
        log(f"[bench] cpu baseline: {cpu}")
    cpu_all = None if a.no_cpu else cpu_baseline_all_cores(sizes, cpu)
    if cpu_all:
        log(f"[bench] cpu baseline, all cores: {cpu_all}")

    gflops = flops / (ms32 * 1e-3) / 1e9
    step_achieved = alg_bytes / (ms32 * 1e-3) / 1e9
    achieved = dom_bytes / (dom_ms * 1e-3) / 1e9
    dom_tag = f"N={dom_n}"
    traffic = read_traffic(dom_tag, "c3")
    step_traffic = read_step_traffic(alg_bytes)
    out = {
        "metric": METRIC,
        "value": round(gflops, 3),
        "unit": "GFLOP/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(ms32, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "8192 x 1D C2C, N=2^6..2^16 log-uniform (seed 3), FP32 data, FP64-generated twiddles",
                   "entries": int(len(sizes)), "points": total_pts, "flop_convention": "5 N log2 N",
                   "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "kernel": f"dominant length group {dom_tag} x{dom_cnt} (fused two-step launch)",
                     "algorithmic_bytes": dom_bytes, "ms_per_launch": round(dom_ms, 5)},
        "step_roofline": {"achieved": round(step_achieved, 2), "frac": round(step_achieved / HBM_PEAK_GBS, 4),
                          "algorithmic_bytes": alg_bytes, "what": "whole batch step, all lengths, one graph launch",
                          "traffic": step_traffic,
                          "traffic_gbs": round(step_traffic / (ms32 * 1e-3) / 1e9, 2) if step_traffic else None},
        "schedule": {"autotuned": os.environ.get("MINFFT_AUTOTUNE", "1") != "0" and "MINFFT_TUNE" not in os.environ,
                     "groups": schedule},
        "accuracy_vs_fp64_oracle": acc,
        "fp64_gpu_same_workload": fp64,
        "speedup_vs_fp64_gpu": round(fp64["ms_per_step"] / ms32, 3) if fp64 else None,
        "cpu_baseline": cpu,
        "cpu_baseline_all_cores": cpu_all,
        "settle": settle_info(),
    }
    print(json.dumps(out), flush=True)
    if a.copy_probe:
        src = torch.empty(512 << 20, dtype=torch.uint8, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        log("[bench] copy probe: 3 copies of 536870912 bytes")


if __name__ == "__main__":
    main()
