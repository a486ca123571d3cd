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
    for _ in range(max(warmup, 1)):
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


def cpu_baseline(sizes, budget_s=12.0):
    """Reference minfft (oracle/_ref, FP32, -O3) on one host core, over the
    workload's entries in order until the time budget is spent."""
    from oracle.oracle import Reference, ref_available
    flags = open("/proc/cpuinfo").read()
    fast = ref_available(32, fast=True) and " avx2 " in flags and " fma " in flags
    if not ref_available(32, fast=fast):
        return None
    ref = Reference(32, fast=fast)
    rng = np.random.default_rng(SEED + 1)
    auxes = {}
    flops = 0.0
    spent = 0.0
    done = 0
    pts = 0
    for n in sizes:
        n = int(n)
        if n not in auxes:
            auxes[n] = ref.mkaux("dft", (n,))
        x = (rng.uniform(-1, 1, n) + 1j * rng.uniform(-1, 1, n)).astype(np.complex64)
        y = np.empty_like(x)
        f = ref.fn("dft")
        t0 = time.perf_counter()
        f(x.ctypes.data, y.ctypes.data, auxes[n])
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
                      f"FP32 {'-O3 -march=x86-64-v3' if fast else '-O2'}, one transform at a time"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-fp64", action="store_true", help="skip the FP64 comparison leg")
    a = ap.parse_args()
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: the batched transforms are independent, so "
              "N GPUs would each run their own share; multi-GPU runs are not part of this build.",
              file=sys.stderr)
        sys.exit(2)
    import torch
    import minfft_amd as mf

    sizes = workload()
    total_pts = int(sizes.sum())
    flops = float(np.sum(5.0 * sizes * np.log2(sizes)))
    alg_bytes = float(total_pts) * 8 * 2  # complex64 read + write
    log(f"[bench] workload: {len(sizes)} entries, {total_pts} points, {alg_bytes / 1e9:.3f} GB")

    ms32, xin, xout, offs, plan = gpu_batch(sizes, mf.PREC_FP32, a.steps, a.warmup, torch, mf)
    log(f"[bench] fp32 batch: {ms32:.4f} ms/step")
    acc = check_sample(xin, xout, sizes, offs)
    log(f"[bench] accuracy: {acc}")
    plan.close()
    del xin, xout
    torch.cuda.empty_cache()

    fp64 = None
    if not a.no_fp64:
        ms64, xi64, xo64, _, p64 = gpu_batch(sizes, mf.PREC_FP64, max(a.steps // 2, 3), 2, torch, mf)
        fp64 = {"ms_per_step": ms64, "gflops": flops / (ms64 * 1e-3) / 1e9,
                "hbm_gbs_algorithmic": 2 * alg_bytes / (ms64 * 1e-3) / 1e9}
        log(f"[bench] fp64 batch: {ms64:.4f} ms/step")
        p64.close()
        del xi64, xo64
        torch.cuda.empty_cache()

    cpu = None if a.no_cpu else cpu_baseline(sizes)
    if cpu:
        log(f"[bench] cpu baseline: {cpu}")

    gflops = flops / (ms32 * 1e-3) / 1e9
    achieved = alg_bytes / (ms32 * 1e-3) / 1e9
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
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                     "kernel": "whole batch step (one graph launch)",
                     "algorithmic_bytes": alg_bytes},
        "accuracy_vs_fp64_oracle": acc,
        "fp64_gpu_same_workload": fp64,
        "speedup_vs_fp64_gpu": round(fp64["ms_per_step"] / ms32, 3) if fp64 else None,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
