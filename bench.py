#!/usr/bin/env python3
"""Headline benchmark: batched RNS negacyclic NTT on one MI355X.

Workload (one step): 256 polynomials x 4 RNS limbs, N = 4096, limb primes the
four largest q < 2^50 with q = 1 (mod 8192).  Every one of the 1024
limb-polynomials runs forward NTT -> Shoup pointwise multiply by a fixed,
device-resident operand -> inverse NTT, all in ONE kernel launch (2048 NTTs).
Inputs are resident in HBM before timing starts; 12 rotating copies of the
batch (384 MiB) keep the working set larger than the 256 MiB Infinity Cache.

Prints one JSON line: NTT/s (value), HBM GB/s, p50/p99 batch latency, the
plain one-launch-per-limb-polynomial GPU path, the HBM roofline of the fused
kernel and the CPU oracle timed on this host.
"""
import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

LOG_N = 12
N = 1 << LOG_N
LIMBS = 4
POLYS = 256
COUNT = POLYS * LIMBS
NTT_PER_STEP = 2 * COUNT
BYTES_PER_STEP = COUNT * N * 8 * 2  # every coefficient read once and written once
ROTATE = 12
HBM_PEAK_GBPS = 8000.0
SEED = 0x5052494D5553 + 1


def workload(rng):
    import primus_amd as pa

    moduli = pa.ntt_primes_below(50, 2 * N, LIMBS)
    polys = np.stack([rng.integers(0, moduli[i % LIMBS], size=N, dtype=np.uint64) for i in range(COUNT)])
    operand = np.stack([rng.integers(0, q, size=N, dtype=np.uint64) for q in moduli])
    return moduli, polys, operand


def cpu_baseline(moduli, polys, operand, budget_s=8.0):
    """Oracle (C restatement of UintNttTable) on this host: all cores, bounded sample."""
    from oracle import oracle as O

    tables = [O.Table(LOG_N, q) for q in moduli]
    opq = np.stack([O.shoup_quotients(operand[i], moduli[i]) for i in range(LIMBS)])
    threads = min(os.cpu_count() or 1, 16)
    sample = polys.copy()
    reps, t0 = 0, time.perf_counter()
    while True:
        O.fwd_mul_inv_batch(tables, sample, operand, opq, threads=threads)
        reps += 1
        el = time.perf_counter() - t0
        if el >= budget_s or reps >= 50:
            break
    # single thread on a quarter of the batch
    one = polys[:COUNT // 4].copy()
    t1 = time.perf_counter()
    O.fwd_mul_inv_batch(tables, one, operand, opq, threads=1)
    el1 = time.perf_counter() - t1
    cpu_model = ""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                cpu_model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {
        "value": reps * NTT_PER_STEP / el,
        "unit": "NTT/s",
        "cores": threads,
        "kind": "port",
        "sample": "%d full batches (1024 limb-polys, fwd+Shoup mul+inv) on %d threads in %.1f s; "
                  "single thread: %.0f NTT/s" % (reps, threads, el, (COUNT // 4) * 2 / el1),
        "single_thread_value": (COUNT // 4) * 2 / el1,
        "cpu_model": cpu_model,
        "nproc": os.cpu_count(),
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--plain-steps", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--profile", action="store_true",
                    help="counter runs: only the fused batch kernel (--steps launches) and one 256 MiB "
                         "device-to-device copy as the byte-count calibration")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path (run with --gpus 1)",
              file=sys.stderr)
        return 2

    import torch

    import primus_amd as pa
    from primus_amd import ntt as pn

    rng = np.random.default_rng(SEED)
    moduli, polys, operand = workload(rng)
    table = pa.UintCrtNttTable.new(LOG_N, moduli)
    op = pa.Operand(table, operand)
    from primus_amd import tuning

    shape, shape_ms = tuning.autotune(table, "fwd_mul_inv", COUNT, reps=20)

    dev = torch.device("cuda:0")
    host = torch.from_numpy(polys.view(np.int64))
    bufs = torch.empty((ROTATE, COUNT, N), dtype=torch.int64, device=dev)
    for r in range(ROTATE):
        bufs[r].copy_(host)
    torch.cuda.synchronize()
    stream = torch.cuda.Stream(device=dev)
    sp = ctypes.c_void_p(stream.cuda_stream)
    ptrs = [ctypes.c_void_p(bufs[r].data_ptr()) for r in range(ROTATE)]

    def launcher(tab, operand):
        def step(i):
            pn._check(pa.lib().ph_dcrt_fwd_mul_inv_device(tab.handle, ptrs[i % ROTATE], COUNT, operand.handle, sp))
        return step

    step = launcher(table, op)

    # correctness gate on the first limb-polynomials before any timing
    from oracle import oracle as O

    otabs = [O.Table(LOG_N, q) for q in moduli]
    int_table = pa.UintCrtNttTable.new(LOG_N, moduli, integer_only=True)
    int_op = pa.Operand(int_table, operand)
    int_shape, _ = tuning.autotune(int_table, "fwd_mul_inv", COUNT, reps=10)
    int_step = launcher(int_table, int_op)
    for st in (step, int_step):
        st(0)
        stream.synchronize()
        got = bufs[0, :8].cpu().numpy().view(np.uint64)
        for i in range(8):
            t = otabs[i % LIMBS]
            want = t.inverse_transform(O.pointwise_mul(t.transform(polys[i]), operand[i % LIMBS], t.q))
            if not np.array_equal(got[i], want):
                print("bench: GPU output differs from the oracle; refusing to time", file=sys.stderr)
                return 1
        bufs[0].copy_(host)
        torch.cuda.synchronize()

    def timed(st, K, W):
        for i in range(W):
            st(i)
        stream.synchronize()
        evs = [torch.cuda.Event(enable_timing=True) for _ in range(K + 1)]
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        evs[0].record(stream)
        for i in range(K):
            st(i)
            evs[i + 1].record(stream)
        stream.synchronize()
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        per = np.array([evs[i].elapsed_time(evs[i + 1]) for i in range(K)])  # ms per launch
        return per, evs[0].elapsed_time(evs[K]), wall

    K = args.steps
    if args.profile:
        per, total_ms, _ = timed(step, K, args.warmup)
        src = torch.empty(32 << 20, dtype=torch.int64, device=dev)
        dst = torch.empty_like(src)
        src.fill_(1)
        torch.cuda.synchronize()
        dst.copy_(src)  # calibration: 256 MiB read + 256 MiB written
        torch.cuda.synchronize()
        print(json.dumps({"profile": True, "kernel_ms": float(np.mean(per)), "launches": K + args.warmup,
                          "calibration_copy_bytes": 2 * src.numel() * 8}))
        return 0
    per, total_ms, wall = timed(step, K, args.warmup)
    ms_per_step = total_ms / K
    value = NTT_PER_STEP / (ms_per_step * 1e-3)
    kernel_ms = float(np.mean(per))
    achieved = BYTES_PER_STEP / (kernel_ms * 1e-3) / 1e9
    int_per, int_total, _ = timed(int_step, max(K // 4, 10), args.warmup)
    int_ms = int_total / max(K // 4, 10)

    # plain path: one launch per limb-polynomial, per-limb single-modulus tables
    limb_tabs = [pa.UintCrtNttTable.new(LOG_N, [q]) for q in moduli]
    limb_ops = [pa.Operand(limb_tabs[l], operand[l]) for l in range(LIMBS)]
    L = pa.lib()
    stride = N * 8

    def plain_step(r):
        base = bufs[r % ROTATE].data_ptr()
        for i in range(COUNT):
            l = i % LIMBS
            pn._check(L.ph_dcrt_fwd_mul_inv_device(limb_tabs[l].handle, ctypes.c_void_p(base + i * stride), 1,
                                                   limb_ops[l].handle, sp))

    plain_step(0)
    stream.synchronize()
    pe = [torch.cuda.Event(enable_timing=True) for _ in range(args.plain_steps + 1)]
    pe[0].record(stream)
    for r in range(args.plain_steps):
        plain_step(r)
        pe[r + 1].record(stream)
    stream.synchronize()
    plain_per = np.array([pe[i].elapsed_time(pe[i + 1]) for i in range(args.plain_steps)])
    plain_ms = float(np.mean(plain_per))

    out = {
        "metric": "NTTs/sec + HBM GB/s at N=4096x4 limbs; p50/p99 latency, batched vs plain",
        "value": value,
        "unit": "NTT/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64",
        "data": "synthetic (seeded uniform coefficients and operand)",
        "config": {
            "workload": "batched RNS NTT: 256 polys x 4 limbs, N=4096, fwd NTT + Shoup pointwise + inv NTT, "
                        "one launch per batch",
            "N": N, "limbs": LIMBS, "polys": POLYS, "moduli": moduli, "ntt_per_step": NTT_PER_STEP,
            "rotating_buffers": ROTATE, "parallelism": "single GPU",
        },
        "hbm_gbps": BYTES_PER_STEP * K / (total_ms * 1e-3) / 1e9,
        "p50_ms": float(np.percentile(per, 50)),
        "p99_ms": float(np.percentile(per, 99)),
        "wall_s": wall,
        "plain": {
            "desc": "one launch per limb-polynomial (1024 launches per batch)",
            "ms_per_batch": plain_ms,
            "ntt_per_s": NTT_PER_STEP / (plain_ms * 1e-3),
            "p50_ms": float(np.percentile(plain_per, 50)),
            "p99_ms": float(np.percentile(plain_per, 99)),
        },
        "batched_vs_plain": plain_ms / ms_per_step,
        "arith": table.arith(),
        "autotune": {"coefficients_per_thread": 1 << shape, "best_ms": shape_ms},
        "int_path": {
            "desc": "same fused kernel on 64-bit integer Shoup arithmetic (32-bit multiply chains)",
            "coefficients_per_thread": 1 << int_shape,
            "ms_per_batch": int_ms,
            "ntt_per_s": NTT_PER_STEP / (int_ms * 1e-3),
            "hbm_frac": BYTES_PER_STEP / (int_ms * 1e-3) / 1e9 / HBM_PEAK_GBPS,
        },
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBPS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBPS,
            "traffic": None,
            "kernel": "chunk_kernel<F64Arith,12,8,FWD|MUL_SHOUP|INV|INV_LAST> (fused fwd -> mul -> inv)",
            "kernel_ms": kernel_ms,
            "bytes_per_launch": BYTES_PER_STEP,
        },
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(moduli, polys, operand)
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
