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
kernel, the CPU oracle timed on this host, and secondary lines for the other
batched forms (plain N=1024 products, uneven RLWE encryption, external product).
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
SEED = 0x5052494D5553
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "traffic_fwd_mul_inv_f64_n4096_l4.json")


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return ""


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
        if el >= budget_s or reps >= 400:
            break
    one = polys[:COUNT // 4].copy()
    t1 = time.perf_counter()
    O.fwd_mul_inv_batch(tables, one, operand, opq, threads=1)
    el1 = time.perf_counter() - t1
    return {
        "value": reps * NTT_PER_STEP / el,
        "unit": "NTT/s",
        "cores": threads,
        "kind": "port",
        "sample": "%d full batches (1024 limb-polys of N=4096, fwd + Shoup mul + inv) on %d threads in %.1f s; "
                  "single thread %.0f NTT/s" % (reps, threads, el, (COUNT // 4) * 2 / el1),
        "single_thread_value": (COUNT // 4) * 2 / el1,
        "cpu_model": cpu_model(),
        "nproc": os.cpu_count(),
    }


LATENCY_SAMPLES = 1000  # p50/p99 always come from at least this many single-step spans


class Timer:
    """HIP events on the stream the kernels are launched on (torch's view of it).

    run() times K steps bracketed by two events only (throughput: nothing is
    queued between the launches), then max(K, LATENCY_SAMPLES) steps with an
    event after every step (per-step latency for p50/p99)."""

    def __init__(self, torch, stream):
        self.torch, self.stream = torch, stream

    def _bracket(self, step, K):
        e0, e1 = (self.torch.cuda.Event(enable_timing=True) for _ in range(2))
        self.torch.cuda.synchronize()
        e0.record(self.stream)
        for i in range(K):
            step(i)
        e1.record(self.stream)
        self.stream.synchronize()
        self.torch.cuda.synchronize()
        return e0.elapsed_time(e1)

    def run(self, step, K, W, latency=True, samples=None):
        for i in range(W):
            step(i)
        self.stream.synchronize()
        total = self._bracket(step, K)
        if not latency:
            return None, total
        S = samples if samples is not None else max(K, LATENCY_SAMPLES)
        evs = [self.torch.cuda.Event(enable_timing=True) for _ in range(S + 1)]
        self.torch.cuda.synchronize()
        evs[0].record(self.stream)
        for i in range(S):
            step(i)
            evs[i + 1].record(self.stream)
        self.stream.synchronize()
        self.torch.cuda.synchronize()
        per = np.array([evs[i].elapsed_time(evs[i + 1]) for i in range(S)])
        return per, total

    def run_pipelined(self, steps, streams, K, W):
        """Step i on streams[i % S] (steps[i % S] launches on that stream), so
        up to S independent batches are in flight: a batch's write-back tail
        overlaps the next batch's loads and arithmetic.  The streams fork from
        and join into streams[0] around the timed region; returns ms per step."""
        torch, S = self.torch, len(streams)
        for i in range(W):
            steps[i % S](i)
        torch.cuda.synchronize()
        e0, e1 = (torch.cuda.Event(enable_timing=True) for _ in range(2))
        e0.record(streams[0])
        for s in streams[1:]:
            s.wait_event(e0)
        for i in range(K):
            steps[i % S](i)
        for s in streams[1:]:
            done = torch.cuda.Event()
            done.record(s)
            streams[0].wait_event(done)
        e1.record(streams[0])
        streams[0].synchronize()
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / K

    def run_graph(self, step, per_graph, K, W):
        """Capture per_graph consecutive steps into one HIP graph and replay it
        K // per_graph times; returns ms per step."""
        torch = self.torch
        g = torch.cuda.CUDAGraph()
        self.stream.synchronize()
        with torch.cuda.graph(g, stream=self.stream):
            for i in range(per_graph):
                step(i)
        self.stream.synchronize()
        reps = max(K // per_graph, 1)
        with torch.cuda.stream(self.stream):
            for _ in range(max(W // per_graph, 1)):
                g.replay()
            total = self._bracket(lambda i: g.replay(), reps)
        return total / (reps * per_graph)


def secondary_plain_products(pa, timer, sp):
    """Plain N=1024 products: 64 pairs, q the largest prime < 2^31 with q = 1 mod 2048."""
    from oracle import oracle as O

    q, log_n, pairs = 2147473409, 10, 64
    n = 1 << log_n
    rng = np.random.default_rng(SEED)
    a = rng.integers(0, q, size=(pairs, n), dtype=np.uint64)
    b = rng.integers(0, q, size=(pairs, n), dtype=np.uint64)
    tab = pa.UintCrtNttTable.new(log_n, [q])
    da, db = pa.DeviceBuffer.from_array(a), pa.DeviceBuffer.from_array(b)
    L = pa.lib()
    pa.negacyclic_mul_device(tab, da.ptr, db.ptr, pairs, stream=sp)
    pa.synchronize(sp)
    got = da.download((pairs, n))
    t = O.Table(log_n, q)
    ok = all(np.array_equal(got[i], O.schoolbook(a[i], b[i], q)) for i in range(2))

    def batched(i):
        L.ph_dcrt_negacyclic_mul_device(tab.handle, da.ptr, db.ptr, pairs, sp)

    def plain(i):
        for k in range(pairs):
            L.ph_dcrt_negacyclic_mul_device(tab.handle, ctypes.c_void_p(da.ptr.value + k * n * 8),
                                            ctypes.c_void_p(db.ptr.value + k * n * 8), 1, sp)

    per, tot = timer.run(batched, 1000, 50)
    pper, ptot = timer.run(plain, 20, 2)
    t0 = time.perf_counter()
    reps = 0
    while time.perf_counter() - t0 < 1.0:
        for i in range(pairs):
            t.inverse_transform(O.pointwise_mul(t.transform(a[i]), t.transform(b[i]), q))
        reps += 1
    cpu_s = (time.perf_counter() - t0) / reps
    return {"workload": "64 pairs N=1024, q=2147473409: fwd a, fwd b, pointwise, inv (one launch)",
            "parity_vs_schoolbook": ok, "ms_per_batch": tot / 1000, "p50_ms": float(np.percentile(per, 50)),
            "p99_ms": float(np.percentile(per, 99)), "ntt_per_s": 3 * pairs / (tot / 1000 * 1e-3),
            "plain_ms_per_batch": ptot / 20, "plain_p99_ms": float(np.percentile(pper, 99)),
            "batched_vs_plain": (ptot / 20) / (tot / 1000),
            "cpu_single_thread_ms_per_batch": cpu_s * 1e3, "arith": tab.arith()}


def secondary_rlwe(pa, timer, sp):
    """Uneven RLWE encryption: 1024 requests, N ~ U{2^10..2^15}, L ~ U{1..12}, 59-bit primes."""
    from primus_amd.fhe import RlweContext, rlwe_layout

    rng = np.random.default_rng(SEED + 3)
    shapes = [(int(rng.integers(10, 16)), int(rng.integers(1, 13))) for _ in range(1024)]
    grid = {ln: pa.ntt_primes_below(59, 2 << ln, 12) for ln in range(10, 16)}
    ctx = RlweContext(10, 15, 12, grid)
    reqs, a_words, small = rlwe_layout(shapes)
    a = np.empty(a_words, dtype=np.uint64)
    for (ln, L, ao, so) in reqs:
        for i in range(L):
            a[ao + (i << ln): ao + ((i + 1) << ln)] = rng.integers(0, grid[ln][i], size=1 << ln, dtype=np.uint64)
    s = rng.integers(-1, 2, size=small).astype(np.int8)
    e = np.clip(np.rint(rng.normal(0, 3.2, size=small)), -19, 19).astype(np.int8)
    m = rng.integers(0, 256, size=small).astype(np.uint8)

    def as_words(v):
        raw = np.ascontiguousarray(v).tobytes()
        return np.frombuffer(raw.ljust((len(raw) + 7) // 8 * 8, b"\0"), dtype=np.uint64)

    da, dout = pa.DeviceBuffer.from_array(a), pa.DeviceBuffer(a_words)
    ds, de, dm = (pa.DeviceBuffer.from_array(as_words(v)) for v in (s, e, m))

    # the batch layout is prepared once (work items, scratch); each step only
    # enqueues the persistent kernel, whose grid the plan's autotuner picks
    plan = ctx.plan(reqs)
    grid, _ = plan.autotune(da.ptr, ds.ptr, de.ptr, dm.ptr, dout.ptr, reps=3)

    def step(i):
        plan.encrypt_device(da.ptr, ds.ptr, de.ptr, dm.ptr, dout.ptr, stream=sp)

    per, tot = timer.run(step, 100, 5)
    ms = tot / 100
    try:
        graph_ms = timer.run_graph(step, 4, 100, 4)
    except Exception:  # capture unsupported: stream launches only
        graph_ms = None
    best = min(ms, graph_ms) if graph_ms else ms
    ntts = 2 * sum(L for _, L in shapes)
    # the same batch with its inputs drawn on the device first (Philox sampler,
    # one more launch from the same plan): sampling alone, and sampling plus
    # encryption per batch
    def sample(i):
        plan.sample_device(0x5052494D5553 + i, da.ptr, ds.ptr, de.ptr, dm.ptr, stream=sp)

    def sample_encrypt(i):
        sample(i)
        step(i)

    _, stot = timer.run(sample, 100, 5, latency=False)
    sper, setot = timer.run(sample_encrypt, 100, 5)
    return {"workload": "1024 RLWE encryptions, N~U{2^10..2^15}, L~U{1..12}, one persistent launch",
            "ms_per_batch": best, "stream_launch_ms": ms, "graph_replay_ms": graph_ms, "persistent_grid": grid,
            "p50_ms": float(np.percentile(per, 50)), "p99_ms": float(np.percentile(per, 99)),
            "ntt_per_s": ntts / (best * 1e-3), "coeff_limbs": int(a_words),
            "algorithmic_gbps": (16 * a_words + 3 * small) / (best * 1e-3) / 1e9,
            "device_sampling": {
                "desc": "s, e, m and a_hat drawn on the device (Philox4x32-10), then the encryption",
                "sample_ms_per_batch": stot / 100,
                "sample_gbps_written": (8 * a_words + 3 * small) / (stot / 100 * 1e-3) / 1e9,
                "sample_and_encrypt_ms_per_batch": setot / 100,
                "p50_ms": float(np.percentile(sper, 50)), "p99_ms": float(np.percentile(sper, 99))}}


def secondary_ext_product(pa, timer, sp):
    """External product: 2048 products, N=1024, q=4294957057, base 2^8, l=3."""
    from primus_amd.fhe import ExtKey

    q, log_n, count = 4294957057, 10, 2048
    n = 1 << log_n
    rng = np.random.default_rng(SEED + 4)
    key = ExtKey(pa.UintNttTable.new(log_n, q), 8, 3, rng.integers(0, q, size=(2, 3, 2, n), dtype=np.uint64))
    x = rng.integers(0, q, size=(count, 2, n), dtype=np.uint64)
    # 6 rotating input/output pairs (384 MiB): the batch is streamed from HBM,
    # not served from the 256 MiB Infinity Cache
    rot = 6
    pairs = [(pa.DeviceBuffer.from_array(x), pa.DeviceBuffer(count * 2 * n)) for _ in range(rot)]

    def step(i):
        din, dout = pairs[i % rot]
        key.product_device(din.ptr, dout.ptr, count, stream=sp)

    per, tot = timer.run(step, 1000, 20)
    ms = tot / 1000
    try:
        graph_ms = timer.run_graph(step, rot, 1000, 20)
    except Exception:  # capture unsupported: stream launches only
        graph_ms = None
    best = min(ms, graph_ms) if graph_ms else ms
    return {"workload": "2048 RLWE x RGSW external products, N=1024, q=4294957057, B=2^8, l=3, one launch",
            "ms_per_batch": best, "stream_launch_ms": ms, "graph_replay_ms": graph_ms,
            "p50_ms": float(np.percentile(per, 50)), "p99_ms": float(np.percentile(per, 99)),
            "ntt_per_s": 8 * count / (best * 1e-3), "algorithmic_gbps": 4 * n * 8 * count / (best * 1e-3) / 1e9,
            "rotating_buffers": rot, "arith": "f64"}


def secondary_ext_product_rns(pa, timer, sp):
    """External product over 2 limbs (Q of 118 bits for 59-bit primes, 100 bits
    for 50-bit ones): 1024 products, N=1024, base 2^8, l=3 (the top 24 bits of
    Q), one fused launch per batch (composition, digits, digit transforms,
    multiply-accumulate with the key, inverse transforms)."""
    from primus_amd.fhe import ExtKeyRns

    log_n, L, l, count = 10, 2, 3, 1024
    n = 1 << log_n
    out = {}
    for bits in (59, 50):
        moduli = pa.ntt_primes_below(bits, 2 * n, L)
        table = pa.UintCrtNttTable.new(log_n, moduli)
        rng = np.random.default_rng(SEED + 8)
        qcol = np.array(moduli, dtype=np.uint64).reshape(-1, 1)
        key = ExtKeyRns(table, 8, l, np.stack([rng.integers(0, 1 << 63, size=(L, n), dtype=np.uint64) % qcol
                                               for _ in range(4 * l)]).reshape(2, l, 2, L, n))
        x = np.stack([rng.integers(0, 1 << 63, size=(L, n), dtype=np.uint64) % qcol for _ in range(2 * count)])
        rot = 4  # 4 x 64 MiB in + out
        pairs = [(pa.DeviceBuffer.from_array(x), pa.DeviceBuffer(count * 2 * L * n)) for _ in range(rot)]

        def step(i, pairs=pairs, key=key):
            din, dout = pairs[i % rot]
            key.product_device(din.ptr, dout.ptr, count, stream=sp)

        per, tot = timer.run(step, 200, 10)
        ms = tot / 200
        try:  # the fused launch allocates nothing: replayed from a HIP graph of rot batches
            graph_ms = timer.run_graph(step, rot, 200, 8)
        except Exception:  # capture unsupported: stream launches only
            graph_ms = None
        best = min(ms, graph_ms) if graph_ms else ms
        out["%d_bit" % bits] = {
            "moduli": moduli, "arith": table.arith() if bits == 50 else "integer (approximate-quotient Shoup)",
            "ms_per_batch": best, "stream_launch_ms": ms, "graph_replay_ms": graph_ms,
            "p50_ms": float(np.percentile(per, 50)), "p99_ms": float(np.percentile(per, 99)),
            "ntt_per_s": (2 * l + 2) * L * count / (best * 1e-3),
            "algorithmic_gbps": 4 * L * n * 8 * count / (best * 1e-3) / 1e9,
            "hbm_frac": 4 * L * n * 8 * count / (best * 1e-3) / 1e9 / HBM_PEAK_GBPS, "rotating_buffers": rot}
    out["workload"] = ("1024 external products over 2 limbs, N=1024, B=2^8, l=3: one fused launch per batch "
                       "(composition, digits, digit transforms, multiply-accumulate, inverse)")
    return out


def secondary_automorphism(pa, timer, sp):
    """NTT-domain automorphism x -> x^5 on the headline batch shape (1024 limb-polys, N=4096)."""
    moduli = pa.ntt_primes_below(50, 2 * N, LIMBS)
    tab = pa.UintCrtNttTable.new(LOG_N, moduli)
    rng = np.random.default_rng(SEED + 7)
    x = np.stack([rng.integers(0, moduli[i % LIMBS], size=N, dtype=np.uint64) for i in range(COUNT)])
    rot = 6  # 6 x 64 MiB in + out: streamed from HBM
    pairs = [(pa.DeviceBuffer.from_array(x), pa.DeviceBuffer(COUNT * N)) for _ in range(rot)]
    L = pa.lib()

    def step(i):
        din, dout = pairs[i % rot]
        L.ph_dcrt_automorphism_device(tab.handle, din.ptr, dout.ptr, COUNT, 5, sp)

    per, tot = timer.run(step, 1000, 20)
    ms = tot / 1000
    try:  # the launch allocates nothing: also replayed from a HIP graph of rot launches
        graph_ms = timer.run_graph(step, rot, 1000, 20)
    except Exception:  # capture unsupported: stream launches only
        graph_ms = None
    best = min(ms, graph_ms) if graph_ms else ms
    return {"workload": "NTT-domain automorphism x -> x^5, 1024 limb-polys of N=4096, out of place, one launch",
            "ms_per_batch": best, "stream_launch_ms": ms, "graph_replay_ms": graph_ms,
            "p50_ms": float(np.percentile(per, 50)), "p99_ms": float(np.percentile(per, 99)),
            "algorithmic_gbps": 16 * COUNT * N / (best * 1e-3) / 1e9,
            "hbm_frac": 16 * COUNT * N / (best * 1e-3) / 1e9 / HBM_PEAK_GBPS, "rotating_buffers": rot}


def secondary_base_conversion(pa, timer, sp):
    """RNS base conversion of 256 CRT polynomials of N=4096: fast 4 -> 4 limbs
    (the headline's 50-bit moduli to four 59-bit ones) and exact 4 -> 1."""
    from primus_amd.rns import BaseConverter

    polys = POLYS
    moduli = pa.ntt_primes_below(50, 2 * N, LIMBS)
    obase = pa.ntt_primes_below(59, 2 * N, LIMBS)
    rng = np.random.default_rng(SEED + 8)
    x = np.stack([np.stack([rng.integers(0, q, size=N, dtype=np.uint64) for q in moduli]) for _ in range(polys)])
    rot = 4  # 4 x (32 + 32) MiB: streamed from HBM
    out = {}
    for name, ob, exact in (("fast_4_to_4", obase, False), ("exact_4_to_1", obase[:1], True)):
        conv = BaseConverter(moduli, ob)
        pairs = [(pa.DeviceBuffer.from_array(x), pa.DeviceBuffer(polys * len(ob) * N)) for _ in range(rot)]
        fn = conv.exact_convert_device if exact else conv.fast_convert_device

        def step(i, pairs=pairs, fn=fn):
            din, dout = pairs[i % rot]
            fn(din.ptr, dout.ptr, LOG_N, polys, stream=sp)

        per, tot = timer.run(step, 500, 20)
        ms = tot / 500
        bytes_ = polys * N * 8 * (LIMBS + len(ob))
        out[name] = {"ms_per_batch": ms, "p50_ms": float(np.percentile(per, 50)),
                     "p99_ms": float(np.percentile(per, 99)), "algorithmic_gbps": bytes_ / (ms * 1e-3) / 1e9,
                     "hbm_frac": bytes_ / (ms * 1e-3) / 1e9 / HBM_PEAK_GBPS}
    out["workload"] = "256 CRT polynomials, N=4096, 4 input limbs (50-bit); one launch per conversion"
    return out


SECONDARY = {
    "plain_n1024": secondary_plain_products,
    "rlwe_uneven": secondary_rlwe,
    "ext_product": secondary_ext_product,
    "ext_product_rns": secondary_ext_product_rns,
    "automorphism": secondary_automorphism,
    "base_conversion": secondary_base_conversion,
}


def kernel_signature(shape):
    """The fused kernel a tuning shape launches, as rocprofv3 names it (without 'void ')."""
    from primus_amd import tuning

    per_thread, streaming = tuning.describe(shape)
    return "ph::%s<ph::F64Arith, %d, %d, 39>" % ("stream_kernel" if streaming else "chunk_kernel", LOG_N,
                                                 LOG_N - (per_thread.bit_length() - 1))


def serving_pool(pa, grid, ext_q, pool, seed):
    """Distinct requests of the serving mix: 50% negacyclic products, 30% RLWE
    encryptions (N ~ U{2^10..2^16}, L ~ U{1..16}, 59-bit primes) and 20%
    external products (N = 1024, q = 4294957057)."""
    from primus_amd import serving as sv

    rng = np.random.default_rng(seed)
    reqs = []
    for i in range(pool):
        u = rng.random()
        if u < 0.8:
            ln, L = int(rng.integers(10, 17)), int(rng.integers(1, 17))
            n = 1 << ln
            a = np.concatenate([rng.integers(0, grid[ln][j], size=n, dtype=np.uint64) for j in range(L)])
            if u < 0.5:
                b = np.concatenate([rng.integers(0, grid[ln][j], size=n, dtype=np.uint64) for j in range(L)])
                reqs.append(sv.Request(sv.POLY_MUL, ln, L, in0=a, in1=b))
            else:
                s = rng.integers(-1, 2, size=n).astype(np.int8)
                e = np.clip(np.rint(rng.normal(0, 3.2, size=n)), -19, 19).astype(np.int8)
                m = rng.integers(0, 256, size=n).astype(np.uint8)
                reqs.append(sv.Request(sv.RLWE_ENCRYPT, ln, L, in0=a, s=s, e=e, m=m))
        else:
            reqs.append(sv.Request(sv.EXT_PRODUCT, 10, 1, in0=rng.integers(0, ext_q, size=2048, dtype=np.uint64)))
    return reqs


def request_bytes(r):
    """(host->device, device->host) bytes of one serving request."""
    up = sum(a.nbytes for a in (r.in0, r.in1, r.s, r.e, r.m) if a is not None)
    return up, r.out.nbytes


def serving_main(args):
    """Serving stream: Poisson arrivals at the offered rates, continuous batching
    against the plain path (one request at a time, one launch per
    limb-polynomial), both through the same request loop and HBM budget."""
    import primus_amd as pa
    from oracle import oracle as O
    from primus_amd import serving as sv
    from primus_amd.fhe import ExtKey

    ext_q = 4294957057
    grid = {ln: pa.ntt_primes_below(59, 2 << ln, 16) for ln in range(10, 17)}
    rng = np.random.default_rng(SEED + 5)
    key_np = rng.integers(0, ext_q, size=(2, 3, 2, 1024), dtype=np.uint64)
    key = ExtKey(pa.UintNttTable.new(10, ext_q), 8, 3, key_np)
    pool = serving_pool(pa, grid, ext_q, args.serving_pool, SEED + 6)
    if not args.serving_pageable:
        pool = [r.pinned_copy() for r in pool]  # page-locked: no staging copies in the server
    budget = 64 << 30

    def run(srv, rate, count, seed):
        reqs = [pool[i % len(pool)] for i in range(count)]
        before = srv.stats()
        lat, elapsed = srv.run_poisson(reqs, rate, seed=seed)
        after = srv.stats()
        return {"offered_per_s": rate, "requests": count, "elapsed_s": elapsed,
                "throughput_per_s": count / elapsed, "p50_ms": float(np.percentile(lat, 50) * 1e3),
                "p99_ms": float(np.percentile(lat, 99) * 1e3), "batches": after["batches"] - before["batches"],
                "launches": after["launches"] - before["launches"],
                "peak_inflight_bytes": after["peak_inflight_bytes"],
                "host_s": {k: after[k] - before[k] for k in ("submit_s", "dispatch_s", "slot_wait_s", "device_wait_s")}}

    batched = sv.Server(grid, 10, 16, 16, ext_key=key, hbm_budget_bytes=budget, max_batch=args.serving_max_batch,
                        streams=args.serving_streams)
    plain = sv.Server(grid, 10, 16, 16, ext_key=key, hbm_budget_bytes=budget, max_batch=1, streams=1,
                      per_poly_launches=True)
    # parity gate: the first request of every (kind, ring degree) in the pool
    # against the oracle, through both servers
    gate = {}
    for r in pool:
        gate.setdefault((r.kind, r.log_n), r)
    for (kind, ln), r in sorted(gate.items()):
        n = 1 << ln
        if kind == sv.POLY_MUL:
            want = []
            for j in range(r.limbs):
                a, b, q = r.in0[j * n:(j + 1) * n], r.in1[j * n:(j + 1) * n], grid[ln][j]
                if ln <= 11:
                    want.append(O.schoolbook(a, b, q))
                else:  # the oracle's transforms, themselves pinned to schoolbook products
                    t = O.Table(ln, q)
                    want.append(t.inverse_transform(O.pointwise_mul(t.transform(a), t.transform(b), q)))
            want = np.concatenate(want)
        elif kind == sv.RLWE_ENCRYPT:
            mods = grid[ln][:r.limbs]
            want = np.concatenate([O.rlwe_encrypt(O.Table(ln, q), r.s, r.e, r.m, O.rlwe_delta(mods, q),
                                                  r.in0[j * n:(j + 1) * n]) for j, q in enumerate(mods)])
        else:
            want = O.ext_product(O.Table(10, ext_q), O.Basis.new(ext_q, 8, 3), key_np,
                                 r.in0.reshape(2, 1024)).ravel()
        for srv in (batched, plain):
            r.out[:] = 0
            srv.wait(srv.submit(r))
            if not np.array_equal(r.out, want):
                print("bench: serving output (kind %d, N=2^%d, %d limbs) differs from the oracle; refusing to time"
                      % (kind, ln, r.limbs), file=sys.stderr)
                return 1
    gate_checked = ["%s N=2^%d L=%d" % (("poly_mul", "rlwe_encrypt", "ext_product")[k], ln, r.limbs)
                    for (k, ln), r in sorted(gate.items())]

    # the host link both paths share: page-locked 256 MiB copies each way
    link = {}
    hbuf = sv.PinnedBuffer((32 << 20,), np.uint64)
    dbuf = pa.DeviceBuffer(32 << 20)
    Lh = pa.lib()
    for name, fn, a, b in (("h2d_gbps", Lh.ph_memcpy_htod, dbuf.ptr, hbuf.array.ctypes.data_as(ctypes.c_void_p)),
                           ("d2h_gbps", Lh.ph_memcpy_dtoh, hbuf.array.ctypes.data_as(ctypes.c_void_p), dbuf.ptr)):
        fn(a, b, 256 << 20)
        t0 = time.perf_counter()
        for _ in range(4):
            fn(a, b, 256 << 20)
        link[name] = 4 * (256 << 20) / (time.perf_counter() - t0) / 1e9
    del hbuf, dbuf

    out = {"metric": "serving stream: p50/p99 latency and throughput, continuous batching vs plain per-request launches",
           "workload": "Poisson arrivals; mix 50% negacyclic product / 30% RLWE encryption (N~U{2^10..2^16}, "
                       "L~U{1..16}, 59-bit primes) / 20% external product (N=1024, q=4294957057, B=2^8, l=3); "
                       "host buffers in and out (PCIe included; " +
                       ("pageable, staged by the server" if args.serving_pageable else "page-locked, copied directly") +
                       "); 64 GiB HBM admission budget",
           "pool_distinct_requests": len(pool), "parity_checked": gate_checked,
           "max_batch": args.serving_max_batch,
           "stream_slots": args.serving_streams, "host_link": link,
           "mean_request_bytes": {"h2d": float(np.mean([request_bytes(r)[0] for r in pool])),
                                  "d2h": float(np.mean([request_bytes(r)[1] for r in pool]))},
           "rates": []}
    for k, rate in enumerate(args.serving_rates):
        count = int(min(args.serving_requests, rate * args.serving_seconds))
        b = run(batched, rate, count, seed=11 + k)
        p = run(plain, rate, int(min(count, args.serving_plain_requests)), seed=11 + k)
        out["rates"].append({"batched": b, "plain": p, "p99_ratio_plain_over_batched": p["p99_ms"] / b["p99_ms"]})
        print("bench: serving rate %g req/s done" % rate, file=sys.stderr, flush=True)
    # saturation: everything offered at once, the highest throughput each path reaches
    count = int(min(args.serving_requests, args.serving_saturation_requests))
    b = run(batched, 1e7, count, seed=99)
    p = run(plain, 1e7, int(min(count, args.serving_plain_requests)), seed=99)
    out["saturation"] = {"batched": b, "plain": p,
                         "throughput_ratio": b["throughput_per_s"] / p["throughput_per_s"]}
    batched.close()
    plain.close()
    print(json.dumps(out))
    return 0


def serving_leg():
    """`bench.py --serving --serving-leg` in a child process; its JSON line, or
    the failure."""
    import subprocess

    cmd = [sys.executable, os.path.abspath(__file__), "--serving-leg"]
    print("bench: serving pass starts", file=sys.stderr, flush=True)
    try:
        res = subprocess.run(cmd, capture_output=True, text=True, timeout=400)
    except subprocess.TimeoutExpired:
        return {"error": "serving pass exceeded 400 s"}
    sys.stderr.write(res.stderr[-2000:])
    lines = [ln for ln in res.stdout.splitlines() if ln.startswith("{")]
    if res.returncode != 0 or not lines:
        return {"error": "serving pass exited with status %d" % res.returncode, "stderr": res.stderr[-500:]}
    leg = json.loads(lines[-1])
    leg["command"] = "python bench.py --serving-leg"
    return leg


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--plain-steps", type=int, default=20)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-secondary", action="store_true")
    ap.add_argument("--profile", action="store_true",
                    help="counter runs: only the fused batch kernel (--steps launches) and one 256 MiB "
                         "device-to-device copy as the byte-count calibration")
    ap.add_argument("--profile-int", action="store_true",
                    help="with --profile: the same batch on the integer-arithmetic table (IntCanonArith kernel)")
    ap.add_argument("--shape", type=int, default=None,
                    help="launch shape of the fused kernel instead of the autotuner's choice "
                         "(3/4: 8/16 coefficients per thread; +16: streaming)")
    ap.add_argument("--serving", action="store_true",
                    help="the serving stream benchmark (Poisson arrivals, batched vs plain) instead of the headline")
    ap.add_argument("--serving-rates", type=float, nargs="+", default=[1000.0, 10000.0, 50000.0])
    ap.add_argument("--serving-requests", type=int, default=100000)
    ap.add_argument("--serving-seconds", type=float, default=20.0,
                    help="cap on offered time per rate: requests = min(--serving-requests, rate * this)")
    ap.add_argument("--serving-plain-requests", type=int, default=3000)
    ap.add_argument("--serving-pool", type=int, default=256)
    ap.add_argument("--serving-max-batch", type=int, default=None,
                    help="requests per batch of the batched server (default 512 with page-locked request "
                         "buffers, 256 with --serving-pageable)")
    ap.add_argument("--serving-streams", type=int, default=None,
                    help="stream slots of the batched server, batches in flight (default 6 with page-locked "
                         "request buffers, 4 with --serving-pageable, whose staging copies run on the host)")
    ap.add_argument("--serving-pageable", action="store_true",
                    help="requests in ordinary (pageable) host memory, staged by the server")
    ap.add_argument("--serving-saturation-requests", type=int, default=20000)
    ap.add_argument("--serving-leg", action="store_true",
                    help="the bounded serving pass of the default run: 1k and 10k req/s for 2 s each and "
                         "saturation, batched and plain on equal request counts (2000, 20000, 20000), a pool "
                         "of 96 requests")
    ap.add_argument("--no-serving", action="store_true", help="skip the serving pass of the default run")
    ap.add_argument("--secondary-only", choices=sorted(SECONDARY), default=None,
                    help="run one secondary workload alone and print its JSON (for kernel profiles)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path (run with --gpus 1)",
              file=sys.stderr)
        return 2

    if args.serving_leg:
        args.serving = True
        args.serving_rates = [1000.0, 10000.0]
        args.serving_seconds = 2.0  # 2000 requests at 1k req/s, 20000 at 10k
        args.serving_requests = 20000
        args.serving_plain_requests = 20000
        args.serving_saturation_requests = 20000
        args.serving_pool = 96
    if args.serving_max_batch is None:
        args.serving_max_batch = 256 if args.serving_pageable else 512
    if args.serving_streams is None:
        args.serving_streams = 4 if args.serving_pageable else 6
    if args.serving:
        # one hardware queue per server stream (slots, input copies, result
        # copies): HIP's default of 4 makes streams share queues, and a copy
        # then waits behind another stream's event wait.  Read when HIP starts,
        # which has not happened yet in this process.
        os.environ.setdefault("GPU_MAX_HW_QUEUES", str(min(32, max(8, args.serving_streams + 4))))
        return serving_main(args)

    # the serving pass runs first, in a child process of its own: it needs a
    # hardware queue per server stream, which HIP reads when it starts, and
    # this process has not touched the GPU yet
    serving = None
    if not (args.no_serving or args.no_secondary or args.profile or args.secondary_only):
        serving = serving_leg()

    # one hardware queue per stream of the pipelined pass (up to 6 streams),
    # read when HIP starts, which has not happened yet in this process
    os.environ.setdefault("GPU_MAX_HW_QUEUES", "8")
    import torch

    import primus_amd as pa
    from oracle import oracle as O
    from primus_amd import tuning

    rng = np.random.default_rng(SEED + 1)
    moduli = pa.ntt_primes_below(50, 2 * N, LIMBS)
    polys = np.stack([rng.integers(0, moduli[i % LIMBS], size=N, dtype=np.uint64) for i in range(COUNT)])
    operand = np.stack([rng.integers(0, q, size=N, dtype=np.uint64) for q in moduli])
    table = pa.UintCrtNttTable.new(LOG_N, moduli)
    op = pa.Operand(table, operand)

    dev = torch.device("cuda:0")
    host = torch.from_numpy(polys.view(np.int64))
    bufs = torch.empty((ROTATE, COUNT, N), dtype=torch.int64, device=dev)
    for r in range(ROTATE):
        bufs[r].copy_(host)
    torch.cuda.synchronize()
    stream = torch.cuda.Stream(device=dev)
    sp = ctypes.c_void_p(stream.cuda_stream)
    ptrs = [ctypes.c_void_p(bufs[r].data_ptr()) for r in range(ROTATE)]
    timer = Timer(torch, stream)
    L = pa.lib()
    if args.secondary_only:
        # one secondary workload alone (counter profiles of its kernels), then
        # the 256 MiB device-to-device copy that calibrates the byte counters
        res = SECONDARY[args.secondary_only](pa, timer, sp)
        src = torch.empty(32 << 20, dtype=torch.int64, device=dev)
        dst = torch.empty_like(src)
        src.fill_(1)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({args.secondary_only: res, "calibration_copy_bytes": src.numel() * 8}))
        return 0

    def launcher(tab, operand_, sp_=sp):
        def step(i):
            L.ph_dcrt_fwd_mul_inv_device(tab.handle, ptrs[i % ROTATE], COUNT, operand_.handle, sp_)
        return step

    if args.shape is None:
        shape, shape_ms = tuning.autotune(table, "fwd_mul_inv", COUNT, reps=20)
    else:  # a fixed launch shape (counter profiles of one kernel)
        shape, shape_ms = args.shape, None
        tuning.set_shape(LOG_N, "fwd_mul_inv", table.arith(), shape)
    step = launcher(table, op)
    int_table = pa.UintCrtNttTable.new(LOG_N, moduli, integer_only=True)
    int_op = pa.Operand(int_table, operand)
    int_shape, _ = tuning.autotune(int_table, "fwd_mul_inv", COUNT, reps=10)
    int_step = launcher(int_table, int_op)

    # correctness gate on the first limb-polynomials before any timing
    otabs = [O.Table(LOG_N, q) for q in moduli]
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

    K = args.steps
    if args.profile:
        if args.profile_int:
            step = int_step
        _, total = timer.run(step, K, args.warmup, latency=False)
        src = torch.empty(32 << 20, dtype=torch.int64, device=dev)
        dst = torch.empty_like(src)
        src.fill_(1)
        torch.cuda.synchronize()
        dst.copy_(src)  # calibration: 256 MiB read + 256 MiB written
        torch.cuda.synchronize()
        print(json.dumps({"profile": True, "kernel": kernel_signature(shape) if not args.profile_int
                          else "chunk_kernel<ph::IntCanonArith, ...> (%d per thread)" % tuning.describe(int_shape)[0],
                          "kernel_ms": total / K,
                          "launches": K + args.warmup, "calibration_copy_bytes": src.numel() * 8}))
        return 0

    # pipelined: consecutive batches on 2, 3, 4 and 6 streams (ROTATE is a
    # multiple of each, so a buffer is only ever reused on the stream that last
    # wrote it)
    pipelined = {}
    extra = [torch.cuda.Stream(device=dev) for _ in range(5)]
    for S in (2, 3, 4, 6):
        streams = [stream] + extra[:S - 1]
        steps = [launcher(table, op, ctypes.c_void_p(s.cuda_stream)) for s in streams]
        pipelined[S] = timer.run_pipelined(steps, streams, K, args.warmup)
    # the same K batches as replays of one captured HIP graph of ROTATE launches
    try:
        graph_ms, graph_err = timer.run_graph(step, ROTATE, K, args.warmup), None
    except Exception as exc:  # capture unsupported: report it, time stream launches only
        graph_ms, graph_err = None, str(exc)[:200]
    # stream launches, timed last: the last K dispatches of a kernel trace of
    # this command are the latency pass, one batch at a time
    t_wall = time.perf_counter()
    per, total_ms = timer.run(step, K, args.warmup)
    wall = time.perf_counter() - t_wall
    stream_ms = total_ms / K
    serial = "hipGraph" if graph_ms is not None and graph_ms < stream_ms else "stream"
    serial_ms = graph_ms if serial == "hipGraph" else stream_ms
    best_s = min(pipelined, key=pipelined.get)
    if pipelined[best_s] < serial_ms:
        launch, ms_per_step = "%d streams" % best_s, pipelined[best_s]
    else:
        launch, ms_per_step = serial, serial_ms
    value = NTT_PER_STEP / (ms_per_step * 1e-3)
    # the kernel's own duration: one launch at a time, nothing else in the timed region
    kernel_ms = serial_ms
    achieved = BYTES_PER_STEP / (kernel_ms * 1e-3) / 1e9
    int_per, int_total = timer.run(int_step, max(K // 4, 10), args.warmup, latency=False)
    int_ms = int_total / max(K // 4, 10)

    # plain path: one launch per limb-polynomial, per-limb single-modulus tables
    limb_tabs = [pa.UintCrtNttTable.new(LOG_N, [q]) for q in moduli]
    limb_ops = [pa.Operand(limb_tabs[l], operand[l]) for l in range(LIMBS)]
    stride = N * 8

    def plain_step(r):
        base = bufs[r % ROTATE].data_ptr()
        for i in range(COUNT):
            l = i % LIMBS
            L.ph_dcrt_fwd_mul_inv_device(limb_tabs[l].handle, ctypes.c_void_p(base + i * stride), 1,
                                         limb_ops[l].handle, sp)

    # 1024 launches per step (about 11 ms): its latency pass is bounded
    plain_per, plain_total = timer.run(plain_step, args.plain_steps, 1, samples=max(args.plain_steps, 100))
    plain_ms = plain_total / args.plain_steps

    # measured HBM bytes per launch (rocprofv3 FETCH_SIZE / WRITE_SIZE, calibrated),
    # used only when they were taken on the kernel this run selected
    kernel_sig = kernel_signature(shape)
    traffic, traffic_note, prof = None, "no counter profile of %s" % kernel_sig, None
    if os.path.exists(TRAFFIC_FILE):
        with open(TRAFFIC_FILE) as f:
            prof = json.load(f).get(kernel_sig)
        if prof:
            traffic, traffic_note = prof.get("traffic_bytes"), "counter profile of %s" % kernel_sig
    # VALU issue time of one launch: SQ_INSTS_VALU of the counter profile, one
    # wave instruction per 4 cycles on each of the 1024 SIMDs, at the engine
    # clock the clock-probe variant measured inside the same kernel (the chip
    # runs this binary64-dense kernel below its 2.4 GHz peak)
    valu = {"valu_frac": None, "valu_note": "no SQ_INSTS_VALU / clock profile of %s" % kernel_sig}
    if prof and prof.get("valu_insts"):
        insts = prof["valu_insts"]
        peak_us = insts * 4 / (1024 * 2.4e9) * 1e6
        valu = {"valu_insts": insts, "valu_floor_us_peak_clock": peak_us,
                "valu_frac_peak_clock": peak_us / (kernel_ms * 1e3)}
        if prof.get("sclk_ghz"):
            at_clk = insts * 4 / (1024 * prof["sclk_ghz"] * 1e9) * 1e6
            valu.update({"sclk_ghz": prof["sclk_ghz"], "valu_issue_us": at_clk,
                         "valu_frac": at_clk / (kernel_ms * 1e3),
                         "valu_note": "valu_frac: SIMD cycles issuing VALU / launch cycles, at the measured "
                                      "engine clock sclk_ghz; valu_frac_peak_clock prices the same "
                                      "instructions at 2.4 GHz"})
        else:
            valu["valu_frac"] = valu["valu_frac_peak_clock"]
            valu["valu_note"] = "valu_frac priced at the 2.4 GHz peak clock (no clock-probe profile)"

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
            "rotating_buffers": ROTATE, "parallelism": "single GPU", "launch": launch,
        },
        "hbm_gbps": BYTES_PER_STEP / (ms_per_step * 1e-3) / 1e9,
        "stream_launch_ms": stream_ms,
        "graph_replay_ms": graph_ms,
        "graph_error": graph_err,
        "pipelined_ms": {"%d streams" % k: v for k, v in pipelined.items()},
        "launch_note": "ms_per_step: K batches back to back; with 'N streams' batch i runs on stream i mod N, so "
                       "up to N independent batches overlap (one launch each). roofline.kernel_ms is the kernel "
                       "alone: one launch at a time (%s)" % serial,
        "p50_ms": float(np.percentile(per, 50)),
        "p99_ms": float(np.percentile(per, 99)),
        "latency_samples": int(len(per)),
        "latency_note": "p50/p99 of max(K, %d) single-batch spans between per-step HIP events on the launch "
                        "stream" % LATENCY_SAMPLES,
        "one_launch_ntt_per_s": NTT_PER_STEP / (kernel_ms * 1e-3),
        "one_launch_note": "NTT/s of one batch at a time (roofline.kernel_ms); value may be the pipelined rate",
        "wall_s": wall,
        "plain": {
            "desc": "one launch per limb-polynomial (1024 launches per batch)",
            "ms_per_batch": plain_ms,
            "ntt_per_s": NTT_PER_STEP / (plain_ms * 1e-3),
            "p50_ms": float(np.percentile(plain_per, 50)),
            "p99_ms": float(np.percentile(plain_per, 99)),
            "latency_samples": int(len(plain_per)),
        },
        "batched_vs_plain": plain_ms / ms_per_step,
        "arith": table.arith(),
        "autotune": dict(tuning.fields(shape), shape=shape, best_ms=shape_ms,
                         candidates=len(tuning.shapes(LOG_N, "fwd_mul_inv"))),
        "int_path": {
            "desc": "same fused kernel on 64-bit integer Shoup arithmetic (32-bit multiply chains)",
            "coefficients_per_thread": tuning.describe(int_shape)[0],
            "streaming": tuning.describe(int_shape)[1],
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
            "traffic": traffic,
            "traffic_note": traffic_note,
            "kernel": kernel_sig + " (MODE 39: fused forward -> Shoup multiply -> inverse)",
            "kernel_ms": kernel_ms,
            "bytes_per_launch": BYTES_PER_STEP,
            **valu,
        },
    }
    if not args.no_secondary:
        out["secondary"] = {name: fn(pa, timer, sp) for name, fn in SECONDARY.items()}
        if serving is not None:
            out["secondary"]["serving"] = serving
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(moduli, polys, operand)
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
