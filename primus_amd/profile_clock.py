"""Engine clock and workgroup timeline of the fused batch kernel, from the
clock-probe variant of the library (libprimus_hip_probe.so, PH_CLOCK_PROBE=1).

    python -m primus_amd.profile_clock [--steps K] [--json FILE]

Workload: the headline batch -- 1024 limb-polynomials (256 polys x 4 limbs of
the four largest 50-bit primes = 1 mod 8192), N = 4096, forward NTT -> Shoup
multiply by a resident operand -> inverse NTT, one launch -- in both
single-pass shapes (8 and 16 coefficients per thread), over 12 rotating copies
of the batch.  The probe variant logs every wave as it ends (shader-clock and
100 MHz counters, hardware slot, workgroup) and keeps the last 32 waves of
each hardware wave slot.  For K back-to-back launches per shape it prints the launch period, the engine clock (the two counters'
advances between successive launches' entries of one hardware wave slot, x
100 MHz) and, per dispatch quarter of the workgroups, when they ended after
the previous launch had ended.  With --json it stores the median clock as
`sclk_ghz` in the file's entry for the kernel (the counter-profile table
bench.py reads; `valu_insts` there comes from an SQ_INSTS_VALU profile), so
that VALU issue time can be priced at the clock the kernel ran at.
"""
import argparse
import ctypes
import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LOG_N, N, LIMBS, COUNT, ROTATE = 12, 4096, 4, 1024, 12
WALL_HZ = 100e6  # s_memrealtime
PROBE_SLOTS, PROBE_DEPTH = 65536, 32  # PH_CLOCK_PROBE_SLOTS, PH_CLOCK_PROBE_DEPTH


def _kernel_name(per_thread):
    return "ph::chunk_kernel<ph::F64Arith, %d, %d, 39>" % (LOG_N, LOG_N - (per_thread.bit_length() - 1))


def _launches(L, _lib, waves_per_launch):
    """The logged launches, one array of entries each (sorted by end time:
    every wave of a launch ends before any wave of the next one on the same
    stream), from the rings of the last PROBE_DEPTH waves of every slot."""
    raw = (ctypes.c_uint64 * (PROBE_SLOTS * PROBE_DEPTH * 4))()
    counts = (ctypes.c_uint32 * PROBE_SLOTS)()
    _lib.check(L.ph_debug_clock_report(raw, counts, 0))
    log = np.frombuffer(raw, dtype=np.uint64).reshape(PROBE_SLOTS, PROBE_DEPTH, 4)
    n = np.frombuffer(counts, dtype=np.uint32).astype(np.int64)
    kept = np.minimum(n, PROBE_DEPTH)
    # the launches every used slot still holds: the last min(counts) of them
    used = n > 0
    last = int(kept[used].min())
    rows = []
    for slot in np.nonzero(used)[0]:
        for c in range(n[slot] - last, n[slot]):
            rows.append(log[slot, c % PROBE_DEPTH])
    st = np.array(rows).astype(np.int64)
    st = st[np.argsort(st[:, 1], kind="stable")]
    k = len(st) // waves_per_launch
    return [st[i * waves_per_launch:(i + 1) * waves_per_launch] for i in range(k)]


def _slot_clocks(launches):
    """Engine clock (GHz) between successive launches' entries of each
    hardware wave slot (HW_ID and XCC_ID)."""
    prev, out = {}, []
    for st in launches:
        for clk, wall, slot in zip(st[:, 0], st[:, 1], st[:, 2]):
            if slot in prev and wall > prev[slot][1]:
                out.append((clk - prev[slot][0]) / (wall - prev[slot][1]) * WALL_HZ / 1e9)
            prev[slot] = (clk, wall)
    return np.array(out)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--steps", type=int, default=100, help="launches per shape (the last 32 are logged)")
    ap.add_argument("--json", default=None)
    args = ap.parse_args()

    import torch  # before the library: torch initialises its HIP runtime first

    from . import _lib
    L = _lib.use_probe()
    # the default library beside it, on the same batch: the probe's clock
    # stands for the default kernel when their launch times agree
    D = ctypes.CDLL(_lib.LIB_PATH)
    _lib._declare(D)

    import primus_amd as pa
    from primus_amd import tuning

    rng = np.random.default_rng(0x5052494D5553 + 1)
    moduli = pa.ntt_primes_below(50, 2 * N, LIMBS)
    polys = np.stack([rng.integers(0, moduli[i % LIMBS], size=N, dtype=np.uint64) for i in range(COUNT)])
    operand = np.stack([rng.integers(0, q, size=N, dtype=np.uint64) for q in moduli])
    table = pa.UintCrtNttTable.new(LOG_N, moduli)
    op = pa.Operand(table, operand)
    _lib._lib = D
    d_table = pa.UintCrtNttTable.new(LOG_N, moduli)
    d_op = pa.Operand(d_table, operand)
    _lib._lib = L
    bufs = torch.empty((ROTATE, COUNT, N), dtype=torch.int64, device="cuda")
    host = torch.from_numpy(polys.view(np.int64))
    for r in range(ROTATE):
        bufs[r].copy_(host)
    stream = torch.cuda.Stream()
    sp = ctypes.c_void_p(stream.cuda_stream)
    ptrs = [ctypes.c_void_p(bufs[r].data_ptr()) for r in range(ROTATE)]
    rows = []
    def timed(lib_, tab, op_, steps):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for i in range(steps):
            lib_.ph_dcrt_fwd_mul_inv_device(tab.handle, ptrs[i % ROTATE], COUNT, op_.handle, sp)
        e1.record(stream)
        stream.synchronize()
        return e0.elapsed_time(e1) / steps * 1e3

    for per_thread, shape in ((8, 3), (16, 4)):
        for lib_ in (L, D):
            _lib._lib = lib_
            tuning.set_shape(LOG_N, "fwd_mul_inv", table.arith(), shape)
        _lib._lib = L
        # default and probe launches interleaved, after a warm-up of each
        timed(D, d_table, d_op, 20)
        timed(L, table, op, 20)
        default_us = [timed(D, d_table, d_op, args.steps), timed(D, d_table, d_op, args.steps)]
        probe_us = [timed(L, table, op, args.steps)]
        default_us.append(timed(D, d_table, d_op, args.steps))
        stream.synchronize()
        _lib.check(L.ph_debug_clock_report(None, None, 1))  # restart the log
        waves = COUNT * (1 << LOG_N) // per_thread // 64
        steps = args.steps
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for i in range(steps):
            L.ph_dcrt_fwd_mul_inv_device(table.handle, ptrs[i % ROTATE], COUNT, op.handle, sp)
        e1.record(stream)
        stream.synchronize()
        launch_us = e0.elapsed_time(e1) / steps * 1e3
        launches = _launches(L, _lib, waves)
        clk = _slot_clocks(launches)
        period = np.diff([st[:, 1].max() for st in launches]) / WALL_HZ * 1e6
        # end times of each dispatch quarter after the previous launch's last end
        rel = {k: [] for k in range(4)}
        for prev, st in zip(launches, launches[1:]):
            t0 = prev[:, 1].max()
            for k in range(4):
                sel = st[:, 3] // (COUNT // 4) == k
                rel[k].append(((np.median(st[sel, 1]) - t0), (st[sel, 1].max() - t0)))
        quarters = [(float(np.median([m for m, _ in rel[k]])) / WALL_HZ * 1e6,
                     float(np.median([x for _, x in rel[k]])) / WALL_HZ * 1e6) for k in range(4)]
        rows.append({"kernel": _kernel_name(per_thread), "per_thread": per_thread, "launch_us": launch_us,
                     "default_launch_us": float(np.median(default_us)), "probe_launch_us_before": probe_us[0],
                     "period_us": float(np.median(period)), "launches_logged": len(launches),
                     "sclk_ghz": float(np.median(clk)), "sclk_p5": float(np.percentile(clk, 5)),
                     "sclk_p95": float(np.percentile(clk, 95)), "quarters": quarters})
    print("| kernel | default library launch us | probe launch us (logged run) | launch period us (median, "
          "logged) | engine clock GHz: median (5th-95th percentile over wave slots and launches) |")
    print("|---|---|---|---|---|")
    for r in rows:
        print("| `%s` | %.2f | %.2f | %.2f | %.3f (%.3f-%.3f) |" % (
            r["kernel"], r["default_launch_us"], r["launch_us"], r["period_us"], r["sclk_ghz"], r["sclk_p5"],
            r["sclk_p95"]))
    print()
    print("Launch times: means over %d back-to-back launches (HIP events); the default library's is the median "
          "of three such runs around the probe's." % steps)
    print()
    print("Workgroups by dispatch quarter (blocks 0-255 are each CU's first, 768-1023 its last): "
          "median / last end of a wave, us after the previous launch's last end (medians over launches)")
    print()
    print("| kernel | blocks 0-255 | 256-511 | 512-767 | 768-1023 |")
    print("|---|---|---|---|---|")
    for r in rows:
        print("| `%s` | %s |" % (r["kernel"], " | ".join("%.2f / %.2f" % q for q in r["quarters"])))
    print()
    print(json.dumps(rows))
    if args.json:
        try:
            with open(args.json) as fh:
                table_ = json.load(fh)
        except (OSError, ValueError):
            table_ = {}
        for r in rows:
            table_.setdefault(r["kernel"], {"kernel": r["kernel"]}).update(
                {"sclk_ghz": r["sclk_ghz"], "sclk_launch_us": r["launch_us"]})
        with open(args.json, "w") as fh:
            json.dump(table_, fh, indent=1)
            fh.write("\n")
    return 0


if __name__ == "__main__":
    sys.exit(main())
