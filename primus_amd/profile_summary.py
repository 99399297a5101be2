"""Summaries of rocprofv3 result databases (rocpd SQLite) for the profiles/ folder.

    python -m primus_amd.profile_summary stats RESULTS_DB
        per-kernel dispatch count, grid, average / min / max duration
    python -m primus_amd.profile_summary sq RESULTS_DB KERNEL_SUBSTRING [JSON_FILE]
        every counter of the run averaged per dispatch of the kernel; with
        SQ_INSTS_VALU among them, the kernel's VALU issue floor, merged into
        JSON_FILE's entry for the kernel when given
    python -m primus_amd.profile_summary pmc FETCH_DB WRITE_DB KERNEL_SUBSTRING COPY_BYTES [JSON_FILE]
        FETCH_SIZE and WRITE_SIZE per dispatch of the kernel, calibrated by the
        byte counts the same counters give for the largest device-to-device
        copy in the run (COPY_BYTES read and COPY_BYTES written); prints a
        markdown table and one JSON line with the calibrated HBM traffic, and
        merges it into JSON_FILE (keyed by kernel) when given.
"""
import json
import sqlite3
import sys


def _rows(db, sql, params=()):
    con = sqlite3.connect(db)
    try:
        return list(con.execute(sql, params))
    finally:
        con.close()


def stats(db, last=0):
    """Per (kernel, grid): dispatches and durations; with last > 0 also the
    average over the last `last` dispatches of each (the timed steps of a run
    whose earlier dispatches are warm-up and tuning)."""
    rows = _rows(db, "select name, grid_x, count(*), avg(duration), min(duration), max(duration) "
                     "from kernels group by name, grid_x order by sum(duration) desc")
    print("| kernel | grid | dispatches | avg us | min us | max us |%s" % (" avg of last %d us |" % last if last else ""))
    print("|---|---|---|---|---|---|%s" % ("---|" if last else ""))
    for name, grid, cnt, avg, mn, mx in rows:
        tail = ""
        if last:
            con = sqlite3.connect(db)
            try:
                d = [r[0] for r in con.execute("select duration from kernels where name = ? and grid_x = ? "
                                               "order by start", (name, grid))]
            finally:
                con.close()
            d = d[-last:]
            tail = " %.2f |" % (sum(d) / len(d) / 1e3)
        print("| %s | %d | %d | %.2f | %.2f | %.2f |%s" % (name[:90], grid, cnt, avg / 1e3, mn / 1e3, mx / 1e3, tail))


def _per_dispatch(db, counter):
    return _rows(db, "select kernel_name, grid_size, value from counters_collection where counter_name = '%s' "
                     "order by dispatch_id" % counter)


def pmc(fetch_db, write_db, kernel, copy_bytes, json_file=None):
    copy_bytes = int(copy_bytes)
    fetch = _per_dispatch(fetch_db, "FETCH_SIZE")
    write = _per_dispatch(write_db, "WRITE_SIZE")
    kf = [v for n, g, v in fetch if kernel in n]
    kw = [v for n, g, v in write if kernel in n]
    # the calibration copy: the largest copy kernel dispatch of the run
    cf = max(v for n, g, v in fetch if "copyBuffer" in n or "copy" in n.lower())
    cw = max(v for n, g, v in write if "copyBuffer" in n or "copy" in n.lower())
    fetch_scale = copy_bytes / (cf * 1024.0)
    write_scale = copy_bytes / (cw * 1024.0)
    f = sum(kf) / len(kf)
    w = sum(kw) / len(kw)
    traffic = f * 1024.0 * fetch_scale + w * 1024.0 * write_scale
    print("| counter | kernel KiB/dispatch | copy KiB (%d B moved each way) | scale | kernel bytes |" % copy_bytes)
    print("|---|---|---|---|---|")
    print("| FETCH_SIZE | %.1f | %.1f | %.4f | %.0f |" % (f, cf, fetch_scale, f * 1024 * fetch_scale))
    print("| WRITE_SIZE | %.1f | %.1f | %.4f | %.0f |" % (w, cw, write_scale, w * 1024 * write_scale))
    print()
    entry = {"kernel": kernel, "fetch_kib": f, "write_kib": w, "fetch_scale": fetch_scale,
             "write_scale": write_scale, "traffic_bytes": traffic, "dispatches": len(kf)}
    print(json.dumps(entry))
    if json_file:
        _merge(json_file, kernel, entry)


def _merge(json_file, kernel, fields):
    """Updates JSON_FILE's entry for the kernel with fields."""
    try:
        with open(json_file) as fh:
            table = json.load(fh)
    except (OSError, ValueError):
        table = {}
    if "kernel" in table:  # a single-entry file: key it by its kernel
        table = {table["kernel"]: table}
    table.setdefault(kernel, {"kernel": kernel}).update(fields)
    with open(json_file, "w") as fh:
        json.dump(table, fh, indent=1)
        fh.write("\n")


# MI355X: 256 compute units x 4 SIMDs; a wave64 VALU instruction occupies its
# 16-lane SIMD for at least 4 cycles, at the 2.4 GHz peak engine clock.
SIMDS = 1024
VALU_CYCLES = 4
CLOCK_HZ = 2.4e9


def valu_floor_us(insts):
    """Shortest time in which the chip issues `insts` wave-level VALU
    instructions: every SIMD busy, one instruction per 4 cycles.  Multi-pass
    instructions (64-bit multiplies, transcendentals) only lengthen it."""
    return insts * VALU_CYCLES / (SIMDS * CLOCK_HZ) * 1e6


def sq(db, kernel, json_file=None):
    """Per-dispatch averages of every counter collected for the kernel."""
    rows = _rows(db, "select counter_name, dispatch_id, sum(value) from counters_collection "
                     "where kernel_name like ? group by counter_name, dispatch_id", ("%" + kernel + "%",))
    per = {}
    for name, _, v in rows:
        per.setdefault(name, []).append(v)
    print("| counter | dispatches | average per dispatch |")
    print("|---|---|---|")
    for name in sorted(per):
        vals = per[name]
        print("| %s | %d | %.0f |" % (name, len(vals), sum(vals) / len(vals)))
    valu = per.get("SQ_INSTS_VALU")
    if valu:
        insts = sum(valu) / len(valu)
        floor = valu_floor_us(insts)
        print()
        print("VALU issue floor: %.0f wave instructions x %d cycles / (%d SIMDs x %.1f GHz) = %.2f us"
              % (insts, VALU_CYCLES, SIMDS, CLOCK_HZ / 1e9, floor))
        if json_file:
            _merge(json_file, kernel, {"valu_insts": insts, "valu_floor_us": floor, "valu_dispatches": len(valu)})


if __name__ == "__main__":
    if len(sys.argv) >= 4 and sys.argv[1] == "sq":
        sq(*sys.argv[2:5])
    elif len(sys.argv) >= 3 and sys.argv[1] == "stats":
        stats(sys.argv[2], int(sys.argv[3]) if len(sys.argv) > 3 else 0)
    elif len(sys.argv) >= 6 and sys.argv[1] == "pmc":
        pmc(*sys.argv[2:7])
    else:
        print(__doc__)
        sys.exit(2)
