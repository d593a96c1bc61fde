"""Summaries of rocprofv3 CSV output for the committed profiles.

    python -m mray_amd.profiling stats <kernel-trace dir>
        per-kernel calls / total / average duration (the --stats table)
    python -m mray_amd.profiling traffic <fetch dir> <write dir> <bench log> <copy MiB>
        DRAM bytes per traced ray of the closest-hit kernel: FETCH_SIZE and
        WRITE_SIZE (KiB per dispatch) summed over the timed frame's KTrace
        dispatches, scaled so that the copy probe (torch.neg over <copy MiB>)
        reads and writes exactly <copy MiB>, divided by the frame's
        extension rays (from the bench log).

Only reads files rocprofv3 and bench.py wrote; runs nothing.
"""
import csv
import glob
import os
import re
import sys

TRACE_KERNEL = re.compile(r"^KTrace<\d+, false[,>]")   # closest hit, not the counting variant
COPY_KERNEL = "neg"


def _find(d, suffix):
    hits = sorted(glob.glob(os.path.join(d, "**", "*" + suffix), recursive=True))
    if not hits:
        raise SystemExit(f"no *{suffix} under {d}")
    return hits


def kernel_stats(d):
    rows = []
    for path in _find(d, "kernel_stats.csv"):
        with open(path) as f:
            for r in csv.DictReader(f):
                rows.append((r["Name"], int(r["Calls"]), float(r["TotalDurationNs"]),
                             float(r["AverageNs"]), float(r["Percentage"])))
    rows.sort(key=lambda r: -r[2])
    return rows


def counter_per_kernel(d, counter):
    """{kernel name: [value per dispatch]} for one counter."""
    out = {}
    for path in _find(d, "counter_collection.csv"):
        with open(path) as f:
            for r in csv.DictReader(f):
                if r["Counter_Name"] != counter:
                    continue
                out.setdefault(r["Kernel_Name"], []).append(float(r["Counter_Value"]))
    return out


def _short(name):
    name = re.sub(r"\(.*$", "", name)
    return name.replace("void ", "").replace("mray::", "")


def _pick(per_kernel, pred):
    vals = []
    for k, v in per_kernel.items():
        if pred(_short(k)):
            vals.extend(v)
    return vals


def cmd_stats(d):
    print(f"{'kernel':48s} {'calls':>6s} {'total ms':>10s} {'avg us':>10s} {'%':>6s}")
    for name, calls, tot, avg, pct in kernel_stats(d):
        print(f"{_short(name)[:48]:48s} {calls:6d} {tot / 1e6:10.3f} {avg / 1e3:10.2f} {pct:6.2f}")


def cmd_traffic(fetch_dir, write_dir, bench_log, copy_mib):
    fetch = counter_per_kernel(fetch_dir, "FETCH_SIZE")
    write = counter_per_kernel(write_dir, "WRITE_SIZE")
    copy_bytes = float(copy_mib) * (1 << 20)
    cf = _pick(fetch, lambda n: COPY_KERNEL in n)
    cw = _pick(write, lambda n: COPY_KERNEL in n)
    if len(cf) != 1 or len(cw) != 1:
        raise SystemExit(f"expected one copy-probe dispatch, found {len(cf)} / {len(cw)}")
    fscale = copy_bytes / (cf[0] * 1024.0)
    wscale = copy_bytes / (cw[0] * 1024.0)
    tf = _pick(fetch, lambda n: TRACE_KERNEL.match(n) is not None)
    tw = _pick(write, lambda n: TRACE_KERNEL.match(n) is not None)
    rays = None
    for line in open(bench_log):
        m = re.search(r"step 1/\d+: .* (\d+) extension rays", line)
        if m:
            rays = int(m.group(1))
    if not tf or not tw or rays is None:
        raise SystemExit("KTrace dispatches or the timed frame's ray count not found")
    fbytes = sum(tf) * 1024.0 * fscale
    wbytes = sum(tw) * 1024.0 * wscale
    print(f"copy probe: {copy_mib} MiB; FETCH_SIZE {cf[0]:.0f} KiB (scale {fscale:.4f}), "
          f"WRITE_SIZE {cw[0]:.0f} KiB (scale {wscale:.4f})")
    print(f"KTrace dispatches: {len(tf)}; extension rays in the frame: {rays}")
    print(f"KTrace DRAM read  per launch (avg): {fbytes / len(tf) / 1e6:.2f} MB")
    print(f"KTrace DRAM write per launch (avg): {wbytes / len(tw) / 1e6:.2f} MB")
    print(f"dram_bytes_per_ray: {(fbytes + wbytes) / rays:.2f}")


def main(argv):
    if len(argv) >= 2 and argv[0] == "stats":
        cmd_stats(argv[1])
    elif len(argv) >= 5 and argv[0] == "traffic":
        cmd_traffic(argv[1], argv[2], argv[3], argv[4])
    else:
        raise SystemExit(__doc__)


if __name__ == "__main__":
    main(sys.argv[1:])
