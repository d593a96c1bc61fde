"""Summaries of rocprofv3 CSV output for the committed profiles.

    python -m mray_amd.profiling stats <kernel-trace dir>
        per-kernel calls / total / average duration (the --stats table)
    python -m mray_amd.profiling traffic <fetch dir> <write dir> <bench log> <copy MiB>
        DRAM bytes per traced ray of the closest-hit kernel: FETCH_SIZE and
        WRITE_SIZE (KiB per dispatch) summed over the timed frame's KTrace
        dispatches, scaled so that the copy probe (torch.neg over <copy MiB>)
        reads and writes exactly <copy MiB>, divided by the frame's
        extension rays (from the bench log).

    python -m mray_amd.profiling counters <sq dir> <tcc dir> [<fetch dir> <write dir> <copy MiB>]
        per hot kernel (closest hit, shading, any hit; the counting variants
        excluded): average per dispatch of the SQ counters (waves, VALU
        instructions, VALU busy cycles, VALU thread cycles, wave cycles,
        busy cycles; GRBM_GUI_ACTIVE), VALU busy %, VALU lane utilisation,
        and the L2 (TCC) hit rate; with the FETCH_SIZE / WRITE_SIZE runs,
        DRAM bytes per dispatch scaled by the copy probe.

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
    # rays per launch of the frame: the counting frame's extension rays over
    # its bounces (the same frame as every other one of the run)
    # the closest-hit launches cover the timed frames and as many profiled
    # ones (bench.py: 2 x steps frames; the counting frame uses the counting
    # kernel, excluded above)
    rays = bounces = steps = None
    for line in open(bench_log):
        m = re.search(r"counting frame: \{.*'ext_rays': (\d+).*'iterations': (\d+)", line)
        if m:
            rays, bounces = int(m.group(1)), int(m.group(2))
        m = re.search(r'"steps": (\d+)', line)
        if m:
            steps = int(m.group(1))
    if not tf or not tw or rays is None or steps is None:
        raise SystemExit("KTrace dispatches, the counting frame or the step count not found")
    fbytes = sum(tf) * 1024.0 * fscale / len(tf)
    wbytes = sum(tw) * 1024.0 * wscale / len(tw)
    print(f"copy probe: {copy_mib} MiB; FETCH_SIZE {cf[0]:.0f} KiB (scale {fscale:.4f}), "
          f"WRITE_SIZE {cw[0]:.0f} KiB (scale {wscale:.4f})")
    print(f"KTrace dispatches: {len(tf)}; a frame: {rays} extension rays in {bounces} bounces")
    print(f"KTrace DRAM read  per launch (avg): {fbytes / 1e6:.2f} MB")
    print(f"KTrace DRAM write per launch (avg): {wbytes / 1e6:.2f} MB")
    print(f"dram_bytes_per_launch: {fbytes + wbytes:.0f}")
    frames = 2 * steps
    print(f"dram_bytes_per_ray ({frames} frames): "
          f"{(fbytes + wbytes) * len(tf) / (frames * rays):.2f}")


HOT = (("KTrace (closest hit)", TRACE_KERNEL),
       ("KShade", re.compile(r"^KShade<")),
       ("KShadow (any hit)", re.compile(r"^KShadow<.*false>$")))
SIMDS = 256 * 4            # MI355X: 256 CUs x 4 SIMDs
XCDS = 8


def _avg(per_kernel, rx):
    v = _pick(per_kernel, lambda n: rx.match(n) is not None)
    return (sum(v) / len(v), len(v)) if v else (None, 0)


def cmd_counters(sq_dir, tcc_dir, fetch_dir=None, write_dir=None, copy_mib=None):
    sq_names = ("SQ_WAVES", "SQ_INSTS_VALU", "SQ_ACTIVE_INST_VALU", "SQ_THREAD_CYCLES_VALU",
                "SQ_WAVE_CYCLES", "SQ_BUSY_CYCLES", "GRBM_GUI_ACTIVE")
    sq = {c: counter_per_kernel(sq_dir, c) for c in sq_names}
    tcc = {c: counter_per_kernel(tcc_dir, c) for c in ("TCC_HIT_sum", "TCC_MISS_sum")}
    dram = None
    if fetch_dir:
        fetch = counter_per_kernel(fetch_dir, "FETCH_SIZE")
        write = counter_per_kernel(write_dir, "WRITE_SIZE")
        copy_bytes = float(copy_mib) * (1 << 20)
        cf = _pick(fetch, lambda n: COPY_KERNEL in n)
        cw = _pick(write, lambda n: COPY_KERNEL in n)
        dram = (fetch, write, copy_bytes / (cf[0] * 1024.0), copy_bytes / (cw[0] * 1024.0))
        print(f"copy probe: {copy_mib} MiB; FETCH_SIZE scale {dram[2]:.4f}, "
              f"WRITE_SIZE scale {dram[3]:.4f}")
    for label, rx in HOT:
        a = {c: _avg(sq[c], rx)[0] for c in sq_names}
        n = _avg(sq["SQ_WAVES"], rx)[1]
        hit, _ = _avg(tcc["TCC_HIT_sum"], rx)
        miss, _ = _avg(tcc["TCC_MISS_sum"], rx)
        print(f"== {label}: {n} dispatches (SQ pass)")
        for c in sq_names:
            print(f"   {c:24s} {a[c]:.4g}" if a[c] is not None else f"   {c:24s} -")
        if a["SQ_ACTIVE_INST_VALU"] and a["GRBM_GUI_ACTIVE"]:
            # GRBM_GUI_ACTIVE reads the sum over the 8 XCDs (8x the kernel's
            # cycles; SQ_BUSY_CYCLES over the 32 shader engines agrees)
            cyc = a["GRBM_GUI_ACTIVE"] / XCDS
            busy = 100.0 * a["SQ_ACTIVE_INST_VALU"] * 4 / SIMDS / cyc
            print(f"   kernel cycles (per XCD)  {cyc:.4g}   (GRBM_GUI_ACTIVE / 8; "
                  f"SQ_BUSY_CYCLES / 32 = {a['SQ_BUSY_CYCLES'] / 32:.4g})")
            print(f"   VALU busy %              {busy:.1f}   (100 * ACTIVE_INST_VALU * 4 / SIMDs / cycles)")
        if a["SQ_THREAD_CYCLES_VALU"] and a["SQ_ACTIVE_INST_VALU"]:
            util = 100.0 * a["SQ_THREAD_CYCLES_VALU"] / (64.0 * a["SQ_ACTIVE_INST_VALU"])
            print(f"   VALU lane utilisation %  {util:.1f}   (THREAD_CYCLES_VALU / (64 * ACTIVE_INST_VALU))")
        if hit is not None and miss is not None and hit + miss > 0:
            print(f"   L2 hit rate %            {100.0 * hit / (hit + miss):.1f}   "
                  f"(TCC_HIT {hit:.4g}, TCC_MISS {miss:.4g} per dispatch)")
        if dram:
            f, _ = _avg(dram[0], rx)
            w, _ = _avg(dram[1], rx)
            if f is not None and w is not None:
                print(f"   DRAM MB per dispatch     read {f * 1024 * dram[2] / 1e6:.2f}, "
                      f"write {w * 1024 * dram[3] / 1e6:.2f}")


def main(argv):
    if len(argv) >= 2 and argv[0] == "stats":
        cmd_stats(argv[1])
    elif len(argv) >= 5 and argv[0] == "traffic":
        cmd_traffic(argv[1], argv[2], argv[3], argv[4])
    elif len(argv) >= 3 and argv[0] == "counters":
        cmd_counters(*argv[1:6])
    else:
        raise SystemExit(__doc__)


if __name__ == "__main__":
    main(sys.argv[1:])
