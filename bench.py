"""Benchmark: wavefront path tracing on one MI355X.

Workload (the 1080p headline configuration): heightfield 1024 x 1024
(0.2 x 6-octave fBm) + 2000 icospheres = 2,253,058 triangles, four material
kinds (Lambert, Unreal metal, Refract 1.5, Unreal plastic), 16 quad lights +
constant Skysphere, 1920 x 1080, 64 spp, max depth 10 (rrRange [10, 10]),
spectral hero wavelengths (4 lambda, hyperbolic sampling), NEE + MIS, fp32.

One step = one full frame: totalSPP x pixels camera paths through
refill -> closest hit -> shade -> shadow/film, all on the GPU with no host
synchronisation inside the frame.  Output: one JSON line (see --help).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Mpaths/s at 1080p; equal-time relMSE: radiance-cache PT vs plain PT, 1 MI355X"
HBM_PEAK_GBS = 8000.0
# bytes a node visit loads: BNode4 = six float4 boxes + uint4 children (112 B);
# CNode8 = origin/exponent header + 48 B of 8-bit boxes + 32 B children (96 B)
NODE_BYTES = {4: 112, 8: 96}
TRI_BYTES = 48            # SoupTri
RAY_BYTES = 32 + 28       # RayGMem in, hit record out (prim/t/bary + keys)


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--spp", type=int, default=64)
    ap.add_argument("--width", type=int, default=1920)
    ap.add_argument("--height", type=int, default=1080)
    ap.add_argument("--grid", type=int, default=1024)
    ap.add_argument("--spheres", type=int, default=2000)
    ap.add_argument("--rgb", action="store_true", help="PathTracerRGB instead of Spectral")
    ap.add_argument("--builder", choices=("sah", "lbvh"), default="sah",
                    help="traversal tree: binned SAH (default) or the collapsed Karras LBVH")
    ap.add_argument("--queues", action="store_true",
                    help="material-keyed shading queues instead of one shading kernel")
    ap.add_argument("--bvh-width", type=int, choices=(4, 8), default=4,
                    help="BVH node width: 4 (fp32 boxes) or 8 (compressed, 8-bit boxes)")
    ap.add_argument("--lane-refill", type=int, choices=(0, 1), default=0,
                    help="closest hit with lane refill (4-wide trees): idle lanes take new rays")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--nrc", type=int, default=1,
                    help="run the radiance-cache equal-time leg (Cornell box, 1 spp per frame)")
    ap.add_argument("--nrc-width", type=int, default=1920)
    ap.add_argument("--nrc-height", type=int, default=1080)
    ap.add_argument("--nrc-only", action="store_true",
                    help="run only the radiance-cache equal-time leg (its own JSON line)")
    ap.add_argument("--nrc-frames", type=int, default=256)
    ap.add_argument("--nrc-ref-spp", type=int, default=16384)
    ap.add_argument("--nrc-scene", choices=("cornell", "terrain"), default="cornell")
    ap.add_argument("--nrc-variants", default="0.01:0:1:3:0;1.0:1:2:4:1",
                    help="cache variants 'spread:neeAtEnd:minDepth:trainShift[:lazyAdam];...'")
    ap.add_argument("--workload", choices=("frame", "batch", "instanced"), default="frame",
                    help="frame: the 1080p headline frame; batch: many scenes in one renderer")
    ap.add_argument("--inst-cache-frames", type=int, default=64)
    ap.add_argument("--inst-cache-ref-spp", type=int, default=256)
    ap.add_argument("--legs", default="batch,instanced,instanced_cache",
                    help="frame workload: extra legs reported in the same JSON line "
                         "(comma list of batch, instanced, instanced_cache; empty for none)")
    ap.add_argument("--inst-spp", type=int, default=16,
                    help="samples per pixel of the instanced workload")
    ap.add_argument("--inst-views", type=int, default=8,
                    help="viewports of the instanced workload")
    ap.add_argument("--batch-scenes", type=int, default=32)
    ap.add_argument("--batch-generator", choices=("cornell", "terrain"), default="cornell",
                    help="batch: tessellated Cornell scenes (the batched workload) or terrains")
    ap.add_argument("--batch-spp", type=int, default=16)
    ap.add_argument("--no-autotune", action="store_true",
                    help="batch: skip the autotuner and use --batch-slots / --batch-tpb, "
                         "4-wide, no queues")
    ap.add_argument("--batch-slots", type=int, default=1 << 22)
    ap.add_argument("--batch-tpb", type=int, choices=(128, 256), default=128)
    ap.add_argument("--no-sequential", action="store_true",
                    help="batch: skip rendering the views one at a time for comparison")
    ap.add_argument("--probe-copy-mib", type=int, default=0,
                    help="run one plain device copy of this size first (HBM counter calibration)")
    return ap.parse_args()


# Algorithmic bytes of one shaded path (KShade, surface hit, NEE+MIS): path
# state read -- active index 4, meta 4, hit 16, ray origin 16, direction 16,
# throughput 16, radiance 16, previous pdf 4, PCG state 4 (+ wavelengths 16
# spectral); scene gathers -- surface key 4, material 52, indices 12, three
# positions 36, three TBN quaternions 48, light triangle 4 + positions 36 +
# radiance 16 (+ colour coefficients 16 spectral); written -- throughput 16,
# previous pdf 4, ray 32, PCG state 4, meta 4, shadow ray 32 + radiance 16 +
# queue index 4.
SHADE_BYTES = {False: 96 + 208 + 112, True: 112 + 224 + 112}
# any-hit ray outside the tree walk: queue index 4, shadow ray 32, its
# radiance 16, path radiance read and written 32
SHADOW_RAY_BYTES = 4 + 32 + 16 + 32


def kernel_rooflines(cst, ptotals, phases, width, spectral, queues):
    """Launch-averaged algorithmic GB/s of the shading and any-hit kernels
    (HIP events around each phase on the render stream)."""
    out = []
    shade_ms, shade_n = phases["shade"]
    if shade_n and not queues:
        per = shade_ms / shade_n
        paths = ptotals["ext_rays"] / shade_n
        gbs = paths * SHADE_BYTES[spectral] / (per * 1e-3) / 1e9
        out.append({"kernel": "KShade", "bound": "hbm", "achieved": round(gbs, 1),
                    "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(gbs / HBM_PEAK_GBS, 4),
                    "bytes_per_path": SHADE_BYTES[spectral], "avg_launch_ms": round(per, 4),
                    "paths_per_launch": round(paths, 1)})
    sh_ms, sh_n = phases["shadow"]
    if sh_n and cst.get("shadow_rays"):
        per = sh_ms / sh_n
        rays = ptotals["shadow_rays"] / sh_n
        nodes = cst["shadow_node_visits"] / cst["shadow_rays"]
        tris = cst["shadow_tri_tests"] / cst["shadow_rays"]
        b = SHADOW_RAY_BYTES + nodes * NODE_BYTES[width] + tris * TRI_BYTES
        gbs = rays * b / (per * 1e-3) / 1e9
        out.append({"kernel": "KShadow", "bound": "hbm", "achieved": round(gbs, 1),
                    "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(gbs / HBM_PEAK_GBS, 4),
                    "bytes_per_ray": round(b, 1), "nodes_per_ray": round(nodes, 2),
                    "tris_per_ray": round(tris, 2), "avg_launch_ms": round(per, 4),
                    "rays_per_launch": round(rays, 1)})
    return out


def probe_copy(torch, mib):
    """One plain copy kernel of known size (torch.neg: reads and writes
    mib MiB once each), the reference point that scales rocprofv3's
    FETCH_SIZE / WRITE_SIZE counters to bytes."""
    n = mib * (1 << 20) // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    torch.neg(src, out=dst)
    torch.cuda.synchronize()
    log(f"probe copy: {mib} MiB read + {mib} MiB written")
    del src, dst


def traffic_from_profiles(rays_per_launch):
    """HBM bytes per KTrace launch from the committed rocprofv3 PMC summary
    (FETCH_SIZE + WRITE_SIZE per launch of the same frame, scaled by the copy
    probe), or None when no summary is present."""
    path = os.path.join(ROOT, "profiles", "ktrace_hbm_traffic_1080p.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("dram_bytes_per_launch:"):
            return float(line.split(":", 1)[1])
    return None


def cpu_baseline(scene, params, budget_s):
    """refcpu (CPU restatement of the reference's TracerCPU semantics) on the
    box's host cores, over a bounded number of paths of the same workload."""
    import numpy as np  # noqa: F401
    from oracle import refcpu
    isa = refcpu.use_native_isa()
    model, _ = refcpu.host_cpu()
    if params.spectral:
        from mray_amd import native
        refcpu.spectral_load(native.spectral_tables())
    threads = os.cpu_count() or 1
    o = refcpu.OracleScene(scene)
    # grow the sample until it takes about the budget (first paths of the frame)
    total = params.width * params.height * params.totalSPP
    sample = 20000
    while True:
        t0 = time.perf_counter()
        _, st = o.render(params, threads=threads, path_limit=sample)
        dt = time.perf_counter() - t0
        if dt >= 0.5 * budget_s or sample >= total:
            break
        sample = int(min(total, max(2 * sample, sample * budget_s / max(dt, 1e-3))))
    return {"value": st["dead_paths"] / dt / 1e6, "unit": "Mpaths/s", "cores": threads,
            "kind": "port", "cpu": model, "build": f"g++ -O3 -march={isa}",
            "sample": f"first {st['dead_paths']} paths of the same frame (reference schedule), "
                      f"{dt:.1f} s on {threads} threads of {model}, g++ -O3 -march={isa}"}


# The cache's MLP: six 64 x 64 layers (the input padded to 64, the output
# layer's 61 zero columns included, as the MFMA tiles compute them).
MLP_FLOP_PER_SAMPLE = 2 * 6 * 64 * 64          # one forward pass
BF16_PEAK_TFLOPS = 2500.0                       # MI355X dense bf16 MFMA


def mlp_roofline(kernel_ms, batch):
    """MFMA roofline of the training GEMM kernels from per-kernel HIP-event
    times: forward (with the hash encoding fused in), data backward and
    weight gradients each do one 64 x 64 product per layer and sample."""
    out = {}
    flops = 0.0
    ms = 0.0
    for k in ("train_forward", "backward", "wgrad"):
        t, n = kernel_ms[k]
        if n:
            f = MLP_FLOP_PER_SAMPLE * batch
            out[k] = {"avg_ms": round(t / n, 4),
                      "tflops": round(f / (t / n * 1e-3) / 1e12, 2)}
            flops += f * n
            ms += t
    if ms <= 0:
        return None
    achieved = flops / (ms * 1e-3) / 1e12
    out.update({"bound": "mfma", "achieved": round(achieved, 2), "peak": BF16_PEAK_TFLOPS,
                "unit": "TFLOP/s", "frac": round(achieved / BF16_PEAK_TFLOPS, 5),
                "flop_per_step": 3 * MLP_FLOP_PER_SAMPLE * batch})
    return out


def relmse(img, ref):
    import numpy as np
    return float(np.mean((img - ref) ** 2 / (ref ** 2 + 1e-2)))


def film_image(film):
    import numpy as np
    return film[:3] / np.maximum(film[3], 1e-12)


def _loglog_at(curve, x_key, y_key, x):
    """y at x on a curve of (x, y) points, log-log interpolated; None outside."""
    import numpy as np
    for a, b in zip(curve, curve[1:]):
        x0, x1, y0, y1 = a[x_key], b[x_key], a[y_key], b[y_key]
        lo, hi = min(x0, x1), max(x0, x1)
        if lo <= x <= hi and x0 != x1 and y0 > 0 and y1 > 0:
            t = (np.log(x) - np.log(x0)) / (np.log(x1) - np.log(x0))
            return float(np.exp(np.log(y0) + t * (np.log(y1) - np.log(y0))))
    return None


def parse_nrc_variants(text):
    """'spread:nee:minDepth:trainShift[:lazyAdam];...' -> list of dicts."""
    out = []
    for part in text.split(";"):
        f = part.split(":")
        out.append({"spread": float(f[0]), "nee_at_end": bool(int(f[1])), "min_depth": int(f[2]),
                    "train_shift": int(f[3]), "lazy_adam": bool(int(f[4])) if len(f) > 4 else True})
    return out


def nrc_equal_time(args):
    """Radiance-cache PT vs plain PT at equal time (1 spp per frame, NEE+MIS,
    RGB), judged against a plain render at --nrc-ref-spp.  Both run frame by
    frame; error curves are sampled at doubling frame counts; times are render
    time only (the film read-back for the error is outside).  Each cache
    variant (termination threshold, NEE at the terminal vertex, minimum
    terminal depth, training-path share) runs against the same judge."""
    import numpy as np
    from mray_amd import abi, native, scenes
    torch = native.require_gpu()
    W, H = args.nrc_width, args.nrc_height
    if args.nrc_scene == "terrain":
        scene = scenes.terrain_spheres(grid=args.grid, spheres=args.spheres, width=W, height=H)
        rr, label = (10, 10), f"terrain {scene.triangle_count} tris"
    else:
        scene = scenes.cornell(aspect=W / H)
        rr, label = (8, 8), "cornell"
    acc = native.Accel(scene)
    t0 = time.perf_counter()
    ref_r = native.Renderer(acc, abi.render_params(W, H, args.nrc_ref_spp, rr=rr, seed=99))
    done = 0
    while done < args.nrc_ref_spp:
        step = min(512, args.nrc_ref_spp - done)
        ref_r.render_frame(step)
        done += step
        log(f"nrc leg: reference {done}/{args.nrc_ref_spp} spp")
    ref = film_image(ref_r.film())
    ref_s = time.perf_counter() - t0
    ref_r.close()
    log(f"nrc leg: reference {args.nrc_ref_spp} spp in {ref_s:.1f} s")

    def run(v):
        r = native.Renderer(acc, abi.render_params(W, H, args.nrc_frames, rr=rr, seed=1))
        nrc = None
        if v is not None:
            nrc = native.Nrc(seed=2, lazy_adam=v["lazy_adam"])
            r.attach_nrc(nrc, train_shift=v["train_shift"], spread=v["spread"], batch=1 << 14,
                         steps=4, min_depth=v["min_depth"], nee_at_end=v["nee_at_end"])
        curve = []
        spent = 0.0
        nxt = 1
        name = "plain" if v is None else f"cached {v}"
        for f in range(1, args.nrc_frames + 1):
            torch.cuda.synchronize()
            t = time.perf_counter()
            r.render_frame(1)
            torch.cuda.synchronize()
            spent += time.perf_counter() - t
            if f == nxt or f == args.nrc_frames:
                curve.append({"frames": f, "seconds": round(spent, 4),
                              "relmse": relmse(film_image(r.film()), ref)})
                log(f"nrc leg: {name} frame {f}: {spent:.3f} s, relMSE {curve[-1]['relmse']:.4e}")
                nxt *= 2
        stats = r.cache_stats() if nrc is not None else None
        if nrc is not None:
            # one more frame with HIP events around every cache kernel (outside
            # the timed curve): the MLP kernels' MFMA roofline
            nrc.set_profiling(True)
            r.render_frame(1)
            stats["kernel_ms"] = nrc.kernel_times()
            stats["mlp_roofline"] = mlp_roofline(stats["kernel_ms"], 1 << 14)
            nrc.set_profiling(False)
        r.close()
        if nrc is not None:
            nrc.close()
        return curve, stats

    plain, _ = run(None)
    e_plain, t_plain_all = plain[-1]["relmse"], plain[-1]["seconds"]
    variants = []
    for v in parse_nrc_variants(args.nrc_variants):
        cached, cst = run(v)
        e_c, t_c = cached[-1]["relmse"], cached[-1]["seconds"]
        # plain error at the cached run's total time
        e_plain_eq = _loglog_at(plain, "seconds", "relmse", t_c)
        if e_plain_eq is None:
            e_plain_eq = e_plain * t_plain_all / t_c          # relMSE ~ 1/t beyond the curve
        # time plain PT needs to reach the cached run's final error
        t_match = _loglog_at(sorted(plain, key=lambda c: c["relmse"]), "relmse", "seconds", e_c)
        if t_match is None and e_plain > e_c:
            t_match = t_plain_all * e_plain / e_c
        # the target: time the cached run needs to reach plain's final error,
        # over plain's time (None when the cached run never gets there)
        t_reach = _loglog_at(sorted(cached, key=lambda c: c["relmse"]), "relmse", "seconds", e_plain)
        variants.append({
            "variant": v,
            "cached": cached,
            "relmse_cached": e_c,
            "relmse_plain_equal_time": e_plain_eq,
            "time_ratio_to_match": None if t_match is None else round(t_c / t_match, 4),
            "time_ratio_to_reach_plain_final": None if t_reach is None else round(t_reach / t_plain_all, 4),
            "cache_stats": cst,
        })
    acc.close()
    best = min(variants, key=lambda x: x["relmse_cached"])
    return {
        "scene": f"{label} {W}x{H}, 1 spp/frame, {args.nrc_frames} frames, NEE+MIS, "
                 f"depth {rr[1]}, RGB",
        "reference_spp": args.nrc_ref_spp,
        "cache": "hash grid 16 x 2^19 x 2 fp16 + MLP 64 x 5 ReLU (bf16 MFMA), 4 x 2^14 pairs/frame",
        "plain": plain,
        "variants": variants,
        "relmse_plain": e_plain,
        "relmse_cached": best["relmse_cached"],
        "relmse_plain_equal_time": best["relmse_plain_equal_time"],
        "time_ratio_to_match": best["time_ratio_to_match"],
        "time_ratio_to_reach_plain_final": best["time_ratio_to_reach_plain_final"],
        "mlp_roofline": (best["cache_stats"] or {}).get("mlp_roofline"),
    }


def bench_batch(args):
    """Batched workload: 32 independent Cornell scenes tessellated to
    triangle counts log-uniform in [10k, 2M] (or terrain scenes), resolutions
    from 256^2 / 512^2 / 720p / 1080p, 16 spp, depth 8, spectral, rendered by
    one batch renderer; settings from the autotuner."""
    import numpy as np
    from mray_amd import abi, native, scenes
    torch = native.require_gpu()
    spectral = not args.rgb
    log(f"generating {args.batch_scenes} scenes")
    t0 = time.perf_counter()
    if args.batch_generator == "cornell":
        scs, views = scenes.cornell_batch_workload(count=args.batch_scenes, spp=args.batch_spp)
    else:
        scs, views = scenes.batch_workload(count=args.batch_scenes, spp=args.batch_spp)
    tris = sum(s.triangle_count for s in scs)
    log(f"scenes: {tris} triangles in {time.perf_counter() - t0:.1f} s")
    t0 = time.perf_counter()
    merged = {}
    singles = []
    for width in ((4,) if args.no_autotune else (4, 8)):
        accs = []
        for s in scs:
            accs.append(native.Accel(s, width=width))
        merged[width] = native.BatchScene(accs)
        if width == 4 and not args.no_sequential:
            singles = accs          # kept for the one-view-at-a-time comparison
        else:
            for a in accs:
                a.close()
        log(f"merged {width}-wide scene: {merged[width].info()}")
    build_s = time.perf_counter() - t0
    params = abi.render_params(1, 1, 1, rr=(8, 8), mode=abi.SAMPLE_NEE_MIS, spectral=spectral,
                               seed=0)
    trials = []
    if args.no_autotune:
        best = {"width": 4, "queues": 0, "slots": args.batch_slots, "trace_tpb": args.batch_tpb}
    else:
        t0 = time.perf_counter()
        best, trials = native.batch_autotune(merged[4], merged[8], views, params)
        log(f"autotune ({len(trials)} trials, {time.perf_counter() - t0:.1f} s): {best}")
    params.maxRayCount = best["slots"]
    params.flags = abi.RENDER_MATERIAL_QUEUES if best["queues"] else 0
    for w in list(merged):
        if w != best["width"]:
            merged.pop(w).close()
    bs = merged[best["width"]]
    r = native.BatchRenderer(bs, views, params)
    r.set_trace_tpb(best["trace_tpb"])
    paths = sum(v["width"] * v["height"] * v["spp"] for v in views)
    m = _timed_render(r, args, torch, "batch")
    totals, elapsed, phases = m["totals"], m["elapsed"], m["phases"]
    lat = r.view_times()
    films = r.film()
    assert totals["dead_paths"] == paths * args.steps, totals
    assert m["ptotals"]["dead_paths"] == paths * args.steps, m["ptotals"]
    assert all(np.isfinite(f).all() for f in films)
    result = {
        "metric": f"Mpaths/s, {len(views)} scenes batched in one renderer, 1 MI355X",
        "value": round(paths * args.steps / elapsed / 1e6, 3),
        "unit": "Mpaths/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(elapsed * 1e3 / args.steps, 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded procedural scenes)",
        "config": {
            "workload": f"{len(views)} {args.batch_generator} scenes ({tris} tris, 10k-2M each), "
                        f"256^2/512^2/720p/1080p views, {args.batch_spp} spp, depth 8, "
                        f"{'spectral 4 lambda' if spectral else 'RGB'}, NEE+MIS",
            "triangles": tris,
            "views": [[v["width"], v["height"]] for v in views],
            "scene_triangles": [s.triangle_count for s in scs],
            "paths_per_step": paths,
        },
        "tuned": best,
        "autotune_trials": trials,
        "mrays_per_s": round((totals["ext_rays"] + totals["shadow_rays"]) / elapsed / 1e6, 3),
        "bounces_per_step": totals["iterations"] // args.steps,
        "accel_build_s": round(build_s, 3),
        "phase_ms_per_step": {k: round(v[0] / args.steps, 3) for k, v in phases.items()},
        "view_latency_ms": [round(float(x), 3) for x in lat],
        "host_waits_per_step": totals["host_waits"] / args.steps,
        "device_drains_per_step": totals["device_drains"] / args.steps,
        "profiled_ms_per_step": round(m["prof_elapsed"] * 1e3 / args.steps, 2),
        "roofline": _trace_roofline(m, best["width"], f"KTrace (closest hit, {best['width']}-wide BVH, batch)"),
        "kernel_rooflines": kernel_rooflines(m["cst"], m["ptotals"], phases, best["width"],
                                             spectral, bool(best["queues"])),
    }
    r.close()
    bs.close()
    if singles:
        # the same views rendered one at a time by one-scene renderers (a
        # slot per pixel, the default pool), for comparison with the batch
        seq_s = 0.0
        for s, a, v in zip(scs, singles, views):
            p1 = abi.render_params(v["width"], v["height"], v["spp"], rr=(8, 8),
                                   mode=abi.SAMPLE_NEE_MIS, spectral=spectral, seed=0)
            r1 = native.Renderer(a, p1)
            r1.render_all()
            r1.reset()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            r1.render_all()
            torch.cuda.synchronize()
            seq_s += time.perf_counter() - t0
            r1.close()
            a.close()
        result["sequential"] = {"ms_per_step": round(seq_s * 1e3, 2),
                                "mpaths_per_s": round(paths / seq_s / 1e6, 3)}
        log(f"one view at a time: {result['sequential']}")
    return result


def _timed_render(r, args, torch, label):
    """Counting frame, warmup, args.steps timed frames of renderer r on the
    production path (hipGraph bursts, no phase events), then the same number
    of frames with HIP events around each phase for per-kernel times."""
    r.set_measurement(profiling=False, counting=True)
    cst = r.render_all()
    log(f"{label} counting frame: {cst}")
    r.set_measurement(profiling=False, counting=False)
    for w in range(args.warmup):
        r.reset()
        r.render_all()
        log(f"{label} warmup {w + 1}/{args.warmup} done")
    totals = {"dead_paths": 0, "ext_rays": 0, "shadow_rays": 0, "iterations": 0,
              "host_waits": 0, "device_drains": 0}
    torch.cuda.synchronize()
    t_all = time.perf_counter()
    for k in range(args.steps):
        t0 = time.perf_counter()
        r.reset()
        st = r.render_all()
        for key in totals:
            totals[key] += st[key]
        log(f"{label} step {k + 1}/{args.steps}: {(time.perf_counter() - t0) * 1e3:.1f} ms, "
            f"{st['dead_paths']} paths, {st['iterations']} bounces, "
            f"{st['host_waits']} host waits, {st['device_drains']} drain")
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t_all
    r.set_measurement(profiling=True, counting=False)
    ptotals = {"dead_paths": 0, "ext_rays": 0, "shadow_rays": 0, "iterations": 0}
    t_prof = time.perf_counter()
    for k in range(args.steps):
        r.reset()
        st = r.render_all()
        for key in ptotals:
            ptotals[key] += st[key]
    torch.cuda.synchronize()
    prof_elapsed = time.perf_counter() - t_prof
    phases = r.phase_times()
    r.set_measurement(profiling=False, counting=False)
    log(f"{label} profiled frames: {prof_elapsed * 1e3 / args.steps:.1f} ms per frame")
    return {"cst": cst, "totals": totals, "elapsed": elapsed, "ptotals": ptotals,
            "phases": phases, "prof_elapsed": prof_elapsed}


def _trace_roofline(m, width, kernel):
    """KTrace's algorithmic GB/s from a _timed_render measurement."""
    cst, phases = m["cst"], m["phases"]
    nodes_per_ray = cst["node_visits"] / max(1, cst["ext_rays"])
    tris_per_ray = cst["tri_tests"] / max(1, cst["ext_rays"])
    trace_ms, trace_launches = phases["trace"]
    trace_avg_ms = trace_ms / max(1, trace_launches)
    rays_per_launch = m["ptotals"]["ext_rays"] / max(1, trace_launches)
    bytes_per_ray = RAY_BYTES + nodes_per_ray * NODE_BYTES[width] + tris_per_ray * TRI_BYTES
    achieved = rays_per_launch * bytes_per_ray / (trace_avg_ms * 1e-3) / 1e9
    return {
        "bound": "hbm",
        "kernel": kernel,
        "achieved": round(achieved, 1),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4),
        "traffic": None,
        "bytes_per_ray": round(bytes_per_ray, 1),
        "nodes_per_ray": round(nodes_per_ray, 2),
        "tris_per_ray": round(tris_per_ray, 2),
        "avg_launch_ms": round(trace_avg_ms, 4),
        "rays_per_launch": round(rays_per_launch, 1),
    }


def bench_instanced_cache(args):
    """Batched + cache: the instanced two-level workload's 8 uneven viewports
    in one renderer, rendered frame by frame (1 spp of every view per frame),
    one radiance cache per viewport trained online after every frame, against
    plain path tracing with the same frames; both judged per view by a plain
    render of --inst-cache-ref-spp spp (frame order).  Spectral, depth 12,
    NEE+MIS, lights chosen by power."""
    import numpy as np
    from mray_amd import abi, native, scenes
    torch = native.require_gpu()
    meshes, inst, env = scenes.instanced_workload(width=args.width, height=args.height)
    accs = [native.Accel(m) for m in meshes]
    isc = native.InstancedScene(accs, inst, env)
    for a in accs:
        a.close()
    views = scenes.instanced_views(spp=1, count=args.inst_views)
    pix = np.array([v["width"] * v["height"] for v in views], np.float64)
    params = abi.render_params(1, 1, 1, rr=(12, 12), mode=abi.SAMPLE_NEE_MIS, spectral=not args.rgb,
                               seed=0, max_ray_count=args.batch_slots, flags=abi.RENDER_LIGHT_POWER)
    t0 = time.perf_counter()
    r = native.BatchRenderer(isc, views, params)
    done = 0
    while done < args.inst_cache_ref_spp:
        step = min(64, args.inst_cache_ref_spp - done)
        r.render_frame(step)
        done += step
    ref = [film_image(f) for f in r.film()]
    r.close()
    log(f"instanced cache leg: judge {args.inst_cache_ref_spp} spp in {time.perf_counter() - t0:.1f} s")

    def err(films):
        e = np.array([relmse(film_image(f), g) for f, g in zip(films, ref)])
        return e, float((e * pix).sum() / pix.sum())

    def run(cached):
        params.seed = 1
        r = native.BatchRenderer(isc, views, params)
        caches = []
        if cached:
            caches = [native.Nrc(seed=11 + k, lazy_adam=True) for k in range(len(views))]
            r.attach_nrcs(caches, train_shift=4, spread=1.0, batch=1 << 14, steps=4, min_depth=2,
                          nee_at_end=True)
        curve = []
        spent, nxt = 0.0, 1
        for f in range(1, args.inst_cache_frames + 1):
            torch.cuda.synchronize()
            t = time.perf_counter()
            r.render_frame(1)
            torch.cuda.synchronize()
            spent += time.perf_counter() - t
            if f == nxt or f == args.inst_cache_frames:
                per, avg = err(r.film())
                curve.append({"frames": f, "seconds": round(spent, 4), "relmse": avg,
                              "relmse_per_view": [round(float(x), 6) for x in per]})
                log(f"instanced cache leg: {'cached' if cached else 'plain'} frame {f}: "
                    f"{spent:.3f} s, relMSE {avg:.4e}")
                nxt *= 2
        cs = r.cache_stats() if cached else None
        r.close()
        for c in caches:
            c.close()
        return curve, cs

    plain, _ = run(False)
    cached, cs = run(True)
    isc.close()
    e_p, t_p = plain[-1]["relmse"], plain[-1]["seconds"]
    e_c, t_c = cached[-1]["relmse"], cached[-1]["seconds"]
    e_p_eq = _loglog_at(plain, "seconds", "relmse", t_c)
    if e_p_eq is None:
        e_p_eq = e_p * t_p / t_c
    t_reach = _loglog_at(sorted(cached, key=lambda c: c["relmse"]), "relmse", "seconds", e_p)
    return {
        "workload": f"instanced two-level scene, {len(views)} viewports "
                    f"({int(pix.sum())} pixels), 1 spp of every view per frame, "
                    f"{args.inst_cache_frames} frames, spectral, depth 12, NEE+MIS, "
                    f"one cache per viewport (16 x 2^19 x 2 grid, 5 x 64 MLP, 4 x 2^14 pairs "
                    f"per frame each)",
        "judge_spp": args.inst_cache_ref_spp,
        "plain": plain,
        "cached": cached,
        "cache_stats": cs,
        "relmse_plain": e_p,
        "relmse_cached": e_c,
        "relmse_plain_equal_time": e_p_eq,
        "time_ratio_to_reach_plain_final": None if t_reach is None else round(t_reach / t_p, 4),
    }


def bench_instanced(args):
    """Two-level workload: 8 terrain meshes (50k-200k triangles each) placed
    1000 times with seeded rotations about +Y, scales and translations, a
    panel of 10k emissive triangles, constant Skysphere; spectral, depth 12.
    Main line: 8 viewports of uneven sizes (1080p down to 512^2) batched in
    one renderer; also one 1080p view alone.  NEE picks lights by power
    through an alias table."""
    import numpy as np
    from mray_amd import abi, native, scenes
    torch = native.require_gpu()
    spectral = not args.rgb
    spp = args.inst_spp
    log("generating instanced scene")
    t0 = time.perf_counter()
    meshes, inst, env = scenes.instanced_workload(width=args.width, height=args.height)
    log(f"meshes: {[m.triangle_count for m in meshes]} in {time.perf_counter() - t0:.1f} s")
    t0 = time.perf_counter()
    accs = [native.Accel(m) for m in meshes]
    isc = native.InstancedScene(accs, inst, env)
    for a in accs:
        a.close()
    torch.cuda.synchronize()
    build_s = time.perf_counter() - t0
    info = isc.info()
    log(f"instanced scene: {info} in {build_s:.2f} s")
    views = scenes.instanced_views(spp=spp, count=args.inst_views)
    paths = sum(v["width"] * v["height"] * v["spp"] for v in views)
    params = abi.render_params(1, 1, 1, rr=(12, 12), mode=abi.SAMPLE_NEE_MIS, spectral=spectral,
                               seed=0, max_ray_count=args.batch_slots, flags=abi.RENDER_LIGHT_POWER)
    r = native.BatchRenderer(isc, views, params)
    m = _timed_render(r, args, torch, "viewports")
    totals, elapsed, phases = m["totals"], m["elapsed"], m["phases"]
    lat = r.view_times()
    films = r.film()
    r.close()
    assert totals["dead_paths"] == paths * args.steps, totals
    assert all(np.isfinite(f).all() for f in films)
    # one 1080p view alone (a slot per pixel)
    p1 = abi.render_params(args.width, args.height, spp, rr=(12, 12), mode=abi.SAMPLE_NEE_MIS,
                           spectral=spectral, seed=0, flags=abi.RENDER_LIGHT_POWER)
    r1 = native.InstancedRenderer(isc, p1)
    m1 = _timed_render(r1, args, torch, "single view")
    tot1, el1, ph1 = m1["totals"], m1["elapsed"], m1["phases"]
    r1.close()
    isc.close()
    single_paths = args.width * args.height * spp
    assert tot1["dead_paths"] == single_paths * args.steps, tot1
    result = {
        "metric": f"Mpaths/s, {len(views)} viewports of an instanced two-level scene in one "
                  f"renderer, 1 MI355X",
        "value": round(paths * args.steps / elapsed / 1e6, 3),
        "unit": "Mpaths/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(elapsed * 1e3 / args.steps, 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded procedural meshes and instances)",
        "config": {
            "workload": f"{info['instances']} instances of {len(meshes) - 1} terrain meshes "
                        f"({info['effective_triangles']} effective tris) + "
                        f"{info['lights']} emissive tris; {len(views)} viewports "
                        f"(1080p..512^2), {spp} spp, depth 12, "
                        f"{'spectral 4 lambda' if spectral else 'RGB'}, NEE+MIS, "
                        f"lights chosen by power (alias table)",
            "mesh_triangles": [m.triangle_count for m in meshes],
            "effective_triangles": info["effective_triangles"],
            "tlas_nodes": info["tlas_nodes"],
            "views": [[v["width"], v["height"]] for v in views],
            "paths_per_step": paths,
            "slots": params.maxRayCount,
        },
        "mrays_per_s": round((totals["ext_rays"] + totals["shadow_rays"]) / elapsed / 1e6, 3),
        "bounces_per_step": totals["iterations"] // args.steps,
        "accel_build_s": round(build_s, 3),
        "phase_ms_per_step": {k: round(v[0] / args.steps, 3) for k, v in phases.items()},
        "view_latency_ms": [round(float(x), 3) for x in lat],
        "host_waits_per_step": totals["host_waits"] / args.steps,
        "device_drains_per_step": totals["device_drains"] / args.steps,
        "profiled_ms_per_step": round(m["prof_elapsed"] * 1e3 / args.steps, 2),
        "roofline": _trace_roofline(m, 4, "KTrace (closest hit, instance tree + 4-wide mesh trees)"),
        "kernel_rooflines": kernel_rooflines(m["cst"], m["ptotals"], phases, 4, spectral, False),
        "single_view": {
            "resolution": [args.width, args.height],
            "mpaths_per_s": round(single_paths * args.steps / el1 / 1e6, 3),
            "ms_per_step": round(el1 * 1e3 / args.steps, 2),
            "phase_ms_per_step": {k: round(v[0] / args.steps, 3) for k, v in ph1.items()},
            "roofline": _trace_roofline(m1, 4, "KTrace (instanced)"),
        },
    }
    return result


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU; multi-GPU sharding is "
                                   "designed (one film reduction per frame) but not run"}))
        sys.exit(2)
    if args.workload == "batch":
        print(json.dumps(bench_batch(args)), flush=True)
        return
    if args.workload == "instanced":
        print(json.dumps(bench_instanced(args)), flush=True)
        return
    if args.nrc_only:
        print(json.dumps({"nrc": nrc_equal_time(args)}), flush=True)
        return
    import numpy as np
    from mray_amd import abi, native, scenes
    torch = native.require_gpu()
    if args.probe_copy_mib:
        probe_copy(torch, args.probe_copy_mib)
    spectral = not args.rgb
    log("generating scene")
    t0 = time.perf_counter()
    scene = scenes.terrain_spheres(grid=args.grid, spheres=args.spheres, width=args.width,
                                   height=args.height)
    log(f"scene: {scene.triangle_count} triangles in {time.perf_counter() - t0:.1f} s")
    # build times of the two builders: the all-GPU LBVH (Karras tree,
    # collapse and soup on the device) and the binned SAH tree (host)
    build_times = {}
    t0 = time.perf_counter()
    accel = native.Accel(scene, builder=args.builder, width=args.bvh_width)
    torch.cuda.synchronize()
    build_s = time.perf_counter() - t0
    build_times[args.builder] = round(build_s, 3)
    if args.builder != "lbvh":
        t0 = time.perf_counter()
        probe = native.Accel(scene, builder="lbvh", width=4)
        torch.cuda.synchronize()
        build_times["lbvh_gpu"] = round(time.perf_counter() - t0, 3)
        probe.close()
    info = accel.info()
    log(f"accel: {info} in {build_s:.2f} s; build times {build_times}")
    params = abi.render_params(args.width, args.height, args.spp, rr=(10, 10),
                               mode=abi.SAMPLE_NEE_MIS, spectral=spectral, seed=0,
                               flags=(abi.RENDER_MATERIAL_QUEUES if args.queues else 0)
                               | (abi.RENDER_LANE_REFILL if args.lane_refill else 0))
    r = native.Renderer(accel, params)
    pixels = args.width * args.height
    paths_per_frame = pixels * args.spp

    m = _timed_render(r, args, torch, "frame")
    cst, totals, ptotals, phases = m["cst"], m["totals"], m["ptotals"], m["phases"]
    elapsed, prof_elapsed = m["elapsed"], m["prof_elapsed"]
    film = r.film()
    assert totals["dead_paths"] == paths_per_frame * args.steps, totals
    assert ptotals["dead_paths"] == paths_per_frame * args.steps, ptotals
    assert np.isfinite(film).all()
    nodes_per_ray = cst["node_visits"] / max(1, cst["ext_rays"])
    tris_per_ray = cst["tri_tests"] / max(1, cst["ext_rays"])

    mpaths = totals["dead_paths"] / elapsed / 1e6
    mrays = (totals["ext_rays"] + totals["shadow_rays"]) / elapsed / 1e6
    trace_ms, trace_launches = phases["trace"]
    trace_avg_ms = trace_ms / max(1, trace_launches)
    rays_per_launch = ptotals["ext_rays"] / max(1, trace_launches)
    bytes_per_ray = RAY_BYTES + nodes_per_ray * NODE_BYTES[args.bvh_width] + tris_per_ray * TRI_BYTES
    achieved = rays_per_launch * bytes_per_ray / (trace_avg_ms * 1e-3) / 1e9
    traffic = traffic_from_profiles(rays_per_launch)
    result = {
        "metric": METRIC,
        "value": round(mpaths, 3),
        "unit": "Mpaths/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(elapsed * 1e3 / args.steps, 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded procedural scene)",
        "config": {
            "workload": f"terrain{args.grid}x{args.grid}+{args.spheres} icospheres "
                        f"({scene.triangle_count} tris), {args.width}x{args.height}, "
                        f"{args.spp} spp, depth 10, "
                        f"{'spectral 4 lambda' if spectral else 'RGB'}, NEE+MIS",
            "triangles": scene.triangle_count,
            "resolution": [args.width, args.height],
            "spp": args.spp,
            "max_depth": 10,
            "spectral": spectral,
            "paths_per_step": paths_per_frame,
        },
        "mrays_per_s": round(mrays, 3),
        "bounces_per_step": totals["iterations"] // args.steps,
        "host_waits_per_step": totals["host_waits"] / args.steps,
        "device_drains_per_step": totals["device_drains"] / args.steps,
        "timed_path": "hipGraph bursts of 8 bounces, scheduler snapshots read 2 bursts behind, "
                      "one device drain per frame",
        "lane_refill": bool(args.lane_refill),
        "profiled_ms_per_step": round(prof_elapsed * 1e3 / args.steps, 2),
        "accel_build_s": round(build_s, 3),
        "accel_build_s_by_builder": build_times,
        "phase_ms_per_step": {k: round(v[0] / args.steps, 3) for k, v in phases.items()},
        "roofline": {
            "bound": "hbm",
            "kernel": f"KTrace (closest hit, {args.bvh_width}-wide BVH)",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": None if traffic is None else round(traffic),
            "bytes": "algorithmic: every node and triangle fetch counted; most are served by "
                     "L2 / Infinity Cache, the DRAM share is `traffic`",
            "dram_frac": None if traffic is None
            else round(traffic / (trace_avg_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
            "bytes_per_ray": round(bytes_per_ray, 1),
            "nodes_per_ray": round(nodes_per_ray, 2),
            "tris_per_ray": round(tris_per_ray, 2),
            "avg_launch_ms": round(trace_avg_ms, 4),
            "rays_per_launch": round(rays_per_launch, 1),
        },
        "kernel_rooflines": kernel_rooflines(cst, ptotals, phases, args.bvh_width, spectral,
                                             args.queues),
    }
    r.close()
    accel.close()
    if not args.no_cpu:
        log("cpu baseline")
        result["cpu_baseline"] = cpu_baseline(scene, params, args.cpu_seconds)
    legs = set(args.legs.split(",")) if args.legs else set()
    if "batch" in legs:
        log("batched leg")
        result["batch"] = bench_batch(args)
    if "instanced" in legs:
        log("instanced leg")
        result["instanced"] = bench_instanced(args)
    if "instanced_cache" in legs:
        log("instanced + per-viewport cache leg")
        result["instanced_cache"] = bench_instanced_cache(args)
    if args.nrc:
        log("radiance-cache equal-time leg")
        result["nrc"] = nrc_equal_time(args)
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
