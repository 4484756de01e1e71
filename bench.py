"""Benchmark of the rendering hot path on one MI355X.

Step = one closest-hit traversal launch (the speculative while-while kernel; the
exact per-lane kernel is timed beside it) over the whole ray set of the traversal
workload (1M-triangle soup SBVH, 8,294,400 rays: 4,147,200 coherent primary rays
from a 1920x1080 pinhole at 2 spp + 4,147,200 incoherent cosine-hemisphere rays),
rays and hits resident in HBM. Prints one JSON line.

    python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--variant", type=int, default=1,
                    help="1: speculative while-while (headline); 0: exact per-lane order")
    ap.add_argument("--lds-stack", type=int, default=16)
    ap.add_argument("--block", type=int, default=256)
    ap.add_argument("--refill", type=int, default=32,
                    help="persistent waves take new rays once this many lanes are idle (0: one wave per group)")
    ap.add_argument("--reorder", type=int, default=1, help="1: rays traced in Morton order of their origins")
    ap.add_argument("--xcd-order", type=int, default=0, help="1: each XCD holds one contiguous range of waves")
    ap.add_argument("--ntris", type=int, default=1_000_000)
    ap.add_argument("--cpu-sample-stride", type=int, default=1)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-render", action="store_true", help="skip the render workloads (secondary metrics)")
    ap.add_argument("--no-counters", action="store_true",
                    help="skip the rocprofv3 counter runs (traffic then from the committed counter profile); "
                         "for running bench.py itself under a profiler")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(scenes are replicated per GPU and rays shard without exchange). "
              "Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    import numpy as np

    import trenchanttracer_amd as tt
    from trenchanttracer_amd import workloads

    mesh, bvh, rays, n_prim = workloads.traversal(ntris=args.ntris, log=log)
    n = len(rays)
    ctx = tt.Context(0)
    scene = tt.DeviceScene(ctx, mesh, bvh)
    info = scene.info()
    log(f"scene on device: {info}")

    d_rays = tt.DeviceBuffer.from_array(ctx, rays)
    d_hits = tt.DeviceBuffer(ctx, n * 16)
    cfg = tt.trace_cfg(variant=args.variant, lds_stack=args.lds_stack, block=args.block, refill=args.refill,
                       xcd_order=args.xcd_order, reorder=args.reorder)

    for _ in range(args.warmup):
        scene.trace_closest(d_rays, n, d_hits, cfg)
    ctx.sync()
    log(f"warmup done ({args.warmup} steps)")

    # HIP events on the context stream, the stream the traversal kernel runs on
    timer = tt.Timer(ctx, 2 * args.steps)
    ctx.sync()
    t0 = time.perf_counter()
    for i in range(args.steps):
        timer.record(2 * i)
        scene.trace_closest(d_rays, n, d_hits, cfg)
        timer.record(2 * i + 1)
    ctx.sync()
    wall = time.perf_counter() - t0
    kernel_ms = [timer.elapsed(2 * i, 2 * i + 1) for i in range(args.steps)]
    kmean = float(np.mean(kernel_ms))
    ms_per_step = wall * 1e3 / args.steps
    mrays = n * args.steps / wall / 1e6
    log(f"timed {args.steps} steps: {ms_per_step:.3f} ms/step wall, kernel {kmean:.3f} ms avg, {mrays:.1f} Mrays/s")

    hits = d_hits.download((n, 4), np.int32)
    hit_frac = float((hits[:, 0] >= 0).mean())

    # the other traversal variant timed beside the headline one (exact per-lane
    # order if the headline is speculative, and vice versa); its hits are compared
    other = 1 - args.variant
    other_cfg = tt.trace_cfg(variant=other, lds_stack=args.lds_stack, block=args.block, refill=args.refill,
                             xcd_order=args.xcd_order, reorder=args.reorder)
    d_hits2 = tt.DeviceBuffer(ctx, n * 16)
    scene.trace_closest(d_rays, n, d_hits2, other_cfg)
    other_ms = []
    for i in range(min(args.steps, 5)):
        timer.record(0)
        scene.trace_closest(d_rays, n, d_hits2, other_cfg)
        timer.record(1)
        other_ms.append(timer.elapsed(0, 1))
    hits2 = d_hits2.download((n, 4), np.int32)
    spec = {"variant": "exact" if other == 0 else "speculative",
            "mrays_kernel": round(n / float(np.median(other_ms)) / 1e3, 1),
            "hit_id_diffs_vs_headline": int((hits2[:, 0] != hits[:, 0]).sum()),
            "hits_bit_identical_to_headline": bool(np.array_equal(hits2, hits))}
    d_hits2.free()
    log(f"other variant: {spec}")

    # each half of the ray set (coherent camera rays, incoherent surface rays) as a
    # call of its own in the same configuration (each call sorts its own rays)
    halves = {}
    for name, lo, hi in (("coherent", 0, n_prim), ("incoherent", n_prim, n)):
        d_half = tt.DeviceBuffer.from_array(ctx, np.ascontiguousarray(rays[lo:hi]))
        d_hh = tt.DeviceBuffer(ctx, (hi - lo) * 16)
        scene.trace_closest(d_half, hi - lo, d_hh, cfg)
        ms_h = []
        for i in range(min(args.steps, 10)):
            timer.record(0)
            scene.trace_closest(d_half, hi - lo, d_hh, cfg)
            timer.record(1)
            ms_h.append(timer.elapsed(0, 1))
        mh = float(np.median(ms_h))
        halves[name] = {"rays": hi - lo, "ms": round(mh, 4), "mrays_per_s": round((hi - lo) / mh / 1e3, 1)}
        d_half.free()
        d_hh.free()
    log(f"halves: {halves}")

    # the traversal's own loop replayed without its box and triangle arithmetic, in
    # the headline's launch form, after the same key pass and sort
    # (tt_bench_fetch_replay); the whole set and each half
    # (the exact loop's counting launch gives the algorithmic bytes: its node visits
    # and triangle tests are the single-lane traversal's)
    exact_cfg = tt.trace_cfg(variant=0, lds_stack=args.lds_stack, block=args.block, refill=args.refill,
                             xcd_order=args.xcd_order, reorder=args.reorder)
    replay = scene.fetch_replay(d_rays, n, cfg, reps=5)
    algo = scene.fetch_replay(d_rays, n, exact_cfg, reps=1)
    log(f"fetch replay of the whole set: {replay['ms']:.3f} ms (call {kmean:.3f} ms)")
    for name, lo, hi in (("coherent", 0, n_prim), ("incoherent", n_prim, n)):
        d_half = tt.DeviceBuffer.from_array(ctx, np.ascontiguousarray(rays[lo:hi]))
        rh = scene.fetch_replay(d_half, hi - lo, cfg, reps=5)
        ah = scene.fetch_replay(d_half, hi - lo, exact_cfg, reps=1)
        d_half.free()
        halves[name].update({
            "algorithmic_bytes_per_ray": round(ah["bytes"] / (hi - lo), 1),
            "effective_gbps": round(ah["bytes"] / (halves[name]["ms"] * 1e-3) / 1e9, 1),
            "node_visits_per_ray": round(ah["node_fetches"] / (hi - lo), 2),
            "tri_tests_per_ray": round(ah["tri_fetches"] / (hi - lo), 2),
            "replay_ms": round(rh["ms"], 4),
            "replay_over_call": round(rh["ms"] / halves[name]["ms"], 4)})
    log(f"halves with their replays: {halves}")

    # fetch-throughput ceiling: independent random 64-byte reads over a buffer the size
    # of the device tree (the same probe the counter runs measure)
    from trenchanttracer_amd.profile_trace import GATHER_FETCHES
    probe = tt.DeviceBuffer(ctx, info["device_bytes"])
    probe.zero()
    gather_ms = []
    for _ in range(3):
        gbps = tt.gather64_gbps(ctx, probe, info["device_bytes"], fetches=GATHER_FETCHES, seed=1)
        gather_ms.append(GATHER_FETCHES * 64 / (gbps * 1e9) * 1e3)
    probe.free()
    gather_ms = float(np.median(gather_ms))
    log(f"random 64-byte gather over {info['device_bytes'] / 1e6:.0f} MB: {gather_ms:.3f} ms for 2^26 reads")

    traffic, traffic_src = (None, "--no-counters") if args.no_counters else pmc_call_traffic(args)
    if traffic is None:
        log(f"live counter runs unavailable ({traffic_src}); using the committed counter profile")
        traffic, traffic_src = committed_traffic(args)
    bpr = algo["bytes"] / n
    algo_gbps = algo["bytes"] / (kmean * 1e-3) / 1e9
    # node-fetch roofline: the call's beyond-L2 bytes per second against those of the
    # random-gather probe (both from the same counter runs); and the call's own loop
    # replayed without its arithmetic (tt_bench_fetch_replay)
    node_fetch = {"replay_ms": round(replay["ms"], 4), "call_ms": round(kmean, 4),
                  "replay_over_call": round(replay["ms"] / kmean, 4),
                  "gather_ms": round(gather_ms, 4)}
    if traffic is not None and traffic.get("gather_bytes"):
        g_rate = traffic["gather_bytes"] / (gather_ms * 1e-3) / 1e9
        c_rate = traffic["call_bytes"] / (kmean * 1e-3) / 1e9
        node_fetch.update({"gather_traffic_gbps": round(g_rate, 1), "call_traffic_gbps": round(c_rate, 1),
                           "frac": round(c_rate / g_rate, 4)})
    roofline = None
    if traffic is not None:
        achieved = traffic["call_bytes"] / (kmean * 1e-3) / 1e9
        roofline = {
            "bound": "hbm", "achieved": round(achieved, 1), "peak": 8000.0, "unit": "GB/s",
            "frac": round(achieved / 8000.0, 4),
            "traffic": round(traffic["call_bytes"] / 1e9, 3),
            "traffic_unit": "GB per tt_trace_closest call (HBM side of L2: FETCH_SIZE + WRITE_SIZE)",
            "traffic_source": traffic_src,
            "traffic_traversal_kernel": round(traffic["trace_bytes"] / 1e9, 3),
            "timed_span": "whole tt_trace_closest call (Morton keys + radix sort + traversal), HIP events"
                          if args.reorder else "traversal launch, HIP events",
        }
    else:
        roofline = {"bound": "hbm", "achieved": None, "peak": 8000.0, "unit": "GB/s", "frac": None,
                    "traffic": None, "traffic_source": traffic_src}
    roofline.update({
        "algorithmic_bytes_per_ray": round(bpr, 1),
        "effective_gbps": round(algo_gbps, 1),
        "algorithmic_over_traffic": round(algo["bytes"] / traffic["call_bytes"], 3) if traffic else None,
        "node_visits_per_ray": round(algo["node_fetches"] / n, 2),
        "tri_tests_per_ray": round(algo["tri_fetches"] / n, 2),
        "speculative_node_visits_per_ray": round(replay["node_fetches"] / n, 2),
        "node_fetch": node_fetch,
        "per_half": halves,
    })

    result = {
        "metric": "Mrays/s closest-hit",
        "value": round(mrays, 2),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"traversal: {args.ntris}-triangle log-normal soup, SBVH alpha 1e-5, {n} rays "
                        f"({n_prim} primary 1920x1080x2spp + {n - n_prim} incoherent), closest hit",
            "rays": n, "triangles": args.ntris, "tri_refs": info["num_tris"], "nodes": info["num_nodes"],
            "variant": "exact" if args.variant == 0 else "speculative", "lds_stack": args.lds_stack,
            "block": args.block, "refill": args.refill, "reorder": args.reorder, "xcd_order": args.xcd_order,
            "hit_fraction": round(hit_frac, 4),
        },
        "kernel_ms": round(kmean, 4),
        "other_variant": spec,
        "roofline": roofline,
        "cpu_baseline": None,
    }

    if not args.no_cpu_baseline:
        # cpu_baseline leg: the oracle (single-lane restatement of the reference
        # traversal) timed on host cores over a strided sample of the same rays;
        # its hits check the GPU's on the sample, and its visit counts the
        # algorithmic bytes of the node-fetch probe.
        from oracle import oracle as O

        stride = args.cpu_sample_stride
        cb = O.CompactBVH(bvh.export_compact())
        cores = min(16, os.cpu_count() or 1)
        # the two halves traced separately (the oracle's visit counts cross-check the
        # recorded fetch counts of the node-fetch probe)
        parts, idxs, st = [], [], {"seconds": 0.0, "inner": 0, "tris": 0}
        for name, lo, hi in (("coherent", 0, n_prim), ("incoherent", n_prim, n)):
            ih = np.arange(lo, hi, stride)
            oh_h, st_h = cb.trace(rays[ih], nthreads=cores)
            parts.append(oh_h)
            idxs.append(ih)
            for key in st:
                st[key] += st_h[key]
        idx = np.concatenate(idxs)
        sample = rays[idx]
        oh = np.concatenate(parts)
        cpu_mrays = len(sample) / st["seconds"] / 1e6
        bpr_o = workloads.bytes_per_ray(st["inner"], st["tris"], len(sample))
        gpu_ids = hits[idx, 0]
        o_ids = O.hit_ids(oh)
        mism = int((gpu_ids != o_ids).sum())
        exact = bool(np.array_equal(hits[idx].view(np.int32), oh.view(np.int32)))
        result["cpu_baseline"] = {
            "value": round(cpu_mrays, 3), "unit": "Mrays/s", "cores": cores, "kind": "port",
            "sample": f"every {stride}th ray of the same set ({len(sample)} rays), oracle single-lane traversal, "
                      f"{st['seconds']:.2f}s wall on {cores} threads",
            "host": host_cpu(),
        }
        result["parity_sample"] = {"rays": int(len(sample)), "hit_id_mismatches": mism, "bit_exact": exact,
                                   "oracle_algorithmic_bytes_per_ray": round(bpr_o, 1)}
        log(f"cpu baseline {cpu_mrays:.2f} Mrays/s on {cores} threads; sample mismatches {mism}; "
            f"oracle bytes/ray {bpr_o:.1f} (probe {bpr:.1f})")

    scene.close()
    d_rays.free()
    d_hits.free()
    if not args.no_render:
        result["render"] = render_bench(tt, ctx, np, args.no_cpu_baseline)

    print(json.dumps(result), flush=True)
    ctx.close()


PMC_COPY_BYTES = 256 << 20   # the calibration copy of trenchanttracer_amd.profile_trace


def pmc_call_traffic(args, timeout=240):
    """HBM-side bytes of one tt_trace_closest call, measured in two child runs of
    trenchanttracer_amd.profile_trace (the same workload and configuration; one call,
    then a 256 MiB device copy) under rocprofv3 --pmc FETCH_SIZE, then WRITE_SIZE.
    Each counter is scaled by what it reads for the copy (known 256 MiB read and
    written). Returns ({call_bytes, trace_bytes}, source) or (None, reason)."""
    import csv
    import glob
    import shutil
    import signal
    import subprocess
    import tempfile
    rocprof = shutil.which("rocprofv3") or "/opt/rocm/bin/rocprofv3"
    if not os.path.exists(rocprof):
        return None, "rocprofv3 not found"
    prog = [sys.executable, "-m", "trenchanttracer_amd.profile_trace", "--variant", str(args.variant),
            "--lds-stack", str(args.lds_stack), "--block", str(args.block), "--refill", str(args.refill),
            "--reorder", str(args.reorder), "--ntris", str(args.ntris)]
    env = dict(os.environ, TMPDIR="/tmp", PYTHONPATH=ROOT + os.pathsep + os.environ.get("PYTHONPATH", ""))
    per = {}
    for counter in ("FETCH_SIZE", "WRITE_SIZE"):
        out = tempfile.mkdtemp(prefix="tt_pmc_", dir="/tmp")
        cmd = [rocprof, "--pmc", counter, "--output-format", "csv", "-d", out, "-o", "run", "--"] + prog
        log(f"counter run: {counter}")
        p = subprocess.Popen(cmd, cwd="/tmp", env=env, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True,
                             start_new_session=True)
        try:
            text, _ = p.communicate(timeout=timeout)
        except subprocess.TimeoutExpired:
            os.killpg(p.pid, signal.SIGKILL)
            p.communicate()
            shutil.rmtree(out, ignore_errors=True)
            return None, f"{counter} run exceeded {timeout} s"
        files = glob.glob(os.path.join(out, "**", "*counter_collection.csv"), recursive=True)
        rows = []
        for f in files:
            with open(f) as fh:
                rows += [r for r in csv.DictReader(fh) if r.get("Counter_Name") == counter]
        shutil.rmtree(out, ignore_errors=True)
        if not rows:
            return None, f"{counter}: no counter rows (rc {p.returncode}): {text[-300:]!r}"
        rows.sort(key=lambda r: int(r["Dispatch_Id"]))
        per[counter] = [(r["Kernel_Name"], float(r["Counter_Value"])) for r in rows]
    kib = {}
    for counter, rows in per.items():
        copy = [v for k, v in rows if "k_copy16" in k]
        gather = [v for k, v in rows if "k_gather64" in k]
        # the call: its dispatches up to its traversal launch (the buffer fills after
        # it belong to the gather probe and the copy)
        last = max((i for i, (k, _) in enumerate(rows) if "k_trace" in k), default=-1)
        call = rows[:last + 1]
        if len(copy) != 1 or len(gather) != 1 or not call:
            return None, f"{counter}: unexpected dispatches {[k for k, _ in rows]}"
        kib[counter] = (copy[0], call, gather[0])
    sr = PMC_COPY_BYTES / 1024 / kib["FETCH_SIZE"][0]
    sw = PMC_COPY_BYTES / 1024 / kib["WRITE_SIZE"][0]
    call = sum(v for _, v in kib["FETCH_SIZE"][1]) * sr + sum(v for _, v in kib["WRITE_SIZE"][1]) * sw
    trace = (sum(v for k, v in kib["FETCH_SIZE"][1] if "k_trace" in k) * sr +
             sum(v for k, v in kib["WRITE_SIZE"][1] if "k_trace" in k) * sw)
    gather = kib["FETCH_SIZE"][2] * sr + kib["WRITE_SIZE"][2] * sw
    src = (f"rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE child runs of trenchanttracer_amd.profile_trace in this "
           f"bench run ({len(kib['FETCH_SIZE'][1])} dispatches per call), scaled by a 256 MiB copy: "
           f"reads x {sr:.5f}, writes x {sw:.5f}")
    return {"call_bytes": call * 1024, "trace_bytes": trace * 1024, "gather_bytes": gather * 1024}, src


def host_cpu():
    """The host CPU model (from /proc/cpuinfo) and the logical CPUs the process may use."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        usable = os.cpu_count() or 1
    return f"{model}, {usable} logical CPUs available to the process"


def committed_traffic(args):
    """The counter profile committed for the headline configuration
    (profiles/k_trace_speculative_soup1M_pmc.md), used only when the live counter runs
    cannot run; None for any other configuration."""
    if (args.variant, args.lds_stack, args.block, args.refill, args.reorder, args.ntris) != (1, 16, 256, 32, 1,
                                                                                          1_000_000):
        return None, "no live counter runs and no committed profile for this configuration"
    sr = 262_144 / 131_079.5
    # keys, 3 scans, 3 scatters, 2 histograms, traversal
    fetch = 129_729.2 + 3_128.5 + 24_564.5 + 32_586.1 + 40_775.8 + 40_738.0 + 6_598_886.9
    write = 49_932.7 + 6_120.0 + 78_975.1 + 32_480.9 + 79_058.8 + 39_029.7 + 214_054.2
    return ({"call_bytes": (fetch * sr + write) * 1024, "trace_bytes": (6_598_886.9 * sr + 214_054.2) * 1024,
             "gather_bytes": (4_157_602.9 * sr + 16_384.0) * 1024},
            "profiles/k_trace_speculative_soup1M_pmc.md (committed counter runs)")


def render_bench(tt, ctx, np, no_cpu):
    """Secondary metrics: samples/s of the path tracer on the full path-trace case
    (1,310,720-tri displaced icosphere, 40/40/20 diffuse/GGX/dielectric, 2048x1024
    sky + sun, 1920x1080, 64 spp, max depth 16, Russian roulette from depth 4); of
    the reference-semantics render of the Cornell case (256x256, 16 spp, depth 8)
    beside the oracle's CPU restatement of it; of the batched case (32 scenes, 16 spp)
    against 32 sequential renders; and guided against plain path tracing on the
    path-trace case (equal-time RMSE against an FP64 8192-spp render)."""
    from trenchanttracer_amd import workloads
    out = {}
    mesh, mats, env, cam = workloads.pathtrace()
    bvh = tt.BVH.build(mesh)
    s = tt.DeviceScene(ctx, mesh, bvh, mats, env=env)
    prm = tt.render_params(spp=64, max_depth=16, rr_depth=4, seed=1, mode=tt.RENDER_PATH)
    s.render(cam, tt.render_params(spp=4, max_depth=16, rr_depth=4, seed=7, mode=tt.RENDER_PATH))
    img, st = s.render(cam, prm)
    rays = st["rays_extension"] + st["rays_shadow"]
    out["pathtrace_1080p_64spp"] = {
        "samples_per_s": round(st["samples"] / st["seconds"], 1), "seconds": round(st["seconds"], 4),
        "mrays_per_s": round(rays / st["seconds"] / 1e6, 1), "rays_extension": st["rays_extension"],
        "rays_shadow": st["rays_shadow"], "mean_radiance": round(float(img.mean()), 5)}
    log(f"path trace 1080p 64spp: {out['pathtrace_1080p_64spp']}")
    s.close()

    mesh, mats, cam = workloads.cornell()
    bvh = tt.BVH.build(mesh)
    s = tt.DeviceScene(ctx, mesh, bvh, mats)
    prm = tt.render_params(spp=16, max_depth=8, seed=1, mode=tt.RENDER_REFERENCE)
    s.render(cam, prm)
    img, st = s.render(cam, prm)
    c1 = {"samples_per_s": round(st["samples"] / st["seconds"], 1), "seconds": round(st["seconds"], 5),
          "rays_extension": st["rays_extension"]}
    if not no_cpu:
        from oracle import oracle as O
        cores = min(16, os.cpu_count() or 1)
        t0 = time.perf_counter()
        ref = O.render_ref(O.CompactBVH(bvh.export_compact()), mesh.arrays(), mats, cam, spp=16, max_depth=8,
                           seed=1, nthreads=cores)
        cpu_s = time.perf_counter() - t0
        c1["cpu_baseline"] = {"samples_per_s": round(st["samples"] / cpu_s, 1), "cores": cores, "kind": "port",
                              "seconds": round(cpu_s, 3)}
        c1["bit_identical_to_cpu"] = bool(np.array_equal(img.view(np.int32), ref.view(np.int32)))
    out["reference_cornell_256_16spp"] = c1
    log(f"reference cornell: {c1}")
    s.close()
    from trenchanttracer_amd import bench_batch
    out["batched_32_scenes_16spp"] = bench_batch.measure(tt, ctx, log=log)
    log(f"batched: {out['batched_32_scenes_16spp']}")
    out["guided_pathtrace_1080p"] = guided_bench(tt, ctx, np)
    return out


# Guided vertices of the path-trace case: the camera bounce's alone. Guiding deeper
# vertices too costs inference on every bounce's queue and, on this scene, buys little or
# nothing: with 4 seeds per point and the same protocol (bench_guide --sweep bounces),
# guiding the first 1 / 2 / all bounces reached the plain 64-spp RMSE in 1.41 / 1.44 /
# 1.53x the plain time trained online and 1.00 / 1.01 / 1.13x frozen. (The door room, lit
# through an opening, gains from every bounce: the library's default guides all of them.)
GUIDED_BOUNCES = 1


def guided_bench(tt, ctx, np):
    """Learned guiding on the path-trace case (1,310,720 tris, 1920x1080, depth 16):
    equal-time RMSE against an FP64-accumulated 8192-spp render of the plain tracer.
    Plain renders at 64 and 128 spp; guided renders (8-lobe vMF MLP, BF16 MFMA, Adam
    lr 1e-3 on 2^18-row batches, one-sample MIS with the BSDF, the selection
    probability learned per vertex, the mixture used at the camera bounce's vertices:
    GUIDED_BOUNCES) trained online from scratch throughout (passes growing 8x from
    1 spp, up to 2M recorded paths each; every bounce recorded), trained during the
    first half of the samples, and frozen after the 128-spp online run; two seeds per
    point.
    The time at which each reaches the plain 64-spp RMSE comes from a log-log fit of
    RMSE against time (trenchanttracer_amd/bench_guide.py, which also runs the longer
    sweep)."""
    from trenchanttracer_amd import bench_guide, workloads
    mesh, mats, env, cam = workloads.pathtrace()
    sc = tt.DeviceScene(ctx, mesh, tt.BVH.build(mesh, nthreads=16), mats, env=env)
    t0 = time.time()
    _, ref, st = sc.render(cam, tt.render_params(spp=8192, max_depth=16, rr_depth=4, seed=1234, mode=tt.RENDER_PATH,
                                                 fp64=True), want_fp64=True)
    log(f"guided leg: FP64 8192-spp oracle render in {time.time() - t0:.1f}s")
    r = bench_guide.compare(tt, ctx, sc, cam, ref, 16, (64, 128), (64, 128), 1920 * 1080, 1e-3, 262144, 2 << 20,
                            seeds=2, pass_growth=8, learn_alpha=True, max_bounce=GUIDED_BOUNCES)
    sc.close()
    p64 = next(x for x in r["plain"] if x["spp"] == 64)
    g64 = next(x for x in r["guided_online"] if x["spp"] == 64)
    f64 = next(x for x in r["guided_pretrained"] if x["spp"] == 64)
    out = {
        "oracle_spp": 8192, "seeds_per_point": 2, "guided_bounces": GUIDED_BOUNCES,
        "plain_64spp": {"rmse": round(p64["rmse"], 5), "seconds": round(p64["seconds"], 4)},
        "wall_ratio_to_plain_64spp_rmse": {
            "online": round(r["guided_online_wall_ratio"], 3),
            "online_first_half": round(r["guided_online_half_wall_ratio"], 3),
            "frozen_after_training": round(r["guided_pretrained_wall_ratio"], 3)},
        "target_wall_ratio": 0.5,
        "breakdown_64spp_s": {
            "plain": round(p64["seconds"], 4),
            "frozen_network_render": round(f64["seconds"], 4),
            "inference_and_guided_shading": round(f64["seconds"] - p64["seconds"], 4),
            "online_render": round(g64["seconds"], 4),
            "online_training": round(g64["train_seconds"], 4),
            "online_rendering_between_steps": round(g64["seconds"] - g64["train_seconds"], 4)},
        "rmse_64spp": {"plain": round(p64["rmse"], 5), "online": round(g64["rmse"], 5),
                       "frozen": round(f64["rmse"], 5)},
        "training": {"steps": g64["train_steps"], "last_loss": round(g64["last_loss"], 4)},
        "points": r,
    }
    log(f"guided path trace: { {k: v for k, v in out.items() if k != 'points'} }")
    return out


if __name__ == "__main__":
    main()
