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
    ap.add_argument("--ntris", type=int, default=1_000_000)
    ap.add_argument("--cpu-sample-stride", type=int, default=1)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-render", action="store_true", help="skip the render workloads (secondary metrics)")
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
                       reorder=args.reorder)

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
                             reorder=args.reorder)
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

    # measured node-fetch roofline: random 64-byte gathers over a buffer the size
    # of the device BVH (nodes + triangle records)
    probe = tt.DeviceBuffer(ctx, info["device_bytes"])
    gather_gbps = max(tt.gather64_gbps(ctx, probe, info["device_bytes"], seed=s) for s in (1, 2, 3))
    probe.free()
    log(f"random 64-byte gather over {info['device_bytes'] / 1e6:.0f} MB: {gather_gbps:.0f} GB/s")

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
            "block": args.block, "refill": args.refill, "reorder": args.reorder, "hit_fraction": round(hit_frac, 4),
        },
        "kernel_ms": round(kmean, 4),
        "other_variant": spec,
        "roofline": None,
        "cpu_baseline": None,
    }

    if not args.no_cpu_baseline:
        # cpu_baseline leg: the oracle (single-lane restatement of the reference
        # traversal) timed on host cores over a strided sample of the same rays;
        # its per-ray visit counts give the algorithmic bytes, and its hits check
        # the GPU's on the sample.
        from oracle import oracle as O

        stride = args.cpu_sample_stride
        cb = O.CompactBVH(bvh.export_compact())
        cores = min(16, os.cpu_count() or 1)
        # the two halves traced separately, so that each half's visit counts give its
        # own algorithmic bytes per ray
        parts, idxs, st = [], [], {"seconds": 0.0, "inner": 0, "tris": 0}
        for name, lo, hi in (("coherent", 0, n_prim), ("incoherent", n_prim, n)):
            ih = np.arange(lo, hi, stride)
            oh_h, st_h = cb.trace(rays[ih], nthreads=cores)
            parts.append(oh_h)
            idxs.append(ih)
            for key in st:
                st[key] += st_h[key]
            bh = workloads.bytes_per_ray(st_h["inner"], st_h["tris"], len(ih))
            gb = bh * halves[name]["rays"] / (halves[name]["ms"] * 1e-3) / 1e9
            halves[name].update({"bytes_per_ray": round(bh, 1), "achieved_gbps": round(gb, 1),
                                 "frac": round(gb / 8000.0, 4),
                                 "node_visits_per_ray": round(st_h["inner"] / len(ih), 2),
                                 "tri_tests_per_ray": round(st_h["tris"] / len(ih), 2)})
        idx = np.concatenate(idxs)
        sample = rays[idx]
        oh = np.concatenate(parts)
        cpu_mrays = len(sample) / st["seconds"] / 1e6
        bpr = workloads.bytes_per_ray(st["inner"], st["tris"], len(sample))
        gpu_ids = hits[idx, 0]
        o_ids = O.hit_ids(oh)
        mism = int((gpu_ids != o_ids).sum())
        exact = bool(np.array_equal(hits[idx].view(np.int32), oh.view(np.int32)))
        achieved = bpr * n / (kmean * 1e-3) / 1e9
        # traffic per launch of the traversal kernel in the default configuration
        # (speculative, LDS 16, block 256, refill 32, reorder 1), from rocprofv3 --pmc
        # FETCH_SIZE and WRITE_SIZE runs of trenchanttracer_amd.profile_trace scaled by
        # a 256 MiB copy's counts (reads x 1.99976, writes x 1.0): 6,597,741 KiB
        # fetched, 214,109 KiB written (the hits, scattered to the rays' own indices)
        default_cfg = (args.variant, args.lds_stack, args.block, args.refill, args.reorder,
                       args.ntris) == (1, 16, 256, 32, 1, 1_000_000)
        traffic = round((6_597_740.5 * 1.99976 + 214_108.7) * 1024 / 1e9, 2) if default_cfg else None
        result["roofline"] = {
            "bound": "hbm", "achieved": round(achieved, 1), "peak": 8000.0, "unit": "GB/s",
            "frac": round(achieved / 8000.0, 4), "traffic": traffic, "traffic_unit": "GB per launch",
            "timed_span": "whole tt_trace_closest call (Morton keys + radix sort + traversal)" if args.reorder
                          else "traversal launch",
            "bytes_per_ray": round(bpr, 1),
            "node_fetch_roofline_gbps": round(gather_gbps, 1),
            "node_fetch_frac": round(achieved / gather_gbps, 4),
            "node_visits_per_ray": round(st["inner"] / len(sample), 2),
            "tri_tests_per_ray": round(st["tris"] / len(sample), 2),
            "per_half": halves,
        }
        result["cpu_baseline"] = {
            "value": round(cpu_mrays, 3), "unit": "Mrays/s", "cores": cores, "kind": "port",
            "sample": f"every {stride}th ray of the same set ({len(sample)} rays), oracle single-lane traversal, "
                      f"{st['seconds']:.2f}s wall on {cores} threads",
        }
        result["parity_sample"] = {"rays": int(len(sample)), "hit_id_mismatches": mism, "bit_exact": exact}
        log(f"cpu baseline {cpu_mrays:.2f} Mrays/s on {cores} threads; sample mismatches {mism}; "
            f"bytes/ray {bpr:.1f}; achieved {achieved:.0f} GB/s")

    scene.close()
    d_rays.free()
    d_hits.free()
    if not args.no_render:
        result["render"] = render_bench(tt, ctx, np, args.no_cpu_baseline)

    print(json.dumps(result), flush=True)
    ctx.close()


def render_bench(tt, ctx, np, no_cpu):
    """Secondary metrics: samples/s of the path tracer on the full path-trace case
    (1,310,720-tri displaced icosphere, 40/40/20 diffuse/GGX/dielectric, 2048x1024
    sky + sun, 1920x1080, 64 spp, max depth 16, Russian roulette from depth 4), and
    of the reference-semantics render of the Cornell case (256x256, 16 spp, depth 8)
    beside the oracle's CPU restatement of it."""
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
    out["guided_door_room_256"] = guided_bench(tt, ctx, np)
    return out


def guided_bench(tt, ctx, np):
    """Learned guiding where light reaches the camera through a door (door room,
    256x256, depth 12): the network is trained online during a 1024-spp render,
    then plain and guided 64-spp renders are scored against an FP64 4096-spp
    render. Equal-time efficiency = (MSE x seconds) plain / (MSE x seconds) guided.
    The full equal-time protocol on the 1080p path-trace case (8192-spp oracle) is
    trenchanttracer_amd/bench_guide.py."""
    from trenchanttracer_amd import workloads
    mesh, mats, cam = workloads.door_room(256, 256)
    s = tt.DeviceScene(ctx, mesh, tt.BVH.build(mesh), mats)
    ref = s.render(cam, tt.render_params(spp=4096, max_depth=12, rr_depth=4, seed=1234, mode=tt.RENDER_PATH,
                                         fp64=True), want_fp64=True)[1]
    g = tt.Guide(ctx, seed=1, lr=1e-2)
    pt = tt.render_params(spp=1024, max_depth=12, rr_depth=4, seed=99, mode=tt.RENDER_PATH, paths_in_flight=131072)
    _, _, gs = g.render(s, cam, pt, alpha=0.5, train=True, train_paths=65536, train_batch=16384)
    p64 = tt.render_params(spp=64, max_depth=12, rr_depth=4, seed=5, mode=tt.RENDER_PATH)
    s.render(cam, p64)
    plain, sp = s.render(cam, p64)
    g.render(s, cam, p64, alpha=0.5, train=False)
    guided, sg, _ = g.render(s, cam, p64, alpha=0.5, train=False)
    mp, mg = float(np.mean((plain - ref) ** 2)), float(np.mean((guided - ref) ** 2))
    r = {"mse_plain_64spp": mp, "mse_guided_64spp": mg, "seconds_plain": round(sp["seconds"], 5),
         "seconds_guided": round(sg["seconds"], 5), "mse_ratio": round(mg / mp, 4),
         "equal_time_efficiency": round((mp * sp["seconds"]) / (mg * sg["seconds"]), 3),
         "training": {"steps": gs["train_steps"], "rows": gs["train_rows"],
                      "seconds": round(gs["train_seconds"], 4), "last_loss": round(gs["last_loss"], 4)}}
    log(f"guided door room: {r}")
    g.close()
    s.close()
    return r


if __name__ == "__main__":
    main()
