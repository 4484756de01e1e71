"""Batched-scenes measurement: 32 seeded scenes (displaced icospheres and soups,
1k-2M triangles, 256^2 / 512^2 / 1024^2), 16 spp path tracing, rendered (a) as 32
sequential tt_render calls and (b) in one batched wavefront with hipGraph-captured
passes and autotuned launch settings. Prints one JSON line.

    python -m trenchanttracer_amd.bench_batch [--scenes 32] [--spp 16]
"""
import argparse
import json
import sys
import time

import numpy as np


def log(m):
    print(f"[batch] {m}", file=sys.stderr, flush=True)


def measure(tt, ctx, scenes=32, spp=16, depth=8, tri_hi=2_000_000, nthreads=16, log=log):
    """The batched case (workloads.batched) rendered as `scenes` sequential tt_render calls
    and as one batched wavefront with autotuned settings; returns the result dict."""
    from trenchanttracer_amd import autotune, workloads

    t0 = time.time()
    meshes, mats, cams, env, counts = workloads.batched(scenes, tri_hi=tri_hi, log=log)
    t1 = time.time()
    bvhs = [tt.BVH.build(m, nthreads=nthreads) for m in meshes]
    t2 = time.time()
    log(f"generated {sum(counts)} tris in {t1 - t0:.1f}s, built {len(bvhs)} SBVHs in {t2 - t1:.1f}s")
    base = dict(spp=spp, max_depth=depth, rr_depth=4, seed=1, mode=tt.RENDER_PATH)

    # (a) sequential: one scene after another, default settings
    singles = [tt.DeviceScene(ctx, m, b, mm, env=env) for m, b, mm in zip(meshes, bvhs, mats)]
    for s, c in zip(singles, cams):                      # warm-up (state allocation)
        s.render(c, tt.render_params(**dict(base, spp=1)))
    seq_gpu = 0.0
    ts = time.perf_counter()
    for s, c in zip(singles, cams):
        _, st = s.render(c, tt.render_params(**base))
        seq_gpu += st["seconds"]
    seq_wall = time.perf_counter() - ts
    for s in singles:
        s.close()
    log(f"sequential: wall {seq_wall:.3f}s, GPU {seq_gpu:.3f}s")

    # (b) one batched wavefront, settings from the autotuner
    batch = tt.Batch(ctx, meshes, bvhs, mats, env=env)

    def run(prm):
        return batch.render(cams, prm)[1]

    tuned, tinfo = autotune.tune_render(run, base)
    log(f"autotuned: {tuned} ({tinfo})")
    prm = tt.render_params(**dict(base, **tuned))
    tb = time.perf_counter()
    imgs, st = batch.render(cams, prm)
    b_wall = time.perf_counter() - tb
    log(f"batched: wall {b_wall:.3f}s, GPU {st['seconds']:.3f}s, stats {st}")
    samples = st["samples"]
    out = {"scenes": scenes, "triangles": int(sum(counts)),
           "pixels": int(sum(c.width * c.height for c in cams)), "spp": spp, "max_depth": depth,
           "sbvh_build_s": round(t2 - t1, 2),
           "sequential": {"wall_s": round(seq_wall, 4), "gpu_s": round(seq_gpu, 4),
                          "samples_per_s": round(samples / seq_wall, 1)},
           "batched": {"wall_s": round(b_wall, 4), "gpu_s": round(st["seconds"], 4),
                       "samples_per_s": round(samples / b_wall, 1), "settings": tuned, "tuner": tinfo,
                       "rays": st["rays_extension"] + st["rays_shadow"], "passes": st["iterations"]},
           "speedup_wall": round(seq_wall / b_wall, 3), "speedup_gpu": round(seq_gpu / st["seconds"], 3)}
    batch.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--scenes", type=int, default=32)
    ap.add_argument("--spp", type=int, default=16)
    ap.add_argument("--depth", type=int, default=8)
    ap.add_argument("--tri-hi", type=int, default=2_000_000)
    args = ap.parse_args()
    import trenchanttracer_amd as tt

    ctx = tt.Context(0)
    out = dict({"metric": "samples/s (batched scenes)"},
               **measure(tt, ctx, args.scenes, args.spp, args.depth, args.tri_hi))
    print(json.dumps(out), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
