"""Autotuner for the traversal and the wavefront renderer.

Times short runs of candidate launch configurations on the actual workload and
keeps the fastest: traversal loop variant x LDS stack entries x block size x
launch form (one wave per 64-ray group, or persistent waves with lane refill) for
`tune_trace`; paths in flight, hipGraph replay, traversal LDS stack, block size,
refill and the pixel tile the path slots walk for `tune_render` (coordinate
descent).
Results can be cached in a JSON file keyed by a workload signature."""
import itertools
import json
import os

from . import render_params, trace_cfg


def _load(path):
    if path and os.path.exists(path):
        with open(path) as f:
            return json.load(f)
    return {}


def _save(path, data):
    if path:
        tmp = path + ".tmp"
        with open(tmp, "w") as f:
            json.dump(data, f, indent=1)
        os.replace(tmp, path)


def tune_trace(ctx, scene, d_rays, n, d_hits, variants=(0, 1), lds=(8, 16, 24), blocks=(64, 128, 256),
               refills=(0, 16, 32), reorders=(0, 1), reps=3, exact_only=True, cache=None, key=None):
    """Fastest trace configuration for this scene and ray buffer: loop variant, LDS
    stack depth, block size, launch form (refill 0: one wave per 64-ray group;
    > 0: persistent waves that take new rays once that many lanes are idle) and ray
    order (reorder 1: traced in Morton order of the origins). With exact_only the
    speculative variant is not considered."""
    from . import Timer
    db = _load(cache)
    if key and key in db:
        c = db[key]
        return trace_cfg(c["variant"], c["lds_stack"], c["block"], refill=c.get("refill", 0),
                         reorder=c.get("reorder", 0)), c["ms"]
    timer = Timer(ctx, 2)
    best = None
    for v, s, b, r, o in itertools.product(variants, lds, blocks, refills, reorders):
        if exact_only and v != 0:
            continue
        cfg = trace_cfg(v, s, b, refill=r, reorder=o)
        scene.trace_closest(d_rays, n, d_hits, cfg)
        ms = []
        for _ in range(reps):
            timer.record(0)
            scene.trace_closest(d_rays, n, d_hits, cfg)
            timer.record(1)
            ms.append(timer.elapsed(0, 1))
        t = sorted(ms)[len(ms) // 2]
        if best is None or t < best[0]:
            best = (t, v, s, b, r, o)
    timer.close()
    if key:
        db[key] = {"variant": best[1], "lds_stack": best[2], "block": best[3], "refill": best[4],
                   "reorder": best[5], "ms": best[0]}
        _save(cache, db)
    return trace_cfg(best[1], best[2], best[3], refill=best[4], reorder=best[5]), best[0]


RENDER_SPACE = {
    "paths_in_flight": (0, 16 << 20),   # 0: up to 2^27 paths per pass
    "use_graph": (False, True),
    "lds_stack": (16, 8, 24),           # traversal LDS stack entries per lane
    "block": (64, 128, 256),            # traversal block size
    "refill": (16, 8, 32),              # persistent-wave refill of the later bounces
    "tile": (8, 4, 16),                 # path slots walk tile x tile pixel blocks
}


def tune_render(render_fn, base, space=None, probe_spp=None, rounds=2, cache=None, key=None):
    """Fastest render configuration by coordinate descent: starting from the first
    value of every parameter, each round tries every value of one parameter at a time
    with the others held at their best so far (paths in flight, hipGraph replay,
    traversal LDS stack, block size, refill of the later bounces, slot tile size),
    `rounds` sweeps over all of them. render_fn(params) -> stats; base: a
    tt_render_params-like dict of the real render. Probes run at the real spp unless
    probe_spp is given: the number of passes, which paths in flight decide, only
    shows at the real sample count. Returns (settings, info): the chosen values of the
    tuned parameters, and {"seconds": the best probe's GPU time, "trials": probes run,
    "default_seconds": the probe at the first values}."""
    space = dict(RENDER_SPACE if space is None else space)
    db = _load(cache)
    if key and key in db:
        return db[key]["settings"], db[key]["info"]
    cur = {k: v[0] for k, v in space.items()}
    seen = {}

    def measure(cfg):
        sig = tuple(sorted(cfg.items()))
        if sig not in seen:
            prm = dict(base)
            prm.update(cfg)
            prm["spp"] = probe_spp or base["spp"]
            render_fn(render_params(**prm))       # warm (allocations, code objects, graph capture)
            seen[sig] = render_fn(render_params(**prm))["seconds"]
        return seen[sig]

    best_t = default_t = measure(cur)
    for _ in range(rounds):
        changed = False
        for name, values in space.items():
            for v in values:
                if v == cur[name]:
                    continue
                cand = dict(cur, **{name: v})
                t = measure(cand)
                if t < best_t:
                    best_t, cur, changed = t, cand, True
        if not changed:
            break
    info = {"seconds": best_t, "trials": len(seen), "default_seconds": default_t}
    if key:
        db[key] = {"settings": cur, "info": info}
        _save(cache, db)
    return cur, info
