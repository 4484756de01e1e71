"""Autotuner for the traversal and the wavefront renderer.

Times short runs of candidate launch configurations on the actual workload and
keeps the fastest: traversal loop variant x LDS stack entries x block size x
launch form (one wave per 64-ray group, or persistent waves with lane refill) for
`tune_trace`; paths in flight x hipGraph replay for `tune_render`.
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


def tune_render(render_fn, base, paths=(0, 16 << 20), lds=(16,), blocks=(64,), graphs=(0, 1),
                probe_spp=None, cache=None, key=None):
    """Fastest render configuration. render_fn(params) -> stats; base: a
    tt_render_params-like dict of the real render. Probes run at the real spp
    unless probe_spp is given: the number of passes, which paths in flight decide,
    only shows at the real sample count."""
    db = _load(cache)
    if key and key in db:
        return db[key]
    best = None
    for pf, s, b, g in itertools.product(paths, lds, blocks, graphs):
        prm = dict(base)
        prm.update(spp=probe_spp or base["spp"], paths_in_flight=pf, lds_stack=s, block=b, use_graph=bool(g))
        render_fn(render_params(**prm))           # warm (allocations, code objects)
        st = render_fn(render_params(**prm))
        t = st["seconds"]
        if best is None or t < best[0]:
            best = (t, {"paths_in_flight": pf, "lds_stack": s, "block": b, "use_graph": bool(g)})
    if key:
        db[key] = best[1]
        _save(cache, db)
    return best[1]
