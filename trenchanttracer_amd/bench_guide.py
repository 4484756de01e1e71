"""Guided-versus-plain measurement. Prints one JSON line.

Equal-time RMSE on the full path-trace case (1,310,720-tri displaced icosphere,
40/40/20 diffuse/GGX/dielectric, 2048x1024 sky + sun, 1920x1080, depth 16,
roulette from depth 4) against an FP64-accumulated oracle render of the plain
tracer (8192 spp by default, a different seed): plain renders at 16..128 spp,
guided renders trained online from scratch (Adam lr 1e-3, 2^18-row batches) at
32..128 spp, the same trained during the first half of the samples and frozen for the
rest (rendered in wide passes), and renders with the network trained by the 128-spp
run and then frozen. Every point is averaged over independent seeds (MSE mean over `--seeds`
renders, each seed a fresh sample set): the error of this scene is carried by rare
sun caustics through glass faces (most of it in 0.1% of the pixels), so one render's
RMSE varies by tens of percent from seed to seed. The time at which guided rendering
reaches the plain 64-spp RMSE is read off a log-log fit of RMSE against time
(RMSE ~ time^-1/2 locally). Each point also reports the share of the squared error
in the worst 0.1% of pixels.

The same protocol on the door room (256x256, depth 12), where all light reaches
the camera's part of the room through a door, with a 4096-spp oracle.

    python -m trenchanttracer_amd.bench_guide [--oracle-spp 8192] [--seeds 4] [--skip-pathtrace] [--quick]
    python -m trenchanttracer_amd.bench_guide --error-classes
    python -m trenchanttracer_amd.bench_guide --sweep [schedules|bounces] [--seeds 2]

--error-classes: where the path-trace case's error sits. Plain 64-spp renders (seeds
100, 101) against a 4096-spp FP64 oracle, pixels classed by the material the centre
ray of the pixel hits first (diffuse, GGX conductor, glass, sky): share of pixels, share
of the squared error, and the classes of the 400 worst pixels.

--sweep: the path-trace case at 64 and 128 spp under the training schedules of
SWEEP (pass growth 2 to 16, fixed or learned selection probability; 2^21 training
paths, one path per pixel per pass), or with `--sweep bounces` under the bench's
schedule with the mixture used at the first 1, 2 or all bounces (BOUNCE_SWEEP): the
wall ratios to the plain 64-spp RMSE and the 64-spp points of each.
"""
import argparse
import json
import sys
import time

import numpy as np


def log(m):
    print(f"[guide] {m}", file=sys.stderr, flush=True)


def _time_to_rmse(points, target):
    """Time at which RMSE reaches `target`, from (time, rmse) pairs: log-log
    interpolation between the bracketing points, or extrapolation with slope -1/2
    from the nearest one."""
    pts = sorted(points)
    for (t0, r0), (t1, r1) in zip(pts, pts[1:]):
        if r0 >= target >= r1 and r0 > r1:
            a = (np.log(target) - np.log(r0)) / (np.log(r1) - np.log(r0))
            return float(np.exp(np.log(t0) + a * (np.log(t1) - np.log(t0))))
    t, r = min(pts, key=lambda p: abs(np.log(p[1] / target)))
    return float(t * (r / target) ** 2)


def compare(tt, ctx, sc, cam, ref, depth, plain_spp, guided_spp, paths, lr, batch, train_paths, seeds=1, **guide_kw):
    def err(img):
        e = ((img.astype(np.float64) - ref) ** 2).sum(-1).ravel()
        k = max(1, e.size // 1000)
        return float(e.mean() / 3), float(np.sort(e)[-k:].sum() / max(e.sum(), 1e-300))

    def point(spp, runs):
        mse = [r[0] for r in runs]
        return {"spp": spp, "rmse": float(np.sqrt(np.mean(mse))), "seconds": float(np.mean([r[2] for r in runs])),
                "rmse_per_seed": [float(np.sqrt(m)) for m in mse],
                "top0.1%_share": float(np.mean([r[1] for r in runs]))}

    def rp(spp, seed, **kw):
        return tt.render_params(spp=spp, max_depth=depth, rr_depth=4, seed=seed, mode=tt.RENDER_PATH, **kw)

    out = {"seeds": seeds, "plain": [], "guided_online": [], "guided_online_half": [], "guided_pretrained": []}
    for spp in plain_spp:
        runs = []
        for j in range(seeds):
            img, st = sc.render(cam, rp(spp, 1000 * spp + j))
            runs.append(err(img) + (st["seconds"],))
        out["plain"].append(point(spp, runs))
        log(f"plain {spp}: {out['plain'][-1]}")
    g = None
    for spp in guided_spp:
        runs, train = [], []
        for j in range(seeds):
            if g is not None:
                g.close()
            g = tt.Guide(ctx, seed=1 + j, lr=lr)
            img, st, gs = g.render(sc, cam, rp(spp, 2000 * spp + j, paths_in_flight=paths), alpha=0.5, train=True,
                                   train_paths=train_paths, train_batch=batch, **guide_kw)
            runs.append(err(img) + (st["seconds"],))
            train.append(gs)
        out["guided_online"].append(dict(point(spp, runs), train_seconds=float(np.mean([t["train_seconds"] for t in train])),
                                         train_steps=train[-1]["train_steps"], last_loss=train[-1]["last_loss"]))
        log(f"guided online {spp}: {out['guided_online'][-1]}")
    # trained from scratch during the first half of the samples (small passes), the
    # second half rendered with the network frozen in passes of the default width
    for spp in guided_spp:
        runs = []
        for j in range(seeds):
            gh = tt.Guide(ctx, seed=1 + j, lr=lr)
            img, st, gs = gh.render(sc, cam, rp(spp, 4000 * spp + j, paths_in_flight=paths), alpha=0.5, train=True,
                                    train_paths=train_paths, train_batch=batch, train_spp=spp // 2, **guide_kw)
            runs.append(err(img) + (st["seconds"],))
            gh.close()
        out["guided_online_half"].append(point(spp, runs))
        log(f"guided online, first half trained {spp}: {out['guided_online_half'][-1]}")
    for spp in plain_spp:
        runs = []
        for j in range(seeds):
            img, st, _ = g.render(sc, cam, rp(spp, 3000 * spp + j), alpha=0.5, train=False,
                                  learn_alpha=guide_kw.get("learn_alpha", False),
                                  max_bounce=guide_kw.get("max_bounce", 0))
            runs.append(err(img) + (st["seconds"],))
        out["guided_pretrained"].append(point(spp, runs))
        log(f"guided pretrained {spp}: {out['guided_pretrained'][-1]}")
    g.close()
    p64 = next(r for r in out["plain"] if r["spp"] == 64)
    for key in ("guided_online", "guided_online_half", "guided_pretrained"):
        t = _time_to_rmse([(r["seconds"], r["rmse"]) for r in out[key]], p64["rmse"])
        out[f"{key}_time_to_plain64_rmse"] = t
        out[f"{key}_wall_ratio"] = t / p64["seconds"]
    return out


def error_classes(tt, ctx, log):
    from trenchanttracer_amd import workloads
    mesh, mats, env, cam = workloads.pathtrace()
    sc = tt.DeviceScene(ctx, mesh, tt.BVH.build(mesh), mats, env=env)
    _, ref, _ = sc.render(cam, tt.render_params(spp=4096, max_depth=16, rr_depth=4, seed=1234, mode=tt.RENDER_PATH,
                                                fp64=True), want_fp64=True)
    ref = ref.astype(np.float64)
    W, H = cam.width, cam.height
    # centre ray of each pixel (the render's pinhole camera: the image plane one unit
    # along the view, spanned by tan(fov / 2) along the side and up axes)
    eye, view, up = (np.array(v, np.float64) for v in (cam.eye, cam.view, cam.up))
    view /= np.linalg.norm(view)
    up /= np.linalg.norm(up)
    hax = np.cross(view, up)
    hax /= np.linalg.norm(hax)
    vax = np.cross(hax, view)
    hor, ver = hax * np.tan(np.radians(cam.fov[0]) / 2), vax * np.tan(-np.radians(cam.fov[1]) / 2)
    Y, X = np.mgrid[0:H, 0:W]
    d = view + hor * (2 * ((X + 0.5) / (W - 1))[..., None] - 1) + ver * (2 * ((Y + 0.5) / (H - 1))[..., None] - 1)
    d /= np.linalg.norm(d, axis=-1, keepdims=True)
    rays = np.zeros((H * W, 8), np.float32)
    rays[:, 0:3] = eye
    rays[:, 4:7] = d.reshape(-1, 3)
    rays[:, 7] = 1e20
    tid = sc.intersect(rays)[0]
    tm = mesh.arrays()["tri_mats"]
    cls = np.full(H * W, 3)
    hit = tid >= 0
    mt = tm[tid[hit]]
    cls[hit] = np.where(mt == 0, 0, np.where(mt == len(mats) - 1, 2, 1))
    names = ["diffuse", "conductor", "glass", "sky"]
    out = {"pixel_share": {names[c]: float(np.mean(cls == c)) for c in range(4)}, "seeds": {}}
    for seed in (100, 101):
        img, _ = sc.render(cam, tt.render_params(spp=64, max_depth=16, rr_depth=4, seed=seed, mode=tt.RENDER_PATH))
        e = ((img.astype(np.float64) - ref) ** 2).sum(-1).ravel()
        worst = np.argsort(e)[::-1][:400]
        out["seeds"][seed] = {
            "mse": float(e.mean() / 3),
            "mse_share": {names[c]: float(e[cls == c].sum() / e.sum()) for c in range(4)},
            "worst400": {names[c]: int(np.sum(cls[worst] == c)) for c in range(4)},
            "worst_0.01%_share": float(np.sort(e)[::-1][: e.size // 10000].sum() / e.sum()),
        }
        log(f"error classes seed {seed}: {out['seeds'][seed]}")
    sc.close()
    return out


SWEEP = (("growth2_fixed_alpha", dict(pass_growth=2)),
         ("growth2_learned_alpha", dict(pass_growth=2, learn_alpha=True)),
         ("growth4_learned_alpha", dict(pass_growth=4, learn_alpha=True)),
         ("growth8_learned_alpha", dict(pass_growth=8, learn_alpha=True)),
         ("growth16_learned_alpha", dict(pass_growth=16, learn_alpha=True)))

# how many bounces use the mixture (max_bounce; 0: every bounce), under the bench's schedule
BOUNCE_SWEEP = tuple((f"guided_bounces_{b or 'all'}", dict(pass_growth=8, learn_alpha=True, max_bounce=b))
                     for b in (0, 1, 2))


def sweep(tt, ctx, oracle_spp, seeds, log, cases=SWEEP):
    """The path-trace case under each setting of `cases` (SWEEP: training schedules;
    BOUNCE_SWEEP: guided bounces), every one compared with the same FP64 oracle."""
    from trenchanttracer_amd import workloads
    mesh, mats, env, cam = workloads.pathtrace()
    sc = tt.DeviceScene(ctx, mesh, tt.BVH.build(mesh, nthreads=16), mats, env=env)
    _, ref, _ = sc.render(cam, tt.render_params(spp=oracle_spp, max_depth=16, rr_depth=4, seed=1234,
                                                mode=tt.RENDER_PATH, fp64=True), want_fp64=True)
    out = {}
    for name, kw in cases:
        r = compare(tt, ctx, sc, cam, ref, 16, (64, 128), (64, 128), cam.width * cam.height, 1e-3, 262144,
                    2 << 20, seeds, **kw)
        out[name] = {k: v for k, v in r.items() if k.endswith("ratio")}
        out[name].update(online64=r["guided_online"][0], frozen64=r["guided_pretrained"][0], plain64=r["plain"][0])
        log(f"{name}: {out[name]}")
    sc.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--oracle-spp", type=int, default=8192)
    ap.add_argument("--seeds", type=int, default=4, help="independent renders averaged per point")
    ap.add_argument("--skip-pathtrace", action="store_true", help="door room only")
    ap.add_argument("--quick", action="store_true", help="door room only, one training run (for profiling)")
    ap.add_argument("--error-classes", action="store_true", help="where the path-trace case's error sits")
    ap.add_argument("--sweep", nargs="?", const="schedules", choices=("schedules", "bounces"),
                    help="path-trace case under the training schedules of SWEEP or the guided bounces of BOUNCE_SWEEP")
    args = ap.parse_args()
    import trenchanttracer_amd as tt
    from trenchanttracer_amd import workloads

    ctx = tt.Context(0)
    if args.error_classes:
        print(json.dumps({"metric": "path-trace case: squared error by pixel class",
                          "pathtrace_1080p": error_classes(tt, ctx, log)}), flush=True)
        ctx.close()
        return
    if args.sweep:
        by = "training schedule" if args.sweep == "schedules" else "guided bounces"
        cases = SWEEP if args.sweep == "schedules" else BOUNCE_SWEEP
        print(json.dumps({"metric": f"path-trace case: wall ratio to the plain 64-spp RMSE by {by}",
                          "oracle_spp": args.oracle_spp, "seeds": args.seeds,
                          "pathtrace_1080p": sweep(tt, ctx, args.oracle_spp, args.seeds, log, cases)}), flush=True)
        ctx.close()
        return
    result = {"metric": "equal-time RMSE vs FP64 oracle, guided vs plain"}

    mesh, mats, cam = workloads.door_room(256, 256)
    sc = tt.DeviceScene(ctx, mesh, tt.BVH.build(mesh), mats)
    _, ref, st = sc.render(cam, tt.render_params(spp=4096, max_depth=12, rr_depth=4, seed=1234, mode=tt.RENDER_PATH,
                                                 fp64=True), want_fp64=True)
    log(f"door oracle 4096 spp in {st['seconds']:.2f}s")
    if args.quick:
        result["door_room_256"] = compare(tt, ctx, sc, cam, ref, 12, (64,), (256,), 131072, 1e-2, 16384, 65536, 1)
    else:
        result["door_room_256"] = compare(tt, ctx, sc, cam, ref, 12, (16, 32, 64, 128), (64, 256, 1024), 131072,
                                          1e-2, 16384, 65536, args.seeds)
    result["door_room_256"]["oracle_spp"] = 4096
    sc.close()

    if not (args.skip_pathtrace or args.quick):
        mesh, mats, env, cam = workloads.pathtrace()
        sc = tt.DeviceScene(ctx, mesh, tt.BVH.build(mesh), mats, env=env)
        t0 = time.time()
        _, ref, st = sc.render(cam, tt.render_params(spp=args.oracle_spp, max_depth=16, rr_depth=4, seed=1234,
                                                     mode=tt.RENDER_PATH, fp64=True), want_fp64=True)
        log(f"path-trace case oracle {args.oracle_spp} spp in {time.time() - t0:.1f}s")
        result["pathtrace_1080p"] = compare(tt, ctx, sc, cam, ref, 16, (16, 32, 64, 128), (32, 64, 128),
                                            2 * 1920 * 1080, 1e-3, 262144, 65536, args.seeds)
        result["pathtrace_1080p"]["oracle_spp"] = args.oracle_spp
        sc.close()
    print(json.dumps(result), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
