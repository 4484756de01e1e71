"""Profiling cases.

traversal (default): one closest-hit traversal call over the traversal workload
(1M-triangle soup SBVH, 8,294,400 rays), one random 64-byte gather probe (2^26
independent reads over a buffer the size of the device tree), then one 256 MiB
device-to-device copy whose known byte counts calibrate the FETCH_SIZE / WRITE_SIZE
hardware counters.
pathtrace: the path-trace case (1,310,720-tri displaced icosphere, sky + sun,
1920x1080, depth 16) rendered once at 4 spp (warm-up) and once at 64 spp, then the
same calibration copy.
train: 13 Adam steps of the guiding network on 262,144 seeded random rows (inputs in
[0, 1], unit directions, Gamma(1) weights), the batch of the online training.
guided: the path-trace case rendered at 64 spp with guiding, the network trained
online during the render (Adam lr 1e-3, 2^18-row batches, 2^21 recorded paths per
training pass, passes growing 8x, learned selection probability), then once more at
64 spp with that network frozen; --guided-bounces N uses the mixture at the first N
bounces only (0: every bounce).
For profiler runs, one counter set per run:

    rocprofv3 --pmc FETCH_SIZE --output-format csv -d OUT -o run -- \\
        python3 -m trenchanttracer_amd.profile_trace [--case traversal|pathtrace|train|guided] [--variant 1] [--lds-stack 16]
            [--block 64] [--refill 0] [--reorder 0] [--paths-in-flight N] [--guided-bounces 0]
"""
import argparse
import sys

GATHER_FETCHES = 1 << 26


def _copy_calibration(ctx, nbytes=256 << 20):
    """One device-to-device copy of known size (k_copy16): its counter readings scale
    the HBM counters of the kernels profiled in the same process."""
    import trenchanttracer_amd as tt
    src, dst = tt.DeviceBuffer(ctx, nbytes), tt.DeviceBuffer(ctx, nbytes)
    src.zero()
    ctx.sync()
    dst.copy_from(src)
    ctx.sync()
    src.free()
    dst.free()
    return nbytes


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--variant", type=int, default=1)
    ap.add_argument("--lds-stack", type=int, default=16)
    ap.add_argument("--block", type=int, default=64)
    ap.add_argument("--refill", type=int, default=0, help="traversal: refill threshold (0: one wave per group)")
    ap.add_argument("--reorder", type=int, default=0, help="traversal: 1 = trace in Morton order of the origins")
    ap.add_argument("--case", choices=("traversal", "pathtrace", "train", "guided"), default="traversal")
    ap.add_argument("--paths-in-flight", type=int, default=0, help="pathtrace: paths per pass (0: the default)")
    ap.add_argument("--ntris", type=int, default=1_000_000, help="traversal: triangles of the soup")
    ap.add_argument("--guided-bounces", type=int, default=0,
                    help="guided: bounces that use the mixture (max_bounce; 0: every bounce)")
    args = ap.parse_args()
    import trenchanttracer_amd as tt
    from trenchanttracer_amd import workloads

    if args.case == "train":
        import numpy as np
        ctx = tt.Context(0)
        g = tt.Guide(ctx, seed=1, lr=1e-3)
        rng = np.random.default_rng(1)
        n = 1 << 18
        x = rng.random((n, 9), dtype=np.float32)
        d = rng.normal(size=(n, 3)).astype(np.float32)
        d /= np.linalg.norm(d, axis=1, keepdims=True)
        w = rng.gamma(1.0, 1.0, n).astype(np.float32)
        for _ in range(13):
            loss = g.train_step(x, d, w)
        ctx.sync()
        print(f"13 training steps of {n} rows, last loss {loss:.4f}", file=sys.stderr, flush=True)
        g.close()
        ctx.close()
        return

    if args.case == "guided":
        mesh, mats, env, cam = workloads.pathtrace()
        ctx = tt.Context(0)
        s = tt.DeviceScene(ctx, mesh, tt.BVH.build(mesh), mats, env=env)
        g = tt.Guide(ctx, seed=1, lr=1e-3)

        def rp(spp, seed, **kw):
            return tt.render_params(spp=spp, max_depth=16, rr_depth=4, seed=seed, mode=tt.RENDER_PATH, **kw)

        _, st, gs = g.render(s, cam, rp(64, 2000, paths_in_flight=cam.width * cam.height), alpha=0.5, train=True,
                             train_paths=2 << 20, train_batch=1 << 18, pass_growth=8, learn_alpha=True,
                             max_bounce=args.guided_bounces)
        print(f"guided 64 spp, trained online: {st} {gs}", file=sys.stderr, flush=True)
        _, st, _ = g.render(s, cam, rp(64, 3000), alpha=0.5, train=False, learn_alpha=True,
                            max_bounce=args.guided_bounces)
        print(f"guided 64 spp, frozen network: {st}", file=sys.stderr, flush=True)
        g.close()
        s.close()
        ctx.close()
        return

    if args.case == "pathtrace":
        mesh, mats, env, cam = workloads.pathtrace()
        ctx = tt.Context(0)
        s = tt.DeviceScene(ctx, mesh, tt.BVH.build(mesh), mats, env=env)
        s.render(cam, tt.render_params(spp=4, max_depth=16, rr_depth=4, seed=7, mode=tt.RENDER_PATH))
        img, st = s.render(cam, tt.render_params(spp=64, max_depth=16, rr_depth=4, seed=1, mode=tt.RENDER_PATH,
                                                 paths_in_flight=args.paths_in_flight))
        print(f"path trace 64 spp: {st}", file=sys.stderr, flush=True)
        print(f"copied {_copy_calibration(ctx)} bytes", file=sys.stderr, flush=True)
        s.close()
        ctx.close()
        return

    mesh, bvh, rays, n_prim = workloads.traversal(ntris=args.ntris)
    n = len(rays)
    ctx = tt.Context(0)
    scene = tt.DeviceScene(ctx, mesh, bvh)
    d_rays = tt.DeviceBuffer.from_array(ctx, rays)
    d_hits = tt.DeviceBuffer(ctx, n * 16)
    cfg = tt.trace_cfg(variant=args.variant, lds_stack=args.lds_stack, block=args.block, refill=args.refill,
                       reorder=args.reorder)
    scene.trace_closest(d_rays, n, d_hits, cfg)
    ctx.sync()
    # independent random 64-byte reads over a buffer the size of the device tree
    # (2^26 reads, seed 1): the fetch-throughput ceiling the traversal's traffic is
    # compared with
    info = scene.info()
    probe = tt.DeviceBuffer(ctx, info["device_bytes"])
    probe.zero()
    tt.gather64_gbps(ctx, probe, info["device_bytes"], fetches=GATHER_FETCHES, seed=1)
    probe.free()
    nbytes = _copy_calibration(ctx)
    print(f"traced {n} rays (variant {args.variant}, refill {args.refill}, reorder {args.reorder}); "
          f"copied {nbytes} bytes", file=sys.stderr, flush=True)
    for b in (d_rays, d_hits):
        b.free()
    scene.close()
    ctx.close()


if __name__ == "__main__":
    main()
