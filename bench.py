"""Benchmark: primary+shadow Mrays/s of the fused HIP render kernel.

Workload (default): 512^3 uint8 grid (fBm terrain scaled x4 in x/z plus 4,000
16^3 noise sprites), one 1920x1080 view, one primary ray per pixel plus one
shadow ray per primary hit toward the sun (0.4, 0.8, 0.3), max 2048 steps.
A step is one render of that frame with the world already resident in HBM.

Prints ONE JSON line.  value = (primary + shadow rays) / kernel time, timed
with HIP events on the stream the kernel is launched on.  roofline.achieved
counts the algorithmic bytes of the launch (per executed lookup step the
traversal's gathers, plus 28 B of shaded + hit-record output per pixel).
cpu_baseline = the oracle (CPU restatement) on the host cores, one frame.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "unnamed-voxel-tracer_amd"))

METRIC = ("Mrays/s primary+shadow, 1080p/512³ grid; fwd+bwd ms per view vs plain kernel")
HBM_PEAK_GBS = 8000.0
OUT_BYTES_PER_PIXEL = 16 + 12  # RGBA32F shaded + 12-byte hit record


def algorithmic_bytes(stats, traversal, npix):
    """Bytes one launch must move: the traversal's gathers plus the outputs.

    flat:    1 byte per executed traceMap step (the voxel gather);
    bitmask: 8 B per 16^3-region word + 8 B per 4^3-brick word fetched,
             + 1 B material per primary hit.
    """
    if traversal == "bitmask":
        b = 8 * stats["l1_fetches"] + 8 * stats["l0_fetches"] + stats["leaf_reads"]
    else:
        b = stats["primary_steps"] + stats["shadow_steps"]
    return b + OUT_BYTES_PER_PIXEL * npix


def cpu_baseline(grid, pal, cam, W, H, max_steps, sun):
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_ffi as O
    threads = min(len(os.sched_getaffinity(0)), 16)
    world = O.DenseWorld(grid, pal)
    t0 = time.perf_counter()
    out = O.render(world, cam, W, H, max_steps, max_steps, sun, threads=threads, gbuffer=False)
    dt = time.perf_counter() - t0
    rays = W * H + int(out["stats"][2])
    return {"value": rays / dt / 1e6, "unit": "Mrays/s", "cores": threads, "kind": "port",
            "sample": f"one full {W}x{H} frame of the same workload, {threads} threads, "
                      f"{dt:.2f} s wall"}


def timed(torch, stream, fn, steps, warmup):
    """Mean ms of fn() over `steps` calls, HIP events on `stream`."""
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record(stream)
    for _ in range(steps):
        fn()
    b.record(stream)
    torch.cuda.synchronize()
    return a.elapsed_time(b) / steps


def bench_batched(args):
    """64 cameras of uneven resolution (8-aligned sides in [64,1024]) over the
    1024^3 world (terrain x8 + 32,000 sprites), primary + shadow, max 2048
    steps: one persistent work-queue launch captured in a hipGraph with the
    autotuned tile / chunk / traversal, against one plain vt_render per view
    on the same GPU."""
    import torch

    import vt
    from vt import scenes as S

    print("[bench] generating 1024^3 world", file=sys.stderr, flush=True)
    grid = S.generate(S.spec("c2"))
    ctx = vt.Context(0)
    world = vt.World.dense(ctx, grid, S.palette())
    del grid
    views = S.views("c2")
    frames = [vt.Frame.alloc(W, H) for _, W, H in views]
    cams = [c for c, _, _ in views]
    pairs = list(zip(cams, frames))
    ms_cap = S.MAX_STEPS["c2"]
    prm = vt.params(0, 0, ms_cap, ms_cap)
    stream = torch.cuda.current_stream()
    npix = sum(W * H for _, W, H in views)

    stats = torch.zeros(8, dtype=torch.int64, device="cuda")
    counted = vt.Batch(ctx, world, pairs, prm, vt.tuning(), stats=stats)
    counted.launch()
    torch.cuda.synchronize()
    st = dict(zip(vt._native.STAT_NAMES, stats.cpu().tolist()))
    counted.close()
    rays = npix + st["primary_hits"]

    print("[bench] plain: one vt_render per view", file=sys.stderr, flush=True)
    vprm = [vt.params(W, H, ms_cap, ms_cap) for _, W, H in views]

    def plain():
        for (cam, f), p in zip(pairs, vprm):
            vt.render(ctx, world, cam, f, p)

    plain_ms = timed(torch, stream, plain, max(1, args.steps // 4), 1)

    print("[bench] autotune", file=sys.stderr, flush=True)
    best, tune_ms = vt.autotune(ctx, world, pairs, prm, reps=2)
    batch = vt.Batch(ctx, world, pairs, prm, best)
    batch.capture_graph()
    print(f"[bench] graph replays x{args.steps}", file=sys.stderr, flush=True)
    ms = timed(torch, stream, batch.graph_launch, args.steps, args.warmup)
    abytes = st["primary_steps"] + st["shadow_steps"] + OUT_BYTES_PER_PIXEL * npix
    achieved = abytes / (ms * 1e-3) / 1e9
    line = {
        "metric": "Mrays/s primary+shadow, 64 uneven cameras on a 1024³ grid, one persistent "
                  "graph launch (vs one plain launch per view)",
        "value": round(rays / (ms * 1e-3) / 1e6, 2),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "1024^3 uint8 grid (fBm terrain x8 + 32000 16^3 sprites), 64 cameras "
                        "with sides in [64,1024], primary + shadow, max 2048 steps",
            "pixels": npix,
            "shadow_rays": st["primary_hits"],
            "dda_steps": st["primary_steps"] + st["shadow_steps"],
            "tuning": {"tile": {3: "8x8", 4: "16x4", 5: "32x2"}[best.tile],
                       "chunk": best.chunk,
                       "traversal": ["flat", "bitmask"][best.traversal],
                       "blocks_per_cu": best.blocks_per_cu},
            "autotune_eager_ms": round(tune_ms, 4),
            "plain_ms_all_views": round(plain_ms, 4),
            "plain_mrays_s": round(rays / (plain_ms * 1e-3) / 1e6, 2),
            "speedup_vs_plain": round(plain_ms / ms, 3),
        },
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": None},
        "cpu_baseline": None,
    }
    print(json.dumps(line), flush=True)
    batch.close()
    world.close()
    ctx.close()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--traversal", choices=["flat", "bitmask"], default="flat")
    ap.add_argument("--workload", choices=["frame", "batched"], default="frame",
                    help="frame: 512^3, one 1080p view (the headline); batched: 1024^3, 64 "
                         "uneven cameras in one persistent graph launch")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported "
              "(rays are independent: a multi-GPU run would shard views/tiles with no exchange)",
              file=sys.stderr)
        sys.exit(2)
    if args.workload == "batched":
        return bench_batched(args)

    import numpy as np
    import torch

    import vt
    from vt import scenes as S

    print("[bench] generating 512^3 world", file=sys.stderr, flush=True)
    grid = S.generate(S.spec("c1"))
    pal = S.palette()
    cam, W, H = S.views("c1")[0]
    max_steps = S.MAX_STEPS["c1"]
    ctx = vt.Context(0)
    world = vt.World.dense(ctx, grid, pal)
    trav = vt.VT_TRAVERSAL_BITMASK if args.traversal == "bitmask" else vt.VT_TRAVERSAL_FLAT

    # one counted launch (statistics kernel variant) for rays and bytes
    fs = vt.Frame.alloc(W, H, stats=True)
    vt.render(ctx, world, cam, fs, vt.params(W, H, max_steps, max_steps, S.SUN, 0, trav))
    torch.cuda.synchronize()
    stats = fs.stats_dict()
    del fs

    f = vt.Frame.alloc(W, H)
    prm = vt.params(W, H, max_steps, max_steps, S.SUN, 0, trav)
    stream = torch.cuda.current_stream()
    print(f"[bench] warmup {args.warmup}", file=sys.stderr, flush=True)
    for _ in range(args.warmup):
        vt.render(ctx, world, cam, f, prm)
    torch.cuda.synchronize()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
          for _ in range(args.steps)]
    t0 = time.perf_counter()
    for a, b in ev:
        a.record(stream)
        vt.render(ctx, world, cam, f, prm)
        b.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kms = [a.elapsed_time(b) for a, b in ev]
    ms = float(np.mean(kms))
    print(f"[bench] {args.steps} steps, kernel mean {ms:.3f} ms, min {min(kms):.3f} ms",
          file=sys.stderr, flush=True)

    npix = W * H
    rays = npix + stats["primary_hits"]
    value = rays / (ms * 1e-3) / 1e6
    abytes = algorithmic_bytes(stats, args.traversal, npix)
    achieved = abytes / (ms * 1e-3) / 1e9
    steps_total = stats["primary_steps"] + stats["shadow_steps"]

    cpu = None
    if not args.no_cpu:
        print("[bench] cpu baseline (oracle, one frame)", file=sys.stderr, flush=True)
        cpu = cpu_baseline(grid, pal, cam, W, H, max_steps, S.SUN)

    line = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall * 1e3 / args.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "512^3 uint8 grid (fBm terrain x4 in x/z + 4000 16^3 sprites), "
                        "1920x1080 primary + shadow rays, sun (0.4,0.8,0.3), max 2048 steps",
            "traversal": args.traversal,
            "kernel_ms": round(ms, 4),
            "primary_rays": npix,
            "shadow_rays": stats["primary_hits"],
            "dda_steps": steps_total,
            "gsteps_per_s": round(steps_total / (ms * 1e-3) / 1e9, 2),
        },
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 2),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 5),
            "traffic": None,
        },
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    world.close()
    ctx.close()


if __name__ == "__main__":
    main()
