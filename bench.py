"""Benchmark: primary+shadow Mrays/s of the fused HIP render kernel.

Workload (default): 512^3 uint8 grid (fBm terrain scaled x4 in x/z plus 4,000
16^3 noise sprites), one 1920x1080 view, one primary ray per pixel plus one
shadow ray per primary hit toward the sun (0.4, 0.8, 0.3), max 2048 steps.
A step is one render of that frame with the world already resident in HBM;
by default 100 untimed frames (the GPU's clocks settle within the first ~50)
precede the 200 timed ones.

Prints ONE JSON line.  value = (primary + shadow rays) / kernel time, timed
with HIP events on the stream the kernel is launched on.  roofline.achieved
counts the algorithmic bytes of the launch (per executed step with a lookup
the traversal's gathers, plus 28 B of shaded + hit-record output per pixel).
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
# VALU issue peak: one wave64 VALU instruction per SIMD every 2 cycles at
# 2.4 GHz on 1,024 SIMDs -- the rate behind the published 157.3 TFLOP/s FP32
# vector peak (64 lanes x 2 flops per FMA instruction)
VALU_PEAK_GIPS = 1024 * 2.4 / 2
LOWER_BYTES_PER_STEP = 1  # every traceMap step reads at least one byte
SUN_DIR = (7.52185881e-01, 6.58950984e-01, 7.52185881e-01)  # camera.glsl:8
TRAVERSALS = ["flat", "bitmask", "bitmask8"]  # VT_TRAVERSAL_FLAT, _BITMASK, _BITMASK8
OUT_BYTES_PER_PIXEL = 16 + 12  # RGBA32F shaded + 12-byte hit record


def lookup_steps(stats, flat=True):
    """Executed steps that made their lookup.  flat: statistics of the flat
    traversal of a dense world, whose l1_fetches slot counts the steps a
    primary ray takes before its line can reach the box of the nonempty
    voxels (those of the steps executed; 0 in traceMap's); otherwise (the
    bitmask traversals, the reference layout) the slot counts word fetches
    and every step counts."""
    steps = stats["primary_steps"] + stats["shadow_steps"]
    return steps - stats["l1_fetches"] if flat else steps


def algorithmic_bytes(stats, traversal, npix):
    """Bytes one launch must move: the traversal's gathers plus the outputs.

    flat:    1 byte per executed traceMap step with a lookup (the voxel
             gather);
    bitmask, bitmask8: 8 B per region word + 8 B per leaf word fetched,
             + 1 B material per primary hit.
    """
    if traversal in ("bitmask", "bitmask8"):
        b = 8 * stats["l1_fetches"] + 8 * stats["l0_fetches"] + stats["leaf_reads"]
    else:
        b = lookup_steps(stats)
    return b + OUT_BYTES_PER_PIXEL * npix


def counted_stats(vt, torch, ctx, world, jobs, free):
    """Statistics summed over views, jobs = [(camera, W, H, params maker)]
    where params maker(flags) returns the view's params.  free=False:
    traceMap's steps (the statistics kernel variant); free=True: the steps
    the free-region kernel executes (VT_FLAG_FREE_MAP_ON: a ray ends where
    nothing is left to hit), the work the timed kernel does when it takes
    the test."""
    stats = torch.zeros(8, dtype=torch.int64, device="cuda")
    for cam, W, H, mk in jobs:
        f = vt.Frame.alloc(W, H, shaded=False)
        f.stats = stats
        vt.render(ctx, world, cam, f, mk(vt.VT_FLAG_FREE_MAP_ON if free else 0))
    torch.cuda.synchronize()
    return dict(zip(vt._native.STAT_NAMES, stats.cpu().tolist()))


def cpu_baseline(grid, pal, cam, W, H, max_steps, sun):
    """The oracle (CPU restatement) on this host's cores, one whole frame.
    Returns (baseline record, the oracle's outputs) -- the outputs serve as
    the checker of the GPU frame."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_ffi as O
    threads = O.cpu_threads()
    world = O.DenseWorld(grid, pal)
    t0 = time.perf_counter()
    out = O.render(world, cam, W, H, max_steps, max_steps, sun, threads=threads, gbuffer=False)
    dt = time.perf_counter() - t0
    rays = W * H + int(out["stats"][2])
    return ({"value": rays / dt / 1e6, "unit": "Mrays/s", "cores": threads, "kind": "port",
             "sample": f"one full {W}x{H} frame of the same workload, {threads} threads, "
                       f"{dt:.2f} s wall"}, out)


def frame_parity(np, hits, shaded, o):
    """The timed kernel's last frame against the oracle's: hit records (bit
    for bit), shaded RGB (max-abs), and how many pixels' rays met a
    crossing-time tie margin under 1e-6 / 1e-4 (reproduced exactly by the
    tie rule, so none is exempted)."""
    bad = int(np.any(hits != o["hits"], axis=-1).sum())
    err = float(np.max(np.abs(shaded - o["shaded"])))
    m = o["tie_margin"]
    return {"checked": "whole frame vs the oracle", "hit_records_differ": bad,
            "shaded_max_abs": err, "tie_margin_lt_1e-6": int((m < 1e-6).sum()),
            "tie_margin_lt_1e-4": int((m < 1e-4).sum()), "exempted": 0}


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
    # the batch takes the free-region test as vt_render does for these frames
    free = vt.free_map_used(world, frames[0], vt.params(views[0][1], views[0][2], ms_cap, ms_cap))
    ex = st
    if free:
        ex = counted_stats(vt, torch, ctx, world,
                           [(c, W, H, lambda fl, W=W, H=H: vt.params(W, H, ms_cap, ms_cap, S.SUN, fl))
                            for c, W, H in views], True)

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
    if TRAVERSALS[best.traversal] != "flat":
        ex = st  # the bitmask traversals run every ray to its end
    flat = TRAVERSALS[best.traversal] == "flat"
    abytes = lookup_steps(ex, flat) + OUT_BYTES_PER_PIXEL * npix
    achieved = abytes / (ms * 1e-3) / 1e9
    bkey = ("k_render_batch/flat_free/1024cube_64views"
            if free and TRAVERSALS[best.traversal] == "flat" else None)
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
            "free_region_test": free and TRAVERSALS[best.traversal] == "flat",
            "executed_steps": ex["primary_steps"] + ex["shadow_steps"],
            "lookup_steps": lookup_steps(ex, flat),
            "tuning": {"tile": {3: "8x8", 4: "16x4", 5: "32x2"}[best.tile],
                       "chunk": best.chunk,
                       "traversal": TRAVERSALS[best.traversal],
                       "blocks_per_cu": best.blocks_per_cu},
            "autotune_eager_ms": round(tune_ms, 4),
            "plain_ms_all_views": round(plain_ms, 4),
            "plain_mrays_s": round(rays / (plain_ms * 1e-3) / 1e6, 2),
            "speedup_vs_plain": round(plain_ms / ms, 3),
        },
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5),
                     # counters of the flat free-region batch kernel (the tuning picked here)
                     "traffic": measured_traffic(bkey) if bkey else None,
                     "issue": measured_issue(bkey, ms) if bkey else None},
        "cpu_baseline": None,
    }
    print(json.dumps(line), flush=True)
    batch.close()
    world.close()
    ctx.close()


def bench_brickmap(args):
    """The drop-in path: the 512^3 headline world in the engine's own layout
    (chunk table + block pool + 256^3 model atlas, one reference block per
    grid voxel with a solid 8^3 atlas model), 1920x1080 primary + shadow, max
    2048 steps, through the literal two-level traceMap kernel.  The hit
    records must equal the dense world's."""
    import torch

    import vt
    from vt import brickmap as BM
    from vt import scenes as S

    print("[bench] generating 512^3 world, reference layout", file=sys.stderr, flush=True)
    grid = S.generate(S.spec("c1"))
    pal = S.palette()
    cam, W, H = S.views("c1")[0]
    cap = S.MAX_STEPS["c1"]
    ctx = vt.Context(0)
    dim, chunks, pool, atlas = BM.from_dense(grid, pal)
    world = vt.World.brickmap(ctx, dim, chunks, pool, atlas)
    dense = vt.World.dense(ctx, grid, pal)
    del grid
    prm = vt.params(W, H, cap, cap, S.SUN)
    fs = vt.Frame.alloc(W, H, stats=True)
    vt.render(ctx, world, cam, fs, prm)
    fd = vt.Frame.alloc(W, H)
    vt.render(ctx, dense, cam, fd, prm)
    torch.cuda.synchronize()
    same = bool(torch.equal(fs.hits, fd.hits))
    st = fs.stats_dict()
    f = vt.Frame.alloc(W, H)
    stream = torch.cuda.current_stream()
    ms = timed(torch, stream, lambda: vt.render(ctx, world, cam, f, prm), args.steps, args.warmup)
    dense_ms = timed(torch, stream, lambda: vt.render(ctx, dense, cam, fd, prm), args.steps,
                     args.warmup)
    npix = W * H
    rays = npix + st["primary_hits"]
    # the steps the timed kernel executes: with the free-region test (the
    # default here) rays end where nothing is left to hit; the block-level
    # steps are those of the dense world with the same voxels (a solid model
    # per block: the literal loop ends at the hit block), whose statistics
    # count them
    free = vt.free_map_used(world, f, prm)
    mkey = "k_render_brickmap/free/512cube_1080p"
    ex = st
    if free:
        ex = counted_stats(vt, torch, ctx, dense,
                           [(cam, W, H, lambda fl: vt.params(W, H, cap, cap, S.SUN, fl))], True)
    # what the timed kernel must read at least: one byte per executed step
    # (an occupancy byte in the empty-space loop, a 4-byte chunk word in the
    # literal loop) + the outputs; the literal loop's own reads -- 4 B chunk,
    # 4 B pool, 4 B atlas per step -- are the upper figure
    # (ex: the dense world's statistics when the test is taken, else this
    # layout's own, whose l1_fetches slot counts chunk-table reads)
    abytes = LOWER_BYTES_PER_STEP * lookup_steps(ex, free) + OUT_BYTES_PER_PIXEL * npix
    achieved = abytes / (ms * 1e-3) / 1e9
    line = {
        "metric": "Mrays/s primary+shadow, 1080p/512³ grid in the engine's chunk/pool/atlas "
                  "layout (drop-in path)",
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
            "workload": "512^3 world of the headline, as chunk table (64^3) + block pool "
                        f"({len(pool) // 512} blocks) + 256^3 atlas; 1920x1080, max 2048 steps",
            "hits_equal_dense": same,
            "dense_ms": round(dense_ms, 4),
            "chunk_reads": st["l1_fetches"],
            "pool_atlas_reads": st["l0_fetches"],
            "dda_steps": st["primary_steps"] + st["shadow_steps"],
            "free_region_test": free,
            "executed_steps": ex["primary_steps"] + ex["shadow_steps"],
            "lookup_steps": lookup_steps(ex, free),
        },
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5),
                     "traffic": measured_traffic(mkey) if free else None,
                     "issue": measured_issue(mkey, ms) if free else None,
                     "bytes": "lower bound: 1 B per executed step with a lookup + 28 B per pixel"},
        "cpu_baseline": None,
    }
    print(json.dumps(line), flush=True)
    world.close()
    dense.close()
    ctx.close()


def bench_diff(args, scene):
    print(json.dumps(measure_diff(args, scene)), flush=True)


def measure_diff(args, scene):
    """Inverse-rendering step: Lambert x shadow forward over every view, then
    dL/dalbedo and dL/dsun of L = sum ||I - I_gt||^2 with the hit structure
    held fixed.  Albedo starts at 0.5; targets are renders of the ground truth
    (material colour + N(0, 0.05)).

    c3: 256^3 world, 16 orbit views of 512x512, dense per-voxel albedo, one
        vt_render per view;
    c4: 1024^3 world, 48 uneven views (sides in [256,1024]), albedo only for
        occupied voxels (compact prefix index), one persistent batch launch
        per pass (captured in a hipGraph).

    Timed modes (ms per view): plain forward (primary + shadow, no shading
    model); stored = forward writing hit records + one backward launch over
    all records; recompute = forward with the fused L2 gradient (no records).
    """
    import torch

    import vt
    from vt import scenes as S

    batched = scene == "c4"
    print(f"[bench] generating {scene} world", file=sys.stderr, flush=True)
    grid = S.generate(S.spec(scene))
    pal = S.palette()
    ctx = vt.Context(0)
    world = vt.World.dense(ctx, grid, pal)
    views = S.views(scene)
    cap = S.MAX_STEPS[scene]
    nviews = len(views)
    npix = sum(W * H for _, W, H in views)
    stream = torch.cuda.current_stream()
    flags = vt.VT_FLAG_LAMBERT
    n_occ = None
    if batched:
        n_occ = vt.build_compact_index(ctx, world)
        flags |= vt.VT_FLAG_ALBEDO_COMPACT
    gt = S.ground_truth_albedo(grid, pal, batched, 0x5EED)
    if batched:
        assert gt.numel() == 3 * n_occ, (gt.numel(), n_occ)
    del grid
    albedo = torch.full_like(gt, 0.5)
    grad = torch.zeros_like(gt)
    gsun = torch.zeros(3, dtype=torch.float32, device="cuda")
    loss = torch.zeros(1, dtype=torch.float32, device="cuda")

    # one buffer per image kind, views are slices: one backward launch covers all
    hits = torch.empty((npix, 3), dtype=torch.int32, device="cuda")
    shaded = torch.empty((npix, 4), dtype=torch.float32, device="cuda")
    target = torch.empty((npix, 4), dtype=torch.float32, device="cuda")
    frames, off = [], 0
    for _, W, H in views:
        f = vt.Frame(W, H)
        f.hits = hits[off:off + W * H].view(H, W, 3)
        f.shaded = shaded[off:off + W * H].view(H, W, 4)
        f.target = target[off:off + W * H].view(H, W, 4)
        frames.append(f)
        off += W * H
    cams = [c for c, _, _ in views]

    def prm(W, H, fl, alb, with_grad=False):
        g = (grad, gsun, loss) if with_grad else (None, None, None)
        return vt.params(W, H, cap, cap, S.SUN, fl, 0, AMBIENT_DIFF, alb, *g)

    def frames_without(*drop):
        out = []
        for f in frames:
            c = vt.Frame(f.width, f.height, shaded=f.shaded, hits=f.hits, target=f.target)
            for k in drop:
                setattr(c, k, None)
            out.append(c)
        return out

    # targets: renders of the ground truth
    print("[bench] rendering targets", file=sys.stderr, flush=True)
    tgt_frames = [vt.Frame(f.width, f.height, shaded=f.target) for f in frames]
    for cam, f in zip(cams, tgt_frames):
        vt.render(ctx, world, cam, f, prm(f.width, f.height, flags, gt))
    torch.cuda.synchronize()
    del gt

    stats = torch.zeros(8, dtype=torch.int64, device="cuda")
    for cam, f in zip(cams, frames):
        fs = vt.Frame(f.width, f.height, hits=f.hits, stats=stats)
        vt.render(ctx, world, cam, fs, prm(f.width, f.height, 0, None))
    torch.cuda.synchronize()
    st = dict(zip(vt._native.STAT_NAMES, stats.cpu().tolist()))
    nhit = st["primary_hits"]
    # the forward kernels take the free-region test as vt_render does for
    # these frames (the batch kernel follows the same rule)
    f0 = frames_without("target")[0]
    free = vt.free_map_used(world, f0, prm(f0.width, f0.height, 0, None))
    ex = st
    if free:
        ex = counted_stats(vt, torch, ctx, world,
                           [(c, f.width, f.height,
                             lambda fl, f=f: prm(f.width, f.height, fl, None))
                            for c, f in zip(cams, frames)], True)

    def zero():
        grad.zero_()
        gsun.zero_()
        loss.zero_()

    if batched:
        plain_b = vt.Batch(ctx, world, list(zip(cams, frames_without("target"))),
                           prm(0, 0, 0, None), vt.tuning())
        fwd_b = vt.Batch(ctx, world, list(zip(cams, frames_without("target"))),
                         prm(0, 0, flags, albedo), vt.tuning())
        rec_b = vt.Batch(ctx, world, list(zip(cams, frames_without("hits"))),
                         prm(0, 0, flags | vt.VT_FLAG_GRAD_L2, albedo, True), vt.tuning())
        for b in (plain_b, fwd_b, rec_b):
            b.capture_graph()
        plain_fwd = plain_b.graph_launch
        lambert_fwd = fwd_b.graph_launch
        fused = rec_b.graph_launch
    else:
        pp = [prm(f.width, f.height, 0, None) for f in frames]
        lp = [prm(f.width, f.height, flags, albedo) for f in frames]
        gp = [prm(f.width, f.height, flags | vt.VT_FLAG_GRAD_L2, albedo, True) for f in frames]
        f_plain, f_rec = frames_without("target"), frames_without("hits")

        def plain_fwd():
            for cam, f, p in zip(cams, f_plain, pp):
                vt.render(ctx, world, cam, f, p)

        def lambert_fwd():
            for cam, f, p in zip(cams, f_plain, lp):
                vt.render(ctx, world, cam, f, p)

        def fused():
            for cam, f, p in zip(cams, f_rec, gp):
                vt.render(ctx, world, cam, f, p)

    bprm = prm(0, 0, flags, albedo)

    def backward():
        vt.lambert_backward(ctx, world, hits, bprm, target, vt.VT_UPSTREAM_L2_TARGET, grad, gsun,
                            loss, stream=stream.cuda_stream)

    def stored():
        zero()
        lambert_fwd()
        backward()

    def recompute():
        zero()
        fused()

    steps, warm = args.steps, args.warmup
    print("[bench] plain forward", file=sys.stderr, flush=True)
    plain_ms = timed(torch, stream, plain_fwd, steps, warm)
    print("[bench] stored-record fwd+bwd", file=sys.stderr, flush=True)
    stored_ms = timed(torch, stream, stored, steps, warm)
    g_stored, s_stored, l_stored = grad.clone(), gsun.clone(), float(loss.item())
    lambert_ms = timed(torch, stream, lambert_fwd, steps, warm)
    lambert_fwd()
    bwd_ms = timed(torch, stream, backward, steps, warm)
    print("[bench] recompute fwd+bwd", file=sys.stderr, flush=True)
    rec_ms = timed(torch, stream, recompute, steps, warm)
    l_rec = float(loss.item())
    agree = bool(torch.allclose(grad, g_stored, rtol=1e-4, atol=1e-4) and
                 abs(l_rec - l_stored) <= 1e-4 * abs(l_stored))
    print(f"[bench] loss stored {l_stored:.6g} recompute {l_rec:.6g} agree={agree} "
          f"gsun {s_stored.tolist()}", file=sys.stderr, flush=True)

    # a few untimed descent steps (Adam on the albedo): the gradients lower the loss
    losses = []
    opt = torch.optim.Adam([albedo], lr=0.05)
    for _ in range(8):
        stored()
        torch.cuda.synchronize()
        losses.append(float(loss.item()))
        albedo.grad = grad
        opt.step()
    torch.cuda.synchronize()

    world_b = world.device_bytes()
    alb_b = albedo.numel() * 4
    img_b = npix * 16
    resident = {
        "stored_bytes": world_b + 2 * alb_b + 2 * img_b + npix * 12,
        "recompute_bytes": world_b + 2 * alb_b + 2 * img_b,
    }
    rays = npix + nhit
    # backward: 12 B record + 16 B target per pixel, 12 B albedo + 12 B gradient per hit
    bwd_bytes = 28 * npix + 24 * nhit
    fwd_bytes = (lookup_steps(ex) + OUT_BYTES_PER_PIXEL * npix + 12 * nhit)
    fwd_gbs = fwd_bytes / (lambert_ms * 1e-3) / 1e9
    line = {
        "metric": f"fwd+bwd ms per view ({scene}: {'1024' if batched else '256'}^3, "
                  f"{nviews} views) vs plain forward",
        "value": round(stored_ms / nviews, 4),
        "unit": "ms/view",
        "n_gpus": 1,
        "steps": steps,
        "warmup": warm,
        "ms_per_step": round(stored_ms, 4),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": (f"{scene}: {'1024^3 terrain x8 + 32000 sprites' if batched else '256^3 terrain x2 + 2400 blocks'}, "
                         f"{nviews} views, {npix} pixels, Lambert x shadow, L2 to ground-truth "
                         f"renders, albedo {'compact (' + str(n_occ) + ' occupied voxels)' if batched else 'dense'}"),
            "launch": "persistent batch, hipGraph" if batched else "one vt_render per view",
            "plain_fwd_ms_per_view": round(plain_ms / nviews, 4),
            "lambert_fwd_ms_per_view": round(lambert_ms / nviews, 4),
            "backward_ms_per_view": round(bwd_ms / nviews, 4),
            "stored_fwd_bwd_ms_per_view": round(stored_ms / nviews, 4),
            "recompute_fwd_bwd_ms_per_view": round(rec_ms / nviews, 4),
            "fwd_bwd_over_plain": round(stored_ms / plain_ms, 3),
            "recompute_over_plain": round(rec_ms / plain_ms, 3),
            "modes_agree": agree,
            "loss_descent": [round(x, 2) for x in losses],
            "resident_hbm_bytes": resident,
            "torch_peak_bytes": int(torch.cuda.max_memory_allocated()),
            "backward_gbs": round(bwd_bytes / (bwd_ms * 1e-3) / 1e9, 2),
            "rays": rays,
            "dda_steps": st["primary_steps"] + st["shadow_steps"],
            "free_region_test": free,
            "executed_steps": ex["primary_steps"] + ex["shadow_steps"],
            "lookup_steps": lookup_steps(ex),
        },
        "roofline": {"bound": "hbm", "achieved": round(fwd_gbs, 2), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(fwd_gbs / HBM_PEAK_GBS, 5), "traffic": None},
        "cpu_baseline": None,
    }
    if batched:
        for b in (plain_b, fwd_b, rec_b):
            b.close()
    world.close()
    ctx.close()
    return line


AMBIENT_DIFF = 0.2

TRAFFIC_FILE = os.path.join(ROOT, "profiles", "traffic.json")


def build_provenance():
    """Which libvt.so this run loaded and whether it is the build of the
    sources beside it (the build record unnamed-voxel-tracer_amd/build.py
    writes next to the library)."""
    import build as B
    p = B.provenance()
    rec = p.get("record") or {}
    return {"lib": p["lib"], "built_from_current_sources": p["built_from_current_sources"],
            "sources_sha256_16": (rec.get("sources_sha256") or "")[:16] or None,
            "libvt_sha256_16": (rec.get("libvt_sha256") or "")[:16] or None,
            "hipcc": rec.get("hipcc")}


def measured_traffic(key):
    """HBM bytes per launch of a kernel/workload as measured with rocprofv3
    FETCH_SIZE and WRITE_SIZE passes, scaled by the copy-kernel calibration
    (profiles/traffic.json, written from the profile summaries); None if the
    workload has not been measured."""
    try:
        with open(TRAFFIC_FILE) as f:
            t = json.load(f)
        e = t["kernels"][key]
        return int(e["fetch_bytes"] * t["fetch_scale"] + e["write_bytes"] * t["write_scale"])
    except (OSError, KeyError, ValueError):
        return None


def measured_issue(key, kernel_ms):
    """The VALU issue roofline of a kernel/workload: its wave64 VALU
    instructions per launch (rocprofv3 SQ_INSTS_VALU, profiles/traffic.json)
    over the kernel time measured in this run, against VALU_PEAK_GIPS; the
    lanes active per VALU instruction from SQ_THREAD_CYCLES_VALU /
    SQ_ACTIVE_INST_VALU.  None if the workload has not been measured."""
    try:
        with open(TRAFFIC_FILE) as f:
            e = json.load(f)["valu"][key]
        gips = e["insts"] / (kernel_ms * 1e-3) / 1e9
        return {"bound": "valu", "achieved": round(gips, 1), "peak": VALU_PEAK_GIPS,
                "unit": "G wave64 VALU instr/s", "frac": round(gips / VALU_PEAK_GIPS, 4),
                "lanes_active": round(e["thread_cycles"] / e["active_inst"] / 64.0, 4),
                "valu_insts": e["insts"]}
    except (OSError, KeyError, ValueError, ZeroDivisionError):
        return None


def bench_procgen(args):
    """The engine's world generator as Game.init runs it, procgen(512,
    &voxels, 0, 0) (game.zig:96-97), into a fresh empty 512^3 brickmap world
    per step (worlds created before the timed region).  vt_world_procgen
    returns when the world is written, so each call is timed on the host
    clock; it includes the device heights, the host walk over the 262,144
    columns, the uploads, the column/tree kernels and the occupancy rebuild.
    cpu_baseline: the oracle's literal restatement (one thread) of the same
    world, whose chunk table and pool must equal the device's."""
    import numpy as np
    import torch

    import vt

    ctx = vt.Context(0)
    atlas = np.zeros(256 ** 3, np.uint32)  # procgen writes blocks only
    worlds = [vt.World.empty_brickmap(ctx, 512, atlas) for _ in range(args.warmup + args.steps)]
    torch.cuda.synchronize()
    for w in worlds[:args.warmup]:
        w.procgen()
    dts = []
    for w in worlds[args.warmup:]:
        t0 = time.perf_counter()
        w.procgen()
        dts.append(time.perf_counter() - t0)
    ms = 1e3 * sum(dts) / len(dts)
    chunks, pool = worlds[-1].read_brickmap()
    cpu = None
    equal = None
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import oracle_ffi as O
        reps = 20
        t0 = time.perf_counter()
        for _ in range(reps):
            rc, oc, op = O.procgen(512)
        cdt = (time.perf_counter() - t0) / reps
        equal = bool(rc == 0 and np.array_equal(oc, chunks) and np.array_equal(op, pool))
        cpu = {"value": round(1e3 * cdt, 3), "unit": "ms per 512^3 world", "cores": 1,
               "kind": "port",
               "sample": f"the same world generated {reps} times by the oracle's literal "
                         f"restatement, 1 thread, {cdt * reps:.2f} s wall"}
    line = {
        "metric": "ms per 512^3 world, reference procgen (FastNoiseLite fBm terrain, "
                  "decorations, trees) into the engine's chunk/pool layout",
        "value": round(ms, 4),
        "unit": "ms",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 4),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+u32",
        "data": "synthetic",
        "config": {
            "workload": "procgen(512, 0, 0) into an empty VoxelBrickmap(512, 8): 262,144 "
                        f"columns, {len(pool) // 512} pool blocks",
            "min_ms": round(1e3 * min(dts), 4),
            "columns_per_s": round(512 * 512 / (ms * 1e-3)),
            "chunks_and_pool_equal_oracle": equal,
        },
        "roofline": None,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    for w in worlds:
        w.close()
    ctx.close()


def bench_engine(args):
    """The engine's own frame configuration: Game.init's world
    (procgen(512, &voxels, 0, 0) into the chunk/pool layout, game.zig:96-97)
    with the model atlas of its .vox assets (the placed texels kept in
    tests/golden/atlas_assets.npz), its start camera (player (256, 22, 256)
    + 3 up, identity rotation, fov pi/2), the G-buffer size
    GBuffer.init(1280, 720) (game.zig:91), the shaders' own step caps (192
    primary, primary.comp.glsl:43; 48 shadow, secondary.comp.glsl:41), its
    rgba8 G-buffer images (gbuffer.zig:10-13: the shadow origin uses the
    stored normal) and SUN_DIR (camera.glsl:8).  A step is one vt_render
    with the world resident in HBM."""
    import math

    import numpy as np
    import torch

    import vt
    from vt import _native as N

    g = np.load(os.path.join(ROOT, "tests", "golden", "atlas_assets.npz"))
    atlas = np.zeros(256 ** 3, np.uint32)
    atlas[g["index"]] = g["value"]
    ctx = vt.Context(0)
    world = vt.World.empty_brickmap(ctx, 512, atlas)
    world.procgen()
    cam = N.Camera()
    cam.position[:] = [256.0, 25.0, 256.0, 0.0]
    cam.view[:] = [1.0, 0, 0, 0, 0, 1.0, 0, 0, 0, 0, 1.0, 0, 0, 0, 0, 1.0]
    cam.fov = float(np.float32(math.pi / 2.0))
    W, H, sun = 1280, 720, SUN_DIR
    prm = vt.params(W, H, 192, 48, sun, vt.VT_FLAG_GBUF_UNORM8)
    fs = vt.Frame.alloc(W, H, stats=True)
    vt.render(ctx, world, cam, fs, prm)
    torch.cuda.synchronize()
    st = fs.stats_dict()
    f = vt.Frame.alloc(W, H)
    stream = torch.cuda.current_stream()
    ms = timed(torch, stream, lambda: vt.render(ctx, world, cam, f, prm), args.steps, args.warmup)
    rays = W * H + st["primary_hits"]
    cpu = None
    same = None
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import oracle_ffi as O
        chunks, pool = world.read_brickmap()
        ow = O.BrickmapWorld(512, chunks, pool, atlas)
        threads = min(len(os.sched_getaffinity(0)), 16)
        t0 = time.perf_counter()
        o = O.render(ow, cam, W, H, 192, 48, sun, flags=vt.VT_FLAG_GBUF_UNORM8,
                     threads=threads, gbuffer=False)
        dt = time.perf_counter() - t0
        same = bool(np.array_equal(f.hits.cpu().numpy().view(np.uint32), o["hits"]))
        cpu = {"value": round((W * H + int(o["stats"][2])) / dt / 1e6, 3), "unit": "Mrays/s",
               "cores": threads, "kind": "port",
               "sample": f"the same full frame, {threads} threads, {dt:.2f} s wall"}
    # lower bound of the timed kernel's reads (see bench_brickmap)
    abytes = LOWER_BYTES_PER_STEP * (st["primary_steps"] + st["shadow_steps"]) + \
        OUT_BYTES_PER_PIXEL * W * H
    literal_bytes = 4 * (st["l1_fetches"] + st["l0_fetches"]) + OUT_BYTES_PER_PIXEL * W * H
    achieved = abytes / (ms * 1e-3) / 1e9
    line = {
        "metric": "Mrays/s primary+shadow, the engine's own frame (Game.init world and camera, "
                  "1280x720 G-buffer, shader step caps 192/48, rgba8 G-buffer, SUN_DIR)",
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
        "data": "the reference's generator and .vox models",
        "config": {
            "workload": "procgen(512, 0, 0) world in the chunk/pool layout + 29-slot model "
                        "atlas; start camera; 1280x720; caps 192/48; rgba8 G-buffer; SUN_DIR",
            "primary_hits": st["primary_hits"],
            "dda_steps": st["primary_steps"] + st["shadow_steps"],
            "hits_equal_oracle": same,
        },
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5),
                     "traffic": measured_traffic("k_render_brickmap/engine/1280x720"),
                     "issue": measured_issue("k_render_brickmap/engine/1280x720", ms),
                     "bytes": "lower bound: 1 B per step + 28 B per pixel",
                     "literal_loop_gbs": round(literal_bytes / (ms * 1e-3) / 1e9, 2)},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    world.close()
    ctx.close()


def bench_copy(args):
    """Calibration for the HBM counters: one plain device-to-device copy of
    1 GiB (torch's copy kernel) per step -- 1 GiB read + 1 GiB written."""
    import torch
    n = 1 << 28
    src = torch.ones(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    stream = torch.cuda.current_stream()
    ms = timed(torch, stream, lambda: dst.copy_(src), args.steps, args.warmup)
    nbytes = 2 * n * 4
    print(json.dumps({"metric": "copy GB/s (HBM counter calibration)",
                      "value": round(nbytes / (ms * 1e-3) / 1e9, 2), "unit": "GB/s",
                      "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
                      "ms_per_step": round(ms, 4), "higher_is_better": True,
                      "bytes_read": n * 4, "bytes_written": n * 4}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None,
                    help="timed steps (default: 200 for single frames, 20 for the batched and "
                         "differentiable workloads, 10 for procgen)")
    ap.add_argument("--warmup", type=int, default=None,
                    help="untimed steps first (default: 100 for single frames -- the GPU's "
                         "clocks settle within the first ~50 frames of 0.8 ms --, 2 for "
                         "procgen, 20 for the others)")
    ap.add_argument("--traversal", choices=TRAVERSALS, default="flat")
    ap.add_argument("--workload",
                    choices=["frame", "batched", "diff", "diff_batched", "brickmap", "procgen",
                             "engine", "copy"],
                    default="frame",
                    help="frame: 512^3, one 1080p view (the headline); batched: 1024^3, 64 "
                         "uneven cameras in one persistent graph launch; diff: 256^3, 16 "
                         "views fwd+bwd; diff_batched: 1024^3 compact albedo, 48 views fwd+bwd; "
                         "brickmap: the headline world in the engine's own layout; "
                         "procgen: the engine's 512^3 world generator; "
                         "engine: the engine's own frame (its world, camera and step caps); "
                         "copy: 1 GiB device copy (HBM counter calibration)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-fwd-bwd", action="store_true",
                    help="skip the C3 fwd+bwd block of the default line")
    args = ap.parse_args()
    frame_like = args.workload in ("frame", "brickmap", "engine")
    if args.steps is None:
        args.steps = 200 if frame_like else (10 if args.workload == "procgen" else 20)
    if args.warmup is None:
        args.warmup = 100 if frame_like else (2 if args.workload == "procgen" else 20)
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported "
              "(rays are independent: a multi-GPU run would shard views/tiles with no exchange)",
              file=sys.stderr)
        sys.exit(2)
    if args.workload == "batched":
        return bench_batched(args)
    if args.workload == "diff":
        return bench_diff(args, "c3")
    if args.workload == "diff_batched":
        return bench_diff(args, "c4")
    if args.workload == "copy":
        return bench_copy(args)
    if args.workload == "brickmap":
        return bench_brickmap(args)
    if args.workload == "procgen":
        return bench_procgen(args)
    if args.workload == "engine":
        return bench_engine(args)

    import numpy as np
    import torch

    import vt
    from vt import scenes as S

    # the metric's second clause, measured in the same run: C3 fwd+bwd per
    # view against the plain forward (its own world and views)
    fwd_bwd = None
    if not args.no_fwd_bwd:
        d = measure_diff(args, "c3")
        c = d["config"]
        fwd_bwd = {
            "workload": c["workload"],
            "plain_fwd_ms_per_view": c["plain_fwd_ms_per_view"],
            "stored_fwd_bwd_ms_per_view": c["stored_fwd_bwd_ms_per_view"],
            "recompute_fwd_bwd_ms_per_view": c["recompute_fwd_bwd_ms_per_view"],
            "backward_ms_per_view": c["backward_ms_per_view"],
            "stored_over_plain": c["fwd_bwd_over_plain"],
            "recompute_over_plain": c["recompute_over_plain"],
            "modes_agree": c["modes_agree"],
        }

    print("[bench] generating 512^3 world", file=sys.stderr, flush=True)
    grid = S.generate(S.spec("c1"))
    pal = S.palette()
    cam, W, H = S.views("c1")[0]
    max_steps = S.MAX_STEPS["c1"]
    ctx = vt.Context(0)
    world = vt.World.dense(ctx, grid, pal)
    trav = TRAVERSALS.index(args.traversal)

    f = vt.Frame.alloc(W, H)
    prm = vt.params(W, H, max_steps, max_steps, S.SUN, 0, trav)
    # counted launches for rays and bytes: traceMap's steps (statistics
    # kernel variant) and, when the timed launch takes the free-region test
    # (the default for this world and these caps), the steps it executes
    free = vt.free_map_used(world, f, prm)
    job = [(cam, W, H, lambda fl: vt.params(W, H, max_steps, max_steps, S.SUN, fl, trav))]
    stats = counted_stats(vt, torch, ctx, world, job, False)
    ex = counted_stats(vt, torch, ctx, world, job, True) if free else stats
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
    # the same frame in the fixed launch order (a frame without a tile
    # schedule), reported beside the timed figure
    ff = vt.Frame.alloc(W, H, schedule=False)
    fixed_ms = timed(torch, stream, lambda: vt.render(ctx, world, cam, ff, prm), args.steps,
                     max(1, args.warmup))
    print(f"[bench] fixed launch order: kernel mean {fixed_ms:.3f} ms", file=sys.stderr,
          flush=True)
    del ff

    npix = W * H
    rays = npix + stats["primary_hits"]
    value = rays / (ms * 1e-3) / 1e6
    abytes = algorithmic_bytes(ex, args.traversal, npix)
    achieved = abytes / (ms * 1e-3) / 1e9
    steps_total = stats["primary_steps"] + stats["shadow_steps"]
    executed = ex["primary_steps"] + ex["shadow_steps"]
    key = "flat_free" if free else args.traversal

    cpu = parity = None
    if not args.no_cpu:
        print("[bench] cpu baseline (oracle, one frame)", file=sys.stderr, flush=True)
        cpu, o = cpu_baseline(grid, pal, cam, W, H, max_steps, S.SUN)
        parity = frame_parity(np, f.hits.cpu().numpy().view(np.uint32), f.shaded.cpu().numpy(), o)
        print(f"[bench] parity {parity}", file=sys.stderr, flush=True)

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
            "free_region_test": free,
            "executed_steps": executed,
            "lookup_steps": lookup_steps(ex) if args.traversal == "flat" else None,
            "tile_schedule": "blocks launched costliest first, in the order the previous "
                             "render into the frame measured (vt_frame.schedule)",
            "kernel_ms_fixed_order": round(fixed_ms, 4),
        },
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 2),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 5),
            "traffic": measured_traffic(f"k_render_dense/{key}/512cube_1080p"),
            "issue": measured_issue(f"k_render_dense/{key}/512cube_1080p", ms),
        },
        "cpu_baseline": cpu,
        "parity": parity,
        "fwd_bwd": fwd_bwd,
        "build": build_provenance(),
    }
    print(json.dumps(line), flush=True)
    world.close()
    ctx.close()


if __name__ == "__main__":
    main()
