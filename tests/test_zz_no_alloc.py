"""vt_batch_render_noalloc through the C-ABI: the batch entry point that
allocates nothing (its view table and queue counters live in the library's
fixed per-GPU device array) must give, view for view, the outputs of
vt_batch_create + vt_batch_launch bit for bit -- on two calls in a row, the
second with fewer and smaller views (the array keeps the first call's
contents), with statistics, and with a Lambert shading mode; more views
than the array holds are refused before any launch."""
import ctypes as C

import numpy as np
import pytest

import vt
from vt import _native as N
from vt import scenes as S

pytestmark = pytest.mark.gpu

FIRST = [(256, 256), (123, 77), (8, 8), (3, 5), (300, 17), (64, 200), (512, 96), (40, 40),
         (17, 300), (96, 96), (1, 1), (200, 120)]
SECOND = [(31, 9), (64, 64), (5, 70)]


def cameras(n, seed):
    rng = np.random.default_rng(seed)
    out = [S.views("c0")[0][0]]
    for _ in range(n - 1):
        eye = rng.uniform(-30, 160, 3)
        eye[1] = rng.uniform(90, 140)
        out.append(vt.look_at(eye, rng.uniform(20, 108, 3), 70.0))
    return out


def table(cams, frames):
    t = (N.View * len(frames))()
    for i, (cam, f) in enumerate(zip(cams, frames)):
        t[i].camera = cam
        t[i].width, t[i].height = f.width, f.height
        t[i].frame = f.struct()
    return t


def fresh(sizes):
    fs = [vt.Frame.alloc(W, H, steps=True) for (W, H) in sizes]
    for f in fs:
        f.shaded.fill_(float("nan"))
        f.hits.fill_(-7)
        f.steps.fill_(-7)
    return fs


def run_pair(ctx, world, sizes, seed, prm, tune):
    import torch
    cams = cameras(len(sizes), seed)
    a, b = fresh(sizes), fresh(sizes)
    sa = torch.zeros(8, dtype=torch.int64, device="cuda")
    sb = torch.zeros(8, dtype=torch.int64, device="cuda")
    t = table(cams, a)
    s = torch.cuda.current_stream().cuda_stream
    print(f"[noalloc] {len(sizes)} views, seed {seed}, flags {prm.flags}", flush=True)
    N.check(N.lib().vt_batch_render_noalloc(ctx.h, world.h, t, len(sizes), C.byref(prm),
                                            C.byref(tune), C.c_void_p(sa.data_ptr()),
                                            C.c_void_p(s)), "vt_batch_render_noalloc")
    del t  # the records went to the device as kernel arguments
    ref = vt.Batch(ctx, world, list(zip(cams, b)), prm, tune, stats=sb)
    ref.launch()
    torch.cuda.synchronize()
    for i, (fa, fb) in enumerate(zip(a, b)):
        for k in ("hits", "steps", "shaded"):
            x = getattr(fa, k).cpu().numpy().view(np.uint32)
            y = getattr(fb, k).cpu().numpy().view(np.uint32)
            assert np.array_equal(x, y), f"view {i} {sizes[i]}: {k} differs"
    assert torch.equal(sa, sb), (sa, sb)
    assert int(sa[0]) > 0
    ref.close()


def test_noalloc_batch_matches_allocating_batch(ctx):
    import torch
    grid = S.generate(S.spec("c0"))
    world = vt.World.dense(ctx, grid, S.palette())
    tune = vt.tuning(tile=vt.VT_TILE_16X4, chunk=2)
    prm = vt.params(0, 0, 512, 512)
    run_pair(ctx, world, FIRST, 5, prm, tune)
    run_pair(ctx, world, SECOND, 6, prm, tune)  # fewer, smaller views: stale table entries
    alb = torch.rand(128 ** 3 * 3, device="cuda")
    lam = vt.params(0, 0, 512, 512, S.SUN, vt.VT_FLAG_LAMBERT, 0, 0.25, alb)
    run_pair(ctx, world, SECOND, 7, lam, vt.tuning(tile=vt.VT_TILE_8X8, chunk=1))
    # more views than the array holds: refused before any launch
    many = [(8, 8)] * (vt.VT_NOALLOC_MAX_VIEWS + 1)
    fs = fresh(many[:1]) * len(many)
    t = table(cameras(1, 8) * len(many), fs)
    st = N.lib().vt_batch_render_noalloc(ctx.h, world.h, t, len(many), C.byref(prm),
                                         C.byref(tune), None, None)
    assert st == N.VT_ERR_INVALID_ARGUMENT
    assert "256" in N.lib().vt_last_error().decode()
    world.close()
