"""Host-side argument checks of the Python layer (vt/render.py), run without
a GPU: every mismatch between a call's sizes and the buffers behind its
device pointers must raise before anything reaches the C-ABI (the library's
entry points take raw pointers and cannot see buffer sizes themselves).

Device tensors are stood in for by a small object with the attributes the
checks read; the C library is replaced by one that fails the test if any
entry point is called."""
from types import SimpleNamespace

import numpy as np
import pytest

import vt
from vt import _native as N


class FakeTensor:
    """What check_tensor reads of a torch tensor."""

    def __init__(self, numel, dtype="float32", device=0, contiguous=True, cuda=True):
        self._n, self.dtype, self._c = numel, dtype, contiguous
        self.is_cuda = cuda
        self.device = SimpleNamespace(index=device if cuda else None, type="cuda" if cuda else "cpu")

    def numel(self):
        return self._n

    def is_contiguous(self):
        return self._c

    def data_ptr(self):
        return 0x1000


@pytest.fixture
def no_launch(monkeypatch):
    class Trap:
        def __getattr__(self, name):
            raise AssertionError(f"{name} reached the C-ABI past a failed check")

    monkeypatch.setattr(N, "lib", lambda: Trap())


CTX = SimpleNamespace(device=0, h=None)


def dense_world(dim=16, n_occupied=None):
    return SimpleNamespace(kind="dense", dim=dim, n_occupied=n_occupied, h=None)


def frame(W, H, **over):
    f = vt.Frame(W, H)
    f.shaded = FakeTensor(W * H * 4)
    f.hits = FakeTensor(W * H * 3, "int32")
    for k, v in over.items():
        setattr(f, k, v)
    return f


def test_well_formed_call_reaches_the_library(monkeypatch):
    called = []

    class Lib:
        def vt_render(self, *a):
            called.append(a)
            return 0

    monkeypatch.setattr(N, "lib", lambda: Lib())
    f = frame(8, 4)
    vt.render(CTX, dense_world(), N.Camera(), f, vt.params(8, 4), stream=0)
    assert len(called) == 1


def test_params_size_must_match_frame(no_launch):
    with pytest.raises(ValueError, match="params are 16x4"):
        vt.render(CTX, dense_world(), N.Camera(), frame(8, 4), vt.params(16, 4))


@pytest.mark.parametrize("name,bad", [
    ("shaded", FakeTensor(8 * 4 * 4 - 1)),
    ("hits", FakeTensor(8 * 4 * 3 - 3, "int32")),
    ("hits", FakeTensor(8 * 4 * 3, "float32")),
    ("shaded", FakeTensor(8 * 4 * 4, contiguous=False)),
    ("shaded", FakeTensor(8 * 4 * 4, cuda=False)),
    ("shaded", FakeTensor(8 * 4 * 4, device=1)),
    ("color", FakeTensor(8 * 4 * 4 - 4)),
    ("steps", FakeTensor(8 * 4 - 1, "int32")),
    ("display", FakeTensor(8 * 4 - 1, "int32")),
    ("stats", FakeTensor(7, "int64")),
    ("target", FakeTensor(8 * 4 * 4 - 1)),
])
def test_frame_buffers_checked(no_launch, name, bad):
    with pytest.raises((ValueError, TypeError), match=name):
        vt.render(CTX, dense_world(), N.Camera(), frame(8, 4, **{name: bad}), vt.params(8, 4))


def test_albedo_length_dense_and_compact(no_launch):
    world = dense_world(dim=16)
    f = frame(8, 4)
    short = FakeTensor(16 ** 3 * 3 - 1)
    with pytest.raises(ValueError, match="albedo"):
        vt.render(CTX, world, N.Camera(), f,
                  vt.params(8, 4, flags=N.VT_FLAG_LAMBERT, albedo=short))
    # compact: the index must be built, and the length is 3 x occupied voxels
    prm = vt.params(8, 4, flags=N.VT_FLAG_LAMBERT | N.VT_FLAG_ALBEDO_COMPACT,
                    albedo=FakeTensor(300))
    with pytest.raises(ValueError, match="compact index"):
        vt.render(CTX, world, N.Camera(), f, prm)
    world.n_occupied = 101
    with pytest.raises(ValueError, match="albedo"):
        vt.render(CTX, world, N.Camera(), f, prm)
    grad = vt.params(8, 4, flags=N.VT_FLAG_LAMBERT | N.VT_FLAG_GRAD_L2,
                     albedo=FakeTensor(16 ** 3 * 3), grad_albedo=FakeTensor(16 ** 3 * 3 - 3),
                     grad_sun=FakeTensor(3), loss=FakeTensor(1))
    with pytest.raises(ValueError, match="grad_albedo"):
        vt.render(CTX, world, N.Camera(), frame(8, 4, target=FakeTensor(128)), grad)
    grad = vt.params(8, 4, flags=N.VT_FLAG_LAMBERT | N.VT_FLAG_GRAD_L2,
                     albedo=FakeTensor(16 ** 3 * 3), grad_albedo=FakeTensor(16 ** 3 * 3),
                     grad_sun=FakeTensor(2), loss=FakeTensor(1))
    with pytest.raises(ValueError, match="grad_sun"):
        vt.render(CTX, world, N.Camera(), frame(8, 4, target=FakeTensor(128)), grad)


def test_hand_made_params_with_pointers_rejected(no_launch):
    prm = N.RenderParams()
    prm.width, prm.height, prm.albedo = 8, 4, 0x1000
    with pytest.raises(ValueError, match="vt.params"):
        vt.render(CTX, dense_world(), N.Camera(), frame(8, 4), prm)


def test_lambert_calls_checked(no_launch):
    world = dense_world(dim=16)
    hits = FakeTensor(3 * 100, "int32")
    prm = vt.params(10, 10, flags=N.VT_FLAG_LAMBERT, albedo=FakeTensor(16 ** 3 * 3))
    with pytest.raises(ValueError, match="out"):
        vt.lambert_shade(CTX, world, hits, prm, FakeTensor(4 * 100 - 1))
    with pytest.raises(ValueError, match="whole 3-word records"):
        vt.lambert_shade(CTX, world, FakeTensor(301, "int32"), prm, FakeTensor(400))
    with pytest.raises(ValueError, match="upstream"):
        vt.lambert_backward(CTX, world, hits, prm, FakeTensor(399), N.VT_UPSTREAM_GRADIENT,
                            FakeTensor(16 ** 3 * 3))
    with pytest.raises(ValueError, match="grad_albedo"):
        vt.lambert_backward(CTX, world, hits, prm, FakeTensor(400), N.VT_UPSTREAM_GRADIENT,
                            FakeTensor(16 ** 3 * 3 - 1))
    with pytest.raises(ValueError, match="loss"):
        vt.lambert_backward(CTX, world, hits, prm, FakeTensor(400), N.VT_UPSTREAM_L2_TARGET,
                            FakeTensor(16 ** 3 * 3), FakeTensor(3), FakeTensor(0))


def test_brickmap_arrays_checked(no_launch):
    atlas = np.zeros(256 ** 3, np.uint32)
    with pytest.raises(ValueError, match="chunk entries"):
        vt.World.brickmap(CTX, 64, np.zeros(511, np.uint32), np.zeros(0, np.uint32), atlas)
    with pytest.raises(ValueError, match="whole 8\\^3 blocks"):
        vt.World.brickmap(CTX, 64, np.zeros(512, np.uint32), np.zeros(100, np.uint32), atlas)
    with pytest.raises(ValueError, match="atlas"):
        vt.World.brickmap(CTX, 64, np.zeros(512, np.uint32), np.zeros(0, np.uint32), atlas[:-1])
    with pytest.raises(ValueError, match="multiple of 8"):
        vt.World.brickmap(CTX, 60, np.zeros(343, np.uint32), np.zeros(0, np.uint32), atlas)


def test_batch_views_checked(no_launch):
    good, bad = frame(8, 8), frame(8, 8, shaded=FakeTensor(10))
    prm = vt.params(8, 8)
    with pytest.raises(ValueError, match="view 1"):
        vt.Batch(CTX, dense_world(), [(N.Camera(), good), (N.Camera(), bad)], prm, vt.tuning())
    with pytest.raises(ValueError, match="view 0"):
        vt.autotune(CTX, dense_world(), [(N.Camera(), bad)], prm)


def test_present_checked(no_launch):
    comp = FakeTensor(4 * 8 * 8)
    comp.shape = (8, 8, 4)
    with pytest.raises(ValueError, match="display"):
        vt.present(CTX, comp, out=None, display=FakeTensor(63, "int32"), stream=0)
