"""MI355X voxel ray traversal: Python host side over libvt.so (include/vt.h)."""
from . import _native, camera, scenes  # noqa: F401
from ._native import (VT_FLAG_GBUF_UNORM8, VT_FLAG_LAMBERT, VT_TRAVERSAL_BITMASK,  # noqa: F401
                      VT_TRAVERSAL_BITMASK8, VT_TRAVERSAL_FLAT, VtError, lib)
from .camera import look_at  # noqa: F401
from ._native import VT_NOALLOC_MAX_VIEWS, VT_TILE_8X8, VT_TILE_16X4, VT_TILE_32X2  # noqa: F401
from ._native import (VT_FLAG_ALBEDO_COMPACT, VT_FLAG_GRAD_L2, VT_FLAG_PRESENT,  # noqa: F401
                      VT_UPSTREAM_GRADIENT, VT_UPSTREAM_L2_TARGET)
from ._native import VT_FLAG_FREE_MAP_OFF, VT_FLAG_FREE_MAP_ON, VT_FREE_MIN_CAP  # noqa: F401
from .render import (Batch, Context, Frame, World, autotune, build_compact_index,  # noqa: F401
                     free_map_used, lambert_backward, lambert_shade, params, present, render,
                     tuning)
