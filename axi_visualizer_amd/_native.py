"""ctypes binding of libaxr.so (the C ABI in include/axr.h).

The library is built in-tree by ``axi_visualizer_amd._build.build_native``.
There is no fallback: if the library is missing or cannot be loaded this
module raises, so nothing silently runs a different path.
"""
from __future__ import annotations

import ctypes
import os
import weakref

import numpy as np

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libaxr.so")

_f = ctypes.c_float
_i = ctypes.c_int
_i64 = ctypes.c_int64
_p = ctypes.c_void_p
# axr_batch_render outputs (include/axr.h AXR_OUT_*)
OUT_RGBA8, OUT_IDS, OUT_HDR = 0, 1, 2

_ctx = ctypes.c_void_p

_u32 = ctypes.c_uint32
_batch = ctypes.c_void_p


class AxrJob(ctypes.Structure):
    """axr_job of include/axr.h."""
    _fields_ = [("scene", ctypes.c_void_p), ("pos", ctypes.c_float * 3),
                ("vm", ctypes.c_float * 9), ("scale", ctypes.c_float),
                ("cax", ctypes.c_float), ("cay", ctypes.c_float), ("width", ctypes.c_int),
                ("height", ctypes.c_int), ("shadows", ctypes.c_int),
                ("pcf_taps", ctypes.c_int), ("exposure", ctypes.c_float)]


# name -> (restype, argtypes); every function declared in include/axr.h.
SIGNATURES = {
    "axr_last_error": (ctypes.c_char_p, []),
    "axr_create": (_i, [_i, _i, ctypes.POINTER(_ctx)]),
    "axr_destroy": (_i, [_ctx]),
    "axr_set_scale_cull": (_i, [_ctx, _f, _f, _f]),
    "axr_add_texture_group": (_i, [_ctx, _p, _p, _p, _i, _p, _p, _p, _i64, _i, _i,
                                   ctypes.POINTER(_i)]),
    "axr_set_group_material": (_i, [_ctx, _i, _f, _i]),
    "axr_set_primary_light": (_i, [_ctx, _p, _p]),
    "axr_vert_light": (_i, [_ctx, _p, _i, _p, _p, _i, _p, _p, _i, _p, _p, _p, _i, _f]),
    "axr_set_pos": (_i, [_ctx, _p]),
    "axr_set_vm": (_i, [_ctx, _p]),
    "axr_clear_zbuffer": (_i, [_ctx]),
    "axr_set_clear_color": (_i, [_ctx, _p]),
    "axr_draw_all": (_i, [_ctx, _p, _p, _i, _p, _i, _i]),
    "axr_add_shadow_map": (_i, [_ctx, _i, _i, _f, _f]),
    "axr_place_shadow_map": (_i, [_ctx, _i, _p, _p, _f]),
    "axr_clear_shadow_map": (_i, [_ctx, _i]),
    "axr_shadow_map": (_i, [_ctx, _i, _p, _i, _f]),
    "axr_gamma": (_i, [_ctx, _f]),
    "axr_blur": (_i, [_ctx]),
    "axr_check_arith": (_i, [ctypes.c_uint64, ctypes.c_uint32, _p]),
    "axr_set_refl_tex": (_i, [_ctx, _i, _p, _p, _p, _i]),
    "axr_get_refl_tex": (_i, [_ctx, _i, _p, _p, _p]),
    "axr_draw_ssr": (_i, [_ctx, _p, _p, _i, _i, _i, _f, _i]),
    "axr_bloom": (_i, [_ctx, _i, _f, _f, _f]),
    "axr_bloom_planes_noalloc": (_i, [_p, _p, _p, _i, _i, _i, _f, _f, _f, _p]),
    "axr_get_frame": (_i, [_ctx, _p, _p, _p, _p]),
    "axr_get_rgba8": (_i, [_ctx, _p]),
    "axr_get_ids": (_i, [_ctx, _p]),
    "axr_get_shadow_map": (_i, [_ctx, _i, _p]),
    "axr_sync": (_i, [_ctx]),
    "axr_host_alloc": (_i, [ctypes.c_uint64, ctypes.POINTER(_p)]),
    "axr_host_free": (_i, [_p]),
    "axr_get_stats": (_i, [_ctx, _p]),
    "axr_set_bin_capacity": (_i, [_ctx, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32,
                                  ctypes.c_uint32]),
    "axr_set_timing": (_i, [_ctx, _i]),
    "axr_set_shaders": (_i, [_ctx, _p, _p, _i]),
    "axr_get_texels": (_i, [_ctx, _i, _p, _p, _p]),
    "axr_get_vertex_light": (_i, [_ctx, _i, _p]),
    "axr_get_bytes": (_i, [_ctx, ctypes.POINTER(ctypes.c_uint64)]),
    "axr_batch_create": (_i, [ctypes.POINTER(_batch)]),
    "axr_batch_destroy": (_i, [_batch]),
    "axr_batch_configure": (_i, [_batch, _i, _u32, _u32, _i]),
    "axr_batch_set_spill_capacity": (_i, [_batch, _u32]),
    "axr_batch_set_timing": (_i, [_batch, _i]),
    "axr_batch_get_timing": (_i, [_batch, _p, _p]),
    "axr_batch_set_stats": (_i, [_batch, _i]),
    "axr_batch_get_visible": (_i, [_batch, _p, _i]),
    "axr_batch_render": (_i, [_batch, ctypes.POINTER(AxrJob), _i, _i]),
    "axr_batch_query": (_i, [_batch]),
    "axr_batch_wait": (_i, [_batch, ctypes.POINTER(ctypes.c_float)]),
    "axr_batch_get_rgba8": (_i, [_batch, _i, _p]),
    "axr_batch_get_depth": (_i, [_batch, _i, _p]),
    "axr_batch_get_ids": (_i, [_batch, _i, _p]),
    "axr_batch_get_hdr": (_i, [_batch, _i, _p, _p, _p]),
    "axr_render_cube": (_i, [_batch, _ctx, _i, _p, _i, _i, _i]),
    "axr_batch_get_bytes": (_i, [_batch, ctypes.POINTER(ctypes.c_uint64)]),
    "axr_get_timing": (_i, [_ctx, _p, _i, _p, _p, _i]),
}

_lib = None


# error codes of include/axr.h
AXR_EINVAL, AXR_EHIP, AXR_ESTATE, AXR_ECAP = -1, -2, -3, -4


class AxrError(RuntimeError):
    """A libaxr call returned an error code (``code``: an AXR_E* value)."""

    def __init__(self, msg: str, code: int = 0):
        super().__init__(msg)
        self.code = code


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is not built; run __graft_entry__.build() (hipcc, gfx950)")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(rc: int, what: str) -> int:
    if rc < 0:
        msg = lib().axr_last_error()
        raise AxrError(f"{what} failed ({rc}): {msg.decode() if msg else ''}", rc)
    return rc


def pinned_array(shape, dtype) -> np.ndarray:
    """A numpy array in page-locked host memory (axr_host_alloc); the memory
    is released when the array and every view of it are gone."""
    dt = np.dtype(dtype)
    n = int(np.prod(shape)) * dt.itemsize
    p = ctypes.c_void_p()
    check(lib().axr_host_alloc(max(n, 1), ctypes.byref(p)), "axr_host_alloc")
    buf = (ctypes.c_uint8 * max(n, 1)).from_address(p.value)
    weakref.finalize(buf, _host_free, p.value)
    return np.frombuffer(buf, dt, count=int(np.prod(shape))).reshape(shape)


def _host_free(addr: int):
    if _lib is not None:
        _lib.axr_host_free(ctypes.c_void_p(addr))


def ptr(a) -> int:
    """Address of a C-contiguous numpy array (or None)."""
    if a is None:
        return None
    if not a.flags["C_CONTIGUOUS"]:
        raise ValueError("array must be C-contiguous")
    return a.ctypes.data
