"""ctypes view of include/swr_hip.h and the loader of libswr_hip.so.

The library is built in-tree by :mod:`softwarerenderer_amd.build`.  Loading fails loudly
when it is missing: there is no CPU fallback for the draw path.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libswr_hip.so")

OK = 0
ERRORS = {-1: "SWR_ERR_INVALID_ARG", -2: "SWR_ERR_UNSUPPORTED", -3: "SWR_ERR_HIP",
          -4: "SWR_ERR_NO_DEVICE", -5: "SWR_ERR_CAPACITY", -6: "SWR_ERR_INDEX_RANGE",
          -7: "SWR_ERR_CLIP_OVERFLOW"}

MAX_AVARS = 16
MAX_PVARS = 16
MAX_VERTEX_ATTRIBS = 8
LOG_WORDS = 7 + MAX_AVARS + MAX_PVARS
HASH_SEED = 2166136261


class VertexShaderDesc(C.Structure):
    _fields_ = [("program", C.c_int32), ("attrib_count", C.c_int32),
                ("avar_count", C.c_int32), ("pvar_count", C.c_int32),
                ("mvp", C.c_float * 16), ("module", C.c_void_p), ("uniforms", C.c_void_p)]


class PixelShaderDesc(C.Structure):
    _fields_ = [("program", C.c_int32), ("interpolate_z", C.c_int32),
                ("interpolate_w", C.c_int32), ("avar_count", C.c_int32),
                ("pvar_count", C.c_int32), ("width", C.c_int32), ("height", C.c_int32),
                ("pitch", C.c_int32), ("target0", C.c_void_p), ("target1", C.c_void_p),
                ("target2", C.c_void_p), ("log", C.c_void_p), ("log_capacity", C.c_uint64),
                ("log_count", C.c_void_p), ("checker_cells", C.c_float),
                ("reserved", C.c_int32), ("texture", C.c_void_p),
                ("module", C.c_void_p), ("uniforms", C.c_void_p)]


PIXEL_MODULE_ABI = 1


VERTEX_MODULE_ABI = 1


class VertexModule(C.Structure):
    """swr_vertex_module: a vertex program compiled outside the library."""
    _fields_ = [("abi", C.c_uint32), ("attrib_count", C.c_int32),
                ("avar_count", C.c_int32), ("pvar_count", C.c_int32),
                ("shade", C.c_void_p), ("name", C.c_char_p)]


class PixelModule(C.Structure):
    """swr_pixel_module: a pixel program compiled outside the library."""
    _fields_ = [("abi", C.c_uint32), ("args_bytes", C.c_uint32),
                ("interpolate_z", C.c_int32), ("interpolate_w", C.c_int32),
                ("avar_count", C.c_int32), ("pvar_count", C.c_int32),
                ("launch_triangles", C.c_void_p), ("launch_fragments", C.c_void_p),
                ("name", C.c_char_p)]


class DrawStats(C.Structure):
    _fields_ = [("input_primitives", C.c_uint64), ("raster_primitives", C.c_uint64),
                ("tile_pairs", C.c_uint64), ("tiles", C.c_uint64),
                ("stage_ms", C.c_float * 8), ("stage_launches", C.c_uint32 * 8),
                ("pairs_emitted", C.c_uint64)]


class Scene(C.Structure):
    """swr_scene: one scene of a batched draw."""
    _fields_ = [("indices", C.c_void_p), ("index_count", C.c_uint64),
                ("attrib0", C.c_void_p), ("attrib0_stride", C.c_int32),
                ("cull_mode", C.c_int32), ("attrib0_bytes", C.c_uint64),
                ("mvp", C.c_float * 16), ("viewport", C.c_int32 * 4),
                ("depth_near", C.c_float), ("depth_far", C.c_float),
                ("raster_mode", C.c_int32), ("scissor", C.c_int32 * 4),
                ("width", C.c_int32), ("height", C.c_int32), ("pitch", C.c_int32),
                ("target0", C.c_void_p), ("target1", C.c_void_p), ("target2", C.c_void_p)]


class DiffDesc(C.Structure):
    """swr_diff_desc: differentiable shading over a visibility buffer."""
    _fields_ = [("views", C.c_int32), ("width", C.c_int32), ("height", C.c_int32),
                ("vertex_count", C.c_int32), ("triangle_count", C.c_int32),
                ("channels", C.c_int32), ("flags", C.c_int32), ("scenes", C.c_int32),
                ("clip", C.c_void_p), ("triangles", C.c_void_p),
                ("adjacency", C.c_void_p), ("attributes", C.c_void_p), ("ids", C.c_void_p),
                ("depth", C.c_void_p), ("bary", C.c_void_p), ("color", C.c_void_p),
                ("image", C.c_void_p), ("scene_views", C.c_void_p),
                ("scene_vertices", C.c_void_p), ("scene_triangles", C.c_void_p)]


class Tuning(C.Structure):
    """swr_tuning: launch shapes of the draw path (0 = default)."""
    _fields_ = [("raster_waves", C.c_int32), ("raster_group", C.c_int32),
                ("xcd_order", C.c_int32), ("expand_chunk", C.c_int32),
                ("vertex_pass", C.c_int32), ("bin_size", C.c_int32),
                ("waves_per_bin", C.c_int32), ("raster_queue", C.c_int32)]


# name -> (restype, argtypes): every entry point include/swr_hip.h declares
_P, _I, _S, _F = C.c_void_p, C.c_int, C.c_size_t, C.c_float
SIGNATURES = {
    "swr_create": (_I, [C.POINTER(C.c_void_p), _I]),
    "swr_destroy": (_I, [_P]),
    "swr_last_error": (C.c_char_p, [_P]),
    "swr_set_stream": (_I, [_P, _P]),
    "swr_synchronize": (_I, [_P]),
    "swr_set_raster_mode": (_I, [_P, _I]),
    "swr_set_scissor_rect": (_I, [_P, _I, _I, _I, _I]),
    "swr_set_pixel_shader": (_I, [_P, C.POINTER(PixelShaderDesc)]),
    "swr_draw_triangle_list": (_I, [_P, _P, _S, _P, _S]),
    "swr_draw_line_list": (_I, [_P, _P, _S, _P, _S]),
    "swr_draw_point_list": (_I, [_P, _P, _S, _P, _S]),
    "swr_set_viewport": (_I, [_P, _I, _I, _I, _I]),
    "swr_set_depth_range": (_I, [_P, _F, _F]),
    "swr_set_cull_mode": (_I, [_P, _I]),
    "swr_set_vertex_shader": (_I, [_P, C.POINTER(VertexShaderDesc)]),
    "swr_set_vertex_attrib_pointer": (_I, [_P, _I, _I, _P, _S]),
    "swr_draw_elements": (_I, [_P, _I, _S, _P]),
    "swr_draw_elements_async": (_I, [_P, _I, _S, _P]),
    "swr_reserve": (_I, [_P, _S, _S]),
    "swr_draw_elements_noalloc": (_I, [_P, _I, _S, _P]),
    "swr_get_draw_stats": (_I, [_P, C.POINTER(DrawStats)]),
    "swr_set_profiling": (_I, [_P, _I]),
    "swr_batch_create": (_I, [_P, C.POINTER(Scene), _S, C.POINTER(C.c_void_p)]),
    "swr_batch_draw": (_I, [_P, _P, _I]),
    "swr_batch_destroy": (_I, [_P]),
    "swr_diff_forward": (_I, [_P, C.POINTER(DiffDesc)]),
    "swr_diff_backward": (_I, [_P, C.POINTER(DiffDesc), _P, _P, _P]),
    "swr_set_tuning": (_I, [_P, C.POINTER(Tuning)]),
    "swr_texture_create": (_I, [_P, _P, C.c_int, C.c_int, C.c_int, C.c_int,
                                C.POINTER(C.c_void_p)]),
    "swr_texture_destroy": (_I, [_P]),
    "swr_get_tuning": (_I, [_P, C.POINTER(Tuning)]),
}

_lib = None


def load(path: str = LIB_PATH) -> C.CDLL:
    """Load libswr_hip.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is missing: run `python -m softwarerenderer_amd.build` "
                           "(the draw path has no CPU fallback)")
    lib = C.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        f = getattr(lib, name)
        f.restype, f.argtypes = res, args
    _lib = lib
    return lib


class SwrError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code
