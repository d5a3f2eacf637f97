"""Python mirror of the reference's host API over the C ABI (include/swr_hip.h).

Names and argument meaning follow the reference: ``Rasterizer`` (Rasterizer.h:52-141),
``VertexProcessor`` (VertexProcessor.h:56-91), and shader classes whose class attributes
are the reference's statics (``InterpolateZ``, ``InterpolateW``, ``AVarCount``,
``PVarCount``; ``AttribCount``, ``AVarCount``, ``PVarCount``).  A shader's body is a device
program id plus its uniforms and render targets, read when a draw is issued, as the
reference reads a shader's static members when ``drawPixel`` runs.

Buffers may be torch tensors (device or host), numpy arrays, or raw pointers with a size.
"""
from __future__ import annotations

import ctypes as C
from typing import Any, Optional

import numpy as np

from . import _abi

# enums (same values as the reference's enum classes)
class RasterMode:
    Span, Block, Adaptive = 0, 1, 2


class DrawMode:
    Point, Line, Triangle = 0, 1, 2


class CullMode:
    None_, CCW, CW = 0, 1, 2


# program ids
VS_DUMMY, VS_RAW, VS_POS_RGB, VS_MVP_POS_RGB_UV, VS_MVP_POS_UV, VS_MODULE = 0, 1, 2, 3, 4, 5
(PS_NULL, PS_DEPTH_CHECKER, PS_RGB, PS_COVERAGE, PS_HASH, PS_LOG, PS_VISIBILITY, PS_TEXTURE,
 PS_MODULE) = range(9)

RASTERIZER_VERTEX = np.dtype([("x", "<f4"), ("y", "<f4"), ("z", "<f4"), ("w", "<f4"),
                              ("avar", "<f4", (_abi.MAX_AVARS,)),
                              ("pvar", "<f4", (_abi.MAX_PVARS,))])


def _buf(obj: Any):
    """(pointer, nbytes, keepalive) of a tensor / array / (ptr, nbytes) tuple / None."""
    if obj is None:
        return None, 0, None
    if isinstance(obj, tuple):
        return obj[0], obj[1], None
    if hasattr(obj, "data_ptr"):  # torch.Tensor
        if not obj.is_contiguous():
            raise ValueError("buffers must be contiguous")
        return obj.data_ptr(), obj.numel() * obj.element_size(), obj
    arr = np.ascontiguousarray(obj)
    return arr.ctypes.data, arr.nbytes, arr


class PixelShaderBase:
    """PixelShaderBase<T> (PixelShaderBase.h:40-137): statics default to 0 / false."""

    InterpolateZ = False
    InterpolateW = False
    AVarCount = 0
    PVarCount = 0
    Program = PS_NULL

    def targets(self):
        """(target0, target1, target2, width, height, pitch) of the program."""
        return None, None, None, 0, 0, 0

    def describe(self, keep: list) -> _abi.PixelShaderDesc:
        d = _abi.PixelShaderDesc()
        d.program = self.Program
        d.interpolate_z = int(bool(self.InterpolateZ))
        d.interpolate_w = int(bool(self.InterpolateW))
        d.avar_count, d.pvar_count = self.AVarCount, self.PVarCount
        t0, t1, t2, w, h, pitch = self.targets()
        ptrs = []
        for t in (t0, t1, t2):
            p, _, k = _buf(t)
            keep.append(k)
            ptrs.append(p)
        d.target0, d.target1, d.target2 = ptrs
        d.width, d.height, d.pitch = w, h, pitch
        d.checker_cells = float(getattr(self, "checker_cells", 8.0))
        tex = getattr(self, "texture", None)
        d.texture = tex.handle if tex is not None else None
        mod = getattr(self, "module", None)
        if mod is not None:
            keep.append(mod)
            d.module = C.addressof(mod.desc)
            p, _, k = _buf(self.uniforms) if self.uniforms is not None else (None, 0, None)
            keep.append(k)
            d.uniforms = p
        log = getattr(self, "log", None)
        if log is not None:
            p, nbytes, k = _buf(log)
            keep.append(k)
            d.log = p
            d.log_capacity = nbytes // (4 * _abi.LOG_WORDS)
            p, _, k = _buf(self.log_count)
            keep.append(k)
            d.log_count = p
        return d


class NullPixelShader(PixelShaderBase):
    """NullPixelShader (PixelShaderBase.h:139)."""


def _hw(t):
    shape = tuple(t.shape)
    return shape[1], shape[0], shape[1]


class DepthCheckerShader(PixelShaderBase):
    """if (z < depth) { depth = z; color = rgb * checker(uv) } into RGBA8 + fp32 depth."""

    InterpolateZ, InterpolateW, AVarCount, PVarCount = True, True, 3, 2
    Program = PS_DEPTH_CHECKER

    def __init__(self, color, depth, checker_cells: float = 8.0):
        self.color, self.depth, self.checker_cells = color, depth, checker_cells

    def targets(self):
        return (self.color, self.depth, None) + _hw(self.color)


class RGBShader(PixelShaderBase):
    """The shader of RasterizerTest.cpp:30-49: (r,g,b)*255 packed 0x00RRGGBB."""

    AVarCount = 3
    Program = PS_RGB

    def __init__(self, color):
        self.color = color

    def targets(self):
        return (self.color, None, None) + _hw(self.color)


class CoverageShader(PixelShaderBase):
    """The shader of Benchmark.cpp:14-26: buffer[p] = 1."""

    AVarCount = 3
    Program = PS_COVERAGE

    def __init__(self, buffer):
        self.buffer = buffer

    def targets(self):
        return (self.buffer, None, None) + _hw(self.buffer)


class HashShader(PixelShaderBase):
    """Order-sensitive per-pixel hash of (key, z, w, invw, avars, pvars) + invocation count."""

    Program = PS_HASH

    def __init__(self, hash_img, count_img, z=True, w=True, avars=3, pvars=2):
        self.hash_img, self.count_img = hash_img, count_img
        self.InterpolateZ, self.InterpolateW, self.AVarCount, self.PVarCount = z, w, avars, pvars

    def targets(self):
        return (self.hash_img, self.count_img, None) + _hw(self.hash_img)


class LogShader(PixelShaderBase):
    """One record per invocation: {x, y, key, seq, z, w, invw, avar[16], pvar[16]}."""

    Program = PS_LOG

    def __init__(self, count_img, log, log_count, z=True, w=True, avars=3, pvars=2):
        self.count_img, self.log, self.log_count = count_img, log, log_count
        self.InterpolateZ, self.InterpolateW, self.AVarCount, self.PVarCount = z, w, avars, pvars

    def targets(self):
        return (self.count_img, None, None) + _hw(self.count_img)


class VisibilityShader(PixelShaderBase):
    """Visibility buffer: id = source triangle + 1 (0 = none) under `z < depth`."""

    InterpolateZ = True
    Program = PS_VISIBILITY

    def __init__(self, ids=None, depth=None):
        self.ids, self.depth = ids, depth

    def targets(self):
        if self.ids is None:
            return None, None, None, 0, 0, 0
        return (self.ids, self.depth, None) + _hw(self.ids)


class Texture:
    """A texture with its mip chain on the rasterizer's device (Texture.h:13-33, 237-296):
    `texels` is uint32 [height, width] 0x00RRGGBB (numpy or torch, host or device)."""

    def __init__(self, rasterizer: "Rasterizer", texels, max_anisotropy: int = 8):
        h, w = texels.shape[:2]
        p, _, keep = _buf(texels)
        out = C.c_void_p()
        self._r = rasterizer
        rasterizer._check(rasterizer._lib.swr_texture_create(rasterizer._ctx, p, w, h, w,
                                                             max_anisotropy, C.byref(out)),
                          "Texture")
        del keep
        self.handle = out.value
        self.width, self.height = w, h

    def close(self):
        if self.handle:
            self._r._lib.swr_texture_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class TextureShader(PixelShaderBase):
    """The textured shader of the Box demo (Box.cpp:27-53): color = texture.sample(u, v,
    perspective derivatives of u and v), written to a u32 0x00RRGGBB target."""

    InterpolateW = True
    PVarCount = 2
    Program = PS_TEXTURE

    def __init__(self, color, texture: Texture):
        self.color, self.texture = color, texture

    def targets(self):
        return (self.color, None, None) + _hw(self.color)


class PixelModule:
    """A pixel program compiled outside the library (include/swr/pixel_program.h): the shared
    library at `path`, whose entry point `entry` (SWR_PIXEL_MODULE) fills the module
    descriptor.  The library stays loaded while this object lives."""

    def __init__(self, path: str, entry: str):
        self._lib = C.CDLL(path)
        fn = getattr(self._lib, entry)
        fn.restype = C.c_int
        fn.argtypes = [C.POINTER(_abi.PixelModule)]
        self.desc = _abi.PixelModule()
        rc = fn(C.byref(self.desc))
        if rc != 0:
            raise _abi.SwrError(rc, f"pixel module entry {entry} failed")
        self.path, self.entry = path, entry

    @property
    def name(self) -> str:
        return (self.desc.name or b"").decode()


class ModuleShader(PixelShaderBase):
    """A pixel module bound with its uniforms (device memory the program reads, e.g. a
    tensor holding a struct) and the size of the target area in pixels.  The statics
    (InterpolateZ, InterpolateW, AVarCount, PVarCount) are the module's."""

    Program = PS_MODULE

    def __init__(self, module: PixelModule, uniforms, width: int, height: int,
                 pitch: Optional[int] = None):
        self.module, self.uniforms = module, uniforms
        self.width, self.height = width, height
        self.pitch = width if pitch is None else pitch
        d = module.desc
        self.InterpolateZ, self.InterpolateW = bool(d.interpolate_z), bool(d.interpolate_w)
        self.AVarCount, self.PVarCount = d.avar_count, d.pvar_count

    def targets(self):
        return None, None, None, self.width, self.height, self.pitch


class VertexShaderBase:
    """VertexShaderBase<T> (VertexShaderBase.h:33-51)."""

    AttribCount = 0
    AVarCount = 0
    PVarCount = 0
    Program = VS_DUMMY
    mvp: Optional[np.ndarray] = None

    def describe(self) -> _abi.VertexShaderDesc:
        d = _abi.VertexShaderDesc()
        d.program = self.Program
        d.attrib_count, d.avar_count, d.pvar_count = self.AttribCount, self.AVarCount, self.PVarCount
        m = np.eye(4, dtype=np.float32) if self.mvp is None else np.asarray(self.mvp, np.float32)
        d.mvp[:] = [float(x) for x in m.reshape(16)]
        return d


class DummyVertexShader(VertexShaderBase):
    """DummyVertexShader (VertexShaderBase.h:53)."""


class MvpVertexShader(VertexShaderBase):
    """clip = M * (p, 1); avar = rgb; pvar = uv over {px,py,pz, r,g,b, u,v}."""

    AttribCount, AVarCount, PVarCount = 1, 3, 2
    Program = VS_MVP_POS_RGB_UV

    def __init__(self, mvp):
        self.mvp = np.asarray(mvp, np.float32)


class PosRGBVertexShader(VertexShaderBase):
    """The vertex shader of Benchmark.cpp:32-49 over {x,y,z,r,g,b}: w = 1, avar = rgb."""

    AttribCount, AVarCount, PVarCount = 1, 3, 0
    Program = VS_POS_RGB


class VertexModule:
    """A vertex program compiled outside the library (include/swr/vertex_program.h): the
    shared library at `path`, whose entry point `entry` (SWR_VERTEX_MODULE) fills the module
    descriptor.  The library stays loaded while this object lives."""

    def __init__(self, path: str, entry: str):
        self._lib = C.CDLL(path)
        fn = getattr(self._lib, entry)
        fn.restype = C.c_int
        fn.argtypes = [C.POINTER(_abi.VertexModule)]
        self.desc = _abi.VertexModule()
        rc = fn(C.byref(self.desc))
        if rc != 0:
            raise _abi.SwrError(rc, f"vertex module entry {entry} failed")
        self.path, self.entry = path, entry

    @property
    def name(self) -> str:
        return (self.desc.name or b"").decode()


class ModuleVertexShader(VertexShaderBase):
    """A vertex module bound with its uniforms (device memory the program reads, e.g. a
    tensor holding a matrix; None for none).  The statics (AttribCount, AVarCount,
    PVarCount) are the module's.  Its attribute buffers must be device memory."""

    Program = VS_MODULE

    def __init__(self, module: VertexModule, uniforms=None):
        self.module, self.uniforms = module, uniforms
        d = module.desc
        self.AttribCount, self.AVarCount, self.PVarCount = (d.attrib_count, d.avar_count,
                                                            d.pvar_count)

    def describe(self) -> _abi.VertexShaderDesc:
        d = super().describe()
        d.module = C.cast(C.pointer(self.module.desc), C.c_void_p)
        d.uniforms = _buf(self.uniforms)[0] if self.uniforms is not None else None
        return d


class RawVertexShader(VertexShaderBase):
    """Clip-space input {x,y,z,w, avar[A], pvar[P]} copied to the output."""

    Program = VS_RAW

    def __init__(self, avars=0, pvars=0):
        self.AttribCount, self.AVarCount, self.PVarCount = 1, avars, pvars


class Rasterizer:
    """Rasterizer (Rasterizer.h:52-141) on one gfx950 device."""

    def __init__(self, device: int = 0):
        self._lib = _abi.load()
        ctx = C.c_void_p()
        rc = self._lib.swr_create(C.byref(ctx), device)
        if rc != 0:
            raise _abi.SwrError(rc, f"swr_create(device={device}) failed")
        self._ctx = ctx
        self._shader: PixelShaderBase = NullPixelShader()
        try:  # order draws after torch's work on the same device
            import torch
            if torch.cuda.is_available():
                self.setStream(torch.cuda.current_stream(device).cuda_stream)
        except ImportError:
            pass

    # lifetime
    def close(self):
        if getattr(self, "_ctx", None):
            self._lib.swr_destroy(self._ctx)
            self._ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc: int, what: str):
        if rc != 0:
            msg = self._lib.swr_last_error(self._ctx)
            raise _abi.SwrError(rc, f"{what}: {msg.decode() if msg else ''}")

    # state (Rasterizer.h:75-96)
    def setRasterMode(self, mode: int):
        self._check(self._lib.swr_set_raster_mode(self._ctx, mode), "setRasterMode")

    def setScissorRect(self, x: int, y: int, width: int, height: int):
        self._check(self._lib.swr_set_scissor_rect(self._ctx, x, y, width, height),
                    "setScissorRect")

    def setPixelShader(self, shader: PixelShaderBase):
        self._shader = shader() if isinstance(shader, type) else shader
        self._apply_shader([])

    def _apply_shader(self, keep: list):
        d = self._shader.describe(keep)
        self._check(self._lib.swr_set_pixel_shader(self._ctx, C.byref(d)), "setPixelShader")

    def setStream(self, stream_handle: int):
        self._check(self._lib.swr_set_stream(self._ctx, C.c_void_p(stream_handle or None)),
                    "setStream")

    def synchronize(self):
        self._check(self._lib.swr_synchronize(self._ctx), "synchronize")

    def setProfiling(self, enable: bool):
        self._check(self._lib.swr_set_profiling(self._ctx, int(enable)), "setProfiling")

    def stats(self) -> _abi.DrawStats:
        s = _abi.DrawStats()
        self._check(self._lib.swr_get_draw_stats(self._ctx, C.byref(s)), "stats")
        return s

    def reserve(self, primitives: int, tile_pairs: int):
        self._check(self._lib.swr_reserve(self._ctx, primitives, tile_pairs), "reserve")

    def setTuning(self, raster_waves: int = 0, raster_group: int = 0, xcd_order: int = 0,
                  expand_chunk: int = 0, vertex_pass: int = 0, bin_size: int = 0,
                  waves_per_bin: int = 0, raster_queue: int = 0):
        """Launch shapes of the draw path (0 = built-in default); results never change."""
        t = _abi.Tuning(raster_waves, raster_group, xcd_order, expand_chunk, vertex_pass,
                        bin_size, waves_per_bin, raster_queue)
        self._check(self._lib.swr_set_tuning(self._ctx, C.byref(t)), "setTuning")

    def tuning(self) -> dict:
        t = _abi.Tuning()
        self._check(self._lib.swr_get_tuning(self._ctx, C.byref(t)), "tuning")
        return {"raster_waves": t.raster_waves, "raster_group": t.raster_group,
                "xcd_order": t.xcd_order, "expand_chunk": t.expand_chunk,
                "vertex_pass": t.vertex_pass, "bin_size": t.bin_size,
                "waves_per_bin": t.waves_per_bin, "raster_queue": t.raster_queue}

    # drawing through the IRasterizer seam (Rasterizer.h:99-141)
    def _list(self, fn, what, vertices, indices):
        keep: list = []
        self._apply_shader(keep)
        if isinstance(vertices, np.ndarray) and vertices.dtype != RASTERIZER_VERTEX:
            raise TypeError("vertices must use the RasterizerVertex layout")
        vp, vbytes, vk = _buf(vertices)
        ip, ibytes, ik = _buf(indices)
        keep += [vk, ik]
        self._check(fn(self._ctx, vp, vbytes // RASTERIZER_VERTEX.itemsize, ip, ibytes // 4), what)

    def drawTriangleList(self, vertices, indices):
        self._list(self._lib.swr_draw_triangle_list, "drawTriangleList", vertices, indices)

    def drawLineList(self, vertices, indices):
        self._list(self._lib.swr_draw_line_list, "drawLineList", vertices, indices)

    def drawPointList(self, vertices, indices):
        self._list(self._lib.swr_draw_point_list, "drawPointList", vertices, indices)

    def drawTriangle(self, v0, v1, v2):
        v = np.zeros(3, RASTERIZER_VERTEX)
        v[0], v[1], v[2] = v0, v1, v2
        self.drawTriangleList(v, np.arange(3, dtype=np.int32))


class VertexProcessor:
    """VertexProcessor (VertexProcessor.h:56-91); shares its Rasterizer's device context."""

    def __init__(self, rasterizer: Rasterizer):
        self.setRasterizer(rasterizer)
        self._attribs: dict = {}

    def setRasterizer(self, rasterizer: Rasterizer):
        if not isinstance(rasterizer, Rasterizer):
            raise TypeError("the gfx950 VertexProcessor feeds a gfx950 Rasterizer")
        self._r = rasterizer

    def setViewport(self, x: int, y: int, width: int, height: int):
        self._r._check(self._r._lib.swr_set_viewport(self._r._ctx, x, y, width, height),
                       "setViewport")

    def setDepthRange(self, n: float, f: float):
        self._r._check(self._r._lib.swr_set_depth_range(self._r._ctx, n, f), "setDepthRange")

    def setCullMode(self, mode: int):
        self._r._check(self._r._lib.swr_set_cull_mode(self._r._ctx, mode), "setCullMode")

    def setVertexShader(self, shader: VertexShaderBase):
        shader = shader() if isinstance(shader, type) else shader
        self._vshader = shader  # a module shader's uniforms stay referenced while bound
        d = shader.describe()
        self._r._check(self._r._lib.swr_set_vertex_shader(self._r._ctx, C.byref(d)),
                       "setVertexShader")

    def setVertexAttribPointer(self, index: int, stride: int, buffer):
        p, nbytes, k = _buf(buffer)
        self._attribs[index] = k  # the reference borrows the buffer until the draw
        self._r._check(self._r._lib.swr_set_vertex_attrib_pointer(self._r._ctx, index, stride,
                                                                  p, nbytes), "setVertexAttribPointer")

    def drawElements(self, mode: int, indices, count: Optional[int] = None):
        keep: list = []
        self._r._apply_shader(keep)
        ip, ibytes, ik = _buf(indices)
        keep.append(ik)
        n = ibytes // 4 if count is None else count
        self._r._check(self._r._lib.swr_draw_elements(self._r._ctx, mode, n, ip), "drawElements")

    def drawElementsAsync(self, mode: int, indices, count: Optional[int] = None):
        """Enqueue on the context stream; device buffers only (see swr_draw_elements_async)."""
        keep: list = []
        self._r._apply_shader(keep)
        ip, ibytes, ik = _buf(indices)
        n = ibytes // 4 if count is None else count
        self._r._check(self._r._lib.swr_draw_elements_async(self._r._ctx, mode, n, ip),
                       "drawElementsAsync")


class Batch:
    """Many independent drawElements(Triangle) calls as one pass (swr_batch_create).

    Each scene is a dict with keys: indices, vertices (attribute 0), stride, mvp (16 floats,
    row-major), viewport (x, y, w, h), depth_range (n, f), cull_mode, raster_mode, scissor
    (x, y, w, h) and targets (t0, t1, t2) of width x height (pitch = width).  Buffers are
    contiguous device tensors.  The vertex and pixel programs are the rasterizer's at draw
    time (set them with setVertexShader / setPixelShader first)."""

    def __init__(self, rasterizer: Rasterizer, scenes: list):
        self._r = rasterizer
        self._keep = []
        arr = (_abi.Scene * max(1, len(scenes)))()
        for i, sc in enumerate(scenes):
            d = arr[i]
            ip, ib, ik = _buf(sc["indices"])
            vp, vb, vk = _buf(sc["vertices"])
            self._keep += [ik, vk]
            d.indices, d.index_count = ip, ib // 4
            d.attrib0, d.attrib0_bytes, d.attrib0_stride = vp, vb, int(sc["stride"])
            d.cull_mode = int(sc.get("cull_mode", CullMode.CW))
            m = np.asarray(sc.get("mvp", np.eye(4)), np.float32).reshape(16)
            for k in range(16):
                d.mvp[k] = float(m[k])
            for k, v in enumerate(sc["viewport"]):
                d.viewport[k] = int(v)
            d.depth_near, d.depth_far = sc.get("depth_range", (0.0, 1.0))
            d.raster_mode = int(sc.get("raster_mode", RasterMode.Span))
            for k, v in enumerate(sc["scissor"]):
                d.scissor[k] = int(v)
            t = list(sc["targets"]) + [None] * (3 - len(sc["targets"]))
            ptrs = []
            for x in t:
                p, _, kx = _buf(x)
                self._keep.append(kx)
                ptrs.append(p)
            d.target0, d.target1, d.target2 = ptrs
            h, w = sc["size"]
            d.width, d.height, d.pitch = int(w), int(h), int(w)
        self._h = C.c_void_p()
        rasterizer._check(rasterizer._lib.swr_batch_create(rasterizer._ctx, arr, len(scenes),
                                                            C.byref(self._h)), "Batch")

    def draw(self, async_: bool = False):
        r = self._r
        r._apply_shader([])
        r._check(r._lib.swr_batch_draw(r._ctx, self._h, int(async_)), "Batch.draw")

    def close(self):
        if getattr(self, "_h", None):
            self._r._lib.swr_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
