"""TEST INFRASTRUCTURE ONLY: ctypes access to the CPU oracle and the reference build.

``liboracle.so`` is the plain-C restatement of trenki2/SoftwareRenderer's draw path
(``swr_oracle.c``); ``_ref/libswr_ref.so`` is the reference's own ``src/renderer`` sources
compiled with ``ref_driver.cpp`` (built only where ``/root/reference`` exists).  Both expose
the same entry points, so tests call them interchangeably through :class:`CpuRenderer`.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline leg import this
module, and only as the checker / the CPU baseline; the product never does.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# SWR_ORACLE_LIB: another build of the same sources (the sanitizer build of `make san`)
ORACLE_SO = os.environ.get("SWR_ORACLE_LIB") or os.path.join(HERE, "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libswr_ref.so")
REF_OMP_SO = os.path.join(HERE, "_ref", "libswr_ref_omp.so")
REF_OBJ_SO = os.path.join(HERE, "_ref", "libswr_refobj.so")
REF_SRC = "/root/reference/src/renderer"

MAX_AVARS = 16
MAX_PVARS = 16
LOG_WORDS = 7 + MAX_AVARS + MAX_PVARS
HASH_SEED = 2166136261

# enums / program ids of include/swr_hip.h
RASTER_SPAN, RASTER_BLOCK, RASTER_ADAPTIVE = 0, 1, 2
DRAW_POINT, DRAW_LINE, DRAW_TRIANGLE = 0, 1, 2
CULL_NONE, CULL_CCW, CULL_CW = 0, 1, 2
VS_DUMMY, VS_RAW, VS_POS_RGB, VS_MVP_POS_RGB_UV, VS_MVP_POS_UV = 0, 1, 2, 3, 4
PS_NULL, PS_DEPTH_CHECKER, PS_RGB, PS_COVERAGE, PS_HASH, PS_LOG, PS_VISIBILITY, PS_TEXTURE = range(8)


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


class PipelineState(C.Structure):
    _fields_ = [("raster_mode", C.c_int32), ("scissor", C.c_int32 * 4),
                ("viewport", C.c_int32 * 4), ("depth_near", C.c_float),
                ("depth_far", C.c_float), ("cull_mode", C.c_int32), ("reserved", C.c_int32),
                ("vs", VertexShaderDesc), ("ps", PixelShaderDesc),
                ("attrib", C.c_void_p * 8), ("stride", C.c_int32 * 8)]


RASTERIZER_VERTEX = np.dtype([("x", "<f4"), ("y", "<f4"), ("z", "<f4"), ("w", "<f4"),
                              ("avar", "<f4", (MAX_AVARS,)), ("pvar", "<f4", (MAX_PVARS,))])


def build(quiet: bool = True) -> None:
    """Compile liboracle.so (and _ref/ when the reference sources are present)."""
    out = subprocess.run(["make", "-C", HERE, "all"], capture_output=True, text=True)
    if out.returncode != 0:
        raise RuntimeError("oracle build failed:\n" + out.stdout + out.stderr)
    if not quiet:
        print(out.stdout)


def ref_available() -> bool:
    return os.path.exists(REF_SO)


class Targets:
    """Host render targets for one pixel program (what the program reads and writes)."""

    def __init__(self, program: int, width: int, height: int, log_capacity: int = 0,
                 depth_clear: float = 1.0):
        self.program, self.width, self.height = program, width, height
        shape = (height, width)
        self.t0 = self.t1 = self.log = None
        self.log_count = np.zeros(1, np.uint64)
        if program == PS_DEPTH_CHECKER:
            self.t0 = np.zeros(shape, np.uint32)
            self.t1 = np.full(shape, depth_clear, np.float32)
        elif program in (PS_RGB, PS_COVERAGE, PS_TEXTURE):
            self.t0 = np.zeros(shape, np.uint32)
        elif program == PS_VISIBILITY:
            self.t0 = np.zeros(shape, np.uint32)
            self.t1 = np.full(shape, depth_clear, np.float32)
        elif program == PS_HASH:
            self.t0 = np.full(shape, HASH_SEED, np.uint32)
            self.t1 = np.zeros(shape, np.uint32)
        elif program == PS_LOG:
            self.t0 = np.zeros(shape, np.uint32)
            self.log = np.zeros((max(log_capacity, 1), LOG_WORDS), np.uint32)
        self.log_capacity = log_capacity

    def log_records(self) -> np.ndarray:
        n = int(min(self.log_count[0], self.log_capacity))
        return self.log[:n]


def _ptr(a):
    return None if a is None else a.ctypes.data


class Pipeline:
    """Everything Rasterizer + VertexProcessor hold for a draw, in reference terms."""

    def __init__(self, width=640, height=480, raster_mode=RASTER_SPAN, cull_mode=CULL_CW,
                 scissor=None, viewport=None, depth_range=(0.0, 1.0),
                 vs_program=VS_MVP_POS_RGB_UV, vs_counts=(1, 3, 2), mvp=None,
                 ps_program=PS_HASH, ps_z=True, ps_w=True, ps_a=3, ps_p=2, checker_cells=8.0,
                 texture=None):
        self.width, self.height = width, height
        self.raster_mode, self.cull_mode = raster_mode, cull_mode
        self.scissor = tuple(scissor) if scissor is not None else (0, 0, width, height)
        self.viewport = tuple(viewport) if viewport is not None else (0, 0, width, height)
        self.depth_range = depth_range
        self.vs_program, self.vs_counts = vs_program, tuple(vs_counts)
        self.mvp = np.eye(4, dtype=np.float32) if mvp is None else np.asarray(mvp, np.float32)
        self.ps_program, self.ps_z, self.ps_w = ps_program, ps_z, ps_w
        self.ps_a, self.ps_p, self.checker_cells = ps_a, ps_p, checker_cells
        self.texture = texture  # OracleTexture for PS_TEXTURE

    def state(self, targets: Targets, attribs=()):
        st = PipelineState()
        st.raster_mode = self.raster_mode
        st.scissor[:] = self.scissor
        st.viewport[:] = self.viewport
        st.depth_near, st.depth_far = self.depth_range
        st.cull_mode = self.cull_mode
        st.vs.program = self.vs_program
        st.vs.attrib_count, st.vs.avar_count, st.vs.pvar_count = self.vs_counts
        st.vs.mvp[:] = [float(x) for x in self.mvp.reshape(16)]
        ps = st.ps
        ps.program = self.ps_program
        ps.interpolate_z, ps.interpolate_w = int(self.ps_z), int(self.ps_w)
        ps.avar_count, ps.pvar_count = self.ps_a, self.ps_p
        ps.width, ps.height, ps.pitch = self.width, self.height, self.width
        ps.target0, ps.target1 = _ptr(targets.t0), _ptr(targets.t1)
        ps.log = _ptr(targets.log)
        ps.log_capacity = targets.log_capacity
        ps.log_count = _ptr(targets.log_count)
        ps.checker_cells = self.checker_cells
        ps.texture = self.texture.handle if self.texture is not None else None
        for i, (buf, stride) in enumerate(attribs):
            st.attrib[i] = buf.ctypes.data
            st.stride[i] = stride
        return st

    def targets(self, log_capacity=0) -> Targets:
        return Targets(self.ps_program, self.width, self.height, log_capacity)


class OracleTexture:
    """The oracle's texture for PS_TEXTURE: level 0 from `texels` (uint32 [h, w],
    0x00RRGGBB) and its mip chain (Texture.h:237-296)."""

    def __init__(self, texels, max_anisotropy: int = 8):
        self.texels = np.ascontiguousarray(texels, np.uint32)
        self.max_anisotropy = max_anisotropy
        h, w = self.texels.shape
        lib = C.CDLL(ORACLE_SO)
        lib.swr_oracle_texture_create.restype = C.c_void_p
        lib.swr_oracle_texture_create.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int]
        lib.swr_oracle_texture_destroy.argtypes = [C.c_void_p]
        self._lib = lib
        self.handle = lib.swr_oracle_texture_create(self.texels.ctypes.data, w, h, w,
                                                    max_anisotropy)
        if not self.handle:
            raise RuntimeError("swr_oracle_texture_create failed")
        lib.swr_oracle_texture_sample.restype = C.c_uint32
        lib.swr_oracle_texture_sample.argtypes = [C.c_void_p] + [C.c_float] * 6
        lib.swr_oracle_texture_level.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p,
                                                 C.c_void_p]

    def sample(self, u, v, dudx, dvdx, dudy, dvdy) -> int:
        """Texture::sample (Texture.h:35-118): 0x00RRGGBB."""
        return int(self._lib.swr_oracle_texture_sample(self.handle, u, v, dudx, dvdx, dudy,
                                                       dvdy))

    def level(self, k: int) -> np.ndarray:
        p, w, h = C.c_void_p(), C.c_int(), C.c_int()
        if self._lib.swr_oracle_texture_level(self.handle, k, C.byref(p), C.byref(w),
                                              C.byref(h)) != 0:
            raise IndexError(k)
        return np.ctypeslib.as_array(C.cast(p, C.POINTER(C.c_uint32)),
                                     (h.value, w.value)).copy()

    @property
    def levels(self) -> int:
        k = 0
        while True:
            try:
                self.level(k)
            except IndexError:
                return k
            k += 1

    def close(self):
        if self.handle:
            self._lib.swr_oracle_texture_destroy(self.handle)
            self.handle = None

    def __del__(self):
        self.close()


class CpuRenderer:
    """The oracle (kind='oracle'), the reference build (kind='ref') or the reference's
    OpenMP build (kind='ref_omp')."""

    def __init__(self, kind: str = "oracle"):
        path = {"oracle": ORACLE_SO, "ref": REF_SO, "ref_omp": REF_OMP_SO}[kind]
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} is not built (run oracle.build())")
        self.kind = kind
        self.lib = C.CDLL(path)
        pre = "swr_oracle_" if kind == "oracle" else "swr_ref_"
        sig = [C.POINTER(PipelineState), C.c_int, C.c_size_t, C.c_void_p]
        self._elements = getattr(self.lib, pre + "draw_elements")
        self._elements.argtypes, self._elements.restype = sig, C.c_int
        lsig = [C.POINTER(PipelineState), C.c_void_p, C.c_size_t, C.c_void_p, C.c_size_t]
        self._lists = {}
        for name in ("triangle", "line", "point"):
            f = getattr(self.lib, pre + f"draw_{name}_list")
            f.argtypes, f.restype = lsig, C.c_int
            self._lists[name] = f
        self._setup = getattr(self.lib, pre + "triangle_setup")
        self._setup.argtypes = [C.c_void_p] * 3 + [C.c_int, C.c_int, C.c_void_p]
        self._setup.restype = C.c_int
        self._clip = getattr(self.lib, pre + "clip_triangle")
        self._clip.argtypes = [C.c_void_p] * 3 + [C.c_int, C.c_int, C.c_void_p, C.c_int]
        self._clip.restype = C.c_int

    def draw_elements(self, pipe: Pipeline, targets: Targets, attribs, mode, indices):
        idx = np.ascontiguousarray(indices, dtype=np.int32)
        attribs = [(np.ascontiguousarray(b), s) for b, s in attribs]
        st = pipe.state(targets, attribs)
        rc = self._elements(C.byref(st), mode, idx.size, idx.ctypes.data)
        if rc != 0:
            raise RuntimeError(f"{self.kind} draw_elements failed: {rc}")
        return targets

    def draw_list(self, pipe: Pipeline, targets: Targets, prim: str, vertices, indices):
        v = np.ascontiguousarray(vertices, dtype=RASTERIZER_VERTEX)
        idx = np.ascontiguousarray(indices, dtype=np.int32)
        st = pipe.state(targets)
        rc = self._lists[prim](C.byref(st), v.ctypes.data, v.size, idx.ctypes.data, idx.size)
        if rc != 0:
            raise RuntimeError(f"{self.kind} draw_{prim}_list failed: {rc}")
        return targets

    def triangle_setup(self, v0, v1, v2, na, np_):
        out = np.zeros(1 + 9 + 3 * (2 + na + np_), np.float32)
        vs = [np.ascontiguousarray(v, dtype=RASTERIZER_VERTEX) for v in (v0, v1, v2)]
        k = self._setup(vs[0].ctypes.data, vs[1].ctypes.data, vs[2].ctypes.data, na, np_,
                        out.ctypes.data)
        return out[: (k if k else 10)]

    def clip_triangle(self, v0, v1, v2, na, np_, max_out=64):
        out = np.zeros(max_out, RASTERIZER_VERTEX)
        vs = [np.ascontiguousarray(v, dtype=RASTERIZER_VERTEX) for v in (v0, v1, v2)]
        n = self._clip(vs[0].ctypes.data, vs[1].ctypes.data, vs[2].ctypes.data, na, np_,
                       out.ctypes.data, max_out)
        if n < 0:
            raise RuntimeError(f"clip failed {n}")
        return out[: min(n, max_out)], n


def log2_rounded(x: float) -> float:
    """The oracle's level-of-detail log2 (one rounding to float)."""
    lib = C.CDLL(ORACLE_SO)
    lib.swr_oracle_log2.restype = C.c_float
    lib.swr_oracle_log2.argtypes = [C.c_float]
    return float(lib.swr_oracle_log2(x))


def set_log2_libm(on: bool) -> None:
    """Level of detail by the C library's log2f (the reference's call) when on, else by
    log2 rounded once (the device's choice, the default); process-wide."""
    lib = C.CDLL(ORACLE_SO)
    lib.swr_oracle_set_log2_libm.argtypes = [C.c_int]
    lib.swr_oracle_set_log2_libm(int(bool(on)))


def ref_obj_available() -> bool:
    return os.path.exists(REF_OBJ_SO)


def ref_obj_load(path: str):
    """The reference's own OBJ loader (ObjData::loadFromFile + toVertexArray, compiled from
    src/examples/ObjData.cpp): (vertices float32 [n, 8], indices int32)."""
    lib = C.CDLL(REF_OBJ_SO)
    f = lib.swr_ref_obj_load
    f.argtypes = [C.c_char_p, C.c_void_p, C.c_size_t, C.c_void_p, C.c_size_t,
                  C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)]
    f.restype = C.c_int
    nv, ni = C.c_size_t(), C.c_size_t()
    f(path.encode(), None, 0, None, 0, C.byref(nv), C.byref(ni))
    v = np.zeros((nv.value, 8), np.float32)
    idx = np.zeros(ni.value, np.int32)
    rc = f(path.encode(), v.ctypes.data, nv.value, idx.ctypes.data, ni.value, C.byref(nv),
           C.byref(ni))
    if rc != 0:
        raise RuntimeError(f"swr_ref_obj_load failed: {rc}")
    return v, idx


def diff_reference(clip, tris, adj, attr, ids, depth, grad_image=None, culled=True):
    """FP64 reference of swr_diff_forward / swr_diff_backward (oracle/swr_diff_ref.c).

    clip [V,N,4] f32, tris [T,3] i32, adj [T,3] i32, attr [N,C] f32, ids [V,H,W] u32,
    depth [V,H,W] f32, grad_image [V,H,W,C] f32 or None.  Returns (image [V,H,W,C] f64,
    grad_clip [V,N,4] f64 or None, grad_attr [N,C] f64 or None)."""
    lib = C.CDLL(ORACLE_SO)
    f = lib.swr_oracle_diff
    f.restype = C.c_int
    f.argtypes = [C.c_int] * 7 + [C.c_void_p] * 10
    clip = np.ascontiguousarray(clip, np.float32)
    tris = np.ascontiguousarray(tris, np.int32)
    adj = np.ascontiguousarray(adj, np.int32)
    attr = np.ascontiguousarray(attr, np.float32)
    ids = np.ascontiguousarray(ids, np.uint32)
    depth = np.ascontiguousarray(depth, np.float32)
    V, H, W = ids.shape
    N, Cc = attr.shape
    image = np.zeros((V, H, W, Cc), np.float64)
    gclip = gattr = gi = None
    if grad_image is not None:
        gi = np.ascontiguousarray(grad_image, np.float32)
        gclip = np.zeros((V, N, 4), np.float64)
        gattr = np.zeros((N, Cc), np.float64)
    rc = f(V, W, H, N, tris.shape[0], Cc, int(bool(culled)), clip.ctypes.data, tris.ctypes.data, adj.ctypes.data,
           attr.ctypes.data, ids.ctypes.data, depth.ctypes.data, image.ctypes.data,
           None if gi is None else gi.ctypes.data, None if gclip is None else gclip.ctypes.data,
           None if gattr is None else gattr.ctypes.data)
    if rc != 0:
        raise RuntimeError(f"swr_oracle_diff failed: {rc}")
    return image, gclip, gattr
