"""Wavefront OBJ input: the mesh loader of the reference's examples (ObjData.cpp:69-200),
for feeding meshes to the draw path.

``load_obj(path)`` (or ``parse_obj(text)``) returns what ``ObjData::loadFromFile(path)
.toVertexArray(vdata, idata)`` produces:

* ``vertices``: float32 ``[n, 8]`` rows ``{x, y, z, nx, ny, nz, u, v}`` — the layout of
  ``ObjData::VertexArrayData`` (ObjData.h:34-38), 32 bytes per vertex, so the array can be
  bound with stride 32 and drawn with the ``SWR_VS_MVP_POS_UV`` vertex program (position at
  byte 0, texture coordinate at byte 24);
* ``indices``: int32 ``[3 * triangles]``, each face fanned around its first corner.

Rules kept from the reference: position, normal and texture-coordinate lists are 1-based
with an implicit zero entry 0, so a missing ``/t`` or ``/n`` component picks the zero vector;
a distinct (position, normal, texcoord) triple becomes one output vertex, numbered in order
of first appearance (ObjData.cpp:166-182); a face ``v0 v1 ... vk`` becomes the triangles
``(v0, v(j-1), vj)`` for j = 2..k (ObjData.cpp:153-163); lines whose first word is not
``v``, ``vn``, ``vt`` or ``f`` are ignored.  Numbers are read like ``std::istream >>``:
the longest numeric prefix of the next word, converted with ``strtof`` (correctly rounded
to float); a missing or unreadable component is 0 (the reference leaves a texture
coordinate's missing components uninitialised, ObjData.cpp:95-98 with vector_math.h:165).

Where the reference has undefined behaviour this module raises ``ValueError``: an index
past its list (or negative, which ``>> unsigned`` wraps), a face with no corners, more than
two slashes in a corner.
"""
from __future__ import annotations

import ctypes
import re
from typing import Tuple

import numpy as np

_libc = ctypes.CDLL(None)
_strtof = _libc.strtof
_strtof.restype = ctypes.c_float
_strtof.argtypes = [ctypes.c_char_p, ctypes.c_void_p]

# what std::num_get accepts for a float: sign, digits, point, digits, exponent
_FLOAT = re.compile(r"[+-]?(?:\d+\.?\d*|\.\d+)(?:[eE][+-]?\d+)?")
_UINT = re.compile(r"[+-]?\d+")


def _floats(words, n):
    """n floats as `iss >> a >> b >> ...` reads them: after the first unreadable or
    partially read word the stream has failed and the rest stay 0."""
    out = [0.0] * n
    for k in range(n):
        if k >= len(words):
            break
        m = _FLOAT.match(words[k])
        if not m:
            break
        out[k] = float(_strtof(m.group(0).encode(), None))
        if m.end() != len(words[k]):
            break
    return out


def _uint(s: str) -> int:
    """`istringstream(s) >> unsigned`, 0 when nothing is read (ObjData.cpp:57-67)."""
    m = _UINT.match(s)
    if not m:
        return 0
    v = int(m.group(0))
    if v < 0 or v > 0xFFFFFFFF:
        raise ValueError(f"OBJ index {m.group(0)} is out of range")
    return v


def _corner(word: str) -> Tuple[int, int, int]:
    """`v/t/n` -> (v, n, t) as getIndicesFromString (ObjData.cpp:39-68) fills them."""
    parts = word.split("/")
    if len(parts) > 3:
        raise ValueError(f"OBJ face corner {word!r} has more than two slashes")
    parts += [""] * (3 - len(parts))
    return _uint(parts[0]), _uint(parts[2]), _uint(parts[1])


def parse_obj(text: str) -> Tuple[np.ndarray, np.ndarray]:
    pos = [(0.0, 0.0, 0.0)]
    nrm = [(0.0, 0.0, 0.0)]
    tex = [(0.0, 0.0)]
    faces = []
    for line in text.split("\n"):
        words = line.split()
        if not words:
            continue
        cmd = words[0]
        if cmd[0] == "#":
            continue
        if cmd == "v":
            pos.append(tuple(_floats(words[1:], 3)))
        elif cmd == "vn":
            nrm.append(tuple(_floats(words[1:], 3)))
        elif cmd == "vt":
            tex.append(tuple(_floats(words[1:], 2)))
        elif cmd == "f":
            face = [_corner(w) for w in words[1:]]
            if not face:
                raise ValueError("OBJ face with no corners")
            faces.append(face)

    index_of = {}
    rows = []

    def add(c):
        k = index_of.get(c)
        if k is not None:
            return k
        vi, ni, ti = c
        if vi >= len(pos) or ni >= len(nrm) or ti >= len(tex):
            raise ValueError(f"OBJ corner {vi}/{ti}/{ni} indexes past its list")
        k = len(index_of)
        index_of[c] = k
        rows.append(pos[vi] + nrm[ni] + tex[ti])
        return k

    idx = []
    for face in faces:
        i1 = add(face[0])
        for j in range(2, len(face)):
            i2 = add(face[j - 1])
            i3 = add(face[j])
            idx += [i1, i2, i3]
    vertices = np.asarray(rows, np.float32).reshape(-1, 8)
    return vertices, np.asarray(idx, np.int32)


def load_obj(path: str) -> Tuple[np.ndarray, np.ndarray]:
    with open(path, "r", errors="surrogateescape") as f:
        return parse_obj(f.read())
