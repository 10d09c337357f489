"""A PLY mesh reader (ply.rs:1-430): ``vertex`` positions (required),
normals (when ``nx``, ``ny``, ``nz`` are all present) and UVs (``u``/``v``,
``s``/``t`` or ``texture_u``/``texture_v``), and ``face`` index lists,
fan-triangulated. ASCII and binary little-endian; big-endian is refused by
name. Unknown properties and elements are read past."""

from __future__ import annotations

import struct
from dataclasses import dataclass
from pathlib import Path

import numpy as np

SCALARS = {  # name -> (struct code, size); both spelling generations
    "char": ("b", 1), "int8": ("b", 1), "uchar": ("B", 1), "uint8": ("B", 1),
    "short": ("h", 2), "int16": ("h", 2), "ushort": ("H", 2), "uint16": ("H", 2),
    "int": ("i", 4), "int32": ("i", 4), "uint": ("I", 4), "uint32": ("I", 4),
    "float": ("f", 4), "float32": ("f", 4), "double": ("d", 8), "float64": ("d", 8),
}

# Past this many vertices or faces a count is treated as corrupt, before
# anything is allocated for it.
MAX_COUNT = 1 << 28


class PlyError(ValueError):
    pass


@dataclass
class Ply:
    positions: np.ndarray          # (V, 3) f32
    normals: np.ndarray | None     # (V, 3) f32
    uvs: np.ndarray | None         # (V, 2) f32
    triangles: np.ndarray          # (T, 3) u32


@dataclass
class _Property:
    name: str
    scalar: str
    count_scalar: str | None = None  # list properties


def _scalar(name):
    if name not in SCALARS:
        raise PlyError(f"unknown property type {name!r}")
    return name


def parse(data: bytes) -> Ply:
    end = data.find(b"end_header")
    if not data.startswith(b"ply") or end < 0:
        raise PlyError("missing the ply magic or end_header")
    nl = data.find(b"\n", end)
    body = data[nl + 1:] if nl >= 0 else b""
    header = data[:end].decode("ascii", "replace").replace("\r", "").split("\n")
    if header[0].strip() != "ply":
        raise PlyError("missing the ply magic")
    binary = None
    elements = []  # (name, count, [props])
    for line in header[1:]:
        words = line.split()
        if not words or words[0] in ("comment", "obj_info"):
            continue
        if words[0] == "format":
            if len(words) < 2:
                raise PlyError("format line without a format")
            if words[1] == "ascii":
                binary = False
            elif words[1] == "binary_little_endian":
                binary = True
            else:
                raise PlyError(f"unsupported format {words[1]}")
        elif words[0] == "element":
            count = int(words[2])
            if count < 0 or count > MAX_COUNT:
                raise PlyError(f"element {words[1]} declares {count} entries")
            elements.append((words[1], count, []))
        elif words[0] == "property":
            if not elements:
                raise PlyError("property before any element")
            if words[1] == "list":
                elements[-1][2].append(_Property(words[4], _scalar(words[3]), _scalar(words[2])))
            else:
                elements[-1][2].append(_Property(words[2], _scalar(words[1])))
    if binary is None:
        raise PlyError("no format line")
    reader = _Binary(body) if binary else _Ascii(body)
    positions = normals = uvs = None
    tris = []
    for name, count, props in elements:
        if name == "vertex":
            names = [p.name for p in props]
            if any(p.count_scalar for p in props) or not all(c in names for c in ("x", "y", "z")):
                raise PlyError("vertex element needs scalar x, y and z")
            values = np.array([[reader.scalar(p.scalar) for p in props] for _ in range(count)], np.float64)
            values = values.reshape(count, len(props))
            col = {n: k for k, n in enumerate(names)}
            positions = values[:, [col["x"], col["y"], col["z"]]].astype(np.float32)
            if all(c in col for c in ("nx", "ny", "nz")):
                normals = values[:, [col["nx"], col["ny"], col["nz"]]].astype(np.float32)
            for u, v in (("u", "v"), ("s", "t"), ("texture_u", "texture_v")):
                if u in col and v in col:
                    uvs = values[:, [col[u], col[v]]].astype(np.float32)
                    break
        elif name == "face":
            for _ in range(count):
                for p in props:
                    if p.count_scalar and p.name in ("vertex_indices", "vertex_index"):
                        n = int(reader.scalar(p.count_scalar))
                        idx = [int(reader.scalar(p.scalar)) for _ in range(n)]
                        for k in range(1, n - 1):
                            tris.append((idx[0], idx[k], idx[k + 1]))
                    else:
                        reader.skip(p)
        else:
            for _ in range(count):
                for p in props:
                    reader.skip(p)
    if positions is None:
        raise PlyError("no vertex element")
    triangles = np.array(tris, np.int64).reshape(-1, 3)
    if triangles.size and (triangles.min() < 0 or triangles.max() >= positions.shape[0]):
        raise PlyError("a face index is out of range")
    return Ply(positions, normals, uvs, triangles.astype(np.uint32))


class _Binary:
    def __init__(self, data):
        self.data, self.at = data, 0

    def scalar(self, ty):
        code, size = SCALARS[ty]
        if self.at + size > len(self.data):
            raise PlyError("data ends early")
        (v,) = struct.unpack_from("<" + code, self.data, self.at)
        self.at += size
        return v

    def skip(self, p):
        if p.count_scalar:
            n = int(self.scalar(p.count_scalar))
            for _ in range(n):
                self.scalar(p.scalar)
        else:
            self.scalar(p.scalar)


class _Ascii:
    def __init__(self, data):
        self.words = data.decode("ascii", "replace").split()
        self.at = 0

    def scalar(self, ty):
        if self.at >= len(self.words):
            raise PlyError("data ends early")
        w = self.words[self.at]
        self.at += 1
        try:
            return float(w) if SCALARS[ty][0] in "fd" else int(w)
        except ValueError as e:
            raise PlyError(f"bad number {w!r}") from e

    def skip(self, p):
        if p.count_scalar:
            n = int(self.scalar(p.count_scalar))
            self.at += n
        else:
            self.at += 1


def read(path) -> Ply:
    path = Path(path)
    try:
        data = path.read_bytes()
    except OSError as e:
        raise PlyError(f'PLY "{path}": {e}') from e
    try:
        return parse(data)
    except PlyError as e:
        raise PlyError(f'PLY "{path}": {e}') from e
