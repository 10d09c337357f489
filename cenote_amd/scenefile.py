"""Scene files: the ``.ron`` front end (format.rs, scene/changeset.rs,
scene/description.rs, scene/lower.rs).

A scene file is ``(version: 1, ops: [...])``: an ordered list of ops applied
to an empty description. Each op patches a named object of one of seven
kinds (Mesh, Instance, Material, Light, Camera, Environment, Settings) —
created with its kind's defaults on first mention, only the patch's
``Some`` fields overwriting — or removes one (``Remove(Kind, "name")``).
References resolve after the whole set; the post-set state is validated
before it replaces the description (changeset.rs:1-23, 510-690). Relative
paths are rebased against the scene file's directory at load
(format.rs:100-113). Unknown fields are errors, never skipped.

``lower`` turns a description into the ``SceneData`` records the library
uploads (lower.rs): instances in name order, colors converted from linear
Rec.709 to ACEScg, materials clamped as lower_material does, powerless
delta lights skipped, exactly one camera, at most one environment (an
equirect EXR, converted to ACEScg with negatives clamped; none is a black
sky). Textured material slots are refused: this build's kernels read
constants only.
"""

from __future__ import annotations

import copy
import dataclasses
import math
from pathlib import Path

import numpy as np

from . import ply as ply_reader
from .ron import RonError, Some, Variant, loads
from .scene import (CameraData, Material, MeshData, OpenPBR, SceneData, _ACESCG_COLS, acescg_from_rec709, f32,
                    luminance, trs)

FORMAT_VERSION = 1
KINDS = ("Mesh", "Instance", "Material", "Light", "Camera", "Environment", "Settings")


class SceneFileError(ValueError):
    """A scene file that does not parse, apply or lower."""


# --- defaults (description.rs) ---------------------------------------------------

def _default(kind):
    if kind == "Mesh":
        return {"source": Variant("Inline", {"positions": [], "triangles": []})}
    if kind == "Instance":
        return {"mesh": "", "material": "", "transform": Variant("Trs", {}), "camera_visible": True}
    if kind == "Material":
        m = dataclasses.asdict(OpenPBR())
        for k in ("base_color", "base_metalness", "specular_roughness", "emission_color", "geometry_opacity"):
            m[k] = Variant("Constant", m[k])
        m["geometry_normal"] = None
        return m
    if kind == "Light":
        return {"light": Variant("Point", {"position": (0.0, 0.0, 0.0), "intensity": (0.0, 0.0, 0.0)})}
    if kind == "Camera":
        return {"position": (0.0, 0.0, 5.0), "look_at": (0.0, 0.0, 0.0), "up": (0.0, 1.0, 0.0),
                "vfov_degrees": 40.0, "focus_distance": None, "aperture_radius": 0.0}
    if kind == "Environment":
        return {"path": ""}
    if kind == "Settings":
        return {"resolution": (1280, 720), "spp": 64, "max_bounces": 8, "seed": 0}
    raise SceneFileError(f"unknown kind {kind}")


_FIELDS = {
    "Mesh": ("source",),
    "Instance": ("mesh", "material", "transform", "camera_visible"),
    "Material": tuple(_default("Material")),
    "Light": ("light",),
    "Camera": ("position", "look_at", "up", "vfov_degrees", "focus_distance", "aperture_radius"),
    "Environment": ("path",),
    "Settings": ("resolution", "spp", "max_bounces", "seed"),
}
# Doubly optional fields: Some(None) clears them.
_DOUBLE = {("Material", "geometry_normal"), ("Camera", "focus_distance")}


# --- parsing ----------------------------------------------------------------------

def from_ron(text: str) -> list:
    """format.rs:78-98: the ops of a scene-file text (version probed first)."""
    try:
        value = loads(text)
    except RonError as e:
        raise SceneFileError(f"not a cenote scene file (expected `(version: {FORMAT_VERSION}, ops: [...])`): {e}")
    if not isinstance(value, dict) or "version" not in value:
        raise SceneFileError(f"not a cenote scene file (expected `(version: {FORMAT_VERSION}, ops: [...])`)")
    if value["version"] != FORMAT_VERSION:
        raise SceneFileError(f"scene file is format version {value['version']}, this build reads {FORMAT_VERSION}")
    extra = set(value) - {"version", "ops"}
    if extra:
        raise SceneFileError(f"scene file parse failed: unknown field {sorted(extra)[0]!r}")
    ops = value.get("ops", [])
    if not isinstance(ops, list):
        raise SceneFileError("scene file parse failed: ops is not a list")
    for op in ops:
        _check_op(op)
    return ops


def _check_op(op):
    if not isinstance(op, Variant):
        raise SceneFileError(f"scene file parse failed: an op is not a variant: {op!r}")
    if op.name == "Remove":
        if not (isinstance(op.value, tuple) and len(op.value) == 2 and isinstance(op.value[0], Variant)
                and op.value[0].name in KINDS and isinstance(op.value[1], str)):
            raise SceneFileError("scene file parse failed: Remove takes (Kind, \"name\")")
        return
    if op.name not in KINDS:
        raise SceneFileError(f"scene file parse failed: unknown op {op.name}")
    if not isinstance(op.value, dict):
        raise SceneFileError(f"scene file parse failed: {op.name} takes a patch struct")
    for field in op.value:
        if field != "name" and field not in _FIELDS[op.name]:
            raise SceneFileError(f"scene file parse failed: unknown field `{field}` in a {op.name} patch")
    if not isinstance(op.value.get("name", ""), str):
        raise SceneFileError("scene file parse failed: name must be a string")


def paths_of(op):
    """The filesystem paths an op references, as (container, key) pairs
    (the containers are the op's own dicts: assigning rewrites the op)."""
    return _paths(op)


def _paths(op):
    """The filesystem paths an op references, as (container, key) pairs."""
    out = []
    v = op.value if isinstance(op, Variant) and isinstance(op.value, dict) else {}
    if op.name == "Mesh":
        src = v.get("source")
        if isinstance(src, Some) and isinstance(src.value, Variant) and src.value.name == "Ply":
            out.append((src.value.value, "path"))
    elif op.name == "Environment":
        if isinstance(v.get("path"), Some):
            out.append((v, "path"))
    elif op.name == "Material":
        for field, val in v.items():
            inner = val.value if isinstance(val, Some) else None
            if isinstance(inner, Some):
                inner = inner.value
            if isinstance(inner, Variant) and inner.name == "Texture" and isinstance(inner.value, dict):
                out.append((inner.value, "path"))
            elif isinstance(inner, dict) and "path" in inner:
                out.append((inner, "path"))
    return out


def load(path) -> list:
    """format.rs:100-113: parse a scene file and rebase relative paths
    against its directory."""
    path = Path(path)
    ops = from_ron(path.read_text())
    base = path.resolve().parent
    for op in ops:
        for holder, key in _paths(op):
            val = holder[key]
            if isinstance(val, Some):
                p = Path(val.value)
                if not p.is_absolute():
                    holder[key] = Some(str(base / p))
            else:
                p = Path(val)
                if not p.is_absolute():
                    holder[key] = str(base / p)
    return ops


# --- the description ----------------------------------------------------------------

class Description:
    """The seven object maps (description.rs:490) with the apply contract."""

    def __init__(self):
        self.objects = {k: {} for k in KINDS}

    def apply(self, ops):
        """changeset.rs:380-470: all ops on a working copy, then validate,
        then swap in. A rejected set leaves the description unchanged."""
        work = copy.deepcopy(self.objects)
        for op in ops:
            _check_op(op)
            if op.name == "Remove":
                kind, name = op.value[0].name, op.value[1]
                if not name:
                    raise SceneFileError(f"a {kind.lower()} op has an empty name")
                if name not in work[kind]:
                    raise SceneFileError(f'Remove targets a {kind.lower()} named "{name}" that does not exist')
                del work[kind][name]
                continue
            name = op.value.get("name", "")
            if not name:
                raise SceneFileError(f"a {op.name.lower()} op has an empty name")
            obj = work[op.name].setdefault(name, _default(op.name))
            for field, val in op.value.items():
                if field == "name" or val is None:
                    continue
                if not isinstance(val, Some):
                    raise SceneFileError(f"{op.name} field `{field}` needs Some(..) or None")
                inner = val.value
                if (op.name, field) in _DOUBLE:
                    obj[field] = inner.value if isinstance(inner, Some) else None
                else:
                    obj[field] = inner
        _validate(work)
        self.objects = work

    def kind(self, kind):
        return dict(sorted(self.objects[kind].items()))


def _vec(v, n, what):
    if not (isinstance(v, (tuple, list)) and len(v) == n and all(isinstance(x, (int, float)) for x in v)):
        raise SceneFileError(f"{what} must be {n} numbers")
    return tuple(float(x) for x in v)


def _transform_rows(t) -> np.ndarray:
    if not isinstance(t, Variant):
        raise SceneFileError("a transform is Trs(..) or Matrix(..)")
    if t.name == "Trs":
        f = t.value or {}
        extra = set(f) - {"translate", "rotate_degrees", "scale"}
        if extra:
            raise SceneFileError(f"unknown field `{sorted(extra)[0]}` in Trs")
        return trs(_vec(f.get("translate", (0, 0, 0)), 3, "translate"),
                   _vec(f.get("rotate_degrees", (0, 0, 0)), 3, "rotate_degrees"),
                   _vec(f.get("scale", (1, 1, 1)), 3, "scale"))
    if t.name == "Matrix":
        rows = t.value
        if not (isinstance(rows, (tuple, list)) and len(rows) == 3):
            raise SceneFileError("Matrix takes three rows of four numbers")
        return np.array([_vec(r, 4, "a matrix row") for r in rows], np.float32)
    raise SceneFileError(f"unknown transform {t.name}")


def _validate(objects):
    for name, mesh in sorted(objects["Mesh"].items()):
        src = mesh["source"]
        if not isinstance(src, Variant) or src.name not in ("Inline", "Ply"):
            raise SceneFileError(f'mesh "{name}": source is Inline(..) or Ply(..)')
        if src.name == "Ply":
            _validate_path(f'mesh "{name}"', (src.value or {}).get("path", ""))
            continue
        f = src.value or {}
        extra = set(f) - {"positions", "normals", "uvs", "triangles"}
        if extra:
            raise SceneFileError(f"unknown field `{sorted(extra)[0]}` in Inline")
        positions, triangles = f.get("positions", []), f.get("triangles", [])
        if not positions or not triangles:
            raise SceneFileError(f'mesh "{name}" has no geometry')
        for key in ("normals", "uvs"):
            val = f.get(key)
            val = val.value if isinstance(val, Some) else val
            if val is not None and len(val) != len(positions):
                raise SceneFileError(f'mesh "{name}" has {len(val)} {key} for {len(positions)} positions')
        if any(int(i) >= len(positions) or int(i) < 0 for tri in triangles for i in tri):
            raise SceneFileError(f'mesh "{name}" has a triangle index out of bounds')
    for name, inst in sorted(objects["Instance"].items()):
        for kind, ref in (("Mesh", inst["mesh"]), ("Material", inst["material"])):
            if not ref:
                raise SceneFileError(f'instance "{name}" was never given a {kind.lower()}')
            if ref not in objects[kind]:
                raise SceneFileError(f'instance "{name}" references a {kind.lower()} "{ref}" that does not exist')
        m = _transform_rows(inst["transform"]).astype(np.float64)
        full = np.vstack([m, [0, 0, 0, 1]])
        if not np.all(np.isfinite(full)) or abs(np.linalg.det(full[:3, :3])) == 0.0:
            raise SceneFileError(f'instance "{name}" has a non-invertible transform')
    for name, mat in sorted(objects["Material"].items()):
        for field, val in mat.items():
            if isinstance(val, Variant) and val.name == "Texture":
                _validate_path(f'a texture of material "{name}"', (val.value or {}).get("path", ""))
    for name, light in sorted(objects["Light"].items()):
        lt = light["light"]
        if not isinstance(lt, Variant) or lt.name not in ("Distant", "Point"):
            raise SceneFileError(f'light "{name}" is Distant(..) or Point(..)')
        if lt.name == "Distant" and _vec(lt.value.get("direction", (0, 0, 0)), 3, "direction") == (0.0, 0.0, 0.0):
            raise SceneFileError(f'distant light "{name}" has a zero direction')
    for name, cam in sorted(objects["Camera"].items()):
        p, a, u = (np.array(_vec(cam[k], 3, k), np.float32) for k in ("position", "look_at", "up"))
        fwd = a - p
        if not np.any(fwd):
            raise SceneFileError(f'camera "{name}": position and look_at coincide')
        if not np.any(np.cross(fwd, u)):
            raise SceneFileError(f'camera "{name}": up is parallel to the view axis')
        v = float(cam["vfov_degrees"])
        if not math.isfinite(v) or v <= 0.0 or v >= 180.0:
            raise SceneFileError(f'camera "{name}": vertical fov must be inside (0, 180) degrees, got {v}')
        fd = cam["focus_distance"]
        if fd is not None and (math.isnan(fd) or fd <= 0.0):
            raise SceneFileError(f'camera "{name}": focus distance must be positive, got {fd}')
        if math.isnan(cam["aperture_radius"]) or cam["aperture_radius"] < 0.0:
            raise SceneFileError(f'camera "{name}": aperture radius must not be negative')
    for name, env in sorted(objects["Environment"].items()):
        _validate_path(f'environment "{name}"', env["path"])
    for name, s in sorted(objects["Settings"].items()):
        if 0 in tuple(s["resolution"]):
            raise SceneFileError(f'settings "{name}": resolution has a zero dimension')
        if s["spp"] == 0 or s["max_bounces"] == 0:
            raise SceneFileError(f'settings "{name}": spp and max_bounces must be at least 1')


def _validate_path(what, path):
    if not path:
        raise SceneFileError(f"{what} has no file path")
    p = Path(path)
    if not p.is_absolute():
        raise SceneFileError(f'{what} references the relative path "{path}" — rebase against the scene directory '
                             "first (scenefile.load does)")
    if not p.is_file():
        raise SceneFileError(f'{what} references "{path}", which does not exist')


# --- lowering (lower.rs) -----------------------------------------------------------

def _constant(val, what):
    if isinstance(val, Variant) and val.name == "Constant":
        return val.value
    if isinstance(val, Variant) and val.name == "Texture":
        raise SceneFileError(f"{what} is textured; this build's kernels read constant material slots only")
    return val


def _smooth_normals(positions, triangles):
    """lower.rs:538-556: area-weighted face normals summed per vertex; a
    vertex whose sum vanishes gets +Y."""
    p = positions.astype(np.float32)
    t = triangles.astype(np.int64)
    face = np.cross(p[t[:, 1]] - p[t[:, 0]], p[t[:, 2]] - p[t[:, 0]])
    sums = np.zeros_like(p)
    for k in range(3):
        np.add.at(sums, t[:, k], face)
    length = np.linalg.norm(sums, axis=1)
    ok = (length > 0) & np.isfinite(length)
    out = np.tile(np.array([0.0, 1.0, 0.0], np.float32), (p.shape[0], 1))
    out[ok] = sums[ok] / length[ok, None]
    return out.astype(np.float32)


def _mesh(name, mesh) -> MeshData:
    src = mesh["source"]
    if src.name == "Ply":
        try:
            data = ply_reader.read(src.value["path"])
        except ply_reader.PlyError as e:
            raise SceneFileError(f'mesh "{name}": {e}') from e
        positions, normals, uvs, triangles = data.positions, data.normals, data.uvs, data.triangles
    else:
        f = src.value
        positions = np.array(f["positions"], np.float32).reshape(-1, 3)
        triangles = np.array(f["triangles"], np.uint32).reshape(-1, 3)
        n = f.get("normals")
        n = n.value if isinstance(n, Some) else n
        u = f.get("uvs")
        u = u.value if isinstance(u, Some) else u
        normals = None if n is None else np.array(n, np.float32).reshape(-1, 3)
        uvs = None if u is None else np.array(u, np.float32).reshape(-1, 2)
    if normals is None:
        normals = _smooth_normals(positions, triangles)
    return MeshData(positions, normals, triangles, uvs)


def _acescg_image(rgba: np.ndarray) -> np.ndarray:
    """environment.rs:87-99: linear Rec.709 texels to ACEScg (glam's
    column-order Mat3 * Vec3 in f32), non-finite texels black, negatives
    clamped, alpha 1."""
    rgb = rgba[..., :3].astype(np.float32)
    finite = np.all(np.isfinite(rgb), axis=-1, keepdims=True)
    rgb = np.where(finite, rgb, np.float32(0.0))
    cx, cy, cz = (np.asarray(c, np.float32) for c in _ACESCG_COLS)
    out = (cx * rgb[..., 0:1] + cy * rgb[..., 1:2]) + cz * rgb[..., 2:3]
    out = np.maximum(out, np.float32(0.0))
    return np.concatenate([out, np.ones_like(out[..., :1])], axis=-1).astype(np.float32)


def lower(desc: Description):
    """The description as (SceneData, settings dict) — lower.rs:95-215."""
    cams = desc.kind("Camera")
    if len(cams) != 1:
        raise SceneFileError(f"a scene needs exactly one camera, this one has {len(cams) or 'none'}")
    envs = desc.kind("Environment")
    if len(envs) > 1:
        raise SceneFileError(f"a scene renders at most one environment, this one has {len(envs)}")
    s = SceneData()
    mesh_index = {}
    for name, mesh in desc.kind("Mesh").items():
        mesh_index[name] = s.add_mesh(_mesh(name, mesh))
    materials = {}
    for name, m in desc.kind("Material").items():
        if m["geometry_normal"] is not None:
            raise SceneFileError(f'material "{name}" has a normal map; this build reads constant slots only')
        fields = {k: _constant(v, f'material "{name}" {k}') for k, v in m.items() if k != "geometry_normal"}
        fields["base_color"] = _vec(fields["base_color"], 3, "base_color")
        for k in ("transmission_color", "coat_color", "fuzz_color", "emission_color"):
            fields[k] = _vec(fields[k], 3, k)
        fields["geometry_thin_walled"] = bool(fields["geometry_thin_walled"])
        materials[name] = OpenPBR(**fields).lower()
    for name, inst in desc.kind("Instance").items():
        s.add_instance(mesh_index[inst["mesh"]], _transform_rows(inst["transform"]), materials[inst["material"]],
                       camera_visible=bool(inst["camera_visible"]))
    for name, light in desc.kind("Light").items():
        lt = light["light"]
        if lt.name == "Distant":
            color = acescg_from_rec709(_vec(lt.value.get("irradiance", (0, 0, 0)), 3, "irradiance"))
            if luminance(color) > 0.0:
                d = np.array(_vec(lt.value["direction"], 3, "direction"), np.float32)
                d = d * (np.float32(1.0) / np.sqrt(np.float32(np.dot(d, d))))
                s.delta_lights.append((1, tuple(float(x) for x in d), color))
        else:
            color = acescg_from_rec709(_vec(lt.value.get("intensity", (0, 0, 0)), 3, "intensity"))
            if luminance(color) > 0.0:
                s.delta_lights.append((2, _vec(lt.value.get("position", (0, 0, 0)), 3, "position"), color))
    (cam,) = cams.values()
    position, look_at = _vec(cam["position"], 3, "position"), _vec(cam["look_at"], 3, "look_at")
    focus = cam["focus_distance"]
    if cam["aperture_radius"] > 0.0 and focus is None:
        focus = float(np.linalg.norm(np.array(position, np.float32) - np.array(look_at, np.float32)))
    s.camera = CameraData(position=position, look_at=look_at, up=_vec(cam["up"], 3, "up"),
                          vfov_degrees=float(cam["vfov_degrees"]), aperture_radius=float(cam["aperture_radius"]),
                          focus_distance=float(focus or 0.0))
    if envs:
        from .output import read_exr

        (env,) = envs.values()
        s.env_texels = _acescg_image(read_exr(env["path"]))
    else:
        s.env_texels = np.array([[[0.0, 0.0, 0.0, 1.0]]], np.float32)  # a black sky
    settings = next(iter(desc.kind("Settings").values()), _default("Settings"))
    return s, dict(settings)


def load_scene(path):
    """A scene file all the way to (SceneData, settings)."""
    desc = Description()
    desc.apply(load(path))
    return lower(desc)
