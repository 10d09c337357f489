"""A pbrt-v4 subset importer (crates/cenote-pbrt): scene text in, the
scene-file change-set out (``scenefile`` ops), plus the import warnings.

The three layers follow the reference crate's:

* tokenizer (tokenize.rs:1-120): whitespace and ``#`` comments separate
  tokens; ``[`` and ``]`` stand alone; ``"..."`` strings may span lines;
* parser (parse.rs:24-580): a closed arity table per directive, ``"type
  name"`` parameter declarations with bracketed or single values (numbers,
  quoted strings or bare ``true``/``false``), ``Include``/``Import`` spliced
  relative to the including file, at most 32 deep;
* mapper (map/mod.rs): pbrt's graphics state — current transform,
  attribute and transform stacks, named materials, the pending area light,
  object recordings — walked front to back, emitting ops. The five traps
  of map/mod.rs:6-40 are kept: RGB light values import verbatim (blackbody
  spectra normalized to luminance 1), roughness remaps to the fourth root
  (square root under ``remaproughness false``), ``fov`` names the shorter
  image axis, the world is conjugated by diag(1, 1, -1) unless the camera
  transform is itself reflective, and orientation flips follow
  ``ReverseOrientation`` XOR a mirroring transform.

Supported content: perspective camera, rgb film resolution, sampler
``pixelsamples``, integrator ``maxdepth``; ``diffuse``, ``coateddiffuse``,
``conductor`` (rgb ``eta``/``k``, ``reflectance`` or the named-metal
table), ``dielectric``/``thindielectric`` materials; ``trianglemesh``,
``plymesh``, ``sphere`` and ``disk`` shapes; ``diffuse`` area lights;
``point``, ``distant`` and constant ``infinite`` lights; object instancing.
Everything else is named in a warning (textures and image skies among
them: this build's kernels read constant material slots only).
"""

from __future__ import annotations

import copy
import math
from pathlib import Path

import numpy as np

from .ron import Some, Variant
from .scenefile import SceneFileError

MAX_INCLUDE_DEPTH = 32
SPHERE_RINGS, SPHERE_SEGMENTS, DISK_SEGMENTS = 32, 64, 64  # map/shape.rs:113-180

# map/mod.rs:58-70: pbrt's left-handed world onto the right-handed one.
FLIP_Z = np.diag(np.array([1.0, 1.0, -1.0, 1.0], np.float32))

# map/color.rs:20-36: normal-incidence linear Rec.709 F0 of pbrt's named metals.
NAMED_METALS = {"Cu": (0.955, 0.638, 0.538), "Au": (1.000, 0.782, 0.344), "Ag": (0.972, 0.960, 0.915),
                "Al": (0.913, 0.921, 0.925)}


class PbrtError(SceneFileError):
    """A pbrt file that does not tokenize, parse or map."""


# --- tokenizer (tokenize.rs) ---------------------------------------------------------

def tokenize(text: str):
    """(token, quoted, line) triples."""
    i, line, n = 0, 1, len(text)
    while True:
        while i < n:
            c = text[i]
            if c == "\n":
                line += 1
                i += 1
            elif c.isspace():
                i += 1
            elif c == "#":
                while i < n and text[i] != "\n":
                    i += 1
            else:
                break
        if i >= n:
            return
        c, start_line = text[i], line
        if c in "[]":
            i += 1
            yield c, False, start_line
        elif c == '"':
            j = i + 1
            while j < n and text[j] != '"':
                if text[j] == "\n":
                    line += 1
                j += 1
            if j >= n:
                raise PbrtError(f"line {start_line}: unterminated string")
            yield text[i + 1:j], True, start_line
            i = j + 1
        else:
            j = i
            while j < n and not text[j].isspace() and text[j] not in '[]"#':
                j += 1
            yield text[i:j], False, start_line
            i = j


# --- parser (parse.rs) -----------------------------------------------------------------

ARITY = {}
for _k in ("WorldBegin", "WorldEnd", "AttributeBegin", "AttributeEnd", "TransformBegin", "TransformEnd",
           "ObjectEnd", "ReverseOrientation", "Identity"):
    ARITY[_k] = ("bare", 0)
for _k, _n in (("Translate", 3), ("Scale", 3), ("Rotate", 4), ("LookAt", 9), ("Transform", 16),
               ("ConcatTransform", 16), ("TransformTimes", 2)):
    ARITY[_k] = ("numbers", _n)
for _k in ("Include", "Import", "ObjectBegin", "ObjectInstance", "NamedMaterial", "CoordinateSystem",
           "CoordSysTransform", "ColorSpace", "ActiveTransform"):
    ARITY[_k] = ("names", 1)
ARITY["MediumInterface"] = ("names", 2)
ARITY["Texture"] = ("names_params", 3)
for _k in ("Shape", "Material", "MakeNamedMaterial", "MakeNamedMedium", "LightSource", "AreaLightSource",
           "Camera", "Film", "Sampler", "Integrator", "PixelFilter", "Accelerator", "Option", "Attribute"):
    ARITY[_k] = ("names_params", 1)


class Param:
    def __init__(self, ty, name, value, location):
        self.ty, self.name, self.value, self.location = ty, name, value, location
        self.consumed = False

    def floats(self):
        if not (isinstance(self.value, list) and all(isinstance(v, float) for v in self.value)):
            raise PbrtError(f'{self.location}: parameter "{self.name}" needs numeric values')
        return self.value

    def scalar(self):
        v = self.floats()
        if len(v) != 1:
            raise PbrtError(f'{self.location}: parameter "{self.name}" needs one value')
        return float(np.float32(v[0]))

    def rgb(self):
        v = self.floats()
        if len(v) != 3:
            raise PbrtError(f'{self.location}: parameter "{self.name}" needs three values')
        return tuple(float(np.float32(x)) for x in v)

    def string(self):
        if not (isinstance(self.value, list) and len(self.value) == 1 and isinstance(self.value[0], str)):
            raise PbrtError(f'{self.location}: parameter "{self.name}" needs one string value')
        return self.value[0]


class Params:
    def __init__(self):
        self.items = []

    def take(self, name, types):
        for p in self.items:
            if p.name == name:
                p.consumed = True
                if p.ty not in types:
                    raise PbrtError(f'{p.location}: parameter "{name}" is declared "{p.ty}", expected one of {types}')
                return p
        return None

    def float(self, name):
        p = self.take(name, ("float", "integer"))
        return None if p is None else p.scalar()

    def integer(self, name):
        p = self.take(name, ("integer",))
        if p is None:
            return None
        v = p.floats()
        if len(v) != 1:
            raise PbrtError(f'{p.location}: parameter "{name}" needs one value')
        return int(v[0])

    def string(self, name):
        p = self.take(name, ("string", "texture"))
        return None if p is None else p.string()

    def boolean(self, name):
        p = self.take(name, ("bool",))
        if p is None:
            return None
        v = p.value
        if isinstance(v, list) and len(v) == 1 and (isinstance(v[0], bool) or v[0] in ("true", "false")):
            return v[0] if isinstance(v[0], bool) else v[0] == "true"
        raise PbrtError(f'{p.location}: parameter "{name}" needs one boolean value')

    def unused(self, context):
        return [f'{p.location}: {context}: parameter "{p.name}" ({p.ty}) is not supported — ignored'
                for p in self.items if not p.consumed]


class Directive:
    def __init__(self, keyword, location):
        self.keyword, self.location = keyword, location
        self.names, self.numbers, self.params = [], [], Params()


class Parser:
    """Directives of a file and everything it includes, in order."""

    def __init__(self, path):
        self.root = Path(path)
        self.stack = [self._open(self.root)]

    @staticmethod
    def _open(path):
        try:
            text = Path(path).read_text()
        except OSError as e:
            raise PbrtError(f"can't read {path}: {e}")
        return {"path": Path(path), "tokens": list(tokenize(text)), "at": 0}

    def resolve(self, name):
        p = Path(name)
        return p if p.is_absolute() else self.stack[-1]["path"].parent / p

    def _peek(self):
        while self.stack:
            f = self.stack[-1]
            if f["at"] < len(f["tokens"]):
                return f["tokens"][f["at"]]
            if len(self.stack) == 1:
                return None
            self.stack.pop()
        return None

    def _next(self):
        tok = self._peek()
        if tok is not None:
            self.stack[-1]["at"] += 1
        return tok

    def _loc(self, tok):
        return f"{self.stack[-1]['path'].name}:{tok[2]}"

    def directives(self):
        while True:
            tok = self._next()
            if tok is None:
                return
            text, quoted, _ = tok
            loc = self._loc(tok)
            if quoted or text in "[]":
                raise PbrtError(f"{loc}: expected a directive keyword")
            if text not in ARITY:
                raise PbrtError(f"{loc}: unknown directive {text}")
            kind, count = ARITY[text]
            d = Directive(text, loc)
            if kind == "numbers":
                self._numbers(d, count)
            elif kind in ("names", "names_params"):
                for _ in range(count):
                    t = self._next()
                    if t is None:
                        raise PbrtError(f"{loc}: {text} ends at end of file, expected a name")
                    if not t[1] and t[0] in "[]":
                        raise PbrtError(f"{self._loc(t)}: expected a name")
                    d.names.append(t[0])
                if kind == "names_params":
                    self._params(d)
            if text in ("Include", "Import"):
                if len(self.stack) >= MAX_INCLUDE_DEPTH:
                    raise PbrtError(f"{loc}: includes nest deeper than {MAX_INCLUDE_DEPTH} — a cycle?")
                self.stack.append(self._open(self.resolve(d.names[0])))
                continue
            yield d

    def _numbers(self, d, count):
        t = self._peek()
        bracketed = t is not None and not t[1] and t[0] == "["
        if bracketed:
            self._next()
        for _ in range(count):
            t = self._next()
            if t is None:
                raise PbrtError(f"{d.location}: {d.keyword} ends at end of file, expected more numbers")
            try:
                d.numbers.append(float(t[0]))
            except ValueError:
                raise PbrtError(f"{self._loc(t)}: {d.keyword} expects numbers")
        if bracketed:
            t = self._next()
            if t is None or t[1] or t[0] != "]":
                raise PbrtError(f'{d.location}: expected "]"')

    def _params(self, d):
        while True:
            t = self._peek()
            if t is None or not t[1] or not any(ch.isspace() for ch in t[0]):
                return
            self._next()
            words = t[0].split()
            if len(words) != 2:
                raise PbrtError(f'{self._loc(t)}: expected a "type name" declaration')
            d.params.items.append(Param(words[0], words[1], self._value(d, words[1]), self._loc(t)))

    def _value(self, d, name):
        first = self._next()
        if first is None:
            raise PbrtError(f"{d.location}: {d.keyword} ends at end of file, expected a parameter value")
        toks = []
        if not first[1] and first[0] == "[":
            while True:
                t = self._next()
                if t is None:
                    raise PbrtError(f'{d.location}: {d.keyword} ends at end of file, expected "]"')
                if not t[1] and t[0] == "]":
                    break
                if not t[1] and t[0] == "[":
                    raise PbrtError(f"{self._loc(t)}: nested brackets")
                toks.append(t)
        else:
            toks.append(first)
        values, kind = [], None
        for text, quoted, line in toks:
            k = "s" if quoted else ("b" if text in ("true", "false") else "n")
            if kind is None:
                kind = k
            if k != kind:
                raise PbrtError(f"{d.location}: mixed value types in one parameter")
            if k == "s":
                values.append(text)
            elif k == "b":
                values.append(text == "true")
            else:
                try:
                    values.append(float(text))
                except ValueError:
                    raise PbrtError(f'{d.location}: parameter "{name}" expects numbers')
        return values


# --- geometry (map/shape.rs) -----------------------------------------------------------

def _f32(x):
    return np.float32(x)


def sphere_mesh(radius):
    """shape.rs:104-146: 32 rings x 64 segments, poles on z, analytic normals."""
    positions, normals, uvs, triangles = [], [], [], []
    pi, tau = _f32(math.pi), _f32(2.0 * math.pi)
    r = _f32(radius)
    for ring in range(SPHERE_RINGS + 1):
        v = _f32(ring) / _f32(SPHERE_RINGS)
        theta = v * pi
        st, ct = np.sin(theta), np.cos(theta)
        for segment in range(SPHERE_SEGMENTS + 1):
            u = _f32(segment) / _f32(SPHERE_SEGMENTS)
            phi = u * tau
            n = (float(st * np.cos(phi)), float(st * np.sin(phi)), float(ct))
            positions.append(tuple(float(_f32(c) * r) for c in n))
            normals.append(n)
            uvs.append((float(u), float(v)))
    row = SPHERE_SEGMENTS + 1
    for ring in range(SPHERE_RINGS):
        for segment in range(SPHERE_SEGMENTS):
            a = ring * row + segment
            b = a + row
            if ring != 0:
                triangles.append((a, b, a + 1))
            if ring != SPHERE_RINGS - 1:
                triangles.append((a + 1, b, b + 1))
    return Variant("Inline", {"positions": positions, "normals": normals, "uvs": uvs, "triangles": triangles})


def disk_mesh(radius, height):
    """shape.rs:150-179: a center vertex and a 64-segment rim facing +z."""
    positions, uvs = [(0.0, 0.0, float(height))], [(0.0, 1.0)]
    tau = _f32(2.0 * math.pi)
    for segment in range(DISK_SEGMENTS + 1):
        u = _f32(segment) / _f32(DISK_SEGMENTS)
        phi = u * tau
        positions.append((float(_f32(radius) * np.cos(phi)), float(_f32(radius) * np.sin(phi)), float(height)))
        uvs.append((float(u), 0.0))
    triangles = [(0, s + 1, s + 2) for s in range(DISK_SEGMENTS)]
    return Variant("Inline", {"positions": positions, "normals": [(0.0, 0.0, 1.0)] * (DISK_SEGMENTS + 2),
                              "uvs": uvs, "triangles": triangles})


def trianglemesh(d: Directive, flip: bool):
    """shape.rs:15-101: the streams verbatim; `flip` negates normals and
    reverses winding."""
    def triples(name, types):
        p = d.params.take(name, types)
        if p is None:
            return None
        f = p.floats()
        if len(f) % 3:
            raise PbrtError(f'{p.location}: "{name}" needs whole (x, y, z) triples')
        return [tuple(float(_f32(x)) for x in f[i:i + 3]) for i in range(0, len(f), 3)]

    positions = triples("P", ("point3", "point"))
    if positions is None:
        raise PbrtError(f'{d.location}: trianglemesh has no "point3 P"')
    normals = triples("N", ("normal", "normal3"))
    if flip and normals is not None:
        normals = [tuple(-c for c in n) for n in normals]
    uvs = None
    p = d.params.take("uv", ("point2", "float", "vector2"))
    if p is not None:
        f = p.floats()
        if len(f) % 2:
            raise PbrtError(f'{p.location}: "uv" needs whole (u, v) pairs')
        uvs = [tuple(float(_f32(x)) for x in f[i:i + 2]) for i in range(0, len(f), 2)]
    p = d.params.take("indices", ("integer",))
    if p is not None:
        f = p.floats()
        if len(f) % 3:
            raise PbrtError(f'{p.location}: "indices" needs whole triangles')
        for v in f:
            if v < 0 or v > 0xFFFFFFFF:
                raise PbrtError(f"{p.location}: index {v} is out of range")
        triangles = [tuple(int(v) for v in f[i:i + 3]) for i in range(0, len(f), 3)]
    elif len(positions) == 3:
        triangles = [(0, 1, 2)]
    else:
        raise PbrtError(f'{d.location}: trianglemesh has no "integer indices"')
    if flip:
        triangles = [(a, c, b) for a, b, c in triangles]
    fields = {"positions": positions, "triangles": triangles}
    if normals is not None:
        fields["normals"] = normals
    if uvs is not None:
        fields["uvs"] = uvs
    return Variant("Inline", fields)


# --- color (map/color.rs) ---------------------------------------------------------------

def conductor_f0(eta, k):
    """color.rs:8-16: normal-incidence reflectance of a complex IOR, per channel."""
    out = []
    for n, kk in zip(eta, k):
        n, kk = _f32(n), _f32(kk)
        a = (n - _f32(1.0)) * (n - _f32(1.0)) + kk * kk
        b = (n + _f32(1.0)) * (n + _f32(1.0)) + kk * kk
        out.append(float(a / b))
    return tuple(out)


def named_metal_f0(spectrum):
    if not spectrum.startswith("metal-"):
        return None
    metal = spectrum[len("metal-"):]
    for suffix in ("-eta", "-k"):
        if metal.endswith(suffix):
            return NAMED_METALS.get(metal[:-len(suffix)])
    return None


def blackbody_rec709(kelvin):
    """color.rs:38-62: Krystek's Planckian locus, to linear Rec.709 at luminance 1."""
    t = min(max(float(kelvin), 1000.0), 15000.0)
    u = (0.860117757 + 1.54118254e-4 * t + 1.28641212e-7 * t * t) / (1.0 + 8.42420235e-4 * t + 7.08145163e-7 * t * t)
    v = (0.317398726 + 4.22806245e-5 * t + 4.20481691e-8 * t * t) / (1.0 - 2.89741816e-5 * t + 1.61456053e-7 * t * t)
    x = 3.0 * u / (2.0 * u - 8.0 * v + 4.0)
    y = 2.0 * v / (2.0 * u - 8.0 * v + 4.0)
    xyz = (x / y, 1.0, (1.0 - x - y) / y)
    rgb = (3.2404542 * xyz[0] - 1.5371385 * xyz[1] - 0.4985314 * xyz[2],
           -0.9692660 * xyz[0] + 1.8760108 * xyz[1] + 0.0415560 * xyz[2],
           0.0556434 * xyz[0] - 0.2040259 * xyz[1] + 1.0572252 * xyz[2])
    rgb = [_f32(max(c, 0.0)) for c in rgb]
    lum = _f32(0.2126) * rgb[0] + _f32(0.7152) * rgb[1] + _f32(0.0722) * rgb[2]
    return tuple(float(c / lum) for c in rgb)


# --- mapper (map/mod.rs) -----------------------------------------------------------------

def _mat4(numbers):
    """Sixteen column-major numbers (pbrt's order) as a 4x4 matrix."""
    return np.array(numbers, np.float64).astype(np.float32).reshape(4, 4).T


def _translate(v):
    m = np.eye(4, dtype=np.float32)
    m[:3, 3] = v
    return m


def _scale(v):
    return np.diag(np.array([v[0], v[1], v[2], 1.0], np.float32))


def _axis_angle(axis, angle):
    """Rotation by `angle` radians about the unit `axis` (column vectors)."""
    x, y, z = (float(a) for a in axis)
    s, c = math.sin(angle), math.cos(angle)
    t = 1.0 - c
    m = np.eye(4, dtype=np.float64)
    m[:3, :3] = [[t * x * x + c, t * x * y - s * z, t * x * z + s * y],
                 [t * x * y + s * z, t * y * y + c, t * y * z - s * x],
                 [t * x * z - s * y, t * y * z + s * x, t * z * z + c]]
    return m.astype(np.float32)


def _inverse(m):
    return np.linalg.inv(m.astype(np.float64)).astype(np.float32)


def _mirrors(m) -> bool:
    return float(np.linalg.det(m[:3, :3].astype(np.float64))) < 0.0


def _rows(m):
    return tuple(tuple(float(x) for x in m[r]) for r in range(3))


def _vec3(numbers, i):
    return np.array(numbers[i:i + 3], np.float64).astype(np.float32)


class _State:
    def __init__(self):
        self.ctm = np.eye(4, dtype=np.float32)
        self.reverse_orientation = False
        self.material = None  # (name, fields)
        self.area_light = None  # dict(color, luminance, two_sided, forks)


def _op(kind, name, fields):
    value = {"name": name}
    value.update({k: Some(v) for k, v in fields.items()})
    return Variant(kind, value)


def _constant(v):
    return Variant("Constant", v)


class Mapper:
    def __init__(self, parser: Parser, generated: Path | None, stem: str):
        self.parser, self.generated, self.stem = parser, generated, stem
        self.ops, self.warnings, self._seen = [], [], set()
        self.options = {"resolution": [1280, 720], "spp": 16, "max_bounces": 5,
                        "camera": (np.eye(4, dtype=np.float32), 90.0, 0.0, 1e6)}
        self.in_world = False
        self.state = _State()
        self.attribute_stack, self.transform_stack = [], []
        self.named_materials, self.named_textures = {}, {}
        self.objects, self.active_object = {}, None
        self.counters = {}
        self.conjugation = FLIP_Z
        self.environment_emitted = False
        self.two_sided_lights = 0
        self.default_material_emitted = False

    def warn(self, text):
        if text not in self._seen:
            self._seen.add(text)
            self.warnings.append(text)

    def fresh(self, prefix):
        n = self.counters.get(prefix, 0)
        self.counters[prefix] = n + 1
        return f"{prefix}-{n}"

    def block(self, d, world):
        if world != self.in_world:
            place = "after" if world else "before"
            raise PbrtError(f"{d.location}: {d.keyword} must appear {place} WorldBegin")

    def run(self):
        for d in self.parser.directives():
            self.dispatch(d)
        if not self.in_world:
            raise PbrtError("the scene never reaches WorldBegin — not a pbrt scene?")
        if self.two_sided_lights:
            self.warn(f"{self.two_sided_lights} area light(s) are two-sided in pbrt; cenote emitters are "
                      "one-sided, so their back faces stay dark")
        return self.ops, self.warnings

    def dispatch(self, d: Directive):
        k, st = d.keyword, self.state
        if k == "Identity":
            st.ctm = np.eye(4, dtype=np.float32)
        elif k == "Translate":
            st.ctm = st.ctm @ _translate(_vec3(d.numbers, 0))
        elif k == "Scale":
            st.ctm = st.ctm @ _scale(_vec3(d.numbers, 0))
        elif k == "Rotate":
            axis = _vec3(d.numbers, 1).astype(np.float64)
            norm = float(np.linalg.norm(axis))
            if not norm > 0.0:
                raise PbrtError(f"{d.location}: Rotate about a zero axis")
            st.ctm = st.ctm @ _axis_angle(axis / norm, math.radians(float(np.float32(d.numbers[0]))))
        elif k == "LookAt":
            self.look_at(d)
        elif k == "Transform":
            st.ctm = _mat4(d.numbers)
        elif k == "ConcatTransform":
            st.ctm = st.ctm @ _mat4(d.numbers)
        elif k == "TransformBegin":
            self.transform_stack.append(st.ctm.copy())
        elif k == "TransformEnd":
            if not self.transform_stack:
                raise PbrtError(f"{d.location}: TransformEnd without TransformBegin")
            st.ctm = self.transform_stack.pop()
        elif k in ("CoordinateSystem", "CoordSysTransform"):
            self.warn(f"{d.location}: named coordinate systems are not supported — {k} ignored")
        elif k in ("ActiveTransform", "TransformTimes"):
            self.warn(f"{d.location}: animated transforms are not supported — the static transform is used")
        elif k == "WorldBegin":
            self.world_begin(d)
        elif k in ("WorldEnd", "PixelFilter", "Accelerator", "Option"):
            pass
        elif k == "AttributeBegin":
            self.attribute_stack.append(copy.deepcopy(self.state))
        elif k == "AttributeEnd":
            if not self.attribute_stack:
                raise PbrtError(f"{d.location}: AttributeEnd without AttributeBegin")
            self.state = self.attribute_stack.pop()
        elif k == "ReverseOrientation":
            self.block(d, True)
            st.reverse_orientation = not st.reverse_orientation
        elif k == "Camera":
            self.camera(d)
        elif k == "Film":
            self.block(d, False)
            x, y = d.params.integer("xresolution"), d.params.integer("yresolution")
            if x is not None:
                self.options["resolution"][0] = max(x, 1)
            if y is not None:
                self.options["resolution"][1] = max(y, 1)
            d.params.string("filename")
            self.unused(d, "Film")
        elif k == "Sampler":
            self.block(d, False)
            spp = d.params.integer("pixelsamples")
            if spp is not None:
                self.options["spp"] = max(spp, 1)
            self.unused(d, "Sampler")
        elif k == "Integrator":
            self.block(d, False)
            depth = d.params.integer("maxdepth")
            if depth is not None:
                self.options["max_bounces"] = max(depth, 1)
            self.unused(d, "Integrator")
        elif k == "ColorSpace":
            if d.names[0] != "srgb":
                self.warn(f'{d.location}: color space "{d.names[0]}" is not supported — colors are read as '
                          "sRGB/Rec.709")
        elif k == "Material":
            self.block(d, True)
            ty = d.names[0]
            fields = self.lower_material(ty, d)
            name = self.fresh(ty)
            self.unused(d, f'material "{ty}"')
            self.ops.append(_op("Material", name, fields))
            st.material = (name, fields)
        elif k == "MakeNamedMaterial":
            self.block(d, True)
            name = d.names[0]
            if name in self.named_materials:
                raise PbrtError(f'{d.location}: named material "{name}" is defined twice')
            ty = d.params.string("type")
            if ty is None:
                raise PbrtError(f'{d.location}: MakeNamedMaterial "{name}" has no "string type"')
            fields = self.lower_material(ty, d)
            self.unused(d, f'material "{name}"')
            self.ops.append(_op("Material", name, fields))
            self.named_materials[name] = fields
        elif k == "NamedMaterial":
            self.block(d, True)
            name = d.names[0]
            if name not in self.named_materials:
                raise PbrtError(f'{d.location}: NamedMaterial "{name}" was never made')
            st.material = (name, self.named_materials[name])
        elif k == "Texture":
            self.warn(f'{d.location}: texture "{d.names[0]}" ({d.names[2]}) is not supported — this build '
                      "reads constant material slots only")
        elif k == "Shape":
            self.shape(d)
        elif k == "AreaLightSource":
            self.area_light(d)
        elif k == "LightSource":
            self.light(d)
        elif k == "ObjectBegin":
            self.block(d, True)
            name = d.names[0]
            if self.active_object is not None:
                raise PbrtError(f"{d.location}: ObjectBegin inside an object definition")
            if name in self.objects:
                raise PbrtError(f'{d.location}: object "{name}" is defined twice')
            self.attribute_stack.append(copy.deepcopy(self.state))
            self.active_object = (name, [])
        elif k == "ObjectEnd":
            if self.active_object is None:
                raise PbrtError(f"{d.location}: ObjectEnd without ObjectBegin")
            name, shapes = self.active_object
            self.active_object = None
            if not self.attribute_stack:
                raise PbrtError(f"{d.location}: ObjectEnd with a mismatched attribute stack")
            self.state = self.attribute_stack.pop()
            self.objects[name] = shapes
        elif k == "ObjectInstance":
            self.object_instance(d)
        elif k in ("MakeNamedMedium", "MediumInterface", "Attribute"):
            self.warn(f"{d.location}: {k} is not supported — ignored")
        else:
            raise PbrtError(f"{d.location}: directive {k} parses but has no mapping")

    def unused(self, d, context):
        for w in d.params.unused(context):
            self.warn(w)

    # camera and options (map/mod.rs:418-541)

    def look_at(self, d):
        eye, look, up = (_vec3(d.numbers, i).astype(np.float64) for i in (0, 3, 6))
        direction = look - eye
        if not np.linalg.norm(direction) > 0.0:
            raise PbrtError(f"{d.location}: LookAt eye and target coincide")
        direction /= np.linalg.norm(direction)
        right = np.cross(up / np.linalg.norm(up), direction)
        if not np.linalg.norm(right) > 0.0:
            raise PbrtError(f"{d.location}: LookAt up is parallel to the view direction")
        right /= np.linalg.norm(right)
        new_up = np.cross(direction, right)
        world_from_camera = np.eye(4)
        world_from_camera[:3, 0], world_from_camera[:3, 1], world_from_camera[:3, 2] = right, new_up, direction
        world_from_camera[:3, 3] = eye
        self.state.ctm = self.state.ctm @ np.linalg.inv(world_from_camera).astype(np.float32)

    def camera(self, d):
        self.block(d, False)
        if d.names[0] != "perspective":
            self.warn(f'{d.location}: camera "{d.names[0]}" is not supported — imported as perspective')
        world_from_camera = _inverse(self.state.ctm)
        if not np.all(np.isfinite(world_from_camera)):
            raise PbrtError(f"{d.location}: the camera transform is not invertible")
        fov = d.params.float("fov")
        lens = d.params.float("lensradius")
        focal = d.params.float("focaldistance")
        self.options["camera"] = (world_from_camera, 90.0 if fov is None else fov, 0.0 if lens is None else lens,
                                  1e6 if focal is None else focal)
        self.unused(d, "Camera")

    def world_begin(self, d):
        self.block(d, False)
        self.in_world = True
        self.state.ctm = np.eye(4, dtype=np.float32)
        width, height = self.options["resolution"]
        self.ops.append(_op("Settings", "main", {"resolution": (width, height), "spp": self.options["spp"],
                                                 "max_bounces": self.options["max_bounces"]}))
        world_from_camera, fov, lens, focal = self.options["camera"]
        self.conjugation = np.eye(4, dtype=np.float32) if _mirrors(world_from_camera) else FLIP_Z
        if width >= height:
            vfov = fov
        else:
            half = math.tan(math.radians(fov) / 2.0) * height / width
            vfov = float(np.float32(2.0 * math.degrees(math.atan(half))))
        to_world = (self.conjugation @ world_from_camera).astype(np.float64)
        position = to_world[:3, 3]
        forward = to_world[:3, 2] / np.linalg.norm(to_world[:3, 2])
        up = to_world[:3, 1]
        f = lambda v: tuple(float(np.float32(x)) for x in v)  # noqa: E731
        fields = {"position": f(position), "look_at": f(position + forward), "up": f(up), "vfov_degrees": vfov,
                  "aperture_radius": lens}
        if lens > 0.0:
            fields["focus_distance"] = Some(focal)
        self.ops.append(_op("Camera", "main", fields))

    # shapes and materials (map/mod.rs:565-710, 965-1180)

    def shape(self, d):
        self.block(d, True)
        ty = d.names[0]
        st = self.state
        flip = st.reverse_orientation ^ _mirrors(st.ctm)
        if ty == "trianglemesh":
            source, prefix = trianglemesh(d, flip), "trianglemesh"
        elif ty == "plymesh":
            file = d.params.string("filename")
            if file is None:
                raise PbrtError(f'{d.location}: plymesh has no "string filename"')
            path = self.parser.resolve(file)
            if flip:
                self.warn(f"{d.location}: ReverseOrientation on a plymesh is ignored (cenote shades and emits "
                          "two-sided)")
            source, prefix = Variant("Ply", {"path": str(path)}), path.stem or "plymesh"
        elif ty == "sphere":
            radius = d.params.float("radius")
            for clip in ("zmin", "zmax", "phimax"):
                if d.params.float(clip) is not None:
                    self.warn(f'{d.location}: partial spheres ("{clip}") are not supported — a full sphere is '
                              "tessellated")
            source, prefix = sphere_mesh(1.0 if radius is None else radius), "sphere"
        elif ty == "disk":
            radius, height = d.params.float("radius"), d.params.float("height")
            for clip in ("innerradius", "phimax"):
                if d.params.float(clip) is not None:
                    self.warn(f'{d.location}: partial disks ("{clip}") are not supported — a full disk is '
                              "tessellated")
            source, prefix = disk_mesh(1.0 if radius is None else radius, 0.0 if height is None else height), "disk"
        else:
            self.warn(f'{d.location}: shape "{ty}" is not supported — skipped')
            return
        mesh = self.fresh(prefix)
        self.ops.append(_op("Mesh", mesh, {"source": source}))
        material = self.shape_material(d.location)
        self.unused(d, f'shape "{ty}"')
        if self.active_object is not None:
            self.active_object[1].append((mesh, material, st.ctm.copy()))
        else:
            self.ops.append(_op("Instance", mesh, {"mesh": mesh, "material": material,
                                                   "transform": Variant("Matrix", _rows(self.conjugation @ st.ctm))}))

    def shape_material(self, location):
        st = self.state
        if st.material is not None:
            base_name, base = st.material
        else:
            base = {"base_color": _constant((0.5, 0.5, 0.5)), "specular_weight": 0.0}
            if not self.default_material_emitted:
                self.default_material_emitted = True
                self.ops.append(_op("Material", "pbrt-default", base))
            base_name = "pbrt-default"
        area = st.area_light
        if area is None:
            return base_name
        if self.active_object is not None:
            self.warn(f"{location}: area lights are not supported with object instancing — the shape imports "
                      "unlit")
            return base_name
        if base_name in area["forks"]:
            return area["forks"][base_name]
        fork = self.fresh(f"{base_name}-glow")
        fields = dict(base)
        fields["emission_color"] = _constant(area["color"])
        fields["emission_luminance"] = area["luminance"]
        self.ops.append(_op("Material", fork, fields))
        if area["two_sided"]:
            self.two_sided_lights += 1
        area["forks"][base_name] = fork
        return fork

    def lower_material(self, ty, d):
        fields = {}
        if ty == "diffuse":
            fields["base_color"] = self.color_slot(d, "reflectance") or _constant((0.5, 0.5, 0.5))
            fields["specular_weight"] = 0.0
        elif ty == "coateddiffuse":
            fields["base_color"] = self.color_slot(d, "reflectance") or _constant((0.5, 0.5, 0.5))
            fields["specular_weight"] = 0.0
            fields["coat_weight"] = 1.0
            fields["coat_ior"] = self.dielectric_eta(d)
            fields["coat_roughness"] = self.roughness(d)
        elif ty == "conductor":
            fields["base_metalness"] = _constant(1.0)
            fields["base_color"] = self.conductor_color(d)
            fields["specular_roughness"] = _constant(self.roughness(d))
        elif ty in ("dielectric", "thindielectric"):
            fields["transmission_weight"] = 1.0
            fields["specular_ior"] = self.dielectric_eta(d)
            if ty == "thindielectric":
                fields["geometry_thin_walled"] = True
            else:
                fields["specular_roughness"] = _constant(self.roughness(d))
        else:
            self.warn(f'{d.location}: material "{ty}" is not supported — OpenPBR defaults used')
        if d.params.string("normalmap") is not None:
            self.warn(f"{d.location}: normal maps are not supported on this path — ignored")
        return fields

    def color_slot(self, d, name):
        p = d.params.take(name, ("rgb", "color", "float", "spectrum", "blackbody", "texture"))
        if p is None:
            return None
        if p.ty in ("rgb", "color"):
            return _constant(p.rgb())
        if p.ty == "float":
            v = p.scalar()
            return _constant((v, v, v))
        self.warn(f'{p.location}: "{p.ty} {name}" is not supported here — mid-gray used')
        return _constant((0.5, 0.5, 0.5))

    def conductor_color(self, d):
        reflectance = self.color_slot(d, "reflectance")
        types = ("rgb", "color", "spectrum", "float", "texture")
        if reflectance is not None:
            if d.params.take("eta", types) is not None or d.params.take("k", types) is not None:
                raise PbrtError(f'{d.location}: conductor has both "reflectance" and "eta"/"k"')
            return reflectance
        eta, k = d.params.take("eta", types), d.params.take("k", types)
        if eta is not None and k is not None and eta.ty in ("rgb", "color") and k.ty in ("rgb", "color"):
            return _constant(conductor_f0(eta.rgb(), k.rgb()))
        p = eta or k
        if p is not None:
            if p.ty == "spectrum" and isinstance(p.value, list) and len(p.value) == 1 and isinstance(p.value[0], str):
                f0 = named_metal_f0(p.value[0])
                if f0 is not None:
                    return _constant(f0)
                self.warn(f'{p.location}: conductor spectrum "{p.value[0]}" is not in the metal table — copper used')
            else:
                self.warn(f'{p.location}: conductor "eta"/"k" of type "{p.ty}" is not supported — copper used')
        return _constant(NAMED_METALS["Cu"])

    def dielectric_eta(self, d):
        p = d.params.take("eta", ("float", "spectrum", "rgb", "color"))
        if p is None:
            return 1.5
        if p.ty == "float" and len(p.floats()) == 1:
            return p.scalar()
        self.warn(f"{p.location}: a spectral IOR (dispersion) is not supported — 1.5 used")
        return 1.5

    def roughness(self, d):
        """Trap 2 (map/mod.rs:1160-1210)."""
        remap = d.params.boolean("remaproughness")
        remap = True if remap is None else remap
        value = d.params.float("roughness")
        aniso = [v for v in (d.params.float("uroughness"), d.params.float("vroughness")) if v is not None]
        if aniso:
            self.warn(f"{d.location}: anisotropic roughness is not supported — the axes are averaged")
            total = np.float32(0.0)
            for v in aniso:
                total = np.float32(total + np.float32(v))
            value = float(total / np.float32(len(aniso)))
        value = 0.0 if value is None else value
        exponent = np.float32(0.25 if remap else 0.5)
        return float(np.power(np.float32(max(value, 0.0)), exponent))

    # lights (map/mod.rs:669-906)

    def area_light(self, d):
        self.block(d, True)
        if d.names[0] != "diffuse":
            self.warn(f'{d.location}: area light "{d.names[0]}" is not supported — shapes import unlit')
            self.state.area_light = None
            return
        luminance = d.params.float("scale")
        luminance = 1.0 if luminance is None else luminance
        if d.params.float("power") is not None:
            self.warn(f'{d.location}: area light "power" needs the shape\'s surface area — the plain photometric '
                      "scale is used instead")
        if d.params.string("filename") is not None:
            self.warn(f"{d.location}: image area lights are not supported on this path — imported white")
            color = (1.0, 1.0, 1.0)
        else:
            color = self.light_color(d, "L")
        two_sided = d.params.boolean("twosided") or False
        self.unused(d, "area light")
        self.state.area_light = {"color": color, "luminance": luminance, "two_sided": two_sided, "forks": {}}

    def light_color(self, d, name):
        p = d.params.take(name, ("rgb", "color", "blackbody", "spectrum", "float"))
        if p is None:
            return (1.0, 1.0, 1.0)
        if p.ty in ("rgb", "color"):
            return p.rgb()
        if p.ty == "blackbody":
            return blackbody_rec709(p.scalar())
        if p.ty == "float":
            v = p.scalar()
            return (v, v, v)
        self.warn(f'{p.location}: spectral light data is not supported — "{name}" imports as white at its '
                  "photometric scale")
        return (1.0, 1.0, 1.0)

    def light(self, d):
        self.block(d, True)
        kind = d.names[0]
        world = self.conjugation @ self.state.ctm

        def point(name, default):
            p = d.params.take(name, ("point3", "point"))
            return np.array(default, np.float32) if p is None else _vec3(p.floats(), 0)

        if kind == "point":
            factor = d.params.float("scale")
            factor = np.float32(1.0 if factor is None else factor)
            power = d.params.float("power")
            if power is not None:
                factor = factor * np.float32(power / (4.0 * math.pi))
            color = self.light_color(d, "I")
            frm = point("from", (0.0, 0.0, 0.0))
            position = (world @ np.append(frm, np.float32(1.0)))[:3]
            self.ops.append(_op("Light", self.fresh("point"), {"light": Variant("Point", {
                "position": tuple(float(x) for x in position),
                "intensity": tuple(float(np.float32(c) * factor) for c in color)})}))
        elif kind == "distant":
            factor = d.params.float("scale")
            factor = np.float32(1.0 if factor is None else factor)
            illuminance = d.params.float("illuminance")
            if illuminance is not None:
                factor = factor * np.float32(illuminance)
            color = self.light_color(d, "L")
            frm, to = point("from", (0.0, 0.0, 0.0)), point("to", (0.0, 0.0, 1.0))
            a = (world @ np.append(frm, np.float32(1.0)))[:3]
            b = (world @ np.append(to, np.float32(1.0)))[:3]
            direction = (b - a).astype(np.float64)
            if not np.linalg.norm(direction) > 0.0:
                raise PbrtError(f'{d.location}: distant light "from" and "to" coincide')
            direction /= np.linalg.norm(direction)
            self.ops.append(_op("Light", self.fresh("distant"), {"light": Variant("Distant", {
                "direction": tuple(float(np.float32(x)) for x in direction),
                "irradiance": tuple(float(np.float32(c) * factor) for c in color)})}))
        elif kind == "infinite":
            self.infinite_light(d)
        else:
            self.warn(f'{d.location}: light "{kind}" is not supported — skipped')
            return
        self.unused(d, f'light "{kind}"')

    def infinite_light(self, d):
        if self.environment_emitted:
            self.warn(f"{d.location}: cenote renders one environment — this infinite light is skipped")
            return
        if d.params.take("portal", ("point3", "point")) is not None:
            self.warn(f"{d.location}: light portals are not supported — the portal is ignored")
        scale = d.params.float("scale")
        factor = np.float32(1.0 if scale is None else scale)
        if d.params.string("filename") is not None:
            self.warn(f"{d.location}: image skies are not supported on this path — the infinite light is skipped")
            return
        illuminance = d.params.float("illuminance")
        if illuminance is not None:
            factor = factor * np.float32(illuminance / math.pi)
        color = self.light_color(d, "L")
        if self.generated is None:
            raise PbrtError(f"{d.location}: a constant sky needs a directory for its generated EXR")
        from .output import write_exr

        self.generated.mkdir(parents=True, exist_ok=True)
        out = self.generated / f"{self.stem}-sky.exr"
        rgba = np.ones((8, 16, 4), np.float32)
        rgba[..., :3] = [float(np.float32(c) * factor) for c in color]
        write_exr(out, rgba)
        self.ops.append(_op("Environment", "sky", {"path": str(out)}))
        self.environment_emitted = True

    def object_instance(self, d):
        self.block(d, True)
        if self.active_object is not None:
            raise PbrtError(f"{d.location}: ObjectInstance inside an object definition")
        name = d.names[0]
        if name not in self.objects:
            raise PbrtError(f'{d.location}: ObjectInstance of "{name}", which was never defined')
        for mesh, material, ctm in self.objects[name]:
            self.ops.append(_op("Instance", self.fresh(name), {
                "mesh": mesh, "material": material,
                "transform": Variant("Matrix", _rows(self.conjugation @ self.state.ctm @ ctm))}))


def import_scene(path, generated=None):
    """lib.rs: a pbrt scene file to (ops, warnings). `generated` is where a
    constant sky's EXR is written (defaults to beside the scene)."""
    path = Path(path)
    generated = Path(generated) if generated is not None else path.parent
    return Mapper(Parser(path), generated, path.stem).run()
