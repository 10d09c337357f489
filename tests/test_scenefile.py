"""The scene-file front end without a GPU: the RON reader, the format's
version and unknown-field rules (format.rs:235-303), the change-set apply
contract and its validation (changeset.rs:1-23, 510-690), PLY reading
(ply.rs:431-579) and lowering to the library's records (lower.rs)."""

import struct

import numpy as np
import pytest

from cenote_amd import ply, ron, scenefile
from cenote_amd.scene import OpenPBR
from cenote_amd.scenefile import Description, SceneFileError


# --- RON ------------------------------------------------------------------------

def test_ron_values():
    v = ron.loads('''(
        // a comment
        a: 1, b: -2.5e3, c: "x\\"y\\n", d: [1, 2, 3,], e: (1.0, 2.0),
        f: Some(None), g: None, h: Some(Some(3)), i: Trs(scale: (1, 2, 3)),
        j: Constant((0.5, 0.5, 0.5)), k: Unit, l: true, /* block */ m: (),
    )''')
    assert v["a"] == 1 and v["b"] == -2500.0 and v["c"] == 'x"y\n'
    assert v["d"] == [1, 2, 3] and v["e"] == (1.0, 2.0)
    assert v["f"] == ron.Some(None) and v["g"] is None and v["h"] == ron.Some(ron.Some(3))
    assert v["i"] == ron.Variant("Trs", {"scale": (1, 2, 3)})
    assert v["j"] == ron.Variant("Constant", (0.5, 0.5, 0.5))
    assert v["k"] == ron.Variant("Unit") and v["l"] is True and v["m"] == {}


def test_ron_errors_name_their_place():
    with pytest.raises(ron.RonError, match="line 2"):
        ron.loads("(a: 1,\n b: ])")
    with pytest.raises(ron.RonError):
        ron.loads("(a: 1, a: 2)")


# --- format.rs tests ----------------------------------------------------------------

def test_a_newer_version_is_refused_by_number():
    with pytest.raises(SceneFileError, match="999.*reads 1"):
        scenefile.from_ron("(version: 999, ops: [])")


def test_non_scene_text_is_refused():
    with pytest.raises(SceneFileError, match="not a cenote scene"):
        scenefile.from_ron("hello")


def test_unknown_fields_are_refused_not_skipped():
    with pytest.raises(SceneFileError):
        scenefile.from_ron("(version: 1, ops: [], extra: 5)")
    with pytest.raises(SceneFileError, match="base_colour"):
        scenefile.from_ron('(version: 1, ops: [Material((name: "m", base_colour: Constant((1.0, 0.0, 0.0))))])')


def test_load_rebases_relative_paths_against_the_file(tmp_path):
    scene = tmp_path / "scene.ron"
    scene.write_text('(version: 1, ops: [Mesh((name: "m", source: Some(Ply(path: "geo/mesh.ply"))))])')
    (op,) = scenefile.load(scene)
    assert op.value["source"].value.value["path"] == str(tmp_path.resolve() / "geo/mesh.ply")


HAND_WRITTEN = '''(
    version: 1,
    ops: [
        Settings((name: "main", spp: Some(16))),
        Camera((name: "main", position: Some((0.0, 1.0, 4.0)))),
        Mesh((name: "tri", source: Some(Inline(
            positions: [(0.0, 0.0, 0.0), (1.0, 0.0, 0.0), (0.0, 1.0, 0.0)],
            triangles: [(0, 1, 2)],
        )))),
        Material((name: "clay", base_color: Some(Constant((0.7, 0.3, 0.2))))),
        Instance((name: "thing", mesh: Some("tri"), material: Some("clay"))),
    ],
)'''


def test_a_hand_written_scene_parses():
    ops = scenefile.from_ron(HAND_WRITTEN)
    assert len(ops) == 5
    d = Description()
    d.apply(ops)
    assert d.objects["Settings"]["main"]["spp"] == 16
    assert tuple(d.objects["Settings"]["main"]["resolution"]) == (1280, 720)


# --- the apply contract ----------------------------------------------------------------

def test_get_or_create_later_fields_win_and_references_resolve_after_the_set():
    d = Description()
    d.apply(scenefile.from_ron('''(version: 1, ops: [
        Instance((name: "i", mesh: Some("tri"), material: Some("m"))),
        Material((name: "m", coat_weight: Some(0.5))),
        Material((name: "m", coat_weight: Some(0.25), fuzz_weight: Some(1.0))),
        Mesh((name: "tri", source: Some(Inline(positions: [(0.0, 0.0, 0.0), (1.0, 0.0, 0.0), (0.0, 1.0, 0.0)],
                                               triangles: [(0, 1, 2)])))),
        Camera((name: "c")),
    ])'''))
    m = d.objects["Material"]["m"]
    assert m["coat_weight"] == 0.25 and m["fuzz_weight"] == 1.0 and m["specular_ior"] == 1.5


def test_a_rejected_set_leaves_the_description_unchanged():
    d = Description()
    d.apply(scenefile.from_ron(HAND_WRITTEN))
    before = repr(d.objects)
    with pytest.raises(SceneFileError, match="does not exist"):
        d.apply(scenefile.from_ron('(version: 1, ops: [Settings((name: "main", spp: Some(3))), '
                                   'Instance((name: "x", mesh: Some("nope"), material: Some("clay")))])'))
    assert repr(d.objects) == before


@pytest.mark.parametrize("ops,message", [
    ('Remove(Material, "ghost")', "does not exist"),
    ('Mesh((name: "e"))', "no geometry"),
    ('Mesh((name: "n", source: Some(Inline(positions: [(0.0, 0.0, 0.0)], normals: Some([]), triangles: [(0, 0, 0)]))))',
     "normals"),
    ('Mesh((name: "o", source: Some(Inline(positions: [(0.0, 0.0, 0.0)], triangles: [(0, 0, 1)]))))', "out of bounds"),
    ('Camera((name: "c", up: Some((0.0, 0.0, -1.0))))', "parallel"),
    ('Camera((name: "c", vfov_degrees: Some(180.0)))', "fov"),
    ('Camera((name: "c", position: Some((0.0, 0.0, 0.0))))', "coincide"),
    ('Light((name: "l", light: Some(Distant(direction: (0.0, 0.0, 0.0), irradiance: (1.0, 1.0, 1.0)))))',
     "zero direction"),
    ('Settings((name: "s", resolution: Some((0, 5))))', "zero dimension"),
    ('Environment((name: "sky", path: Some("/no/such/sky.exr")))', "does not exist"),
    ('Material((name: "m")), Mesh((name: "t", source: Some(Inline(positions: [(0.0, 0.0, 0.0), (1.0, 0.0, 0.0), '
     '(0.0, 1.0, 0.0)], triangles: [(0, 1, 2)])))), Instance((name: "i", mesh: Some("t"), material: Some("m"), '
     'transform: Some(Trs(scale: (1.0, 0.0, 1.0)))))', "non-invertible"),
    ('Instance((name: "i"))', "never given"),
    ('Camera((name: ""))', "empty name"),
])
def test_validation_refuses_bad_sets(ops, message):
    with pytest.raises(SceneFileError, match=message):
        Description().apply(scenefile.from_ron(f"(version: 1, ops: [{ops}])"))


# --- PLY (ply.rs tests) ------------------------------------------------------------

QUAD_ASCII = b"""ply
format ascii 1.0
comment a unit quad
element vertex 4
property float x
property float y
property float z
property float nx
property float ny
property float nz
property float u
property float v
element face 1
property list uchar int vertex_indices
end_header
0 0 0 0 0 1 0 0
1 0 0 0 0 1 1 0
1 1 0 0 0 1 1 1
0 1 0 0 0 1 0 1
4 0 1 2 3
"""


def test_ascii_reads_streams_and_fan_triangulates():
    m = ply.parse(QUAD_ASCII)
    assert m.positions.shape == (4, 3) and m.normals is not None and m.uvs is not None
    assert m.triangles.tolist() == [[0, 1, 2], [0, 2, 3]]
    assert m.uvs[2].tolist() == [1.0, 1.0]


def test_binary_little_endian_reads():
    head = (b"ply\r\nformat binary_little_endian 1.0\r\nelement vertex 3\r\nproperty float x\r\nproperty float y\r\n"
            b"property float z\r\nelement face 1\r\nproperty list uint ushort vertex_indices\r\nend_header\r\n")
    body = struct.pack("<9f", 0, 0, 0, 1, 0, 0, 0, 1, 0) + struct.pack("<I3H", 3, 0, 1, 2)
    m = ply.parse(head + body)
    assert m.triangles.tolist() == [[0, 1, 2]] and m.normals is None and m.uvs is None
    assert m.positions[1].tolist() == [1.0, 0.0, 0.0]


def test_unknown_elements_are_read_past():
    text = QUAD_ASCII.replace(b"element face 1", b"element material 2\nproperty float shine\nproperty list uchar "
                                                   b"int stuff\nelement face 1")
    text = text.replace(b"4 0 1 2 3", b"0.5 2 7 8\n0.25 0\n4 0 1 2 3")
    assert ply.parse(text).triangles.tolist() == [[0, 1, 2], [0, 2, 3]]


@pytest.mark.parametrize("text", [
    b"obj\nformat ascii 1.0\nend_header\n",
    b"ply\nformat binary_big_endian 1.0\nend_header\n",
    b"ply\nformat ascii 1.0\nelement vertex 1\nproperty float w\nend_header\n1\n",
    b"ply\nformat ascii 1.0\nelement vertex 3\nproperty float x\nproperty float y\nproperty float z\nelement face 1\n"
    b"property list uchar int vertex_indices\nend_header\n0 0 0\n1 0 0\n0 1 0\n3 0 1 9\n",
    b"ply\nformat ascii 1.0\nelement vertex 999999999999\nproperty float x\nend_header\n",
])
def test_malformed_files_are_refused(text):
    with pytest.raises(ply.PlyError):
        ply.parse(text)


def test_a_missing_file_names_its_path(tmp_path):
    with pytest.raises(ply.PlyError, match="nope.ply"):
        ply.read(tmp_path / "nope.ply")


# --- lowering ------------------------------------------------------------------------

def test_lowering_builds_the_library_records(tmp_path):
    from cenote_amd import output

    (tmp_path / "quad.ply").write_bytes(QUAD_ASCII)
    sky = np.zeros((4, 8, 4), np.float32)
    sky[..., 0] = 1.0
    sky[..., 3] = 1.0
    sky[0, 0, 1] = np.nan
    output.write_exr(tmp_path / "sky.exr", sky)
    (tmp_path / "s.ron").write_text('''(version: 1, ops: [
        Settings((name: "main", resolution: Some((320, 200)), spp: Some(8), max_bounces: Some(5))),
        Camera((name: "main", position: Some((0.0, 0.5, 3.0)), look_at: Some((0.0, 0.5, 0.0)),
                aperture_radius: Some(0.1))),
        Environment((name: "sky", path: Some("sky.exr"))),
        Mesh((name: "quad", source: Some(Ply(path: "quad.ply")))),
        Mesh((name: "tri", source: Some(Inline(positions: [(0.0, 0.0, 0.0), (1.0, 0.0, 0.0), (0.0, 1.0, 0.0)],
                                               triangles: [(0, 1, 2)])))),
        Material((name: "gold", base_color: Some(Constant((1.0, 0.8, 0.3))), base_metalness: Some(Constant(1.0)),
                  coat_weight: Some(2.0))),
        Material((name: "lamp", emission_luminance: Some(5.0))),
        Instance((name: "b", mesh: Some("tri"), material: Some("lamp"))),
        Instance((name: "a", mesh: Some("quad"), material: Some("gold"),
                  transform: Some(Trs(translate: (1.0, 0.0, 0.0))))),
        Light((name: "sun", light: Some(Distant(direction: (0.0, -2.0, 0.0), irradiance: (1.0, 1.0, 1.0))))),
        Light((name: "dark", light: Some(Point(position: (0.0, 1.0, 0.0), intensity: (0.0, 0.0, 0.0))))),
    ])''')
    data, settings = scenefile.load_scene(tmp_path / "s.ron")
    assert settings["resolution"] == (320, 200) and settings["spp"] == 8 and settings["max_bounces"] == 5
    # Instances in name order: "a" (the quad) then "b".
    assert [i.mesh for i in data.instances] == [0, 1] and data.meshes[0].indices.shape == (2, 3)
    assert data.instances[0].transform[0, 3] == 1.0
    gold = OpenPBR(base_color=(1.0, 0.8, 0.3), base_metalness=1.0, coat_weight=2.0).lower()
    assert bytes(data.instances[0].material.record()) == bytes(gold.record())
    assert data.instances[1].material.emission[1] > 0.0
    # The powerless point light is skipped; the distant light is normalized.
    assert len(data.delta_lights) == 1 and data.delta_lights[0][0] == 1
    assert np.allclose(data.delta_lights[0][1], (0.0, -1.0, 0.0))
    # Thin lens focused at look_at when no focus distance is given.
    assert data.camera.aperture_radius == pytest.approx(0.1) and data.camera.focus_distance == pytest.approx(3.0)
    # The sky: ACEScg, non-finite texels black, alpha 1.
    assert data.env_texels.shape == (4, 8, 4) and np.all(data.env_texels[0, 0, :3] == 0.0)
    assert np.all(data.env_texels[..., 3] == 1.0) and data.env_texels[1, 1, 0] > 0.5


def test_the_singleton_rules_hold():
    d = Description()
    d.apply(scenefile.from_ron(HAND_WRITTEN.replace('Camera((name: "main", position: Some((0.0, 1.0, 4.0)))),', '')))
    with pytest.raises(SceneFileError, match="exactly one camera"):
        scenefile.lower(d)
    d.apply(scenefile.from_ron('(version: 1, ops: [Camera((name: "a")), Camera((name: "b"))])'))
    with pytest.raises(SceneFileError, match="exactly one camera"):
        scenefile.lower(d)


def test_textured_slots_are_refused_at_lowering(tmp_path):
    (tmp_path / "wood.png").write_bytes(b"\x89PNG")
    (tmp_path / "s.ron").write_text(HAND_WRITTEN.replace('base_color: Some(Constant((0.7, 0.3, 0.2)))',
                                                         'base_color: Some(Texture((path: "wood.png")))'))
    with pytest.raises(SceneFileError, match="textured"):
        scenefile.load_scene(tmp_path / "s.ron")
