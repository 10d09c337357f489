"""Incremental scene updates (Scene::update, scene/prep.rs:122-177): a scene
updated in place for what an edit touched renders, bit for bit, as a scene
freshly built from the edited description — its BVH triangles, light
records, environment tables and a multi-sample film all equal — and a
rejected edit leaves the scene as it was (prep.rs:113-121)."""

import copy
import ctypes
import dataclasses

import numpy as np
import pytest

import cenote_amd as cn
from cenote_amd import _abi, workloads
from cenote_amd.scene import Material, dirty_bits, trs

pytestmark = pytest.mark.gpu

W, H, SPP = 160, 90, 4


def render(ctx, renderer, scene):
    film = cn.Film(ctx, W, H)
    renderer.accumulate(scene, film, count=SPP)
    out = film.averages()
    film.close()
    return out


def state(scene):
    _, tris = scene.bvh()
    marginal, conditional, pdfs = scene.env_tables()
    info = scene.info()
    return tris, scene.light_records(), marginal, conditional, pdfs, info.env_select_prob, info.light_count


def assert_same(ctx, renderer, updated, data):
    fresh = cn.Scene(ctx, data)
    try:
        for a, b in zip(state(updated), state(fresh)):
            assert np.array_equal(np.asarray(a), np.asarray(b))
        for a, b in zip(render(ctx, renderer, updated), render(ctx, renderer, fresh)):
            assert np.array_equal(a, b)
    finally:
        fresh.close()


def edit_materials(data):
    d = copy.deepcopy(data)
    d.instances[0].material = dataclasses.replace(d.instances[0].material, base_color=(0.1, 0.7, 0.2))
    d.instances[3].material = Material.emitter((4.0, 2.0, 1.0))
    d.instances[7].material = d.instances[7].material.with_opacity(0.4)
    d.instances[9].camera_visible = False
    d.instances[-1].material = Material.matte((0.5, 0.5, 0.5))  # the key light goes dark
    return d


def edit_lights(data):
    d = copy.deepcopy(data)
    d.delta_lights.append((_abi.LIGHT_DISTANT, (0.3, -1.0, -0.2), (2.0, 1.8, 1.5)))
    d.delta_lights.append((_abi.LIGHT_POINT, (0.0, 3.0, 2.0), (30.0, 30.0, 30.0)))
    return d


def edit_environment(data):
    d = copy.deepcopy(data)
    d.env_texels = workloads.sky(512, 256, seed=5)
    return d


def edit_geometry(data):
    d = copy.deepcopy(data)
    d.instances[12].transform = trs((0.4, 1.5, -0.3), (0.0, 30.0, 0.0), (0.7, 0.7, 0.7))
    return d


EDITS = {
    "materials": (edit_materials, _abi.DIRTY_MATERIALS),
    "lights": (edit_lights, _abi.DIRTY_LIGHTS),
    "environment": (edit_environment, _abi.DIRTY_ENVIRONMENT),
    "geometry": (edit_geometry, _abi.DIRTY_GEOMETRY),
}


@pytest.mark.parametrize("name", list(EDITS))
def test_an_updated_scene_renders_as_a_fresh_one(ctx, name):
    edit, bit = EDITS[name]
    data = workloads.build("chart")
    edited = edit(data)
    assert dirty_bits(data, edited) == bit
    scene = cn.Scene(ctx, data)
    r = cn.Renderer(ctx, max_bounces=12)
    try:
        render(ctx, r, scene)  # graphs captured against the old scene
        scene.update(edited, bit)
        assert_same(ctx, r, scene, edited)
    finally:
        r.close()
        scene.close()


def test_updates_compose(ctx):
    """Three edits in a row, each its own update, end where one fresh build
    of the final description starts."""
    data = workloads.build("chart")
    scene = cn.Scene(ctx, data)
    r = cn.Renderer(ctx, max_bounces=12)
    try:
        current = data
        for name in ("environment", "materials", "lights"):
            edited = EDITS[name][0](current)
            scene.update(edited, dirty_bits(current, edited))
            current = edited
        assert_same(ctx, r, scene, current)
    finally:
        r.close()
        scene.close()


def test_a_rejected_update_leaves_the_scene_untouched(ctx):
    data = workloads.build("chart")
    scene = cn.Scene(ctx, data)
    r = cn.Renderer(ctx, max_bounces=12)
    try:
        before = render(ctx, r, scene)
        before_state = state(scene)
        # A placement edit presented as a material edit.
        with pytest.raises(_abi.CenoteError) as e:
            scene.update(edit_geometry(data), _abi.DIRTY_MATERIALS)
        assert e.value.status == _abi.CN_E_SCENE
        # A material the path cannot render (a texture slot set).
        desc, keep = edit_materials(data).desc()
        desc.instances[5].material.base_color_texture = 0
        status = scene._lib.cn_scene_update(scene.handle, ctypes.byref(desc), _abi.DIRTY_MATERIALS)
        del keep
        assert status == _abi.CN_E_SCENE
        # An unknown delta-light kind.
        bad = edit_lights(data)
        bad.delta_lights[0] = (7, (0.0, -1.0, 0.0), (1.0, 1.0, 1.0))
        with pytest.raises(_abi.CenoteError) as e:
            scene.update(bad, _abi.DIRTY_LIGHTS | _abi.DIRTY_MATERIALS)
        assert e.value.status == _abi.CN_E_SCENE
        for a, b in zip(state(scene), before_state):
            assert np.array_equal(np.asarray(a), np.asarray(b))
        for a, b in zip(render(ctx, r, scene), before):
            assert np.array_equal(a, b)
    finally:
        r.close()
        scene.close()


def test_a_camera_update_moves_only_the_camera(ctx):
    data = workloads.build("chart")
    edited = copy.deepcopy(data)
    edited.camera = dataclasses.replace(data.camera, position=(1.0, 4.0, 10.0))
    assert dirty_bits(data, edited) == _abi.DIRTY_CAMERA
    scene = cn.Scene(ctx, data)
    r = cn.Renderer(ctx, max_bounces=12)
    try:
        scene.update(edited, _abi.DIRTY_CAMERA)
        assert_same(ctx, r, scene, edited)
    finally:
        r.close()
        scene.close()


# --- the edit walk of scene/prep.rs:270-410 through the scene-file front end ---

WALK_SCENE = '''(version: 1, ops: [
    Camera((name: "main", position: Some((0.0, 2.0, 6.0)), look_at: Some((0.0, 0.5, 0.0)))),
    Mesh((name: "floor", source: Some(Inline(
        positions: [(-4.0, 0.0, -4.0), (-4.0, 0.0, 4.0), (4.0, 0.0, 4.0), (4.0, 0.0, -4.0)],
        triangles: [(0, 1, 2), (0, 2, 3)])))),
    Mesh((name: "key", source: Some(Inline(
        positions: [(-0.5, 3.0, -0.5), (0.5, 3.0, -0.5), (0.5, 3.0, 0.5), (-0.5, 3.0, 0.5)],
        triangles: [(0, 1, 2), (0, 2, 3)])))),
    Mesh((name: "sphere", source: Some(Inline(
        positions: [(0.0, 0.5, 0.0), (-0.5, 0.0, 0.0), (0.0, 0.0, 0.5), (0.5, 0.0, 0.0), (0.0, 0.0, -0.5),
                    (0.0, -0.5, 0.0)],
        triangles: [(0, 1, 2), (0, 2, 3), (0, 3, 4), (0, 4, 1), (5, 2, 1), (5, 3, 2), (5, 4, 3), (5, 1, 4)])))),
    Material((name: "floor", base_color: Some(Constant((0.6, 0.6, 0.6))))),
    Material((name: "key", emission_luminance: Some(12.0))),
    Material((name: "chart_a", base_color: Some(Constant((0.8, 0.35, 0.2))), base_metalness: Some(Constant(1.0)))),
    Material((name: "chart_b", base_color: Some(Constant((0.2, 0.35, 0.8))), specular_roughness: Some(Constant(0.3)))),
    Instance((name: "floor", mesh: Some("floor"), material: Some("floor"))),
    Instance((name: "key", mesh: Some("key"), material: Some("key"))),
    Instance((name: "chart_a", mesh: Some("sphere"), material: Some("chart_a"),
              transform: Some(Trs(translate: (-1.0, 0.5, 0.0))))),
    Instance((name: "chart_b", mesh: Some("sphere"), material: Some("chart_b"),
              transform: Some(Trs(translate: (1.0, 0.5, 0.0))))),
])'''

WALK = [
    ("material", 'Material((name: "floor", base_color: Some(Constant((0.1, 0.6, 0.2)))))'),
    ("emission", 'Material((name: "key", emission_luminance: Some(30.0)))'),
    ("transform", 'Instance((name: "chart_a", transform: Some(Trs(translate: (0.0, 1.5, 0.0), '
                  'scale: (0.75, 0.75, 0.75)))))'),
    ("topology", 'Mesh((name: "sphere", source: Some(Inline(positions: [(0.0, 0.6, 0.0), (-0.6, 0.0, 0.0), '
                 '(0.0, 0.0, 0.6), (0.6, 0.0, 0.0)], triangles: [(0, 1, 2), (0, 2, 3), (0, 3, 1), (1, 3, 2)]))))'),
    ("removal", 'Remove(Instance, "chart_b"), Remove(Material, "chart_b")'),
    ("environment", 'Environment((name: "sky", path: Some("SKY")))'),
    ("camera", 'Camera((name: "main", position: Some((2.0, 4.0, 9.0))))'),
    ("delta light", 'Light((name: "sun", light: Some(Distant(direction: (0.2, -1.0, 0.1), '
                    'irradiance: (1.5, 1.4, 1.2)))))'),
    ("camera visibility", 'Instance((name: "key", camera_visible: Some(false)))'),
    ("closure and opacity", 'Material((name: "floor", coat_weight: Some(1.0), coat_roughness: Some(0.2), '
                            'geometry_opacity: Some(Constant(0.5))))'),
]


def test_incremental_updates_match_a_fresh_build(ctx, tmp_path):
    """scene/prep.rs:411-457: after every edit of the walk, the scene
    updated in place for what the edit dirtied renders bit for bit as a
    fresh build of the edited description (the walk's texture steps are
    outside this path: material textures are rejected)."""
    from cenote_amd import output, scenefile

    sky = np.full((2, 2, 4), 0.3, np.float32)
    output.write_exr(tmp_path / "sky.exr", sky)
    d = scenefile.Description()
    d.apply(scenefile.from_ron(WALK_SCENE))
    data, _ = scenefile.lower(d)
    scene = cn.Scene(ctx, data)
    r = cn.Renderer(ctx, max_bounces=8)
    try:
        for label, ops in WALK:
            text = ops.replace("SKY", str(tmp_path / "sky.exr"))
            d.apply(scenefile.from_ron(f"(version: 1, ops: [{text}])"))
            edited, _ = scenefile.lower(d)
            bits = dirty_bits(data, edited)
            assert bits, label
            print(f"edit {label}: dirty 0x{bits:x}", flush=True)
            scene.update(edited, bits)
            data = edited
            assert_same(ctx, r, scene, data)
    finally:
        r.close()
        scene.close()
