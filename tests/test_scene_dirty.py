"""dirty_bits: which parts of a resident scene an edited description
invalidates (the Dirty set of scene/changeset.rs:351-356 at the grain of
cn_scene_update)."""

import copy
import dataclasses

import numpy as np

from cenote_amd import _abi, workloads
from cenote_amd.scene import Material, dirty_bits, trs


def test_an_unchanged_description_is_clean():
    data = workloads.build("cornell")
    assert dirty_bits(data, copy.deepcopy(data)) == 0


def test_each_kind_of_edit_names_its_bit():
    data = workloads.build("cornell")
    d = copy.deepcopy(data)
    d.instances[1].material = Material.matte((0.2, 0.2, 0.9))
    assert dirty_bits(data, d) == _abi.DIRTY_MATERIALS
    d = copy.deepcopy(data)
    d.instances[1].camera_visible = False
    assert dirty_bits(data, d) == _abi.DIRTY_MATERIALS
    d = copy.deepcopy(data)
    d.delta_lights.append((_abi.LIGHT_POINT, (0.0, 1.0, 0.0), (1.0, 1.0, 1.0)))
    assert dirty_bits(data, d) == _abi.DIRTY_LIGHTS
    d = copy.deepcopy(data)
    d.env_texels = d.env_texels * np.float32(2.0)
    assert dirty_bits(data, d) == _abi.DIRTY_ENVIRONMENT
    d = copy.deepcopy(data)
    d.camera = dataclasses.replace(d.camera, vfov_degrees=30.0)
    assert dirty_bits(data, d) == _abi.DIRTY_CAMERA


def test_geometry_edits_supersede_material_edits():
    data = workloads.build("cornell")
    d = copy.deepcopy(data)
    d.instances[0].transform = trs((0.0, 0.1, 0.0))
    d.instances[1].material = Material.matte((0.2, 0.2, 0.9))
    assert dirty_bits(data, d) == _abi.DIRTY_GEOMETRY
    d = copy.deepcopy(data)
    d.meshes[0].positions = d.meshes[0].positions * np.float32(1.01)
    assert dirty_bits(data, d) == _abi.DIRTY_GEOMETRY
    d = copy.deepcopy(data)
    d.instances.append(copy.deepcopy(d.instances[0]))
    assert dirty_bits(data, d) == _abi.DIRTY_GEOMETRY
