"""cenote_amd — an MI355X (gfx950) wavefront path tracer with the drop-in
C ABI of include/cenote_hip.h, for the hot path of joseph-wardle/cenote.

The compute path is libcenote_hip.so (hand-written HIP kernels, built
in-tree by ``cenote_amd.build``); this package is the Python host harness
over it.
"""

from ._abi import CenoteError, DEFAULT_CAPACITY, DEFAULT_MAX_BOUNCES, POOL_CAPACITY  # noqa: F401
from .render import Context, Film, LightSampling, Renderer, Scene  # noqa: F401
from .scene import CameraData, Material, MeshData, OpenPBR, SceneData  # noqa: F401
