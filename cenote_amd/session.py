"""The progressive render loop as an actor (render/session.rs): a thread
accumulates as fast as the device allows while a consumer peeks at its
output. Four lanes cross the thread boundary, each behind its own short
lock (session.rs:1-51):

* inputs in — camera, target size, a generation counter, a running flag;
  latest wins, snapshotted once per sample;
* edits in — change-sets (``apply`` overlays, ``replace`` swaps in a whole
  scene); they land at the next sample boundary, the resident scene is
  updated in place for exactly what the edit touched (cn_scene_update:
  materials, lights, the environment, or — for meshes and placements — the
  geometry and BVH) and accumulation restarts at sample 0 when anything but
  settings changed;
* frames out — the resolved linear averages (beauty, albedo, normal,
  depth) with their sample count and last sample time, published at most
  every 15 ms;
* faults out — a rejected edit is posted for ``take_edit_error`` while the
  last good scene keeps rendering; any other failure ends the thread and
  surfaces through ``check``.
"""

from __future__ import annotations

import dataclasses
import threading
import time

import numpy as np

from ._abi import CN_E_SCENE as STATUS_SCENE
from ._abi import DIRTY_CAMERA, CenoteError
from .render import Context, Film, LightSampling, Renderer, Scene
from .scene import CameraData, dirty_bits
from .scenefile import Description, SceneFileError, lower

PUBLISH_INTERVAL = 0.015  # session.rs:65-70
IDLE_NAP = 0.016


@dataclasses.dataclass
class Frame:
    beauty: np.ndarray
    albedo: np.ndarray
    normal: np.ndarray
    depth: np.ndarray
    width: int
    height: int
    samples: int
    sample_time: float


@dataclasses.dataclass
class _Inputs:
    camera: CameraData
    size: tuple
    generation: int = 0
    running: bool = True


class Session:
    """Render ``description`` progressively on its own thread."""

    def __init__(self, ctx: Context, description: Description, width: int, height: int,
                 camera: CameraData | None = None, max_bounces: int = 8,
                 light_sampling: LightSampling = LightSampling.MIS):
        self.ctx = ctx
        data, _ = lower(description)
        self._lock_inputs = threading.Lock()
        self._lock_edits = threading.Lock()
        self._lock_error = threading.Lock()
        self._lock_frame = threading.Lock()
        self._inputs = _Inputs(camera or data.camera, (width, height))
        self._edits = []
        self._edit_error = None
        self._frame = None
        self._failure = None
        self._renderer = Renderer(ctx, max_bounces=max_bounces, light_sampling=light_sampling)
        self._description = description
        self._data = data
        self._scene = Scene(ctx, data)
        self._thread = threading.Thread(target=self._run, name="cenote-render", daemon=True)
        self._thread.start()

    # --- the consumer's side --------------------------------------------------------

    def set_camera(self, camera: CameraData):
        with self._lock_inputs:
            self._inputs.camera = camera
            self._inputs.generation += 1

    def resize(self, width: int, height: int):
        with self._lock_inputs:
            self._inputs.size = (width, height)

    def apply(self, ops):
        with self._lock_edits:
            self._edits.append(("apply", list(ops)))

    def replace(self, ops):
        with self._lock_edits:
            self._edits.append(("replace", list(ops)))

    def take_edit_error(self):
        with self._lock_error:
            err, self._edit_error = self._edit_error, None
        return err

    def take_frame(self) -> Frame | None:
        with self._lock_frame:
            frame, self._frame = self._frame, None
        return frame

    def check(self):
        """Raise the render thread's failure, if it has ended with one."""
        if not self._thread.is_alive() and self._failure is not None:
            failure, self._failure = self._failure, None
            raise failure

    def close(self):
        with self._lock_inputs:
            self._inputs.running = False
        self._thread.join()
        self._scene.close()
        self._renderer.close()
        failure, self._failure = self._failure, None
        if failure is not None:
            raise failure

    # --- the render thread ------------------------------------------------------------

    def _run(self):
        try:
            self._loop()
        except BaseException as e:  # surfaces through check()/close()
            self._failure = e

    def _apply_edits(self) -> bool:
        with self._lock_edits:
            edits, self._edits = self._edits, []
        if not edits:
            return False
        changed = False
        for verb, ops in edits:
            try:
                if verb == "apply":
                    fresh = Description()
                    fresh.objects = {k: dict(v) for k, v in self._description.objects.items()}
                    fresh.apply(ops)
                else:
                    fresh = Description()
                    fresh.apply(ops)
                data, _ = lower(fresh)
            except (SceneFileError, OSError, ValueError) as e:
                with self._lock_error:
                    self._edit_error = e
                continue
            # Scene::update (scene/prep.rs:122-177): only what the edit
            # touched is rebuilt; the camera follows the session's inputs.
            dirty = dirty_bits(self._data, data) & ~DIRTY_CAMERA
            if dirty:
                try:
                    self._scene.update(data, dirty)
                except CenoteError as e:
                    if e.status != STATUS_SCENE:
                        raise
                    with self._lock_error:
                        self._edit_error = e
                    continue
                changed = True
            self._description = fresh
            self._data = data
        return changed

    def _loop(self):
        film = None
        applied_size = (0, 0)
        applied_generation = 0
        last_publish = None
        while True:
            with self._lock_inputs:
                inputs = dataclasses.replace(self._inputs)
            if not inputs.running:
                break
            width, height = inputs.size
            if width == 0 or height == 0:
                time.sleep(IDLE_NAP)
                continue
            if inputs.size != applied_size:
                if film is not None:
                    film.close()
                film = Film(self.ctx, width, height)
                self._scene.set_camera(inputs.camera)
                applied_size = inputs.size
                applied_generation = inputs.generation
                last_publish = None
            if self._apply_edits():
                self._scene.set_camera(inputs.camera)
                film.reset()
                last_publish = None
            if inputs.generation != applied_generation:
                self._scene.set_camera(inputs.camera)
                film.reset()
                applied_generation = inputs.generation
            t0 = time.perf_counter()
            self._renderer.accumulate(self._scene, film)
            sample_time = time.perf_counter() - t0
            if last_publish is None or time.perf_counter() - last_publish >= PUBLISH_INTERVAL:
                b, a, n, d = self._renderer.resolve(film)
                frame = Frame(b, a, n, d, width, height, film.samples, sample_time)
                with self._lock_frame:
                    self._frame = frame
                last_publish = time.perf_counter()
        if film is not None:
            film.close()
