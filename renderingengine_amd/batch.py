"""Host mirror of the batch engine (``rex_batch_*``): many render jobs of
any resident scene, resolution, spp and max bounce through one shared set
of wavefront launches per step, with continuous admission.

``ShadedSceneHandle`` uploads one scene and its shading tables (the scene
half of ``Wave2Runtime::SetScene``, src/renderers/Wave2Runtime.cpp:462-666);
``BatchEngine.add_scene`` makes it resident, ``submit`` queues one frame
(the ``rex_frame_desc`` of ``WavefrontRuntime.update_frame``), ``step`` /
``run`` advance the engine on a HIP stream.  A job's Film equals the frame
rendered alone by ``WavefrontRuntime``.  Errors raise ``_lib.RexError``.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from .render import _stream_handle, make_camera
from .traversal import Scene

STATE_QUEUED, STATE_RUNNING, STATE_RESOLVING, STATE_FINISHED = 0, 1, 2, 3


class ShadedSceneHandle:
    """A rex scene with its shading tables (no renderer)."""

    def __init__(self, tris, materials, lights, light_selection=_lib.REX_LIGHT_SELECTION_POWER, fingerprint=1,
                 generation=1, max_leaf_size=4, build_threads=0, environment=None, env_sampler=0, alpha=None):
        self.scene = Scene(tris, fingerprint=fingerprint, generation=generation, max_leaf_size=max_leaf_size,
                           build_threads=build_threads, alpha=alpha)
        desc, _keep = _lib.shading_desc(materials, lights, light_selection, environment, env_sampler)
        _lib.check(_lib.load().rex_scene_set_shading(self.scene.handle, ctypes.byref(desc)))

    @property
    def handle(self):
        return self.scene.handle

    def destroy(self):
        if self.scene is not None:
            self.scene.destroy()
            self.scene = None


def frame_desc(camera, width, height, sample_index=0, spp=1, max_bounce=4, estimator=_lib.REX_ESTIMATOR_MIS,
               seed=0, stream_tag=0, rr=(0.0, 0.0, 0.0), ray_epsilon=0.0):
    f = _lib.FrameDesc()
    f.camera = make_camera(camera)
    f.width, f.height, f.spp = int(width), int(height), int(spp)
    f.maxBounce, f.sampleIndex, f.estimator = int(max_bounce), int(sample_index), int(estimator)
    f.seed = int(seed)
    f.streamTag, f.walkMode = int(stream_tag), _lib.REX_WALK_WIDE_COMPRESSED
    f.rrStart, f.rrMinimum, f.rrMaximum = (float(x) for x in rr)
    f.rayEpsilon = float(ray_epsilon)
    return f


class BatchEngine:
    def __init__(self, path_capacity=0, max_jobs=0, use_graphs=False, shade_blocks_per_cu=0,
                 trace_blocks_per_cu=0, output_mib=0, autotune=False):
        self._lib = _lib.load()
        d = _lib.BatchDesc()
        d.pathCapacity, d.maxJobs, d.useGraphs = int(path_capacity), int(max_jobs), 1 if use_graphs else 0
        d.shadeBlocksPerCu, d.traceBlocksPerCu = int(shade_blocks_per_cu), int(trace_blocks_per_cu)
        d.outputMiB = int(output_mib)
        d.autotune = 1 if autotune else 0
        h = ctypes.c_void_p()
        _lib.check(self._lib.rex_batch_create(ctypes.byref(d), ctypes.byref(h)))
        self.handle = h.value
        self._sizes = {}

    def add_scene(self, shaded):
        idx = ctypes.c_uint32()
        _lib.check(self._lib.rex_batch_add_scene(self.handle, shaded.handle, ctypes.byref(idx)))
        return idx.value

    def submit(self, scene_index, frame):
        """frame: a ``_lib.FrameDesc`` (see ``frame_desc``); returns the job id."""
        jid = ctypes.c_uint64()
        _lib.check(self._lib.rex_batch_submit(self.handle, int(scene_index), ctypes.byref(frame), ctypes.byref(jid)))
        self._sizes[jid.value] = frame.width * frame.height
        return jid.value

    def step(self, stream=None):
        _lib.check(self._lib.rex_batch_step(self.handle, _stream_handle(stream)))

    def run(self, stream=None):
        _lib.check(self._lib.rex_batch_run(self.handle, _stream_handle(stream)))

    def job(self, job_id):
        info = _lib.BatchJobInfo()
        _lib.check(self._lib.rex_batch_job(self.handle, int(job_id), ctypes.byref(info)))
        return {"state": info.state, "width": info.width, "height": info.height, "submitStep": info.submitStep,
                "finishStep": info.finishStep, "film": info.film}

    def film(self, job_id):
        """Host copy of a finished job's Film: (width * height, 4) float32."""
        info = self.job(job_id)
        out = np.zeros((info["width"] * info["height"], 4), dtype=np.float32)
        _lib.check(self._lib.rex_synchronize())
        _lib.check(self._lib.rex_copy_to_host(out.ctypes.data, info["film"], out.nbytes))
        return out

    def release(self, job_id):
        _lib.check(self._lib.rex_batch_release(self.handle, int(job_id)))

    def stats(self):
        s = _lib.BatchStats()
        _lib.check(self._lib.rex_batch_get_stats(self.handle, ctypes.byref(s)))
        return {k: getattr(s, k) for k, _ in s._fields_ if k != "reserved"}

    def destroy(self):
        if self.handle:
            self._lib.rex_batch_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass
