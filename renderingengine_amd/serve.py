"""Host mirror of the request loop (``rex_server_*``): render jobs arrive at
any time, are admitted in arrival order while their HBM footprint fits the
budget, and run overlapped on the device -- frames in the batch engine's
shared launches, ReSTIR DI sequences on renderer slots of their own streams
with reservoir and Film history kept on the device between frames.

A job's result equals the same job run alone: a frame's Film equals
``WavefrontRuntime``'s frame, a sequence's last Film equals the same frames
rendered back to back by a fresh ``WavefrontRuntime`` with the ReSTIR
estimator.  Errors raise ``_lib.RexError``.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from .render import make_camera

STATE_QUEUED, STATE_ADMITTED, STATE_FINISHED = 0, 1, 2


def restir_desc(candidates=32, neighbors=5, max_m=20, max_age=20, temporal=True,
                neighbor_mode=_lib.REX_RESTIR_NEIGHBORS_RING, radius=30, validation=None):
    d = _lib.RestirDesc()
    d.candidates, d.neighbors, d.maxM, d.maxAge = int(candidates), int(neighbors), int(max_m), int(max_age)
    d.temporal, d.neighborMode, d.radius = 1 if temporal else 0, int(neighbor_mode), int(radius)
    if validation is not None:
        d.validation[:] = [float(x) for x in validation]
    return d


class Server:
    def __init__(self, budget_bytes=0, path_capacity=0, max_jobs=0, use_graphs=False, shade_blocks_per_cu=0,
                 output_mib=0, sequence_slots=0, sequence_max_pixels=0, sequence_output_mib=0):
        self._lib = _lib.load()
        d = _lib.ServerDesc()
        d.budgetBytes, d.pathCapacity, d.maxJobs = int(budget_bytes), int(path_capacity), int(max_jobs)
        d.useGraphs, d.shadeBlocksPerCu, d.outputMiB = 1 if use_graphs else 0, int(shade_blocks_per_cu), int(output_mib)
        d.sequenceSlots, d.sequenceMaxPixels = int(sequence_slots), int(sequence_max_pixels)
        d.sequenceOutputMiB = int(sequence_output_mib)
        h = ctypes.c_void_p()
        _lib.check(self._lib.rex_server_create(ctypes.byref(d), ctypes.byref(h)))
        self.handle = h.value

    def add_scene(self, shaded):
        """shaded: a ``batch.ShadedSceneHandle`` (must outlive the server)."""
        idx = ctypes.c_uint32()
        _lib.check(self._lib.rex_server_add_scene(self.handle, shaded.handle, ctypes.byref(idx)))
        return idx.value

    def _submit(self, kind, scene_index, frame, frames=1, cameras=None, restir=None):
        j = _lib.ServerJobDesc()
        j.kind, j.sceneIndex, j.frames = int(kind), int(scene_index), int(frames)
        j.frame = frame
        cams = None
        if cameras is not None:
            cams = (_lib.Camera * len(cameras))(*[make_camera(c) for c in cameras])
            j.cameras = ctypes.cast(cams, ctypes.c_void_p)
        if restir is not None:
            j.restir = ctypes.cast(ctypes.pointer(restir), ctypes.c_void_p)
        jid = ctypes.c_uint64()
        _lib.check(self._lib.rex_server_submit(self.handle, ctypes.byref(j), ctypes.byref(jid)))
        return jid.value

    def submit_frame(self, scene_index, frame):
        """frame: a ``_lib.FrameDesc`` (``batch.frame_desc``); returns the job id."""
        return self._submit(_lib.REX_JOB_FRAME, scene_index, frame)

    def submit_sequence(self, scene_index, frame, cameras, restir=None):
        """A ReSTIR DI sequence: len(cameras) frames at one sample per pixel,
        frame k with cameras[k] and sample index frame.sampleIndex + k."""
        frame.estimator, frame.spp = _lib.REX_ESTIMATOR_RESTIR, 1
        return self._submit(_lib.REX_JOB_RESTIR_SEQUENCE, scene_index, frame, len(cameras), cameras, restir)

    def pump(self):
        n = ctypes.c_uint32()
        _lib.check(self._lib.rex_server_pump(self.handle, ctypes.byref(n)))
        return n.value

    def drain(self):
        _lib.check(self._lib.rex_server_drain(self.handle))

    def job(self, job_id):
        i = _lib.ServerJobInfo()
        _lib.check(self._lib.rex_server_job(self.handle, int(job_id), ctypes.byref(i)))
        return {k: getattr(i, k) for k, _ in i._fields_}

    def film(self, job_id):
        """Host copy of a finished job's Film: (width * height, 4) float32."""
        info = self.job(job_id)
        if info["state"] != STATE_FINISHED:
            raise _lib.RexError(f"job {job_id} has not finished")
        out = np.zeros((info["width"] * info["height"], 4), dtype=np.float32)
        _lib.check(self._lib.rex_synchronize())
        _lib.check(self._lib.rex_copy_to_host(out.ctypes.data, info["film"], out.nbytes))
        return out

    def release(self, job_id):
        _lib.check(self._lib.rex_server_release(self.handle, int(job_id)))

    def stats(self):
        s = _lib.ServerStats()
        _lib.check(self._lib.rex_server_get_stats(self.handle, ctypes.byref(s)))
        return {k: getattr(s, k) for k, _ in s._fields_}

    def destroy(self):
        if self.handle:
            self._lib.rex_server_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass
