"""Reconstruction boundary: the ABI-v2 records a frame publishes and the
post passes over them.

* ``ReconstructionTargets`` owns the device buffers a renderer publishes
  into (the set-4 bindings of
  resources/shaders/integrators/pbr_reconstruction_export_v2.hlsli:17-32):
  G-buffer (208 B), motion inputs (32 B), raw split signal (64 B), the ReSTIR
  primary-surface copy (96 B) and the terminal signal (16 B) per pixel.
* ``motion_vectors`` runs MotionVectorsCS
  (reconstruction/shaders/MotionVectorsCS.hlsl:35-90) over a G-buffer.
* ``compose`` runs ComposeCS (reconstruction/shaders/ComposeCS.hlsl:26-126)
  for the raw, progressive-mean, motion and material / surface debug outputs.

Record layouts follow include/contracts/ReconstructionAbiV2.hpp:45-116.
Everything runs through librex.so; there is no host fallback.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from .restir_abi import SURFACE as PRIMARY_SURFACE
from .traversal import DeviceBuffer

_F4 = (np.float32, (4,))
_U4 = (np.uint32, (4,))
MOTION_VECTOR = np.dtype([("motionExpectedDepth", *_F4), ("currentPreviousUv", *_F4), ("identity", *_U4)])
SIGNAL = np.dtype([("directDiffuse", *_F4), ("directSpecular", *_F4), ("indirectDiffuse", *_F4),
                   ("indirectSpecular", *_F4)])
GBUFFER = np.dtype([("primary", PRIMARY_SURFACE), ("motion", MOTION_VECTOR), ("signal", SIGNAL)])
MOTION_INPUT = np.dtype([("objectPosition", np.float32, (3,)), ("currentTransformIndex", np.uint32),
                         ("previousTransformIndex", np.uint32), ("reserved", np.uint32, (3,))])
assert (MOTION_VECTOR.itemsize, SIGNAL.itemsize, GBUFFER.itemsize, MOTION_INPUT.itemsize) == (48, 64, 208, 32)

PRIMARY_SURFACE_VALID = _lib.REX_PRIMARY_SURFACE_VALID
MOTION_VALID = _lib.REX_MOTION_VALID
COMPOSE_RAW = _lib.REX_COMPOSE_RAW
COMPOSE_MOTION = _lib.REX_COMPOSE_MOTION
COMPOSE_BASE_COLOR = _lib.REX_COMPOSE_BASE_COLOR
COMPOSE_NORMAL = _lib.REX_COMPOSE_NORMAL
COMPOSE_ROUGHNESS = _lib.REX_COMPOSE_ROUGHNESS
COMPOSE_METALLIC = _lib.REX_COMPOSE_METALLIC
COMPOSE_EMISSIVE = _lib.REX_COMPOSE_EMISSIVE
COMPOSE_PROGRESSIVE_MEAN = _lib.REX_COMPOSE_PROGRESSIVE_MEAN

_KINDS = (("gbuffer", GBUFFER), ("motionInputs", MOTION_INPUT), ("rawSignal", SIGNAL),
          ("primarySurface", PRIMARY_SURFACE), ("terminal", np.dtype((np.float32, (4,)))))


class ReconstructionTargets:
    """Device buffers of ``pixels`` records each; ``kinds`` names the ones
    to allocate (default: all five)."""

    def __init__(self, pixels, kinds=None):
        self.pixels = int(pixels)
        want = set(kinds) if kinds is not None else {k for k, _ in _KINDS}
        self.buffers = {k: DeviceBuffer(self.pixels * dt.itemsize) for k, dt in _KINDS if k in want}

    def struct(self):
        t = _lib.ReconstructionTargets()
        for k, _ in _KINDS:
            setattr(t, k, self.buffers[k].ptr if k in self.buffers else None)
        t.pixels = self.pixels
        return t

    def ptr(self, kind):
        return self.buffers[kind].ptr

    def download(self, pixels=None):
        """Host copies of the allocated records (dict by kind)."""
        n = self.pixels if pixels is None else int(pixels)
        _lib.check(_lib.load().rex_synchronize())
        out = {}
        for k, dt in _KINDS:
            if k in self.buffers:
                out[k] = self.buffers[k].download(dt, (n,))
        return out

    def free(self):
        for b in self.buffers.values():
            b.free()
        self.buffers = {}


def camera_matrices(camera):
    """The renderer's (view-projection, world-to-view) rows of a camera dict
    or ``_lib.Camera``: float32 (4, 4) each."""
    from .render import make_camera
    cam = camera if isinstance(camera, _lib.Camera) else make_camera(camera)
    vp = np.zeros(16, dtype=np.float32)
    wv = np.zeros(16, dtype=np.float32)
    _lib.check(_lib.load().rex_camera_matrices(ctypes.byref(cam), vp.ctypes.data, wv.ctypes.data))
    return vp.reshape(4, 4), wv.reshape(4, 4)


def motion_constants(width, height, current_vp, previous_vp, previous_world_to_view, current_jitter=(0.0, 0.0),
                     previous_jitter=(0.0, 0.0), current_count=1, previous_count=1):
    c = _lib.MotionConstants()
    c.currentViewProjection[:] = [float(x) for x in np.asarray(current_vp, np.float32).reshape(16)]
    c.previousViewProjection[:] = [float(x) for x in np.asarray(previous_vp, np.float32).reshape(16)]
    c.previousWorldToView[:] = [float(x) for x in np.asarray(previous_world_to_view, np.float32).reshape(16)]
    c.currentJitterUv[:] = [float(x) for x in current_jitter]
    c.previousJitterUv[:] = [float(x) for x in previous_jitter]
    c.width, c.height = int(width), int(height)
    c.currentTransformCount, c.previousTransformCount = int(current_count), int(previous_count)
    return c


def motion_vectors(gbuffer_ptr, inputs_ptr, current_transforms, previous_transforms, constants, stream=None):
    """MotionVectorsCS over device pointers; transforms are device buffers
    (DeviceBuffer or pointer) of float32 rows, 16 per instance."""
    def p(x):
        return x.ptr if isinstance(x, DeviceBuffer) else x
    _lib.check(_lib.load().rex_motion_vectors_v2(gbuffer_ptr, inputs_ptr, p(current_transforms),
                                                 p(previous_transforms), ctypes.byref(constants), stream))


def compose(width, height, mode, output_ptr, gbuffer=None, raw_signal=None, terminal=None, materials=None,
            material_count=0, film=None, film_sample_count=0, stream=None):
    """ComposeCS into a float4 output buffer (device pointers)."""
    d = _lib.ComposeDesc()
    d.width, d.height, d.outputMode, d.filmSampleCount = int(width), int(height), int(mode), int(film_sample_count)
    d.gbuffer, d.rawSignal, d.terminal, d.materials = gbuffer, raw_signal, terminal, materials
    d.materialCount = int(material_count)
    d.film, d.output = film, output_ptr
    _lib.check(_lib.load().rex_compose_v2(ctypes.byref(d), stream))
