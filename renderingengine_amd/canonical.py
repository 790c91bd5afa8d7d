"""Host mirror of scene ingestion from ABI-v0 canonical scenes
(include/contracts/SceneAbiV0.hpp; Scene::CanonicalSceneView): the records
as uint32 word arrays (floats stored by bit pattern), the expansion into
world-space canonical triangles (BuildCanonicalTraversalScene), the L6
material / light conversion (the host half of Wave2Runtime::SetScene) and
one-call scene creation.  Errors raise ``_lib.RexError``."""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib

CONSTANTS_WORDS, VERTEX_WORDS, GEOMETRY_WORDS = 24, 16, 16
INSTANCE_WORDS, MATERIAL_WORDS, LIGHT_WORDS = 56, 32, 24
TRIANGLE_WORDS = 44

INSTANCE_VISIBLE, INSTANCE_CASTS_SHADOW = 1, 2
GEOMETRY_OPAQUE, GEOMETRY_DOUBLE_SIDED = 1, 4
MATERIAL_MODEL_METALLIC_ROUGHNESS, MATERIAL_MODEL_SMOOTH_DIELECTRIC = 0, 1
MATERIAL_FLAG_DOUBLE_SIDED, MATERIAL_FLAG_EMISSIVE, MATERIAL_FLAG_THIN_WALLED = 1, 4, 8
LIGHT_POINT, LIGHT_DIRECTIONAL, LIGHT_EMISSIVE_TRIANGLE, LIGHT_ENVIRONMENT = 0, 1, 4, 5
LIGHT_ENABLED, LIGHT_TWO_SIDED, LIGHT_DELTA = 1, 2, 8


def _words(a, width):
    a = np.ascontiguousarray(a)
    if a.dtype == np.float32:
        a = a.view(np.uint32)
    return np.ascontiguousarray(a, dtype=np.uint32).reshape(-1, width)


class CanonicalScene:
    """The arrays of one canonical scene (word arrays of the ABI-v0 record
    widths) and its identity."""

    def __init__(self, constants, vertices, indices, geometries, instances, materials, lights,
                 fingerprint=1, generation=1):
        self.constants = _words(constants, CONSTANTS_WORDS)
        self.vertices = _words(vertices, VERTEX_WORDS)
        self.indices = np.ascontiguousarray(indices, dtype=np.uint32).reshape(-1)
        self.geometries = _words(geometries, GEOMETRY_WORDS)
        self.instances = _words(instances, INSTANCE_WORDS)
        self.materials = _words(materials, MATERIAL_WORDS)
        self.lights = _words(lights, LIGHT_WORDS)
        self.fingerprint, self.generation = int(fingerprint), int(generation)

    def view(self):
        v = _lib.CanonicalScene()
        v.fingerprint, v.generation = self.fingerprint, self.generation
        v.constants = self.constants.ctypes.data
        v.vertices, v.indices = self.vertices.ctypes.data, self.indices.ctypes.data
        v.geometries, v.instances = self.geometries.ctypes.data, self.instances.ctypes.data
        v.materials = self.materials.ctypes.data
        v.lights = self.lights.ctypes.data if self.lights.shape[0] else None
        v.vertexCount, v.indexCount = self.vertices.shape[0], self.indices.shape[0]
        v.geometryCount, v.instanceCount = self.geometries.shape[0], self.instances.shape[0]
        v.materialCount, v.lightCount = self.materials.shape[0], self.lights.shape[0]
        return v


def expand(scene):
    """World-space canonical triangles: (n, 44) uint32 (rex_triangle)."""
    lib = _lib.load()
    v = scene.view()
    n = ctypes.c_uint32()
    _lib.check(lib.rex_canonical_expand(ctypes.byref(v), None, 0, ctypes.byref(n)))
    out = np.zeros((n.value, TRIANGLE_WORDS), dtype=np.uint32)
    _lib.check(lib.rex_canonical_expand(ctypes.byref(v), out.ctypes.data, n.value, ctypes.byref(n)))
    return out


def shading(scene, triangles):
    """(materials (m, 32), lights (l, 24)) uint32: the L6 records."""
    lib = _lib.load()
    v = scene.view()
    tris = np.ascontiguousarray(triangles, dtype=np.uint32)
    mats = np.zeros((scene.materials.shape[0], MATERIAL_WORDS), dtype=np.uint32)
    lts = np.zeros((scene.lights.shape[0], LIGHT_WORDS), dtype=np.uint32)
    _lib.check(lib.rex_canonical_shading(ctypes.byref(v), tris.ctypes.data, tris.shape[0], mats.ctypes.data,
                                         lts.ctypes.data))
    return mats, lts


class CanonicalSceneHandle:
    """A rex scene created from a canonical scene (rex_scene_create_canonical),
    usable wherever a ``batch.ShadedSceneHandle`` is (``.handle``)."""

    def __init__(self, scene, max_leaf_size=4, light_selection=_lib.REX_LIGHT_SELECTION_POWER):
        lib = _lib.load()
        v = scene.view()
        h = ctypes.c_void_p()
        _lib.check(lib.rex_scene_create_canonical(ctypes.byref(v), int(max_leaf_size), int(light_selection),
                                                  ctypes.byref(h)))
        self.handle = h.value

    def refit(self, scene, generation):
        """Moves the instances to those of ``scene`` (the same canonical scene
        with new instance transforms): rex_scene_refit_canonical."""
        v = scene.view()
        _lib.check(_lib.load().rex_scene_refit_canonical(self.handle, ctypes.byref(v), int(generation)))

    def destroy(self):
        if self.handle:
            _lib.load().rex_scene_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass
