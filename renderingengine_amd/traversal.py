"""Host-side mirror of the reference's traversal backend interface.

``FlattenedSahBackend`` follows ``Rt::Gpu::IGpuTraversalBackend``
(include/rt/gpu/IGpuTraversalBackend.hpp:72-84): ``descriptor()``,
``build_or_update_scene()``, ``record_trace_closest_batch()`` and
``record_trace_any_batch()`` return a ``TraversalStatus`` instead of raising,
with the reference's status codes (:15-33).  Batches name the scene by
fingerprint + generation and use record-index offsets (:56-70).  Every call
goes through librex.so's C ABI; buffers are device memory.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np

from . import _lib

SUCCESS, INVALID_ARGUMENT, UNSUPPORTED, MISSING_SCENE, RECORDING_FAILED, DEVICE_ERROR = range(6)

WALK_HIT_DTYPE = np.dtype([
    ("result", np.uint32), ("t", np.float32), ("u", np.float32), ("v", np.float32),
    ("frontFace", np.uint32), ("triangle", np.uint32), ("position", np.float32, 3),
    ("geometricNormal", np.float32, 3), ("shadingNormal", np.float32, 3),
    ("instanceId", np.uint32), ("primitiveId", np.uint32), ("materialIndex", np.uint32)])


@dataclass
class TraversalStatus:
    code: int = SUCCESS
    message: str = ""

    def __bool__(self):
        return self.code == SUCCESS


@dataclass(frozen=True)
class BackendDescriptor:
    stable_token: str = "software-flattened-sah-hip"
    display_name: str = "Flattened SAH BVH (HIP, gfx950)"
    supports_closest: bool = True
    supports_any: bool = True
    requires_acceleration_structure: bool = False


class DeviceBuffer:
    """A device allocation owned by the caller (the reference's callers own
    their Vulkan buffers the same way)."""

    def __init__(self, nbytes):
        lib = _lib.load()
        ptr = ctypes.c_void_p()
        _lib.check(lib.rex_device_alloc(max(int(nbytes), 1), ctypes.byref(ptr)))
        self.ptr = ptr.value
        self.nbytes = int(nbytes)

    @classmethod
    def from_array(cls, arr):
        arr = np.ascontiguousarray(arr)
        buf = cls(arr.nbytes)
        buf.upload(arr)
        return buf

    def upload(self, arr):
        arr = np.ascontiguousarray(arr)
        assert arr.nbytes <= self.nbytes
        _lib.check(_lib.load().rex_copy_to_device(self.ptr, arr.ctypes.data, arr.nbytes))

    def download(self, dtype, shape):
        out = np.empty(shape, dtype=dtype)
        assert out.nbytes <= self.nbytes
        _lib.check(_lib.load().rex_copy_to_host(out.ctypes.data, self.ptr, out.nbytes))
        return out

    def zero(self):
        self.upload(np.zeros(self.nbytes, dtype=np.uint8))

    def free(self):
        if self.ptr:
            _lib.load().rex_device_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Scene:
    """One uploaded scene (BVH + triangle streams resident in HBM)."""

    def __init__(self, triangles, fingerprint=1, generation=1, max_leaf_size=None, build_threads=0,
                 build_method=_lib.REX_BUILD_BINNED_SAH, alpha=None):
        """max_leaf_size: None selects 4 for the SAH build and 1 for the LBVH
        (one triangle per leaf).  alpha: None, or (materials_v0 (k, 32) uint32
        ABI-v0 records[, layer_alpha[, sampler_id]]) for alpha-masked
        geometry (rex_alpha_desc; the renderers' atlas by default)."""
        if max_leaf_size is None:
            max_leaf_size = 1 if build_method == _lib.REX_BUILD_KARRAS_LBVH else 4
        lib = _lib.load()
        tris = np.ascontiguousarray(triangles, dtype=np.uint32)
        assert tris.ndim == 2 and tris.shape[1] == 44, "expects (n, 44) canonical triangle records"
        desc = _lib.SceneDesc(fingerprint, generation, max_leaf_size, build_threads, build_method)
        if alpha is not None:
            ad, self._alpha_keep = _lib.alpha_desc(*((alpha,) if isinstance(alpha, np.ndarray) else alpha))
            self._alpha_desc = ad
            desc.alpha = ctypes.addressof(ad)
        handle = ctypes.c_void_p()
        _lib.check(lib.rex_scene_create(tris.ctypes.data, tris.shape[0], ctypes.byref(desc),
                                        ctypes.byref(handle)))
        self.handle = handle.value
        self.fingerprint = fingerprint
        self.generation = generation

    def info(self):
        info = _lib.SceneInfo()
        _lib.check(_lib.load().rex_scene_get_info(self.handle, ctypes.byref(info)))
        return info

    def export_flat(self):
        info = self.info()
        nodes = np.zeros((info.nodeCount, 12), dtype=np.uint32)
        order = np.zeros(info.triangleCount, dtype=np.uint32)
        _lib.check(_lib.load().rex_scene_export_flat(self.handle, nodes.ctypes.data, order.ctypes.data))
        return nodes, order

    def destroy(self):
        if self.handle:
            _lib.load().rex_scene_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


def build_host(triangles, max_leaf_size=4, build_threads=0):
    """Host-only SAH build + flatten: (nodes (m, 12) uint32, order, depth)."""
    lib = _lib.load()
    tris = np.ascontiguousarray(triangles, dtype=np.uint32)
    n = tris.shape[0]
    nodes = np.zeros((max(2 * n - 1, 1), 12), dtype=np.uint32)
    order = np.zeros(max(n, 1), dtype=np.uint32)
    count = ctypes.c_uint32()
    depth = ctypes.c_uint32()
    _lib.check(lib.rex_bvh_build_host(tris.ctypes.data, n, max_leaf_size, build_threads, nodes.ctypes.data,
                                      order.ctypes.data, ctypes.byref(count), ctypes.byref(depth)))
    return nodes[:count.value].copy(), order[:n].copy(), depth.value


class FlattenedSahBackend:
    """IGpuTraversalBackend over librex (software flattened SAH on gfx950)."""

    def __init__(self):
        self.scene = None

    @staticmethod
    def descriptor():
        return BackendDescriptor()

    def build_or_update_scene(self, triangles, fingerprint, generation, max_leaf_size=4, alpha=None,
                              build_method=_lib.REX_BUILD_BINNED_SAH):
        """alpha: ABI-v0 materials for alpha-masked geometry (see Scene)."""
        try:
            if self.scene is not None:
                self.scene.destroy()
            self.scene = Scene(triangles, fingerprint, generation,
                               1 if build_method == _lib.REX_BUILD_KARRAS_LBVH else max_leaf_size,
                               build_method=build_method, alpha=alpha)
            return TraversalStatus()
        except _lib.RexError as e:
            self.scene = None
            return TraversalStatus(e.status, str(e))

    def _record(self, any_hit, rays, hits, ray_count, ray_offset=0, hit_offset=0, fingerprint=None,
                generation=None, stack_capacity=64, counters=None, stream=None):
        if self.scene is None:
            return TraversalStatus(MISSING_SCENE, "no scene was built")
        batch = _lib.TraceBatch(
            self.scene.fingerprint if fingerprint is None else fingerprint,
            self.scene.generation if generation is None else generation,
            ray_offset, hit_offset, ray_count, stack_capacity, 0,
            rays.ptr, hits.ptr, counters.ptr if counters is not None else None)
        lib = _lib.load()
        fn = lib.rex_trace_any if any_hit else lib.rex_trace_closest
        st = fn(self.scene.handle, ctypes.byref(batch), stream)
        if st != SUCCESS:
            return TraversalStatus(st, lib.rex_last_error().decode())
        return TraversalStatus()

    def record_trace_closest_batch(self, rays, hits, ray_count, **kw):
        return self._record(False, rays, hits, ray_count, **kw)

    def record_trace_any_batch(self, rays, hits, ray_count, **kw):
        return self._record(True, rays, hits, ray_count, **kw)

    def trace_walk(self, rays, keys, count, shadow, hits, mode=_lib.REX_WALK_SHORT_STACK, counters=None,
                   stream=None):
        """The integrator's traversal seam (WfTraceClosest / WfTraceAny)."""
        lib = _lib.load()
        _lib.check(lib.rex_trace_walk(self.scene.handle, rays.ptr, keys.ptr if keys is not None else None,
                                      count, 1 if shadow else 0, mode, hits.ptr,
                                      counters.ptr if counters is not None else None, stream))
