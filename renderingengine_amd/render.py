"""Host-side mirror of the reference's wavefront runtime.

``WavefrontRuntime`` follows ``Wave2Runtime``
(include/renderers/Wave2Runtime.hpp:85-135) for the Wavefront architecture:
``set_scene()`` uploads the canonical triangles, materials and lights and
builds the light tables (Wave2Runtime.cpp:462-666), ``invalidate_progressive_film()``
restarts the Film mean (Wave2Runtime.hpp:106), ``update_frame()`` fixes the
frame constants (camera, sample index, estimator; Wave2Runtime.cpp:1346-1440)
and ``record_frame()`` launches RayGen, the per-bounce Intersect / Shade /
TraceShadow stages and Resolve + Film on a HIP stream
(Wave2Runtime.cpp:1442-1585).  ``collect_completed_frame()`` returns the
queue counters after the stream has finished (Wave2Runtime.hpp:128-131).

Errors raise ``_lib.RexError`` with the library's status, as the reference's
``Require`` throws; there is no host fallback.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib, restir_abi
from .traversal import Scene

SIGNAL_NAMES = ("cameraEmission", "directDiffuse", "directSpecular", "indirectDiffuse", "indirectSpecular")
ALIAS_DTYPE = np.dtype([("q", np.float32), ("pmf", np.float32), ("alias", np.uint32), ("item", np.uint32)])

WALK_SHORT_STACK = _lib.REX_WALK_SHORT_STACK
WALK_NEAR_FIRST = _lib.REX_WALK_NEAR_FIRST
WALK_WIDE = _lib.REX_WALK_WIDE
WALK_WIDE_COMPRESSED = _lib.REX_WALK_WIDE_COMPRESSED


def _stream_handle(stream):
    """Accepts None (the null stream), an int handle or a torch stream."""
    if stream is None:
        return None
    if isinstance(stream, int):
        return stream
    return int(stream.cuda_stream)


def make_camera(cam):
    c = _lib.Camera()
    c.position[:] = [float(x) for x in cam["position"]]
    c.forward[:] = [float(x) for x in cam["forward"]]
    c.right[:] = [float(x) for x in cam["right"]]
    c.up[:] = [float(x) for x in cam["up"]]
    c.tanHalfFov = float(cam["tanHalfFov"])
    c.aspect = float(cam["aspect"])
    return c


class WavefrontRuntime:
    """One scene's device tables plus the path state, queues and Film of up
    to ``max_pixels * max_spp`` paths per launch."""

    def __init__(self, max_pixels, max_spp=1):
        self.max_pixels = int(max_pixels)
        self.max_spp = int(max_spp)
        self.scene = None
        self._renderer = None
        self._frame = None
        self._keep_signals = False
        self._light_count = 0
        self._lib = _lib.load()

    # -- scene ------------------------------------------------------------
    def set_scene(self, tris, materials, lights, light_selection=_lib.REX_LIGHT_SELECTION_POWER,
                  fingerprint=1, generation=1, max_leaf_size=4, build_threads=0, environment=None, env_sampler=0,
                  alpha=None):
        """environment: (height, width, 4) float32 lat-long map (row 0 at +y)
        or None (constant (1, 1, 1)); env_sampler REX_ENVIRONMENT_SAMPLER_*;
        alpha: ABI-v0 materials for alpha-masked geometry (traversal.Scene)."""
        self._release_renderer()
        if self.scene is not None:
            self.scene.destroy()
        self.scene = Scene(tris, fingerprint=fingerprint, generation=generation, max_leaf_size=max_leaf_size,
                           build_threads=build_threads, alpha=alpha)
        desc, keep = _lib.shading_desc(materials, lights, light_selection, environment, env_sampler)
        _lib.check(self._lib.rex_scene_set_shading(self.scene.handle, ctypes.byref(desc)))
        self._light_count = keep[1].shape[0]
        self._env_shape = (1, 1) if environment is None else tuple(np.shape(environment)[:2])
        r = ctypes.c_void_p()
        _lib.check(self._lib.rex_renderer_create(self.scene.handle, self.max_pixels, self.max_spp, ctypes.byref(r)))
        self._renderer = r.value

    def attach(self, shaded):
        """Renders on an existing shaded scene (``batch.ShadedSceneHandle``)
        instead of uploading one; the scene stays owned by the caller."""
        self._release_renderer()
        self.scene = None
        self._attached = shaded
        r = ctypes.c_void_p()
        _lib.check(self._lib.rex_renderer_create(shaded.handle, self.max_pixels, self.max_spp, ctypes.byref(r)))
        self._renderer = r.value

    def bind(self, shaded, stream=None):
        """Points the existing renderer at another shaded scene (buffers kept,
        Film restarted, ReSTIR history dropped)."""
        self._require()
        _lib.check(self._lib.rex_renderer_bind_scene(self._renderer, shaded.handle, _stream_handle(stream)))
        self._attached = shaded

    def light_tables(self):
        """(alias table, per-light selection pmf) as uploaded."""
        n = ctypes.c_uint32()
        _lib.check(self._lib.rex_scene_export_light_tables(self.scene.handle, None, None, ctypes.byref(n)))
        alias = np.zeros(n.value, dtype=ALIAS_DTYPE)
        pmf = np.zeros(self._light_count, dtype=np.float32)
        _lib.check(self._lib.rex_scene_export_light_tables(self.scene.handle, alias.ctypes.data, pmf.ctypes.data,
                                                           ctypes.byref(n)))
        return alias, pmf

    def environment_tables(self):
        """(rows, columns, integrated luminance) of the environment
        distribution as built (rows: height alias entries, columns:
        height x width; the constant map is 1 x 1)."""
        h, w = self._env_shape
        rows = np.zeros(h, dtype=ALIAS_DTYPE)
        cols = np.zeros(h * w, dtype=ALIAS_DTYPE)
        integ = ctypes.c_double()
        _lib.check(self._lib.rex_scene_export_environment_tables(self.scene.handle, rows.ctypes.data, cols.ctypes.data,
                                                                 ctypes.byref(integ)))
        return rows, cols, integ.value

    def is_ready(self):
        return self._renderer is not None

    # -- frames -----------------------------------------------------------
    def invalidate_progressive_film(self, stream=None):
        self._require()
        _lib.check(self._lib.rex_renderer_reset_film(self._renderer, _stream_handle(stream)))

    def update_frame(self, camera, width, height, sample_index=0, spp=1, max_bounce=4,
                     estimator=_lib.REX_ESTIMATOR_MIS, seed=0, walk_mode=WALK_WIDE_COMPRESSED, stream_tag=0,
                     rr=(0.0, 0.0, 0.0), ray_epsilon=0.0, keep_signals=False, queue_mode=0):
        """queue_mode: REX_QUEUE_ATOMIC_APPEND (0) or REX_QUEUE_PREFIX_SCAN (1)."""
        f = _lib.FrameDesc()
        f.camera = make_camera(camera)
        f.width, f.height, f.spp = int(width), int(height), int(spp)
        f.maxBounce, f.sampleIndex, f.estimator = int(max_bounce), int(sample_index), int(estimator)
        f.seed = int(seed)
        f.streamTag, f.walkMode = int(stream_tag), int(walk_mode)
        f.rrStart, f.rrMinimum, f.rrMaximum = (float(x) for x in rr)
        f.rayEpsilon = float(ray_epsilon)
        f.queueMode = int(queue_mode)
        self._frame = f
        self._keep_signals = bool(keep_signals)
        return f

    def record_frame(self, stream=None):
        self._require()
        if self._frame is None:
            raise _lib.RexError(1, "update_frame() must precede record_frame()")
        _lib.check(self._lib.rex_render_frame(self._renderer, ctypes.byref(self._frame),
                                              1 if self._keep_signals else 0, _stream_handle(stream)))

    def set_profiling(self, counters=False, timing=False):
        """Traversal counters and per-stage GPU times for later frames."""
        self._require()
        flags = (_lib.REX_PROFILE_COUNTERS if counters else 0) | (_lib.REX_PROFILE_TIMING if timing else 0)
        _lib.check(self._lib.rex_renderer_set_profiling(self._renderer, flags))

    def collect_completed_frame(self):
        """Synchronizes the frame's stream; returns the queue counters (and,
        when profiling is on, traversal counters and stage times)."""
        self._require()
        st = _lib.FrameStats()
        _lib.check(self._lib.rex_renderer_stats(self._renderer, ctypes.byref(st)))
        b = st.bounces
        return {"closestRays": st.closestRays, "shadowRays": st.shadowRays, "pathCount": st.pathCount,
                "raysPerBounce": list(st.raysPerBounce[:b]), "shadowsPerBounce": list(st.shadowsPerBounce[:b]),
                "closestCounters": list(st.closestCounters), "shadowCounters": list(st.shadowCounters),
                "stageMs": list(st.stageMs), "traceMs": list(st.traceMs[:st.traceLaunches])}

    # -- ReSTIR DI (ReSTIRDIRuntime, include/renderers/ReSTIRDIRuntime.hpp:205-243) --
    def set_restir(self, candidates=32, neighbors=5, max_m=20, max_age=20, temporal=True,
                   neighbor_mode=_lib.REX_RESTIR_NEIGHBORS_RING, radius=30, generations=None, validation=None,
                   estimator_mode=_lib.REX_RESTIR_ESTIMATOR_BIASED, object_motion=False):
        """Settings of later ReSTIR frames (estimator REX_ESTIMATOR_RESTIR);
        drops the history.  generations: dict of scene / light / config /
        resource / history generation overrides; validation: (normal cosine,
        relative depth, position, thin) or None for the production values;
        estimator_mode: biased or reference correction; object_motion:
        temporal reuse through the instances' current and previous
        object-to-world (canonical scenes)."""
        self._require()
        d = _lib.RestirDesc()
        d.candidates, d.neighbors, d.maxM, d.maxAge = int(candidates), int(neighbors), int(max_m), int(max_age)
        d.temporal, d.neighborMode, d.radius = 1 if temporal else 0, int(neighbor_mode), int(radius)
        g = generations or {}
        d.sceneGeneration = int(g.get("scene", 0))
        d.lightGeneration = int(g.get("light", 0))
        d.configGeneration = int(g.get("config", 0))
        d.resourceGeneration = int(g.get("resource", 0))
        d.historyGeneration = int(g.get("history", 0))
        if validation is not None:
            d.validation[:] = [float(x) for x in validation]
        d.estimatorMode = int(estimator_mode)
        d.objectMotion = 1 if object_motion else 0
        _lib.check(self._lib.rex_renderer_set_restir(self._renderer, ctypes.byref(d)))

    def set_reconstruction_targets(self, targets):
        """Binds a ``reconstruction.ReconstructionTargets`` (None unbinds):
        every later frame publishes its ABI-v2 records into it."""
        self._require()
        t = targets.struct() if targets is not None else None
        _lib.check(self._lib.rex_renderer_set_reconstruction_v2(self._renderer,
                                                                ctypes.byref(t) if t is not None else None))

    def reset_restir(self):
        self._require()
        _lib.check(self._lib.rex_renderer_reset_restir(self._renderer))

    def restir_state(self):
        """Host copies of the last ReSTIR frame's buffers: surfaces, initial /
        temporal / spatial (resolved) / published reservoirs, direct light
        (pixels, 2, 4), rejection reasons, statistics, frame index and
        whether history was reused."""
        self._require()
        b = _lib.RestirBuffers()
        _lib.check(self._lib.rex_renderer_restir_buffers(self._renderer, ctypes.byref(b)))
        n = self._frame.width * self._frame.height
        _lib.check(self._lib.rex_synchronize())

        def fetch(ptr, dtype, shape):
            out = np.zeros(shape, dtype=dtype)
            _lib.check(self._lib.rex_copy_to_host(out.ctypes.data, ptr, out.nbytes))
            return out

        return {"surfaces": fetch(b.surfaces, restir_abi.SURFACE, n),
                "initial": fetch(b.initial, restir_abi.RESERVOIR, n),
                "temporal": fetch(b.temporal, restir_abi.RESERVOIR, n),
                "final": fetch(b.spatial, restir_abi.RESERVOIR, n),
                "published": fetch(b.history, restir_abi.RESERVOIR, n),
                "direct": fetch(b.direct, np.float32, (n, 2, 4)),
                "reasons": fetch(b.reasons, np.uint32, n),
                "stats": fetch(b.statistics, restir_abi.STATISTICS, 1)[0],
                "frameIndex": int(b.frameIndex), "historyValid": bool(b.historyValid)}

    def buffers(self):
        """Device pointers (film, signals)."""
        self._require()
        film, sig = ctypes.c_void_p(), ctypes.c_void_p()
        _lib.check(self._lib.rex_renderer_buffers(self._renderer, ctypes.byref(film), ctypes.byref(sig)))
        return film.value, sig.value

    def film(self):
        """Host copy of the Film: (width * height, 4) float32 (rgb mean, count)."""
        f = self._frame
        n = f.width * f.height
        out = np.zeros((n, 4), dtype=np.float32)
        film, _ = self.buffers()
        _lib.check(self._lib.rex_synchronize())
        _lib.check(self._lib.rex_copy_to_host(out.ctypes.data, film, out.nbytes))
        return out

    def signals(self):
        """Host copy of the last frame's per-sample signals: (paths, 5, 3)."""
        f = self._frame
        n = f.width * f.height * f.spp
        out = np.zeros((n, 5, 4), dtype=np.float32)
        _, sig = self.buffers()
        _lib.check(self._lib.rex_synchronize())
        _lib.check(self._lib.rex_copy_to_host(out.ctypes.data, sig, out.nbytes))
        return out[:, :, :3].copy()

    # -- lifetime ---------------------------------------------------------
    def _require(self):
        if self._renderer is None:
            raise _lib.RexError(3, "no scene: call set_scene() first")

    def _release_renderer(self):
        if self._renderer is not None:
            self._lib.rex_renderer_destroy(self._renderer)
            self._renderer = None

    def destroy(self):
        self._release_renderer()
        if self.scene is not None:
            self.scene.destroy()
            self.scene = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass
