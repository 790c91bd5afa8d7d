"""ctypes binding of librex.so (include/rex.h).

The library is built in-tree by ``renderingengine_amd.build``.  Loading fails
loudly when it is missing: there is no CPU fallback for any entry point.
"""
from __future__ import annotations

import ctypes
import os

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "librex.so")

REX_SUCCESS = 0
STATUS_NAMES = {0: "SUCCESS", 1: "INVALID_ARGUMENT", 2: "UNSUPPORTED", 3: "MISSING_SCENE",
                4: "RECORDING_FAILED", 5: "DEVICE_ERROR"}

REX_BUILD_BINNED_SAH = 0
REX_BUILD_KARRAS_LBVH = 1
REX_WALK_SHORT_STACK = 0
REX_WALK_STACKLESS = 1
REX_WALK_NEAR_FIRST = 2
REX_WALK_WIDE = 3
REX_WALK_WIDE_COMPRESSED = 4
REX_QUEUE_ATOMIC_APPEND = 0
REX_QUEUE_PREFIX_SCAN = 1
REX_COMPACT_FATAL_SCAN_OVERFLOW = 1
REX_COMPACT_FATAL_NEXT_OVERFLOW = 2
REX_COMPACT_FATAL_SHADOW_OVERFLOW = 4
REX_COUNTER_COUNT = 10
REX_LIGHT_SELECTION_UNIFORM = 0
REX_LIGHT_SELECTION_POWER = 1
REX_ENVIRONMENT_SAMPLER_UNIFORM_SPHERE, REX_ENVIRONMENT_SAMPLER_IMPORTANCE_MAP = 0, 1
REX_ESTIMATOR_BSDF_ONLY = 0
REX_ESTIMATOR_NEE = 1
REX_ESTIMATOR_MIS = 2
REX_ESTIMATOR_RESTIR = 3
REX_JOB_FRAME = 0
REX_JOB_RESTIR_SEQUENCE = 1
REX_RESTIR_NEIGHBORS_RING = 0
REX_RESTIR_NEIGHBORS_DISK = 1
REX_RESTIR_ESTIMATOR_BIASED = 0
REX_RESTIR_ESTIMATOR_REFERENCE_CORRECTION = 1
REX_PROFILE_COUNTERS = 1
REX_PROFILE_TIMING = 2
REX_PRIMARY_SURFACE_VALID = 1
REX_PRIMARY_SURFACE_FRONT_FACE = 2
REX_PRIMARY_SURFACE_HAS_DIFFUSE = 4
REX_PRIMARY_SURFACE_HAS_SPECULAR = 8
REX_MOTION_VALID = 1
REX_COMPOSE_RAW = 0
REX_COMPOSE_MOTION = 4
REX_COMPOSE_BASE_COLOR = 10
REX_COMPOSE_NORMAL = 11
REX_COMPOSE_ROUGHNESS = 12
REX_COMPOSE_METALLIC = 13
REX_COMPOSE_EMISSIVE = 14
REX_COMPOSE_PROGRESSIVE_MEAN = 15

c_u32 = ctypes.c_uint32
c_u64 = ctypes.c_uint64
c_vp = ctypes.c_void_p


class RexError(RuntimeError):
    def __init__(self, status, message):
        super().__init__(f"rex status {STATUS_NAMES.get(status, status)}: {message}")
        self.status = status


class AlphaDesc(ctypes.Structure):
    _fields_ = [("materials", c_vp), ("materialCount", c_u32), ("atlasLayerCount", c_u32),
                ("layerAlpha", c_vp), ("samplerId", c_u32), ("reserved", c_u32)]


class SceneDesc(ctypes.Structure):
    _fields_ = [("fingerprint", c_u64), ("generation", c_u32), ("maxLeafSize", c_u32),
                ("buildThreads", c_u32), ("buildMethod", c_u32), ("alpha", c_vp)]


REX_ALPHA_ATLAS_RENDERER_LAYERS = 2
REX_ALPHA_ATLAS_RENDERER_SAMPLER = 0
RENDERER_ALPHA_ATLAS = (0.0, 1.0)  # layer alphas of the renderers' 1x1 two-layer atlas


def alpha_desc(materials_v0, layer_alpha=RENDERER_ALPHA_ATLAS, sampler_id=REX_ALPHA_ATLAS_RENDERER_SAMPLER):
    """rex_alpha_desc over (k, 32) ABI-v0 material records (rex_material_v0)
    and an atlas of uniform layers; returns (desc, arrays to keep alive)."""
    import numpy as np
    mats = np.ascontiguousarray(materials_v0, dtype=np.uint32).reshape(-1, 32)
    layers = np.ascontiguousarray(layer_alpha, dtype=np.float32).reshape(-1)
    d = AlphaDesc(mats.ctypes.data if mats.shape[0] else None, mats.shape[0], layers.shape[0],
                  layers.ctypes.data if layers.shape[0] else None, sampler_id, 0)
    return d, (mats, layers)


class SceneInfo(ctypes.Structure):
    _fields_ = [("fingerprint", c_u64), ("generation", c_u32), ("triangleCount", c_u32),
                ("nodeCount", c_u32), ("wideNodeCount", c_u32), ("maximumDepth", c_u32),
                ("maxLeafSize", c_u32), ("compressedNodeCount", c_u32), ("buildMilliseconds", ctypes.c_double),
                ("deviceBytes", c_u64), ("deviceMilliseconds", ctypes.c_double)]


class TraceBatch(ctypes.Structure):
    _fields_ = [("sceneFingerprint", c_u64), ("sceneGeneration", c_u32), ("rayOffset", c_u32),
                ("hitOffset", c_u32), ("rayCount", c_u32), ("stackCapacity", c_u32),
                ("reserved", c_u32), ("rays", c_vp), ("hits", c_vp), ("counters", c_vp)]


c_f32 = ctypes.c_float


class ShadingDesc(ctypes.Structure):
    _fields_ = [("materials", c_vp), ("materialCount", c_u32), ("lightSelection", c_u32),
                ("lights", c_vp), ("lightCount", c_u32), ("environmentSampler", c_u32),
                ("environmentTexels", c_vp), ("environmentWidth", c_u32), ("environmentHeight", c_u32)]


def shading_desc(materials, lights, light_selection=1, environment=None, env_sampler=0):
    """(ShadingDesc, arrays it points into) for rex_scene_set_shading;
    environment: (height, width, 4) float32 lat-long map or None."""
    import numpy as np
    mats = np.ascontiguousarray(materials, dtype=np.uint32).reshape(-1, 32)
    lts = np.ascontiguousarray(lights, dtype=np.uint32).reshape(-1, 24)
    keep = [mats, lts]
    d = ShadingDesc(mats.ctypes.data, mats.shape[0], int(light_selection), lts.ctypes.data, lts.shape[0],
                    int(env_sampler), None, 0, 0)
    if environment is not None:
        env = np.ascontiguousarray(environment, dtype=np.float32)
        keep.append(env)
        d.environmentTexels, d.environmentWidth, d.environmentHeight = env.ctypes.data, env.shape[1], env.shape[0]
    return d, keep


class Camera(ctypes.Structure):
    _fields_ = [("position", c_f32 * 3), ("tanHalfFov", c_f32), ("forward", c_f32 * 3),
                ("aspect", c_f32), ("right", c_f32 * 3), ("up", c_f32 * 3)]


class FrameDesc(ctypes.Structure):
    _fields_ = [("camera", Camera), ("width", c_u32), ("height", c_u32), ("spp", c_u32),
                ("maxBounce", c_u32), ("sampleIndex", c_u32), ("estimator", c_u32), ("seed", c_u64),
                ("streamTag", c_u32), ("walkMode", c_u32), ("rrStart", c_f32), ("rrMinimum", c_f32),
                ("rrMaximum", c_f32), ("rayEpsilon", c_f32), ("queueMode", c_u32), ("reserved", c_u32)]


class FrameStats(ctypes.Structure):
    _fields_ = [("closestRays", c_u64), ("shadowRays", c_u64), ("raysPerBounce", c_u32 * 16),
                ("shadowsPerBounce", c_u32 * 16), ("pathCount", c_u32), ("bounces", c_u32),
                ("closestCounters", c_u64 * REX_COUNTER_COUNT), ("shadowCounters", c_u64 * REX_COUNTER_COUNT),
                ("stageMs", c_f32 * 6), ("traceMs", c_f32 * 17), ("traceLaunches", c_u32)]


class BatchDesc(ctypes.Structure):
    _fields_ = [("pathCapacity", c_u32), ("maxJobs", c_u32), ("useGraphs", c_u32), ("shadeBlocksPerCu", c_u32),
                ("traceBlocksPerCu", c_u32), ("outputMiB", c_u32), ("autotune", c_u32), ("reserved", c_u32)]


class BatchJobInfo(ctypes.Structure):
    _fields_ = [("state", c_u32), ("width", c_u32), ("height", c_u32), ("reserved", c_u32),
                ("submitStep", c_u64), ("finishStep", c_u64), ("film", c_vp)]


class BatchStats(ctypes.Structure):
    _fields_ = [("steps", c_u64), ("jobsFinished", c_u64), ("pathsInFlight", c_u64), ("queuedJobs", c_u64),
                ("shadeBlocksPerCu", c_u32), ("traceBlocksPerCu", c_u32), ("graphs", c_u32), ("autotune", c_u32),
                ("outputBytesInUse", c_u64), ("outputBytesTotal", c_u64), ("raysTraced", c_u64),
                ("pathBudget", c_u64)]


class CanonicalScene(ctypes.Structure):
    _fields_ = [("fingerprint", c_u64), ("generation", c_u32), ("reserved", c_u32),
                ("constants", c_vp), ("vertices", c_vp), ("indices", c_vp), ("geometries", c_vp),
                ("instances", c_vp), ("materials", c_vp), ("lights", c_vp),
                ("vertexCount", c_u32), ("indexCount", c_u32), ("geometryCount", c_u32), ("instanceCount", c_u32),
                ("materialCount", c_u32), ("lightCount", c_u32)]


class ServerDesc(ctypes.Structure):
    _fields_ = [("budgetBytes", c_u64), ("pathCapacity", c_u32), ("maxJobs", c_u32), ("useGraphs", c_u32),
                ("shadeBlocksPerCu", c_u32), ("outputMiB", c_u32), ("sequenceSlots", c_u32),
                ("sequenceMaxPixels", c_u32), ("sequenceOutputMiB", c_u32), ("reserved", c_u32 * 2)]


class ServerJobDesc(ctypes.Structure):
    _fields_ = [("kind", c_u32), ("sceneIndex", c_u32), ("frames", c_u32), ("reserved", c_u32),
                ("frame", FrameDesc), ("cameras", c_vp), ("restir", c_vp)]


class ServerJobInfo(ctypes.Structure):
    _fields_ = [("state", c_u32), ("kind", c_u32), ("width", c_u32), ("height", c_u32),
                ("footprintBytes", c_u64), ("submitNs", ctypes.c_int64), ("admitNs", ctypes.c_int64),
                ("finishNs", ctypes.c_int64), ("rays", c_u64), ("film", c_vp)]


class ServerStats(ctypes.Structure):
    _fields_ = [(k, c_u64) for k in ("submitted", "admitted", "finished", "released", "queuedJobs", "residentBytes",
                                     "admittedBytes", "peakAdmittedBytes", "budgetBytes", "engineSteps",
                                     "engineRays", "sequenceRays", "sequenceFrames")]


class RestirDesc(ctypes.Structure):
    _fields_ = [("candidates", c_u32), ("neighbors", c_u32), ("maxM", c_u32), ("maxAge", c_u32),
                ("temporal", c_u32), ("neighborMode", c_u32), ("radius", c_u32), ("sceneGeneration", c_u32),
                ("lightGeneration", c_u32), ("configGeneration", c_u32), ("resourceGeneration", c_u32),
                ("historyGeneration", c_u32), ("validation", c_f32 * 4), ("estimatorMode", c_u32),
                ("objectMotion", c_u32), ("reserved", c_u32 * 2)]


class RestirBuffers(ctypes.Structure):
    _fields_ = [("surfaces", c_vp), ("initial", c_vp), ("temporal", c_vp), ("spatial", c_vp), ("history", c_vp),
                ("direct", c_vp), ("reasons", c_vp), ("statistics", c_vp), ("frameIndex", c_u64),
                ("historyValid", c_u32), ("reserved", c_u32)]


class ReconstructionTargets(ctypes.Structure):
    _fields_ = [("gbuffer", c_vp), ("motionInputs", c_vp), ("rawSignal", c_vp), ("primarySurface", c_vp),
                ("terminal", c_vp), ("pixels", c_u32), ("reserved", c_u32)]


class MotionConstants(ctypes.Structure):
    _fields_ = [("currentViewProjection", c_f32 * 16), ("previousViewProjection", c_f32 * 16),
                ("previousWorldToView", c_f32 * 16), ("currentJitterUv", c_f32 * 2),
                ("previousJitterUv", c_f32 * 2), ("width", c_u32), ("height", c_u32),
                ("currentTransformCount", c_u32), ("previousTransformCount", c_u32)]


class ComposeDesc(ctypes.Structure):
    _fields_ = [("width", c_u32), ("height", c_u32), ("outputMode", c_u32), ("filmSampleCount", c_u32),
                ("gbuffer", c_vp), ("rawSignal", c_vp), ("terminal", c_vp), ("materials", c_vp),
                ("materialCount", c_u32), ("reserved", c_u32), ("film", c_vp), ("output", c_vp)]


# (name, restype, argtypes) for every entry point declared in include/rex.h.
PROTOTYPES = [
    ("rex_last_error", ctypes.c_char_p, []),
    ("rex_abi_version", c_u32, []),
    ("rex_device_alloc", ctypes.c_int, [ctypes.c_size_t, ctypes.POINTER(c_vp)]),
    ("rex_device_free", ctypes.c_int, [c_vp]),
    ("rex_copy_to_device", ctypes.c_int, [c_vp, c_vp, ctypes.c_size_t]),
    ("rex_copy_to_host", ctypes.c_int, [c_vp, c_vp, ctypes.c_size_t]),
    ("rex_synchronize", ctypes.c_int, []),
    ("rex_scene_create", ctypes.c_int, [c_vp, c_u32, ctypes.POINTER(SceneDesc), ctypes.POINTER(c_vp)]),
    ("rex_scene_destroy", ctypes.c_int, [c_vp]),
    ("rex_scene_get_info", ctypes.c_int, [c_vp, ctypes.POINTER(SceneInfo)]),
    ("rex_bvh_build_host", ctypes.c_int, [c_vp, c_u32, c_u32, c_u32, c_vp, c_vp,
                                          ctypes.POINTER(c_u32), ctypes.POINTER(c_u32)]),
    ("rex_scene_export_flat", ctypes.c_int, [c_vp, c_vp, c_vp]),
    ("rex_scene_refit", ctypes.c_int, [c_vp, c_vp, c_u32, c_u32]),
    ("rex_bvh_refit_host", ctypes.c_int, [c_vp, c_u32, c_vp, c_vp, c_u32]),
    ("rex_bvh_build_lbvh_host", ctypes.c_int, [c_vp, c_u32, c_vp, c_vp, ctypes.POINTER(c_u32), ctypes.POINTER(c_u32)]),
    ("rex_trace_closest", ctypes.c_int, [c_vp, ctypes.POINTER(TraceBatch), c_vp]),
    ("rex_trace_any", ctypes.c_int, [c_vp, ctypes.POINTER(TraceBatch), c_vp]),
    ("rex_trace_no_alloc", ctypes.c_int, [c_vp, ctypes.POINTER(TraceBatch), ctypes.c_int, c_vp]),
    ("rex_compact_scratch_bytes", ctypes.c_uint64, [ctypes.c_uint32]),
    ("rex_compact_stable", ctypes.c_int, [c_vp, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, c_vp, c_vp, c_vp,
                                          c_vp, c_vp, c_vp]),
    ("rex_trace_walk", ctypes.c_int, [c_vp, c_vp, c_vp, c_u32, ctypes.c_int, c_u32, c_vp, c_vp, c_vp]),
    ("rex_scene_set_shading", ctypes.c_int, [c_vp, ctypes.POINTER(ShadingDesc)]),
    ("rex_scene_export_light_tables", ctypes.c_int, [c_vp, c_vp, c_vp, ctypes.POINTER(c_u32)]),
    ("rex_scene_export_environment_tables", ctypes.c_int, [c_vp, c_vp, c_vp, ctypes.POINTER(ctypes.c_double)]),
    ("rex_renderer_create", ctypes.c_int, [c_vp, c_u32, c_u32, ctypes.POINTER(c_vp)]),
    ("rex_renderer_destroy", ctypes.c_int, [c_vp]),
    ("rex_renderer_reset_film", ctypes.c_int, [c_vp, c_vp]),
    ("rex_renderer_bind_scene", ctypes.c_int, [c_vp, c_vp, c_vp]),
    ("rex_canonical_expand", ctypes.c_int, [ctypes.POINTER(CanonicalScene), c_vp, c_u32, ctypes.POINTER(c_u32)]),
    ("rex_canonical_shading", ctypes.c_int, [ctypes.POINTER(CanonicalScene), c_vp, c_u32, c_vp, c_vp]),
    ("rex_scene_create_canonical", ctypes.c_int, [ctypes.POINTER(CanonicalScene), c_u32, c_u32,
                                                  ctypes.POINTER(c_vp)]),
    ("rex_scene_refit_canonical", ctypes.c_int, [c_vp, ctypes.POINTER(CanonicalScene), c_u32]),
    ("rex_server_create", ctypes.c_int, [ctypes.POINTER(ServerDesc), ctypes.POINTER(c_vp)]),
    ("rex_server_destroy", ctypes.c_int, [c_vp]),
    ("rex_server_add_scene", ctypes.c_int, [c_vp, c_vp, ctypes.POINTER(c_u32)]),
    ("rex_server_submit", ctypes.c_int, [c_vp, ctypes.POINTER(ServerJobDesc), ctypes.POINTER(c_u64)]),
    ("rex_server_pump", ctypes.c_int, [c_vp, ctypes.POINTER(c_u32)]),
    ("rex_server_drain", ctypes.c_int, [c_vp]),
    ("rex_server_job", ctypes.c_int, [c_vp, c_u64, ctypes.POINTER(ServerJobInfo)]),
    ("rex_server_release", ctypes.c_int, [c_vp, c_u64]),
    ("rex_server_get_stats", ctypes.c_int, [c_vp, ctypes.POINTER(ServerStats)]),
    ("rex_render_frame", ctypes.c_int, [c_vp, ctypes.POINTER(FrameDesc), ctypes.c_int, c_vp]),
    ("rex_renderer_stats", ctypes.c_int, [c_vp, ctypes.POINTER(FrameStats)]),
    ("rex_renderer_buffers", ctypes.c_int, [c_vp, ctypes.POINTER(c_vp), ctypes.POINTER(c_vp)]),
    ("rex_renderer_set_profiling", ctypes.c_int, [c_vp, c_u32]),
    ("rex_renderer_set_restir", ctypes.c_int, [c_vp, ctypes.POINTER(RestirDesc)]),
    ("rex_batch_create", ctypes.c_int, [ctypes.POINTER(BatchDesc), ctypes.POINTER(c_vp)]),
    ("rex_batch_destroy", ctypes.c_int, [c_vp]),
    ("rex_batch_add_scene", ctypes.c_int, [c_vp, c_vp, ctypes.POINTER(c_u32)]),
    ("rex_batch_submit", ctypes.c_int, [c_vp, c_u32, c_vp, ctypes.POINTER(c_u64)]),
    ("rex_batch_step", ctypes.c_int, [c_vp, c_vp]),
    ("rex_batch_run", ctypes.c_int, [c_vp, c_vp]),
    ("rex_batch_job", ctypes.c_int, [c_vp, c_u64, ctypes.POINTER(BatchJobInfo)]),
    ("rex_batch_release", ctypes.c_int, [c_vp, c_u64]),
    ("rex_batch_get_stats", ctypes.c_int, [c_vp, ctypes.POINTER(BatchStats)]),
    ("rex_renderer_reset_restir", ctypes.c_int, [c_vp]),
    ("rex_renderer_restir_buffers", ctypes.c_int, [c_vp, ctypes.POINTER(RestirBuffers)]),
    ("rex_renderer_set_reconstruction_v2", ctypes.c_int, [c_vp, ctypes.POINTER(ReconstructionTargets)]),
    ("rex_camera_matrices", ctypes.c_int, [ctypes.POINTER(Camera), c_vp, c_vp]),
    ("rex_motion_vectors_v2", ctypes.c_int, [c_vp, c_vp, c_vp, c_vp, ctypes.POINTER(MotionConstants), c_vp]),
    ("rex_compose_v2", ctypes.c_int, [ctypes.POINTER(ComposeDesc), c_vp]),
]

_lib = None


def load():
    """Loads librex.so (once); raises if it was not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: run `python -m renderingengine_amd.build` "
                           "(there is no CPU fallback)")
    lib = ctypes.CDLL(LIB_PATH)
    for name, res, args in PROTOTYPES:
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(status):
    if status != REX_SUCCESS:
        msg = load().rex_last_error()
        raise RexError(status, msg.decode() if msg else "")
    return status
