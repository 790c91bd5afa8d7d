"""Alpha-masked geometry on the GPU against the oracle's restatement of the
reference's alpha test (pbr_software_traversal.hlsli:207-248,
software_trace_v1.hlsl:127-170): every walk mode, the batch tracer (with its
alpha-tested hit flag), the device LBVH build and a refit, and a wavefront
frame, bit-exact with the oracle walking the same triangles with the alpha
outcomes resolved; skipped triangles give the same hits as a scene without
them."""
import numpy as np
import pytest

from renderingengine_amd import _lib, scenes, traversal
from renderingengine_amd.render import WALK_WIDE_COMPRESSED, WavefrontRuntime

from test_configs_gpu import _compare, _threads
from test_oracle_alpha import _masked_soup
from test_trace_gpu import ORACLE_ORDER, assert_walk_equal, run_v1, run_walk

pytestmark = pytest.mark.gpu

MODES = [_lib.REX_WALK_SHORT_STACK, _lib.REX_WALK_STACKLESS, _lib.REX_WALK_NEAR_FIRST, _lib.REX_WALK_WIDE,
         _lib.REX_WALK_WIDE_COMPRESSED]


def _oracle_order(mode):
    return 2 if mode == _lib.REX_WALK_WIDE_COMPRESSED else ORACLE_ORDER[mode]


def _backend(tris, mats, build_method=_lib.REX_BUILD_BINNED_SAH):
    be = traversal.FlattenedSahBackend()
    st = be.build_or_update_scene(tris, 0xA1FA, 1, alpha=mats, build_method=build_method)
    assert st, st.message
    return be


def _oracle_on(oracle, be, resolved):
    """The oracle walking the scene's own hierarchy (an LBVH, or refit
    boxes) over the resolved triangles."""
    nodes, order = be.scene.export_flat()
    ob = oracle.OracleBvh(resolved[:2])
    ob.nodes, ob.order = nodes, order
    ob.tris = np.ascontiguousarray(resolved[order])
    return ob


@pytest.mark.parametrize("build_method", [_lib.REX_BUILD_BINNED_SAH, _lib.REX_BUILD_KARRAS_LBVH])
def test_walks_resolve_alpha_like_the_oracle(oracle, build_method):
    tris, mats = _masked_soup(oracle, 20000, seed=31)
    resolved, outcome = oracle.alpha_resolve(tris, mats)
    assert (outcome == 1).sum() > 4000 and (outcome == 0).sum() > 10000
    leaf = 1 if build_method == _lib.REX_BUILD_KARRAS_LBVH else 4
    be = _backend(tris, mats, build_method)
    if build_method == _lib.REX_BUILD_KARRAS_LBVH:
        ob = _oracle_on(oracle, be, resolved)
    else:
        ob = oracle.OracleBvh(resolved, leaf_size=leaf)
    rays = scenes.rays_toward(tris, 20000, seed=32, extent=12.0)
    for mode in MODES:
        for shadow in (False, True):
            exp = ob.walk(rays, shadow=shadow, order=_oracle_order(mode))
            got, _ = run_walk(be, rays, shadow=shadow, mode=mode)
            assert_walk_equal(got, exp, shadow=shadow)
    # the same hits as a scene without the skipped triangles
    kept = tris[outcome != 1].copy()
    kept[:, 40] &= ~np.uint32(scenes.GEOMETRY_FLAG_ALPHA_MASK)
    kept[:, 41] &= ~np.uint32(scenes.MATERIAL_FLAG_ALPHA_MASK)
    be2 = traversal.FlattenedSahBackend()
    assert be2.build_or_update_scene(kept, 0xA1FB, 1)
    a, _ = run_walk(be, rays, mode=_lib.REX_WALK_WIDE_COMPRESSED)
    b, _ = run_walk(be2, rays, mode=_lib.REX_WALK_WIDE_COMPRESSED)
    assert np.array_equal(a["result"], b["result"])
    h = a["result"] == 1
    assert h.sum() > 5000
    for f in ("t", "instanceId", "primitiveId", "materialIndex"):
        assert np.array_equal(a[f][h], b[f][h]), f


def test_batch_tracer_alpha_and_hit_flag(oracle):
    tris, mats = _masked_soup(oracle, 20000, seed=33)
    resolved, _ = oracle.alpha_resolve(tris, mats)
    ob = oracle.OracleBvh(resolved)
    be = _backend(tris, mats)
    rays_v1 = oracle.rays_v1_from(scenes.rays_toward(tris, 20000, seed=34, extent=12.0))
    for any_hit in (False, True):
        exp = ob.batch_v1(rays_v1, 1 if any_hit else 0, 64)
        got, _ = run_v1(be, rays_v1, any_hit=any_hit)
        mism = np.nonzero((got != exp).any(axis=1))[0]
        assert mism.size == 0, f"any={any_hit}: {mism.size} mismatching records, first {mism[:5]}"
        if not any_hit:
            hits = got[:, 17] == 1
            assert ((got[hits, 18] & 2) != 0).sum() > 1000  # accepted masked winners are flagged


def test_refit_keeps_alpha_outcomes(oracle):
    tris, mats = _masked_soup(oracle, 5000, seed=35)
    be = _backend(tris, mats)
    moved = tris.copy()
    pos = moved[:, 0:12].view(np.float32).reshape(-1, 3, 4)
    pos[:, :, 0:3] += np.float32(0.01)
    lib = _lib.load()
    _lib.check(lib.rex_scene_refit(be.scene.handle, np.ascontiguousarray(moved).ctypes.data, moved.shape[0], 2))
    be.scene.generation = 2
    resolved, _ = oracle.alpha_resolve(moved, mats)
    ob = _oracle_on(oracle, be, resolved)
    rays = scenes.rays_toward(moved, 8000, seed=36, extent=12.0)
    for mode in (_lib.REX_WALK_SHORT_STACK, _lib.REX_WALK_WIDE_COMPRESSED):
        exp = ob.walk(rays, order=_oracle_order(mode))
        got, _ = run_walk(be, rays, mode=mode)
        assert_walk_equal(got, exp)


def test_frame_with_alpha_masked_geometry(oracle):
    """A soup frame (1 spp, 4 bounces, MIS) whose triangles are partly
    masked: every per-sample signal within 1e-4 of the oracle integrator
    walking the resolved triangles."""
    w, h = 192, 108
    sc = scenes.soup_scene(n=30000, seed=5, quads=4, palette=16, width=w, height=h)
    tris = sc.tris.copy()
    rng = np.random.default_rng(6)
    mats_v0 = np.stack([np.zeros(32, np.uint32) for _ in range(int(tris[:, 39].max()) + 1)])
    f = mats_v0.view(np.float32)
    f[:, 3] = rng.uniform(0.0, 1.0, mats_v0.shape[0]).astype(np.float32)  # alpha
    f[:, 11] = np.float32(0.5)                                             # cutoff
    mats_v0[:, 20:28] = 0xFFFFFFFF
    masked = rng.random(tris.shape[0]) < 0.4
    tris[masked, 40] |= np.uint32(scenes.GEOMETRY_FLAG_ALPHA_MASK)
    resolved, outcome = oracle.alpha_resolve(tris, mats_v0)
    assert (outcome == 1).sum() > 1000
    rt = WavefrontRuntime(w * h, 1)
    rt.set_scene(tris, sc.materials, sc.lights, light_selection=_lib.REX_LIGHT_SELECTION_POWER, alpha=mats_v0)
    rt.update_frame(sc.camera, w, h, sample_index=0, spp=1, max_bounce=4, estimator=_lib.REX_ESTIMATOR_MIS,
                    seed=0, walk_mode=WALK_WIDE_COMPRESSED, keep_signals=True)
    rt.record_frame()
    stats = rt.collect_completed_frame()
    orr = oracle.OracleRender(resolved, sc.materials, sc.lights, selection=_lib.REX_LIGHT_SELECTION_POWER)
    fr = orr.frame(sc.camera, w, h, spp=1, max_bounce=4, estimator=_lib.REX_ESTIMATOR_MIS, seed=0)
    res = orr.render(fr, order=2, threads=_threads())
    _compare(rt.signals(), res, oracle, "alpha-masked soup 192x108")
    assert stats["closestRays"] == int(res["bounces"].sum())
    assert stats["shadowRays"] == int(res["shadowRays"].sum())
    rt.destroy()
