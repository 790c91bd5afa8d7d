"""rex_scene_refit on the GPU: after the triangles move, every walk (the
reference's left-first, near-first, exact and compressed 4-wide) over the
refit scene returns the oracle's hits over the same refit hierarchy, bit for
bit (the compressed walk against the oracle's exact 4-wide walk), for
closest and any-hit queries; the scene's exported nodes are the host refit's
(and so the reference RefitBottomUp's, tests/test_refit.py)."""
import ctypes

import numpy as np
import pytest

from renderingengine_amd import _lib, scenes, traversal
from test_refit import _build, _moved, _refit
from test_trace_gpu import assert_walk_equal, run_walk

pytestmark = pytest.mark.gpu


def test_refit_walks_match_oracle(oracle):
    tris = scenes.triangle_soup(20000, seed=4)
    be = traversal.FlattenedSahBackend()
    assert be.build_or_update_scene(tris, fingerprint=0x77, generation=1)
    nodes, order = _build(tris)
    lib = _lib.load()
    ob = oracle.OracleBvh(tris)
    assert np.array_equal(ob.nodes, nodes) and np.array_equal(ob.order, order)
    # The compressed walk's traversal hierarchy (one triangle per leaf) is
    # refit on the device with the reference one and keeps its node count.
    compressed = be.scene.info().compressedNodeCount
    assert compressed > len(nodes) // 6
    for step in (1, 2):
        moved = _moved(tris, 10 + step)
        _lib.check(lib.rex_scene_refit(be.scene.handle, moved.ctypes.data, moved.shape[0], 1 + step))
        assert be.scene.info().compressedNodeCount == compressed
        fn, fo = be.scene.export_flat()
        refit = _refit(nodes, moved, order)
        assert np.array_equal(fn.reshape(-1, 12), refit) and np.array_equal(fo, order)
        ob.nodes = refit
        ob.tris = np.ascontiguousarray(moved[order])
        rays = scenes.rays_toward(moved, 3000, seed=step)
        for mode, oorder in ((_lib.REX_WALK_SHORT_STACK, 0), (_lib.REX_WALK_NEAR_FIRST, 1), (_lib.REX_WALK_WIDE, 2),
                             (_lib.REX_WALK_WIDE_COMPRESSED, 2)):
            print(f"refit step {step} walk mode {mode}", flush=True)
            for shadow in (False, True):
                got, _ = run_walk(be, rays, shadow=shadow, mode=mode)
                exp = ob.walk(rays, shadow=shadow, order=oorder)
                assert_walk_equal(got, exp, shadow=shadow)
        hits, _ = run_walk(be, rays, mode=_lib.REX_WALK_WIDE_COMPRESSED)
        assert np.count_nonzero(hits["result"] == 1) > 1000
    # Identity and count checks.
    bad = tris.copy()
    bad[0, 42] += 12345
    assert lib.rex_scene_refit(be.scene.handle, bad.ctypes.data, bad.shape[0], 9) == 1
    assert lib.rex_scene_refit(be.scene.handle, tris.ctypes.data, tris.shape[0] - 1, 9) == 1
    assert lib.rex_scene_refit(be.scene.handle, tris.ctypes.data, tris.shape[0], 0) == 1


def test_lbvh_scene_walks_match_oracle(oracle):
    """A scene built with REX_BUILD_KARRAS_LBVH: its exported nodes are the
    host LBVH's (the reference's BuildKarrasLbvh, tests/test_refit.py) and
    every walk returns the oracle's hits over the same hierarchy; then a
    refit keeps both true."""
    from test_refit import _lbvh
    tris = scenes.triangle_soup(20000, seed=6)
    be = traversal.FlattenedSahBackend()
    be.scene = traversal.Scene(tris, fingerprint=0x78, generation=1, build_method=_lib.REX_BUILD_KARRAS_LBVH)
    nodes, order, _ = _lbvh(tris)
    fn, fo = be.scene.export_flat()
    assert np.array_equal(fn, nodes) and np.array_equal(fo, order)
    ob = oracle.OracleBvh(tris)
    lib = _lib.load()
    for step in (0, 1):
        cur = tris if step == 0 else _moved(tris, 21)
        if step:
            _lib.check(lib.rex_scene_refit(be.scene.handle, cur.ctypes.data, cur.shape[0], 2))
            nodes = _refit(nodes, cur, order)
            assert np.array_equal(be.scene.export_flat()[0], nodes)
        ob.nodes, ob.tris = nodes, np.ascontiguousarray(cur[order])
        rays = scenes.rays_toward(cur, 3000, seed=7 + step)
        for mode, oorder in ((_lib.REX_WALK_SHORT_STACK, 0), (_lib.REX_WALK_NEAR_FIRST, 1), (_lib.REX_WALK_WIDE, 2),
                             (_lib.REX_WALK_WIDE_COMPRESSED, 2)):
            print(f"lbvh step {step} walk mode {mode}", flush=True)
            for shadow in (False, True):
                got, _ = run_walk(be, rays, shadow=shadow, mode=mode)
                exp = ob.walk(rays, shadow=shadow, order=oorder)
                assert_walk_equal(got, exp, shadow=shadow)


@pytest.mark.parametrize("n", [1, 2, 777, 30_000, 1_000_000])
def test_device_lbvh_equals_host_lbvh(n):
    """The device LBVH build (Morton codes, radix sort, Karras splits,
    pre-order numbering, bottom-up boxes) equals the host BuildKarrasLbvh
    restatement -- pinned node for node against the reference's compiled
    build in tests/test_refit.py -- node for node, with the same leaf order."""
    from test_refit import _lbvh
    tris = scenes.triangle_soup(n, seed=31 + n % 7)
    sc = traversal.Scene(tris, fingerprint=0x79, generation=1, build_method=_lib.REX_BUILD_KARRAS_LBVH)
    nodes, order, depth = _lbvh(tris)
    fn, fo = sc.export_flat()
    info = sc.info()
    print(f"LBVH n={n}: {info.nodeCount} nodes, depth {info.maximumDepth}, device {info.deviceMilliseconds:.3f} ms, "
          f"rex_scene_create {info.buildMilliseconds:.1f} ms", flush=True)
    assert np.array_equal(fo, order)
    assert np.array_equal(fn.reshape(-1, 12), nodes.reshape(-1, 12))
    assert info.maximumDepth == depth
    if n >= 30_000:
        moved = _moved(tris, 5)
        lib = _lib.load()
        _lib.check(lib.rex_scene_refit(sc.handle, moved.ctypes.data, moved.shape[0], 2))
        print(f"  refit: device {sc.info().deviceMilliseconds:.3f} ms", flush=True)
        assert np.array_equal(sc.export_flat()[0].reshape(-1, 12), _refit(nodes, moved, order))
    sc.destroy()

