"""The stable compaction on the MI355X (rex_compact_stable and the
PrefixScan queue mode of frames) against the oracle's CompactStable.

Bar: exclusive prefixes and compacted index lists bit-identical to the
oracle on the reference's vectors and on random flags up to a 1080p x 4 spp
path count; totals, active counts and overflow / fatal bits as the
reference's ScanBlocksCS / ScatterCS define them; frames rendered with the
PrefixScan queue mode bit-identical to atomic-append frames and to the
oracle (signals, queue sizes, Film).
"""
import ctypes

import numpy as np
import pytest

from renderingengine_amd import _lib, scenes, traversal
from renderingengine_amd.render import WALK_WIDE_COMPRESSED, WavefrontRuntime

pytestmark = pytest.mark.gpu


def device_compact(flags, cap_next=None, cap_shadow=None):
    lib = _lib.load()
    flags = np.ascontiguousarray(flags, dtype=np.uint32).reshape(-1, 2)
    n = flags.shape[0]
    cap_next = n if cap_next is None else cap_next
    cap_shadow = n if cap_shadow is None else cap_shadow
    d_flags = traversal.DeviceBuffer.from_array(flags) if n else None
    d_prefix = traversal.DeviceBuffer(max(n, 1) * 8)
    d_next = traversal.DeviceBuffer(max(cap_next, 1) * 4)
    d_shadow = traversal.DeviceBuffer(max(cap_shadow, 1) * 4)
    d_status = traversal.DeviceBuffer(5 * 4)
    d_scratch = traversal.DeviceBuffer(int(lib.rex_compact_scratch_bytes(n)))
    st = lib.rex_compact_stable(d_flags.ptr if d_flags else None, n, cap_next, cap_shadow, d_scratch.ptr,
                                d_prefix.ptr, d_next.ptr, d_shadow.ptr, d_status.ptr, None)
    assert st == 0, lib.rex_last_error()
    status = d_status.download(np.uint32, (5,))
    prefix = d_prefix.download(np.uint32, (max(n, 1), 2))[:n]
    nxt = d_next.download(np.uint32, (max(cap_next, 1),))[:min(status[2], cap_next)]
    shd = d_shadow.download(np.uint32, (max(cap_shadow, 1),))[:min(status[3], cap_shadow)]
    return status, prefix, nxt, shd


def _check(oracle, flags):
    rc, prefix, nxt, shd = oracle.compact_stable(flags)
    assert rc == 0
    status, gp, gn, gs = device_compact(flags)
    assert status.tolist() == [nxt.size, shd.size, nxt.size, shd.size, 0]
    assert np.array_equal(gp, prefix)
    assert np.array_equal(gn, nxt) and np.array_equal(gs, shd)


def test_reference_vectors(oracle):
    _check(oracle, np.zeros((0, 2), np.uint32))
    _check(oracle, np.array([[1, 0], [0, 1], [1, 1], [0, 0], [1, 0]], np.uint32))
    for count in (1, 127, 128, 129, 4097):
        idx = np.arange(count, dtype=np.uint32)
        _check(oracle, np.stack([(idx % 3 == 0), (idx % 5 == 0)], axis=1).astype(np.uint32))


@pytest.mark.parametrize("count,density", [(1023, 0.5), (1024, 0.5), (1025, 0.99), (1_048_577, 0.3),
                                           (1920 * 1080 * 4, 0.6), (3_000_000, 0.0), (2_000_000, 1.0)])
def test_random_flags(oracle, count, density):
    rng = np.random.default_rng(count)
    flags = (rng.random((count, 2)) < density).astype(np.uint32)
    _check(oracle, flags)


def test_capacity_overflow_and_non_binary_flags():
    # Totals above the capacities: active counts clamp, indices at or past
    # the capacity are not written, the overflow bits are set.
    flags = np.tile(np.array([[1, 1]], np.uint32), (5000, 1))
    status, prefix, nxt, shd = device_compact(flags, cap_next=3000, cap_shadow=4999)
    assert status[:4].tolist() == [5000, 5000, 3000, 4999]
    assert status[4] == _lib.REX_COMPACT_FATAL_NEXT_OVERFLOW | _lib.REX_COMPACT_FATAL_SHADOW_OVERFLOW
    assert np.array_equal(nxt, np.arange(3000, dtype=np.uint32))
    assert np.array_equal(shd, np.arange(4999, dtype=np.uint32))
    # A flag above 1 is a fatal scan overflow and counts as (0, 0).
    flags = np.array([[1, 0], [2, 1], [1, 1]], np.uint32)
    status, prefix, nxt, shd = device_compact(flags)
    assert status[4] == _lib.REX_COMPACT_FATAL_SCAN_OVERFLOW
    assert status[:2].tolist() == [2, 1] and nxt.tolist() == [0, 2] and shd.tolist() == [2]


def _frame(rt, sc, w, h, spp, bounce, mode, queue_mode, sample_index=0):
    rt.update_frame(sc.camera, w, h, sample_index=sample_index, spp=spp, max_bounce=bounce,
                    estimator=_lib.REX_ESTIMATOR_MIS, seed=13, walk_mode=mode, keep_signals=True,
                    queue_mode=queue_mode)
    rt.record_frame()
    st = rt.collect_completed_frame()
    return st, rt.signals(), rt.film()


@pytest.mark.parametrize("name", ["cornell", "soup"])
def test_prefix_scan_frames_equal_append_frames_and_oracle(oracle, name):
    if name == "cornell":
        sc, w, h, spp, bounce = scenes.cornell_box(64, 36), 64, 36, 2, 5
    else:
        sc, w, h, spp, bounce = scenes.soup_scene(n=20000, seed=4, quads=6, palette=32, width=96, height=54), 96, 54, 2, 8
    orr = oracle.OracleRender(sc.tris, sc.materials, sc.lights, selection=1)
    rt_a = WavefrontRuntime(w * h, spp)
    rt_a.set_scene(sc.tris, sc.materials, sc.lights, light_selection=1)
    rt_s = WavefrontRuntime(w * h, spp)
    rt_s.set_scene(sc.tris, sc.materials, sc.lights, light_selection=1)
    for k in range(2):
        sa, siga, filma = _frame(rt_a, sc, w, h, spp, bounce, WALK_WIDE_COMPRESSED, _lib.REX_QUEUE_ATOMIC_APPEND,
                                 k * spp)
        ss, sigs, films = _frame(rt_s, sc, w, h, spp, bounce, WALK_WIDE_COMPRESSED, _lib.REX_QUEUE_PREFIX_SCAN,
                                 k * spp)
        assert sigs.tobytes() == siga.tobytes()
        assert films.tobytes() == filma.tobytes()
        assert ss["raysPerBounce"] == sa["raysPerBounce"] and ss["shadowsPerBounce"] == sa["shadowsPerBounce"]
        fr = orr.frame(sc.camera, w, h, spp=spp, max_bounce=bounce, sample_index=k * spp,
                       estimator=_lib.REX_ESTIMATOR_MIS, seed=13)
        res = orr.render(fr, order=2)
        np.testing.assert_allclose(sigs, oracle.signals_of(res), rtol=1e-4, atol=1e-6)
        assert ss["closestRays"] == int(res["bounces"].sum()) and ss["shadowRays"] == int(res["shadowRays"].sum())
    rt_a.destroy()
    rt_s.destroy()


def test_batch_rejects_prefix_scan():
    from renderingengine_amd.batch import BatchEngine, frame_desc

    sc = scenes.cornell_box(32, 18)
    rt = WavefrontRuntime(32 * 18, 1)
    rt.set_scene(sc.tris, sc.materials, sc.lights)
    eng = BatchEngine(path_capacity=4096, max_jobs=4)
    eng.add_scene(rt.scene)
    f = frame_desc(sc.camera, 32, 18)
    f.queueMode = _lib.REX_QUEUE_PREFIX_SCAN
    with pytest.raises(_lib.RexError):
        eng.submit(0, f)
    eng.destroy()
    rt.destroy()
