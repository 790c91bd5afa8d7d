"""Reconstruction boundary on the host (no GPU): the oracle's MotionVectorsCS,
ComposeCS and motion-input restatements pinned against the reference's own
CPU motion vectors (reconstruction/src/MotionVectors.cpp, built into
oracle/_ref where the reference tree is present), the known answers of the
reference's test (reconstruction/tests/ReconstructionTests.cpp:80-127), matrix
inverses and the closed forms of the compose outputs; and the C-ABI's
argument checks (host-side, no kernel launch)."""
import ctypes

import numpy as np
import pytest

from renderingengine_amd import _lib, reconstruction as R

F32_MAX = np.finfo(np.float32).max


def _t(x=0.0, y=0.0, z=0.0):
    m = np.eye(4, dtype=np.float32)
    m[0:3, 3] = (x, y, z)
    return m


def _pixel():
    gb = np.zeros(1, R.GBUFFER)
    gb["primary"]["identity"][0] = (3, 5, 7, R.PRIMARY_SURFACE_VALID)
    return gb, np.zeros(1, R.MOTION_INPUT)


def _motion(oracle, obj=(0, 0, 0), cur_o2w=None, prev_o2w=None, vpc=None, vpp=None, wvp=None, jc=(0, 0),
            jp=(0, 0)):
    gb, mi = _pixel()
    mi["objectPosition"][0] = obj
    I = np.eye(4, dtype=np.float32)
    mats = [I if m is None else np.asarray(m, np.float32) for m in (cur_o2w, prev_o2w, vpc, vpp, wvp)]
    out = oracle.motion_vectors(gb, mi, mats[0], mats[1], mats[2], mats[3], mats[4], jc, jp)[0]["motion"]
    ref = oracle.ref_motion_vector(np.asarray(obj, np.float32), *mats, jc, jp)
    return out, ref


def test_motion_vector_known_answers(oracle):
    """TestMotionVectorSignAndJitter, case by case."""
    near = lambda a, b: abs(float(a) - b) < 1e-5  # noqa: E731
    m, ref = _motion(oracle, vpp=_t(0.2))
    assert m["identity"][3] == R.MOTION_VALID and tuple(m["identity"][:3]) == (3, 5, 7)
    assert near(m["currentPreviousUv"][0], 0.5) and near(m["currentPreviousUv"][2], 0.6)
    assert near(m["motionExpectedDepth"][0], 0.1) and near(m["motionExpectedDepth"][2], 0.0)
    cases = [(m, ref)]
    m, ref = _motion(oracle, cur_o2w=_t(0.2))
    assert near(m["currentPreviousUv"][0], 0.6) and near(m["currentPreviousUv"][2], 0.5)
    assert near(m["motionExpectedDepth"][0], -0.1)
    cases.append((m, ref))
    jit = dict(jc=(0.01, -0.02), jp=(-0.02, 0.03))
    m, ref = _motion(oracle, **jit)
    assert near(m["motionExpectedDepth"][0], -0.03) and near(m["motionExpectedDepth"][1], 0.05)
    cases.append((m, ref))
    m, ref = _motion(oracle, wvp=_t(0, 0, -2), **jit)
    assert m["identity"][3] == R.MOTION_VALID and near(m["motionExpectedDepth"][2], 2.0)
    cases.append((m, ref))
    bad = np.eye(4, dtype=np.float32)
    bad[3, 3] = 0.0
    m0, ref0 = _motion(oracle, wvp=_t(0, 0, -2), vpc=bad, **jit)
    bad[3, 3] = -1.0
    m1, ref1 = _motion(oracle, vpc=bad)
    tiny = np.eye(4, dtype=np.float32)
    tiny[3, 3] = 2.0e-8
    m2, ref2 = _motion(oracle, obj=(F32_MAX, 0, 0), vpc=tiny, vpp=tiny)
    for m in (m0, m1, m2):
        assert m["identity"][3] == 0 and not m["motionExpectedDepth"].any()
    if ref is not None:
        for m, r in cases:
            assert r[7] == 1.0
            np.testing.assert_allclose(m["currentPreviousUv"], r[0:4], atol=1e-6)
            np.testing.assert_allclose(m["motionExpectedDepth"][:3], r[4:7], atol=1e-6)
        assert ref0[7] == 0.0 and ref1[7] == 0.0 and ref2[7] == 0.0


def _camera(rng):
    pos = rng.uniform(-3, 3, 3)
    fwd = rng.normal(size=3)
    fwd /= np.linalg.norm(fwd)
    right = np.cross(fwd, (0, 1, 0))
    right /= np.linalg.norm(right)
    up = np.cross(right, fwd)
    return dict(position=pos, forward=fwd, right=right, up=up, tanHalfFov=float(rng.uniform(0.3, 1.0)),
                aspect=float(rng.uniform(1.0, 2.0)))


def _rigid(rng):
    a = rng.normal(size=3)
    q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
    m = np.eye(4)
    m[0:3, 0:3] = q * rng.uniform(0.5, 2.0)
    m[0:3, 3] = a
    return m.astype(np.float32)


def test_motion_vectors_against_reference_cpu(oracle):
    """Random perspective cameras and rigid instances: the shader restatement
    agrees with the reference's CPU ComputeMotionVector on validity and
    within 1e-6 (+1e-6 relative) on UVs, motion and depth (the CPU divides
    by clip w through its reciprocal, the shader directly)."""
    if oracle.ref() is None:
        pytest.skip("oracle/_ref not built (no reference tree)")
    rng = np.random.default_rng(5)
    n = 3000
    gb = np.zeros(n, R.GBUFFER)
    gb["primary"]["identity"][:, 3] = R.PRIMARY_SURFACE_VALID
    mi = np.zeros(n, R.MOTION_INPUT)
    cur = np.stack([_rigid(rng) for _ in range(8)])
    prev = np.stack([_rigid(rng) for _ in range(8)])
    ca, cb = _camera(rng), _camera(rng)
    vpc, _ = R.camera_matrices(ca)
    vpp, wvp = R.camera_matrices(cb)
    idx = rng.integers(0, 8, n)
    mi["currentTransformIndex"] = idx
    mi["previousTransformIndex"] = idx
    mi["objectPosition"] = rng.uniform(-6, 6, (n, 3))
    jc, jp = (0.001, -0.002), (-0.0015, 0.0005)
    got = oracle.motion_vectors(gb, mi, cur, prev, vpc, vpp, wvp, jc, jp)["motion"]
    valid = 0
    for i in range(n):
        r = oracle.ref_motion_vector(mi["objectPosition"][i], cur[idx[i]], prev[idx[i]], vpc, vpp, wvp, jc, jp)
        m = got[i]
        assert (m["identity"][3] == 1) == (r[7] == 1.0), i
        if r[7] == 1.0:
            valid += 1
            np.testing.assert_allclose(m["currentPreviousUv"], r[0:4], rtol=1e-6, atol=1e-6)
            np.testing.assert_allclose(m["motionExpectedDepth"][:3], r[4:7], rtol=1e-6, atol=1e-6)
    assert 0.1 * n < valid < n   # both behind-camera rejections and valid points occur


def test_motion_inputs_invert_instances(oracle):
    """Object-space positions of the first instance with the surface's stable
    ID (a world point mapped back through world-to-object recovers the
    object point), the instance count for unknown IDs, the world position
    without a table, zeros for invalid surfaces."""
    rng = np.random.default_rng(2)
    k, n = 6, 500
    inst = np.zeros((k, 56), np.uint32)
    o2w = np.stack([_rigid(rng) for _ in range(k)])
    ids = np.array([40, 11, 40, 7, 99, 11], np.uint32)  # duplicates: the first wins
    for i in range(k):
        inst[i, 0:16] = o2w[i].reshape(-1).view(np.uint32)
        inst[i, 16:32] = np.linalg.inv(o2w[i].astype(np.float64)).astype(np.float32).reshape(-1).view(np.uint32)
        inst[i, 50] = ids[i]
    gb = np.zeros(n, R.GBUFFER)
    which = rng.integers(0, k, n)
    obj = rng.uniform(-2, 2, (n, 3))
    world = np.einsum("nij,nj->ni", o2w[which].astype(np.float64), np.c_[obj, np.ones(n)])[:, :3]
    gb["primary"]["worldPositionLinearDepth"][:, :3] = world
    gb["primary"]["identity"][:, 1] = ids[which]
    gb["primary"]["identity"][:, 3] = R.PRIMARY_SURFACE_VALID
    gb["primary"]["identity"][0, 1] = 12345             # unknown ID
    gb["primary"]["identity"][1, 3] = 0                 # invalid surface
    got = oracle.motion_inputs(gb, inst)
    first = {int(v): int(np.nonzero(ids == v)[0][0]) for v in ids}
    for i in range(2, n):
        t = first[int(ids[which[i]])]
        assert got["currentTransformIndex"][i] == t and got["previousTransformIndex"][i] == t
        if t == which[i]:
            np.testing.assert_allclose(got["objectPosition"][i], obj[i], atol=2e-5 * (1 + np.abs(world[i]).max()))
    assert got["currentTransformIndex"][0] == k and not got["objectPosition"][0].any()
    assert got[1].tobytes() == bytes(32)
    none = oracle.motion_inputs(gb, None)
    assert np.array_equal(none["objectPosition"][2:], gb["primary"]["worldPositionLinearDepth"][2:, :3])
    assert not none["currentTransformIndex"].any()


def test_compose_closed_forms(oracle):
    rng = np.random.default_rng(9)
    n = 257
    raw = np.zeros(n, R.SIGNAL)
    for f in R.SIGNAL.names:
        raw[f][:, :3] = rng.uniform(0, 4, (n, 3))
    term = np.zeros((n, 4), np.float32)
    term[:, :3] = rng.uniform(0, 2, (n, 3))
    total = sum(raw[f][:, :3].astype(np.float64) for f in R.SIGNAL.names) + term[:, :3]
    out, _ = oracle.compose(R.COMPOSE_RAW, n, raw=raw, terminal=term)
    np.testing.assert_allclose(out[:, :3], total, rtol=1e-6)
    assert np.all(out[:, 3] == 1.0)
    # Progressive mean over five frames = the arithmetic mean.
    film = np.zeros((n, 4), np.float32)
    frames = []
    for k in range(5):
        for f in R.SIGNAL.names:
            raw[f][:, :3] = rng.uniform(0, 4, (n, 3))
        frames.append(sum(raw[f][:, :3].astype(np.float64) for f in R.SIGNAL.names) + term[:, :3])
        out, film = oracle.compose(R.COMPOSE_PROGRESSIVE_MEAN, n, raw=raw, terminal=term, film=film, film_count=k)
    np.testing.assert_allclose(film[:, :3], np.mean(frames, axis=0), rtol=1e-5)
    assert np.all(film[:, 3] == 5.0) and np.array_equal(out[:, :3], film[:, :3])
    # Surface and material outputs.
    gb = np.zeros(n, R.GBUFFER)
    nrm = rng.normal(size=(n, 3))
    gb["primary"]["shadingNormalMetallic"][:, :3] = nrm * rng.uniform(0.1, 3, (n, 1))
    gb["primary"]["shadingNormalMetallic"][0, :3] = 0.0
    gb["primary"]["shadingNormalMetallic"][:, 3] = rng.uniform(-0.5, 1.5, n)
    gb["primary"]["geometricNormalRoughness"][:, 3] = rng.uniform(-0.5, 1.5, n)
    gb["motion"]["motionExpectedDepth"][:, :2] = rng.uniform(-0.1, 0.1, (n, 2))
    mats = np.zeros((4, 32), np.uint32)
    mf = mats.view(np.float32)
    mf[:, 0:3] = rng.uniform(-0.5, 1.5, (4, 3))
    mf[:, 4:7] = rng.uniform(-0.5, 2, (4, 3))
    mf[:, 7] = rng.uniform(0, 3, 4)
    mat = rng.integers(0, 6, n)                      # 4, 5: out of range
    gb["primary"]["identity"][:, 0] = mat
    valid = rng.random(n) < 0.8
    gb["primary"]["identity"][:, 3] = valid
    ok = valid & (mat < 4)
    out, _ = oracle.compose(R.COMPOSE_NORMAL, n, gbuffer=gb)
    unit = nrm / np.linalg.norm(nrm, axis=1, keepdims=True)
    np.testing.assert_allclose(out[1:, :3], unit[1:] * 0.5 + 0.5, atol=1e-6)
    assert tuple(out[0, :3]) == (0.5, 0.5, 1.0)
    out, _ = oracle.compose(R.COMPOSE_ROUGHNESS, n, gbuffer=gb)
    assert np.array_equal(out[:, 0], np.clip(gb["primary"]["geometricNormalRoughness"][:, 3], 0, 1))
    out, _ = oracle.compose(R.COMPOSE_METALLIC, n, gbuffer=gb)
    assert np.array_equal(out[:, 2], np.clip(gb["primary"]["shadingNormalMetallic"][:, 3], 0, 1))
    out, _ = oracle.compose(R.COMPOSE_MOTION, n, gbuffer=gb)
    np.testing.assert_allclose(out[:, :2], 0.5 + 16 * gb["motion"]["motionExpectedDepth"][:, :2], atol=1e-7)
    assert np.all(out[:, 2] == 0.5)
    out, _ = oracle.compose(R.COMPOSE_BASE_COLOR, n, gbuffer=gb, materials=mats)
    safe = np.minimum(mat, 3)
    np.testing.assert_array_equal(out[:, :3], np.where(ok[:, None], np.clip(mf[safe, 0:3], 0, 1), 0))
    out, _ = oracle.compose(R.COMPOSE_EMISSIVE, n, gbuffer=gb, materials=mats)
    em = np.maximum(mf[safe, 4:7] * mf[safe, 7:8], 0)
    np.testing.assert_array_equal(out[:, :3], np.where(ok[:, None], em, 0))


def test_abi_records_and_compose_argument_errors():
    """Record sizes of the public header, and ComposeCS / MotionVectorsCS
    argument checks (rejected before any launch)."""
    assert ctypes.sizeof(_lib.ComposeDesc) == 72 and ctypes.sizeof(_lib.MotionConstants) == 224
    lib = _lib.load()
    out = ctypes.c_void_p(0x1000)

    def compose(mode, **kw):
        d = _lib.ComposeDesc()
        d.width, d.height, d.outputMode = kw.get("w", 4), 4, mode
        d.output = kw.get("output", out)
        d.rawSignal = kw.get("raw", 0x2000)
        d.terminal = kw.get("terminal", 0x3000)
        d.gbuffer = kw.get("gbuffer", None)
        d.film = kw.get("film", None)
        return lib.rex_compose_v2(ctypes.byref(d), None)

    for mode in (1, 2, 3, 5, 6, 7, 8, 9):
        assert compose(mode) == 2                                 # UNSUPPORTED
    assert compose(16) == 1
    assert compose(0, output=None) == 1
    assert compose(0, w=0) == 1
    assert compose(15) == 1                                       # no film
    assert compose(0, raw=None) == 1
    assert compose(11) == 1                                       # no G-buffer
    assert compose(0, terminal=0x3004) == 1                       # misaligned
    c = R.motion_constants(4, 4, np.eye(4), np.eye(4), np.eye(4), current_count=2)
    assert lib.rex_motion_vectors_v2(ctypes.c_void_p(0x1000), ctypes.c_void_p(0x2000), None, None,
                                     ctypes.byref(c), None) == 1  # NULL transforms with a count
    assert lib.rex_motion_vectors_v2(None, ctypes.c_void_p(0x2000), None, None, ctypes.byref(c), None) == 1
