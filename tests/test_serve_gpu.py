"""The request loop on the GPU: frames and ReSTIR DI sequences submitted
into one server, admitted by footprint, run overlapped; each result equals
the same job run alone, bit for bit -- a frame's Film equals the plain
per-frame runtime's, a sequence's last Film equals the same frames rendered
back to back by a fresh WavefrontRuntime with the ReSTIR estimator."""
import numpy as np
import pytest

from renderingengine_amd import _lib, scenes
from renderingengine_amd.batch import ShadedSceneHandle, frame_desc
from renderingengine_amd.render import WALK_WIDE_COMPRESSED, WavefrontRuntime
from renderingengine_amd.serve import STATE_FINISHED, Server, restir_desc

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def scene_set():
    out = []
    for sc in (scenes.cornell_box(64, 36), scenes.soup_scene(n=20000, seed=4, quads=5, palette=16),
               scenes.restir_room(emitters=24), scenes.restir_scene(n=20000, seed=2, emitters=300, palette=16,
                                                                   width=48, height=32, extent=2.0)):
        out.append((sc, ShadedSceneHandle(sc.tris, sc.materials, sc.lights)))
    yield out
    for _, h in out:
        h.destroy()


FRAMES = [  # (scene, width, height, spp, max bounce, sample index, seed, estimator)
    (0, 64, 36, 2, 4, 0, 1, _lib.REX_ESTIMATOR_MIS),
    (1, 80, 45, 3, 6, 5, 2, _lib.REX_ESTIMATOR_MIS),
    (2, 40, 30, 1, 3, 0, 3, _lib.REX_ESTIMATOR_NEE),
    (3, 48, 32, 2, 5, 1, 4, _lib.REX_ESTIMATOR_MIS),
]
SEQUENCES = [  # (scene, width, height, frames, max bounce, seed, candidates, neighbours)
    (2, 64, 36, 4, 2, 7, 8, 5),
    (3, 48, 32, 3, 1, 9, 32, 5),
    (2, 32, 16, 2, 3, 11, 4, 2),
]


def _plain_frame(sc, job):
    _, w, h, spp, bounce, si, seed, est = job
    rt = WavefrontRuntime(w * h, spp)
    rt.set_scene(sc.tris, sc.materials, sc.lights)
    rt.update_frame(sc.camera, w, h, sample_index=si, spp=spp, max_bounce=bounce, estimator=est, seed=seed,
                    walk_mode=WALK_WIDE_COMPRESSED)
    rt.record_frame()
    film = rt.film()
    rt.destroy()
    return film


def _alone_sequence(sc, seq):
    _, w, h, frames, bounce, seed, cand, nb = seq
    rt = WavefrontRuntime(w * h, 1)
    rt.set_scene(sc.tris, sc.materials, sc.lights)
    rt.set_restir(candidates=cand, neighbors=nb)
    for k, cam in enumerate(scenes.camera_walk(sc.camera, frames, step=0.02, seed=seed)):
        rt.update_frame(cam, w, h, sample_index=k, spp=1, max_bounce=bounce, estimator=_lib.REX_ESTIMATOR_RESTIR,
                        seed=seed, walk_mode=WALK_WIDE_COMPRESSED)
        rt.record_frame()
    film = rt.film()
    rt.destroy()
    return film


def _submit_frame(srv, scene_set, job):
    s, w, h, spp, bounce, si, seed, est = job
    sc = scene_set[s][0]
    return srv.submit_frame(s, frame_desc(sc.camera, w, h, sample_index=si, spp=spp, max_bounce=bounce,
                                          estimator=est, seed=seed))


def _submit_sequence(srv, scene_set, seq):
    s, w, h, frames, bounce, seed, cand, nb = seq
    sc = scene_set[s][0]
    cams = scenes.camera_walk(sc.camera, frames, step=0.02, seed=seed)
    return srv.submit_sequence(s, frame_desc(cams[0], w, h, max_bounce=bounce, seed=seed), cams,
                               restir=restir_desc(candidates=cand, neighbors=nb))


def _check(srv, scene_set, frame_ids, seq_ids):
    for jid, job in zip(frame_ids, FRAMES):
        info = srv.job(jid)
        assert info["state"] == STATE_FINISHED and info["finishNs"] >= info["admitNs"] >= info["submitNs"] > 0
        exp = _plain_frame(scene_set[job[0]][0], job)
        assert np.array_equal(srv.film(jid).view(np.uint32), exp.view(np.uint32)), f"frame {job} differs"
    for jid, seq in zip(seq_ids, SEQUENCES):
        info = srv.job(jid)
        assert info["state"] == STATE_FINISHED and info["rays"] > seq[1] * seq[2] * seq[3]
        exp = _alone_sequence(scene_set[seq[0]][0], seq)
        assert np.array_equal(srv.film(jid).view(np.uint32), exp.view(np.uint32)), f"sequence {seq} differs"


def test_mixed_jobs_equal_alone(scene_set):
    """Frames and sequences interleaved, three sequence slots (so sequences
    overlap each other and the engine), continuous pumping while submitting."""
    srv = Server(path_capacity=20000, sequence_slots=3, sequence_max_pixels=64 * 36, sequence_output_mib=16,
                 output_mib=64)
    for _, h in scene_set:
        srv.add_scene(h)
    frame_ids, seq_ids = [], []
    for k in range(max(len(FRAMES), len(SEQUENCES))):
        if k < len(FRAMES):
            frame_ids.append(_submit_frame(srv, scene_set, FRAMES[k]))
        if k < len(SEQUENCES):
            seq_ids.append(_submit_sequence(srv, scene_set, SEQUENCES[k]))
        srv.pump()
    print(f"{len(frame_ids)} frames, {len(seq_ids)} sequences submitted", flush=True)
    srv.drain()
    st = srv.stats()
    assert st["finished"] == len(FRAMES) + len(SEQUENCES) and st["queuedJobs"] == 0
    assert st["sequenceFrames"] == sum(s[3] for s in SEQUENCES) and st["engineRays"] > 0
    _check(srv, scene_set, frame_ids, seq_ids)
    for jid in frame_ids + seq_ids:
        srv.release(jid)
    assert srv.stats()["admittedBytes"] == 0
    srv.destroy()


def test_budget_admission(scene_set):
    """A budget a little above the resident scenes: jobs wait for earlier
    ones to be released, the admitted footprint never exceeds what the budget
    leaves, and the results are unchanged."""
    probe = Server(path_capacity=20000, sequence_slots=2, sequence_max_pixels=64 * 36, sequence_output_mib=16,
                   output_mib=64)
    for _, h in scene_set:
        probe.add_scene(h)
    resident = probe.stats()["residentBytes"]
    probe.destroy()
    room = 400 << 10
    srv = Server(budget_bytes=resident + room, path_capacity=20000, sequence_slots=2,
                 sequence_max_pixels=64 * 36, sequence_output_mib=16, output_mib=64)
    for _, h in scene_set:
        srv.add_scene(h)
    with pytest.raises(_lib.RexError):   # 48x32 sequence: ~3 MiB of reservoirs, above the room
        _submit_sequence(srv, scene_set, SEQUENCES[1])
    frame_ids = [_submit_frame(srv, scene_set, j) for j in FRAMES]
    results = {}
    while len(results) < len(frame_ids):
        srv.pump()
        st = srv.stats()
        assert st["admittedBytes"] <= room
        for jid in frame_ids:
            if jid not in results and srv.job(jid)["state"] == STATE_FINISHED:
                results[jid] = srv.film(jid)
                srv.release(jid)
    assert srv.stats()["peakAdmittedBytes"] <= room
    for jid, job in zip(frame_ids, FRAMES):
        exp = _plain_frame(scene_set[job[0]][0], job)
        assert np.array_equal(results[jid].view(np.uint32), exp.view(np.uint32)), f"frame {job} differs"
    srv.destroy()


def test_server_argument_errors(scene_set):
    srv = Server(path_capacity=2000, sequence_slots=1, sequence_max_pixels=32 * 16, sequence_output_mib=4,
                 output_mib=4)
    srv.add_scene(scene_set[2][1])
    sc = scene_set[2][0]
    with pytest.raises(_lib.RexError):   # larger than the renderer slot
        srv.submit_sequence(0, frame_desc(sc.camera, 64, 36), [sc.camera])
    with pytest.raises(_lib.RexError):   # 65 frames
        srv.submit_sequence(0, frame_desc(sc.camera, 8, 8), [sc.camera] * 65)
    with pytest.raises(_lib.RexError):   # frames do not take the ReSTIR estimator
        srv.submit_frame(0, frame_desc(sc.camera, 8, 8, estimator=_lib.REX_ESTIMATOR_RESTIR))
    with pytest.raises(_lib.RexError):   # not resident
        srv.submit_frame(3, frame_desc(sc.camera, 8, 8))
    with pytest.raises(_lib.RexError):   # outputs above the 4 MiB engine arena
        srv.submit_frame(0, frame_desc(sc.camera, 512, 512, spp=2))
    jid = srv.submit_frame(0, frame_desc(sc.camera, 8, 8))
    with pytest.raises(_lib.RexError):   # not finished yet
        srv.release(jid)
    srv.drain()
    srv.release(jid)
    srv.destroy()
