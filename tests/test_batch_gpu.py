"""Batched jobs on the GPU: every job's Film from the batch engine equals
the same frame rendered alone by the plain per-frame runtime, bit for bit
(paths are keyed by (pixel, sample) and resolved in sample order, so
sharing queues, launches and path slots with other jobs changes nothing).
Covers several resident scenes, uneven resolutions / spp / bounces, jobs
larger than the path pool (their paths stream through it), continuous
submission between steps, and hipGraph replay."""
import numpy as np
import pytest

from renderingengine_amd import _lib, scenes
from renderingengine_amd.batch import BatchEngine, ShadedSceneHandle, frame_desc
from renderingengine_amd.render import WALK_WIDE_COMPRESSED, WavefrontRuntime

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def scene_set():
    out = []
    for sc in (scenes.cornell_box(64, 36), scenes.soup_scene(n=20000, seed=4, quads=5, palette=16),
               scenes.restir_room(emitters=24)):
        out.append((sc, ShadedSceneHandle(sc.tris, sc.materials, sc.lights)))
    yield out
    for _, h in out:
        h.destroy()


JOBS = [  # (scene, width, height, spp, max bounce, sample index, seed, estimator)
    (0, 64, 36, 2, 4, 0, 1, _lib.REX_ESTIMATOR_MIS),
    (1, 80, 45, 3, 6, 5, 2, _lib.REX_ESTIMATOR_MIS),
    (2, 40, 30, 1, 3, 0, 3, _lib.REX_ESTIMATOR_NEE),
    (0, 32, 32, 5, 8, 2, 4, _lib.REX_ESTIMATOR_MIS),
    (1, 48, 27, 1, 2, 0, 5, _lib.REX_ESTIMATOR_BSDF_ONLY),
    (2, 64, 48, 4, 5, 9, 6, _lib.REX_ESTIMATOR_MIS),
]


def _plain_film(sc, job):
    _, w, h, spp, bounce, si, seed, est = job
    rt = WavefrontRuntime(w * h, spp)
    rt.set_scene(sc.tris, sc.materials, sc.lights)
    rt.update_frame(sc.camera, w, h, sample_index=si, spp=spp, max_bounce=bounce, estimator=est, seed=seed,
                    walk_mode=WALK_WIDE_COMPRESSED)
    rt.record_frame()
    film = rt.film()
    rt.destroy()
    return film


def _submit(eng, idx, scene_set, job):
    s, w, h, spp, bounce, si, seed, est = job
    sc = scene_set[s][0]
    return eng.submit(idx[s], frame_desc(sc.camera, w, h, sample_index=si, spp=spp, max_bounce=bounce,
                                         estimator=est, seed=seed))


@pytest.mark.parametrize("capacity,graphs", [(0, False), (6000, False), (6000, True)])
def test_batch_films_equal_plain_frames(scene_set, capacity, graphs):
    """A pool of 6000 paths: every job streams through it over many steps;
    0 selects the default pool (everything in flight at once)."""
    eng = BatchEngine(path_capacity=capacity, max_jobs=8, use_graphs=graphs)
    idx = [eng.add_scene(h) for _, h in scene_set]
    ids = [_submit(eng, idx, scene_set, j) for j in JOBS]
    print(f"capacity={capacity} graphs={graphs}: {len(ids)} jobs", flush=True)
    eng.run()
    st = eng.stats()
    assert st["jobsFinished"] == len(JOBS) and st["pathsInFlight"] == 0 and st["queuedJobs"] == 0
    for jid, job in zip(ids, JOBS):
        info = eng.job(jid)
        assert info["state"] == 3
        got = eng.film(jid)
        exp = _plain_film(scene_set[job[0]][0], job)
        assert np.array_equal(got.view(np.uint32), exp.view(np.uint32)), f"job {job}: film differs"
        eng.release(jid)
    eng.destroy()


def test_continuous_submission(scene_set):
    """Jobs submitted while others are in flight join at the next step."""
    eng = BatchEngine(path_capacity=12000, max_jobs=4)
    idx = [eng.add_scene(h) for _, h in scene_set]
    first = [_submit(eng, idx, scene_set, j) for j in JOBS[:3]]
    for _ in range(3):
        eng.step()
    later = [_submit(eng, idx, scene_set, j) for j in JOBS[3:]]
    eng.step()
    assert eng.job(later[0])["state"] in (0, 1)
    eng.run()
    for jid, job in zip(first + later, JOBS):
        got = eng.film(jid)
        exp = _plain_film(scene_set[job[0]][0], job)
        assert np.array_equal(got.view(np.uint32), exp.view(np.uint32)), f"job {job}: film differs"
    eng.destroy()


def test_output_arena_admission(scene_set):
    """Jobs take their Film and samples from a 1 MiB output arena: the six
    jobs (~0.75 MiB) fit; six more wait for room and the run stops with an
    error naming the full arena; after the first six are released they run,
    and every Film still equals the plain frame."""
    eng = BatchEngine(path_capacity=5000, max_jobs=16, output_mib=1)
    idx = [eng.add_scene(h) for _, h in scene_set]
    first = [_submit(eng, idx, scene_set, j) for j in JOBS]
    eng.run()
    st = eng.stats()
    assert st["outputBytesTotal"] == 1 << 20 and 0 < st["outputBytesInUse"] <= 1 << 20
    second = [_submit(eng, idx, scene_set, j) for j in JOBS]
    with pytest.raises(_lib.RexError):
        eng.run()
    assert eng.job(second[-1])["state"] == 0 and eng.job(second[-1])["film"] is None
    films = [eng.film(j) for j in first]
    for j in first:
        eng.release(j)
    eng.run()
    assert eng.stats()["jobsFinished"] == 2 * len(JOBS)
    for jid, film, job in zip(second, films, JOBS):
        got = eng.film(jid)
        exp = _plain_film(scene_set[job[0]][0], job)
        assert np.array_equal(got.view(np.uint32), exp.view(np.uint32)), f"job {job}: film differs"
        assert np.array_equal(film.view(np.uint32), exp.view(np.uint32)), f"job {job}: first film differs"
    eng.destroy()


def test_batch_argument_errors(scene_set):
    eng = BatchEngine(path_capacity=1000)
    idx = eng.add_scene(scene_set[0][1])
    sc = scene_set[0][0]
    with pytest.raises(_lib.RexError):
        eng.submit(idx, frame_desc(sc.camera, 64, 36, max_bounce=17))
    with pytest.raises(_lib.RexError):
        eng.submit(idx + 5, frame_desc(sc.camera, 8, 8))         # not resident
    with pytest.raises(_lib.RexError):
        eng.submit(idx, frame_desc(sc.camera, 8, 8, estimator=_lib.REX_ESTIMATOR_RESTIR))
    small = BatchEngine(path_capacity=1000, output_mib=1)
    small.add_scene(scene_set[0][1])
    with pytest.raises(_lib.RexError):
        small.submit(0, frame_desc(sc.camera, 256, 256, spp=2))   # 3 MiB of outputs, arena 1 MiB
    small.destroy()
    eng.destroy()
