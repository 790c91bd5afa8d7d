"""The batched and served configurations on the MI355X.

C4: the 64-job mix (resolutions U{256^2, 512^2, 1280x720, 1920x1080}, soups of
LogU(1e4, 2e6) triangles, spp U{1..16}, max bounce U{4..8}, seeded cameras)
through one batch engine with its autotuner on; every job's Film
bit-identical to the same frame rendered alone by rex_render_frame, and the
four cheapest jobs' Films within 1e-4 of the oracle's per-path integrator.

C5: a Poisson arrival window through the request loop with 20% of the jobs
8-frame ReSTIR DI sequences (device-resident reservoir and Film history on
renderer slots) and the rest frames in the batch engine; every frame's Film
bit-identical to the frame alone, every sequence's Film bit-identical to the
sequence rendered alone, and sequences' Films within 1e-4 of the oracle's
ReSTIR restatement (OracleRestir).
"""
import time

import numpy as np
import pytest

from renderingengine_amd import _lib, scenes
from renderingengine_amd.batch import BatchEngine, ShadedSceneHandle, frame_desc
from renderingengine_amd.render import WALK_WIDE_COMPRESSED, WavefrontRuntime
from renderingengine_amd.serve import STATE_FINISHED, Server, restir_desc

pytestmark = pytest.mark.gpu

RTOL = 1.0e-4
ATOL = 1.0e-6


def _camera(j):
    return scenes.look_at_camera(j["pos"], (0.0, 0.0, 0.0), vfov_deg=45.0, aspect=j["width"] / j["height"])


def _plain(handle, cam, j):
    rt = WavefrontRuntime(j["width"] * j["height"], j["spp"])
    rt.attach(handle)
    rt.update_frame(cam, j["width"], j["height"], sample_index=0, spp=j["spp"], max_bounce=j["bounce"],
                    estimator=_lib.REX_ESTIMATOR_MIS, seed=j["seed"], walk_mode=WALK_WIDE_COMPRESSED)
    rt.record_frame()
    film = rt.film()
    rt.destroy()
    return film


def test_c4_batch_64_job_mix(oracle):
    jobs = scenes.batch_job_mix(64, 7, 2_000_000)
    t0 = time.perf_counter()
    built = []
    for j in jobs:
        sc = scenes.soup_scene(n=j["tris"], seed=j["scene_seed"], quads=16, palette=64, width=j["width"],
                               height=j["height"])
        built.append((sc, _camera(j), ShadedSceneHandle(sc.tris, sc.materials, sc.lights)))
    print(f"64 scenes ({sum(j['tris'] for j in jobs)} triangles) built in {time.perf_counter() - t0:.1f}s",
          flush=True)
    eng = BatchEngine(path_capacity=16 << 20, max_jobs=256, use_graphs=True, shade_blocks_per_cu=16, autotune=True)
    idx = [eng.add_scene(b[2]) for b in built]
    ids = [eng.submit(idx[k], frame_desc(built[k][1], j["width"], j["height"], sample_index=0, spp=j["spp"],
                                         max_bounce=j["bounce"], estimator=_lib.REX_ESTIMATOR_MIS, seed=j["seed"]))
           for k, j in enumerate(jobs)]
    t1 = time.perf_counter()
    eng.run()
    st = eng.stats()
    print(f"batched: {time.perf_counter() - t1:.2f}s, {st['steps']} steps, autotune {st['autotune']}: shade "
          f"{st['shadeBlocksPerCu']}/CU graphs {st['graphs']} budget {st['pathBudget']}", flush=True)
    assert st["jobsFinished"] == len(jobs) and st["autotune"] == 2
    for k, j in enumerate(jobs):
        exp = _plain(built[k][2], built[k][1], j)
        got = eng.film(ids[k])
        assert np.array_equal(got.view(np.uint32), exp.view(np.uint32)), f"job {k} {j} differs from the plain frame"
    cost = [j["width"] * j["height"] * j["spp"] * j["bounce"] * np.log(j["tris"]) for j in jobs]
    for k in np.argsort(cost)[:4]:
        j = jobs[int(k)]
        sc, cam, _ = built[int(k)]
        orr = oracle.OracleRender(sc.tris, sc.materials, sc.lights, selection=1)
        fr = orr.frame(cam, j["width"], j["height"], spp=j["spp"], max_bounce=j["bounce"], sample_index=0,
                       estimator=_lib.REX_ESTIMATOR_MIS, seed=j["seed"])
        res = orr.render(fr, order=2, threads=16)
        exp = oracle.film_accumulate(res, j["width"] * j["height"], j["spp"])
        got = eng.film(ids[int(k)])
        print(f"oracle check of job {int(k)}: {j['width']}x{j['height']} spp {j['spp']} bounce {j['bounce']} "
              f"{j['tris']} triangles", flush=True)
        np.testing.assert_allclose(got, exp, rtol=RTOL, atol=ATOL)
    eng.destroy()
    for b in built:
        b[2].destroy()


SERVE_TEMPLATES = [  # (triangles, width, height, spp, max bounce, scene seed)
    (20_000, 64, 36, 2, 4, 1), (40_000, 96, 54, 1, 5, 2), (30_000, 80, 45, 4, 6, 3),
    (50_000, 64, 64, 2, 4, 4), (25_000, 128, 72, 1, 4, 5),
]
SEQ_FRAMES = 8
SEQ_RESTIR = dict(candidates=8, neighbors=5)


def test_c5_poisson_window_with_restir_sequences(oracle):
    built = []
    for n, w, h, spp, bounce, seed in SERVE_TEMPLATES:
        sc = scenes.restir_scene(n=n, seed=seed, emitters=200, palette=32, width=w, height=h, extent=3.0)
        built.append((sc, ShadedSceneHandle(sc.tris, sc.materials, sc.lights)))
    trace = scenes.serve_trace(len(SERVE_TEMPLATES), 120.0, 0.5, 17, sequence_share=0.2,
                               sequence_frames=SEQ_FRAMES)
    n_seq = sum(1 for t in trace if t[2] == 1)
    print(f"{len(trace)} arrivals, {n_seq} ReSTIR sequences", flush=True)
    assert n_seq >= 3 and len(trace) - n_seq >= 20

    def walk(k, n):
        return scenes.camera_walk(built[k][0].camera, SEQ_FRAMES, step=0.02, seed=1000 + n)

    srv = Server(path_capacity=1 << 20, use_graphs=True, shade_blocks_per_cu=16, output_mib=512, sequence_slots=4,
                 sequence_max_pixels=128 * 72, sequence_output_mib=64)
    for _, h in built:
        srv.add_scene(h)
    submitted = []
    t0 = time.perf_counter()
    nxt = 0
    while nxt < len(trace):
        now = time.perf_counter() - t0
        while nxt < len(trace) and trace[nxt][0] <= now:
            _, k, kind = trace[nxt]
            n, w, h, spp, bounce, seed = SERVE_TEMPLATES[k]
            cam = built[k][0].camera
            if kind == 0:
                jid = srv.submit_frame(k, frame_desc(cam, w, h, sample_index=0, spp=spp, max_bounce=bounce,
                                                     estimator=_lib.REX_ESTIMATOR_MIS, seed=seed + nxt))
            else:
                jid = srv.submit_sequence(k, frame_desc(cam, w, h, max_bounce=bounce, seed=seed + nxt),
                                          walk(k, nxt), restir=restir_desc(**SEQ_RESTIR))
            submitted.append((jid, nxt))
            nxt += 1
        srv.pump()
    srv.drain()
    st = srv.stats()
    print(f"served {st['finished']} jobs in {time.perf_counter() - t0:.2f}s", flush=True)
    assert st["finished"] == len(trace)
    oracle_checks = 0
    for jid, m in submitted:
        info = srv.job(jid)
        assert info["state"] == STATE_FINISHED
        _, k, kind = trace[m]
        n, w, h, spp, bounce, seed = SERVE_TEMPLATES[k]
        sc, handle = built[k]
        got = srv.film(jid)
        if kind == 0:
            exp = _plain(handle, sc.camera, {"width": w, "height": h, "spp": spp, "bounce": bounce, "seed": seed + m})
            assert np.array_equal(got.view(np.uint32), exp.view(np.uint32)), f"frame arrival {m} differs"
            continue
        rt = WavefrontRuntime(w * h, 1)
        rt.attach(handle)
        rt.set_restir(**SEQ_RESTIR)
        for f, cam in enumerate(walk(k, m)):
            rt.update_frame(cam, w, h, sample_index=f, spp=1, max_bounce=bounce,
                            estimator=_lib.REX_ESTIMATOR_RESTIR, seed=seed + m, walk_mode=WALK_WIDE_COMPRESSED)
            rt.record_frame()
        exp = rt.film()
        rt.destroy()
        assert np.array_equal(got.view(np.uint32), exp.view(np.uint32)), f"sequence arrival {m} differs"
        if oracle_checks < 2:
            orr = oracle.OracleRender(sc.tris, sc.materials, sc.lights, selection=1)
            ors = oracle.OracleRestir(orr, w, h, selection=1, **SEQ_RESTIR)
            for f, cam in enumerate(walk(k, m)):
                ors.render(cam, sample_index=f, seed=seed + m, max_bounce=bounce)
            np.testing.assert_allclose(got, ors.film, rtol=RTOL, atol=ATOL)
            oracle_checks += 1
            print(f"sequence arrival {m}: Film within 1e-4 of the oracle over {SEQ_FRAMES} frames", flush=True)
    assert oracle_checks == 2
    for jid, _ in submitted:
        srv.release(jid)
    srv.destroy()
    for _, h in built:
        h.destroy()
