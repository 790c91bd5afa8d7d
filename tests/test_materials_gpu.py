"""BSDF models 1-4 (GGX conductor with Schlick and complex-IOR Fresnel, GGX
dielectric reflection, smooth glass, rough dielectric), transmission with
interior (Beer-Lambert) attenuation and delta paths on the GPU against the
oracle's one-sample integrator (whose BSDF is pinned against the reference's
compiled double-precision ReferenceBsdfL6).

Bar: the five per-sample signals within 1e-4 relative (atol 1e-6), queue
sizes equal, and the Film over two launches within 1e-4.
"""
import numpy as np
import pytest

from renderingengine_amd import _lib, scenes
from renderingengine_amd.batch import BatchEngine, ShadedSceneHandle, frame_desc
from renderingengine_amd.render import WALK_WIDE_COMPRESSED, WavefrontRuntime

pytestmark = pytest.mark.gpu

RTOL = 1.0e-4
ATOL = 1.0e-6


def _material_room(variant):
    """The ReSTIR test room with its four surfaces' materials replaced:
    variant 0 glass boxes (smooth, attenuating; rough) on a complex-IOR
    conductor floor and a dielectric-reflection wall; variant 1 a Schlick
    conductor box, a rough-dielectric wall seen through, transmissive
    metallic-roughness floor."""
    sc = scenes.restir_room(emitters=16, width=96, height=54, seed=3)
    mats = sc.materials.copy()
    if variant == 0:
        mats[0] = scenes.conductor_material((0.9, 0.6, 0.5), roughness=0.35, eta=(0.2, 0.9, 1.1), k=(3.9, 2.4, 2.3))
        mats[1] = scenes.dielectric_material(scenes.BSDF_MODEL_GGX_DIELECTRIC_REFLECTION, ior=1.7, roughness=0.25)
        mats[2] = scenes.dielectric_material(scenes.BSDF_MODEL_SMOOTH_GLASS, ior=1.5, transmission=1.0,
                                             attenuation=(0.8, 0.5, 0.3), attenuation_distance=0.7)
        mats[3] = scenes.dielectric_material(scenes.BSDF_MODEL_ROUGH_DIELECTRIC, ior=1.33, roughness=0.3,
                                             transmission=0.9, attenuation=(0.6, 0.9, 0.7), attenuation_distance=1.5)
    else:
        mats[0] = scenes.material((0.6, 0.6, 0.6), roughness=0.6)
        mats[0].view(np.float32)[8] = 0.3          # transmission weight of a metallic-roughness floor
        mats[1] = scenes.dielectric_material(scenes.BSDF_MODEL_ROUGH_DIELECTRIC, ior=1.5, roughness=0.005)  # -> smooth
        mats[2] = scenes.conductor_material((0.95, 0.8, 0.4), roughness=0.2)
        mats[3] = scenes.dielectric_material(scenes.BSDF_MODEL_SMOOTH_GLASS, ior=2.2, transmission=0.6)
    out = scenes.ShadedScene(sc.tris, mats, sc.lights, sc.camera, f"material_room_{variant}")
    return scenes.with_lights(out, [scenes.environment_light((0.4, 0.45, 0.5))])


def _check(oracle, rt, orr, sc, w, h, spp, bounce, est, seed):
    rt.update_frame(sc.camera, w, h, sample_index=0, spp=spp, max_bounce=bounce, estimator=est, seed=seed,
                    walk_mode=WALK_WIDE_COMPRESSED, keep_signals=True)
    rt.record_frame()
    stats = rt.collect_completed_frame()
    fr = orr.frame(sc.camera, w, h, spp=spp, max_bounce=bounce, estimator=est, seed=seed)
    res = orr.render(fr, order=2)
    exp = oracle.signals_of(res)
    got = rt.signals()
    bad = ~np.all(np.isclose(got, exp, rtol=RTOL, atol=ATOL), axis=(1, 2))
    exact = np.mean(np.all(got.view(np.uint32) == exp.view(np.uint32), axis=(1, 2)))
    assert not bad.any(), (f"{bad.sum()} of {len(bad)} samples differ (exact {exact:.4f}); first path "
                           f"{np.nonzero(bad)[0][:5]}: got {got[bad][0]} expected {exp[bad][0]}")
    assert stats["closestRays"] == int(res["bounces"].sum())
    assert stats["shadowRays"] == int(res["shadowRays"].sum())
    return res, exact


@pytest.mark.parametrize("variant", [0, 1])
def test_material_models_signals(oracle, variant):
    sc = _material_room(variant)
    w, h = 96, 54
    rt = WavefrontRuntime(w * h, 2)
    rt.set_scene(sc.tris, sc.materials, sc.lights)
    orr = oracle.OracleRender(sc.tris, sc.materials, sc.lights, selection=1)
    exacts = []
    for k, est in enumerate((_lib.REX_ESTIMATOR_BSDF_ONLY, _lib.REX_ESTIMATOR_NEE, _lib.REX_ESTIMATOR_MIS)):
        print(f"variant {variant} estimator {est}", flush=True)
        res, exact = _check(oracle, rt, orr, sc, w, h, 2, 8, est, seed=31 + k)
        exacts.append(exact)
        # paths continue through the specular and transmissive surfaces
        assert (res["bounces"] >= 3).sum() > 500
    print(f"exact-match fractions {exacts}", flush=True)
    rt.destroy()


def test_material_film_and_batch(oracle):
    """Film of two MIS launches against the oracle, and the same frames
    through the batch engine bit-identical to the frames alone."""
    sc = _material_room(0)
    w, h, spp = 64, 36, 2
    rt = WavefrontRuntime(w * h, spp)
    rt.set_scene(sc.tris, sc.materials, sc.lights)
    orr = oracle.OracleRender(sc.tris, sc.materials, sc.lights, selection=1)
    film = None
    for k in range(2):
        rt.update_frame(sc.camera, w, h, sample_index=k * spp, spp=spp, max_bounce=8,
                        estimator=_lib.REX_ESTIMATOR_MIS, seed=5, walk_mode=WALK_WIDE_COMPRESSED)
        rt.record_frame()
        fr = orr.frame(sc.camera, w, h, spp=spp, max_bounce=8, sample_index=k * spp, estimator=2, seed=5)
        film = oracle.film_accumulate(orr.render(fr, order=2), w * h, spp, film)
    np.testing.assert_allclose(rt.film(), film, rtol=RTOL, atol=ATOL)
    rt.destroy()

    handles = [ShadedSceneHandle(s.tris, s.materials, s.lights) for s in (sc, _material_room(1))]
    eng = BatchEngine(path_capacity=1 << 17, max_jobs=8, use_graphs=True)
    idx = [eng.add_scene(hd) for hd in handles]
    jobs = [(k % 2, eng.submit(idx[k % 2], frame_desc(sc.camera, w, h, spp=spp, max_bounce=8, seed=40 + k,
                                                      estimator=_lib.REX_ESTIMATOR_MIS))) for k in range(4)]
    eng.run()
    for k, (s, jid) in enumerate(jobs):
        alone = WavefrontRuntime(w * h, spp)
        alone.attach(handles[s])
        alone.update_frame(sc.camera, w, h, spp=spp, max_bounce=8, estimator=_lib.REX_ESTIMATOR_MIS, seed=40 + k,
                           walk_mode=WALK_WIDE_COMPRESSED)
        alone.record_frame()
        exp = alone.film()
        alone.destroy()
        assert np.array_equal(eng.film(jid).view(np.uint32), exp.view(np.uint32)), f"job {k}"
    eng.destroy()
    for hd in handles:
        hd.destroy()
