"""Benchmark of the rex hot path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME]

Prints one JSON line (the driver's contract).  Workloads:

  trace_bounce_1m  -- one step = one closest-hit launch of the integrator's
                      traversal walk over 1920x1080x4 = 8,294,400 incoherent
                      bounce rays (origins on random surface points of the
                      1M-triangle soup, cosine-distributed directions) with the
                      scene's BVH and triangles resident in HBM.

`value` is Grays/s = rays traced / wall time of the timed steps.  The roofline
object prices the traversal kernel with the algorithmic byte model
48*N_node + 48*N_tri + 64 + 32 bytes per ray (N_node = slab tests, N_tri =
triangle tests, counted on the same rays), divided by the kernel's average
duration from HIP events recorded on the launch stream.  cpu_baseline times
the C oracle's walk on a bounded sample of the same rays on the host cores.
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


class HipEvents:
    """hipEvent timing on the library's launch stream (the default stream)."""

    def __init__(self):
        self.hip = ctypes.CDLL("libamdhip64.so")
        self.start = ctypes.c_void_p()
        self.stop = ctypes.c_void_p()
        assert self.hip.hipEventCreate(ctypes.byref(self.start)) == 0
        assert self.hip.hipEventCreate(ctypes.byref(self.stop)) == 0

    def record_start(self, stream=None):
        assert self.hip.hipEventRecord(self.start, ctypes.c_void_p(stream)) == 0

    def record_stop(self, stream=None):
        assert self.hip.hipEventRecord(self.stop, ctypes.c_void_p(stream)) == 0

    def elapsed_ms(self):
        assert self.hip.hipEventSynchronize(self.stop) == 0
        ms = ctypes.c_float()
        assert self.hip.hipEventElapsedTime(ctypes.byref(ms), self.start, self.stop) == 0
        return ms.value

    def sync(self):
        assert self.hip.hipDeviceSynchronize() == 0


def bounce_rays(tris, n, seed):
    """Incoherent secondary rays: origins on random points of random
    triangles, offset 1e-4 along the face normal toward a random side,
    cosine-weighted directions about that side's normal."""
    rng = np.random.default_rng(seed)
    pos = np.ascontiguousarray(tris).view(np.float32)[:, 0:12].reshape(-1, 3, 4)[:, :, :3]
    out = np.empty((n, 8), dtype=np.float32)
    chunk = 1 << 20
    for s in range(0, n, chunk):
        m = min(chunk, n - s)
        pick = rng.integers(0, pos.shape[0], size=m)
        p = pos[pick].astype(np.float64)
        b = rng.dirichlet([1.0, 1.0, 1.0], size=m)
        x = np.einsum('nk,nkj->nj', b, p)
        nrm = np.cross(p[:, 1] - p[:, 0], p[:, 2] - p[:, 0])
        nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
        nrm *= np.where(rng.random(m) < 0.5, -1.0, 1.0)[:, None]
        # cosine hemisphere
        u1, u2 = rng.random(m), rng.random(m)
        r, phi = np.sqrt(u1), 2 * np.pi * u2
        lx, ly, lz = r * np.cos(phi), r * np.sin(phi), np.sqrt(1 - u1)
        t = np.where(np.abs(nrm[:, :1]) > 0.9, np.array([[0.0, 1.0, 0.0]]), np.array([[1.0, 0.0, 0.0]]))
        tx = np.cross(t, nrm)
        tx /= np.linalg.norm(tx, axis=1, keepdims=True)
        ty = np.cross(nrm, tx)
        d = tx * lx[:, None] + ty * ly[:, None] + nrm * lz[:, None]
        d = d.astype(np.float32)
        d = d / np.sqrt((d * d).sum(axis=1, keepdims=True, dtype=np.float32))
        out[s:s + m, 0:3] = (x + nrm * 1e-4).astype(np.float32)
        out[s:s + m, 3] = 1e-4
        out[s:s + m, 4:7] = d
        out[s:s + m, 7] = 1e30
    return out


def cpu_baseline(tris, rays, sample, threads):
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib
    ob = oracle_lib.OracleBvh(tris)
    sub = np.ascontiguousarray(rays[:sample])
    parts = np.array_split(np.arange(sample), threads * 4)

    def work(idx):
        return ob.walk(sub[idx])

    ob.walk(sub[:1000])  # warm
    t0 = time.perf_counter()
    with ThreadPoolExecutor(max_workers=threads) as ex:
        list(ex.map(work, parts))
    dt = time.perf_counter() - t0
    return sample / dt, dt


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="trace_bounce_1m")
    ap.add_argument("--triangles", type=int, default=1_000_000)
    ap.add_argument("--mode", type=int, default=2, help="0 left-first short stack, 1 stackless, 2 near-first")
    ap.add_argument("--cpu-sample", type=int, default=400_000)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1 "
                                   "(multi-GPU sharding is described in DESIGN.md, not run)"}))
        sys.exit(2)

    from renderingengine_amd import scenes, traversal

    t0 = time.perf_counter()
    tris = scenes.triangle_soup(args.triangles, seed=0)
    width, height, spp = 1920, 1080, 4
    nrays = width * height * spp
    rays = bounce_rays(tris, nrays, seed=1)
    print(f"[bench] scene + rays generated in {time.perf_counter() - t0:.1f}s", file=sys.stderr, flush=True)

    be = traversal.FlattenedSahBackend()
    st = be.build_or_update_scene(tris, fingerprint=0xC2, generation=1)
    assert st, st.message
    info = be.scene.info()
    print(f"[bench] scene built: {info.nodeCount} nodes, depth {info.maximumDepth}, "
          f"{info.buildMilliseconds:.0f} ms host build", file=sys.stderr, flush=True)
    d_rays = traversal.DeviceBuffer.from_array(rays)
    d_hits = traversal.DeviceBuffer(nrays * 72)
    d_cnt = traversal.DeviceBuffer(80)
    d_cnt.zero()

    ev = HipEvents()
    # counting pass (untimed): slab and triangle tests on these rays
    be.trace_walk(d_rays, None, nrays, False, d_hits, mode=args.mode, counters=d_cnt)
    ev.sync()
    cnt = d_cnt.download(np.uint64, (10,))
    box_tests, tri_tests = int(cnt[9]), int(cnt[1])
    bytes_per_launch = 48 * box_tests + 48 * tri_tests + (64 + 32) * nrays

    for _ in range(args.warmup):
        be.trace_walk(d_rays, None, nrays, False, d_hits, mode=args.mode)
    ev.sync()
    t1 = time.perf_counter()
    ev.record_start()
    for _ in range(args.steps):
        be.trace_walk(d_rays, None, nrays, False, d_hits, mode=args.mode)
    ev.record_stop()
    ev.sync()
    wall = time.perf_counter() - t1
    kernel_ms = ev.elapsed_ms() / args.steps
    hits = d_hits.download(traversal.WALK_HIT_DTYPE, (nrays,))
    hit_rate = float((hits["result"] == 1).mean())

    grays = nrays * args.steps / wall / 1e9
    achieved = bytes_per_launch / (kernel_ms * 1e-3) / 1e9

    cpu = None
    if not args.no_cpu:
        threads = min(16, os.cpu_count() or 1)
        rate, dt = cpu_baseline(tris, rays, args.cpu_sample, threads)
        cpu = {"value": rate / 1e9, "unit": "Grays/s", "cores": threads, "kind": "port",
               "sample": f"first {args.cpu_sample} of the same bounce rays through the C oracle's walk "
                         f"({dt:.1f}s wall on {threads} threads)"}

    out = {
        "metric": "Grays/s on 1M-tri 1080p scene (incoherent bounce rays, closest hit)",
        "value": round(grays, 4), "unit": "Grays/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(wall / args.steps * 1e3, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": args.workload, "triangles": args.triangles, "rays_per_step": nrays,
                   "film": f"{width}x{height}x{spp}spp", "walk": {0: "left_first_short_stack", 1: "stackless", 2: "near_first"}[args.mode],
                   "bvh_nodes": info.nodeCount, "bvh_depth": info.maximumDepth, "hit_rate": round(hit_rate, 4)},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                     "kernel": "k_walk (closest)", "kernel_ms": round(kernel_ms, 4),
                     "bytes_per_launch": bytes_per_launch,
                     "slab_tests_per_ray": round(box_tests / nrays, 3),
                     "tri_tests_per_ray": round(tri_tests / nrays, 3)},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
