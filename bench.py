"""Benchmark of the rex hot path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME]

Prints one JSON line (the driver's contract).  Workloads:

  frame_1m (default) -- one step = one wavefront frame of the 1M-triangle
                      soup scene at 1920x1080, 4 spp, max bounce 8, MIS
                      (RayGen, per bounce Intersect / Shade / TraceShadow,
                      Resolve + Film), scene BVH, triangles, materials and
                      light tables resident in HBM.  Frames advance the
                      sample index (progressive Film).
  trace_bounce_1m  -- one step = one closest-hit walk launch over 8,294,400
                      incoherent bounce rays of the same triangle soup.
  cornell          -- one step = one 256x256, 64 spp, 8-bounce frame of the
                      reference CPU path tracer's Cornell fixture; its
                      cpu_baseline is the oracle's bit-exact restatement of
                      that tracer on the whole image, beside the reference's
                      published 10.17 Mrays/s.
  restir_1m, batch64, serve, reconstruction -- see their functions.

`value` is Grays/s = traversal queries (closest-hit + any-hit, counted on
the device) / wall time of the timed steps.  The roofline object prices the
dominant kernel (the queue traversal kernel: closest-hit rays and any-hit
shadow queries) with the algorithmic byte model 48*N_node + 48*N_tri + 64 + 32
bytes per query (N_node = slab tests, N_tri = triangle tests, counted by the
device on a counting frame of the same workload), divided by that kernel's
GPU time from HIP events recorded on the launch stream during the timed
frames.
cpu_baseline times the C oracle's per-path integrator (or walk) on a bounded
sample of the same workload on the host cores.
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import re
import sqlite3
import sys
import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# VALU of one MI355X: 256 CUs x 4 SIMDs, 16 lanes each (a wave64 instruction
# takes 4 cycles), 2.4 GHz peak engine clock -> 39.3 T lane-operations / s.
SIMDS = 1024
SIMD_LANES = 16
CLOCK_HZ = 2.4e9
VALU_PEAK_LANE_OPS = SIMDS * SIMD_LANES * CLOCK_HZ
FRAME_COUNTERS = os.path.join(ROOT, "profiles", "frame_1m_1080p_4spp_b8_counters.md")
RESTIR_COUNTERS = os.path.join(ROOT, "profiles", "restir_1m_10k_1080p_counters.md")
# The ReSTIR passes of a frame: the reservoir kernels and the visibility
# queries (the any-hit ABI-v1 instantiation of the traversal kernel).
RESTIR_KERNELS = ("k_restir_", "k_trace_dense<5, 3, true, false>")


# ---- rocprofv3 summaries -----------------------------------------------------

def _dispatches(db):
    """(dispatch id, kernel name, duration ns) in start order from a rocpd
    database of `rocprofv3 --kernel-trace`."""
    c = sqlite3.connect(db)
    return c.execute("select dispatch_id, name, duration from kernels order by start").fetchall()


def _counters(db):
    """{dispatch id: {counter: value summed over dimensions}}, {dispatch: kernel}."""
    c = sqlite3.connect(db)
    vals, names = {}, {}
    for d, k, cn, v in c.execute("select dispatch_id, kernel_name, counter_name, value from counters_collection"):
        vals.setdefault(d, {})
        vals[d][cn] = vals[d].get(cn, 0.0) + v
        names[d] = k
    return vals, names


def _frame_launches(seq):
    """Dispatch ids of the first plain frame's traversal launches after the
    counting frame: from the first production closest-hit launch of bounce 0
    to the next production shadow launch."""
    out, on = [], False
    for d, name in seq:
        if not on and "k_trace_dense<2, 3, false, false>" in name:
            on = True
        if on and "k_trace" in name:
            out.append(d)
            if "k_trace_dense<3, 3, true, false>" in name:
                break
    return out


def summarize_profiles(args):
    """Writes the counter summary that the frame workload's roofline reads:
    the traversal launches of the first timed plain frame in each rocprofv3
    pass (FETCH_SIZE, WRITE_SIZE, SQ VALU counters, TCC hit / miss), the copy
    probe's FETCH_SIZE / WRITE_SIZE for scaling, and the kernel-trace
    durations."""
    launches = {}
    probe = {}
    for key, db in (("fetch", args.fetch_db), ("write", args.write_db), ("sq", args.sq_db), ("tcc", args.tcc_db)):
        vals, names = _counters(db)
        seq = sorted(names.items())
        ids = _frame_launches(seq)
        launches[key] = [vals[d] for d in ids]
        for d, name in seq:
            if "copyBuffer" in name and vals[d].get("FETCH_SIZE", 0) + vals[d].get("WRITE_SIZE", 0) > 1e5:
                probe.update(vals[d])
                break
    kt = _dispatches(args.kt_db)
    ids = _frame_launches([(d, n) for d, n, _ in kt])
    dur = {d: t for d, _, t in kt}
    names = {d: n for d, n, _ in kt}
    n = len(launches["fetch"])
    assert all(len(v) == n for v in launches.values()) and len(ids) == n, {k: len(v) for k, v in launches.items()}
    probe_kib = args.copy_probe_mib * 1024
    block = {
        "workload": "frame_1m_1080p_4spp_b8 (first timed plain frame)",
        "launches": n,
        "kernels": [names[d].split("(")[0].replace("void rex::", "") for d in ids],
        "kernel_trace_us": [round(dur[d] / 1e3, 2) for d in ids],
        "fetch_kib": [round(v["FETCH_SIZE"]) for v in launches["fetch"]],
        "write_kib": [round(v["WRITE_SIZE"]) for v in launches["write"]],
        "copy_probe_kib": probe_kib,
        "fetch_factor": round(probe_kib / probe.get("FETCH_SIZE", probe_kib), 4),
        "write_factor": round(probe_kib / probe.get("WRITE_SIZE", probe_kib), 4),
        "sq_insts_valu": [round(v["SQ_INSTS_VALU"]) for v in launches["sq"]],
        "sq_active_inst_valu": [round(v["SQ_ACTIVE_INST_VALU"]) for v in launches["sq"]],
        "sq_thread_cycles_valu": [round(v["SQ_THREAD_CYCLES_VALU"]) for v in launches["sq"]],
        "sq_busy_cycles": [round(v["SQ_BUSY_CYCLES"]) for v in launches["sq"]],
        "tcc_hit": [round(v["TCC_HIT_sum"]) for v in launches["tcc"]],
        "tcc_miss": [round(v["TCC_MISS_sum"]) for v in launches["tcc"]],
    }
    fetch_b = sum(block["fetch_kib"]) * 1024 * block["fetch_factor"]
    write_b = sum(block["write_kib"]) * 1024 * block["write_factor"]
    hit, miss = sum(block["tcc_hit"]), sum(block["tcc_miss"])
    lanes = sum(block["sq_thread_cycles_valu"]) / max(sum(block["sq_active_inst_valu"]), 1)
    t_s = sum(block["kernel_trace_us"]) * 1e-6
    issue = sum(block["sq_insts_valu"]) * 4 / (SIMDS * CLOCK_HZ * t_s)
    lines = [
        "# Traversal counters of the 1M-triangle 1080p frame (4 spp, 8 bounces, MIS)",
        "",
        "Traversal launches of the first timed plain frame (bounce 0's closest-hit",
        "rays, one launch per bounce holding its shadow queries and the next bounce's",
        "rays, the last bounce's shadow queries), one rocprofv3 pass per counter set.",
        "Commands (repo root, one MI355X, `export TMPDIR=/tmp`):",
        "",
        "    rocprofv3 --kernel-trace --stats -d gpurun_out/kt -o kt -- python3 -u bench.py --steps 3 --warmup 1 --no-cpu",
        f"    rocprofv3 --pmc FETCH_SIZE -d gpurun_out/pf -o pf -- python3 -u bench.py --steps 1 --warmup 0 --no-cpu --copy-probe-mib {args.copy_probe_mib}",
        f"    rocprofv3 --pmc WRITE_SIZE -d gpurun_out/pw -o pw -- python3 -u bench.py --steps 1 --warmup 0 --no-cpu --copy-probe-mib {args.copy_probe_mib}",
        "    rocprofv3 --pmc SQ_WAVES SQ_INSTS_VALU SQ_ACTIVE_INST_VALU SQ_THREAD_CYCLES_VALU SQ_BUSY_CYCLES SQ_WAVE_CYCLES SQ_INSTS_VMEM_RD SQ_INSTS_LDS -d gpurun_out/psq -o psq -- python3 -u bench.py --steps 1 --warmup 0 --no-cpu",
        "    rocprofv3 --pmc TCC_HIT_sum TCC_MISS_sum GRBM_GUI_ACTIVE -d gpurun_out/ptcc -o ptcc -- python3 -u bench.py --steps 1 --warmup 0 --no-cpu",
        "    python3 bench.py --summarize-profiles --kt-db gpurun_out/kt/kt_results.db --fetch-db gpurun_out/pf/pf_results.db \\",
        "        --write-db gpurun_out/pw/pw_results.db --sq-db gpurun_out/psq/psq_results.db --tcc-db gpurun_out/ptcc/ptcc_results.db",
        "",
        f"Copy probe ({args.copy_probe_mib} MiB read and written by `__amd_rocclr_copyBuffer`): FETCH_SIZE factor "
        f"{block['fetch_factor']}, WRITE_SIZE factor {block['write_factor']} (bytes = KiB x 1024 x factor).",
        "",
        "| launch | kernel | trace us | FETCH KiB | WRITE KiB | L2 hit | VALU insts | lanes active |",
        "|---|---|---:|---:|---:|---:|---:|---:|",
    ]
    for i in range(n):
        h, m = block["tcc_hit"][i], block["tcc_miss"][i]
        la = block["sq_thread_cycles_valu"][i] / max(block["sq_active_inst_valu"][i], 1) / 64
        lines.append(f"| {i} | `{block['kernels'][i]}` | {block['kernel_trace_us'][i]} | {block['fetch_kib'][i]:,} | "
                     f"{block['write_kib'][i]:,} | {h / max(h + m, 1):.2f} | {block['sq_insts_valu'][i]:,} | {la:.3f} |")
    lines += [
        "",
        f"Frame: {(fetch_b + write_b) / 1e9:.2f} GB of L2-to-fabric traffic ({(fetch_b + write_b) / n / 1e9:.2f} GB per "
        f"launch; FETCH_SIZE and WRITE_SIZE count MALL hits as well as DRAM), L2 hit rate {hit / max(hit + miss, 1):.3f}, "
        f"{lanes / 64:.3f} of the lanes active per VALU instruction, VALU issue "
        f"{issue:.3f} of {SIMDS} SIMDs x one wave64 instruction per 4 cycles at {CLOCK_HZ / 1e9:.1f} GHz over the "
        f"kernel-trace time.",
        "",
        "```json",
        json.dumps(block),
        "```",
    ]
    text = "\n".join(lines) + "\n"
    if args.summarize_out:
        with open(args.summarize_out, "w") as f:
            f.write(text)
    print(text)
    if args.trace_out:
        with open(args.trace_out, "w") as f:
            f.write(kernel_trace_table(args.kt_db, "bench.py --steps 3 --warmup 1 --no-cpu"))


def _short_kernel(name):
    return name.split("(")[0].replace("void rex::", "").replace("rex::", "")


def summarize_restir_profiles(args):
    """Writes the ReSTIR workload's counter summary: per ReSTIR kernel of
    the profiled frames, VALU instructions, active-lane cycles and
    FETCH_SIZE / WRITE_SIZE (scaled by a copy probe), plus the kernel-trace
    average duration per call."""
    per = {}
    probe = {}
    calls = {}
    for key, db in (("fetch", args.fetch_db), ("write", args.write_db), ("sq", args.sq_db)):
        vals, names = _counters(db)
        for d, name in sorted(names.items()):
            if "copyBuffer" in name and vals[d].get("FETCH_SIZE", 0) + vals[d].get("WRITE_SIZE", 0) > 1e5:
                probe.update(vals[d])
                continue
            k = _short_kernel(name)
            if not k.startswith(RESTIR_KERNELS):
                continue
            e = per.setdefault(k, {})
            for cn, v in vals[d].items():
                e[cn] = e.get(cn, 0.0) + v
            if key == "sq":
                calls[k] = calls.get(k, 0) + 1
    dur = {}
    for _, name, t in _dispatches(args.kt_db):
        k = _short_kernel(name)
        if k.startswith(RESTIR_KERNELS):
            dur.setdefault(k, []).append(t)
    names = sorted(per, key=lambda k: -per[k].get("SQ_THREAD_CYCLES_VALU", 0))
    frames = max(calls.values())
    probe_kib = args.copy_probe_mib * 1024
    block = {
        "workload": "restir_1m_10k_1080p (biased estimator), per frame",
        "frames_profiled": frames,
        "kernels": names,
        "calls_per_frame": [round(calls[k] / frames, 3) for k in names],
        "kernel_trace_us_per_call": [round(sum(dur[k]) / len(dur[k]) / 1e3, 2) for k in names],
        "sq_insts_valu": [round(per[k]["SQ_INSTS_VALU"] / frames) for k in names],
        "sq_active_inst_valu": [round(per[k]["SQ_ACTIVE_INST_VALU"] / frames) for k in names],
        "sq_thread_cycles_valu": [round(per[k]["SQ_THREAD_CYCLES_VALU"] / frames) for k in names],
        "fetch_kib": [round(per[k]["FETCH_SIZE"] / frames) for k in names],
        "write_kib": [round(per[k]["WRITE_SIZE"] / frames) for k in names],
        "copy_probe_kib": probe_kib,
        "fetch_factor": round(probe_kib / probe.get("FETCH_SIZE", probe_kib), 4),
        "write_factor": round(probe_kib / probe.get("WRITE_SIZE", probe_kib), 4),
    }
    lines = [
        "# ReSTIR DI passes: counters per frame (restir_1m_10k_1080p)",
        "",
        "The reservoir kernels and the visibility queries of the ReSTIR stage, summed per frame over the",
        "profiled frames; one rocprofv3 pass per counter set (repo root, one MI355X, `export TMPDIR=/tmp`):",
        "",
        "    rocprofv3 --kernel-trace --stats -d gpurun_out/rk -o rk -- python3 -u bench.py --workload restir_1m --steps 2 --warmup 0 --no-cpu",
        f"    rocprofv3 --pmc FETCH_SIZE -d gpurun_out/rf -o rf -- python3 -u bench.py --workload restir_1m --steps 2 --warmup 0 --no-cpu --copy-probe-mib {args.copy_probe_mib}",
        f"    rocprofv3 --pmc WRITE_SIZE -d gpurun_out/rw -o rw -- python3 -u bench.py --workload restir_1m --steps 2 --warmup 0 --no-cpu --copy-probe-mib {args.copy_probe_mib}",
        "    rocprofv3 --pmc SQ_WAVES SQ_INSTS_VALU SQ_ACTIVE_INST_VALU SQ_THREAD_CYCLES_VALU SQ_BUSY_CYCLES SQ_WAVE_CYCLES -d gpurun_out/rs -o rs -- python3 -u bench.py --workload restir_1m --steps 2 --warmup 0 --no-cpu",
        "    python3 bench.py --summarize-restir-profiles --kt-db gpurun_out/rk/rk_results.db --fetch-db gpurun_out/rf/rf_results.db \\",
        "        --write-db gpurun_out/rw/rw_results.db --sq-db gpurun_out/rs/rs_results.db",
        "",
        "| kernel | calls | trace us / call | VALU insts | lanes active | lane-op/s (T) | HBM GB | GB/s |",
        "|---|---:|---:|---:|---:|---:|---:|---:|",
    ]
    for j, k in enumerate(names):
        t = block["kernel_trace_us_per_call"][j] * block["calls_per_frame"][j] * 1e-6
        lanes = block["sq_thread_cycles_valu"][j] / max(block["sq_active_inst_valu"][j], 1) / 64
        gb = (block["fetch_kib"][j] * block["fetch_factor"] + block["write_kib"][j] * block["write_factor"]) * 1024 / 1e9
        lines.append(f"| `{k}` | {block['calls_per_frame'][j]:g} | {block['kernel_trace_us_per_call'][j]} | "
                     f"{block['sq_insts_valu'][j]:,} | {lanes:.3f} | {block['sq_thread_cycles_valu'][j] / t / 1e12:.2f} | "
                     f"{gb:.2f} | {gb / t:.0f} |")
    lines += ["", "```json", json.dumps(block), "```"]
    text = "\n".join(lines) + "\n"
    with open(args.summarize_out if args.summarize_out != FRAME_COUNTERS else RESTIR_COUNTERS, "w") as f:
        f.write(text)
    print(text)


def kernel_trace_table(db, bench_args):
    """The --stats table of a kernel-trace database: calls, average and total
    duration per kernel, largest total first."""
    agg = {}
    for _, name, dur in _dispatches(db):
        a = agg.setdefault(name.split("(")[0].replace("void rex::", "").replace("rex::", ""), [0, 0.0])
        a[0] += 1
        a[1] += dur
    rows = sorted(agg.items(), key=lambda kv: -kv[1][1])
    lines = [
        "# Kernel trace: wavefront frames, 1M-triangle soup, 1920x1080, 4 spp, max bounce 8, MIS, compressed 4-wide walk",
        "",
        "Command (repo root, one MI355X):",
        "",
        "    export TMPDIR=/tmp",
        f"    rocprofv3 --kernel-trace --stats -d gpurun_out/kt -o kt -- python3 -u {bench_args}",
        "",
        "The run renders one counting frame (traversal counters on: the `<..., true>` instantiations), the",
        "warm-up and timed plain frames (`k_raygen`; `k_trace_dense<2,3,...>` for bounce 0's rays; 8 x",
        "`k_shade<false>`; 7 x `k_trace_dense<4,3,...>`, each a bounce's shadow queries plus the next",
        "bounce's rays; `k_trace_dense<3,3,true,...>` for the last bounce's shadow queries; `k_resolve_film`),",
        "then frames in flight through the pooled batch engine (`k_trace_dense<4,3,...>` with per-item",
        "scene roots, `k_shade<true>`, `k_batch_*`).",
        "",
        "| calls | avg (us) | total (ms) | kernel |",
        "|---:|---:|---:|---|",
    ]
    for name, (n, t) in rows:
        lines.append(f"| {n} | {t / n / 1e3:.1f} | {t / 1e6:.3f} | `{name}` |")
    return "\n".join(lines) + "\n"


def counter_block(path=FRAME_COUNTERS):
    """The JSON block of a committed counter summary (None if absent)."""
    try:
        text = open(path).read()
    except OSError:
        return None
    m = re.search(r"```json\n(.*?)\n```", text, re.S)
    return json.loads(m.group(1)) if m else None


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


def trace_bounce(args, ev):
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
        threads = cpu_threads()
        rate, dt = cpu_baseline(tris, rays, args.cpu_sample, threads)
        cpu = {"value": rate / 1e9, "unit": "Grays/s", "cores": threads, "kind": "port",
               "sample": f"first {args.cpu_sample} of the same bounce rays through the C oracle's walk "
                         f"({dt:.1f}s wall on {threads} threads)"}
    walk = {0: "left_first_short_stack", 1: "stackless", 2: "near_first", 3: "near_first_wide4",
            4: "near_first_wide4_compressed"}[args.mode]
    return {
        "metric": "Grays/s on 1M-tri 1080p scene (incoherent bounce rays, closest hit)",
        "value": round(grays, 4), "unit": "Grays/s", "ms_per_step": round(wall / args.steps * 1e3, 3),
        "config": {"workload": "trace_bounce_1m", "triangles": args.triangles, "rays_per_step": nrays,
                   "film": f"{width}x{height}x{spp}spp", "walk": walk,
                   "bvh_nodes": info.nodeCount, "bvh_depth": info.maximumDepth, "hit_rate": round(hit_rate, 4)},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                     "kernel": "k_trace (walk, closest)", "kernel_ms": round(kernel_ms, 4),
                     "bytes_per_launch": bytes_per_launch,
                     "slab_tests_per_ray": round(box_tests / nrays, 3),
                     "tri_tests_per_ray": round(tri_tests / nrays, 3)},
        "cpu_baseline": cpu,
    }


def cpu_threads():
    """Host threads for the CPU baseline: every core this process may use, at
    most 16 (the host share of one GPU on the box)."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def host_cpu():
    """(logical CPUs of the host, CPU model name)."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return os.cpu_count() or 1, model


def frame_cpu_baseline(sc, camera, width, height, spp, bounce, seed, rows):
    """The oracle's per-path integrator with the reference's left-first walk
    (order 0) on `rows` rows of the same frame (same pixels, samples, seed):
    the CPU baseline, and the reference walk's traversal counters on the very
    queries the GPU traced in that frame (slab tests = node records read,
    triangle tests)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib
    orr = oracle_lib.OracleRender(sc.tris, sc.materials, sc.lights, selection=1)
    fr = orr.frame(camera, width, height, spp=spp, max_bounce=bounce, estimator=2, seed=seed)
    threads = cpu_threads()
    begin = (height - rows) // 2
    orr.render_rows(fr, begin, 1, order=0, threads=threads)  # warm
    cnt = np.zeros(10, np.uint64)
    t0 = time.perf_counter()
    res = orr.render_rows(fr, begin, rows, order=0, threads=threads, counters=cnt)
    dt = time.perf_counter() - t0
    nrays = int(res["bounces"].sum() + res["shadowRays"].sum())
    nproc, model = host_cpu()
    cpu = {"value": nrays / dt / 1e9, "unit": "Grays/s", "cores": threads, "kind": "port",
           "host_logical_cpus": nproc, "host_cpu_model": model,
           "sample": f"rows {begin}..{begin + rows - 1} of the same frame ({width * rows * spp} paths, "
                     f"{nrays} traversal queries) through the C oracle's per-path integrator with the "
                     f"reference's left-first walk, {dt:.1f}s wall on {threads} threads"}
    return cpu, {"queries": nrays, "slab_tests": int(cnt[9]), "tri_tests": int(cnt[1]), "rows": rows}


def frame(args, ev):
    from renderingengine_amd import _lib, scenes
    from renderingengine_amd.render import WavefrontRuntime

    width, height, spp, bounce, seed = 1920, 1080, 4, 8, 0
    t0 = time.perf_counter()
    sc = scenes.soup_scene(n=args.triangles, seed=0, width=width, height=height)
    print(f"[bench] scene generated in {time.perf_counter() - t0:.1f}s", file=sys.stderr, flush=True)
    rt = WavefrontRuntime(width * height, spp)
    rt.set_scene(sc.tris, sc.materials, sc.lights, light_selection=_lib.REX_LIGHT_SELECTION_POWER,
                 fingerprint=0xC2, generation=1)
    info = rt.scene.info()
    print(f"[bench] scene built: {info.nodeCount} nodes, depth {info.maximumDepth}, "
          f"{info.buildMilliseconds:.0f} ms host build", file=sys.stderr, flush=True)
    walk = {0: _lib.REX_WALK_SHORT_STACK, 2: _lib.REX_WALK_NEAR_FIRST, 3: _lib.REX_WALK_WIDE,
            4: _lib.REX_WALK_WIDE_COMPRESSED}[args.mode]

    def launch(k):
        rt.update_frame(sc.camera, width, height, sample_index=k * spp, spp=spp, max_bounce=bounce,
                        estimator=_lib.REX_ESTIMATOR_MIS, seed=seed, walk_mode=walk)
        rt.record_frame()

    # Counting frame (untimed): the device walk's own slab / triangle tests.
    rt.set_profiling(counters=True)
    launch(0)
    cstat = rt.collect_completed_frame()
    cc, sc_ = cstat["closestCounters"], cstat["shadowCounters"]
    rt.set_profiling(timing=True)
    rt.invalidate_progressive_film()
    for k in range(args.warmup):
        launch(k)
        rt.collect_completed_frame()
    rt.invalidate_progressive_film()
    ev.sync()
    closest = shadow = primary_rays = 0
    stage = np.zeros(6)
    trace_ms = primary_ms = 0.0
    launches = 0
    t1 = time.perf_counter()
    ev.record_start()
    for k in range(args.steps):
        launch(k)
        st = rt.collect_completed_frame()
        closest += st["closestRays"]
        shadow += st["shadowRays"]
        primary_rays += st["raysPerBounce"][0]
        stage += np.array(st["stageMs"])
        trace_ms += sum(st["traceMs"])
        primary_ms += st["traceMs"][0]
        launches = len(st["traceMs"])
    ev.record_stop()
    ev.sync()
    wall = time.perf_counter() - t1
    gpu_ms = ev.elapsed_ms()
    print(f"[bench] {args.steps} frames: wall {wall * 1e3:.1f} ms, GPU {gpu_ms:.1f} ms, stages {stage.round(2)}",
          file=sys.stderr, flush=True)
    rays = closest + shadow
    plain_grays = rays / wall / 1e9
    # Incoherent queries: closest-hit rays of bounce >= 1 and every shadow
    # query, over the launches after bounce 0's (each launch timed by events
    # on the frame's stream).
    incoherent = closest - primary_rays + shadow
    incoherent_grays = incoherent / ((trace_ms - primary_ms) * 1e-3) / 1e9
    primary_grays = primary_rays / (primary_ms * 1e-3) / 1e9
    trace_ms_frame = trace_ms / args.steps

    cpu, ref_counts = None, None
    if not args.no_cpu:
        cpu, ref_counts = frame_cpu_baseline(sc, sc.camera, width, height, spp, bounce, seed, args.cpu_rows)
    film = rt.film()

    # Frames in flight: the same frames through the pooled engine (a pool of
    # --inflight-pool frames' paths), so one frame's last bounces overlap the
    # next one's first instead of leaving the device to a few thousand rays.
    from renderingengine_amd.batch import BatchEngine, frame_desc
    eng = BatchEngine(path_capacity=args.inflight_pool * width * height * spp, max_jobs=64, use_graphs=True,
                      shade_blocks_per_cu=16)
    eng.add_scene(rt.scene)

    def engine_frames(count, first):
        ids = [eng.submit(0, frame_desc(sc.camera, width, height, sample_index=(first + k) * spp, spp=spp,
                                        max_bounce=bounce, estimator=_lib.REX_ESTIMATOR_MIS, seed=seed))
               for k in range(count)]
        eng.run()
        for j in ids:
            eng.release(j)

    engine_frames(max(args.warmup, 2), 0)
    rays0 = eng.stats()["raysTraced"]
    ev.sync()
    t2 = time.perf_counter()
    engine_frames(args.steps, 0)
    inflight_wall = time.perf_counter() - t2
    inflight_rays = eng.stats()["raysTraced"] - rays0
    eng.destroy()
    rt.destroy()
    inflight_grays = inflight_rays / inflight_wall / 1e9
    print(f"[bench] {args.steps} frames in flight: wall {inflight_wall * 1e3:.1f} ms, "
          f"{inflight_grays:.4f} Grays/s (plain {plain_grays:.4f}); incoherent {incoherent_grays:.4f}, "
          f"primary {primary_grays:.4f} Grays/s", file=sys.stderr, flush=True)

    # Rooflines of the traversal kernel (every launch of a frame together).
    cb = counter_block()
    traffic = None
    valu = None
    if cb is not None and cb.get("launches") == launches:
        traffic = int((sum(cb["fetch_kib"]) * 1024 * cb["fetch_factor"] +
                       sum(cb["write_kib"]) * 1024 * cb["write_factor"]) / launches)
        active = sum(cb["sq_thread_cycles_valu"])
        lane_eff = active / max(sum(cb["sq_active_inst_valu"]), 1) / 64.0
        achieved_ops = active / (trace_ms_frame * 1e-3)
        issue = sum(cb["sq_insts_valu"]) * 4 / (SIMDS * CLOCK_HZ * trace_ms_frame * 1e-3)
        valu = {"achieved": round(achieved_ops / 1e12, 3), "peak": round(VALU_PEAK_LANE_OPS / 1e12, 3),
                "unit": "Tlane-op/s", "frac": round(achieved_ops / VALU_PEAK_LANE_OPS, 4),
                "issue_utilization": round(issue, 4), "lanes_active_per_instruction": round(lane_eff, 4)}
    hbm = None
    if ref_counts is not None and ref_counts["rows"] == height:
        # 48 B per node record and per triangle-position block read by the
        # reference walk, 64 B of ray read and 32 B of hit written per query
        # (frame 0's queries, counted by the oracle on the same rays).
        alg = 48 * ref_counts["slab_tests"] + 48 * ref_counts["tri_tests"] + 96 * ref_counts["queries"]
        achieved = alg / (trace_ms_frame * 1e-3) / 1e9
        hbm = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
               "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
               "bytes_per_launch": int(alg // launches),
               "slab_tests_per_query": round(ref_counts["slab_tests"] / ref_counts["queries"], 3),
               "tri_tests_per_query": round(ref_counts["tri_tests"] / ref_counts["queries"], 3),
               "measured_traffic_gbs": round(traffic / (trace_ms_frame / launches * 1e-3) / 1e9, 1)
               if traffic else None}
    roofline = {"bound": "valu", "achieved": valu["achieved"] if valu else None,
                "peak": round(VALU_PEAK_LANE_OPS / 1e12, 3), "unit": "Tlane-op/s",
                "frac": valu["frac"] if valu else None, "traffic": traffic,
                "kernel": "k_trace (queue traversal: closest-hit rays + any-hit shadow queries)",
                "launches_per_frame": launches, "kernel_ms_per_launch": round(trace_ms_frame / launches, 4),
                "valu": valu, "hbm": hbm,
                "device_slab_tests_per_query": round((cc[9] + sc_[9]) / max(cstat["closestRays"] + cstat["shadowRays"], 1), 3),
                "device_tri_tests_per_query": round((cc[1] + sc_[1]) / max(cstat["closestRays"] + cstat["shadowRays"], 1), 3)}
    return {
        "metric": "Grays/s on 1M-tri 1080p scene",
        "value": round(max(inflight_grays, plain_grays), 4), "unit": "Grays/s",
        "ms_per_step": round(min(inflight_wall, wall) / args.steps * 1e3, 3),
        "config": {"workload": "frame_1m_1080p_4spp_b8", "triangles": args.triangles,
                   "form": "frames_in_flight" if inflight_grays > plain_grays else "plain_per_frame",
                   "frames_in_flight_grays_per_s": round(inflight_grays, 4),
                   "frames_in_flight_pool_frames": args.inflight_pool,
                   "plain_grays_per_s": round(plain_grays, 4),
                   "incoherent_bounce_grays_per_s": round(incoherent_grays, 4),
                   "incoherent_bounce_queries_per_frame": incoherent // args.steps,
                   "primary_grays_per_s": round(primary_grays, 4),
                   "frames_in_flight_rays": int(inflight_rays), "plain_rays": int(rays),
                   "film": f"{width}x{height}", "spp": spp, "max_bounce": bounce, "estimator": "mis",
                   "light_selection": "power",
                   "walk": {0: "left_first_short_stack", 2: "near_first", 3: "near_first_wide4",
                            4: "near_first_wide4_compressed"}[args.mode],
                   "bvh_nodes": info.nodeCount, "bvh_depth": info.maximumDepth,
                   "closest_rays_per_frame": closest // args.steps, "shadow_rays_per_frame": shadow // args.steps,
                   "paths_per_s": round(width * height * spp * args.steps / wall, 1),
                   "gpu_ms_per_frame": round(gpu_ms / args.steps, 3),
                   "trace_ms_per_frame": round(trace_ms_frame, 3),
                   "stage_ms_per_frame": {n: round(v / args.steps, 3) for n, v in
                                          zip(("raygen", "intersect", "shade", "shadow", "resolve"), stage[:5])},
                   "film_mean": [round(float(x), 5) for x in film[:, :3].mean(axis=0)]},
        "roofline": roofline,
        "cpu_baseline": cpu,
    }


def restir(args, ev):
    """ReSTIR DI frames of the 1M-triangle soup with 10,000 emissive
    triangles at 1920x1080: 32 candidates per pixel, temporal reuse with M
    cap 20, 5 spatial neighbours in a 30-pixel disk, winner visibility by
    any-hit, along a seeded 0.02-unit/frame camera walk (one step = one
    frame: primary rays, bounce shading, the ReSTIR passes, Resolve + Film)."""
    from renderingengine_amd import _lib, scenes
    from renderingengine_amd.render import WavefrontRuntime

    width, height, bounce, seed = 1920, 1080, args.restir_bounces, 0
    t0 = time.perf_counter()
    sc = scenes.restir_scene(n=args.triangles, seed=0, emitters=10_000, width=width, height=height)
    print(f"[bench] scene generated in {time.perf_counter() - t0:.1f}s", file=sys.stderr, flush=True)
    rt = WavefrontRuntime(width * height, 1)
    rt.set_scene(sc.tris, sc.materials, sc.lights, light_selection=_lib.REX_LIGHT_SELECTION_POWER,
                 fingerprint=0xC3, generation=1)
    rt.set_restir(candidates=32, neighbors=5, max_m=20, max_age=20, temporal=True,
                  neighbor_mode=_lib.REX_RESTIR_NEIGHBORS_DISK, radius=30,
                  estimator_mode=_lib.REX_RESTIR_ESTIMATOR_REFERENCE_CORRECTION if args.restir_reference
                  else _lib.REX_RESTIR_ESTIMATOR_BIASED)
    frames = args.warmup + args.steps
    cams = scenes.camera_walk(sc.camera, frames, step=0.02, seed=1)
    rt.set_profiling(timing=True)

    def launch(k):
        rt.update_frame(cams[k], width, height, sample_index=k, spp=1, max_bounce=bounce,
                        estimator=_lib.REX_ESTIMATOR_RESTIR, seed=seed, walk_mode=_lib.REX_WALK_WIDE_COMPRESSED)
        rt.record_frame()
        return rt.collect_completed_frame()

    for k in range(args.warmup):
        launch(k)
    ev.sync()
    stage = np.zeros(6)
    queries = 0
    t1 = time.perf_counter()
    ev.record_start()
    for k in range(args.warmup, frames):
        st = launch(k)
        stage += np.array(st["stageMs"])
        queries += st["closestRays"] + st["shadowRays"]
    ev.record_stop()
    ev.sync()
    wall = time.perf_counter() - t1
    state = rt.restir_state()
    stats = state["stats"]
    visibility = int(stats["visibilityCounts"][0])
    rt.destroy()
    n = width * height
    restir_ms = stage[5] / args.steps
    # Roofline of the ReSTIR stage (VALU-bound): active-lane VALU operations
    # per frame from the committed counter summary over the stage time
    # measured here; k_restir_initial alone over its kernel-trace time.
    roofline = None
    cb = counter_block(RESTIR_COUNTERS)
    if cb is not None and not args.restir_reference and restir_ms > 0:
        lane_ops = sum(cb["sq_thread_cycles_valu"])
        traffic = int(sum(k * 1024 * cb["fetch_factor"] for k in cb["fetch_kib"]) +
                      sum(k * 1024 * cb["write_factor"] for k in cb["write_kib"]))
        achieved = lane_ops / (restir_ms * 1e-3)
        roofline = {"bound": "valu", "achieved": round(achieved / 1e12, 3),
                    "peak": round(VALU_PEAK_LANE_OPS / 1e12, 3), "unit": "Tlane-op/s",
                    "frac": round(achieved / VALU_PEAK_LANE_OPS, 4), "traffic": traffic,
                    "kernel": "ReSTIR stage (initial, temporal, spatial, visibility, resolve)",
                    "stage_ms_per_frame": round(restir_ms, 3),
                    "measured_traffic_gbs": round(traffic / (restir_ms * 1e-3) / 1e9, 1)}
        if "k_restir_initial<false>" in cb["kernels"]:
            j = cb["kernels"].index("k_restir_initial<false>")
            t = cb["kernel_trace_us_per_call"][j] * cb["calls_per_frame"][j] * 1e-6
            ops = cb["sq_thread_cycles_valu"][j]
            roofline["initial"] = {"achieved": round(ops / t / 1e12, 3), "frac": round(ops / t / VALU_PEAK_LANE_OPS, 4),
                                   "lanes_active_per_instruction":
                                       round(ops / max(cb["sq_active_inst_valu"][j], 1) / 64, 4),
                                   "kernel_ms": round(t * 1e3, 3)}
    return {
        "metric": "ReSTIR DI frames/s on 1M-tri + 10k-emitter 1080p scene",
        "value": round(args.steps / wall, 3), "unit": "frames/s", "ms_per_step": round(wall / args.steps * 1e3, 3),
        "config": {"workload": "restir_1m_10k_1080p", "triangles": args.triangles, "emitters": 10_000,
                   "film": f"{width}x{height}", "candidates": 32, "neighbors": 5, "neighbor_mode": "disk30",
                   "max_m": 20, "max_bounce": bounce, "camera_walk": "0.02 units/frame",
                   "estimator": "reference_correction" if args.restir_reference else "biased",
                   "restir_ms_per_frame": round(restir_ms, 3),
                   "stage_ms_per_frame": {n_: round(v / args.steps, 3) for n_, v in
                                          zip(("raygen", "intersect", "shade", "shadow", "resolve", "restir"),
                                              stage)},
                   "traversal_queries_per_frame": queries // args.steps,
                   "visibility_rays_last_frame": visibility,
                   "temporal_accepted_last_frame": int(stats["reuseCounts"][1]),
                   "spatial_accepted_last_frame": int(stats["reuseCounts"][3]),
                   "reservoir_bytes_per_frame": n * 256 * 10},
        "roofline": roofline,
        "cpu_baseline": None,
    }


def build_mix_scenes(jobs):
    """(scene, camera, ShadedSceneHandle) of every job of the mix."""
    from renderingengine_amd import scenes
    from renderingengine_amd.batch import ShadedSceneHandle

    t0 = time.perf_counter()
    built = []
    for j in jobs:
        sc = scenes.soup_scene(n=j["tris"], seed=j["scene_seed"], quads=16, palette=64, width=j["width"],
                               height=j["height"])
        cam = scenes.look_at_camera(j["pos"], (0.0, 0.0, 0.0), vfov_deg=45.0, aspect=j["width"] / j["height"])
        built.append((sc, cam, ShadedSceneHandle(sc.tris, sc.materials, sc.lights, build_threads=cpu_threads())))
    print(f"[bench] {len(jobs)} scenes built in {time.perf_counter() - t0:.1f}s "
          f"({sum(j['tris'] for j in jobs)} triangles)", file=sys.stderr, flush=True)
    return built


def batch(args, ev):
    """64 render jobs: the batch engine (one launch sequence per step for all
    jobs in flight, continuous admission, hipGraphs, autotuned grid and path
    capacity) against 64 sequential plain per-frame launches of the same
    frames; films compared bit for bit."""
    from renderingengine_amd import _lib, scenes
    from renderingengine_amd.batch import BatchEngine, ShadedSceneHandle, frame_desc
    from renderingengine_amd.render import WALK_WIDE_COMPRESSED, WavefrontRuntime

    jobs = scenes.batch_job_mix(args.batch_jobs, 7, args.batch_max_tris)
    built = build_mix_scenes(jobs)

    def frame_of(k):
        j = jobs[k]
        return frame_desc(built[k][1], j["width"], j["height"], sample_index=0, spp=j["spp"],
                          max_bounce=j["bounce"], estimator=_lib.REX_ESTIMATOR_MIS, seed=j["seed"])

    # Plain: one rex_render_frame per job, sequentially (renderer setup untimed).
    plain_ms, films, rays = [], [], 0
    for k, j in enumerate(jobs):
        sc, cam, h = built[k]
        rt = WavefrontRuntime(j["width"] * j["height"], j["spp"])
        rt.attach(h)
        rt.update_frame(cam, j["width"], j["height"], sample_index=0, spp=j["spp"], max_bounce=j["bounce"],
                        estimator=_lib.REX_ESTIMATOR_MIS, seed=j["seed"], walk_mode=WALK_WIDE_COMPRESSED)
        ev.sync()
        t1 = time.perf_counter()
        rt.record_frame()
        st = rt.collect_completed_frame()
        plain_ms.append((time.perf_counter() - t1) * 1e3)
        rays += st["closestRays"] + st["shadowRays"]
        films.append(rt.film())
        rt.destroy()
    plain_s = sum(plain_ms) / 1e3
    print(f"[bench] plain: {plain_s * 1e3:.1f} ms for {len(jobs)} jobs", file=sys.stderr, flush=True)

    def run_batch(capacity, shade_bpc, graphs, subset, autotune):
        eng = BatchEngine(path_capacity=capacity, max_jobs=256, use_graphs=graphs, shade_blocks_per_cu=shade_bpc,
                          autotune=autotune)
        idx = [eng.add_scene(b[2]) for b in built]
        ev.sync()
        t1 = time.perf_counter()
        ids = [eng.submit(idx[k], frame_of(k)) for k in subset]
        t_sub = time.perf_counter() - t1
        eng.run()
        dt = time.perf_counter() - t1
        st = eng.stats()
        print(f"[bench] batch capacity {capacity}: {len(subset)} jobs {dt * 1e3:.1f} ms (submit {t_sub * 1e3:.1f} ms), "
              f"setting shade {st['shadeBlocksPerCu']}/CU graphs {st['graphs']} budget {st['pathBudget']} "
              f"(autotune state {st['autotune']})", file=sys.stderr, flush=True)
        return eng, ids, dt

    # The engine's own autotuner picks the Shade grid, graphs and the path
    # budget on its first busy steps (--batch-fixed pins a setting instead).
    if args.batch_fixed:
        capacity, shade_bpc, graphs = (int(x) for x in args.batch_fixed.split(","))
        autotune = False
    else:
        capacity, shade_bpc, graphs, autotune = 32 << 20, 16, True, True
    best = None
    for _ in range(max(1, args.steps)):
        eng, ids, dt = run_batch(capacity, shade_bpc, bool(graphs), list(range(len(jobs))), autotune)
        if best is None or dt < best[0]:
            if best is not None:
                best[1].destroy()
            best = (dt, eng, ids)
        else:
            eng.destroy()
    batch_s, eng, ids = best
    identical = all(np.array_equal(eng.film(ids[k]).view(np.uint32), films[k].view(np.uint32))
                    for k in range(len(jobs)))
    st = eng.stats()
    eng.destroy()
    for b in built:
        b[2].destroy()
    speedup = plain_s / batch_s
    return {
        "metric": "batched job throughput vs plain per-frame launches (64-job mix)",
        "value": round(len(jobs) / batch_s, 3), "unit": "jobs/s", "ms_per_step": round(batch_s * 1e3, 3),
        "config": {"workload": "batch64_mixed", "jobs": len(jobs), "triangles_total": sum(j["tris"] for j in jobs),
                   "resolutions": "U{256^2, 512^2, 1280x720, 1920x1080}", "spp": "U{1..16}", "max_bounce": "U{4..8}",
                   "plain_ms": round(plain_s * 1e3, 3), "batched_ms": round(batch_s * 1e3, 3),
                   "speedup_vs_plain": round(speedup, 3), "films_bit_identical": bool(identical),
                   "plain_grays_per_s": round(rays / plain_s / 1e9, 4),
                   "batched_grays_per_s": round(rays / batch_s / 1e9, 4),
                   "plain_job_ms_p50": round(float(np.percentile(plain_ms, 50)), 3),
                   "plain_job_ms_p99": round(float(np.percentile(plain_ms, 99)), 3),
                   "engine_steps": int(st["steps"]),
                   "path_capacity": capacity,
                   "setting": {"autotuned": bool(autotune), "shade_blocks_per_cu": int(st["shadeBlocksPerCu"]),
                               "graphs": bool(st["graphs"]), "path_budget": int(st["pathBudget"])}},
        "roofline": None,
        "cpu_baseline": None,
    }


def serve(args, ev):
    """Served form: one Poisson arrival trace of the batched mix (20% of the
    jobs 8-frame ReSTIR DI sequences on the job's scene, resolution and
    camera walk) replayed twice -- through the request loop (rex_server:
    admission by footprint under the budget, frames in the batch engine,
    sequences on renderer slots, all overlapped) and through the plain form
    (one renderer, jobs one at a time in arrival order, each rendered by
    rex_render_frame and waited for).  Latency = finish - arrival."""
    from renderingengine_amd import _lib, scenes
    from renderingengine_amd.batch import frame_desc
    from renderingengine_amd.render import WALK_WIDE_COMPRESSED, WavefrontRuntime
    from renderingengine_amd.serve import STATE_FINISHED, Server

    jobs = scenes.batch_job_mix(args.batch_jobs, 7, args.batch_max_tris)
    built = build_mix_scenes(jobs)
    trace = scenes.serve_trace(len(jobs), args.serve_rate, args.serve_seconds, 11)
    frames_per_seq = 8

    def frame_of(k):
        j = jobs[k]
        return frame_desc(built[k][1], j["width"], j["height"], sample_index=0, spp=j["spp"], max_bounce=j["bounce"],
                          estimator=_lib.REX_ESTIMATOR_MIS, seed=j["seed"])

    def walk_of(k, n):
        return scenes.camera_walk(built[k][1], frames_per_seq, step=0.02, seed=1000 + n)

    # Served.
    srv = Server(path_capacity=args.serve_capacity, use_graphs=True, shade_blocks_per_cu=16,
                 output_mib=32768, sequence_slots=4)
    for b in built:
        srv.add_scene(b[2])
    lat = {0: [], 1: []}
    pending = {}
    t0 = time.perf_counter()
    t0n = time.monotonic_ns()  # the server's job times are on the same monotonic clock
    nxt = 0
    report = 5.0
    while nxt < len(trace) or pending:
        now = time.perf_counter() - t0
        if now >= report:
            print(f"[bench] served t={now:.0f}s: {nxt} of {len(trace)} submitted, {len(pending)} pending",
                  file=sys.stderr, flush=True)
            report += 5.0
        while nxt < len(trace) and trace[nxt][0] <= now:
            _, k, kind = trace[nxt]
            if kind == 0:
                jid = srv.submit_frame(k, frame_of(k))
            else:
                j = jobs[k]
                jid = srv.submit_sequence(k, frame_desc(built[k][1], j["width"], j["height"],
                                                        max_bounce=j["bounce"], seed=j["seed"]), walk_of(k, nxt))
            pending[jid] = (kind, trace[nxt][0])
            nxt += 1
        srv.pump()
        for jid in list(pending):
            info = srv.job(jid)
            if info["state"] == STATE_FINISHED:
                kind, arr = pending.pop(jid)
                lat[kind].append((info["finishNs"] - (t0n + arr * 1e9)) / 1e6)
                srv.release(jid)
        if nxt < len(trace) and not pending:
            time.sleep(max(0.0, min(0.002, trace[nxt][0] - (time.perf_counter() - t0))))
    served_s = time.perf_counter() - t0
    st = srv.stats()
    srv.destroy()
    served_rays = st["engineRays"] + st["sequenceRays"]
    print(f"[bench] served: {len(trace)} jobs in {served_s:.2f}s, {served_rays / served_s / 1e9:.3f} Grays/s",
          file=sys.stderr, flush=True)

    # Plain: one renderer per kind, jobs in arrival order, each waited for.
    maxp = max(j["width"] * j["height"] for j in jobs)
    plain = WavefrontRuntime(maxp, max(j["spp"] for j in jobs))
    plain.attach(built[0][2])
    seqr = WavefrontRuntime(maxp, 1)
    seqr.attach(built[0][2])
    plat = {0: [], 1: []}
    plain_rays = 0
    t0 = time.perf_counter()
    report = 5.0
    for n, (arr, k, kind) in enumerate(trace):
        now = time.perf_counter() - t0
        if now >= report:
            print(f"[bench] plain t={now:.0f}s: {n} of {len(trace)} done", file=sys.stderr, flush=True)
            report += 5.0
        if now < arr:
            time.sleep(arr - now)
        j = jobs[k]
        if kind == 0:
            plain.bind(built[k][2])
            plain.update_frame(built[k][1], j["width"], j["height"], sample_index=0, spp=j["spp"],
                               max_bounce=j["bounce"], estimator=_lib.REX_ESTIMATOR_MIS, seed=j["seed"],
                               walk_mode=WALK_WIDE_COMPRESSED)
            plain.record_frame()
            s = plain.collect_completed_frame()
            plain_rays += s["closestRays"] + s["shadowRays"]
        else:
            seqr.bind(built[k][2])
            seqr.set_restir()
            for f, cam in enumerate(walk_of(k, n)):
                seqr.update_frame(cam, j["width"], j["height"], sample_index=f, spp=1, max_bounce=j["bounce"],
                                  estimator=_lib.REX_ESTIMATOR_RESTIR, seed=j["seed"], walk_mode=WALK_WIDE_COMPRESSED)
                seqr.record_frame()
                s = seqr.collect_completed_frame()
                plain_rays += s["closestRays"] + s["shadowRays"] + j["width"] * j["height"]
        plat[kind].append((time.perf_counter() - t0 - arr) * 1e3)
    plain_s = time.perf_counter() - t0
    plain.destroy()
    seqr.destroy()
    for b in built:
        b[2].destroy()
    print(f"[bench] plain: {len(trace)} jobs in {plain_s:.2f}s, {plain_rays / plain_s / 1e9:.3f} Grays/s",
          file=sys.stderr, flush=True)

    def pct(v, q):
        return round(float(np.percentile(v, q)), 2) if v else None

    allserved = lat[0] + lat[1]
    allplain = plat[0] + plat[1]
    return {
        "metric": "served job latency p99 (Poisson arrivals of the 64-job mix, 20% 8-frame ReSTIR DI sequences)",
        "value": pct(allserved, 99), "unit": "ms", "ms_per_step": round(served_s * 1e3, 3),
        "higher_is_better": False,
        "config": {"workload": "serve_poisson", "rate_jobs_per_s": args.serve_rate,
                   "arrival_seconds": args.serve_seconds, "jobs": len(trace),
                   "sequences": sum(1 for t in trace if t[2] == 1), "sequence_frames": frames_per_seq,
                   "served_p50_ms": pct(allserved, 50), "served_p99_ms": pct(allserved, 99),
                   "plain_p50_ms": pct(allplain, 50), "plain_p99_ms": pct(allplain, 99),
                   "served_frame_p50_ms": pct(lat[0], 50), "served_frame_p99_ms": pct(lat[0], 99),
                   "served_sequence_p50_ms": pct(lat[1], 50), "served_sequence_p99_ms": pct(lat[1], 99),
                   "plain_frame_p50_ms": pct(plat[0], 50), "plain_frame_p99_ms": pct(plat[0], 99),
                   "plain_sequence_p50_ms": pct(plat[1], 50), "plain_sequence_p99_ms": pct(plat[1], 99),
                   "served_wall_s": round(served_s, 3), "plain_wall_s": round(plain_s, 3),
                   "served_grays_per_s": round(served_rays / served_s / 1e9, 4),
                   "plain_grays_per_s": round(plain_rays / plain_s / 1e9, 4),
                   "rays_served": int(served_rays), "rays_plain": int(plain_rays),
                   "served_jobs_per_s": round(len(trace) / served_s, 2),
                   "plain_jobs_per_s": round(len(trace) / plain_s, 2),
                   "budget_bytes": int(st["budgetBytes"]), "resident_bytes": int(st["residentBytes"]),
                   "peak_admitted_bytes": int(st["peakAdmittedBytes"]), "engine_steps": int(st["engineSteps"]),
                   "path_capacity": args.serve_capacity, "sequence_slots": 4},
        "roofline": None,
        "cpu_baseline": None,
    }


REFERENCE_CORNELL_MRAYS = 10.17  # rt/cpu/references/wave1/cornell-256x256-64spp-seed1.json:13-16


def cornell(args, ev):
    """The reference CPU path tracer's Cornell fixture (36 triangles,
    Lambert, one 0.6 x 0.6 light) at 256x256, 64 spp, 8 bounces through the
    wavefront path (MIS, one 64-spp frame per step), beside the oracle's
    restatement of RenderCornellReference (bit-identical to the reference's
    golden) timed on the host's cores: the same workload as the reference's
    only published CPU timing, 10.17 Mrays/s."""
    from renderingengine_amd import _lib, scenes
    from renderingengine_amd.render import WavefrontRuntime

    width, height, spp, bounce, seed = 256, 256, 64, 8, 1
    sc = scenes.reference_cornell_fixture(width, height)
    rt = WavefrontRuntime(width * height, spp)
    rt.set_scene(sc.tris, sc.materials, sc.lights, light_selection=_lib.REX_LIGHT_SELECTION_POWER)

    def launch(k):
        rt.update_frame(sc.camera, width, height, sample_index=k * spp, spp=spp, max_bounce=bounce,
                        estimator=_lib.REX_ESTIMATOR_MIS, seed=seed, walk_mode=_lib.REX_WALK_WIDE_COMPRESSED)
        rt.record_frame()

    rt.set_profiling(counters=True)
    launch(0)
    cstat = rt.collect_completed_frame()
    cc, scn = cstat["closestCounters"], cstat["shadowCounters"]
    queries0 = cstat["closestRays"] + cstat["shadowRays"]
    slab, tri = int(cc[9] + scn[9]), int(cc[1] + scn[1])
    rt.set_profiling(timing=True)
    for k in range(args.warmup):
        launch(k + 1)
        rt.collect_completed_frame()
    rt.invalidate_progressive_film()
    ev.sync()
    rays = 0
    trace_ms = 0.0
    launches = 0
    t0 = time.perf_counter()
    for k in range(args.steps):
        launch(k)
        st = rt.collect_completed_frame()
        rays += st["closestRays"] + st["shadowRays"]
        trace_ms += sum(st["traceMs"])
        launches += len(st["traceMs"])
    ev.sync()
    wall = time.perf_counter() - t0
    film = rt.film()
    rt.destroy()
    gpu_mean = film[:, :3].mean(axis=0)
    grays = rays / wall / 1e9
    print(f"[bench] cornell 256x256 64spp b8: {args.steps} frames, wall {wall * 1e3:.1f} ms, "
          f"{grays:.4f} Grays/s, film mean {gpu_mean}", file=sys.stderr, flush=True)

    cpu, golden_mean = None, None
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import oracle_lib
        threads = cpu_threads()
        oracle_lib.cornell_render(16, 16, 1, bounce, seed, threads=threads)  # warm
        px, cnt, scene_hash, dt = oracle_lib.cornell_render(width, height, spp, bounce, seed, threads=threads)
        image_hash = oracle_lib.linear_rgb_hash(px)
        golden_mean = px.reshape(-1, 3).astype(np.float64).mean(axis=0)
        nproc, model = host_cpu()
        cpu_rays = int(cnt[0] + cnt[1])
        cpu = {"value": round(cpu_rays / dt / 1e9, 6), "unit": "Grays/s", "cores": threads, "kind": "port",
               "host_logical_cpus": nproc, "host_cpu_model": model,
               "reference_published_mrays_per_s": REFERENCE_CORNELL_MRAYS,
               "sample": f"the whole 256x256, 64 spp, 8-bounce, seed-1 image through the C oracle's "
                         f"RenderCornellReference ({int(cnt[0])} rays + {int(cnt[1])} shadow rays, image hash "
                         f"0x{image_hash:016x}, scene hash 0x{scene_hash:016x}), {dt:.2f}s wall on {threads} threads"}
        print(f"[bench] cornell CPU oracle: {cpu_rays / dt / 1e6:.2f} Mrays/s on {threads} threads "
              f"(reference published {REFERENCE_CORNELL_MRAYS} Mrays/s)", file=sys.stderr, flush=True)
    # 48 B per node record and per triangle-position block read, 96 B of ray
    # and hit per query (the device walk's own counts on the first frame).
    alg = (48 * slab + 48 * tri + 96 * queries0) * (rays / max(queries0 * args.steps, 1))
    achieved = alg / (trace_ms / args.steps * 1e-3) / 1e9
    return {
        "metric": "Grays/s on the reference's CPU Cornell fixture (256x256, 64 spp, 8 bounces)",
        "value": round(grays, 4), "unit": "Grays/s",
        "ms_per_step": round(wall / args.steps * 1e3, 3),
        "config": {"workload": "reference_cornell_256_64spp_b8", "triangles": int(sc.tris.shape[0]),
                   "film": f"{width}x{height}", "spp": spp, "max_bounce": bounce, "estimator": "mis",
                   "rays_per_frame": rays // args.steps,
                   "trace_ms_per_frame": round(trace_ms / args.steps, 3),
                   "film_mean": [round(float(x), 5) for x in gpu_mean],
                   "reference_cpu_image_mean": [round(float(x), 5) for x in golden_mean]
                   if golden_mean is not None else None},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                     "kernel": "k_trace (queue traversal: closest-hit rays + any-hit shadow queries)",
                     "kernel_ms_per_launch": round(trace_ms / max(launches, 1), 4),
                     "slab_tests_per_query": round(slab / max(queries0, 1), 3),
                     "tri_tests_per_query": round(tri / max(queries0, 1), 3)},
        "cpu_baseline": cpu,
    }


def reconstruction(args, ev):
    """The reconstruction boundary at 1080p on the 1M-triangle frame: the
    ABI-v2 export Resolve adds to a frame, then MotionVectorsCS and the
    progressive-mean ComposeCS over the exported records, timed with HIP
    events; step = one motion-vector pass + one compose pass."""
    from renderingengine_amd import _lib, reconstruction as R, scenes
    from renderingengine_amd.render import WavefrontRuntime
    from renderingengine_amd.traversal import DeviceBuffer

    width, height, spp, bounce = 1920, 1080, 1, 8
    n = width * height
    sc = scenes.soup_scene(n=args.triangles, seed=0, width=width, height=height)
    rt = WavefrontRuntime(n, spp)
    rt.set_scene(sc.tris, sc.materials, sc.lights, light_selection=_lib.REX_LIGHT_SELECTION_POWER)
    rt.set_profiling(timing=True)

    def resolve_ms(frames):
        ms = []
        for k in range(frames):
            rt.update_frame(sc.camera, width, height, sample_index=k, spp=spp, max_bounce=bounce,
                            estimator=_lib.REX_ESTIMATOR_MIS, seed=0, walk_mode=_lib.REX_WALK_WIDE_COMPRESSED)
            rt.record_frame()
            st = rt.collect_completed_frame()
            ms.append(st["stageMs"][4])
        return float(np.median(ms)), st

    plain_ms, _ = resolve_ms(3)
    tg = R.ReconstructionTargets(n)
    rt.set_reconstruction_targets(tg)
    export_ms, st = resolve_ms(3)
    rec = tg.download()
    gb, mi, raw, term = rec["gbuffer"], rec["motionInputs"], rec["rawSignal"], rec["terminal"]
    valid = int(np.count_nonzero(gb["primary"]["identity"][:, 3] & 1))
    eye = DeviceBuffer.from_array(np.eye(4, dtype=np.float32))
    vp, wv = R.camera_matrices(sc.camera)
    cam2 = dict(sc.camera)
    cam2["position"] = np.asarray(sc.camera["position"], np.float64) + (0.05, 0.0, 0.02)
    vp2, wv2 = R.camera_matrices(cam2)
    c = R.motion_constants(width, height, vp, vp2, wv2)
    out, film = DeviceBuffer(n * 16), DeviceBuffer(n * 16)

    def mv():
        R.motion_vectors(tg.ptr("gbuffer"), tg.ptr("motionInputs"), eye, eye, c)

    def comp(k):
        R.compose(width, height, R.COMPOSE_PROGRESSIVE_MEAN, out.ptr, raw_signal=tg.ptr("rawSignal"),
                  terminal=tg.ptr("terminal"), film=film.ptr, film_sample_count=k)

    for k in range(args.warmup):
        mv()
        comp(k)
    ev.sync()
    ev.record_start()
    for _ in range(args.steps):
        mv()
    ev.record_stop()
    mv_ms = ev.elapsed_ms() / args.steps
    ev.record_start()
    for k in range(args.steps):
        comp(k)
    ev.record_stop()
    comp_ms = ev.elapsed_ms() / args.steps
    t0 = time.perf_counter()
    ev.sync()
    for k in range(args.steps):
        mv()
        comp(k)
    ev.sync()
    wall = time.perf_counter() - t0
    # Algorithmic bytes: motion reads each record's 16-B surface identity,
    # and for valid surfaces the 32-B input, writing 48 B (36 B for a
    # cleared record); compose reads 64 + 16 + 16 B and writes 16 + 16 B.
    mv_bytes = 16 * n + valid * (32 + 48) + (n - valid) * 36
    comp_bytes = 128 * n
    mv_gbs = mv_bytes / (mv_ms * 1e-3) / 1e9
    comp_gbs = comp_bytes / (comp_ms * 1e-3) / 1e9
    export_bytes = (208 + 32 + 64 + 16) * n
    print(f"[bench] resolve {plain_ms:.3f} ms plain, {export_ms:.3f} ms with export; motion {mv_ms:.4f} ms "
          f"({mv_gbs:.0f} GB/s), compose {comp_ms:.4f} ms ({comp_gbs:.0f} GB/s)", file=sys.stderr, flush=True)
    cpu = None
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import oracle_lib
        t1 = time.perf_counter()
        oracle_lib.motion_vectors(gb, mi, np.eye(4), np.eye(4), vp, vp2, wv2)
        oracle_lib.compose(R.COMPOSE_PROGRESSIVE_MEAN, n, raw=raw, terminal=term, film=np.zeros((n, 4), np.float32))
        cpu_s = time.perf_counter() - t1
        cpu = {"value": round(n / cpu_s / 1e9, 5), "unit": "Gpixels/s", "cores": 1, "kind": "port",
               "sample": f"one motion-vector + progressive-mean compose pass over the full {width}x{height} frame"}
    dominant = (mv_ms, mv_gbs, mv_bytes, "k_motion_vectors") if mv_ms > comp_ms else \
        (comp_ms, comp_gbs, comp_bytes, "k_compose (progressive mean)")
    rt.destroy()
    tg.free()
    return {
        "metric": "reconstruction passes Gpixels/s (motion vectors + progressive-mean compose, 1080p)",
        "value": round(n / ((mv_ms + comp_ms) * 1e-3) / 1e9, 4), "unit": "Gpixels/s",
        "ms_per_step": round(wall / args.steps * 1e3, 4),
        "config": {"workload": "reconstruction_1080p", "triangles": args.triangles, "film": f"{width}x{height}",
                   "valid_surfaces": valid, "motion_ms": round(mv_ms, 4), "compose_ms": round(comp_ms, 4),
                   "resolve_ms_plain": round(plain_ms, 4), "resolve_ms_with_export": round(export_ms, 4),
                   "export_bytes_per_frame": export_bytes,
                   "export_gbs": round(export_bytes / (max(export_ms - plain_ms, 1e-6) * 1e-3) / 1e9, 1)},
        "roofline": {"bound": "hbm", "achieved": round(dominant[1], 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(dominant[1] / HBM_PEAK_GBS, 4), "traffic": None, "kernel": dominant[3],
                     "kernel_ms_per_launch": round(dominant[0], 4), "bytes_per_launch": int(dominant[2])},
        "cpu_baseline": cpu,
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="frame_1m",
                    choices=("frame_1m", "trace_bounce_1m", "restir_1m", "batch64", "serve", "reconstruction",
                             "cornell"))
    ap.add_argument("--inflight-pool", type=int, default=8, help="frame_1m: path pool of the in-flight form, in frames "
                                                                 "(8 x 8.3M paths, about 19 GB of path and queue state)")
    ap.add_argument("--serve-rate", type=float, default=30.0, help="Poisson arrival rate, jobs/s")
    ap.add_argument("--serve-seconds", type=float, default=20.0, help="arrival window, s")
    ap.add_argument("--serve-capacity", type=int, default=32 << 20, help="batch engine path pool of the server")
    ap.add_argument("--batch-jobs", type=int, default=64)
    ap.add_argument("--batch-max-tris", type=int, default=2_000_000)
    ap.add_argument("--batch-fixed", default="", help="CAPACITY,SHADE_BLOCKS_PER_CU,GRAPHS: skip the autotuner")
    ap.add_argument("--restir-bounces", type=int, default=1, help="max bounce of restir_1m frames")
    ap.add_argument("--restir-reference", action="store_true",
                    help="restir_1m: the reference-correction estimator instead of the biased one")
    ap.add_argument("--triangles", type=int, default=1_000_000)
    ap.add_argument("--mode", type=int, default=4, help="0 left-first short stack, 1 stackless (walk only), "
                                                       "2 near-first, 3 near-first 4-wide, "
                                                       "4 near-first compressed 4-wide (production)")
    ap.add_argument("--cpu-sample", type=int, default=400_000, help="rays (trace_bounce_1m)")
    ap.add_argument("--cpu-rows", type=int, default=1080, help="frame rows (frame_1m; 1080 = the whole frame, ~7 s)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--summarize-profiles", action="store_true",
                    help="write the frame workload's counter summary from rocprofv3 databases and exit")
    ap.add_argument("--summarize-restir-profiles", action="store_true",
                    help="write the ReSTIR workload's counter summary from rocprofv3 databases and exit")
    ap.add_argument("--kt-db")
    ap.add_argument("--fetch-db")
    ap.add_argument("--write-db")
    ap.add_argument("--sq-db")
    ap.add_argument("--tcc-db")
    ap.add_argument("--summarize-out", default=FRAME_COUNTERS)
    ap.add_argument("--trace-out", default=os.path.join(ROOT, "profiles", "frame_1m_1080p_4spp_b8_kernel_trace.md"))
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="also run one device-to-device copy of this many MiB (untimed), the reference "
                         "kernel for scaling rocprofv3 FETCH_SIZE / WRITE_SIZE to bytes")
    args = ap.parse_args()
    if args.summarize_profiles or args.summarize_restir_profiles:
        if args.copy_probe_mib <= 0:
            args.copy_probe_mib = 1024
        (summarize_restir_profiles if args.summarize_restir_profiles else summarize_profiles)(args)
        return
    if args.gpus != 1:
        print(json.dumps({"error": "rex benchmarks one GPU; --gpus must be 1 (frames shard by image tile, "
                                   "sample range or job across GPUs, which is not run here)"}))
        sys.exit(2)
    ev = HipEvents()
    if args.copy_probe_mib > 0:
        import torch
        n = args.copy_probe_mib << 20
        src = torch.ones(n, dtype=torch.uint8, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        print(f"[bench] copy probe: {n} bytes read, {n} bytes written", file=sys.stderr, flush=True)
        del src, dst
    run = {"frame_1m": frame, "trace_bounce_1m": trace_bounce, "restir_1m": restir, "batch64": batch,
           "serve": serve, "reconstruction": reconstruction, "cornell": cornell}[args.workload]
    out = run(args, ev)
    line = {"metric": out["metric"], "value": out["value"], "unit": out["unit"], "n_gpus": 1,
            "steps": args.steps, "warmup": args.warmup, "ms_per_step": out["ms_per_step"],
            "higher_is_better": out.get("higher_is_better", True), "scaling": "weak", "vs_baseline": None,
            "dtype": "f32",
            "data": "synthetic", "config": out["config"], "roofline": out["roofline"],
            "cpu_baseline": out["cpu_baseline"]}
    print(json.dumps(line))


if __name__ == "__main__":
    main()
