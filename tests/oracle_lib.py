"""ctypes bindings for the CPU oracle (oracle/liboracle.so) and, when it was
built, the reference's own CPU code (oracle/_ref/libref.so).

Test infrastructure only: the product path never imports this module.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
ORACLE_SO = os.path.join(ORACLE_DIR, "liboracle.so")
REF_SO = os.path.join(ORACLE_DIR, "_ref", "libref.so")

INVALID = 0xFFFFFFFF
P = ctypes.c_void_p


def _ptr(a):
    return ctypes.c_void_p(a.ctypes.data) if a is not None else None


_oracle = None
_ref = None


def build_oracle():
    subprocess.run(["make", "-s", "-C", ORACLE_DIR], check=True)


def oracle():
    global _oracle
    if _oracle is None:
        if not os.path.exists(ORACLE_SO):
            build_oracle()
        _oracle = ctypes.CDLL(ORACLE_SO)
    return _oracle


def ref():
    """The reference's compiled CPU code, or None where it was not built."""
    global _ref
    if _ref is None and os.path.exists(REF_SO):
        _ref = ctypes.CDLL(REF_SO)
    return _ref


WALK_HIT = np.dtype([
    ("result", np.uint32), ("t", np.float32), ("u", np.float32), ("v", np.float32),
    ("frontFace", np.uint32), ("triangle", np.uint32), ("position", np.float32, 3),
    ("geometricNormal", np.float32, 3), ("shadingNormal", np.float32, 3),
    ("instanceId", np.uint32), ("primitiveId", np.uint32), ("materialIndex", np.uint32)])
assert WALK_HIT.itemsize == 72


class OracleBvh:
    """Flattened BVH built by the oracle: nodes (m, 12) uint32 words,
    order (n,) leaf-order source indices, leaf-order triangle records."""

    def __init__(self, tris, leaf_size=4):
        tris = np.ascontiguousarray(tris, dtype=np.uint32)
        n = tris.shape[0]
        nodes = np.zeros((max(2 * n - 1, 1), 12), dtype=np.uint32)
        order = np.zeros(max(n, 1), dtype=np.uint32)
        count = ctypes.c_uint32()
        depth = ctypes.c_uint32()
        rc = oracle().ora_bvh_build(_ptr(tris), ctypes.c_uint32(n), ctypes.c_uint32(leaf_size),
                                    _ptr(nodes), ctypes.byref(count), _ptr(order),
                                    ctypes.byref(depth))
        if rc != 0:
            raise ValueError(f"oracle BVH build failed ({rc})")
        self.nodes = nodes[:count.value].copy()
        self.order = order[:n].copy()
        self.depth = depth.value
        self.tris = np.ascontiguousarray(tris[self.order]) if n else tris

    def walk(self, rays, keys=None, shadow=False, counters=None):
        rays = np.ascontiguousarray(rays, dtype=np.float32)
        n = rays.shape[0]
        hits = np.zeros(n, dtype=WALK_HIT)
        if keys is not None:
            keys = np.ascontiguousarray(keys, dtype=np.uint32)
        oracle().ora_trace_walk_batch(_ptr(self.nodes), ctypes.c_uint32(self.nodes.shape[0]),
                                      _ptr(self.tris), ctypes.c_uint32(self.tris.shape[0]),
                                      _ptr(rays), _ptr(keys), ctypes.c_uint32(n),
                                      ctypes.c_int(1 if shadow else 0), _ptr(hits),
                                      _ptr(counters))
        return hits

    def batch_v1(self, rays_v1, query_mode=0, stack_capacity=64, counters=None):
        rays_v1 = np.ascontiguousarray(rays_v1, dtype=np.uint32)
        n = rays_v1.shape[0]
        hits = np.zeros((n, 24), dtype=np.uint32)
        oracle().ora_trace_batch_v1(_ptr(self.nodes), ctypes.c_uint32(self.nodes.shape[0]),
                                    _ptr(self.tris), ctypes.c_uint32(self.tris.shape[0]),
                                    _ptr(rays_v1), ctypes.c_uint32(n), ctypes.c_uint32(query_mode),
                                    ctypes.c_uint32(stack_capacity), _ptr(hits), _ptr(counters))
        return hits


def brute(tris_leaf_order, rays):
    rays = np.ascontiguousarray(rays, dtype=np.float32)
    tris_leaf_order = np.ascontiguousarray(tris_leaf_order, dtype=np.uint32)
    out = np.zeros(rays.shape[0], dtype=WALK_HIT)
    for i in range(rays.shape[0]):
        r = rays[i]
        o = np.ascontiguousarray(r[0:3])
        d = np.ascontiguousarray(r[4:7])
        h = np.zeros(1, dtype=WALK_HIT)
        oracle().ora_trace_brute(_ptr(tris_leaf_order), ctypes.c_uint32(tris_leaf_order.shape[0]),
                                 _ptr(o), _ptr(d), ctypes.c_float(r[3]), ctypes.c_float(r[7]),
                                 _ptr(h))
        out[i] = h[0]
    return out


def rays_v1_from(rays, mask=0xFF):
    """(n, 8) (o, tMin, d, tMax) -> (n, 16) ABI-v1 queue records."""
    rays = np.ascontiguousarray(rays, dtype=np.float32)
    n = rays.shape[0]
    rec = np.zeros((n, 16), dtype=np.uint32)
    rec.view(np.float32)[:, 0:8] = rays
    rec[:, 8] = np.arange(n, dtype=np.uint32)        # ray ID
    rec[:, 9] = np.arange(n, dtype=np.uint32) + 7   # path ID
    rec[:, 10] = 0
    rec[:, 11] = mask
    return rec


def ref_build(tris, leaf_size=4):
    """Reference Bvh<float> + FlattenCanonicalL3BinnedSah on the positions of
    tris, with primitive IDs 0..n-1.  Returns (nodes (m, 12), order, depth)."""
    lib = ref()
    if lib is None:
        return None
    tris = np.ascontiguousarray(tris, dtype=np.uint32)
    n = tris.shape[0]
    pos = np.ascontiguousarray(tris.view(np.float32)[:, 0:12].reshape(n, 3, 4)[:, :, :3]).reshape(n, 9)
    nodes = np.zeros((2 * n - 1, 12), dtype=np.uint32)
    order = np.zeros(n, dtype=np.uint32)
    count = ctypes.c_uint32()
    depth = ctypes.c_uint32()
    rc = lib.ref_build_flatten(_ptr(pos), ctypes.c_uint32(n), ctypes.c_uint32(leaf_size),
                               _ptr(nodes), ctypes.byref(count), _ptr(order), ctypes.byref(depth))
    if rc != 0:
        raise ValueError(f"reference build failed ({rc})")
    return nodes[:count.value].copy(), order, depth.value


def ref_trace(rays, mode=0):
    """Reference SoftwareGpu::Trace over the BVH of the last ref_build."""
    lib = ref()
    rays = np.ascontiguousarray(rays, dtype=np.float32)
    n = rays.shape[0]
    tuvf = np.zeros((n, 4), dtype=np.float32)
    prim = np.zeros(n, dtype=np.uint32)
    status = np.zeros(n, dtype=np.uint32)
    rc = lib.ref_trace(_ptr(rays), ctypes.c_uint32(n), ctypes.c_uint32(mode), _ptr(tuvf),
                       _ptr(prim), _ptr(status), None, None)
    assert rc == 0
    return tuvf, prim, status
