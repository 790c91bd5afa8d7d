"""Synthetic scene generators for the rex path-tracing core.

Every scene is produced as a canonical traversal-triangle stream: one
176-byte record per world-space triangle, the record the reference builds in
``BuildCanonicalTraversalScene`` (src/rt/gpu/CanonicalTraversalScene.cpp:123-240;
layout include/rt/gpu/CanonicalTraversalScene.hpp:15-23), plus a material
table and light list.  The generators are seeded and deterministic.

Scenes:
  * ``triangle_soup``  -- N fp32 triangles: 70 % of centroids uniform in
    [-10, 10]^3, 30 % in 8 Gaussian blobs (sigma 0.5, centres uniform in
    [-8, 8]^3); edge lengths log-normal(mu=-3, sigma=0.7); random orientation.
  * ``canonical_triangle`` -- the reference's canonical triangle
    (include/scene/CanonicalScene.hpp:26-30) used by FixedHitFixtureV1.
  * ``grid12`` -- the L4 12-triangle grid (tests/gpu/L4/TestMain.cpp:29-79).
"""
from __future__ import annotations

import numpy as np

TRI_WORDS = 44  # 176 bytes / 4

# include/contracts/SceneAbiV0.hpp:16-44
GEOMETRY_FLAG_OPAQUE = 1 << 0
GEOMETRY_FLAG_ALPHA_MASK = 1 << 1
GEOMETRY_FLAG_DOUBLE_SIDED = 1 << 2
INSTANCE_FLAG_VISIBLE = 1 << 0
INSTANCE_FLAG_CASTS_SHADOW = 1 << 1
MATERIAL_FLAG_DOUBLE_SIDED = 1 << 0
MATERIAL_FLAG_ALPHA_MASK = 1 << 1
MATERIAL_FLAG_EMISSIVE = 1 << 2


def pack_triangles(positions, normals=None, texcoords=None, identity=None, metadata=None):
    """Pack (n,3,3) positions (+ optional attributes) into (n, 44) uint32 records.

    identity = (instance, primitive, geometry, material);
    metadata = (geometry flags, material flags, traversal ID, instance flags).
    Defaults: face normals, zero texcoords, identity (0, i, 0, 0), metadata
    (opaque | double-sided, 0, i, visible | casts-shadow).
    """
    positions = np.ascontiguousarray(positions, dtype=np.float32).reshape(-1, 3, 3)
    n = positions.shape[0]
    rec = np.zeros((n, TRI_WORDS), dtype=np.uint32)
    f = rec.view(np.float32)
    pos4 = np.ones((n, 3, 4), dtype=np.float32)
    pos4[:, :, :3] = positions
    f[:, 0:12] = pos4.reshape(n, 12)
    if normals is None:
        e1 = positions[:, 1] - positions[:, 0]
        e2 = positions[:, 2] - positions[:, 0]
        fn = np.cross(e1.astype(np.float64), e2.astype(np.float64))
        fn /= np.linalg.norm(fn, axis=1, keepdims=True)
        normals = np.repeat(fn[:, None, :], 3, axis=1).astype(np.float32)
    nrm4 = np.zeros((n, 3, 4), dtype=np.float32)
    nrm4[:, :, :3] = np.asarray(normals, dtype=np.float32).reshape(n, 3, 3)
    f[:, 12:24] = nrm4.reshape(n, 12)
    if texcoords is not None:
        uv4 = np.zeros((n, 3, 4), dtype=np.float32)
        uv4[:, :, :2] = np.asarray(texcoords, dtype=np.float32).reshape(n, 3, 2)
        f[:, 24:36] = uv4.reshape(n, 12)
    idx = np.arange(n, dtype=np.uint32)
    if identity is None:
        identity = np.stack([np.zeros(n, np.uint32), idx, np.zeros(n, np.uint32),
                             np.zeros(n, np.uint32)], axis=1)
    rec[:, 36:40] = np.asarray(identity, dtype=np.uint32).reshape(n, 4)
    if metadata is None:
        metadata = np.stack([
            np.full(n, GEOMETRY_FLAG_OPAQUE | GEOMETRY_FLAG_DOUBLE_SIDED, np.uint32),
            np.zeros(n, np.uint32), idx,
            np.full(n, INSTANCE_FLAG_VISIBLE | INSTANCE_FLAG_CASTS_SHADOW, np.uint32)], axis=1)
    rec[:, 40:44] = np.asarray(metadata, dtype=np.uint32).reshape(n, 4)
    return rec


def _random_rotations(rng, n):
    # Uniform random rotations from unit quaternions (Shoemake 1992).
    u1, u2, u3 = rng.random(n), rng.random(n), rng.random(n)
    q = np.stack([np.sqrt(1 - u1) * np.sin(2 * np.pi * u2), np.sqrt(1 - u1) * np.cos(2 * np.pi * u2),
                  np.sqrt(u1) * np.sin(2 * np.pi * u3), np.sqrt(u1) * np.cos(2 * np.pi * u3)], axis=1)
    x, y, z, w = q[:, 0], q[:, 1], q[:, 2], q[:, 3]
    return np.stack([
        np.stack([1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)], axis=1),
        np.stack([2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)], axis=1),
        np.stack([2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)], axis=1)], axis=1)


def soup_positions(n, seed=0, extent=10.0, blob_fraction=0.3, blobs=8, blob_sigma=0.5,
                   edge_mu=-3.0, edge_sigma=0.7):
    """(n, 3, 3) float32 vertex positions of the triangle-soup scene."""
    rng = np.random.default_rng(seed)
    n_blob = int(round(n * blob_fraction))
    n_uni = n - n_blob
    cent = np.empty((n, 3), dtype=np.float64)
    cent[:n_uni] = rng.uniform(-extent, extent, size=(n_uni, 3))
    centres = rng.uniform(-0.8 * extent, 0.8 * extent, size=(blobs, 3))
    which = rng.integers(0, blobs, size=n_blob)
    cent[n_uni:] = centres[which] + rng.normal(0.0, blob_sigma, size=(n_blob, 3))
    perm = rng.permutation(n)
    cent = cent[perm]
    l1 = rng.lognormal(edge_mu, edge_sigma, size=n)
    l2 = rng.lognormal(edge_mu, edge_sigma, size=n)
    theta = rng.uniform(np.pi / 6, 5 * np.pi / 6, size=n)
    local = np.zeros((n, 3, 3), dtype=np.float64)
    local[:, 1, 0] = l1
    local[:, 2, 0] = l2 * np.cos(theta)
    local[:, 2, 1] = l2 * np.sin(theta)
    local -= local.mean(axis=1, keepdims=True)
    rot = _random_rotations(rng, n)
    world = np.einsum('nij,nkj->nki', rot, local) + cent[:, None, :]
    return world.astype(np.float32)


def triangle_soup(n, seed=0, **kw):
    """Packed 176-byte records for the triangle-soup scene (identity defaults)."""
    return pack_triangles(soup_positions(n, seed, **kw))


def canonical_triangle():
    """The canonical triangle (include/scene/CanonicalScene.hpp:26-30) with
    identity (0, 0, 0, 0) as FixedHitFixtureV1 uses it
    (src/rt/gpu/GpuTraversalContracts.cpp:240-252)."""
    pos = np.array([[[-0.75, -0.5, 0.0], [0.75, -0.5, 0.0], [0.0, 0.75, 0.0]]], dtype=np.float32)
    nrm = np.array([[[0.0, 0.0, 1.0]] * 3], dtype=np.float32)
    identity = np.zeros((1, 4), dtype=np.uint32)
    metadata = np.array([[GEOMETRY_FLAG_OPAQUE, 0, 0, INSTANCE_FLAG_VISIBLE | INSTANCE_FLAG_CASTS_SHADOW]],
                        dtype=np.uint32)
    return pack_triangles(pos, normals=nrm, identity=identity, metadata=metadata)


def fixed_hit_rays_v1():
    """FixedHitFixtureV1's three ABI-v1 rays (GpuTraversalContracts.cpp:253-270):
    a front-face hit, a miss and an invalid (zero-direction) ray; (3, 16) uint32."""
    rec = np.zeros((3, 16), dtype=np.uint32)
    f = rec.view(np.float32)
    f[0, 0:8] = [0.0, 0.0, 2.0, 0.001, 0.0, 0.0, -1.0, 100.0]
    f[1, 0:8] = [2.0, 2.0, 2.0, 0.001, 0.0, 0.0, -1.0, 100.0]
    f[2, 0:8] = [0.0, 0.0, 0.0, 1.0, 0.0, 0.0, 0.0, 1.0]
    rec[:, 8:12] = [[0, 0, 0, 0xFF], [1, 1, 0, 0xFF], [2, 2, 0, 0xFF]]
    rec[:, 12:16] = [1, 0, 0, 0x46495831]
    return rec


def grid(count=12):
    """The L4 grid fixture: triangle i centred at ((i%4)*2, (i/4)*2, 0) with
    vertices (x-.5, y-.5), (x+.5, y-.5), (x, y+.5) and stable primitive ID
    1000 + i (tests/gpu/L4/TestMain.cpp:29-45)."""
    tris = []
    for i in range(count):
        x = float(i % 4) * 2.0
        y = float(i // 4) * 2.0
        tris.append([[x - 0.5, y - 0.5, 0.0], [x + 0.5, y - 0.5, 0.0], [x, y + 0.5, 0.0]])
    n = len(tris)
    idx = np.arange(n, dtype=np.uint32)
    identity = np.stack([np.zeros(n, np.uint32), 1000 + idx, np.zeros(n, np.uint32),
                         np.zeros(n, np.uint32)], axis=1)
    return pack_triangles(np.array(tris, dtype=np.float32), identity=identity)


def grid_corpus(hit_count=12):
    """The L4 ray corpus (tests/gpu/L4/TestMain.cpp:47-79): one downward ray per
    grid triangle centre plus three misses; (n, 8) = (o, tMin, d, tMax)."""
    rays = []
    for i in range(hit_count):
        x = float(i % 4) * 2.0
        y = float(i // 4) * 2.0
        rays.append([x, y, 2.0, 0.001, 0.0, 0.0, -1.0, 100.0])
    rays.append([100.0, 100.0, 2.0, 0.001, 0.0, 0.0, -1.0, 100.0])
    rays.append([-2.0, 0.0, 0.0, 0.001, 1.0, 0.0, 0.0, 100.0])
    rays.append([0.0, 0.0, -2.0, 0.001, 0.0, 0.0, 1.0, 100.0])
    return np.array(rays, dtype=np.float32)


def shared_edge_quad():
    """Two triangles sharing the diagonal of a quad at z = -2 with primitive IDs
    17 and 11 (tests/gpu/L4/TestMain.cpp:284-308); a ray down the shared edge
    must report ID 11."""
    pos = np.array([[[-1.0, -1.0, -2.0], [1.0, -1.0, -2.0], [1.0, 1.0, -2.0]],
                    [[-1.0, -1.0, -2.0], [1.0, 1.0, -2.0], [-1.0, 1.0, -2.0]]], dtype=np.float32)
    identity = np.array([[0, 17, 0, 0], [0, 11, 0, 0]], dtype=np.uint32)
    metadata = np.array([[GEOMETRY_FLAG_OPAQUE | GEOMETRY_FLAG_DOUBLE_SIDED, 0, 17, 3],
                         [GEOMETRY_FLAG_OPAQUE | GEOMETRY_FLAG_DOUBLE_SIDED, 0, 11, 3]],
                        dtype=np.uint32)
    return pack_triangles(pos, identity=identity, metadata=metadata)


def rays_toward(tris, n, seed=0, extent=12.0, t_max=1.0e30):
    """Seeded rays aimed at random points of random triangles of `tris`
    ((m, 44) records), from origins uniform in a cube: (n, 8) float32."""
    rng = np.random.default_rng(seed)
    pos = np.ascontiguousarray(tris).view(np.float32)[:, 0:12].reshape(-1, 3, 4)[:, :, :3].astype(np.float64)
    pick = rng.integers(0, pos.shape[0], size=n)
    b = rng.dirichlet([1.0, 1.0, 1.0], size=n)
    target = np.einsum('nk,nkj->nj', b, pos[pick])
    o = rng.uniform(-extent, extent, size=(n, 3))
    d = target - o
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    d = d.astype(np.float32)
    d = d / np.sqrt((d * d).sum(axis=1, keepdims=True, dtype=np.float32))
    r = np.zeros((n, 8), dtype=np.float32)
    r[:, 0:3] = o.astype(np.float32)
    r[:, 4:7] = d
    r[:, 7] = t_max
    return r


def random_rays(n, seed=0, extent=12.0, t_max=1.0e30):
    """Seeded rays (n, 8) float32 = (o.xyz, tMin, d.xyz, tMax): origins uniform in
    a cube, directions uniform on the sphere (normalized in fp32)."""
    rng = np.random.default_rng(seed)
    o = rng.uniform(-extent, extent, size=(n, 3)).astype(np.float32)
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    d = d.astype(np.float32)
    # Re-normalize in fp32 so |d|^2 is within the 1e-4 validity window.
    d = d / np.sqrt((d * d).sum(axis=1, keepdims=True, dtype=np.float32))
    r = np.zeros((n, 8), dtype=np.float32)
    r[:, 0:3] = o
    r[:, 3] = 0.0
    r[:, 4:7] = d
    r[:, 7] = t_max
    return r
