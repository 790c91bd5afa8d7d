"""The oracle's stable compaction (CompactStable / BlellochExclusiveScan,
integrators/wavefront/src/WavefrontCompactionOracle.cpp:9-81) against the
reference's own test vectors (integrators/wavefront/tests/WavefrontCpuOracleTests.cpp:30-89)
and the closed form of an exclusive scan (the serial prefix sum)."""
import numpy as np
import pytest


def test_reference_example_and_empty(oracle):
    rc, prefix, nxt, shd = oracle.compact_stable(np.zeros((0, 2), np.uint32))
    assert rc == 0 and nxt.size == 0 and shd.size == 0
    flags = np.array([[1, 0], [0, 1], [1, 1], [0, 0], [1, 0]], np.uint32)
    rc, prefix, nxt, shd = oracle.compact_stable(flags)
    assert rc == 0
    assert nxt.tolist() == [0, 2, 4] and shd.tolist() == [1, 2]


def test_reference_stale_capacity_tail(oracle):
    # Only the active prefix (3 of 6) is compacted.
    flags = np.array([[1, 0], [0, 1], [1, 0], [1, 1], [1, 1], [1, 1]], np.uint32)
    rc, _, nxt, shd = oracle.compact_stable(flags[:3])
    assert rc == 0 and nxt.tolist() == [0, 2] and shd.tolist() == [1]


@pytest.mark.parametrize("count", [127, 128, 129, 4097])
def test_reference_boundary_counts(oracle, count):
    idx = np.arange(count, dtype=np.uint32)
    flags = np.stack([(idx % 3 == 0), (idx % 5 == 0)], axis=1).astype(np.uint32)
    rc, _, nxt, shd = oracle.compact_stable(flags)
    assert rc == 0
    assert nxt.tolist() == idx[idx % 3 == 0].tolist()
    assert shd.tolist() == idx[idx % 5 == 0].tolist()


@pytest.mark.parametrize("count,density", [(1, 0.5), (1000, 0.3), (65537, 0.7), (300_001, 0.05)])
def test_blelloch_equals_serial_prefix_sum(oracle, count, density):
    rng = np.random.default_rng(count)
    flags = (rng.random((count, 2)) < density).astype(np.uint32)
    rc, prefix, nxt, shd = oracle.compact_stable(flags)
    assert rc == 0
    serial = np.cumsum(flags, axis=0, dtype=np.uint64) - flags
    assert np.array_equal(prefix, serial.astype(np.uint32))
    assert np.array_equal(nxt, np.nonzero(flags[:, 0])[0].astype(np.uint32))
    assert np.array_equal(shd, np.nonzero(flags[:, 1])[0].astype(np.uint32))


def test_non_binary_flag_is_rejected(oracle):
    flags = np.array([[1, 0], [2, 0], [0, 1]], np.uint32)
    rc, *_ = oracle.compact_stable(flags)
    assert rc == -1
