"""GPU test of arclength resampling (hair_loader.cpp:261-307) against the
oracle: bit-exact, both use the reference's fp32 order of operations."""
import numpy as np
import pytest

import oracle
from conftest import report
from strns2cards_amd import strand_io as sio
from strns2cards_amd import synth

pytestmark = pytest.mark.gpu


def test_resample_bit_exact(dev):
    report("strand io: arclength resampling vs oracle")
    rng = np.random.default_rng(0)
    counts = rng.integers(2, 300, size=3000)
    pts = np.cumsum(rng.normal(size=(int(counts.sum()), 3)).astype(np.float32) * 0.01, 0)
    pts[5] = pts[4]  # a zero-length segment
    for target in (2, 17, 100):
        got = sio.resample_strands_by_arclength(counts, pts, target)
        want = oracle.resample_strands_by_arclength(counts, pts, target)
        np.testing.assert_array_equal(got, want)


def test_resample_workload_strands(dev):
    x = synth.generate_strands(500, 64, 0.004, 3)
    ns = np.full(500, 64, np.int32)
    got = sio.resample_strands_by_arclength(ns, x.reshape(-1, 3), 21)
    want = oracle.resample_strands_by_arclength(ns, x.reshape(-1, 3), 21)
    np.testing.assert_array_equal(got, want)
    assert got.shape == (500, 21, 3)
    with pytest.raises(ValueError):
        sio.resample_strands_by_arclength([1, 3], np.zeros((4, 3), np.float32), 5)
    empty = sio.resample_strands_by_arclength(np.zeros(0, np.int32), np.zeros((0, 3)), 5)
    assert empty.shape == (0, 5, 3)


def test_core_resample_host_arrays(dev):
    """_core.resample_strands_by_arclength takes and returns host arrays, as
    the reference's binding does (pystrns2cards.cpp:818-839)."""
    from strns2cards_amd import _core

    report("strand io: _core.resample_strands_by_arclength on host arrays")
    rng = np.random.default_rng(4)
    counts = rng.integers(2, 90, size=700).astype(np.int32)
    pts = np.cumsum(rng.normal(size=(int(counts.sum()), 3)).astype(np.float32) * 0.02, 0)
    got = _core.resample_strands_by_arclength(counts, pts, 33)
    assert isinstance(got, np.ndarray) and got.shape == (700, 33, 3) and got.dtype == np.float32
    np.testing.assert_array_equal(got, oracle.resample_strands_by_arclength(counts, pts, 33))
    with pytest.raises(ValueError):
        _core.resample_strands_by_arclength(np.array([1, 3], np.int32),
                                           np.zeros((4, 3), np.float32), 5)
    with pytest.raises(TypeError):
        _core.resample_strands_by_arclength(counts.astype(np.int64), pts, 5)


def test_npz_arc_length(dev, tmp_path):
    rng = np.random.default_rng(1)
    pts = rng.normal(size=(10, 3)).astype(np.float32)
    p = tmp_path / "s.npz"
    np.savez(p, num_samples=np.array([4, 6], np.int32), points=pts)
    got = sio.load_npz_as_strands_arc_length(str(p), 8)
    np.testing.assert_array_equal(got, oracle.resample_strands_by_arclength([4, 6], pts, 8))
