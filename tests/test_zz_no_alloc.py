"""s2c_depth_test_noalloc (the depth test with a static key buffer, for
callers that cannot allocate) against s2c_depth_test, through the C ABI.

Bar: bit-exact (the depth keys are unique, so the winners do not depend on
the order of the atomics).  Two calls in a row share the static buffer; the
second is smaller, so a stale key left by the first would show.
"""
import numpy as np
import pytest
import torch

import oracle
from capi import CapiError, call
from conftest import report
from strns2cards_amd import synth

pytestmark = pytest.mark.gpu


def _fragments(dev, res, B, seed):
    """Triangle fragments of B views of the small workload's cards, each view
    shifted and rescaled so the views differ."""
    w = synth.small_workload()
    rng = np.random.default_rng(seed)
    pos = np.repeat(w.card_pos[:1], B, axis=0).copy()
    for b in range(B):
        pos[b, :, :2] = pos[b, :, :2] * rng.uniform(0.7, 1.2) + rng.uniform(-0.2, 0.2, 2)
        pos[b, :, 2] += rng.uniform(-0.1, 0.1, pos.shape[1]).astype(np.float32)
    tri = w.cards.tri
    fp, fa = oracle.compute_fragments(res, pos, tri)
    return (torch.from_numpy(fp).to(dev), torch.from_numpy(fa).to(dev),
            oracle.depth_test(B, res, fp, fa))


def _both(dev, B, res, fp, fa):
    got = torch.zeros((B, *res, 4), device=dev)  # rast_out is zero-filled by the caller
    want = torch.zeros((B, *res, 4), device=dev)
    call("s2c_depth_test_noalloc", B, res[0], res[1], fp.shape[0], fp, fa, got)
    call("s2c_depth_test", B, res[0], res[1], fp.shape[0], fp, fa, want)
    torch.cuda.synchronize()
    return got.cpu().numpy(), want.cpu().numpy()


def test_depth_test_noalloc_matches_allocating_form(dev):
    report("no-alloc depth test: 3 views at 256 x 256, then 2 views at 128 x 96")
    res1 = (256, 256)
    fp1, fa1, o1 = _fragments(dev, res1, 3, 1)
    assert fp1.shape[0] > 1000
    got, want = _both(dev, 3, res1, fp1, fa1)
    np.testing.assert_array_equal(got, want)
    np.testing.assert_array_equal(got, o1)
    res2 = (128, 96)
    fp2, fa2, o2 = _fragments(dev, res2, 2, 2)
    assert fp2.shape[0] > 100
    got, want = _both(dev, 2, res2, fp2, fa2)
    np.testing.assert_array_equal(got, want)
    np.testing.assert_array_equal(got, o2)
    # no fragments at all: every pixel empty
    empty_fp = torch.zeros((0, 3), dtype=torch.int32, device=dev)
    empty_fa = torch.zeros((0, 4), dtype=torch.float32, device=dev)
    got, want = _both(dev, 1, (64, 64), empty_fp, empty_fa)
    assert not got.any() and not want.any()


def test_depth_test_noalloc_full_buffer_and_limit(dev):
    report("no-alloc depth test: 4 views at 1024 x 1024 (the buffer's size), then one pixel more")
    res = (1024, 1024)
    fp, fa, o = _fragments(dev, res, 4, 3)
    got, want = _both(dev, 4, res, fp, fa)
    np.testing.assert_array_equal(got, want)
    np.testing.assert_array_equal(got, o)
    rast = torch.zeros((1, 4, 4), device=dev)  # never written: the call is refused first
    with pytest.raises(CapiError) as e:
        call("s2c_depth_test_noalloc", 4, 1024, 1025, 0, None, None, rast)
    assert e.value.code == -1 and "exceeds the key buffer" in str(e.value)
