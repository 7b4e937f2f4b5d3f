"""The GPU curve helpers against vectors the reference's own curve_utils.py
wrote on the host (tests/golden/make_curve_vectors.py): the Bishop-frame
based alignment rotations and tubes, align_wisp (1, 2, 3 and 8 strands,
the one-strand case with its rank-0 roots), align_wisps (wisps of 1, 2, 5
and 9 strands) and the outlier mask.  Float outputs agree to 2e-6 absolute
(float32 chains of frame recurrences, rotations and Gaussian-weighted sums,
values of order 1, summed in another order on the GPU: about 16 ulp;
measured at most 7.4e-7); indices, faces and masks exactly."""
import os

import numpy as np
import pytest
import torch

from conftest import report
from strns2cards_amd import curve_utils as cu

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "curve_vectors.npz")
ATOL = 2e-6


def close(got, want, what):
    got = got.detach().cpu().numpy() if torch.is_tensor(got) else np.asarray(got)
    err = float(np.abs(got.astype(np.float64) - want).max()) if want.size else 0.0
    report(f"  {what}: max |got - reference| {err:.2e}")
    assert got.shape == want.shape, (what, got.shape, want.shape)
    assert err <= ATOL, (what, err)


@pytest.fixture(scope="module")
def g():
    return np.load(GOLDEN, allow_pickle=False)


def test_alignment_rotation_and_tubes(dev, g):
    report("curves vs reference vectors: alignment rotations, tubes")
    R = cu.compute_alignment_rotation(torch.from_numpy(g["ar_strands"]).to(dev),
                                      torch.from_numpy(g["ar_T"]).to(dev))
    close(R, g["ar_R"], "compute_alignment_rotation")
    v, f, a = cu.strands_to_tube(g["tube_strands"], radius=0.5, n_circle=8, attr=g["tube_attr"])
    close(v, g["tube_verts"], "strands_to_tube verts")
    np.testing.assert_array_equal(f, g["tube_faces"])
    np.testing.assert_array_equal(a, g["tube_attr_out"])
    v, f = cu.strands_to_tube(g["tube_strands"], radius=0.25, n_circle=5)
    close(v, g["tube5_verts"], "strands_to_tube verts (5 sides)")
    np.testing.assert_array_equal(f, g["tube5_faces"])


@pytest.mark.parametrize("n", [1, 2, 3, 8])
def test_align_wisp(dev, g, n):
    report(f"curves vs reference vectors: align_wisp, {n} strand(s)")
    al, hd, li = cu.align_wisp(torch.from_numpy(g[f"aw{n}_in"]).to(dev), float(g["aw_beta"]))
    assert int(li) == int(g[f"aw{n}_longest"])
    close(hd, g[f"aw{n}_handles"], "handles")
    close(al, g[f"aw{n}_aligned"], "aligned")


def test_align_wisps(dev, g):
    report("curves vs reference vectors: align_wisps (1, 2, 5, 9 strands)")
    wisps = [g[f"aws_in{k}"] for k in range(4)]
    al, hd, li = cu.align_wisps(wisps, float(g["aw_beta"]))
    np.testing.assert_array_equal(li.cpu().numpy(), g["aws_longest"])
    close(hd, g["aws_handles"], "handles")
    close(al, g["aws_aligned"], "aligned")


def test_statistical_outlier_mask(dev, g):
    p = g["som_points"]
    np.testing.assert_array_equal(cu.statistical_outlier_mask(p, 5, 2.0), g["som_mask"])
    np.testing.assert_array_equal(cu.statistical_outlier_mask(p, 3, 1.5), g["som_mask_k3"])
    np.testing.assert_array_equal(cu.statistical_outlier_mask(p[:1], 5, 2.0), g["som_one"])
