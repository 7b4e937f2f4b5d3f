"""The torch-only curve helpers against vectors the reference's own
curve_utils.py wrote (tests/golden/make_curve_vectors.py; the reference is
not read here): tangent_frame and minimal_rotation run on host tensors and
must match bit for bit (same float32 operations in the same order)."""
import os

import numpy as np
import torch

from strns2cards_amd import curve_utils as cu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "curve_vectors.npz")


def golden():
    return np.load(GOLDEN, allow_pickle=False)


def test_golden_file_loads_without_pickle():
    g = golden()
    assert "tf_in" in g.files and "aws_aligned" in g.files
    assert str(g["versions"][0]).startswith("numpy")


def test_tangent_frame_matches_reference():
    """curve_utils.py:76-102, including z = 0 (sign forced to +1), z = -1,
    unnormalized and tiny tangents."""
    g = golden()
    n, b = cu.tangent_frame(torch.from_numpy(g["tf_in"]))
    np.testing.assert_array_equal(n.numpy(), g["tf_normal"])
    np.testing.assert_array_equal(b.numpy(), g["tf_binormal"])


def test_minimal_rotation_matches_reference():
    """curve_utils.py:9-74: random pairs, parallel (identity), anti-parallel
    (the canonical 180-degree matrix), near anti-parallel, zero source."""
    g = golden()
    R = cu.minimal_rotation(torch.from_numpy(g["mr_src"]), torch.from_numpy(g["mr_dst"]))
    np.testing.assert_array_equal(R.numpy(), g["mr_R"])
    np.testing.assert_array_equal(g["mr_R"][1], np.diag([-1.0, -1.0, 1.0]).astype(np.float32))
