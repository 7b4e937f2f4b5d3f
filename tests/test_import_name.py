"""The reference's import name resolves to this package (no GPU needed)."""
import importlib

# Every s2c.* name the reference scripts use (1_cluster.py, 2_texture.py,
# 3_fitting.py, texture_utils.py), except the sklearn clustering helpers of
# pystrns2cards/cluster.py, which are not provided.
SCRIPT_NAMES = [
    "accumulate_bitset_rgb", "align_wisps", "antialiased_cluster_bitset_loss",
    "build_edge_indices_from_wisp", "build_edges_from_triangles", "cluster_mask_from_fragments",
    "compute_fragments", "dda_compute_fragments", "dda_interpolate_attributes", "depth_test",
    "interpolate_triangle_attributes", "load_npz_as_strands", "load_npz_as_strands_arc_length",
    "popcount_bitset", "print_version", "rasterize", "resample_single_strand_by_arclength",
    "save_obj", "statistical_outlier_mask", "strands_to_tube",
]
# The rest of the API of the five hot-path modules (pystrns2cards/*.py).
MODULE_NAMES = [
    "mark_discontinuity_edges", "compute_arclength", "smooth_strands", "skinning",
    "minimal_rotation", "tangent_frame", "compute_bishop_frame", "compute_aligned_strands",
    "compute_alignment_rotation", "align_wisp", "load_strands", "resample_strands_by_arclength",
    "load_strands_arc_length",
]


def test_reference_import_name():
    import strns2cards_amd
    import pystrns2cards as s2c

    for name in SCRIPT_NAMES + MODULE_NAMES:
        assert getattr(s2c, name) is getattr(strns2cards_amd, name), name
    assert s2c.optimizer.UniformAdam is strns2cards_amd.optimizer.UniformAdam
    assert callable(s2c.optimizer.laplacian_edges)
    assert s2c._core is strns2cards_amd._core


def test_reference_submodule_paths():
    import strns2cards_amd
    import pystrns2cards

    for sub in ("dda", "rasterize", "antialias", "bitset_utils", "curve_utils", "optimizer"):
        mod = importlib.import_module(f"pystrns2cards.{sub}")
        assert mod is importlib.import_module(f"strns2cards_amd.{sub}")
    from pystrns2cards.rasterize import InterpolateTriangleAttributes, cluster_mask_from_fragments

    assert cluster_mask_from_fragments is strns2cards_amd.cluster_mask_from_fragments
    assert InterpolateTriangleAttributes is importlib.import_module(
        "strns2cards_amd.rasterize").InterpolateTriangleAttributes
    # the package attribute is the function, as the reference's star import leaves it
    assert callable(pystrns2cards.rasterize) and pystrns2cards.rasterize is strns2cards_amd.rasterize
