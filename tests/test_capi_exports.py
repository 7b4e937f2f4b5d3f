"""CPU checks of the boundary: the HIP library loads, exports every entry
point the C header declares, and the Python package exposes the reference's
API names."""
import ctypes
import subprocess

import capi


def test_header_declares_reference_launchers():
    names = set(capi.declarations())
    # pystrns2cards.cpp:13-192, one entry point per launcher
    reference = {
        "compute_triangle_rects", "compute_fragments", "depth_test", "filter_valid_fragments",
        "interpolate_triangle_attributes", "backward_interpolate_triangle_attributes",
        "cluster_mask_from_fragments", "accumulate_bitset_rgb", "popcount_bitset",
        "dda_compute_span", "dda_compute_fragments", "dda_interpolate_attributes",
        "backward_dda_interpolate_attributes", "mark_discontinuity_edges",
        "backward_antialiased_cluster_bitset", "compact_valid_ints", "smooth_strands",
        "skinning",
    }
    assert {"s2c_" + n for n in reference} <= names


def test_library_exports_every_declared_symbol():
    out = subprocess.run(["nm", "-D", "--defined-only", capi.LIB], check=True,
                         capture_output=True, text=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if line.strip()}
    missing = set(capi.declarations()) - exported
    assert not missing, missing


def test_library_loads_and_reports_version():
    lib = capi.lib()
    for name in capi.declarations():
        assert hasattr(lib, name)
    assert lib.s2c_version() == 100
    assert isinstance(lib.s2c_last_error(), bytes)


def test_argument_errors_need_no_gpu():
    # Negative sizes are rejected before any HIP call.
    lib = capi.lib()
    n = ctypes.c_int(5)
    rc = lib.s2c_compact_valid_ints(-1, None, None, ctypes.byref(n), None)
    assert rc == -1 and n.value == 0
    assert b"negative" in lib.s2c_last_error()


def test_python_api_surface():
    import strns2cards_amd as s2c

    for name in [
        "dda_compute_fragments", "DDAInterpolateAttributes", "dda_interpolate_attributes",
        "compute_fragments", "depth_test", "rasterize", "InterpolateTriangleAttributes",
        "interpolate_triangle_attributes", "cluster_mask_from_fragments",
        "accumulate_bitset_rgb", "popcount_bitset", "mark_discontinuity_edges",
        "antialiased_cluster_bitset_loss", "compute_arclength", "smooth_strands", "skinning",
        "print_version",
    ]:
        assert hasattr(s2c, name), name


def test_core_exports_every_reference_binding_name():
    """The names NB_MODULE(_core) defines at pystrns2cards.cpp:204-924."""
    from strns2cards_amd import _core

    reference = [
        "print_version", "build_edges_from_triangles", "compute_triangle_rects",
        "compute_fragments", "depth_test", "filter_valid_fragments",
        "interpolate_triangle_attributes", "backward_interpolate_triangle_attributes",
        "cluster_mask_from_fragments", "accumulate_bitset_rgb", "popcount_bitset",
        "dda_compute_span", "dda_compute_fragments", "dda_interpolate_attributes",
        "backward_dda_interpolate_attributes", "mark_discontinuity_edges",
        "backward_antialiased_cluster_bitset", "compact_valid_ints", "load_strands",
        "resample_strands_by_arclength", "smooth_strands", "skinning",
    ]
    missing = [n for n in reference if not hasattr(_core, n)]
    assert not missing, missing


def test_build_id_matches_the_sources():
    """Both libraries carry the hash of the sources they were built from, and
    it is that of the sources in the tree (the package refuses a stale
    build at import)."""
    import ctypes
    import importlib.util
    import os

    import strns2cards_amd as s2c

    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    spec = importlib.util.spec_from_file_location(
        "_s2c_build_check", os.path.join(root, "strns2cards_amd", "build.py"))
    build = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(build)
    lib = ctypes.CDLL(s2c.LIBRARY_PATH)
    lib.s2c_build_id.restype = ctypes.c_char_p
    want = build.source_hash()
    assert len(want) == 16
    assert lib.s2c_build_id().decode() == want
    assert s2c._core.build_id == want
