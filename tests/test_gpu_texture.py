"""GPU parity of the software-MSAA texture patches against the oracle.

The oracle rasterizes each jitter separately (as the reference loops over
64 rasterize calls) with the C raster oracle and accumulates in numpy in
sample order; the GPU path depth-tests the 64 jitters in one fused kernel
(or rasterizes them as one batched call) and resolves in one kernel.  rgb, the weighted coverage and its blur (in
the oracle's summation order) are bit-exact, and so are the atlases.
"""
import numpy as np
import pytest
import torch

import oracle
from strns2cards_amd import texture

pytestmark = pytest.mark.gpu


def _wisp(seed, n=4, s=14):
    rng = np.random.default_rng(seed)
    base = np.cumsum(rng.normal(size=(s, 3)) * [0.02, 0.08, 0.02] + [0, 0.1, 0], 0)
    return np.stack([base + rng.normal(size=3) * 0.05 for _ in range(n)]).astype(np.float32)


def _mesh(seed, uv=(0.1, 0.2, 0.4, 0.9)):
    wisp = _wisp(seed)
    p = wisp.reshape(-1, 3)
    bbox = p[:, :2].max(0) - p[:, :2].min(0)
    scale = 2.0 * np.ones(3) / (p.max(0) - p.min(0))
    scale[2] *= 0.2
    return texture.wisp_ndc_mesh(wisp, uv, bbox, scale, 0.01, 6)


@pytest.mark.parametrize("fused", [True, False], ids=["fused", "batched_rasterize"])
@pytest.mark.parametrize("seed,H,W", [(0, 24, 16), (1, 40, 33), (2, 7, 5)])
def test_patch_matches_oracle(dev, seed, H, W, fused):
    print(f"[msaa patch] seed={seed} {H}x{W} fused={fused}", flush=True)
    verts, faces, attr = _mesh(seed)
    rgb, alpha = texture.render_patch(torch.from_numpy(verts.astype(np.float32)).to(dev),
                                      torch.from_numpy(faces).to(dev),
                                      torch.from_numpy(attr.astype(np.float32)).to(dev), (H, W),
                                      fused=fused)
    offs, w = texture.msaa_samples()
    dx = np.array([o[0] * 2.0 / W for o in offs], np.float32)
    dy = np.array([o[1] * 2.0 / H for o in offs], np.float32)
    want_rgb, want_alpha = oracle.msaa_patch(verts, faces, attr, H, W, dx, dy,
                                             np.array(w, np.float32), texture.gaussian_kernel3(1.5))
    np.testing.assert_array_equal(rgb.cpu().numpy(), want_rgb)
    np.testing.assert_array_equal(alpha.cpu().numpy(), want_alpha)
    assert want_alpha.max() > 0.1  # the mesh is in view


def _paste_np(rgb, alpha, rect, R, atlas_rgb, atlas_alpha):
    """texture_utils.py:256-276 in numpy."""
    cx0, cy0, x0, y0, x1, y1 = rect
    f_rgb = np.clip(rgb[::-1], 0, 1)
    f_alpha = np.clip(alpha[::-1], 0, 1)
    atlas_rgb[y0:y1, x0:x1] = (f_rgb * 255).astype(np.uint8)[cy0:cy0 + y1 - y0, cx0:cx0 + x1 - x0]
    atlas_alpha[y0:y1, x0:x1] = (f_alpha * 255).astype(np.uint8)[cy0:cy0 + y1 - y0,
                                                                 cx0:cx0 + x1 - x0]


def test_render_texture_atlas(dev):
    wisps = [_wisp(s, 3, 10) for s in range(4)]
    uv = [(0, 0.0, 0.0, 0.3, 0.6), (1, 0.3, 0.0, 0.55, 0.5), (2, 0.55, 0.0, 0.9, 0.62),
          (3, 0.0, 0.62, 0.28, 0.98)]
    R = 64
    got_rgb, got_alpha = texture.render_texture(wisps, uv, 0.01, 6, resolution=R,
                                                output_dirname=None, device=dev)
    # the same per-wisp meshes (host glue shared), oracle patches, numpy paste
    max_u = max(u1 - u0 for _, u0, _, u1, _ in uv)
    max_v = max(v1 - v0 for _, _, v0, _, v1 in uv)
    H, W = int(np.ceil(max_v * R)), int(np.ceil(max_u * R))
    pts = np.concatenate([w.reshape(-1, 3) for w in wisps])
    bbox = pts[:, :2].max(0) - pts[:, :2].min(0)
    size = np.zeros(3)
    for w in wisps:
        p = w.reshape(-1, 3)
        size = np.maximum(size, p.max(0) - p.min(0))
    scale = 2.0 * np.ones(3) / size
    scale[2] *= 0.2
    offs, wts = texture.msaa_samples()
    dx = np.array([o[0] * 2.0 / W for o in offs], np.float32)
    dy = np.array([o[1] * 2.0 / H for o in offs], np.float32)
    want_rgb = np.zeros((R, R, 3), np.uint8)
    want_alpha = np.zeros((R, R), np.uint8)
    for wisp, (_, u0, v0, u1, v1) in zip(wisps, uv):
        verts, faces, attr = texture.wisp_ndc_mesh(wisp, (u0, v0, u1, v1), bbox, scale, 0.01, 6)
        rgb, alpha = oracle.msaa_patch(verts, faces, attr, H, W, dx, dy, np.array(wts, np.float32),
                                       texture.gaussian_kernel3(1.5))
        _paste_np(rgb, alpha, texture.atlas_rect((u0, v0, u1, v1), R, (H, W)), R, want_rgb,
                  want_alpha)
    np.testing.assert_array_equal(got_rgb, want_rgb)
    np.testing.assert_array_equal(got_alpha, want_alpha)
    assert got_alpha.max() > 128


def test_paste_bounds_are_checked(dev):
    rgb = torch.zeros((8, 8, 3), device=dev)
    alpha = torch.zeros((8, 8), device=dev)
    a_rgb = torch.zeros((16, 16, 3), dtype=torch.uint8, device=dev)
    a_alpha = torch.zeros((16, 16), dtype=torch.uint8, device=dev)
    from strns2cards_amd import _core
    with pytest.raises(RuntimeError, match="crop"):
        _core.atlas_paste(rgb, alpha, 4, 0, 0, 0, 10, 4, a_rgb, a_alpha)
    with pytest.raises(RuntimeError, match="outside"):
        _core.atlas_paste(rgb, alpha, 0, 0, 12, 0, 20, 4, a_rgb, a_alpha)
