"""Writes tests/golden/curve_vectors.npz: input/output vectors of the
reference's torch-only curve helpers (pystrns2cards/curve_utils.py:9-102,
242-478, 481-644), computed by the reference module itself.

Run in a container that holds the reference checkout (default
/root/reference; pass another root as the first argument).  The module is
loaded by path under a stand-in package whose `_core` carries the two
kernels the helpers call (smooth_strands, skinning), backed by this repo's
CPU oracle, so the module needs no build.  align_wisp(s) insist on CUDA
tensors; they get host tensors of a subclass that reports is_cuda.  Only
the vectors are committed; neither this script's outputs nor the tests
read the reference at run time.
"""
import importlib.util
import os
import sys
import types

import numpy as np
import torch

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import oracle  # noqa: E402


class _Core(types.ModuleType):
    """The two reference kernels align_wisp(s) reach (curve_utils.py:205, 239)
    on host tensors, through the CPU oracle (curve_utils.cu:24-131)."""

    @staticmethod
    def smooth_strands(strands, arclength, out, beta):
        out.copy_(torch.from_numpy(oracle.smooth_strands(strands.numpy(), arclength.numpy(),
                                                         float(beta))))

    @staticmethod
    def skinning(strands, arclength, handles, canonical, R, out, beta):
        out.copy_(torch.from_numpy(oracle.skinning(strands.numpy(), arclength.numpy(),
                                                   handles.numpy(), canonical.numpy(),
                                                   R.numpy(), float(beta))))


class HostCuda(torch.Tensor):
    """A host tensor that passes the reference's `is_cuda` checks."""

    @property
    def is_cuda(self):
        return True

    def cpu(self):
        return torch.Tensor(self.as_subclass(torch.Tensor))


def load_reference(root):
    pkg = types.ModuleType("_s2c_reference")
    pkg.__path__ = []
    sys.modules["_s2c_reference"] = pkg
    sys.modules["_s2c_reference._core"] = _Core("_s2c_reference._core")
    path = os.path.join(root, "pystrns2cards", "curve_utils.py")
    spec = importlib.util.spec_from_file_location("_s2c_reference.curve_utils", path)
    mod = importlib.util.module_from_spec(spec)
    sys.modules[spec.name] = mod
    spec.loader.exec_module(mod)
    return mod


def strands(n, v, seed):
    """Seeded random-walk strands (n, v, 3) float32 with roots spread on a
    small patch and a different step per strand (lengths 0.06-0.3 apart by
    far more than rounding, so the longest strand is well defined)."""
    rng = np.random.default_rng(seed)
    roots = rng.normal(scale=0.02, size=(n, 3))
    step = rng.uniform(0.004, 0.02, size=(n, 1, 1))
    d = np.array([0.0, -1.0, 0.2]) + rng.normal(scale=0.3, size=(n, v - 1, 3))
    d /= np.linalg.norm(d, axis=-1, keepdims=True)
    x = np.concatenate([roots[:, None], roots[:, None] + np.cumsum(step * d, axis=1)], 1)
    return x.astype(np.float32)


def main(root="/root/reference"):
    cu = load_reference(root)
    out = {}
    rng = np.random.default_rng(2024)
    # tangent_frame (curve_utils.py:76-102): random directions, z = 0 (sign
    # forced to +1), z = -1, unnormalized and tiny vectors
    t = rng.normal(size=(64, 3)).astype(np.float32)
    t[0] = [0.3, -0.4, 0.0]
    t[1] = [0.0, 0.0, -1.0]
    t[2] = [0.0, 0.0, 1.0]
    t[3] = [5.0, 1.0, -2.0]
    t[4] = [1e-9, 0.0, 0.0]
    nrm, bnm = cu.tangent_frame(torch.from_numpy(t))
    out.update(tf_in=t, tf_normal=nrm.numpy(), tf_binormal=bnm.numpy())
    # minimal_rotation (curve_utils.py:9-74): random pairs, parallel,
    # anti-parallel (the canonical 180-degree matrix), near anti-parallel,
    # a zero source
    a = rng.normal(size=(48, 3)).astype(np.float32)
    b = rng.normal(size=(48, 3)).astype(np.float32)
    b[0] = 3 * a[0]
    b[1] = -a[1]
    b[2] = -2 * a[2]
    b[3] = -a[3] + np.float32(1e-4) * rng.normal(size=3).astype(np.float32)
    a[4] = 0.0
    R = cu.minimal_rotation(torch.from_numpy(a), torch.from_numpy(b))
    out.update(mr_src=a, mr_dst=b, mr_R=R.numpy())
    # compute_alignment_rotation (curve_utils.py:272-308)
    x = strands(6, 16, 7)
    Td = np.array([[0.0, -1.0, 0.0]], np.float32)
    out.update(ar_strands=x, ar_T=Td,
               ar_R=cu.compute_alignment_rotation(torch.from_numpy(x), torch.from_numpy(Td)).numpy())
    # strands_to_tube (curve_utils.py:310-372), with an attribute
    x = strands(3, 10, 8)
    attr = rng.normal(size=(3, 10, 2)).astype(np.float32)
    v, f, ea = cu.strands_to_tube(x, radius=0.5, n_circle=8, attr=attr)
    out.update(tube_strands=x, tube_attr=attr, tube_verts=v, tube_faces=f, tube_attr_out=ea)
    v2, f2 = cu.strands_to_tube(x, radius=0.25, n_circle=5)
    out.update(tube5_verts=v2, tube5_faces=f2)
    # align_wisp (curve_utils.py:374-478): 1 strand (rank-0 roots), 2, 3 and 8
    beta = 0.05
    for n in (1, 2, 3, 8):
        x = strands(n, 16, 100 + n)
        al, hd, li = cu.align_wisp(torch.from_numpy(x).as_subclass(HostCuda), beta)
        out[f"aw{n}_in"] = x
        out[f"aw{n}_aligned"] = torch.Tensor(al.as_subclass(torch.Tensor)).numpy()
        out[f"aw{n}_handles"] = torch.Tensor(hd.as_subclass(torch.Tensor)).numpy()
        out[f"aw{n}_longest"] = np.int64(li)
    # align_wisps (curve_utils.py:481-607): wisps of 1, 2, 5 and 9 strands
    wisps = [strands(n, 16, 200 + n) for n in (1, 2, 5, 9)]
    to_cuda = torch.Tensor.to

    def host_to(self, *args, **kw):  # the reference moves its inputs to "cuda"
        dev = str(kw.pop("device", ""))
        cuda = dev.startswith("cuda") or any(isinstance(a, str) and a.startswith("cuda")
                                             for a in args)
        args = tuple(a for a in args if not (isinstance(a, str) and a.startswith("cuda")))
        r = to_cuda(self, *args, **kw)
        return r.as_subclass(HostCuda) if cuda else r

    orig_tensor = torch.tensor

    def host_tensor(*args, **kw):
        kw.pop("device", None)
        return orig_tensor(*args, **kw)

    torch.Tensor.to, torch.tensor = host_to, host_tensor
    try:
        al, hd, li = cu.align_wisps(wisps, beta)
    finally:
        torch.Tensor.to, torch.tensor = to_cuda, orig_tensor
    for k, w in enumerate(wisps):
        out[f"aws_in{k}"] = w
    plain = lambda t: torch.Tensor(t.as_subclass(torch.Tensor)).numpy()  # noqa: E731
    out.update(aws_aligned=plain(al), aws_handles=plain(hd), aws_longest=plain(li),
               aw_beta=np.float32(beta))
    # statistical_outlier_mask (curve_utils.py:610-644): a cluster with
    # outliers, and the N <= 1 case
    p = rng.normal(scale=0.05, size=(200, 3)).astype(np.float32)
    p[:6] += rng.normal(scale=1.0, size=(6, 3)).astype(np.float32)
    out.update(som_points=p, som_mask=cu.statistical_outlier_mask(p, 5, 2.0),
               som_mask_k3=cu.statistical_outlier_mask(p, 3, 1.5),
               som_one=cu.statistical_outlier_mask(p[:1], 5, 2.0))
    out["versions"] = np.array([f"numpy {np.__version__}", f"torch {torch.__version__}"])
    np.savez_compressed(os.path.join(HERE, "curve_vectors.npz"), **out)
    print(f"wrote {len(out)} arrays to {os.path.join(HERE, 'curve_vectors.npz')}")


if __name__ == "__main__":
    main(*sys.argv[1:])
