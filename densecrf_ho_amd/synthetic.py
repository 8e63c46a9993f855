"""Seeded synthetic inputs for dense-CRF inference.

The workloads follow the reference's data layout (``src/densecrf.cpp:171-188``):
pixel index ``y*W + x``, RGB interleaved per pixel, unaries ``L x N`` stored
label-contiguous per pixel (``U[pixel*L + label]``), features ``d x N`` stored
feature-contiguous per pixel (``f[pixel*d + k]``).

Generator: a Voronoi image with ``K`` cells (centres uniform in
``[0,W) x [0,H)``, nearest centre wins), a uniform random RGB colour per cell plus
Gaussian noise of sigma 10, rounded and clipped to uint8.  Unaries are
``-log softmax`` of per-cell logits ``N(0, 2)`` plus per-pixel ``N(0, 0.5)``.
All randomness comes from ``numpy.random.default_rng(seed)``.
"""

from __future__ import annotations

import numpy as np

try:  # scipy is optional: a chunked brute-force search is the fallback
    from scipy.spatial import cKDTree as _KDTree
except Exception:  # pragma: no cover
    _KDTree = None


def voronoi_regions(W: int, H: int, K: int, seed: int) -> np.ndarray:
    """Region id (int32, length W*H) of the nearest of K random centres."""
    rng = np.random.default_rng(seed)
    cx = rng.uniform(0.0, W, K)
    cy = rng.uniform(0.0, H, K)
    centres = np.stack([cx, cy], axis=1)
    ys, xs = np.mgrid[0:H, 0:W]
    pts = np.stack([xs.ravel(), ys.ravel()], axis=1).astype(np.float64)
    if _KDTree is not None:
        _, idx = _KDTree(centres).query(pts, k=1)
        return idx.astype(np.int32)
    out = np.empty(len(pts), np.int32)
    for s in range(0, len(pts), 1 << 16):
        p = pts[s:s + (1 << 16)]
        d2 = ((p[:, None, :] - centres[None, :, :]) ** 2).sum(-1)
        out[s:s + len(p)] = d2.argmin(1)
    return out


def voronoi_image(W: int, H: int, K: int, seed: int, noise: float = 10.0):
    """Return ``(img uint8 [H*W*3], region int32 [H*W])``."""
    region = voronoi_regions(W, H, K, seed)
    rng = np.random.default_rng(seed + 0x5EED)
    colours = rng.integers(0, 256, size=(K, 3)).astype(np.float64)
    img = colours[region] + rng.normal(0.0, noise, size=(W * H, 3))
    img = np.clip(np.rint(img), 0, 255).astype(np.uint8)
    return img.reshape(-1), region


def voronoi_unaries(region: np.ndarray, K: int, L: int, seed: int,
                    cell_sigma: float = 2.0, pixel_sigma: float = 0.5) -> np.ndarray:
    """``-log softmax`` unaries, float32, shape (N, L) (label-contiguous)."""
    rng = np.random.default_rng(seed + 0xC0FFEE)
    cell = rng.normal(0.0, cell_sigma, size=(K, L))
    logits = cell[region] + rng.normal(0.0, pixel_sigma, size=(len(region), L))
    logits -= logits.max(axis=1, keepdims=True)
    lse = np.log(np.exp(logits).sum(axis=1, keepdims=True))
    return (-(logits - lse)).astype(np.float32)


def gaussian_features(W: int, H: int, sx: float, sy: float) -> np.ndarray:
    """``DenseCRF2D::addPairwiseGaussian`` features (src/densecrf.cpp:170-178).

    Shape (N, 2) float32: ``x/sx, y/sy`` with a float32 division."""
    ys, xs = np.mgrid[0:H, 0:W]
    f = np.empty((H * W, 2), np.float32)
    f[:, 0] = xs.ravel().astype(np.float32) / np.float32(sx)
    f[:, 1] = ys.ravel().astype(np.float32) / np.float32(sy)
    return f


def bilateral_features(W: int, H: int, sx: float, sy: float, sr: float, sg: float, sb: float,
                       img: np.ndarray) -> np.ndarray:
    """``DenseCRF2D::addPairwiseBilateral`` features (src/densecrf.cpp:179-190).

    Shape (N, 5) float32: ``x/sx, y/sy, r/sr, g/sg, b/sb``."""
    ys, xs = np.mgrid[0:H, 0:W]
    im = np.asarray(img, np.uint8).reshape(H * W, 3).astype(np.float32)
    f = np.empty((H * W, 5), np.float32)
    f[:, 0] = xs.ravel().astype(np.float32) / np.float32(sx)
    f[:, 1] = ys.ravel().astype(np.float32) / np.float32(sy)
    f[:, 2] = im[:, 0] / np.float32(sr)
    f[:, 3] = im[:, 1] / np.float32(sg)
    f[:, 4] = im[:, 2] / np.float32(sb)
    return f


def softmax_init(U: np.ndarray) -> np.ndarray:
    """``expAndNormalize(-U)`` (src/densecrf.cpp:321-336), float32."""
    b = -U.astype(np.float32)
    b = b - b.max(axis=1, keepdims=True)
    e = np.exp(b)
    return (e / e.sum(axis=1, keepdims=True, dtype=np.float32)).astype(np.float32)


class Workload:
    """One synthetic image: image, unaries and the two standard kernels.

    Kernels are those of the example inference: spatial Gaussian (w=3, sxy=3)
    and bilateral (w=5, sxy=50, srgb=15), Potts compatibility."""

    def __init__(self, W: int, H: int, K: int, seed: int, L: int = 21,
                 spatial=(3.0, 3.0), bilateral=(5.0, 50.0, 15.0)):
        self.W, self.H, self.K, self.seed, self.L = W, H, K, seed, L
        self.N = W * H
        self.img, self.region = voronoi_image(W, H, K, seed)
        self.U = voronoi_unaries(self.region, K, L, seed)
        ws, ss = spatial
        wb, sxy, srgb = bilateral
        self.kernels = [
            (gaussian_features(W, H, ss, ss), float(ws)),
            (bilateral_features(W, H, sxy, sxy, srgb, srgb, srgb, self.img), float(wb)),
        ]


def superpixel_layers(W: int, H: int, cells=(300, 600, 1200), seeds=(21, 22, 23)):
    """Voronoi region maps standing in for mean-shift superpixel layers."""
    return [(voronoi_regions(W, H, k, s), k) for k, s in zip(cells, seeds)]


def batch_size(i: int):
    """(W, H) of image i of the batched workload: seed 3000+i,
    H ~ U[240, 1080], W ~ U[320, 1920]."""
    rng = np.random.default_rng(3000 + i)
    H = int(rng.integers(240, 1081))
    W = int(rng.integers(320, 1921))
    return W, H


def batch_image(i: int, L: int = 21):
    """Image i of the batched workload of 64 uneven images (sizes from
    batch_size), one Voronoi cell per 4000 pixels.  Returns (W, H, RGB image,
    unaries)."""
    W, H = batch_size(i)
    wl = Workload(W, H, max(1, round(W * H / 4000)), 3000 + i, L=L)
    return W, H, wl.img, wl.U


def batch_images(n: int, workers: int = 16):
    """The first n images of the batched workload, generated in worker
    processes (at most `workers`, and no more than the host's CPUs)."""
    import multiprocessing as mp
    import os
    w = max(1, min(workers, os.cpu_count() or 1, n))
    if w == 1:
        return [batch_image(i) for i in range(n)]
    with mp.get_context("spawn").Pool(w) as pool:
        return pool.map(batch_image, range(n))

