"""Seeded synthetic inputs of the benchmark workloads (no datasets offline).

tube_volume: float32 volume of random 3-D polyline tubes (Gaussian cross
section, peak 1, combined by maximum) plus N(0, noise) -- the 3-D Frangi
workloads.  rgb_tiles: uint8 RGB tiles of uneven square sizes with 8x8
block-smoothed structure -- the batched 2-D workload.

Tube parameters come from numpy's seeded generator; the distance fields are
evaluated on the device inside each segment's bounding box.
"""

import numpy as np
import torch


def tube_volume(shape, ntubes=40, seed=0, noise=0.05, radius=(1.0, 6.0), segments=(4, 8),
                device=None):
    shape = tuple(int(s) for s in shape)
    dev = torch.device("cuda", torch.cuda.current_device()) if device is None else device
    rng = np.random.default_rng(seed)
    vol = torch.zeros(shape, dtype=torch.float32, device=dev)
    ext = np.array(shape, dtype=np.float64)
    for _ in range(ntubes):
        r = rng.uniform(*radius)
        nseg = int(rng.integers(segments[0], segments[1] + 1))
        p = rng.uniform(0, 1, 3) * ext
        pts = [p]
        for _ in range(nseg):
            d = rng.normal(size=3)
            d /= np.linalg.norm(d)
            step = rng.uniform(0.15, 0.4) * ext.min()
            p = np.clip(p + d * step, 0, ext - 1)
            pts.append(p)
        for a, b in zip(pts[:-1], pts[1:]):
            pad = 4.0 * r
            lo = np.maximum(np.floor(np.minimum(a, b) - pad), 0).astype(int)
            hi = np.minimum(np.ceil(np.maximum(a, b) + pad) + 1, ext).astype(int)
            if np.any(hi <= lo):
                continue
            axes = [torch.arange(lo[k], hi[k], device=dev, dtype=torch.float32) for k in range(3)]
            z, y, x = torch.meshgrid(*axes, indexing="ij")
            P = torch.stack([z, y, x], -1)
            A = torch.tensor(a, dtype=torch.float32, device=dev)
            D = torch.tensor(b - a, dtype=torch.float32, device=dev)
            dd = float((b - a) @ (b - a))
            t = ((P - A) @ D) / dd if dd > 0 else torch.zeros_like(z)
            t = t.clamp(0, 1)
            C = A + t[..., None] * D
            d2 = ((P - C) ** 2).sum(-1)
            prof = torch.exp(-d2 / (2 * r * r))
            sl = tuple(slice(lo[k], hi[k]) for k in range(3))
            vol[sl] = torch.maximum(vol[sl], prof)
    if noise > 0:
        g = torch.Generator(device=dev)
        g.manual_seed(int(seed) + 1)
        vol += noise * torch.randn(shape, generator=g, device=dev, dtype=torch.float32)
    return vol


def rgb_tiles(n=512, edges=(256, 384, 512, 768, 1024), seed=0, device=None):
    """List of (e, e, 3) uint8 tiles: 8x8 block means of U[0, 255] noise,
    upsampled, plus U noise."""
    dev = torch.device("cuda", torch.cuda.current_device()) if device is None else device
    rng = np.random.default_rng(seed)
    sizes = rng.choice(np.asarray(edges), size=n)
    g = torch.Generator(device=dev)
    g.manual_seed(int(seed) + 7)
    tiles = []
    for e in sizes:
        e = int(e)
        coarse = torch.rand((e // 8, e // 8, 3), generator=g, device=dev) * 255
        up = coarse.repeat_interleave(8, 0).repeat_interleave(8, 1)
        fine = torch.rand((e, e, 3), generator=g, device=dev) * 32 - 16
        tiles.append((up + fine).clamp(0, 255).to(torch.uint8).contiguous())
    return tiles
