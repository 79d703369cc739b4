"""Batched rgb2gray -> gaussian -> sobel over uneven images in one launch.

Every output must equal, bit for bit, the three separate skimage-level calls
on the same image (each of which is itself checked bit-exactly against the
oracle), and, on small images, the oracle's float32 emulation of the three
reference calls directly.
"""

import functools

import numpy as np
import pytest
import torch

from oracle import ndimage as on
from oracle import skimage as osk

pytestmark = pytest.mark.gpu


def _rgb(shape, seed):
    rng = np.random.default_rng(seed)
    return torch.from_numpy(rng.integers(0, 256, shape + (3,), dtype=np.uint8)).cuda()


def _separate(x, sigma=1.5, gauss_mode="nearest", gauss_cval=0.0, edge_mode="reflect"):
    from cucim_amd.skimage.color import rgb2gray
    from cucim_amd.skimage.filters import gaussian, sobel

    g = gaussian(rgb2gray(x), sigma=sigma, mode=gauss_mode, cval=gauss_cval)
    return sobel(g, mode=edge_mode)


def _bits(t):
    torch.cuda.synchronize()
    return t.cpu().numpy().view(np.uint32)


SHAPES = [(1, 1), (1, 300), (300, 1), (5, 7), (17, 250), (33, 238), (32, 239), (239, 241),
          (300, 600), (256, 256)]


@pytest.mark.parametrize("rows", [16, 32])
def test_batch_equals_separate_calls(gpu, rows):
    from cucim_amd.batch import GrayGaussSobelBatch

    xs = [_rgb(s, i) for i, s in enumerate(SHAPES)]
    plan = GrayGaussSobelBatch(xs, rows=rows)
    outs = plan.run()
    for x, o in zip(xs, outs):
        np.testing.assert_array_equal(_bits(o), _bits(_separate(x)))


@pytest.mark.parametrize("gauss_mode", ["reflect", "mirror", "nearest", "wrap", "constant"])
@pytest.mark.parametrize("edge_mode", ["reflect", "mirror", "nearest", "constant"])
@pytest.mark.parametrize("sigma", [0.7, 1.5, 3.0])
def test_batch_modes_and_sigmas(gpu, gauss_mode, edge_mode, sigma):
    from cucim_amd.batch import GrayGaussSobelBatch

    xs = [_rgb(s, 10 + i) for i, s in enumerate([(2, 3), (40, 250), (70, 33), (19, 500)])]
    plan = GrayGaussSobelBatch(xs, sigma=sigma, gauss_mode=gauss_mode, gauss_cval=0.3,
                               edge_mode=edge_mode)
    outs = plan.run()
    for x, o in zip(xs, outs):
        exp = _separate(x, sigma, gauss_mode, 0.3, edge_mode)
        np.testing.assert_array_equal(_bits(o), _bits(exp))


class _EmuNdi:
    """The oracle's ndimage with the reference's float32 arithmetic."""

    gaussian_filter = staticmethod(functools.partial(on.gaussian_filter, acc="emulate"))
    convolve1d = staticmethod(functools.partial(on.convolve1d, acc="emulate"))


@pytest.mark.parametrize("shape", [(20, 30), (45, 7), (3, 64)])
def test_batch_vs_oracle(gpu, shape):
    from cucim_amd.batch import GrayGaussSobelBatch

    x = _rgb(shape, 99)
    out = GrayGaussSobelBatch([x]).run()[0]
    g = osk.rgb2gray(x.cpu().numpy())
    g = osk.gaussian(g, sigma=1.5, mode="nearest", ndi=_EmuNdi)
    exp = osk.sobel(g, mode="reflect", ndi=_EmuNdi)
    assert exp.dtype == np.float32
    np.testing.assert_array_equal(_bits(out), exp.view(np.uint32))


def test_batch_graph_replay(gpu):
    """Replays of a captured batch, with new data copied into the bound
    inputs and eager work between replays, each equal to an eager run."""
    from cucim_amd.batch import GrayGaussSobelBatch

    xs = [_rgb(s, 50 + i) for i, s in enumerate([(64, 300), (128, 77), (500, 260)])]
    plan = GrayGaussSobelBatch(xs)
    g = plan.capture()
    for rep in range(3):
        for i, x in enumerate(xs):
            x.copy_(_rgb(tuple(x.shape[:2]), 200 + 10 * rep + i))
        g.replay()
        torch.cuda.synchronize()
        got = [_bits(o).copy() for o in plan.outputs]
        junk = torch.ones(1000, device="cuda").cumsum(0)  # eager work between replays
        del junk
        for x, b in zip(xs, got):
            np.testing.assert_array_equal(b, _bits(_separate(x)))


def test_batch_autotune(gpu):
    from cucim_amd.batch import GrayGaussSobelBatch

    xs = [_rgb((256, 384), 7), _rgb((384, 256), 8)]
    plan = GrayGaussSobelBatch(xs)
    res = plan.autotune(reps=2)
    assert res["rows"] in (16, 32) and res["ms"] > 0
    again = GrayGaussSobelBatch(xs)
    assert again.rows == res["rows"]
    outs = again.run()
    for x, o in zip(xs, outs):
        np.testing.assert_array_equal(_bits(o), _bits(_separate(x)))


def test_batch_errors(gpu):
    from cucim_amd import _lib
    from cucim_amd.batch import GrayGaussSobelBatch

    x = _rgb((10, 10), 1)
    with pytest.raises(NotImplementedError):
        GrayGaussSobelBatch([x], edge_mode="wrap").run()
    with pytest.raises(_lib.ResourceLimitError):
        GrayGaussSobelBatch([x], sigma=5.0)
    with pytest.raises(ValueError):
        GrayGaussSobelBatch([x.float()])
    with pytest.raises(ValueError):
        GrayGaussSobelBatch([])


def test_frangi_batch3d_graph_equals_eager(gpu):
    """Batched plain Frangi over uneven volumes: every volume of a graph
    replay equals the skimage-level frangi() call on it, bit for bit."""
    from cucim_amd import synthetic
    from cucim_amd.batch import FrangiBatch3D
    from cucim_amd.skimage.filters import frangi

    shapes = [(48, 60, 52), (70, 49, 81), (55, 55, 55), (64, 96, 50)]
    vols = [synthetic.tube_volume(s, ntubes=3, seed=i, noise=0.08) for i, s in enumerate(shapes)]
    b = FrangiBatch3D(vols, sigmas=(1, 2, 3))
    g = b.capture()
    for rep in range(2):
        for i, v in enumerate(b.volumes):
            v.copy_(synthetic.tube_volume(tuple(v.shape), ntubes=3, seed=100 + 10 * rep + i,
                                          noise=0.08))
        g.replay()
        torch.cuda.synchronize()
        got = [_bits(o).copy() for o in b.outputs]
        for v, gb in zip(b.volumes, got):
            exp = frangi(v, sigmas=(1, 2, 3), black_ridges=False)
            np.testing.assert_array_equal(gb, _bits(exp))
    lat = b.latencies_ms()
    assert len(lat) == len(shapes) and all(t > 0 for t in lat)


@pytest.mark.parametrize("streams", [1, 3, 8])
def test_frangi_batch3d_branches_equal_eager(gpu, streams):
    """The batch captured as parallel graph branches (volume i on stream
    i % streams): three replays, with eager Frangi calls and new inputs
    between them, each equal to the eager result bit for bit."""
    from cucim_amd import synthetic
    from cucim_amd.batch import FrangiBatch3D
    from cucim_amd.skimage.filters import frangi

    shapes = [(48, 60, 52), (70, 49, 81), (55, 55, 55), (64, 96, 50), (50, 48, 64)]
    vols = [synthetic.tube_volume(s, ntubes=3, seed=i, noise=0.08) for i, s in enumerate(shapes)]
    b = FrangiBatch3D(vols, sigmas=(1, 2), gamma=None)
    g = b.capture(streams=streams)
    for rep in range(3):
        for i, v in enumerate(b.volumes):
            v.copy_(synthetic.tube_volume(tuple(v.shape), ntubes=3, seed=200 + 10 * rep + i,
                                          noise=0.08))
        exp = [_bits(frangi(v, sigmas=(1, 2), black_ridges=False)) for v in b.volumes]
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        for o, e in zip(b.outputs, exp):
            np.testing.assert_array_equal(_bits(o), e)
