"""Whole-slide region reads (the reference's CuImage.read_region over cuslide).

CPU tests: the test writer and the oracle are pinned by libtiff and
libjpeg-turbo (through Pillow), the library's host decode path
(``device='cpu'``, host code only) equals the oracle bit for bit, metadata,
argument checks, and the reference's own read_region tests
(``python/cucim/tests/unit/clara/test_tiff_read_region.py``) restated on the
same stripe images (``tests/util/gen_tiff.py:87-124``).
GPU tests: the device path (``device='cuda'``: JPEG tiles decoded by device
kernels) equals the oracle bit for bit, and the region feeds the filters.
"""

import numpy as np
import pytest

from oracle.wsi import TiffFile, lzw_decode
from tests.wsi_tiff import lzw_encode, pyramid, write_tiff

PIL = pytest.importorskip("PIL")
from PIL import Image  # noqa: E402

CASES = {
    "raw": dict(compression="raw"),
    "deflate": dict(compression="deflate"),
    "lzw": dict(compression="lzw"),
    "deflate_pred": dict(compression="deflate", predictor=2),
    "lzw_pred": dict(compression="lzw", predictor=2),
    "jpeg420": dict(compression="jpeg", subsampling=2),
    "jpeg422": dict(compression="jpeg", subsampling=1),
    "jpeg444": dict(compression="jpeg", subsampling=0),
    "jpeg_rgb": dict(compression="jpeg", photometric="rgb", subsampling=0),
    "jpeg_notables": dict(compression="jpeg", tables=False, quality=75),
    "jpeg_restart": dict(compression="jpeg", restart=3),
    "jpeg_q100": dict(compression="jpeg", quality=100, subsampling=0),
    "jpeg_big_be": dict(compression="jpeg", big=True, byteorder=">"),
    "deflate_big": dict(compression="deflate", big=True),
    "lzw_be": dict(compression="lzw", byteorder=">"),
    "jpeg_strips": dict(compression="jpeg", strip_rows=40),
    "lzw_strips": dict(compression="lzw", strip_rows=37),
    "jpeg_gray": dict(compression="jpeg", gray=True),
    "deflate_gray": dict(compression="deflate", gray=True),
}

REGIONS = [((-13, -7), (50, 40)), ((200, 250), (90, 80)), ((33, 65), (1, 1)),
           ((-500, -500), (20, 10)), ((250, 290), (64, 48))]


def _make(tmp_path_factory, name, h=300, w=260, nlevels=3, tile=(64, 48), seed=1):
    kw = dict(CASES[name])
    gray = kw.pop("gray", False)
    lv = pyramid(h, w, nlevels, seed=seed, spp=1 if gray else 3)
    p = tmp_path_factory.mktemp("wsi") / f"{name}.tif"
    return write_tiff(p, lv, tile=tile, **kw), lv


@pytest.fixture(scope="module", params=sorted(CASES))
def case(request, tmp_path_factory):
    path, lv = _make(tmp_path_factory, request.param)
    return request.param, path, lv


def _cuimage(path):
    from cucim_amd.clara import CuImage

    return CuImage(path)


# ----------------------------------------------------------- oracle pins --
def test_lzw_roundtrip():
    rng = np.random.default_rng(0)
    for data in [b"", b"a", bytes(range(256)) * 40, rng.integers(0, 4, 50000).astype(np.uint8).tobytes(),
                 rng.integers(0, 256, 20000).astype(np.uint8).tobytes(), b"ab" * 30000]:
        assert lzw_decode(lzw_encode(data), len(data)) == data


@pytest.mark.parametrize("kw", [dict(compression="raw"), dict(compression="deflate"),
                                dict(compression="lzw"), dict(compression="lzw", predictor=2),
                                dict(compression="deflate", byteorder=">"),
                                dict(compression="deflate", big=True)])
def test_writer_pinned_by_libtiff(tmp_path, kw):
    """libtiff (through Pillow) reads the test writer's tiled lossless files
    back to the array written."""
    lv = pyramid(150, 130, 1, seed=3)
    p = write_tiff(tmp_path / "w.tif", lv, tile=(64, 48), **kw)
    assert np.array_equal(np.asarray(Image.open(p)), lv[0])


@pytest.mark.parametrize("comp", ["tiff_lzw", "tiff_adobe_deflate", "raw", "jpeg"])
def test_oracle_pinned_by_libtiff(tmp_path, comp):
    """Files libtiff wrote (strips; JPEG with JPEGTables, photometric RGB):
    the oracle and the library read what libtiff decodes; lossless files read
    back the array written."""
    lv = pyramid(150, 130, 1, seed=4)
    p = str(tmp_path / "l.tif")
    Image.fromarray(lv[0]).save(p, "TIFF", compression=comp,
                                **({"quality": 85} if comp == "jpeg" else {}))
    ref = np.asarray(Image.open(p))
    o = TiffFile(p).read_region((0, 0), (130, 150), 0)
    assert np.array_equal(o, ref)
    if comp != "jpeg":
        assert np.array_equal(ref, lv[0])
    assert np.array_equal(_cuimage(p).read_region((0, 0), (130, 150), 0), ref)


def test_oracle_jpeg_tiles_match_libjpeg(tmp_path):
    """Each JPEG tile of the oracle's level image is libjpeg-turbo's decode
    of that tile's stream; lossy error against the source stays small."""
    lv = pyramid(128, 128, 1, seed=5)
    p = write_tiff(tmp_path / "j.tif", lv, tile=(64, 64), compression="jpeg", tables=False,
                   quality=95)
    o = TiffFile(p)
    img = o.level_image(0)
    f = o.ifds[0]
    raw = o.data[f.offsets[3]:f.offsets[3] + f.counts[3]]
    import io

    dec = np.asarray(Image.open(io.BytesIO(raw)).convert("RGB"))
    assert np.array_equal(img[64:, 64:], dec)
    assert np.abs(img.astype(int) - lv[0]).mean() < 8  # 4:2:0 chroma of a noisy image


# ------------------------------------------------------ host path = oracle --
def test_host_region_matches_oracle(case):
    name, path, lv = case
    o, im = TiffFile(path), _cuimage(path)
    res = im.resolutions
    assert res["level_count"] == len(lv)
    for level in range(res["level_count"]):
        dims = res["level_dimensions"][level]
        assert dims == (lv[level].shape[1], lv[level].shape[0])
        for loc, size in [((0, 0), dims)] + REGIONS:
            for nw in (0, 3):
                got = im.read_region(loc, size, level, num_workers=nw)
                exp = o.read_region(loc, size, level)
                assert got.shape == exp.shape
                assert np.array_equal(got, exp), (name, level, loc, size, nw)
        if CASES[name]["compression"] != "jpeg":
            assert np.array_equal(im.read_region((0, 0), dims, level), lv[level])


def test_host_buf_and_batches(tmp_path_factory):
    path, lv = _make(tmp_path_factory, "jpeg420")
    im, o = _cuimage(path), TiffFile(path)
    buf = np.empty((40, 50, 3), np.uint8)
    r = im.read_region((10, 20), (50, 40), 0, buf=buf)
    assert r is buf and np.array_equal(buf, o.read_region((10, 20), (50, 40), 0))
    locs = [(0, 0), (40, 30), (-5, 100), (200, 200), (7, 9)]
    items = list(im.read_region(locs, (32, 24), 0, num_workers=2))
    assert len(items) == 5
    for loc, a in zip(locs, items):
        assert np.array_equal(a, o.read_region(loc, (32, 24), 0))
    batches = list(im.read_region(locs, (32, 24), 1, batch_size=2))
    assert [b.shape for b in batches] == [(2, 24, 32, 3), (2, 24, 32, 3), (1, 24, 32, 3)]
    assert np.array_equal(batches[1][1], o.read_region(locs[3], (32, 24), 1))


# ------------------------------------------------------------- metadata --
def test_metadata_levels_and_downsamples(tmp_path):
    lv = pyramid(301, 257, 3, seed=6)  # odd sizes: downsamples are not powers of 2
    p = write_tiff(tmp_path / "m.tif", lv, tile=(64, 64), compression="deflate",
                   resolution=(40000, 20000, 3), description=b"generic slide")
    im = _cuimage(p)
    r = im.resolutions
    assert r["level_dimensions"] == ((257, 301), (128, 150), (64, 75))
    ds = [(np.float32(257) / np.float32(w) + np.float32(301) / np.float32(h)) / np.float32(2)
          for w, h in r["level_dimensions"]]
    assert r["level_downsamples"] == tuple(float(np.float32(d)) for d in ds)
    assert r["level_tile_sizes"] == ((64, 64),) * 3
    assert im.shape == [301, 257, 3] and im.size("XY") == [257, 301] and im.dims == "YXC"
    assert im.dtype == np.uint8 and im.channel_names == ["R", "G", "B"]
    assert im.spacing("XY") == [1 / 40000, 1 / 20000]
    assert im.spacing_units() == ["centimeter", "centimeter", "color"]
    assert im.metadata["tiff"]["image_description"] == "generic slide"
    assert im.background == 0
    # level-0 locations divide by the float32 downsample and truncate
    o = TiffFile(p)
    for loc in [(0, 0), (101, 203), (255, 17), (-37, -3), (13, 299)]:
        for level in (1, 2):
            assert np.array_equal(im.read_region(loc, (30, 20), level),
                                  o.read_region(loc, (30, 20), level))


def test_aperio_background_and_associated(tmp_path):
    """Aperio files: background 255 and untiled IFDs after the first are
    associated images, not levels (tiff.cpp:670-707)."""
    lv = pyramid(200, 180, 2, seed=7)
    label = np.full((40, 60, 3), 77, np.uint8)
    p = write_tiff(tmp_path / "a.svs", lv, tile=(64, 64), compression="jpeg",
                   photometric="rgb", subsampling=0,
                   description=b"Aperio Image Library v10.0.50\r\n180x200 [0,0 180x200] (64x64)",
                   associated=[(label, 1), (label, 9)])
    im = _cuimage(p)
    assert im.resolutions["level_count"] == 2
    assert im.background == 255 and im.metadata["tiff"]["vendor"] == "aperio"
    r = im.read_region((-10, -10), (30, 30), 0)
    assert (r[:10] == 255).all() and (r[:, :10] == 255).all()
    assert np.array_equal(r, TiffFile(p).read_region((-10, -10), (30, 30), 0))
    # a generic file keeps every IFD as a level, sorted largest first
    p2 = write_tiff(tmp_path / "g.tif", [lv[1], lv[0]], tile=(64, 64), compression="deflate")
    assert _cuimage(p2).resolutions["level_dimensions"] == ((180, 200), (90, 100))


@pytest.mark.parametrize("comp", ["lzw", "jpeg", "deflate"])
def test_associated_images(tmp_path, comp):
    """Aperio associated images: the untiled IFD 1 of subfile type 0 is the
    thumbnail, subfile 1 the label, subfile 9 the macro; each read whole
    equals the oracle's decode (the reference reads JPEG and LZW strips,
    tiff.cpp:790-900); generic files have none."""
    lv = pyramid(200, 180, 1, seed=8)
    rng = np.random.default_rng(9)
    thumb = pyramid(50, 45, 1, seed=10)[0]
    label = rng.integers(0, 256, (40, 60, 3)).astype(np.uint8)
    macro = pyramid(37, 70, 1, seed=11)[0]
    p = write_tiff(tmp_path / "a.svs", lv, tile=(64, 64), compression=comp,
                   **({"photometric": "rgb", "subsampling": 0} if comp == "jpeg" else {}),
                   description=b"Aperio Image Library v10.0.50",
                   associated=[(thumb, 0), (label, 1), (macro, 9)])
    im, o = _cuimage(p), TiffFile(p)
    assert im.associated_images == {"thumbnail", "label", "macro"} == set(o.associated)
    assert im.resolutions["level_count"] == 1
    for name, src in [("thumbnail", thumb), ("label", label), ("macro", macro)]:
        a = im.associated_image(name, num_workers=2)
        e = o.ifd_image(o.associated[name])
        assert a.shape == src.shape and np.array_equal(a, e), name
        if comp != "jpeg":
            assert np.array_equal(a, src)
    with pytest.raises(ValueError, match="no associated image"):
        im.associated_image("barcode")
    g = write_tiff(tmp_path / "g.tif", lv, tile=(64, 64), compression=comp,
                   associated=[(label, 1)])
    assert _cuimage(g).associated_images == set()


def test_argument_errors(tmp_path):
    lv = pyramid(64, 48, 1)
    p = write_tiff(tmp_path / "e.tif", lv, tile=(16, 16), compression="raw")
    im = _cuimage(p)
    with pytest.raises(ValueError, match="'level' should be less than "):
        im.read_region(level=-1)
    with pytest.raises(ValueError, match="'level' should be less than "):
        im.read_region(level=7)
    with pytest.raises(ValueError, match="Should be > 0"):
        im.read_region((0, 0), (0, 5))
    with pytest.raises(ValueError, match="exceeds the original image width"):
        im.read_region((0, 0), (49, 5))
    with pytest.raises(ValueError, match="exceeds the original image height"):
        im.read_region((0, 0), (5, 65))
    with pytest.raises(ValueError, match="buf"):
        im.read_region((0, 0), (5, 5), buf=np.empty((5, 6, 3), np.uint8))
    bad = tmp_path / "x.tif"
    bad.write_bytes(b"not a tiff at all")
    with pytest.raises(ValueError, match="not a TIFF"):
        _cuimage(bad)
    with pytest.raises(ValueError, match="cannot open"):
        _cuimage(tmp_path / "missing.tif")
    # PackBits (32773) is outside the supported codecs
    Image.fromarray(lv[0]).save(tmp_path / "pb.tif", "TIFF", compression="packbits")
    with pytest.raises(NotImplementedError, match="compression"):
        _cuimage(tmp_path / "pb.tif").read_region((0, 0), (8, 8))


# ------------------------- the reference's read_region tests, restated --
def _stripe_levels(width, height):
    """gen_tiff.py:87-124: rows in pairs of red / green / blue at 255 - level;
    levels halve while the width is at least 256."""
    sizes = [(width, height)]
    if 256 <= width <= 4096:
        while sizes[-1][0] >= 256:
            sizes.append((sizes[-1][0] // 2, sizes[-1][1] // 2))
    out = []
    for level, (w, h) in enumerate(sizes):
        a = np.zeros((h, w, 3), np.uint8)
        for r in range(6):
            a[r::6, :, r // 2] = 255 - level
        out.append(a)
    return out


@pytest.fixture(scope="module", params=["jpeg", "deflate", "raw"])
def stripe_32x24(request, tmp_path_factory):
    p = tmp_path_factory.mktemp("stripe") / f"s_{request.param}.tif"
    kw = dict(quality=95) if request.param == "jpeg" else {}
    return write_tiff(p, _stripe_levels(32, 24), tile=(16, 16), compression=request.param, **kw)


@pytest.fixture(scope="module")
def stripe_4096(tmp_path_factory):
    p = tmp_path_factory.mktemp("stripe") / "s4096.tif"
    return write_tiff(p, _stripe_levels(4096, 4096), tile=(256, 256), compression="jpeg",
                      quality=95)


def test_tiff_stripe_inner(stripe_32x24):
    im = _cuimage(stripe_32x24)
    width, height = im.size("XY")
    tw, th = im.resolutions["level_tile_sizes"][0]
    for start, size in [((0, 0), (width, height)), ((0, 0), (tw // 2, th // 2)),
                        ((tw // 2, th // 2), (tw, th)),
                        ((width - tw // 2, height - th // 2), (tw // 2, th // 2))]:
        a = im.read_region(start, size)
        assert np.all(np.count_nonzero(a, axis=2) > 0)


def test_tiff_stripe_boundary(stripe_32x24):
    im = _cuimage(stripe_32x24)
    width, height = im.size("XY")
    tw, th = im.resolutions["level_tile_sizes"][0]
    for start in [(-(tw // 2), -(th // 2)), (width - tw // 2, -(th // 2)),
                  (-(tw // 2), height - th // 2), (width - tw // 2, height - th // 2)]:
        a = im.read_region(start, (tw, th))
        zero = np.count_nonzero(np.count_nonzero(a, axis=2) == 0)
        assert zero - (tw * th * 0.75) < 5


def test_tiff_stripe_outside(stripe_32x24):
    im = _cuimage(stripe_32x24)
    width, height = im.size("XY")
    tw, th = im.resolutions["level_tile_sizes"][0]
    for start in [(-width - tw, -height - th), (width + th, height + th)]:
        a = im.read_region(start, (tw, th))
        assert np.count_nonzero(np.count_nonzero(a, axis=2) == 0) == tw * th


def test_tiff_stripe_multiresolution(stripe_4096):
    im = _cuimage(stripe_4096)
    assert im.resolutions["level_count"] == 6
    for level in range(6):
        a = im.read_region((0, 0), (256, 256), level)
        zero = np.count_nonzero(np.count_nonzero(a, axis=2) == 0)
        w, h = im.resolutions["level_dimensions"][level]
        assert zero == 256 * 256 - min(w, 256) * min(h, 256)


def test_tiff_iterator(stripe_4096):
    im = _cuimage(stripe_4096)
    x = im.read_region([(0, 0), (0, 0)], (256, 256), 0, num_workers=1)
    x = im.read_region([(0, 0), (0, 0)], (256, 256), 0, num_workers=1)
    first = next(x)
    x = im.read_region([(0, 0), (0, 0), (0, 0)], (256, 256), 0, num_workers=1)
    assert all(np.array_equal(first, b) for b in x)


# ------------------------------------------------------------------- GPU --
@pytest.mark.gpu
def test_device_region_matches_oracle(gpu, case):
    import torch

    name, path, lv = case
    o, im = TiffFile(path), _cuimage(path)
    for level in range(im.resolutions["level_count"]):
        dims = im.resolutions["level_dimensions"][level]
        for loc, size in [((0, 0), dims)] + REGIONS:
            got = im.read_region(loc, size, level, device="cuda", num_workers=2)
            assert got.is_cuda and got.dtype == torch.uint8
            exp = o.read_region(loc, size, level)
            assert np.array_equal(got.cpu().numpy(), exp), (name, level, loc, size)


@pytest.mark.gpu
@pytest.mark.parametrize("restart,sub", [(1, 2), (3, 0), (5, 1), (64, 2)])
def test_device_restart_intervals(gpu, tmp_path, restart, sub):
    """JPEG tiles with restart markers (DRI every `restart` MCUs): decoded on
    the device interval by interval (each starts byte-aligned, DC predictors
    at zero), equal to the oracle's libjpeg-turbo decode."""
    lv = pyramid(512, 384, 1, seed=14)
    p = write_tiff(tmp_path / "r.tif", lv, tile=(128, 128), compression="jpeg",
                   restart=restart, subsampling=sub, quality=92)
    o, im = TiffFile(p), _cuimage(p)
    got = im.read_region((0, 0), (384, 512), 0, device="cuda", num_workers=2)
    assert np.array_equal(got.cpu().numpy(), o.read_region((0, 0), (384, 512), 0))
    got = im.read_region((50, 70), (200, 300), 0, device="cuda")
    assert np.array_equal(got.cpu().numpy(), o.read_region((50, 70), (200, 300), 0))


@pytest.mark.gpu
def test_device_associated_image(gpu, tmp_path):
    import torch

    lv = pyramid(128, 128, 1, seed=12)
    label = pyramid(40, 56, 1, seed=13)[0]
    p = write_tiff(tmp_path / "a.svs", lv, tile=(64, 64), compression="jpeg",
                   description=b"Aperio Image Library", associated=[(label, 1)])
    im = _cuimage(p)
    d = im.associated_image("label", device="cuda")
    assert d.is_cuda and d.dtype == torch.uint8
    assert np.array_equal(d.cpu().numpy(), im.associated_image("label"))


@pytest.mark.gpu
def test_device_region_4096_pyramid(gpu, stripe_4096):
    """Every level of the 4096^2 stripe pyramid (256 x 256 JPEG tiles; 256
    tiles decoded in one launch at level 0) equals the oracle."""
    o, im = TiffFile(stripe_4096), _cuimage(stripe_4096)
    for level in range(6):
        dims = im.resolutions["level_dimensions"][level]
        got = im.read_region((0, 0), dims, level, device="cuda")
        assert np.array_equal(got.cpu().numpy(), o.read_region((0, 0), dims, level)), level
    got = im.read_region((1000, 700), (1500, 1300), 0, device="cuda")
    assert np.array_equal(got.cpu().numpy(), o.read_region((1000, 700), (1500, 1300), 0))


@pytest.mark.gpu
def test_device_region_feeds_filters(gpu, tmp_path_factory):
    """Region in device memory -> DLPack -> rgb2gray -> gaussian: the same
    result as the filters on the oracle's region uploaded, bit for bit."""
    import torch

    from cucim_amd.skimage.color import rgb2gray
    from cucim_amd.skimage.filters import gaussian

    path, lv = _make(tmp_path_factory, "jpeg420", h=512, w=448, nlevels=2, tile=(128, 128))
    im = _cuimage(path)
    reg = im.read_region((37, 51), (300, 260), 0, device="cuda")
    view = torch.from_dlpack(reg)
    assert view.data_ptr() == reg.data_ptr()
    ref = torch.from_numpy(TiffFile(path).read_region((37, 51), (300, 260), 0)).to(gpu)
    a = gaussian(rgb2gray(view), sigma=1.5)
    b = gaussian(rgb2gray(ref), sigma=1.5)
    assert torch.equal(a, b)
    buf = torch.empty((260, 300, 3), dtype=torch.uint8, device=gpu)
    assert im.read_region((37, 51), (300, 260), 0, device="cuda", buf=buf) is buf
    assert torch.equal(buf, reg)
