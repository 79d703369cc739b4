"""Whole-slide image regions: the reference's ``cucim.CuImage`` reader for
tiled TIFF / BigTIFF pyramids (Aperio SVS, Philips TIFF, generic pyramidal
TIFF), feeding the filters.

``CuImage(path).read_region(location, size, level, device=...)`` returns the
region as ``(height, width, channels)`` uint8: a numpy array for
``device='cpu'`` (tiles decoded on host threads), a torch tensor in device
memory for ``device='cuda'`` (JPEG tiles entropy-decoded, inverse-transformed
and colour-converted by device kernels).  Both expose DLPack
(``__dlpack__``), so ``torch.from_dlpack`` / ``cupy.from_dlpack`` take them
without a copy.

Reference: ``python/pybind11/cucim_py.cpp:146-158`` (read_region's
arguments), ``cpp/src/cuimage.cpp:649`` (CuImage::read_region),
``cpp/plugins/cucim.kit.cuslide/src/cuslide/cuslide.cpp:190-282`` (metadata),
``.../tiff/tiff.cpp:720-770`` (argument checks, level-0 location).
"""

import ctypes
import os

import numpy as np

from .. import _lib

__all__ = ["CuImage"]

_COMPRESSION = {1: "raw", 5: "lzw", 7: "jpeg", 8: "deflate", 32946: "deflate"}
_UNITS = {1: "", 2: "inch", 3: "centimeter"}


class CuImage:
    """An open whole-slide image (the reference's ``cucim.CuImage``).

    Levels are the file's IFDs sorted by size, largest first (for Aperio and
    Philips files the untiled IFDs after the first are associated images, not
    levels).  ``read_region`` takes its location in level-0 pixels and its
    size in pixels of the requested level, as the reference does.
    """

    def __init__(self, path):
        self._h = None
        lib = _lib.load_library()
        self.path = os.fspath(path)
        h = ctypes.c_void_p()
        _lib.check(lib.hc_wsi_open(self.path.encode(), ctypes.byref(h)), "CuImage")
        self._h = h
        n = lib.hc_wsi_level_count(h)
        info = (ctypes.c_int64 * 8)()
        self._levels = []
        for lv in range(n):
            _lib.check(lib.hc_wsi_level_info(h, lv, info), "CuImage")
            self._levels.append(tuple(int(v) for v in info))
        res = (ctypes.c_double * 3)()
        bg = ctypes.c_int()
        desc = ctypes.create_string_buffer(1 << 16)
        vendor = ctypes.create_string_buffer(32)
        _lib.check(lib.hc_wsi_properties(h, res, ctypes.byref(bg), desc, len(desc), vendor,
                                         len(vendor)), "CuImage")
        self._res = (res[0], res[1], int(res[2]))
        self._background = bg.value
        self._description = desc.value.decode(errors="replace")
        self._vendor = vendor.value.decode()
        # level_downsamples: mean of the width and height ratios in float32
        # (cuslide.cpp:227-234)
        self._assoc = {}
        name = ctypes.create_string_buffer(64)
        for k in range(lib.hc_wsi_associated_count(h)):
            _lib.check(lib.hc_wsi_associated_info(h, k, name, len(name), info), "CuImage")
            self._assoc[name.value.decode()] = (k, int(info[0]), int(info[1]), int(info[2]))
        w0, h0 = np.float32(self._levels[0][0]), np.float32(self._levels[0][1])
        self._downsamples = [
            float((w0 / np.float32(L[0]) + h0 / np.float32(L[1])) / np.float32(2))
            for L in self._levels]

    # ------------------------------------------------------------ metadata --
    def close(self):
        if self._h is not None:
            _lib.load_library().hc_wsi_close(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def is_loaded(self):
        return self._h is not None

    @property
    def ndim(self):
        return 3

    @property
    def dims(self):
        return "YXC"

    @property
    def dtype(self):
        return np.dtype(np.uint8)

    @property
    def shape(self):
        w, h, _, _, c = self._levels[0][:5]
        return [h, w, c]

    def size(self, dim_order=""):
        d = dict(zip("YXC", self.shape))
        return [d[c] for c in (dim_order or self.dims)]

    @property
    def channel_names(self):
        return ["R", "G", "B", "A"][:self._levels[0][4]] if self._levels[0][4] > 1 else ["I"]

    @property
    def resolutions(self):
        return {
            "level_count": len(self._levels),
            "level_dimensions": tuple((L[0], L[1]) for L in self._levels),
            "level_downsamples": tuple(self._downsamples),
            "level_tile_sizes": tuple((L[2], L[3]) for L in self._levels),
        }

    def spacing(self, dim_order=""):
        """Pixel spacing per dimension from the TIFF resolution tags (1.0
        where absent)."""
        xr, yr, _ = self._res
        d = {"Y": 1.0 / yr if yr else 1.0, "X": 1.0 / xr if xr else 1.0, "C": 1.0}
        return [d[c] for c in (dim_order or self.dims)]

    def spacing_units(self, dim_order=""):
        u = _UNITS.get(self._res[2], "")
        d = {"Y": u, "X": u, "C": "color"}
        return [d[c] for c in (dim_order or self.dims)]

    @property
    def metadata(self):
        return {
            "cucim": {
                "path": self.path,
                "ndim": 3,
                "dims": self.dims,
                "shape": self.shape,
                "dtype": {"code": 1, "bits": 8, "lanes": 1},
                "channel_names": self.channel_names,
                "spacing": self.spacing(),
                "spacing_units": self.spacing_units(),
                "resolutions": {k: list(v) if isinstance(v, tuple) else v
                                for k, v in self.resolutions.items()},
                "associated_images": sorted(self._assoc),
            },
            "tiff": {
                "vendor": self._vendor,
                "image_description": self._description,
                "resolution_unit": _UNITS.get(self._res[2], ""),
                "x_resolution": self._res[0],
                "y_resolution": self._res[1],
                "compression": [_COMPRESSION.get(L[5], str(L[5])) for L in self._levels],
            },
        }

    @property
    def associated_images(self):
        """Names of the associated images (Aperio: ``thumbnail``, ``label``,
        ``macro``); an empty set for generic TIFF files."""
        return set(self._assoc)

    def associated_image(self, name, device="cpu", num_workers=0):
        """The whole associated image ``name`` as ``(height, width,
        channels)`` uint8: a numpy array, or a torch tensor in device memory
        for ``device='cuda'`` (decoded on the host, then uploaded)."""
        if name not in self._assoc:
            raise ValueError(f"no associated image named {name!r}; the file has "
                             f"{sorted(self._assoc)}")
        k, w, h, c = self._assoc[name]
        out = np.empty((h, w, c), np.uint8)
        _lib.check(_lib.load_library().hc_wsi_associated_read(
            self._h, k, out.ctypes.data_as(ctypes.c_void_p), int(num_workers)),
            "associated_image")
        if _is_cuda(device):
            import torch

            from .. import _array

            return torch.from_numpy(out).to(_array.device())
        return out

    @property
    def background(self):
        """Value of region pixels outside the level (255 for Aperio /
        Philips, else 0)."""
        return self._background

    # ------------------------------------------------------------- regions --
    def _check(self, size, level):
        if not 0 <= level < len(self._levels):
            # cuimage.cpp:104-131
            raise ValueError(f"'level' should be less than {len(self._levels)}")
        w0, h0 = self._levels[0][:2]
        for v in size:
            if v <= 0:
                raise ValueError(f"Invalid size ({v}) in the request! (Should be > 0)")
        if size[0] > w0:
            raise ValueError(f"Invalid size (it exceeds the original image width {w0})")
        if size[1] > h0:
            raise ValueError(f"Invalid size (it exceeds the original image height {h0})")

    def _level_xy(self, location, level):
        # int64 /= float in the reference: float32 division, truncation
        ds = np.float32(self._downsamples[level])
        return [int(np.float32(np.float32(v) / ds)) for v in location]

    def read_region(self, location=(), size=(), level=0, num_workers=0, batch_size=1,
                    device="cpu", buf=None):
        """The region of ``level`` at ``location`` (level-0 pixels, x, y;
        may be negative or reach past the edge: such pixels take the
        background value) of ``size`` (level pixels, width, height).

        ``location`` may also be a list of locations: the result is then an
        iterator over the regions, ``batch_size`` of them stacked per item
        (the reference's batch loader).  ``num_workers`` host threads decode
        lossless tiles (and JPEG tiles for ``device='cpu'``).  ``buf``: an
        existing output array / tensor of the right shape and dtype.
        """
        if not location:
            location = (0, 0)
        lv = self._levels[level] if 0 <= level < len(self._levels) else None
        if not size:
            size = (lv[0], lv[1]) if lv else (1, 1)
        size = tuple(int(v) for v in size)
        self._check(size, level)
        locs = np.asarray(location, dtype=np.int64)
        if locs.ndim == 2:
            return self._iter_regions(locs, size, level, num_workers, batch_size, device)
        return self._read_one(tuple(int(v) for v in locs), size, level, num_workers, device, buf)

    def _iter_regions(self, locs, size, level, num_workers, batch_size, device):
        for i in range(0, len(locs), batch_size):
            regs = [self._read_one(tuple(int(v) for v in p), size, level, num_workers, device,
                                   None) for p in locs[i:i + batch_size]]
            if batch_size == 1:
                yield regs[0]
            elif _is_cuda(device):
                import torch

                yield torch.stack(regs)
            else:
                yield np.stack(regs)

    def _read_one(self, location, size, level, num_workers, device, buf):
        x, y = self._level_xy(location, level)
        w, h = size
        nch = self._levels[level][4]
        lib = _lib.load_library()
        if _is_cuda(device):
            import torch

            from .. import _array

            out = buf
            if out is None:
                out = torch.empty((h, w, nch), dtype=torch.uint8, device=_array.device())
            if (tuple(out.shape) != (h, w, nch) or out.dtype != torch.uint8
                    or not out.is_cuda or not out.is_contiguous()):
                raise ValueError("buf must be a contiguous uint8 device tensor of shape "
                                 f"{(h, w, nch)}")
            _lib.lib()
            _lib.check(lib.hc_wsi_read_region_device(
                self._h, level, x, y, w, h, ctypes.c_void_p(out.data_ptr()), int(num_workers),
                ctypes.c_void_p(_array.stream())), "read_region")
            return out
        if device not in ("cpu", None):
            raise ValueError(f"unsupported device {device!r} (use 'cpu' or 'cuda')")
        out = buf if buf is not None else np.empty((h, w, nch), np.uint8)
        if out.shape != (h, w, nch) or out.dtype != np.uint8 or not out.flags.c_contiguous:
            raise ValueError(f"buf must be a C-contiguous uint8 array of shape {(h, w, nch)}")
        _lib.check(lib.hc_wsi_read_region(self._h, level, x, y, w, h,
                                          out.ctypes.data_as(ctypes.c_void_p), int(num_workers)),
                   "read_region")
        return out


def _is_cuda(device):
    return isinstance(device, str) and device.startswith("cuda") or (
        getattr(device, "type", None) == "cuda")
