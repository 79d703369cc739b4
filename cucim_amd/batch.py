"""Batched pipelines over many images of uneven sizes in one launch.

The reference filters one image per call (every skimage function takes one
array); a caller with many tiles loops in Python, one set of kernel launches
per tile.  Here a plan binds a list of input and output images, builds a
work-item table once (``hc_batch2d_plan``), and runs the whole batch as one
persistent launch that can be captured into a HIP graph and replayed.  Tile
shapes (output rows per work item) are chosen by measuring the candidates
(``autotune``); choices are remembered per workload signature, in memory and
in a JSON file that later processes read.

GrayGaussSobelBatch: per image, ``filters.sobel(filters.gaussian(
color.rgb2gray(rgb), sigma, mode=gauss_mode, cval=gauss_cval,
truncate=truncate), mode=edge_mode)`` (reference:
color/colorconv.py:1037-1095, _shared/filters.py:21-166,
filters/edges.py:148-241), each output equal to the three separate calls.
"""

import ctypes
import hashlib
import json
import math
import os

import numpy as np
import torch

from . import _array, _lib, _trace
from .skimage._vendored import _ndimage_filters as _F
from .skimage._vendored._ndimage_util import _mode_code


class HCTile2d(ctypes.Structure):
    _fields_ = [
        ("in_", ctypes.c_void_p),
        ("out", ctypes.c_void_p),
        ("H", ctypes.c_int64),
        ("W", ctypes.c_int64),
        ("item0", ctypes.c_int64),
    ]


# workload signature -> chosen item rows; persisted across processes in a
# JSON file (CUCIM_AMD_TUNE_CACHE, default ~/.cache/cucim_amd/autotune.json;
# "off" disables the file)
_TUNED = {}
_TUNED_LOADED = False


def _tune_path():
    p = os.environ.get("CUCIM_AMD_TUNE_CACHE")
    if p is not None:
        return None if p.strip().lower() in ("", "0", "off", "none") else p
    return os.path.join(os.path.expanduser("~"), ".cache", "cucim_amd", "autotune.json")


def _tune_read(path):
    try:
        with open(path) as f:
            data = json.load(f)
    except (OSError, ValueError):
        return {}
    ent = data.get("entries") if isinstance(data, dict) else None
    if not isinstance(ent, dict):
        return {}
    return {k: v for k, v in ent.items() if isinstance(k, str) and isinstance(v, int)}


def tune_load():
    """Merge the persisted table into the in-memory one (entries chosen in
    this process win)."""
    global _TUNED_LOADED
    _TUNED_LOADED = True
    path = _tune_path()
    if path is None:
        return
    for k, v in _tune_read(path).items():
        _TUNED.setdefault(k, v)


def tune_save():
    """Write the table, merged with what other processes saved meanwhile
    (atomic replace)."""
    path = _tune_path()
    if path is None:
        return
    merged = _tune_read(path)
    merged.update(_TUNED)
    try:
        os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
        tmp = f"{path}.{os.getpid()}.tmp"
        with open(tmp, "w") as f:
            json.dump({"version": 1, "entries": merged}, f, indent=1, sort_keys=True)
        os.replace(tmp, path)
    except OSError:
        pass


def _tuned(sig, default):
    if not _TUNED_LOADED:
        tune_load()
    return _TUNED.get(sig, default)


def _arch():
    try:
        return torch.cuda.get_device_properties(torch.cuda.current_device()).gcnArchName
    except (RuntimeError, AssertionError, AttributeError):
        return "unknown"


def _time_ms(fn, reps):
    s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    fn()
    torch.cuda.synchronize()
    s.record()
    for _ in range(reps):
        fn()
    e.record()
    torch.cuda.synchronize()
    return s.elapsed_time(e) / reps


class GrayGaussSobelBatch:
    """Plan of rgb2gray -> gaussian -> sobel over bound images.

    inputs: list of (H, W, 3) uint8 CUDA tensors (C-contiguous);
    outputs: list of (H, W) float32 CUDA tensors, or None to allocate.
    """

    ROWS = (16, 32)

    def __init__(self, inputs, outputs=None, *, sigma=1.5, truncate=4.0, gauss_mode="nearest",
                 gauss_cval=0.0, edge_mode="reflect", rows=None):
        if len(inputs) == 0:
            raise ValueError("no images")
        self.inputs = []
        for x in inputs:
            x = _array.as_tensor(x)
            if x.dtype != torch.uint8 or x.dim() != 3 or x.shape[2] != 3:
                raise ValueError("inputs must be (H, W, 3) uint8 images")
            if not x.is_contiguous():
                raise ValueError("inputs must be C-contiguous")
            self.inputs.append(x)
        if outputs is None:
            outputs = [_array.empty(tuple(x.shape[:2]), torch.float32) for x in self.inputs]
        self.outputs = []
        for x, o in zip(self.inputs, outputs):
            o = _array.as_tensor(o)
            if o.dtype != torch.float32 or tuple(o.shape) != tuple(x.shape[:2]) \
                    or not o.is_contiguous():
                raise ValueError("outputs must be C-contiguous float32 of the inputs' (H, W)")
            self.outputs.append(o)
        if len(self.outputs) != len(self.inputs):
            raise ValueError("inputs and outputs differ in length")
        if sigma < 0:
            raise ValueError("Sigma values less than zero are not valid")
        radius = int(float(truncate) * float(sigma) + 0.5)
        w = _F._gaussian_kernel1d(float(sigma), 0, radius).astype(np.float32)
        self.taps = _F._Taps(w, radius, np.float32)
        if self.taps.w.size > 33:
            raise _lib.ResourceLimitError("gaussian longer than 33 taps")
        self.gauss_mode = _mode_code(gauss_mode)
        self.edge_mode = _mode_code(edge_mode)
        self.gauss_cval = float(gauss_cval)
        # sobel magnitude weights as the edge filter forms them (float32), as
        # correlation weights (the convolution weights reversed)
        edge = (np.array([1, 0, -1], dtype=np.float32) / math.sqrt(2)).astype(np.float32)
        smooth = (np.array([1, 2, 1]) / 4).astype(np.float32)
        self.edge3 = np.ascontiguousarray(edge[::-1], dtype=np.float64)
        self.smooth3 = np.ascontiguousarray(smooth[::-1], dtype=np.float64)
        self.pixels = sum(int(x.shape[0]) * int(x.shape[1]) for x in self.inputs)
        self._tables = {}
        self.rows = rows if rows is not None else _tuned(self.signature(), 16)
        if self.rows not in self.ROWS:
            raise ValueError("rows must be 16 or 32")

    def signature(self):
        """Workload key of the tuning table: pipeline, kernel length, modes,
        device architecture and a digest of the sorted image shapes."""
        shapes = sorted((int(x.shape[0]), int(x.shape[1])) for x in self.inputs)
        digest = hashlib.sha1(repr(shapes).encode()).hexdigest()[:16]
        return (f"gray_gauss_sobel|K={self.taps.w.size}|modes={self.gauss_mode},"
                f"{self.edge_mode}|{_arch()}|n={len(shapes)}|{digest}")

    def _table(self, rows):
        if rows in self._tables:
            return self._tables[rows]
        lib = _lib.lib()
        n = len(self.inputs)
        tiles = (HCTile2d * n)()
        for i, (x, o) in enumerate(zip(self.inputs, self.outputs)):
            tiles[i].in_ = x.data_ptr()
            tiles[i].out = o.data_ptr()
            tiles[i].H = int(x.shape[0])
            tiles[i].W = int(x.shape[1])
        K = self.taps.w.size
        total = lib.hc_batch2d_plan(tiles, n, K, rows, None, 0)
        if total < 0:
            _lib.check(_lib.ERR_VALUE, "batch plan")
        item_tile = np.empty(max(total, 1), dtype=np.int32)
        got = lib.hc_batch2d_plan(tiles, n, K, rows,
                                  item_tile.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), total)
        if got != total:
            _lib.check(_lib.ERR_VALUE, "batch plan")
        raw = np.frombuffer(bytes(tiles), dtype=np.uint8)
        dev = _array.device()
        tab = (torch.from_numpy(raw.copy()).to(dev), torch.from_numpy(item_tile).to(dev), total)
        self._tables[rows] = tab
        return tab

    @_trace.annotate("batch2d.run")
    def run(self):
        """Launch the whole batch on the current stream; returns the outputs."""
        tiles_dev, items_dev, total = self._table(self.rows)
        lib = _lib.lib()
        P = ctypes.POINTER(ctypes.c_double)
        rc = lib.hc_batch2d_gray_gauss_sobel(
            tiles_dev.data_ptr(), len(self.inputs), items_dev.data_ptr(), total,
            self.taps.struct, self.gauss_mode, self.gauss_cval,
            self.edge3.ctypes.data_as(P), self.smooth3.ctypes.data_as(P), self.edge_mode,
            self.rows, _array.stream())
        _lib.check(rc, "batched gray/gaussian/sobel")
        return self.outputs

    def autotune(self, reps=3):
        """Time every candidate tile shape on the bound images and keep the
        fastest (remembered for plans with the same signature)."""
        best = None
        for rows in self.ROWS:
            self.rows = rows
            ms = _time_ms(self.run, reps)
            if best is None or ms < best[1]:
                best = (rows, ms)
        self.rows = best[0]
        _TUNED[self.signature()] = self.rows
        tune_save()
        return {"rows": self.rows, "ms": best[1]}

    def capture(self):
        """A torch CUDA (HIP) graph replaying run(); inputs are read from the
        bound tensors at each replay."""
        self._table(self.rows)
        s = torch.cuda.Stream()
        s.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(s):
            self.run()  # warm-up outside capture
        torch.cuda.current_stream().wait_stream(s)
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            self.run()
        return g


def _capture_branches(n, run_one, streams):
    """Capture run_one(0..n-1) into one graph, item i on side stream
    i % streams (default priority), after a warm-up of the same schedule
    outside the graph."""
    k = max(1, min(int(streams), n))
    cur = torch.cuda.current_stream()
    side = [torch.cuda.Stream() for _ in range(k)]
    for s in side:
        s.wait_stream(cur)
    for i in range(n):
        with torch.cuda.stream(side[i % k]):
            run_one(i)
    for s in side:
        cur.wait_stream(s)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    cap = torch.cuda.Stream()
    cap.wait_stream(cur)
    with torch.cuda.graph(g, stream=cap):
        for s in side:
            s.wait_stream(cap)
        for i in range(n):
            with torch.cuda.stream(side[i % k]):
                run_one(i)
        for s in side:
            cap.wait_stream(s)
    cur.wait_stream(cap)
    return g


class FrangiBatch3D:
    """Plain Frangi over many uneven 3-D volumes: one fused plan per volume
    (work arrays kept), the whole multi-sigma schedule of every volume
    captured into one HIP graph (``capture``) and replayed as one launch
    sequence.  Inputs are the bound float32 volumes; outputs are allocated.

    Reference behaviour per volume: filters/ridges.py:322-476
    (``frangi(vol, sigmas, alpha, beta, gamma, black_ridges, mode)``).
    """

    def __init__(self, volumes, sigmas=(1, 2, 3), alpha=0.5, beta=0.5, gamma=None,
                 black_ridges=False, mode="reflect"):
        from .skimage.filters.ridges import FrangiPlan

        if len(volumes) == 0:
            raise ValueError("no volumes")
        self.volumes = [_array.as_tensor(v).contiguous() for v in volumes]
        for v in self.volumes:
            if v.dtype != torch.float32 or v.dim() != 3:
                raise ValueError("volumes must be 3-D float32")
        self.sigmas = [float(s) for s in sigmas]
        self.alpha, self.beta, self.gamma = alpha, beta, gamma
        self.black_ridges, self.mode = black_ridges, mode
        self.plans = [FrangiPlan(tuple(v.shape), np.float32, self.sigmas, mode)
                      for v in self.volumes]
        self.outputs = [torch.empty_like(v) for v in self.volumes]
        self.voxels = sum(v.numel() for v in self.volumes)

    def run_one(self, i):
        rc = self.plans[i].run(self.volumes[i], self.outputs[i], self.alpha, self.beta,
                               self.gamma, self.black_ridges)
        _lib.check(rc, "frangi")

    @_trace.annotate("batch3d.run")
    def run(self):
        for i in range(len(self.volumes)):
            self.run_one(i)
        return self.outputs

    def capture(self, streams=8):
        """One HIP graph of the whole batch.  Volume i's schedule is recorded
        on stream i % streams (forked from and joined back to the capture
        stream), so the graph holds ``streams`` independent branches that
        the GPU runs concurrently: small volumes, whose launches alone do not
        fill the GPU, overlap (200 uneven volumes of 48..192^3: 103 ms with
        one branch, 68 ms with 8).  Each volume has its own plan, work
        arrays and scratch, so the branches share no memory."""
        return _capture_branches(len(self.volumes), self.run_one, streams)

    def latencies_ms(self, reps=1):
        """Per-volume latency of one volume's schedule run alone (events)."""
        out = []
        for i in range(len(self.volumes)):
            out.append(_time_ms(lambda: self.run_one(i), reps))
        return out
