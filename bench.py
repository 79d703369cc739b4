"""Benchmark of the Frangi / separable-filter path on one MI355X.

Workload (the headline, one "step"): plain Frangi vesselness of a 256^3
float32 volume of 40 random polyline tubes (radius U[1, 6], Gaussian cross
section, peak 1) + N(0, 0.05) noise, sigmas (1, 2, 3, 4, 5), alpha = beta =
0.5, black_ridges=False, mode='reflect', inputs resident in HBM.

Also measured: correlate1d and the fused 2-D Gaussian on a 2048^2 float32
image (K = 17, sigma 2, reflect) with rotating buffers larger than the
256 MB Infinity Cache, reported as achieved HBM GB/s of algorithmic bytes
(in + out).

Prints one JSON line.
"""

import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

SIGMAS = (1.0, 2.0, 3.0, 4.0, 5.0)
SHAPE = (256, 256, 256)
HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3      # published MI355X FP32
FP32_VALU_NOPACK_TFLOPS = 78.6  # without packed FP32 (disabled at compile time)


def _parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=20)
    p.add_argument("--warmup", type=int, default=3)
    p.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    p.add_argument("--cpu-edge", type=int, default=160,
                   help="edge of the cubic crop timed on the CPU (a bounded sample, "
                        "about 20 s on one core; 256: the whole volume)")
    p.add_argument("--parity-edge", type=int, default=96,
                   help="edge of the cubic crop compared against the FP64 oracle")
    p.add_argument("--only", choices=["all", "frangi", "correlate", "copy", "batch2d",
                                      "surrogate", "batch3d", "vesselness", "wsi"],
                   default="all",
                   help="restrict the run (profiler passes); 'copy' runs the plain copy "
                        "kernel that calibrates the HBM byte counters")
    p.add_argument("--hessian-valu", action="store_true",
                   help="run the fused Hessian passes on the vector ALUs only")
    p.add_argument("--no-check", action="store_true",
                   help="skip the fused-vs-reference-order cross-check")
    p.add_argument("--cpu-tiles", type=int, default=128,
                   help="tiles of the batched 2-D workload timed on the CPU pool")
    p.add_argument("--corr-sizes", default="2048,8192",
                   help="image edges of the correlate1d / Gaussian measurements")
    p.add_argument("--summarize", metavar="DIR",
                   help="condense the rocprofv3 CSV output under DIR and exit")
    return p.parse_args()


def _events_time(fn, reps):
    import torch

    s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    s.record()
    for i in range(reps):
        fn(i)
    e.record()
    torch.cuda.synchronize()
    return s.elapsed_time(e) / reps


def frangi_flops_per_voxel(plan):
    """Algorithmic FLOPs per voxel of the fused schedule: per sigma
    27 taps-sweeps (axis 0: 5, axis 1: 10, axis 2: 12 one-kernel sweeps of
    the kernel's K taps), 2 FLOPs per tap."""
    total = 0
    nd = plan.ndim
    for s in range(len(plan.sigmas)):
        K = [plan.ts[s * nd + x].K for x in range(nd)]
        total += 2 * (5 * K[0] + 10 * K[1] + 12 * K[2])
    return total


def per_kernel_stats(records):
    from collections import defaultdict

    agg = defaultdict(list)
    for name, ms in records:
        agg[name].append(ms)
    return {k: (len(v), sum(v) / len(v), sum(v)) for k, v in agg.items()}


DROPIN = {}


def bench_frangi(steps, warmup, check=True):
    import numpy as np
    import torch

    from cucim_amd import _lib, synthetic
    from cucim_amd.skimage.filters import ridges

    vol = synthetic.tube_volume(SHAPE, ntubes=40, seed=2024, noise=0.05)
    out = torch.empty_like(vol)
    plan = ridges.FrangiPlan(SHAPE, np.float32, SIGMAS, "reflect")

    def step(_):
        rc = plan.run(vol, out, alpha=0.5, beta=0.5, gamma=None, black_ridges=False)
        if rc != 0:
            _lib.check(rc, "frangi")

    for i in range(warmup):
        step(i)
    ms = _events_time(step, steps)
    # per-kernel launch times over a second run of the timed steps
    with _lib.KernelTimer() as kt:
        for i in range(steps):
            step(i)
        torch.cuda.synchronize()
    stats = per_kernel_stats(kt.records)
    # the drop-in call (a plan built per call, as frangi() does) for its wall
    # time next to the plan-reusing step
    from cucim_amd.skimage.filters import frangi

    frangi(vol, sigmas=SIGMAS, black_ridges=False)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(5):
        frangi(vol, sigmas=SIGMAS, black_ridges=False)
    torch.cuda.synchronize()
    DROPIN["frangi_wall_ms_per_call"] = (time.perf_counter() - t0) / 5 * 1e3
    if not check:
        return vol, plan, ms, stats, None, None
    # fused schedule vs the reference's pass order, both on the GPU
    ref = torch.empty_like(vol)
    ridges._frangi_sequential(vol, list(SIGMAS), 0.5, 0.5, None, False, "reflect", 0, ref)
    step(0)
    torch.cuda.synchronize()
    a, b = out.double(), ref.double()
    rel = float(torch.linalg.vector_norm(a - b) / torch.linalg.vector_norm(b))
    mx = float((a - b).abs().max())
    return vol, plan, ms, stats, rel, mx


def bench_correlate(sizes=(2048, 8192)):
    """correlate1d along the contiguous (-1) and the strided (0) axis, and the
    2-D Gaussian of skimage.filters.gaussian (sigma 2, reflect, truncate 4,
    preserve_range) on float32 images of 2048^2 (the Gaussian workload) and
    8192^2, K = 17; rotating buffers totalling > 256 MB so the Infinity Cache
    does not serve the reads.  GB/s count algorithmic bytes (in + out once)."""
    import numpy as np
    import torch

    from cucim_amd import _array, _lib
    from cucim_amd.skimage import filters
    from cucim_amd.skimage._vendored import _ndimage_filters as F

    res = {}
    lib = _lib.lib()
    st = _array.stream()
    w = F._gaussian_kernel1d(2.0, 0, 8).astype(np.float32)
    taps = F._Taps(w, 8, np.float32)
    for n in sizes:
        nbuf = max(2, (384 << 20) // (n * n * 8))
        g = torch.Generator(device="cuda")
        g.manual_seed(0)
        ins = [torch.rand((n, n), generator=g, device="cuda") for _ in range(nbuf)]
        outs = [torch.empty((n, n), device="cuda") for _ in range(nbuf)]
        hin = [_array.hc_array(t) for t in ins]
        hout = [_array.hc_array(t) for t in outs]
        algo_bytes = n * n * 8
        for axis, key in ((1, "correlate1d_axis-1"), (0, "correlate1d_axis0")):
            def run(i, axis=axis):
                k = i % nbuf
                lib.hc_correlate1d(hin[k], hout[k], taps.struct, axis, 0, 0.0, 0, st)

            for i in range(nbuf):
                run(i)
            with _lib.KernelTimer() as kt:
                ms_wall = _events_time(run, 8 * nbuf)
            st_ = per_kernel_stats(kt.records)
            kname, (cnt, avg, _) = max(st_.items(), key=lambda kv: kv[1][2])
            res[f"{key}_{n}"] = {"kernel": kname, "avg_ms": avg,
                                 "GBps": algo_bytes / (avg * 1e-3) / 1e9,
                                 "wall_ms_per_call": ms_wall}

        # the drop-in call as written by a user (new output each call), and
        # into a preallocated output; wall time measured without launch
        # tracing (which adds an event pair per launch), kernel time with it
        def run_g(i):
            filters.gaussian(ins[i % nbuf], sigma=2.0, mode="reflect")

        def run_go(i):
            k = i % nbuf
            filters.gaussian(ins[k], sigma=2.0, mode="reflect", out=outs[k])

        for i in range(nbuf):
            run_g(i)
            run_go(i)
        ms_wall = _events_time(run_g, 8 * nbuf)
        ms_wall_out = _events_time(run_go, 8 * nbuf)
        with _lib.KernelTimer() as kt:
            _events_time(run_g, 8 * nbuf)
        kern_ms = sum(m for _, m in kt.records) / (8 * nbuf)
        res[f"gaussian2d_{n}"] = {
            "kernels": sorted(set(k for k, _ in kt.records)),
            "kernel_ms_per_call": kern_ms,
            "GBps_kernels": algo_bytes / (kern_ms * 1e-3) / 1e9,
            "wall_ms_per_call": ms_wall,
            "wall_ms_per_call_out": ms_wall_out,
            "wall_over_kernel": ms_wall / kern_ms,
            "GBps_wall": algo_bytes / (ms_wall * 1e-3) / 1e9}

        # skimage sobel magnitude (reflect), one fused launch: 4 B in + 4 B out
        def run_s(i):
            k = i % nbuf
            filters.sobel(ins[k], mode="reflect")

        for i in range(nbuf):
            run_s(i)
        ms_wall = _events_time(run_s, 8 * nbuf)
        with _lib.KernelTimer() as kt:
            _events_time(run_s, 8 * nbuf)
        kern_ms = sum(m for _, m in kt.records) / (8 * nbuf)
        res[f"sobel2d_{n}"] = {
            "kernels": sorted(set(k for k, _ in kt.records)),
            "kernel_ms_per_call": kern_ms,
            "GBps_kernels": algo_bytes / (kern_ms * 1e-3) / 1e9,
            "wall_ms_per_call": ms_wall}
        del ins, outs, hin, hout
        # rgb2gray of uint8 RGB: 3 B in + 4 B out per pixel
        rgbs = [torch.randint(0, 256, (n, n, 3), generator=g, device="cuda",
                              dtype=torch.uint8) for _ in range(max(2, nbuf * 8 // 7))]
        from cucim_amd.skimage import color

        def run_c(i):
            color.rgb2gray(rgbs[i % len(rgbs)])

        for i in range(len(rgbs)):
            run_c(i)
        ms_wall = _events_time(run_c, 8 * len(rgbs))
        with _lib.KernelTimer() as kt:
            _events_time(run_c, 8 * len(rgbs))
        kern_ms = sum(m for _, m in kt.records) / (8 * len(rgbs))
        res[f"rgb2gray_{n}"] = {
            "kernels": sorted(set(k for k, _ in kt.records)),
            "kernel_ms_per_call": kern_ms,
            "GBps_kernels": n * n * 7 / (kern_ms * 1e-3) / 1e9,
            "wall_ms_per_call": ms_wall}
        del rgbs
    return res


def _cpu_tiles_worker(tiles):
    """CPU leg of the batched 2-D workload: the oracle's rgb2gray and the
    skimage-level gaussian / sobel over scipy.ndimage, one process."""
    import time as _time

    import numpy as np
    import scipy.ndimage as sndi

    from oracle import skimage as osk

    t0 = _time.perf_counter()
    for x in tiles:
        g = osk.gaussian(osk.rgb2gray(x), sigma=1.5, mode="nearest", ndi=sndi, dtype=np.float32)
        osk.sobel(g, mode="reflect", ndi=sndi, dtype=np.float32)
    return _time.perf_counter() - t0


def bench_batch2d(steps, cpu_tiles):
    """Batched uneven 2-D tiles: 512 RGB uint8 tiles, edges from {256, 384,
    512, 768, 1024}, rgb2gray -> gaussian(sigma 1.5, nearest) -> sobel
    magnitude (reflect), one persistent launch replayed from a HIP graph,
    tile shape autotuned.  Compulsory bytes: 3 in + 4 out per pixel."""
    import torch

    from cucim_amd import synthetic
    from cucim_amd.batch import GrayGaussSobelBatch
    from cucim_amd.skimage.color import rgb2gray
    from cucim_amd.skimage.filters import gaussian, sobel

    tiles = synthetic.rgb_tiles(512, seed=0)
    px = sum(int(t.shape[0]) * int(t.shape[1]) for t in tiles)
    plan = GrayGaussSobelBatch(tiles)
    tune = plan.autotune(reps=3)
    g = plan.capture()
    g.replay()
    ms = _events_time(lambda i: g.replay(), steps)
    # the same work as separate per-tile calls (the reference's usage)
    n_sep = 32
    sep_ms = _events_time(
        lambda i: [sobel(gaussian(rgb2gray(t), sigma=1.5)) for t in tiles[:n_sep]], 1)
    sep_px = sum(int(t.shape[0]) * int(t.shape[1]) for t in tiles[:n_sep])
    # parity: batched == separate calls, bit for bit, on a sample
    g.replay()
    torch.cuda.synchronize()
    mism = 0
    for t, o in zip(tiles[:16], plan.outputs[:16]):
        e = sobel(gaussian(rgb2gray(t), sigma=1.5))
        mism += int((o.view(torch.int32) != e.view(torch.int32)).sum())
    res = {"tiles": len(tiles), "pixels": px, "ms_per_batch": ms,
           "pixels_per_s": px / (ms * 1e-3), "GBps_compulsory": 7 * px / (ms * 1e-3) / 1e9,
           "autotune": tune, "rows": plan.rows,
           "separate_calls_pixels_per_s": sep_px / (sep_ms * 1e-3),
           "mismatched_pixels_vs_separate_calls_16_tiles": mism,
           "traffic_bytes_per_launch": read_traffic("k_tiles_ggs",
                                                    "batch2d_512tiles_counters.txt"),
           "compulsory_bytes_per_launch": 7 * px}
    if cpu_tiles > 0:
        import multiprocessing as mp

        host = [t.cpu().numpy() for t in tiles[:cpu_tiles]]
        # one GPU's share of the host: 16 cores on the MI355X boxes, whatever
        # os.cpu_count() reports for the whole machine
        cores = min(16, os.cpu_count() or 1)
        chunks = [host[i::cores] for i in range(cores)]
        ctx = mp.get_context("spawn")
        t0 = time.perf_counter()
        # workers end on their own (close + join): terminating them at the
        # pool's exit would SIGTERM processes a profiler is attached to
        pool = ctx.Pool(cores)
        try:
            pool.map(_cpu_tiles_worker, chunks)
            pool.close()
            pool.join()
        except BaseException:
            pool.terminate()
            raise
        wall = time.perf_counter() - t0
        cpx = sum(x.shape[0] * x.shape[1] for x in host)
        res["cpu_baseline"] = {"value": cpx / wall, "unit": "pixels/s", "cores": cores,
                               "kind": "port",
                               "sample": f"first {cpu_tiles} tiles ({cpx} px), process pool "
                                         f"of {cores}, {wall:.1f} s incl. pool start"}
    return res


SURROGATE_WEIGHTS = os.path.join(ROOT, "cucim_amd", "data", "frangi_surrogate.npz")


def bench_surrogate(steps, edge=512, ntubes=320):
    """Learned Frangi surrogate vs plain Frangi on an edge^3 float32 tube
    volume (the 256^3 generator at the same tube density: 320 tubes at
    512^3), sigmas 1..5, black_ridges=False, gamma fixed to the model's.
    Times the plain fused Frangi, the raw surrogate, and the checked
    surrogate (sampled per-brick check + plain fallback); rel-L2 and Dice
    (threshold 0.5 max of plain) against plain."""
    import numpy as np
    import torch

    from cucim_amd import synthetic
    from cucim_amd.skimage.filters.ridges import FrangiPlan
    from cucim_amd.surrogate import CheckedSurrogate, FrangiSurrogate, dice, rel_l2

    if not os.path.exists(SURROGATE_WEIGHTS):
        return {"error": "no trained weights"}
    model = FrangiSurrogate.load(SURROGATE_WEIGHTS)
    shape = (edge,) * 3
    vol = synthetic.tube_volume(shape, ntubes=ntubes, seed=4096)
    nvox = edge ** 3
    plan = FrangiPlan(shape, np.float32, list(model.sigmas), "reflect")
    plain = torch.empty_like(vol)

    def run_plain(_):
        rc = plan.run(vol, plain, gamma=model.gamma, black_ridges=False)
        if rc != 0:
            from cucim_amd import _lib
            _lib.check(rc, "frangi")

    run_plain(0)
    ms_plain = _events_time(run_plain, steps)
    sur = model(vol)
    ms_sur = _events_time(lambda i: model(vol), steps)
    res = {"workload": f"{edge}^3 float32, {ntubes} tubes, sigmas 1..5, gamma "
                       f"{model.gamma:.4f} (fixed)",
           "plain_ms": ms_plain, "plain_voxels_per_s": nvox / (ms_plain * 1e-3),
           "surrogate_ms": ms_sur, "surrogate_voxels_per_s": nvox / (ms_sur * 1e-3),
           "surrogate_speedup": ms_plain / ms_sur,
           "surrogate_rel_l2": rel_l2(sur, plain),
           "surrogate_dice": dice(sur, plain, 0.5 * float(plain.max())),
           "tflops_surrogate": 95904 * nvox / (ms_sur * 1e-3) / 1e12,
           "conv_traffic_bytes_per_launch": read_traffic("k_conv3d_mfma",
                                                         "surrogate_512cubed_counters.txt")}
    del sur
    cs = CheckedSurrogate(model, shape, brick=64, tol=0.02)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    out, info = cs(vol)
    torch.cuda.synchronize()
    ms_chk = (time.perf_counter() - t0) * 1e3
    res.update({"checked_ms": ms_chk, "checked_voxels_per_s": nvox / (ms_chk * 1e-3),
                "checked_speedup": ms_plain / ms_chk, "checked_rel_l2": rel_l2(out, plain),
                "checked_dice": dice(out, plain, 0.5 * float(plain.max())),
                "check": info})
    model.release()
    return res


def bench_batch3d(steps, n=200, seed=77):
    """Batched uneven 3-D volumes: n float32 tube volumes, each edge drawn
    independently from 48..192, tube density of the 256^3 generator (40
    tubes per 256^3), noise sigma 0.08; plain Frangi sigmas (1, 2, 3),
    black_ridges=False, reflect; the whole batch captured in one HIP graph.
    Per-volume latency: each volume's schedule run alone."""
    import numpy as np
    import torch

    from cucim_amd import synthetic
    from cucim_amd.batch import FrangiBatch3D

    rng = np.random.default_rng(seed)
    shapes = [tuple(int(v) for v in rng.integers(48, 193, 3)) for _ in range(n)]
    vols = []
    for i, s in enumerate(shapes):
        nt = max(2, int(round(40 * (s[0] * s[1] * s[2]) / 256 ** 3)))
        vols.append(synthetic.tube_volume(s, ntubes=nt, seed=seed * 1000 + i, noise=0.08))
    res = {"volumes": n, "sigmas": [1, 2, 3]}
    weights = os.path.join(ROOT, "cucim_amd", "data", "frangi_surrogate_s123.npz")
    sb = None
    # the configuration's plain Frangi: gamma=None, i.e. gamma from each
    # volume's own first-sigma Hessian norm
    b = FrangiBatch3D(vols, sigmas=(1, 2, 3), gamma=None)
    g1 = b.capture(streams=1)
    g1.replay()
    ms1 = _events_time(lambda i: g1.replay(), steps)
    del g1
    g = b.capture()
    g.replay()
    ms = _events_time(lambda i: g.replay(), steps)
    lat = np.asarray(b.latencies_ms())
    res.update({"voxels": b.voxels, "gamma": "per volume",
                "plain": {"ms_per_batch": ms, "voxels_per_s": b.voxels / (ms * 1e-3),
                          "graph_branches": 8, "ms_per_batch_one_branch": ms1,
                          "latency_ms_p50": float(np.percentile(lat, 50)),
                          "latency_ms_p99": float(np.percentile(lat, 99)),
                          "eager_sum_ms": float(lat.sum())}})
    if os.path.exists(weights):
        from cucim_amd.surrogate import FrangiSurrogate, SurrogateBatch3D, dice, rel_l2

        model = FrangiSurrogate.load(weights)
        sb = SurrogateBatch3D(model, vols)
        # the surrogate is trained for one fixed gamma: its comparison runs
        # the plain schedule with that gamma
        del g
        b = FrangiBatch3D(vols, sigmas=(1, 2, 3), gamma=model.gamma)
        g = b.capture()
        g.replay()
        ms = _events_time(lambda i: g.replay(), steps)
        res["plain_fixed_gamma"] = {"gamma": model.gamma, "ms_per_batch": ms}
    if sb is not None:
        gs = sb.capture()
        gs.replay()
        ms_s = _events_time(lambda i: gs.replay(), steps)
        lat_s = []
        for i in range(len(vols)):
            lat_s.append(_events_time(lambda j, i=i: sb.run_one(i), 1))
        lat_s = np.asarray(lat_s)
        gs.replay()
        b.run()
        torch.cuda.synchronize()
        allp = torch.cat([o.flatten() for o in b.outputs])
        alls = torch.cat([o.flatten() for o in sb.outputs])
        res["surrogate"] = {"ms_per_batch": ms_s, "voxels_per_s": b.voxels / (ms_s * 1e-3),
                            "speedup_vs_plain": ms / ms_s,
                            "latency_ms_p50": float(np.percentile(lat_s, 50)),
                            "latency_ms_p99": float(np.percentile(lat_s, 99)),
                            "rel_l2": rel_l2(alls, allp),
                            "dice": dice(alls, allp, 0.5 * float(allp.max()))}
    return res


COPY_BYTES = 256 << 20  # per direction


def bench_copy(reps=4):
    """The plain copy kernel (``k_copy_vec``, reached through ``hc_copy_cast``) over
    256 MiB -> 256 MiB: a kernel of known HBM traffic that calibrates the
    FETCH_SIZE / WRITE_SIZE counters."""
    import torch

    from cucim_amd import _array, _lib

    n = COPY_BYTES // 4
    src = torch.rand(n, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(2):
        _array.copy_cast(src, dst)
    with _lib.KernelTimer() as kt:
        for _ in range(reps):
            _array.copy_cast(src, dst)
        torch.cuda.synchronize()
    ms = sum(m for _, m in kt.records) / len(kt.records)
    return {"kernel": "k_copy_vec", "bytes_each_way": COPY_BYTES, "avg_ms": ms,
            "GBps": 2 * COPY_BYTES / (ms * 1e-3) / 1e9}


def summarize(root):
    """Condense rocprofv3 CSV output (``--output-format csv``) under ``root``:
    kernel statistics, and per-dispatch FETCH_SIZE / WRITE_SIZE (KiB) scaled
    by what the same counter reads for the plain copy kernel of known size."""
    import collections
    import csv
    import glob

    lines = []
    for f in sorted(glob.glob(os.path.join(root, "**", "*kernel_stats.csv"), recursive=True)):
        lines.append(f"# kernel stats: {os.path.relpath(f, root)}")
        with open(f) as fh:
            rows = list(csv.reader(fh))
        lines.append("Name,Calls,TotalDurationNs,AverageNs,Percentage,MinNs,MaxNs")
        for row in rows[1:]:
            if row and (row[0].startswith("void hc::") or row[0].startswith("hc::")):
                lines.append(",".join(row[:7]))
    counters = collections.defaultdict(list)
    for f in sorted(glob.glob(os.path.join(root, "**", "*counter_collection.csv"),
                              recursive=True)):
        with open(f) as fh:
            for row in csv.DictReader(fh):
                name = row.get("Kernel_Name", "?")
                short = name.split("(")[0].replace("void ", "").split("<")[0]
                try:
                    counters[(short, row.get("Counter_Name", "?"))].append(
                        float(row.get("Counter_Value", "nan")))
                except ValueError:
                    pass
    if counters:
        known_kib = COPY_BYTES / 1024
        factor = {}
        for c in ("FETCH_SIZE", "WRITE_SIZE"):
            v = counters.get(("hc::k_copy_vec", c))
            factor[c] = known_kib / (sum(v) / len(v)) if v else None
        lines.append("# counters (KiB per dispatch, mean over dispatches); calibration: "
                     f"k_copy_vec moves {COPY_BYTES} B each way")
        lines.append("# scale factors (known / counter): " + json.dumps(factor))
        traffic = {}
        for (k, c), v in sorted(counters.items()):
            if not k.startswith("hc::"):
                continue
            mean = sum(v) / len(v)
            lines.append(f"{k}, {c}, dispatches={len(v)}, mean_KiB={mean:.1f}")
            if factor.get(c):
                traffic.setdefault(k.replace("hc::", ""), {})[c] = mean * 1024 * factor[c]
        for k, d in traffic.items():
            if len(d) == 2:
                d["bytes"] = d["FETCH_SIZE"] + d["WRITE_SIZE"]
        lines.append("TRAFFIC " + json.dumps(traffic))
    return "\n".join(lines)


def read_traffic(kernel, profile):
    """HBM bytes per launch of ``kernel`` from a committed profile summary
    (the ``TRAFFIC`` line written by ``--summarize``), or None."""
    path = os.path.join(ROOT, "profiles", profile)
    try:
        with open(path) as fh:
            for ln in fh:
                if ln.startswith("TRAFFIC "):
                    d = json.loads(ln[len("TRAFFIC "):])
                    return d.get(kernel, {}).get("bytes")
    except OSError:
        return None
    return None


def host_info():
    """The host the CPU legs run on: core count, CPU model, library versions."""
    import numpy as np
    import scipy

    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("model name"):
                    model = ln.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"os_cpu_count": os.cpu_count(), "cpu_model": model, "numpy": np.__version__,
            "scipy": scipy.__version__}


def cpu_baseline(vol, edge, parity_edge):
    """The oracle's restatement of the reference path (scipy.ndimage passes,
    closed-form eigenvalues in numpy) on the volume (or a crop of edge^3), one
    thread, one repetition; parity of the GPU on a parity_edge^3 crop against
    the FP64 oracle."""
    import numpy as np
    import scipy.ndimage as sndi

    from cucim_amd.skimage.filters import frangi
    from oracle import skimage as osk

    x = vol[:edge, :edge, :edge].contiguous().cpu().numpy()
    t0 = time.perf_counter()
    osk.frangi(x, sigmas=SIGMAS, black_ridges=False, mode="reflect", ndi=sndi)
    t = time.perf_counter() - t0
    del x
    crop = vol[:parity_edge, :parity_edge, :parity_edge].contiguous()
    xc = crop.cpu().numpy()
    exp64 = osk.frangi(xc.astype(np.float64), sigmas=SIGMAS, black_ridges=False,
                       mode="reflect", ndi=sndi)
    got = frangi(crop, sigmas=SIGMAS, black_ridges=False, mode="reflect").cpu().numpy()
    err = np.abs(got - exp64)
    rel = float(np.linalg.norm(got - exp64) / np.linalg.norm(exp64))
    whole = "the whole 256^3 volume" if edge >= SHAPE[0] else f"{edge}^3 crop of the volume"
    return {"value": edge**3 / t, "unit": "voxels/s", "cores": 1, "kind": "port",
            "sample": f"{whole}, float32, same sigmas, one repetition: {t:.1f} s; "
                      "scipy.ndimage passes + numpy closed-form eigenvalues (one thread)",
            "host": host_info()}, \
        {"sample": f"{parity_edge}^3 crop", "rel_l2_vs_fp64": rel,
         "max_abs_vs_fp64": float(err.max()),
         "q9999_abs_vs_fp64": float(np.quantile(err, 0.9999))}


def bench_vesselness(reps=5, cpu=True):
    """The reference's published panel (notebooks/vesselness_example.ipynb,
    cell 1): meijering, sato, frangi and hessian, each with black_ridges 1 and
    0, sigmas=[2], mode='reflect', on rgb2gray(retina)[200:-200, 200:-200]
    (1011 x 1011 float32; the RGB crop is a committed fixture).  Timed as
    the filter calls alone, and with the notebook's per-result
    percentile(1, 99.5) and device-to-host copy (the published 0.125 s also
    includes matplotlib drawing, which is not run here)."""
    import numpy as np
    import torch

    from cucim_amd.skimage import color, filters

    rgb = np.load(os.path.join(ROOT, "tests", "golden", "retina_crop.npz"))["rgb"]
    img = color.rgb2gray(torch.from_numpy(rgb).cuda()).to(torch.float32).contiguous()
    funcs = [filters.meijering, filters.sato, filters.frangi, filters.hessian]
    q = torch.tensor([0.01, 0.995], device=img.device, dtype=torch.float32)

    def panel(with_host):
        for br in (1, 0):
            for f in funcs:
                r = f(img, sigmas=[2], mode="reflect", black_ridges=br)
                if with_host:
                    torch.quantile(r.reshape(-1), q).cpu()
                    r.cpu()

    res = {"image": list(img.shape), "published_gpu_s": 0.125, "published_cpu_s": 0.666}
    for with_host in (False, True):
        panel(with_host)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            panel(with_host)
        torch.cuda.synchronize()
        key = "panel_s_with_percentile_and_copy" if with_host else "panel_s_filters"
        res[key] = (time.perf_counter() - t0) / reps
    res["speedup_vs_published_gpu"] = 0.125 / res["panel_s_with_percentile_and_copy"]
    if cpu:
        import scipy.ndimage as sndi

        from oracle import skimage as osk

        x = img.cpu().numpy()
        t0 = time.perf_counter()
        for br in (1, 0):
            for name in ("meijering", "sato", "frangi", "hessian"):
                getattr(osk, name)(x, sigmas=[2], mode="reflect", black_ridges=bool(br),
                                   ndi=sndi)
        res["cpu_baseline"] = {"value": time.perf_counter() - t0, "unit": "s", "cores": 1,
                               "kind": "port",
                               "sample": "the same 8 filter calls on the host (one thread)"}
    return res


def bench_wsi(reps=5, cpu=True):
    """Whole-slide region read (the reference's CuImage.read_region): a
    synthetic 4096 x 4096 RGB pyramid (3 levels, 256 x 256 baseline-JPEG tiles,
    quality 90, 4:2:0, shared JPEGTables) written to a temporary file; the
    whole level 0 read into device memory (JPEG tiles decoded by the device
    kernels), alone and followed by rgb2gray + gaussian(sigma=2); the host
    decode path on the host threads; the CPU baseline is the oracle's
    libjpeg-turbo decode of every tile (one thread)."""
    import tempfile

    import numpy as np
    import torch

    from cucim_amd import _lib
    from cucim_amd.clara import CuImage
    from cucim_amd.skimage import color, filters
    from tests.wsi_tiff import pyramid, write_tiff

    edge = 4096
    res = {"image": [edge, edge, 3], "tile": [256, 256], "codec": "jpeg q90 4:2:0"}
    with tempfile.TemporaryDirectory() as d:
        path = write_tiff(os.path.join(d, "slide.tif"), pyramid(edge, edge, 3, seed=11),
                          tile=(256, 256), compression="jpeg", quality=90, subsampling=2)
        res["file_mb"] = os.path.getsize(path) / 1e6
        im = CuImage(path)
        out = torch.empty((edge, edge, 3), dtype=torch.uint8, device="cuda")

        def timed(fn, n):
            fn()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(n):
                fn()
            torch.cuda.synchronize()
            return (time.perf_counter() - t0) / n

        # reader threads for the device read: the tile bytes and headers
        # need few, and the library starts its threads per call (8 measure
        # faster than 16)
        nw = min(8, os.cpu_count() or 1)
        rd = lambda: im.read_region((0, 0), (edge, edge), 0, device="cuda", buf=out,  # noqa: E731
                                    num_workers=nw)
        t = timed(rd, reps)
        res["device_read_ms"] = t * 1e3
        res["device_read_mpx_s"] = edge * edge / t / 1e6
        with _lib.KernelTimer() as kt:
            rd()
            torch.cuda.synchronize()
        res["kernels_ms"] = {k: v for k, v in kt.collect()}
        pipe = lambda: filters.gaussian(color.rgb2gray(rd()), sigma=2.0)  # noqa: E731
        t = timed(pipe, reps)
        res["read_gray_gaussian_ms"] = t * 1e3
        res["host_threads_for_tile_reads"] = nw
        nwh = min(16, os.cpu_count() or 1)
        hb = np.empty((edge, edge, 3), np.uint8)
        im.read_region((0, 0), (edge, edge), 0, num_workers=nwh, buf=hb)
        t0 = time.perf_counter()
        im.read_region((0, 0), (edge, edge), 0, num_workers=nwh, buf=hb)
        th = time.perf_counter() - t0
        res["host_read"] = {"ms": th * 1e3, "mpx_s": edge * edge / th / 1e6, "threads": nwh}
        res["device_equals_host"] = bool(np.array_equal(out.cpu().numpy(), hb))
        if cpu:
            from oracle.wsi import TiffFile

            o = TiffFile(path)
            t0 = time.perf_counter()
            o.level_image(0)
            tc = time.perf_counter() - t0
            res["cpu_baseline"] = {"value": edge * edge / tc / 1e6, "unit": "Mpx/s",
                                   "cores": 1, "kind": "port",
                                   "sample": "every level-0 tile decoded by libjpeg-turbo "
                                             "(Pillow) and assembled, one thread"}
            # the same tile decodes on every host thread (Pillow releases the
            # GIL while libjpeg-turbo decodes)
            from concurrent.futures import ThreadPoolExecutor

            f0 = o.ifds[o.levels[0]]
            with ThreadPoolExecutor(nwh) as ex:
                list(ex.map(lambda i: o.decode_tile(f0, i), range(len(f0.offsets))))
                t0 = time.perf_counter()
                list(ex.map(lambda i: o.decode_tile(f0, i), range(len(f0.offsets))))
                tm = time.perf_counter() - t0
            res["cpu_baseline_all_threads"] = {"value": edge * edge / tm / 1e6, "unit": "Mpx/s",
                                               "cores": nwh, "kind": "port",
                                               "sample": "every level-0 tile decoded by "
                                                         "libjpeg-turbo (Pillow) on a thread "
                                                         "pool, not assembled"}
        im.close()
    return res


def _corr_roofline(corr):
    """HBM roofline of the correlate1d line kernel (contiguous axis, 8192^2
    float32, K = 17): algorithmic bytes (in + out) per launch over its
    average launch time."""
    if not corr or "correlate1d_axis-1_8192" not in corr:
        return None
    c = corr["correlate1d_axis-1_8192"]
    return {"bound": "hbm", "kernel": c["kernel"], "achieved": c["GBps"],
            "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": c["GBps"] / HBM_PEAK_GBS,
            "traffic": read_traffic(c["kernel"], "correlate1d_8192sq_counters.txt"),
            "algorithmic_bytes": 8192 * 8192 * 8, "avg_launch_ms": c["avg_ms"]}


def main():
    args = _parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this benchmark runs on one GPU; --gpus must be 1",
                          "n_gpus": args.gpus}))
        sys.exit(2)
    if args.summarize:
        print(summarize(args.summarize))
        return
    import torch

    torch.cuda.set_device(0)
    if args.hessian_valu:
        from cucim_amd import _lib

        _lib.check(_lib.lib().hc_tune_set(_lib.TUNE_HESSIAN_VALU, 1))
    if args.only == "copy":
        print(json.dumps({"copy": bench_copy()}))
        return
    if args.only == "batch3d":
        print(json.dumps({"batch3d": bench_batch3d(args.steps)}))
        return
    if args.only == "surrogate":
        print(json.dumps({"surrogate": bench_surrogate(max(2, args.steps // 4))}))
        return
    if args.only == "batch2d":
        print(json.dumps({"batch2d": bench_batch2d(args.steps, 0 if args.no_cpu
                                                   else args.cpu_tiles)}))
        return
    if args.only == "vesselness":
        print(json.dumps({"vesselness": bench_vesselness(max(2, args.steps // 4),
                                                         not args.no_cpu)}))
        return
    if args.only == "wsi":
        print(json.dumps({"wsi_read": bench_wsi(max(2, args.steps // 4), not args.no_cpu)}))
        return
    if args.only == "correlate":
        sizes = tuple(int(v) for v in args.corr_sizes.split(","))
        print(json.dumps({"copy": bench_copy(), "correlate1d": bench_correlate(sizes)}))
        return
    vol, plan, ms, stats, rel_fused, max_fused = bench_frangi(args.steps, args.warmup,
                                                               not args.no_check)
    nvox = math.prod(SHAPE)
    value = nvox / (ms * 1e-3)
    # dominant kernel of the step by total time
    dom, (cnt, avg_ms, _) = max(stats.items(), key=lambda kv: kv[1][2])
    flops_vox = frangi_flops_per_voxel(plan)
    # per-launch algorithmic FLOPs of the dominant kernel
    nd = 3
    Ks = [[plan.ts[s * nd + x].K for x in range(nd)] for s in range(len(plan.sigmas))]
    if dom in ("k_final_contig", "k_final_mfma"):
        fl = sum(2 * 12 * k[2] for k in Ks) * nvox / len(Ks)
    elif dom in ("k_chain_strided", "k_chain_mfma"):
        fl = sum(2 * (5 * k[0] + 10 * k[1]) for k in Ks) * nvox / (2 * len(Ks))
    else:
        fl = 0.0
    achieved = fl / (avg_ms * 1e-3) / 1e12 if fl else None
    on_mfma = dom.endswith("_mfma")
    corr = None
    if args.only == "all":
        corr = bench_correlate(tuple(int(v) for v in args.corr_sizes.split(",")))
    batch2d = surrogate = batch3d = vessel = wsi = None
    if args.only == "all":
        vessel = bench_vesselness(max(2, args.steps // 4), not args.no_cpu)
        wsi = bench_wsi(max(2, args.steps // 4), not args.no_cpu)
        batch2d = bench_batch2d(args.steps, 0 if args.no_cpu else args.cpu_tiles)
        batch3d = bench_batch3d(max(2, args.steps // 4))
        surrogate = bench_surrogate(max(2, args.steps // 4))
    cpu, parity = (None, None)
    if not args.no_cpu and args.only == "all":
        cpu, parity = cpu_baseline(vol, args.cpu_edge, args.parity_edge)
    line = {
        "metric": "Frangi voxels/s plain vs surrogate at rel-L2<=2%; correlate1d HBM GB/s",
        "value": value,
        "unit": "voxels/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded tube volume, generated on the device)",
        "config": {"workload": "plain 3-D Frangi, 256^3 float32, 40 polyline tubes + "
                               "N(0,0.05), sigmas 1..5, alpha=beta=0.5, "
                               "black_ridges=False, reflect",
                   "voxels": nvox},
        "roofline": {
            "bound": "mfma" if on_mfma else "valu",
            "kernel": dom,
            "achieved": achieved,
            "peak": FP32_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": (achieved / FP32_PEAK_TFLOPS) if achieved else None,
            "traffic": read_traffic(dom, "frangi_256cubed_counters.txt"),
            "note": ("separable passes as Toeplitz products on v_mfma_f32_16x16x4f32 "
                     "(dense FP32 matrix peak); achieved counts the K real taps per "
                     "output, the matrix cores also multiply the zero band of each "
                     "16-output block (4*ceil((K+15)/4) products per output)")
                    if on_mfma else
                    ("FP32 FMA sweeps of the separable passes; packed FP32 is disabled at "
                     "compile time, so the reachable VALU ceiling is "
                     f"{FP32_VALU_NOPACK_TFLOPS} TFLOP/s"),
            "avg_launch_ms": avg_ms,
            "launches_per_step": cnt / args.steps,
        },
        "roofline_correlate1d": _corr_roofline(corr),
        "kernels_ms_per_step": {k: v[2] / args.steps for k, v in stats.items()},
        "frangi_flops_per_voxel": flops_vox,
        "frangi_tflops": flops_vox * nvox / (ms * 1e-3) / 1e12,
        "fused_vs_reference_order": {"rel_l2": rel_fused, "max_abs": max_fused},
        "dropin_frangi_wall_ms_per_call": DROPIN.get("frangi_wall_ms_per_call"),
        "correlate1d": corr,
        "cpu_baseline": cpu,
        "parity": parity,
        "vesselness_panel": vessel,
        "wsi_read": wsi,
        "batch2d": batch2d,
        "batch3d": batch3d,
        "surrogate": surrogate,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
