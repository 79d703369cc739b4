"""Learned surrogate of the multi-scale Frangi stack (3-D CNN on the matrix
cores).

Five valid 3x3x3 convolutions, 24 channels, GELU (tanh form) between
layers, bf16
weights and activations with float32 accumulation
(``hc_cnn3d_conv``); the receptive field is 11^3, so a volume is padded by
5 voxels in the Frangi call's boundary mode and the network's output has the
volume's shape.  The target is ``frangi(volume, sigmas, black_ridges=False,
gamma=gamma)`` with gamma FIXED (a local operator cannot see the image-wide
maximum that ``gamma=None`` derives); ``gamma`` is part of the model.

Training uses hand-written kernels only: the matrix-core forward (with the
pre-activations kept), float32 input / weight gradients, GELU backward, MSE
and Adam (``hc_cnn3d_*``).  Patches are 64^3 crops of tube volumes labelled
by the plain GPU Frangi on the whole volume; the loss covers the 54^3
interior that the valid convolutions produce.

Per-brick fallback (``infer``): the volume is cut into bricks (64^3 by
default); each brick is checked on sampled voxels against the plain Frangi
of the brick's neighbourhood and recomputed with the plain kernels when the
sampled relative L2 error exceeds the threshold.
"""

import ctypes
import math
import os

import numpy as np
import torch

from . import _array, _lib, _trace

# (in channels, out channels, activation): 1 = GELU, 0 = none (training),
# 2 = clamp to [0, 1] (inference of the last layer)
LAYERS = ((1, 24, 1), (24, 24, 1), (24, 24, 1), (24, 24, 1), (24, 1, 0))
HALO = len(LAYERS)  # one voxel per valid 3x3x3 layer


def _dev():
    return _array.device()


def _check(rc, what):
    _lib.check(rc, what)


class FrangiSurrogate:
    def __init__(self, sigmas=(1, 2, 3, 4, 5), gamma=None, seed=0):
        self.sigmas = tuple(float(s) for s in sigmas)
        self.gamma = gamma
        g = torch.Generator().manual_seed(seed)
        self.w, self.b = [], []
        for ci, co, _ in LAYERS:
            fan_in = 27 * ci
            bound = math.sqrt(6.0 / fan_in) if co > 1 else math.sqrt(1.0 / fan_in)
            self.w.append(((torch.rand((27, ci, co), generator=g) * 2 - 1) * bound)
                          .to(_dev()).contiguous())
            self.b.append(torch.zeros(co, device=_dev()))
        self.packed = [None] * len(LAYERS)
        self.repack()

    # ------------------------------------------------------------ weights --
    def repack(self):
        lib = _lib.lib()
        for i, (ci, co, _) in enumerate(LAYERS):
            n = lib.hc_cnn3d_packed_size(ci, co)
            if self.packed[i] is None:
                self.packed[i] = torch.empty(n, dtype=torch.uint8, device=_dev())
            _check(lib.hc_cnn3d_pack(self.w[i].data_ptr(), self.packed[i].data_ptr(), ci, co,
                                     _array.stream()), "pack")

    def n_params(self):
        return sum(w.numel() + b.numel() for w, b in zip(self.w, self.b))

    def save(self, path):
        arrs = {f"w{i}": w.cpu().numpy() for i, w in enumerate(self.w)}
        arrs.update({f"b{i}": b.cpu().numpy() for i, b in enumerate(self.b)})
        arrs["sigmas"] = np.asarray(self.sigmas)
        arrs["gamma"] = np.asarray(-1.0 if self.gamma is None else float(self.gamma))
        np.savez(path, **arrs)

    @classmethod
    def load(cls, path):
        with np.load(path, allow_pickle=False) as z:
            gamma = float(z["gamma"])
            m = cls(sigmas=tuple(z["sigmas"].tolist()), gamma=None if gamma < 0 else gamma)
            for i in range(len(LAYERS)):
                m.w[i] = torch.from_numpy(z[f"w{i}"]).to(_dev()).contiguous()
                m.b[i] = torch.from_numpy(z[f"b{i}"]).to(_dev()).contiguous()
        m.repack()
        return m

    # ------------------------------------------------------------ forward --
    def pad_input(self, vol, mode="reflect", cval=0.0):
        """(D, H, W) float32 -> (D+10, H+10, W+10, 1) bfloat16, boundary-extended."""
        from .skimage._vendored._ndimage_util import _mode_code

        vol = _array.as_tensor(vol).contiguous()
        D, H, W = vol.shape
        out = torch.empty((D + 2 * HALO, H + 2 * HALO, W + 2 * HALO, 1), dtype=torch.bfloat16,
                          device=_dev())
        _check(_lib.lib().hc_cnn3d_pad(_array.hc_array(vol), _array.hc_array(out), HALO,
                                       _mode_code(mode), float(cval), _array.stream()), "pad")
        return out

    def forward(self, x, keep=False, final_act=2, bufs=None, out=None):
        """x: (D, H, W, 1) bf16 -> (D-10, H-10, W-10) float32.  keep: return
        the per-layer inputs and pre-activations (training)."""
        lib = _lib.lib()
        acts, pres = [x], []
        cur = x
        for i, (ci, co, act) in enumerate(LAYERS):
            D, H, W = (int(s) - 2 for s in cur.shape[:3])
            last = i == len(LAYERS) - 1
            if last:
                dst = out.view(D, H, W, 1) if out is not None else \
                    torch.empty((D, H, W, 1), dtype=torch.float32, device=_dev())
                act = final_act
            elif bufs is not None:
                dst = bufs[i % 2][: D * H * W * co].view(D, H, W, co)
            else:
                dst = torch.empty((D, H, W, co), dtype=torch.bfloat16, device=_dev())
            pre = torch.empty((D, H, W, co), dtype=torch.float32, device=_dev()) if keep else None
            _check(lib.hc_cnn3d_conv(_array.hc_array(cur), _array.hc_array(dst),
                                     self.packed[i].data_ptr(), self.b[i].data_ptr(), act,
                                     None if pre is None else pre.data_ptr(),
                                     _array.stream()), "conv3d")
            if keep:
                pres.append(pre)
                acts.append(dst)
            cur = dst
        y = cur.view(cur.shape[:3])
        return (y, acts, pres) if keep else y

    def __call__(self, vol, mode="reflect", cval=0.0):
        """Surrogate Frangi of a (D, H, W) float32 volume.  Work buffers
        (the padded input and two activation volumes) are kept for the last
        shape seen."""
        from .skimage._vendored._ndimage_util import _mode_code

        vol = _array.as_tensor(vol).contiguous()
        D, H, W = (int(s) + 2 * HALO for s in vol.shape)
        key = (D, H, W)
        if getattr(self, "_wkey", None) != key:
            self._wkey = None
            self._wbufs = None
            n1 = (D - 2) * (H - 2) * (W - 2) * 24
            self._wx = torch.empty((D, H, W, 1), dtype=torch.bfloat16, device=_dev())
            self._wbufs = [torch.empty(n1, dtype=torch.bfloat16, device=_dev())
                           for _ in range(2)]
            self._wkey = key
        _check(_lib.lib().hc_cnn3d_pad(_array.hc_array(vol), _array.hc_array(self._wx), HALO,
                                       _mode_code(mode), float(cval), _array.stream()), "pad")
        return self.forward(self._wx, bufs=self._wbufs)

    def release(self):
        """Free the inference work buffers."""
        self._wkey = self._wx = self._wbufs = None

    # ----------------------------------------------------------- training --
    def _grads(self):
        if not hasattr(self, "_gw"):
            self._gw = [torch.zeros_like(w) for w in self.w]
            self._gb = [torch.zeros_like(b) for b in self.b]
            self._mw = [torch.zeros_like(w) for w in self.w]
            self._vw = [torch.zeros_like(w) for w in self.w]
            self._mb = [torch.zeros_like(b) for b in self.b]
            self._vb = [torch.zeros_like(b) for b in self.b]
            self._step = 0
        return self._gw, self._gb

    def train_step(self, xs, ts, lr=1e-3):
        """One Adam step on a batch: xs list of (64, 64, 64, 1) bf16 inputs,
        ts list of (54, 54, 54) float32 targets.  Returns the mean loss
        (a device tensor)."""
        lib = _lib.lib()
        st = _array.stream()
        gw, gb = self._grads()
        for g in gw + gb:
            g.zero_()
        loss = torch.zeros(1, dtype=torch.float32, device=_dev())
        for x, t in zip(xs, ts):
            y, acts, pres = self.forward(x, keep=True, final_act=0)
            g = torch.empty_like(y)
            _check(lib.hc_cnn3d_mse_grad(y.data_ptr(), t.data_ptr(), g.data_ptr(), y.numel(),
                                         loss.data_ptr(), st), "mse")
            g = g.view(*y.shape, 1)
            for i in range(len(LAYERS) - 1, -1, -1):
                ci, co, act = LAYERS[i]
                xin = acts[i]
                D, H, W = (int(s) for s in xin.shape[:3])
                if act == 1:
                    _check(lib.hc_cnn3d_gelu_bwd(g.data_ptr(), pres[i].data_ptr(), g.numel(),
                                                 st), "gelu bwd")
                _check(lib.hc_cnn3d_wgrad(xin.data_ptr(), g.data_ptr(), gw[i].data_ptr(),
                                          gb[i].data_ptr(), D, H, W, ci, co, st), "wgrad")
                if i > 0:
                    dx = torch.empty((D, H, W, ci), dtype=torch.float32, device=_dev())
                    _check(lib.hc_cnn3d_dgrad(g.data_ptr(), self.w[i].data_ptr(),
                                              dx.data_ptr(), D, H, W, ci, co, st), "dgrad")
                    g = dx
        self._step += 1
        inv = 1.0 / len(xs)
        for i in range(len(LAYERS)):
            for p, gr, m, v in ((self.w[i], gw[i], self._mw[i], self._vw[i]),
                                (self.b[i], gb[i], self._mb[i], self._vb[i])):
                gr.mul_(inv)
                _check(lib.hc_cnn3d_adam(p.data_ptr(), gr.data_ptr(), m.data_ptr(),
                                         v.data_ptr(), p.numel(), lr, 0.9, 0.999, 1e-8,
                                         self._step, st), "adam")
        self.repack()
        return loss * inv


# ------------------------------------------------------------------ data --
def frangi_label(vol, sigmas, gamma, mode="reflect"):
    """Plain GPU Frangi (fused schedule) with a fixed gamma."""
    from .skimage.filters import frangi

    return frangi(vol, sigmas=sigmas, black_ridges=False, gamma=gamma, mode=mode)


def gamma_of(vol, sigmas):
    """The gamma that frangi(gamma=None) derives for ``vol`` (first sigma)."""
    from .skimage.filters.ridges import FrangiPlan

    vol = _array.as_tensor(vol).contiguous()
    plan = FrangiPlan(tuple(vol.shape), np.float32, list(sigmas), "reflect")
    out = torch.empty_like(vol)
    _check(plan.run(vol, out, black_ridges=False), "frangi")
    torch.cuda.synchronize()
    gamma_sq = float(plan.scratch[16:24].view(torch.float64).item())
    return math.sqrt(gamma_sq / 2.0)


def sample_patches(vols, labels, n, edge=64, seed=0, vessel_frac=0.75):
    """n random (edge^3 input bf16, (edge-10)^3 label) pairs; a fraction
    vessel_frac of them is centred on a voxel whose label exceeds a quarter
    of the volume's maximum (tube voxels are a few percent of a volume)."""
    rng = np.random.default_rng(seed)
    xs, ts = [], []
    hot = []
    for lab in labels:
        idx = torch.nonzero(lab > 0.25 * lab.max()).cpu().numpy()
        hot.append(idx)
    for _ in range(n):
        k = int(rng.integers(len(vols)))
        v, lab = vols[k], labels[k]
        if len(hot[k]) and rng.random() < vessel_frac:
            c = hot[k][int(rng.integers(len(hot[k])))]
            z, y, x = (int(min(max(int(c[d]) - edge // 2, 0), v.shape[d] - edge))
                       for d in range(3))
        else:
            z, y, x = (int(rng.integers(0, s - edge + 1)) for s in v.shape)
        p = v[z:z + edge, y:y + edge, x:x + edge]
        xs.append(p.to(torch.bfloat16).contiguous().view(edge, edge, edge, 1))
        ts.append(lab[z + HALO:z + edge - HALO, y + HALO:y + edge - HALO,
                      x + HALO:x + edge - HALO].contiguous())
    return xs, ts


def rel_l2(a, b):
    a, b = a.double(), b.double()
    return float(torch.linalg.vector_norm(a - b) / torch.linalg.vector_norm(b))


def dice(a, b, thr):
    ma, mb = a > thr, b > thr
    inter = (ma & mb).sum().item()
    tot = ma.sum().item() + mb.sum().item()
    return 1.0 if tot == 0 else 2.0 * inter / tot


def train(n_vols=8, edge=256, ntubes=40, n_patches=4000, batch=8, epochs=1, lr=2e-3,
          sigmas=(1, 2, 3, 4, 5), seed=0, log=None, max_seconds=None):
    """Train a surrogate on patches of seeded tube volumes (the benchmark's
    generator) labelled by the plain GPU Frangi.  gamma: the value
    frangi(gamma=None) derives on the first training volume, then fixed.
    Returns (model, history)."""
    import time

    from . import synthetic

    vols = [synthetic.tube_volume((edge,) * 3, ntubes=ntubes, seed=seed + 1000 + k)
            for k in range(n_vols)]
    gamma = gamma_of(vols[0], sigmas)
    labels = [frangi_label(v, sigmas, gamma) for v in vols]
    model = FrangiSurrogate(sigmas=sigmas, gamma=gamma, seed=seed)
    hist = []
    t0 = time.perf_counter()
    step = 0
    steps_per_epoch = max(1, n_patches // batch)
    for ep in range(epochs):
        xs_all, ts_all = sample_patches(vols, labels, steps_per_epoch * batch, seed=seed + ep)
        for it in range(steps_per_epoch):
            # cosine decay over the whole run
            frac = step / max(1, epochs * steps_per_epoch - 1)
            cur_lr = lr * 0.5 * (1 + math.cos(math.pi * frac))
            loss = model.train_step(xs_all[it * batch:(it + 1) * batch],
                                    ts_all[it * batch:(it + 1) * batch], lr=cur_lr)
            step += 1
            if step % 50 == 0 or step == 1:
                lv = float(loss)
                hist.append((step, lv, time.perf_counter() - t0))
                if log:
                    log(f"step {step} loss {lv:.3e} lr {cur_lr:.2e} "
                        f"{time.perf_counter() - t0:.1f} s")
            if max_seconds and time.perf_counter() - t0 > max_seconds:
                break
        if max_seconds and time.perf_counter() - t0 > max_seconds:
            break
    torch.cuda.synchronize()
    return model, hist


def evaluate(model, vol, mode="reflect"):
    """rel-L2 and Dice (threshold 0.5 * max of the plain result) of the
    surrogate against the plain Frangi with the model's gamma."""
    plain = frangi_label(vol, model.sigmas, model.gamma, mode)
    sur = model(vol, mode)
    thr = 0.5 * float(plain.max())
    return {"rel_l2": rel_l2(sur, plain), "dice": dice(sur, plain, thr),
            "threshold": thr}


def _main(argv=None):
    import argparse
    import json

    p = argparse.ArgumentParser(description="train the Frangi surrogate")
    p.add_argument("--out", required=True)
    p.add_argument("--epochs", type=int, default=1)
    p.add_argument("--patches", type=int, default=4000)
    p.add_argument("--batch", type=int, default=8)
    p.add_argument("--lr", type=float, default=2e-3)
    p.add_argument("--max-seconds", type=float, default=None)
    p.add_argument("--sigmas", default="1,2,3,4,5")
    a = p.parse_args(argv)
    torch.cuda.set_device(0)
    sig = tuple(float(v) for v in a.sigmas.split(","))
    model, hist = train(n_patches=a.patches, batch=a.batch, epochs=a.epochs, lr=a.lr,
                        sigmas=sig, log=lambda s: print(s, flush=True),
                        max_seconds=a.max_seconds)
    model.save(a.out)
    from . import synthetic

    held = synthetic.tube_volume((256,) * 3, ntubes=40, seed=2024)
    ev = evaluate(model, held)
    print(json.dumps({"gamma": model.gamma, "history": hist[-5:], "heldout_256": ev}),
          flush=True)


if __name__ == "__main__":
    _main()


# ------------------------------------------------------------- fallback --
def _composite_kernels(sigmas):
    """Per sigma the composite 1-D kernels (SS, SD, DD) of the plain Hessian's
    two passes along one axis (full convolution of the float32 pass kernels,
    computed in float64), their length L and anchor A."""
    from .skimage.feature.corner import hessian_kernels

    out = []
    for s in sigmas:
        w0, w1, anc = hessian_kernels(s, 3, np.float32)[0]
        a, d = w0.astype(np.float64), w1.astype(np.float64)
        ks = (np.convolve(a, a), np.convolve(a, d), np.convolve(d, d))
        out.append((ks, ks[0].size, 2 * int(anc)))
    return out


class CheckedSurrogate:
    """Surrogate Frangi of a (D, H, W) float32 volume with the per-brick plain
    fallback.  Every brick (``brick``^3) is checked on up to four lines of 64
    voxels along axis 0 against the plain Frangi evaluated exactly there
    (``hc_frangi_check``): its sampled relative L2 error is
    ||sur - plain|| / max(||plain||, floor * sqrt(n)) over its n samples, the
    floor being ``floor`` times the largest sampled plain value of the volume
    (so that bricks whose plain values are all ~0 pass when the surrogate is
    ~0 too).  A brick above ``tol`` -- or without a sample far enough from the
    volume's border -- is recomputed with the plain kernels on the brick plus
    the plain filter's full reach rounded up to 16 voxels (exact), or, when that would cost more than
    one plain pass over the volume, copied from one such pass."""

    def __init__(self, model, shape, brick=64, tol=0.02, mode="reflect", alpha=0.5, beta=0.5,
                 floor=0.05):
        if model.gamma is None:
            raise ValueError("the surrogate needs a fixed gamma")
        if int(brick) <= 0 or int(brick) % 16:
            raise ValueError("brick must be a positive multiple of 16")
        self.model, self.shape = model, tuple(int(s) for s in shape)
        self.brick, self.tol, self.mode = int(brick), float(tol), mode
        self.alpha, self.beta, self.floor = float(alpha), float(beta), float(floor)
        comp = _composite_kernels(model.sigmas)
        kc, info, off = [], [], 0
        amax = rmax = 0
        for ks, L, A in comp:
            kc += list(ks)
            info += [off, L, A]
            off += 3 * L
            amax, rmax = max(amax, A), max(rmax, L - 1 - A)
        self.reach = max(amax, rmax)
        # region halo: the reach rounded up to 16 voxels, so that a region
        # starts at a multiple of 16 along every axis (the matrix-core line
        # passes group a line's outputs in blocks of 16 from its start, and
        # the grouping fixes the rounding)
        self.halo = -(-self.reach // 16) * 16
        self.kc = torch.tensor(np.concatenate(kc).astype(np.float32), device=_dev())
        self.kinfo_host = (ctypes.c_int32 * len(info))(*info)
        self.kinfo = torch.tensor(info, dtype=torch.int32, device=_dev())
        D, H, W = self.shape
        b = self.brick
        self.grid = tuple(-(-n // b) for n in self.shape)
        self.nb = self.grid[0] * self.grid[1] * self.grid[2]
        lines, owner = [], []
        for i in range(self.nb):
            bz, by, bx = i // (self.grid[1] * self.grid[2]), (i // self.grid[2]) % self.grid[1], \
                i % self.grid[2]
            z0, y0, x0 = bz * b, by * b, bx * b
            z1, y1, x1 = min(z0 + b, D), min(y0 + b, H), min(x0 + b, W)
            za, zb = max(z0, amax), min(z1, D - rmax, max(z0, amax) + 64)
            seen = set()
            for fy, fx in ((1, 1), (1, 3), (3, 1), (3, 3)):
                xc = min(max(x0 + (x1 - x0) * fx // 4, amax), W - 1 - rmax)
                yc = min(max(y0 + (y1 - y0) * fy // 4, amax), H - 1 - rmax)
                if not (x0 <= xc < x1 and y0 <= yc < y1) or zb <= za or (xc, yc) in seen:
                    continue
                seen.add((xc, yc))
                lines.append((xc, yc, za, zb))
                owner.append(i)
        if not lines:
            lines, owner = [(0, 0, 0, 0)], [0]
        self.owner = np.asarray(owner)
        self.nl = len(lines)
        self.lines = torch.tensor(lines, dtype=torch.int32, device=_dev()).contiguous()
        self.plain_s = torch.empty(self.nl * 64, dtype=torch.float32, device=_dev())
        self.err = torch.empty(self.nl * 3, dtype=torch.float32, device=_dev())
        self._plans = {}

    def check(self, vol, sur):
        """Per-brick sampled relative L2 error (NaN: no sample in the brick)."""
        lib = _lib.lib()
        _check(lib.hc_frangi_check(_array.hc_array(vol), _array.hc_array(sur),
                                   self.lines.data_ptr(), self.nl, self.kc.data_ptr(),
                                   self.kinfo.data_ptr(), self.kinfo_host,
                                   len(self.model.sigmas), 0, self.alpha, self.beta,
                                   float(self.model.gamma), self.plain_s.data_ptr(),
                                   self.err.data_ptr(), _array.stream()), "frangi check")
        e = self.err.cpu().numpy().astype(np.float64).reshape(-1, 3)
        ps = self.plain_s.cpu().numpy()
        top = float(np.nanmax(ps)) if np.isfinite(ps).any() else 0.0
        dn = np.bincount(self.owner, e[:, 0], minlength=self.nb)
        nn = np.bincount(self.owner, e[:, 1], minlength=self.nb)
        n = np.bincount(self.owner, e[:, 2], minlength=self.nb)
        with np.errstate(divide="ignore", invalid="ignore"):
            den = np.maximum(np.sqrt(nn), self.floor * top * np.sqrt(n))
            rel = np.where(n > 0, np.sqrt(dn) / np.where(den > 0, den, 1.0), np.nan)
            rel = np.where((n > 0) & (den == 0), np.where(dn > 0, np.inf, 0.0), rel)
        return rel

    def _plain(self, region, center=None):
        from .skimage.filters.ridges import FrangiPlan

        shp = tuple(int(s) for s in region.shape)
        plan = self._plans.get(shp)
        if plan is None:
            plan = FrangiPlan(shp, np.float32, list(self.model.sigmas), self.mode)
            self._plans[shp] = plan
        out = torch.empty_like(region)
        _check(plan.run(region, out, alpha=self.alpha, beta=self.beta,
                        gamma=self.model.gamma, black_ridges=False, center=center),
               "plain frangi")
        return out

    @staticmethod
    def _center(vol):
        """The midrange the plain filter subtracts from the whole volume
        (0 with non-finite values), as a device scalar: recompute regions use
        it so that they reproduce the whole volume's values bit for bit."""
        from .skimage.exposure.exposure import _range

        lo, hi = _range(vol.reshape(-1))
        c = 0.5 * (lo + hi) if (np.isfinite(lo) and np.isfinite(hi)) else 0.0
        return torch.tensor([c], dtype=torch.float64, device=vol.device)

    def _region(self, i):
        """Brick i's bounds and those of its recompute region (the brick plus
        the halo, clipped to the volume)."""
        D, H, W = self.shape
        b, r = self.brick, self.halo
        gy, gx = self.grid[1], self.grid[2]
        bz, by, bx = i // (gy * gx), (i // gx) % gy, i % gx
        z0, y0, x0 = bz * b, by * b, bx * b
        z1, y1, x1 = min(z0 + b, D), min(y0 + b, H), min(x0 + b, W)
        return ((z0, z1, y0, y1, x0, x1),
                (max(0, z0 - r), min(D, z1 + r), max(0, y0 - r), min(H, y1 + r),
                 max(0, x0 - r), min(W, x1 + r)))

    def __call__(self, vol, sur=None, force=None):
        """Returns (output, info).  sur: a precomputed surrogate output;
        force: brick indices to recompute regardless of the check."""
        vol = _array.as_tensor(vol).contiguous()
        if sur is None:
            sur = self.model(vol, self.mode)
        rel = self.check(vol, sur)
        if force is not None:
            bad = sorted(set(int(i) for i in force))
        else:
            bad = [i for i in range(self.nb) if not (rel[i] <= self.tol)]
        D, H, W = self.shape
        regions = [self._region(i) for i in bad]
        region_vox = sum((rz1 - rz0) * (ry1 - ry0) * (rx1 - rx0)
                         for _, (rz0, rz1, ry0, ry1, rx0, rx1) in regions)
        mass = region_vox > D * H * W
        full = self._plain(vol) if (bad and mass) else None
        center = self._center(vol) if (bad and not mass) else None
        for (z0, z1, y0, y1, x0, x1), (rz0, rz1, ry0, ry1, rx0, rx1) in regions:
            if full is not None:
                src = full[z0:z1, y0:y1, x0:x1]
            else:
                reg_in = torch.empty((rz1 - rz0, ry1 - ry0, rx1 - rx0), dtype=vol.dtype,
                                     device=vol.device)
                _array.copy_cast(vol[rz0:rz1, ry0:ry1, rx0:rx1], reg_in)
                reg = self._plain(reg_in, center)
                src = reg[z0 - rz0:z1 - rz0, y0 - ry0:y1 - ry0, x0 - rx0:x1 - rx0]
            _array.copy_cast(src, sur[z0:z1, y0:y1, x0:x1])
        finite = rel[np.isfinite(rel)]
        info = {"bricks": self.nb, "fallback_bricks": len(bad), "mass_fallback": bool(mass),
                "sampled_rel_l2_median": float(np.median(finite)) if finite.size else None,
                "bricks_without_samples": int(np.isnan(rel).sum())}
        return sur, info


class SurrogateBatch3D:
    """Surrogate Frangi over many uneven volumes as one launch sequence that
    can be captured into a HIP graph: per volume the padding and the five
    layers, work buffers sized for the largest volume and reused in stream
    order."""

    def __init__(self, model, volumes, mode="reflect", cval=0.0):
        self.model, self.mode, self.cval = model, mode, float(cval)
        self.volumes = [_array.as_tensor(v).contiguous() for v in volumes]
        pads = [tuple(int(s) + 2 * HALO for s in v.shape) for v in self.volumes]
        self.xbuf = torch.empty(max(d * h * w for d, h, w in pads), dtype=torch.bfloat16,
                                device=_dev())
        n1 = max((d - 2) * (h - 2) * (w - 2) * 24 for d, h, w in pads)
        self.bufs = [torch.empty(n1, dtype=torch.bfloat16, device=_dev()) for _ in range(2)]
        self.outputs = [torch.empty_like(v) for v in self.volumes]
        self.voxels = sum(v.numel() for v in self.volumes)

    def run_one(self, i):
        from .skimage._vendored._ndimage_util import _mode_code

        v = self.volumes[i]
        D, H, W = (int(s) + 2 * HALO for s in v.shape)
        x = self.xbuf[: D * H * W].view(D, H, W, 1)
        _check(_lib.lib().hc_cnn3d_pad(_array.hc_array(v), _array.hc_array(x), HALO,
                                       _mode_code(self.mode), self.cval, _array.stream()),
               "pad")
        self.model.forward(x, bufs=self.bufs, out=self.outputs[i])

    @_trace.annotate("surrogate.run")
    def run(self):
        for i in range(len(self.volumes)):
            self.run_one(i)
        return self.outputs

    def capture(self):
        s = torch.cuda.Stream()
        s.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(s):
            self.run()
        torch.cuda.current_stream().wait_stream(s)
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            self.run()
        return g
