"""Learned surrogate of the multi-scale Frangi stack (3-D CNN on the matrix
cores).

Five valid 3x3x3 convolutions, 24 channels, GELU between layers, bf16
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

from . import _array, _lib

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

    def forward(self, x, keep=False, final_act=2, bufs=None):
        """x: (D, H, W, 1) bf16 -> (D-10, H-10, W-10) float32.  keep: return
        the per-layer inputs and pre-activations (training)."""
        lib = _lib.lib()
        acts, pres = [x], []
        cur = x
        for i, (ci, co, act) in enumerate(LAYERS):
            D, H, W = (int(s) - 2 for s in cur.shape[:3])
            last = i == len(LAYERS) - 1
            if last:
                out = torch.empty((D, H, W, 1), dtype=torch.float32, device=_dev())
                act = final_act
            elif bufs is not None:
                out = bufs[i % 2][: D * H * W * co].view(D, H, W, co)
            else:
                out = torch.empty((D, H, W, co), dtype=torch.bfloat16, device=_dev())
            pre = torch.empty((D, H, W, co), dtype=torch.float32, device=_dev()) if keep else None
            _check(lib.hc_cnn3d_conv(_array.hc_array(cur), _array.hc_array(out),
                                     self.packed[i].data_ptr(), self.b[i].data_ptr(), act,
                                     None if pre is None else pre.data_ptr(),
                                     _array.stream()), "conv3d")
            if keep:
                pres.append(pre)
                acts.append(out)
            cur = out
        y = cur.view(cur.shape[:3])
        return (y, acts, pres) if keep else y

    def __call__(self, vol, mode="reflect", cval=0.0):
        """Surrogate Frangi of a (D, H, W) float32 volume."""
        x = self.pad_input(vol, mode, cval)
        D, H, W = (int(s) for s in x.shape[:3])
        n1 = (D - 2) * (H - 2) * (W - 2) * 24
        bufs = [torch.empty(n1, dtype=torch.bfloat16, device=_dev()) for _ in range(2)]
        return self.forward(x, bufs=bufs)

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


def sample_patches(vols, labels, n, edge=64, seed=0):
    """n random (edge^3 input bf16, (edge-10)^3 label) pairs."""
    rng = np.random.default_rng(seed)
    xs, ts = [], []
    for _ in range(n):
        k = int(rng.integers(len(vols)))
        v, lab = vols[k], labels[k]
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
