"""The surrogate network's kernels against a float64 PyTorch reference on
the CPU (inputs and weights rounded to bfloat16 as the kernels see them).

Pre-activations (float32 accumulation of 27 * 24 products) within 1e-5 of
the reference relative to the layer's scale; activations within bfloat16
rounding (2^-8 relative); gradients within 1e-4 relative.
"""

import math

import numpy as np
import pytest
import torch
import torch.nn.functional as Fnn

pytestmark = pytest.mark.gpu


def _lib():
    from cucim_amd import _lib as L

    return L


def _ref_conv(x_bf, w, b):
    """x_bf: (D, H, W, ci) bf16 tensor; w: (27, ci, co) float32 (rounded to
    bf16 here); returns (D-2, H-2, W-2, co) float64 pre-activation."""
    x = x_bf.float().double().cpu().permute(3, 0, 1, 2)[None]
    wq = w.to(torch.bfloat16).double().cpu()
    ci, co = w.shape[1], w.shape[2]
    k = wq.view(3, 3, 3, ci, co).permute(4, 3, 0, 1, 2)  # (co, ci, dz, dy, dx)
    y = Fnn.conv3d(x, k) + b.double().cpu().view(1, co, 1, 1, 1)
    return y[0].permute(1, 2, 3, 0)


def _run_conv(x, w, b, act, out_f32=False):
    from cucim_amd import _array
    from cucim_amd.surrogate import FrangiSurrogate  # noqa: F401

    L = _lib()
    lib = L.lib()
    ci, co = w.shape[1], w.shape[2]
    packed = torch.empty(lib.hc_cnn3d_packed_size(ci, co), dtype=torch.uint8, device="cuda")
    L.check(lib.hc_cnn3d_pack(w.data_ptr(), packed.data_ptr(), ci, co, _array.stream()))
    D, H, W = (int(s) - 2 for s in x.shape[:3])
    out = torch.empty((D, H, W, co), dtype=torch.float32 if out_f32 else torch.bfloat16,
                      device="cuda")
    pre = torch.empty((D, H, W, co), dtype=torch.float32, device="cuda")
    L.check(lib.hc_cnn3d_conv(_array.hc_array(x), _array.hc_array(out), packed.data_ptr(),
                              b.data_ptr(), act, pre.data_ptr(), _array.stream()))
    torch.cuda.synchronize()
    return out, pre


@pytest.mark.parametrize("ci,co,act,f32", [(1, 24, 1, False), (24, 24, 1, False),
                                           (24, 1, 0, True), (24, 1, 2, True)])
@pytest.mark.parametrize("shape", [(4, 10, 18), (7, 13, 40), (20, 3, 3), (3, 21, 35)])
def test_conv_layer(gpu, ci, co, act, f32, shape):
    g = torch.Generator().manual_seed(1)
    x = (torch.randn(shape + (ci,), generator=g)).to(torch.bfloat16).cuda()
    w = (torch.randn((27, ci, co), generator=g) / math.sqrt(27 * ci)).cuda()
    b = (torch.randn(co, generator=g) * 0.1).cuda()
    out, pre = _run_conv(x, w, b, act, f32)
    ref = _ref_conv(x, w, b)
    scale = float(ref.abs().max())
    np.testing.assert_allclose(pre.double().cpu().numpy(), ref.numpy(), rtol=0,
                               atol=1e-5 * max(scale, 1.0))
    if act == 1:
        exp = Fnn.gelu(pre.double().cpu(), approximate="tanh")
    elif act == 2:
        exp = pre.double().cpu().clamp(0, 1)
    else:
        exp = pre.double().cpu()
    got = out.double().cpu()
    if f32:
        np.testing.assert_allclose(got.numpy(), exp.numpy(), rtol=1e-6, atol=1e-6)
    else:
        # float32 GELU (hardware exp / reciprocal, ~1e-6 relative) then
        # round-to-nearest bf16: one bf16 ulp at most
        np.testing.assert_allclose(got.numpy(), exp.to(torch.bfloat16).double().numpy(),
                                   rtol=2.0 ** -7, atol=1e-5)


def test_backward_kernels(gpu):
    """dgrad / wgrad / gelu backward / mse against autograd (float64, CPU)."""
    from cucim_amd import _array

    L = _lib()
    lib = L.lib()
    st = _array.stream()
    g = torch.Generator().manual_seed(2)
    D, H, W, ci, co = 6, 7, 9, 24, 24
    x = torch.randn((D, H, W, ci), generator=g).to(torch.bfloat16)
    w = torch.randn((27, ci, co), generator=g) / 25
    dy = torch.randn((D - 2, H - 2, W - 2, co), generator=g)
    # autograd reference
    xr = x.double().permute(3, 0, 1, 2)[None].requires_grad_(True)
    wr = w.double().view(3, 3, 3, ci, co).permute(4, 3, 0, 1, 2).contiguous().requires_grad_(True)
    br = torch.zeros(co, dtype=torch.float64, requires_grad=True)
    y = Fnn.conv3d(xr, wr, br)
    y.backward(dy.double().permute(3, 0, 1, 2)[None])
    dx_ref = xr.grad[0].permute(1, 2, 3, 0)
    dw_ref = wr.grad.permute(2, 3, 4, 1, 0).reshape(27, ci, co)
    db_ref = br.grad
    xd, wd, dyd = x.cuda(), w.cuda().contiguous(), dy.cuda().contiguous()
    dx = torch.empty((D, H, W, ci), device="cuda")
    L.check(lib.hc_cnn3d_dgrad(dyd.data_ptr(), wd.data_ptr(), dx.data_ptr(), D, H, W, ci, co, st))
    dw = torch.zeros((27, ci, co), device="cuda")
    db = torch.zeros(co, device="cuda")
    L.check(lib.hc_cnn3d_wgrad(xd.data_ptr(), dyd.data_ptr(), dw.data_ptr(), db.data_ptr(), D,
                               H, W, ci, co, st))
    torch.cuda.synchronize()
    for got, ref in ((dx, dx_ref), (dw, dw_ref), (db, db_ref)):
        r = ref.detach()
        np.testing.assert_allclose(got.double().cpu().numpy(), r.numpy(), rtol=0,
                                   atol=1e-4 * float(r.abs().max()))
    # gelu backward
    pre = torch.randn(1000, generator=g).cuda()
    gg = torch.randn(1000, generator=g).cuda()
    exp = gg.double().cpu()
    p = pre.double().cpu().requires_grad_(True)
    Fnn.gelu(p, approximate="tanh").backward(exp)
    g2 = gg.clone()
    L.check(lib.hc_cnn3d_gelu_bwd(g2.data_ptr(), pre.data_ptr(), g2.numel(), st))
    torch.cuda.synchronize()
    np.testing.assert_allclose(g2.double().cpu().numpy(), p.grad.numpy(), rtol=1e-4, atol=1e-5)
    # mse
    yv, tv = torch.randn(5000, generator=g).cuda(), torch.randn(5000, generator=g).cuda()
    gm = torch.empty_like(yv)
    loss = torch.zeros(1, device="cuda")
    L.check(lib.hc_cnn3d_mse_grad(yv.data_ptr(), tv.data_ptr(), gm.data_ptr(), 5000,
                                  loss.data_ptr(), st))
    torch.cuda.synchronize()
    d = (yv - tv).double().cpu()
    assert abs(float(loss) - float((d * d).mean())) <= 1e-5 * float((d * d).mean())
    np.testing.assert_allclose(gm.double().cpu().numpy(), (2 * d / 5000).numpy(), rtol=1e-6)


def test_adam(gpu):
    from cucim_amd import _array

    L = _lib()
    lib = L.lib()
    g = torch.Generator().manual_seed(3)
    p = torch.randn(300, generator=g).cuda()
    p0 = p.clone()
    grad = torch.randn(300, generator=g).cuda()
    m = torch.zeros(300, device="cuda")
    v = torch.zeros(300, device="cuda")
    L.check(lib.hc_cnn3d_adam(p.data_ptr(), grad.data_ptr(), m.data_ptr(), v.data_ptr(), 300,
                              1e-2, 0.9, 0.999, 1e-8, 1, _array.stream()))
    torch.cuda.synchronize()
    # first step: update = lr * g / (|g| + eps)
    exp = p0 - 1e-2 * grad / (grad.abs() + 1e-8)
    np.testing.assert_allclose(p.cpu().numpy(), exp.cpu().numpy(), rtol=1e-5, atol=1e-7)


def test_surrogate_forward_shape_and_training_descends(gpu):
    from cucim_amd import synthetic
    from cucim_amd.surrogate import FrangiSurrogate, frangi_label, sample_patches

    vol = synthetic.tube_volume((80, 80, 80), ntubes=6, seed=5)
    lab = frangi_label(vol, (1, 2, 3), gamma=0.5)
    m = FrangiSurrogate(sigmas=(1, 2, 3), gamma=0.5, seed=1)
    y = m(vol)
    assert y.shape == vol.shape and y.dtype == torch.float32
    assert bool(torch.isfinite(y).all()) and float(y.min()) >= 0 and float(y.max()) <= 1
    xs, ts = sample_patches([vol], [lab], 4, edge=32, seed=0)
    losses = [float(m.train_step(xs, ts, lr=3e-3)) for _ in range(30)]
    assert losses[-1] < 0.7 * losses[0], losses


def _sample_positions(cs):
    lines = cs.lines.cpu().numpy()
    pos = []
    for li, (x0, y0, za, zb) in enumerate(lines):
        for k, z in enumerate(range(za, zb)):
            pos.append((li, k, z, y0, x0))
    return pos


def test_frangi_check_samples_match_plain(gpu):
    """The sampled exact plain values equal the plain Frangi (fused kernels,
    same fixed gamma) at those voxels within float32 rounding; the per-brick
    error follows from the samples."""
    from cucim_amd import synthetic
    from cucim_amd.surrogate import CheckedSurrogate, FrangiSurrogate, frangi_label

    vol = synthetic.tube_volume((176, 160, 168), ntubes=12, seed=3)
    m = FrangiSurrogate(sigmas=(1, 2, 3, 4, 5), gamma=0.17, seed=2)
    cs = CheckedSurrogate(m, vol.shape, brick=64)
    sur = m(vol)
    rel = cs.check(vol, sur)
    plain = frangi_label(vol, m.sigmas, m.gamma).cpu().numpy()
    samp = cs.plain_s.cpu().numpy().reshape(-1, 64)
    pos = _sample_positions(cs)
    assert len(pos) > 0
    got = np.array([samp[li, k] for li, k, z, y, x in pos])
    exp = np.array([plain[z, y, x] for li, k, z, y, x in pos])
    np.testing.assert_allclose(got, exp, rtol=1e-3, atol=1e-5 * max(1.0, exp.max()))
    s = sur.cpu().numpy()
    top = got.max()
    for b in range(cs.nb):
        pb = [(z, y, x) for li, k, z, y, x in pos if cs.owner[li] == b]
        if not pb:
            assert np.isnan(rel[b])
            continue
        e = np.array([plain[p] for p in pb])
        q = np.array([s[p] for p in pb])
        den = max(np.linalg.norm(e), cs.floor * top * np.sqrt(len(pb)))
        assert abs(rel[b] - np.linalg.norm(q - e) / den) <= 1e-2 * rel[b] + 1e-5


@pytest.mark.parametrize("path", ["region", "mass"])
def test_fallback_is_plain(gpu, path):
    """Recomputed bricks equal the plain Frangi of the whole volume bit for
    bit (per-brick regions carry the plain filter's full reach)."""
    from cucim_amd import synthetic
    from cucim_amd.surrogate import CheckedSurrogate, FrangiSurrogate, frangi_label

    vol = synthetic.tube_volume((256, 192, 200), ntubes=16, seed=4)
    m = FrangiSurrogate(sigmas=(1, 2, 3), gamma=0.2, seed=3)
    cs = CheckedSurrogate(m, vol.shape, brick=64)
    plain = frangi_label(vol, m.sigmas, m.gamma).cpu().numpy()
    raw = m(vol).cpu().numpy()
    force = [0, 5, cs.nb - 1] if path == "region" else list(range(cs.nb))
    out, info = cs(vol, force=force)
    assert info["mass_fallback"] == (path == "mass")
    out = out.cpu().numpy()
    gz, gy, gx = cs.grid
    for i in range(cs.nb):
        bz, by, bx = i // (gy * gx), (i // gx) % gy, i % gx
        sl = (slice(bz * 64, bz * 64 + 64), slice(by * 64, by * 64 + 64),
              slice(bx * 64, bx * 64 + 64))
        ref = plain[sl] if i in force else raw[sl]
        np.testing.assert_array_equal(out[sl].view(np.uint32), ref.view(np.uint32))


def test_surrogate_batch_graph_equals_eager(gpu):
    from cucim_amd import synthetic
    from cucim_amd.surrogate import FrangiSurrogate, SurrogateBatch3D

    m = FrangiSurrogate(sigmas=(1, 2, 3), gamma=0.2, seed=7)
    vols = [synthetic.tube_volume(s, ntubes=3, seed=i, noise=0.08)
            for i, s in enumerate([(48, 70, 52), (90, 49, 61), (64, 64, 64)])]
    sb = SurrogateBatch3D(m, vols)
    g = sb.capture()
    g.replay()
    torch.cuda.synchronize()
    got = [o.clone() for o in sb.outputs]
    for v, o in zip(vols, got):
        exp = m(v)
        torch.cuda.synchronize()
        assert torch.equal(o.view(torch.int32), exp.view(torch.int32))
