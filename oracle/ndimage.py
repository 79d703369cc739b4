"""Oracle: numpy restatement of the reference's vendored ndimage filter core.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

Reference files (relative to ``python/cucim/src/cucim/skimage/``):
  _vendored/_ndimage_filters.py       correlate1d/convolve1d, gaussian_filter, sobel
  _vendored/_ndimage_filters_core.py  elementwise n-D kernel, cast<>, _run_1d_filters
  _vendored/_ndimage_util.py          boundary modes, weights/output dtype rules

Accumulation:
  ``acc="f64"``      float64 weights and float64 accumulation (the FP64 oracle)
  ``acc="emulate"``  the reference's own arithmetic: weights cast to the weights
                     dtype W, ``sum = fma(cast<W>(x), w, sum)`` over the taps in
                     order, skipping zero weights, intermediates of multi-pass
                     filters stored in the output dtype.  For W = float32 the
                     fused multiply-add is emulated exactly (float64 product,
                     TwoSum error term, correct final rounding); for
                     W = float64 a plain float64 multiply-add is used.
"""

import numpy as np

__all__ = [
    "canonical_mode",
    "remap_index",
    "weights_dtype",
    "supported_float_type",
    "cast_to",
    "fma_f32",
    "correlate1d",
    "convolve1d",
    "run_1d_filters",
    "gaussian_kernel1d",
    "gaussian_filter1d",
    "gaussian_filter",
    "sobel",
    "prewitt",
    "uniform_filter1d",
    "uniform_filter",
    "correlate",
]

# _ndimage_util.py:95-108 lists the accepted names.  For filters 'wrap' is
# 'grid-wrap' (_ndimage_filters.py:310-311, _ndimage_filters_core.py:137-138);
# 'grid-mirror' shares reflect's remap (_ndimage_util.py:198);
# 'grid-constant' and 'grid-reflect' follow scipy.ndimage's filter semantics
# (same as 'constant' and 'reflect').
_MODE_ALIASES = {
    "reflect": "reflect",
    "grid-mirror": "reflect",
    "grid-reflect": "reflect",
    "mirror": "mirror",
    "nearest": "nearest",
    "wrap": "wrap",
    "grid-wrap": "wrap",
    "constant": "constant",
    "grid-constant": "constant",
}


def canonical_mode(mode):
    """Map a boundary-mode name to one of the five remap rules.

    Raises RuntimeError for unknown names, like ``_ndimage_util.py:95-108``.
    """
    try:
        return _MODE_ALIASES[mode]
    except (KeyError, TypeError):
        raise RuntimeError(f"boundary mode not supported (actual: {mode})")


def remap_index(i, n, mode):
    """Boundary-extend integer positions ``i`` of an axis of length ``n``.

    Follows ``_ndimage_util.py:186-308`` (the non-separate form):
      reflect : i<0 -> -1-i ; i %= 2n ; i = min(i, 2n-1-i)
      mirror  : n==1 -> 0 ; i<0 -> -i ; i = 1 + (i-1) % (2n-2) ; min(i, 2n-2-i)
      nearest : clamp to [0, n-1]
      wrap    : i %= n ; i<0 -> i+n           (grid-wrap)
      constant: out of range -> -1 (caller substitutes cval)
    C's truncating ``%`` is reproduced with ``np.fmod``.
    """
    mode = canonical_mode(mode)
    i = np.asarray(i, dtype=np.int64).copy()
    n = int(n)
    if mode == "reflect":
        i = np.where(i < 0, -1 - i, i)
        i = np.fmod(i, 2 * n)
        i = np.minimum(i, 2 * n - 1 - i)
    elif mode == "mirror":
        if n == 1:
            return np.zeros_like(i)
        i = np.where(i < 0, -i, i)
        i = 1 + np.fmod(i - 1, 2 * n - 2)
        i = np.minimum(i, 2 * n - 2 - i)
    elif mode == "nearest":
        i = np.clip(i, 0, n - 1)
    elif mode == "wrap":
        i = np.fmod(i, n)
        i = np.where(i < 0, i + n, i)
    else:  # constant
        i = np.where((i < 0) | (i >= n), -1, i)
    return i


# ``_shared/utils.py:1131-1149``: cuCIM maps 8/16-bit ints and bool to float32.
_NEW_FLOAT_TYPE = {
    "f": np.float32,
    "d": np.float64,
    "e": np.float32,
    "g": np.float64,
    "b": np.float32,
    "B": np.float32,
    "h": np.float32,
    "H": np.float32,
    "?": np.float32,
}


def supported_float_type(dtype):
    """``_shared/utils.py:1151-1179`` (real dtypes only)."""
    dtype = np.dtype(dtype)
    if dtype.kind == "c":
        raise ValueError("complex valued input is not supported")
    return np.dtype(_NEW_FLOAT_TYPE.get(dtype.char, np.float64))


def weights_dtype(input_dtype, w_dtype):
    """``_ndimage_util.py:35-46`` with ``use_cucim_casting=True``."""
    input_dtype = np.dtype(input_dtype)
    w_dtype = np.dtype(w_dtype)
    if w_dtype.kind in "iub":
        return supported_float_type(w_dtype)
    return np.promote_types(input_dtype, np.float32)


def cast_to(values, dtype):
    """Convert accumulator values to the output dtype like ``cast<Y>``.

    ``_ndimage_filters_core.py:227-244``: a C cast, except float -> unsigned,
    which is ``a >= 0 ? (B)a : -(B)(-a)``.  Float -> integer C casts truncate
    toward zero; on the GPU the conversion saturates to the destination range
    and maps NaN to 0, which is what is reproduced here.  For unsigned outputs
    the negative branch negates modulo 2**bits.
    """
    dtype = np.dtype(dtype)
    v = np.asarray(values)
    if dtype.kind == "f":
        return v.astype(dtype)
    if dtype.kind == "b":
        return v != 0
    vf = v.astype(np.float64)
    info = np.iinfo(dtype)
    bits = info.bits

    def sat_trunc(a, lo, hi):
        a = np.where(np.isnan(a), 0.0, a)
        a = np.trunc(a)
        a = np.clip(a, lo, hi)
        return a

    if dtype.kind == "i":
        if bits == 64:
            # float64 cannot hold 2**63-1; clip below it then convert exactly.
            a = np.where(np.isnan(vf), 0.0, np.trunc(vf))
            out = np.empty(a.shape, dtype=np.int64)
            hi = a >= 9.223372036854775807e18
            lo = a <= -9.223372036854775808e18
            mid = ~(hi | lo)
            out[mid] = a[mid].astype(np.int64)
            out[hi] = np.iinfo(np.int64).max
            out[lo] = np.iinfo(np.int64).min
            return out
        return sat_trunc(vf, info.min, info.max).astype(dtype)
    # unsigned
    pos = vf >= 0
    if bits == 64:
        a = np.where(np.isnan(vf), 0.0, np.trunc(np.abs(vf)))
        out = np.empty(a.shape, dtype=np.uint64)
        big = a >= 1.8446744073709552e19
        out[big] = np.iinfo(np.uint64).max
        out[~big] = a[~big].astype(np.uint64)
        neg = ~pos & ~np.isnan(vf)
        out[neg] = (np.uint64(0) - out[neg]).astype(np.uint64)
        return out
    mag = sat_trunc(np.abs(vf), 0, info.max).astype(np.uint64)
    neg = ~pos & ~np.isnan(vf)
    mag = np.where(neg, (np.uint64(1 << bits) - mag) % np.uint64(1 << bits), mag)
    return mag.astype(dtype)


def fma_f32(a, b, c):
    """Exactly rounded float32 fused multiply-add, elementwise.

    ``a*b`` is exact in float64 (24+24 bit significands).  The float64 sum
    ``s = p + c`` is made exact with Knuth's TwoSum error ``e``; ``s`` rounded
    to float32 is correct unless ``s`` sits exactly on a float32 rounding
    boundary, where the sign of ``e`` decides.
    """
    a = np.asarray(a, dtype=np.float32).astype(np.float64)
    b = np.asarray(b, dtype=np.float32).astype(np.float64)
    c = np.asarray(c, dtype=np.float32).astype(np.float64)
    p = a * b
    with np.errstate(invalid="ignore", over="ignore"):
        s = p + c
        bb = s - p
        e = (p - (s - bb)) + (c - bb)  # exact: p + c == s + e
        r = s.astype(np.float32)
        rd = r.astype(np.float64)
        # float32 midpoints are float64 numbers, so RN32(s + e) differs from
        # RN32(s) only when s is exactly a midpoint and e != 0.
        above = s > rd
        other = np.where(above, np.nextafter(r, np.float32(np.inf)),
                         np.nextafter(r, np.float32(-np.inf)))
        od = other.astype(np.float64)
        mid = np.isfinite(s) & np.isfinite(od) & (s != rd) & ((s - rd) * 2 == (od - rd))
        # at a midpoint r and other straddle s; pick the side e points to
        go_other = mid & (((e > 0) & (od > rd)) | ((e < 0) & (od < rd)))
        r = np.where(go_other, other, r)
    return r.astype(np.float32)


def _as_acc(x, acc_dtype):
    x = np.asarray(x)
    if x.dtype == np.bool_:
        return x.astype(acc_dtype)
    return x.astype(acc_dtype)


def correlate1d(
    input,
    weights,
    axis=-1,
    output=None,
    mode="reflect",
    cval=0.0,
    origin=0,
    *,
    acc="f64",
):
    """One-dimensional correlation along ``axis``.

    Restates the elementwise kernel ``_ndimage_filters_core.py:247-408`` for a
    1-D weight vector reshaped along ``axis`` (``_convert_1d_args``
    ``_ndimage_filters_core.py:61-70``):
        offset = K//2 + origin               (``_origins_to_offsets`` :20-21)
        y[i]   = sum_{k : w[k] != 0} w[k] * x~[i + k - offset]
    where ``x~`` is the boundary extension (``remap_index``) and constant mode
    substitutes ``cast<W>(cval)`` (:334-349).  ``output`` is None (input
    dtype) or a dtype (``_ndimage_util.py:49-71``).
    """
    x = np.asarray(input)
    w = np.asarray(weights)
    if w.ndim != 1 or w.size < 1:
        raise RuntimeError("incorrect filter size")
    K = w.size
    if (K // 2 + origin < 0) or (K // 2 + origin >= K):
        raise ValueError("invalid origin")
    axis = axis % x.ndim
    mode = canonical_mode(mode)
    out_dtype = x.dtype if output is None else np.dtype(output)
    if acc == "f64":
        W = np.float64
        wv = w.astype(np.float64)
    else:
        W = weights_dtype(x.dtype, w.dtype).type
        wv = w.astype(W)
    n = x.shape[axis]
    xa = np.moveaxis(_as_acc(x, W), axis, -1)
    if n == 0 or x.size == 0:
        return np.zeros(x.shape, out_dtype)
    offset = K // 2 + origin
    pos = np.arange(n, dtype=np.int64)
    s = np.zeros(xa.shape, dtype=W)
    cv = W(cval)
    for k in range(K):
        wk = wv[k]
        if wk == 0:
            continue  # _ndimage_filters_core.py:308-309 nonzero(wval)
        idx = remap_index(pos + k - offset, n, mode)
        if mode == "constant":
            valid = idx >= 0
            vals = np.where(valid, xa[..., np.where(valid, idx, 0)], cv)
        else:
            vals = xa[..., idx]
        if acc == "emulate" and W == np.float32:
            s = fma_f32(vals, np.float32(wk), s)
        else:
            s = s + vals * wk
    y = np.moveaxis(s, -1, axis)
    return cast_to(y, out_dtype)


def convolve1d(input, weights, axis=-1, output=None, mode="reflect", cval=0.0,
               origin=0, *, acc="f64"):
    """Convolution = reversed weights with origin -> -origin - (K even).

    ``_ndimage_filters.py:236-243``.
    """
    w = np.asarray(weights)[::-1]
    origin = -origin
    if w.size % 2 == 0:
        origin -= 1
    return correlate1d(input, w, axis, output, mode, cval, origin, acc=acc)


def run_1d_filters(filters, input, axes, args, output, modes, cval, origins,
                   out_dtype):
    """Sequential 1-D passes; ``_ndimage_filters_core.py:124-163``.

    Each pass writes an array of the output dtype; axes whose filter is None
    are skipped; no filter at all returns a copy in the output dtype.
    """
    x = np.asarray(input)
    cur = x
    n = 0
    for fltr, axis, arg, mode, origin in zip(filters, axes, args, modes, origins):
        if fltr is None:
            continue
        cur = fltr(cur, arg, axis, out_dtype, mode, cval, origin)
        n += 1
    if n == 0:
        return cast_to(x, out_dtype) if np.dtype(out_dtype).kind != "f" else x.astype(out_dtype)
    return cur


def gaussian_kernel1d(sigma, order, radius):
    """1-D Gaussian correlation kernel, ``_ndimage_filters.py:689-717``.

    phi(x) = exp(-x^2 / (2 sigma^2)) / sum, x in [-radius, radius]; order n > 0
    applies ``Q = D + P`` n times to the polynomial q and returns
    ``(q(x) * phi(x))[::-1]`` (the reversal makes it a correlation kernel).
    Computed in float64.
    """
    if order < 0:
        raise ValueError("order must be non-negative")
    sigma2 = sigma * sigma
    x = np.arange(-radius, radius + 1)
    phi_x = np.exp(-0.5 / sigma2 * x**2)
    phi_x /= phi_x.sum()
    if order == 0:
        return phi_x
    exponent_range = np.arange(order + 1)
    q = np.zeros(order + 1)
    q[0] = 1
    D = np.diag(exponent_range[1:], 1)
    P = np.diag(np.ones(order) / -sigma2, -1)
    Q_deriv = D + P
    for _ in range(order):
        q = Q_deriv.dot(q)
    q = (x[:, None] ** exponent_range).dot(q)
    return (q * phi_x)[::-1].copy()


def _fix_sequence_arg(arg, ndim, name, conv=lambda v: v):
    """``_ndimage_util.py:74-85``."""
    if isinstance(arg, str):
        return [conv(arg)] * ndim
    try:
        it = iter(arg)
    except TypeError:
        return [conv(arg)] * ndim
    lst = [conv(v) for v in it]
    if len(lst) != ndim:
        raise RuntimeError(f"{name} must have length equal to input rank")
    return lst


def _pass(acc):
    def f(cur, w, axis, out_dtype, mode, cval, origin):
        return correlate1d(cur, w, axis, out_dtype, mode, cval, origin, acc=acc)
    return f


def gaussian_filter1d(input, sigma, axis=-1, order=0, output=None,
                      mode="reflect", cval=0.0, truncate=4.0, *, acc="f64"):
    """``_ndimage_filters.py:524-577``: radius = int(truncate*sigma + 0.5)."""
    x = np.asarray(input)
    radius = int(float(truncate) * float(sigma) + 0.5)
    w = gaussian_kernel1d(sigma, int(order), radius)
    if acc != "f64":
        w = w.astype(np.promote_types(x.dtype, np.float32))
    return correlate1d(x, w, axis, output, mode, cval, acc=acc)


def gaussian_filter(input, sigma, order=0, output=None, mode="reflect",
                    cval=0.0, truncate=4.0, radius=None, axes=None, *,
                    acc="f64"):
    """``_ndimage_filters.py:580-686``.

    Axes with sigma < 1e-15 are dropped; radius = int(truncate*sigma + 0.5)
    unless given; a radius <= 0 skips the axis; passes run in axis order.
    """
    x = np.asarray(input)
    out_dtype = x.dtype if output is None else np.dtype(output)
    axes = tuple(range(x.ndim)) if axes is None else tuple(a % x.ndim for a in axes)
    num_axes = len(axes)
    sigmas = _fix_sequence_arg(sigma, num_axes, "sigma", float)
    if num_axes == 0 or all(s < 1e-15 for s in sigmas):
        return x.astype(out_dtype) if out_dtype.kind == "f" else cast_to(x, out_dtype)
    orders = _fix_sequence_arg(order, num_axes, "order", int)
    modes = _fix_sequence_arg(mode, num_axes, "mode", str)
    radii = _fix_sequence_arg(radius, num_axes, "radius")
    truncate = float(truncate)
    wdt = np.promote_types(x.dtype, np.float32)
    params = [(axes[i], sigmas[i], orders[i], modes[i], radii[i])
              for i in range(num_axes) if sigmas[i] > 1e-15]
    ws = []
    for _, s, o, _, r in params:
        if r is None:
            r = int(truncate * float(s) + 0.5)
        if r <= 0:
            ws.append(None)
            continue
        w = gaussian_kernel1d(s, o, r)
        ws.append(w if acc == "f64" else w.astype(wdt))
    filters = [None if w is None else _pass(acc) for w in ws]
    return run_1d_filters(filters, x, [p[0] for p in params], ws, None,
                          [p[3] for p in params], cval, [0] * len(params),
                          out_dtype)


def _prewitt_or_sobel(x, axis, output, mode, cval, smooth, acc):
    """``_ndimage_filters.py:785-803``: [-1,0,1] along ``axis``, ``smooth``
    along every other axis, each pass in ascending axis order."""
    x = np.asarray(x)
    axis = axis % x.ndim
    out_dtype = x.dtype if output is None else np.dtype(output)
    wdt = np.promote_types(x.dtype, np.float32)
    diff = np.array([-1, 0, 1], dtype=np.float64)
    ws = [diff if a == axis else np.asarray(smooth, np.float64) for a in range(x.ndim)]
    if acc != "f64":
        ws = [w.astype(wdt) for w in ws]
    return run_1d_filters([_pass(acc)] * x.ndim, x, list(range(x.ndim)), ws,
                          None, [mode] * x.ndim, cval, [0] * x.ndim, out_dtype)


def sobel(input, axis=-1, output=None, mode="reflect", cval=0.0, *, acc="f64"):
    """``_ndimage_filters.py:756-782``: smoothing weights [1, 2, 1]."""
    return _prewitt_or_sobel(input, axis, output, mode, cval, [1, 2, 1], acc)


def prewitt(input, axis=-1, output=None, mode="reflect", cval=0.0, *, acc="f64"):
    """``_ndimage_filters.py:720-753``: smoothing weights [1, 1, 1]."""
    return _prewitt_or_sobel(input, axis, output, mode, cval, [1, 1, 1], acc)


def uniform_filter1d(input, size, axis=-1, output=None, mode="reflect",
                     cval=0.0, origin=0, *, acc="f64"):
    """``_ndimage_filters.py:399-444``: weights full(size, 1/size)."""
    x = np.asarray(input)
    wdt = np.promote_types(x.dtype, np.float32)
    w = np.full(size, 1 / size, dtype=np.float64 if acc == "f64" else wdt)
    return correlate1d(x, w, axis, output, mode, cval, origin, acc=acc)


def uniform_filter(input, size=3, output=None, mode="reflect", cval=0.0,
                   origin=0, axes=None, *, acc="f64"):
    """``_ndimage_filters.py:447-521``: sizes <= 1 skip the axis."""
    x = np.asarray(input)
    out_dtype = x.dtype if output is None else np.dtype(output)
    axes = tuple(range(x.ndim)) if axes is None else tuple(a % x.ndim for a in axes)
    sizes = _fix_sequence_arg(size, len(axes), "size", int)
    origins = _fix_sequence_arg(origin, len(axes), "origin", int)
    modes = _fix_sequence_arg(mode, len(axes), "mode", str)
    wdt = np.promote_types(x.dtype, np.float32)
    ws = [None if s <= 1 else np.full(s, 1 / s, dtype=np.float64 if acc == "f64" else wdt)
          for s in sizes]
    filters = [None if w is None else _pass(acc) for w in ws]
    return run_1d_filters(filters, x, list(axes), ws, None, modes, cval,
                          origins, out_dtype)


def correlate(input, weights, output=None, mode="reflect", cval=0.0, origin=0,
              *, acc="f64"):
    """n-D correlation with a full weights array.

    ``_ndimage_filters.py:224-257`` + ``_ndimage_filters_core.py:247-408``:
    taps are visited in C order of the weights array, zero weights skipped,
    per-axis offsets ``w_shape//2 + origin``.
    """
    x = np.asarray(input)
    w = np.asarray(weights)
    if w.ndim != x.ndim:
        raise RuntimeError("filter weights array has incorrect shape")
    out_dtype = x.dtype if output is None else np.dtype(output)
    origins = _fix_sequence_arg(origin, x.ndim, "origin", int)
    modes = _fix_sequence_arg(mode, x.ndim, "mode", str)
    for o, k in zip(origins, w.shape):
        if (k // 2 + o < 0) or (k // 2 + o >= k):
            raise ValueError("invalid origin")
    if acc == "f64":
        W = np.float64
    else:
        W = weights_dtype(x.dtype, w.dtype).type
    wv = w.astype(W)
    xa = _as_acc(x, W)
    s = np.zeros(x.shape, dtype=W)
    grids = np.meshgrid(*[np.arange(n) for n in x.shape], indexing="ij")
    all_const = all(canonical_mode(m) == "constant" for m in modes)
    for tap in np.ndindex(*w.shape):
        wk = wv[tap]
        if wk == 0:
            continue
        idx = []
        valid = np.ones(x.shape, bool)
        for d in range(x.ndim):
            off = w.shape[d] // 2 + origins[d]
            ii = remap_index(grids[d] + tap[d] - off, x.shape[d], modes[d])
            valid &= ii >= 0
            idx.append(np.where(ii >= 0, ii, 0))
        vals = xa[tuple(idx)]
        if all_const:
            vals = np.where(valid, vals, W(cval))
        if acc == "emulate" and W == np.float32:
            s = fma_f32(vals, np.float32(wk), s)
        else:
            s = s + vals * wk
    return cast_to(s, out_dtype)


def _wrap_or_round(a, b, op):
    """``output += tmp`` / ``output *= output`` of CuPy's in-place ufuncs on
    arrays of one dtype: one rounding per element for floats, wrap-around
    modulo 2**bits for integers (numpy's integer array ops wrap too)."""
    with np.errstate(over="ignore", invalid="ignore"):
        r = a + b if op == "add" else a * b
    return r.astype(a.dtype)


def _sqrt_unsafe(a):
    """``cupy.sqrt(output, output, casting='unsafe')``
    (``_ndimage_filters.py:1097``): floats in place; integers through the
    float type the ufunc picks (float16 for 8-bit, float32 for 16-bit,
    float64 wider), truncated back, with the NaN of a negative value stored
    as 0 (the device conversion)."""
    a = np.asarray(a)
    if a.dtype.kind == "f":
        with np.errstate(invalid="ignore"):
            return np.sqrt(a)
    F = {1: np.float16, 2: np.float32}.get(a.dtype.itemsize, np.float64)
    with np.errstate(invalid="ignore"):
        r = np.sqrt(a.astype(F))
    r = np.where(np.isnan(r), 0, r)
    return np.trunc(r).astype(a.dtype)


def _axes(axes, ndim):
    if axes is None:
        return tuple(range(ndim))
    if np.isscalar(axes):
        axes = (axes,)
    return tuple(int(a) % ndim for a in axes)


def generic_laplace_arrays(derivs):
    """``_ndimage_filters.py:806-890``: the first axis's derivative is the
    output, the others are added in, in axis order."""
    out = derivs[0]
    for t in derivs[1:]:
        out = _wrap_or_round(out, t, "add")
    return out


def generic_gradient_magnitude_arrays(derivs):
    """``_ndimage_filters.py:1012-1097``: output *= output, then per further
    axis tmp *= tmp; output += tmp; finally the unsafe in-place sqrt."""
    out = _wrap_or_round(derivs[0], derivs[0], "mul")
    for t in derivs[1:]:
        out = _wrap_or_round(out, _wrap_or_round(t, t, "mul"), "add")
    return _sqrt_unsafe(out)


def laplace(input, output=None, mode="reflect", cval=0.0, *, axes=None, acc="f64"):
    """``_ndimage_filters.py:893-930``: correlate1d with [1, -2, 1] (weights of
    dtype promote(input, float32)) along each axis, summed in the output
    dtype."""
    x = np.asarray(input)
    out_dtype = x.dtype if output is None else np.dtype(output)
    ax = _axes(axes, x.ndim)
    if len(ax) == 0:
        return cast_to(x, out_dtype) if out_dtype.kind != "f" else x.astype(out_dtype)
    modes = _fix_sequence_arg(mode, len(ax), "mode", str)
    wdt = np.float64 if acc == "f64" else np.promote_types(x.dtype, np.float32)
    w = np.array([1, -2, 1], dtype=wdt)
    derivs = [correlate1d(x, w, a, out_dtype, m, cval, acc=acc) for a, m in zip(ax, modes)]
    return generic_laplace_arrays(derivs)


def gaussian_laplace(input, sigma, output=None, mode="reflect", cval=0.0, *, axes=None,
                     truncate=4.0, acc="f64"):
    """``_ndimage_filters.py:933-1009``: per axis a, gaussian_filter with
    order 2 along a and 0 elsewhere (sigma on the listed axes, 0 on the
    others when ``axes`` is a subset), summed in the output dtype."""
    x = np.asarray(input)
    out_dtype = x.dtype if output is None else np.dtype(output)
    ax = _axes(axes, x.ndim)
    if len(ax) == 0:
        return cast_to(x, out_dtype) if out_dtype.kind != "f" else x.astype(out_dtype)
    modes = _fix_sequence_arg(mode, len(ax), "mode", str)
    sig = _fix_sequence_arg(sigma, len(ax), "sigma", float)
    if len(ax) < x.ndim:
        full = [0.0] * x.ndim
        for s, a in zip(sig, ax):
            full[a] = s
        sig = full
    derivs = []
    for a, m in zip(ax, modes):
        order = [0] * x.ndim
        order[a] = 2
        derivs.append(gaussian_filter(x, sig, order, out_dtype, m, cval, truncate, acc=acc))
    return generic_laplace_arrays(derivs)


def gaussian_gradient_magnitude(input, sigma, output=None, mode="reflect", cval=0.0, *,
                                axes=None, truncate=4.0, acc="f64"):
    """``_ndimage_filters.py:1100-1157``: per axis a, gaussian_filter with
    order 1 along a (sigma as given, over every axis), combined by
    generic_gradient_magnitude."""
    x = np.asarray(input)
    out_dtype = x.dtype if output is None else np.dtype(output)
    ax = _axes(axes, x.ndim)
    if x.ndim == 0:
        return x.astype(out_dtype)
    modes = _fix_sequence_arg(mode, len(ax), "mode", str)
    derivs = []
    for a, m in zip(ax, modes):
        order = [0] * x.ndim
        order[a] = 1
        derivs.append(gaussian_filter(x, sigma, order, out_dtype, m, cval, truncate, acc=acc))
    return generic_gradient_magnitude_arrays(derivs)


def convolve(input, weights, output=None, mode="reflect", cval=0.0, origin=0, *,
             acc="f64"):
    """n-D convolution (``_ndimage_filters.py:85-132, 236-243``): the weights
    reversed along every axis and, per axis, origin -> -origin, minus 1 more
    for an even-sized axis; then ``correlate``."""
    x = np.asarray(input)
    w = np.asarray(weights)
    origins = _fix_sequence_arg(origin, x.ndim, "origin", int)
    wr = w[(slice(None, None, -1),) * w.ndim]
    o2 = [-o - (1 if s % 2 == 0 else 0) for o, s in zip(origins, w.shape)]
    return correlate(x, wr, output, mode, cval, o2, acc=acc)
