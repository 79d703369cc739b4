"""roctx ranges: off by default, balanced push/pop, transparent to results
and exceptions (CPU; the roctx library is loaded when present)."""

import pytest

from cucim_amd import _trace


def test_annotate_passes_results_and_errors():
    calls = []

    @_trace.annotate("unit")
    def f(a, b=1):
        calls.append((a, b))
        if a < 0:
            raise ValueError("neg")
        return a + b

    was = _trace.enabled()
    try:
        for on in (False, True):
            _trace.enable(on)
            assert f(2, b=3) == 5
            with pytest.raises(ValueError):
                f(-1)
            with _trace.range_("block"):
                assert f(1) == 2
    finally:
        _trace.enable(was)
    assert calls == [(2, 3), (-1, 1), (1, 1)] * 2
    assert f.__name__ == "f"


def test_public_entry_points_are_annotated():
    from cucim_amd.skimage.feature import hessian_matrix
    from cucim_amd.skimage.filters import frangi, threshold_otsu
    from cucim_amd.skimage.measure import label

    for fn in (frangi, hessian_matrix, threshold_otsu, label):
        assert hasattr(fn, "__wrapped__"), fn
