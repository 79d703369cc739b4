"""The C ABI's host code under AddressSanitizer and UndefinedBehaviorSanitizer.

tests/host_selftest.cpp is linked against a host-only, sanitized build of
every library source (cucim_amd._build.build_host_sanitized) and run without a
GPU: Gaussian tap construction, batch planning, knob and tracing state and the
argument checks of the entry points.  A sanitizer report or a failed check
ends the run with a nonzero status.  The taps and plans it prints are compared
with the production library's (same host code, optimised build).
"""

import ctypes
import os
import shutil
import subprocess

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))

pytestmark = pytest.mark.skipif(
    shutil.which("hipcc") is None and not os.path.exists("/opt/rocm/bin/hipcc"),
    reason="hipcc not available")


@pytest.fixture(scope="module")
def selftest_output(tmp_path_factory):
    from cucim_amd import _build

    out = tmp_path_factory.mktemp("sanitized")
    exe = _build.build_host_sanitized(os.path.join(HERE, "host_selftest.cpp"), str(out))
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:abort_on_error=0",
               UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([exe], capture_output=True, text=True, env=env, timeout=120)
    return r


def test_no_reports_and_checks_pass(selftest_output):
    r = selftest_output
    fails = [ln for ln in r.stdout.splitlines() if ln.startswith("FAIL")]
    assert r.returncode == 0, (fails, r.stderr[-4000:])
    assert "ERROR: AddressSanitizer" not in r.stderr
    assert "runtime error:" not in r.stderr
    assert r.stdout.rstrip().endswith("DONE 0")


def test_taps_match_production_library(selftest_output):
    from cucim_amd import _lib

    lib = _lib.load_library()
    lines = [ln.split() for ln in selftest_output.stdout.splitlines() if ln.startswith("K ")]
    assert len(lines) == 32
    for ln in lines:
        sigma, order, radius = float(ln[1]), int(ln[2]), int(ln[3])
        got = np.array([float.fromhex(v) for v in ln[4:]])
        exp = np.empty(2 * radius + 1)
        rc = lib.hc_gaussian_kernel1d(sigma, order, radius,
                                      exp.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
        assert rc == 0
        np.testing.assert_array_equal(got, exp)


def test_plans_match_production_library(selftest_output):
    from cucim_amd import _lib
    from cucim_amd.batch import HCTile2d

    lib = _lib.load_library()
    shapes = [(512, 512), (1, 1), (0, 7), (33, 1000), (1023, 17), (5, 0)]
    lines = [ln.split() for ln in selftest_output.stdout.splitlines() if ln.startswith("P ")]
    assert len(lines) == 8
    for ln in lines:
        K, ty, total = int(ln[1]), int(ln[2]), int(ln[3])
        tiles = (HCTile2d * len(shapes))()
        for i, (h, w) in enumerate(shapes):
            tiles[i].H, tiles[i].W = h, w
        assert lib.hc_batch2d_plan(tiles, len(shapes), K, ty, None, 0) == total
        assert [t.item0 for t in tiles] == [int(v) for v in ln[4:]]
