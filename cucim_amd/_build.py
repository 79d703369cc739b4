"""Compile the HIP sources into ``cucim_amd/lib/libcucim_amd.so`` (gfx950).

Run ``python -m cucim_amd._build`` or call :func:`build`.
``python -m cucim_amd._build --sanitize-host DRIVER.cpp OUTDIR`` builds the
host code with AddressSanitizer and UndefinedBehaviorSanitizer instead
(:func:`build_host_sanitized`; tests/host_selftest.cpp is such a driver).  Objects go to
``_build/`` at the repository root; the shared library is built in-tree so
that it travels with the source tree.
"""

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
SRC = os.path.join(PKG, "csrc")
INC = os.path.join(ROOT, "include")
LIBDIR = os.path.join(PKG, "lib")
LIB = os.path.join(LIBDIR, "libcucim_amd.so")
OBJ = os.path.join(ROOT, "_build")

SOURCES = ["hc_host.hip", "k_correlate.hip", "k_hessian.hip", "k_pointwise.hip",
           "k_stencil2d.hip", "k_ridges.hip", "k_cnn3d.hip",
           "k_frangi_check.hip", "k_eig_nd.hip", "k_measure.hip",
           "hc_taps.hip", "hc_wsi.hip"]

# system libraries the library links (zlib: deflate-compressed TIFF tiles)
LIBS = ["-lz"]

ARCH = "gfx950"
# Packed FP32 instructions are disabled for every HIP translation unit.
HIPFLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    f"-I{INC}",
]


def _hipcc():
    cand = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(cand):
        raise RuntimeError("hipcc not found; the HIP library cannot be built")
    return cand


def _compile(src):
    obj = os.path.join(OBJ, os.path.splitext(src)[0] + ".o")
    cmd = [_hipcc(), *HIPFLAGS, "-c", os.path.join(SRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        msg = "\n".join(line for line in r.stderr.splitlines()
                        if "packed-fp32-ops" not in line)
        raise RuntimeError(f"hipcc failed for {src}:\n{msg}")
    return obj


def build(verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    os.makedirs(LIBDIR, exist_ok=True)
    with ThreadPoolExecutor(max_workers=len(SOURCES)) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [_hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, *LIBS, "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


SANITIZE = ["-fsanitize=address", "-fsanitize=undefined", "-fno-sanitize-recover=undefined"]


def build_host_sanitized(driver, outdir):
    """Host-only build of every source with AddressSanitizer and
    UndefinedBehaviorSanitizer, linked with ``driver`` (a C++ file with a
    main()) into ``outdir/<driver stem>``.  No device code is compiled: the
    executable runs the host logic of the C ABI (tap construction, planning,
    argument checks) on a machine without a GPU; a kernel launch from it
    fails with a HIP error.  Run it with ASAN_OPTIONS=detect_leaks=0 (the HIP
    runtime keeps allocations until exit)."""
    os.makedirs(outdir, exist_ok=True)
    host = [x for f in SANITIZE for x in ("-Xarch_host", f)]

    def one(src):
        obj = os.path.join(outdir, os.path.splitext(src)[0] + ".o")
        cmd = [_hipcc(), "-fgpu-rdc", "--offload-host-only", "-O1", "-g", "-std=c++17",
               "-fPIC", f"-I{INC}", *host, "-c", os.path.join(SRC, src), "-o", obj]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"hipcc (host, sanitized) failed for {src}:\n{r.stderr}")
        return obj

    with ThreadPoolExecutor(max_workers=len(SOURCES)) as ex:
        objs = list(ex.map(one, SOURCES))
    clang = os.path.join(os.path.dirname(os.path.realpath(_hipcc())), "..", "lib", "llvm", "bin",
                         "clang++")
    if not os.path.exists(clang):
        clang = "/opt/rocm/lib/llvm/bin/clang++"
    stem = os.path.splitext(os.path.basename(driver))[0]
    dobj = os.path.join(outdir, stem + ".o")
    r = subprocess.run([clang, "-O1", "-g", "-std=c++17", f"-I{INC}", *SANITIZE, "-c", driver,
                        "-o", dobj], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"clang++ failed for {driver}:\n{r.stderr}")
    exe = os.path.join(outdir, stem)
    r = subprocess.run([_hipcc(), "-fgpu-rdc", "--hip-link", f"--offload-arch={ARCH}",
                        "-fsanitize=address,undefined", *objs, dobj, *LIBS, "-o", exe],
                       capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link (sanitized) failed:\n{r.stderr}")
    return exe


if __name__ == "__main__":
    if len(sys.argv) == 4 and sys.argv[1] == "--sanitize-host":
        print(build_host_sanitized(sys.argv[2], sys.argv[3]))
        sys.exit(0)
    build(verbose=True)
    sys.exit(0)
