"""Compile the gfx950 library ``libcosmokdtree_amd.so`` in-tree with hipcc.

    python -m cosmokdtree_amd.build          # or: python cosmokdtree_amd/build.py
    python -m cosmokdtree_amd.build --debug

Every ``csrc/*.hip`` file is compiled to an object in ``_build/`` (in
parallel) and linked into ``cosmokdtree_amd/libcosmokdtree_amd.so``.
Always recompiles (the build check runs this on every call).

``--debug`` builds ``libcosmokdtree_amd_debug.so`` (objects in
``_build_debug/``): device code with the CKD_DEBUG bounds checks (a failed
check prints its condition and stops the kernel), host code under
AddressSanitizer and UndefinedBehaviorSanitizer (``-Xarch_host``, so no
device code is instrumented).  Load it with ``CKD_LIB`` and the clang ASan
runtime preloaded (``asan_runtime()``).
"""
from __future__ import annotations

import concurrent.futures as cf
import glob
import os
import subprocess
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libcosmokdtree_amd.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    # distances must round once per operation, as the reference source reads
    "-ffp-contract=off",
    # no packed FP32 instructions (v_pk_*_f32) anywhere in device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
]


DEBUG_FLAGS = ["-DCKD_DEBUG", "-g",
               "-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined",
               "-Xarch_host", "-fno-sanitize-recover=all", "-Xarch_host", "-fno-omit-frame-pointer"]
DEBUG_LIB = os.path.join(HERE, "libcosmokdtree_amd_debug.so")


def asan_runtime() -> str:
    """The clang AddressSanitizer runtime the debug library's host code needs
    (to be preloaded into the process that loads it)."""
    p = subprocess.run([HIPCC.replace("hipcc", "amdclang++"), "-print-file-name=libclang_rt.asan-x86_64.so"],
                       capture_output=True, text=True)
    path = p.stdout.strip()
    if p.returncode != 0 or not os.path.isabs(path) or not os.path.exists(path):
        raise RuntimeError("the clang AddressSanitizer runtime was not found")
    return path


def _compile(src: str, obj_dir: str = OBJ, extra=()) -> tuple[str, str]:
    obj = os.path.join(obj_dir, os.path.basename(src).replace(".hip", ".o"))
    cmd = [HIPCC, *FLAGS, *extra, "-c", src, "-o", obj]
    p = subprocess.run(cmd, capture_output=True, text=True)
    if p.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{p.stderr[-6000:]}")
    return obj, p.stderr


def build(verbose: bool = False, out: str | None = None, objdir: str | None = None, debug: bool = False) -> str:
    """Compile and link; `out`/`objdir` redirect the library and objects;
    debug builds the checked variant (module docstring)."""
    obj_dir = objdir or (OBJ + "_debug" if debug else OBJ)
    lib_path = out or (DEBUG_LIB if debug else LIB)
    extra = DEBUG_FLAGS if debug else []
    os.makedirs(obj_dir, exist_ok=True)
    srcs = sorted(glob.glob(os.path.join(CSRC, "*.hip")))
    t0 = time.time()
    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        results = list(ex.map(lambda f: _compile(f, obj_dir, extra), srcs))
    objs = [o for o, _ in results]
    if verbose:
        for (o, err) in results:
            errs = "\n".join(l for l in err.splitlines() if "packed-fp32-ops" not in l and l.strip())
            if errs:
                print(errs, file=sys.stderr)
    tmp = lib_path + ".tmp"
    link_extra = ["-Xarch_host", "-fsanitize=address,undefined", "-shared-libsan"] if debug else []
    p = subprocess.run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", *link_extra, *objs, "-o", tmp],
                       capture_output=True, text=True)
    if p.returncode != 0:
        raise RuntimeError(f"link failed:\n{p.stderr[-4000:]}")
    os.replace(tmp, lib_path)
    if verbose:
        print(f"built {lib_path} in {time.time() - t0:.1f}s", file=sys.stderr)
    return lib_path


if __name__ == "__main__":
    build(verbose=True, debug="--debug" in sys.argv[1:])
