"""Builds the native libraries in-tree.

* ``method_amd/lib/libmethod_gpu.so`` -- HIP kernels and the C ABI of
  ``include/method_gpu.h``, compiled for gfx950 only.
* ``method_amd/lib/libmethod_host.so`` -- the C++ host API (Data, Simulation,
  SRMHD, FVS, RK2, ...) layered on the C ABI.

``python -m method_amd.build`` builds both.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "method_amd")
CSRC = os.path.join(PKG, "csrc")
HOST = os.path.join(PKG, "host")
LIBDIR = os.path.join(PKG, "lib")
INCLUDE = os.path.join(ROOT, "include")

GPU_LIB = os.path.join(LIBDIR, "libmethod_gpu.so")
HOST_LIB = os.path.join(LIBDIR, "libmethod_host.so")

GPU_SOURCES = ["mgpu_kernels.hip", "mgpu_engine.hip", "mgpu_mlp.hip", "mgpu_batch.hip", "mgpu_regime.hip",
               "mgpu_cells.hip"]
ARCH = "gfx950"
# Per-source extra flags. The MLP's MFMA accumulators stay in ordinary VGPRs,
# so the activations read them without a copy out of the accumulation
# registers.
SOURCE_FLAGS = {"mgpu_mlp.hip": ["-mllvm", "-amdgpu-mfma-vgpr-form"]}


def hipcc() -> str:
    for cand in (shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def _run(cmd, cwd=ROOT):
    r = subprocess.run(cmd, cwd=cwd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("build failed: " + " ".join(cmd) + "\n" + r.stdout[-4000:] + r.stderr[-8000:])
    return r


def gpu_flags():
    return [
        f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
        # no packed FP32 instructions in any HIP code
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        "-I", INCLUDE, "-I", CSRC,
    ]


def build_gpu() -> str:
    os.makedirs(LIBDIR, exist_ok=True)
    objs = [os.path.join(LIBDIR, src.replace(".hip", ".o")) for src in GPU_SOURCES]
    # one hipcc per source, run side by side (each is single-threaded)
    with ThreadPoolExecutor(max_workers=min(len(GPU_SOURCES), os.cpu_count() or 1, 8)) as pool:
        list(pool.map(lambda so: _run([hipcc()] + gpu_flags() + SOURCE_FLAGS.get(so[0], [])
                                      + ["-c", os.path.join(CSRC, so[0]), "-o", so[1]]),
                      zip(GPU_SOURCES, objs)))
    tmp = GPU_LIB + ".tmp"
    _run([hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs)
    os.replace(tmp, GPU_LIB)
    for o in objs:
        os.remove(o)
    return GPU_LIB


def build_host() -> str | None:
    srcs = sorted(os.path.join(HOST, f) for f in os.listdir(HOST) if f.endswith(".cc")) if os.path.isdir(HOST) else []
    if not srcs:
        return None
    os.makedirs(LIBDIR, exist_ok=True)
    tmp = HOST_LIB + ".tmp"
    _run(["g++", "-std=c++14", "-O2", "-fPIC", "-shared", "-Wall", "-I", INCLUDE, "-I", HOST, "-I", os.path.join(HOST, "include"), "-o", tmp] + srcs
         + ["-L", LIBDIR, "-lmethod_gpu", "-Wl,-rpath,$ORIGIN"])
    os.replace(tmp, HOST_LIB)
    return HOST_LIB


EXAMPLES = os.path.join(HOST, "examples")


def build_examples() -> list[str]:
    """Host-API drivers (written like the reference's example mains)."""
    out = []
    if not os.path.isdir(EXAMPLES):
        return out
    for f in sorted(os.listdir(EXAMPLES)):
        if not f.endswith(".cc"):
            continue
        exe = os.path.join(LIBDIR, f[:-3])
        _run(["g++", "-std=c++14", "-O2", "-Wall", "-I", INCLUDE, "-I", os.path.join(HOST, "include"),
              os.path.join(EXAMPLES, f), "-o", exe, "-L", LIBDIR, "-lmethod_host", "-lmethod_gpu",
              "-Wl,-rpath,$ORIGIN"])
        out.append(exe)
    return out


def build_all(verbose: bool = True) -> list[str]:
    out = [build_gpu()]
    h = build_host()
    if h:
        out.append(h)
        out += build_examples()
    if verbose:
        for p in out:
            print("built", p)
    return out


if __name__ == "__main__":
    build_all(verbose=True)
    sys.exit(0)
