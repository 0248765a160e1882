"""Build recipe for the native parts of henon_map_cpp_amd (in-tree, no JIT cache).

  libhenon_hip.so                 hipcc, gfx950: every HIP kernel + the C ABI
                                  declared in include/henon_hip.h
  henon_map_engine.<abi>.so       g++, pybind11: the drop-in for the reference's
                                  henon_map_engine module, calling only the C ABI

Run ``python -m henon_map_cpp_amd.build`` (or ``__graft_entry__.build()``).
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys
import sysconfig

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
INCLUDE = os.path.join(ROOT, "include")
BUILD = os.path.join(PKG, "_build")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17",
    # exact mode relies on no contraction; the fast kernels place fma() by hand
    "-ffp-contract=off",
    # no packed FP32 instructions in any HIP object
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
    f"-I{INCLUDE}", f"-I{CSRC}",
]
HIP_SOURCES = ["hm_runtime.hip", "hm_track.hip", "hm_tangent.hip", "hm_fft.hip", "hm_adjoint.hip", "hm_batch.hip",
               "hm_ghost.hip", "hm_kick.hip", "hm_chirpz.hip", "hm_fft_mixed.hip"]

LIB_NAME = "libhenon_hip.so"
ENGINE_SOURCES = ["engine.cpp"]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l and "hip-link" not in l)
    if r.returncode != 0:
        raise RuntimeError(f"command failed ({r.returncode}): {' '.join(cmd)}\n{out}")
    return out


def build_hip(jobs: int = 8) -> str:
    os.makedirs(BUILD, exist_ok=True)
    objs = []
    cmds = []
    for src in HIP_SOURCES:
        obj = os.path.join(BUILD, src.replace(".hip", ".o"))
        objs.append(obj)
        cmds.append([HIPCC, *HIP_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj])
    with cf.ThreadPoolExecutor(max_workers=max(1, min(jobs, len(cmds)))) as ex:
        list(ex.map(_run, cmds))
    lib = os.path.join(PKG, LIB_NAME)
    tmp = lib + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs,
          f"-Wl,-soname,{LIB_NAME}"])
    os.replace(tmp, lib)
    return lib


def build_engine() -> str | None:
    srcs = [os.path.join(CSRC, s) for s in ENGINE_SOURCES if os.path.exists(os.path.join(CSRC, s))]
    if not srcs:
        return None
    import pybind11

    suffix = sysconfig.get_config_var("EXT_SUFFIX")
    out = os.path.join(PKG, "henon_map_engine" + suffix)
    tmp = out + ".tmp"
    py_inc = sysconfig.get_paths()["include"]
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-fvisibility=hidden",
          f"-I{INCLUDE}", f"-I{pybind11.get_include()}", f"-I{py_inc}", *srcs,
          f"-L{PKG}", "-lhenon_hip", "-Wl,-rpath,$ORIGIN", "-o", tmp])
    os.replace(tmp, out)
    return out


def build_all() -> None:
    build_hip()
    build_engine()


if __name__ == "__main__":
    build_all()
    print("built", LIB_NAME, file=sys.stderr)
