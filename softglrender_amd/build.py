"""Builds librhip (gfx950) in-tree: softglrender_amd/lib/librhip.so.

Every HIP translation unit is compiled with -ffp-contract=off (fused
multiply-adds are written out where the reference's arithmetic has them) and
with packed-FP32 instructions disabled.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB_DIR = os.path.join(HERE, "lib")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

HIP_FLAGS = [f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-I", os.path.join(ROOT, "include"), "-I", CSRC]

TARGETS = {
    "librhip.so": ["rhip_api.hip", "rhip_batch.hip", "rhip_batch_vis1.hip", "rhip_batch_vis4.hip", "rhip_libm.hip"],
}
OBJ_DIR = os.path.join(LIB_DIR, "obj")

HOST = os.path.join(HERE, "host")
# RendererHIP (the C++ Renderer implementation) and its stream driver, on librhip
HOST_FLAGS = ["-std=c++17", "-O2", "-fPIC", "-shared", "-Wall", "-Wno-unused-function",
              "-I", os.path.join(HOST, "compat"), "-I", HOST, "-I", os.path.join(ROOT, "include")]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"build failed: {' '.join(cmd[:3])} ...")


def build(verbose: bool = False) -> None:
    os.makedirs(OBJ_DIR, exist_ok=True)
    for lib, srcs in TARGETS.items():
        out = os.path.join(LIB_DIR, lib)
        tmp = out + ".tmp"
        # translation units compile in parallel, then link
        objs, procs = [], []
        for src in srcs:
            obj = os.path.join(OBJ_DIR, os.path.splitext(src)[0] + ".o")
            cmd = [HIPCC, *HIP_FLAGS, "-c", "-o", obj, os.path.join(CSRC, src)]
            if verbose:
                print(" ".join(cmd), flush=True)
            procs.append((cmd, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)))
            objs.append(obj)
        failed = None
        for cmd, pr in procs:
            txt, _ = pr.communicate()
            if pr.returncode != 0:
                sys.stderr.write(txt)
                failed = failed or cmd
        if failed:
            raise RuntimeError(f"build failed: {' '.join(failed[-2:])}")
        _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs])
        os.replace(tmp, out)
    out = os.path.join(LIB_DIR, "librenderer_hip.so")
    cmd = ["g++", *HOST_FLAGS, "-o", out + ".tmp", os.path.join(HOST, "RendererHIP.cpp"), os.path.join(HOST, "replay.cpp"),
           "-L", LIB_DIR, "-lrhip", "-Wl,-rpath,$ORIGIN"]
    if verbose:
        print(" ".join(cmd), flush=True)
    _run(cmd)
    os.replace(out + ".tmp", out)


if __name__ == "__main__":
    build(verbose=True)
