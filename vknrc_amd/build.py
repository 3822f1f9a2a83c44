"""Builds vknrc_amd/libvknrc.so (the C-ABI library of include/nrc.h) with hipcc for gfx950.

Each source compiles to its own object in parallel, then hipcc links the shared
library.

    python -m vknrc_amd.build            # or vknrc_amd.build.build()
"""
from __future__ import annotations

import os
import subprocess
import sys
import tempfile
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libvknrc.so")
SOURCES = ["nrc_api.cpp", "nrc_plain.hip", "nrc_fused.hip", "nrc_fp8.hip", "nrc_optim.hip", "nrc_records.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    # no packed FP32 instructions in device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # MFMA results in VGPRs when they fit, so the fp16 pack reads them without
    # v_accvgpr_read copies
    "-mllvm", "-amdgpu-mfma-vgpr-form=1",
    "-Wno-unused-command-line-argument",
]


def _run(cmd):
    res = subprocess.run(cmd, capture_output=True, text=True)
    out = "\n".join(l for l in (res.stdout + res.stderr).splitlines() if "packed-fp32-ops" not in l)
    return res.returncode, out


def build(verbose: bool = False) -> str:
    """Compile every HIP source for gfx950 into one shared library; returns its path."""
    inc = ["-I", os.path.join(ROOT, "include"), "-I", CSRC]
    with tempfile.TemporaryDirectory(prefix="vknrc_build_") as tmp:
        jobs = []
        for src in SOURCES:
            obj = os.path.join(tmp, os.path.splitext(src)[0] + ".o")
            jobs.append((obj, [HIPCC, *FLAGS, *inc, "-c", os.path.join(CSRC, src), "-o", obj]))
        with ThreadPoolExecutor(max_workers=len(jobs)) as ex:
            results = list(ex.map(lambda j: _run(j[1]), jobs))
        for (obj, cmd), (rc, out) in zip(jobs, results):
            if verbose:
                print(" ".join(cmd), flush=True)
            if rc != 0:
                raise RuntimeError(f"hipcc failed on {cmd[-3]} ({rc}):\n{out}")
            if verbose and out.strip():
                print(out, file=sys.stderr)
        tmp_lib = LIB + ".tmp"
        rc, out = _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", *[o for o, _ in jobs], "-o", tmp_lib])
        if rc != 0:
            raise RuntimeError(f"link failed ({rc}):\n{out}")
        os.replace(tmp_lib, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
