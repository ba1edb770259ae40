"""Builds lib_et_array_interface.so in-tree with hipcc for gfx950.

    python -m niftyrec_amd.build          # or: python niftyrec_amd/build.py

Every HIP source is compiled with packed-FP32 instruction selection disabled
(-Xclang -target-feature -Xclang -packed-fp32-ops); the host layer keeps the
reference's single-precision rotation arithmetic by compiling without
floating-point contraction.
"""
import os
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB_NAME = "lib_et_array_interface.so"
LIB_PATH = os.path.join(HERE, LIB_NAME)
ARCH = os.environ.get("ET_OFFLOAD_ARCH", "gfx950")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
COMMON = ["-O3", "-fPIC", "-std=c++17", "-I" + CSRC, "-I" + os.path.join(ROOT, "include"),
          "--offload-arch=" + ARCH, "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-Wno-unused-command-line-argument"]
SOURCES = [("et_kernels.hip", ["-ffp-contract=fast"]),
           ("et_host.cpp", ["-x", "hip", "-ffp-contract=off"])]


def build(verbose=True):
    """Compile every HIP source and link the shared library (always rebuilds)."""
    with tempfile.TemporaryDirectory(prefix="et_build_") as tmp:
        objs = []
        for src, extra in SOURCES:
            obj = os.path.join(tmp, os.path.splitext(src)[0] + ".o")
            cmd = [HIPCC] + COMMON + extra + ["-c", os.path.join(CSRC, src), "-o", obj]
            if verbose:
                print("[build]", " ".join(cmd), flush=True)
            r = subprocess.run(cmd, capture_output=True, text=True)
            # the host pass does not know the packed-fp32 feature: drop that warning
            err = "\n".join(l for l in r.stderr.splitlines() if "packed-fp32-ops" not in l)
            if r.returncode != 0:
                sys.stderr.write(err + "\n")
                raise RuntimeError("hipcc failed on " + src)
            if err.strip() and verbose:
                sys.stderr.write(err + "\n")
            objs.append(obj)
        out_tmp = LIB_PATH + ".tmp"
        cmd = [HIPCC, "-shared", "-fPIC", "--offload-arch=" + ARCH, "-o", out_tmp] + objs
        if verbose:
            print("[build]", " ".join(cmd), flush=True)
        subprocess.check_call(cmd)
        os.replace(out_tmp, LIB_PATH)
    return LIB_PATH


if __name__ == "__main__":
    build()
