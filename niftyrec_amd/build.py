"""Builds lib_et_array_interface.so in-tree with hipcc for gfx950.

    python -m niftyrec_amd.build          # or: python niftyrec_amd/build.py

Every HIP source is compiled with packed-FP32 instruction selection disabled
(-Xclang -target-feature -Xclang -packed-fp32-ops); the kernels contract
multiply-adds except where a pragma forbids it (-ffp-contract=fast-honor-pragmas),
and the host layer keeps the reference's single-precision rotation arithmetic
by compiling without floating-point contraction.
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
RMAX = 10  # largest PSF radius of the sweep kernels (ET_RMAX in et_internal.h)
# Kernels contract multiply-adds into FMAs except where a function body says
# `#pragma clang fp contract(off)` (sample positions, TwoSum): plain
# -ffp-contract=fast lets the AMDGPU backend fuse those as well, which moves a
# sample position by an ulp against the reference's rounding.
KERNEL_FP = ["-ffp-contract=fast-honor-pragmas"]
# (source, object name, extra flags); the sweep launchers are instantiated once
# per PSF radius in their own translation unit so that hipcc runs in parallel
SOURCES = ([("et_kernels.hip", "et_kernels", KERNEL_FP)] +
           [("et_sweep_r.hip", "et_sweep_r%d" % r, KERNEL_FP + ["-DET_SWEEP_R=%d" % r])
            for r in range(RMAX, -1, -1)] +
           [("et_host.cpp", "et_host", ["-x", "hip", "-ffp-contract=off"])])


def _jobs():
    try:
        n = int(os.environ.get("MAX_JOBS", "0"))
    except ValueError:
        n = 0
    return max(1, min(n or (os.cpu_count() or 1), 16))


def build(verbose=True):
    """Compile every HIP source (in parallel) and link the shared library
    (always rebuilds)."""
    from concurrent.futures import ThreadPoolExecutor
    with tempfile.TemporaryDirectory(prefix="et_build_") as tmp:
        def compile_one(item):
            src, name, extra = item
            obj = os.path.join(tmp, name + ".o")
            cmd = [HIPCC] + COMMON + extra + ["-c", os.path.join(CSRC, src), "-o", obj]
            r = subprocess.run(cmd, capture_output=True, text=True)
            # the host pass does not know the packed-fp32 feature: drop that warning
            err = "\n".join(l for l in r.stderr.splitlines() if "packed-fp32-ops" not in l)
            return cmd, obj, r.returncode, err

        objs = []
        with ThreadPoolExecutor(_jobs()) as ex:
            for cmd, obj, rc, err in ex.map(compile_one, SOURCES):
                if verbose:
                    print("[build]", " ".join(cmd), flush=True)
                if rc != 0:
                    sys.stderr.write(err + "\n")
                    raise RuntimeError("hipcc failed: " + " ".join(cmd))
                if err.strip() and verbose:
                    sys.stderr.write(err + "\n")
                objs.append(obj)
        out_tmp = LIB_PATH + ".tmp"
        cmd = [HIPCC, "-shared", "-fPIC", "--offload-arch=" + ARCH, "-o", out_tmp] + objs
        if verbose:
            print("[build]", " ".join(cmd), flush=True)
        subprocess.check_call(cmd)
        os.replace(out_tmp, LIB_PATH)
    build_tools(verbose)
    return LIB_PATH


APPS = os.path.join(HERE, "apps")
REC_SPECT = os.path.join(HERE, "rec_spect")


def build_tools(verbose=True):
    """The command line tool rec_spect (host C++, links the library in place:
    its run path is the directory it sits in)."""
    out_tmp = REC_SPECT + ".tmp"
    cmd = [os.environ.get("CXX", "g++"), "-O2", "-std=c++17", "-Wall", os.path.join(APPS, "rec_spect.cpp"),
           os.path.join(APPS, "nifti_io.cpp"), "-o", out_tmp, "-L" + HERE, "-l:" + LIB_NAME,
           "-Wl,-rpath,$ORIGIN", "-Wl,-rpath-link,/opt/rocm/lib", "-lz"]
    if verbose:
        print("[build]", " ".join(cmd), flush=True)
    subprocess.check_call(cmd)
    os.replace(out_tmp, REC_SPECT)
    return REC_SPECT


if __name__ == "__main__":
    build()
