"""Builds libnqes.so (HIP, gfx950) and the rbm-matfree drop-in executable in-tree.

    python nqes-trapped-ion_amd/build.py [--jobs N]

Outputs (git-ignored, shipped with the tree):
    nqes-trapped-ion_amd/lib/libnqes.so
    nqes-trapped-ion_amd/bin/rbm-matfree
"""
import argparse
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIBDIR = os.path.join(HERE, "lib")
BINDIR = os.path.join(HERE, "bin")
OBJDIR = os.path.join(HERE, "_build")
ARCH = os.environ.get("NQES_OFFLOAD_ARCH", "gfx950")

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"

COMMON = [
    "-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-function",
    f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", os.path.join(REPO, "include"),
]
# FP64 path: no contraction anywhere (host or device) -- canonical operation order
EXACT = ["-ffp-contract=off"]

# mixed path: FMA contraction allowed in its kernels (FP32 work, FP64 sums)
SOURCES = [
    ("kernels_fp64.hip", EXACT),
    ("kernels_mixed.hip", []),
    ("sampler_mx.hip", ["-DNQES_SWEEP_K=1"], "sampler_mx_k1"),
    ("sampler_mx.hip", ["-DNQES_SWEEP_K=2"], "sampler_mx_k2"),
    ("sampler_mx.hip", ["-DNQES_SWEEP_K=3"], "sampler_mx_k3"),
    ("sampler_mx.hip", ["-DNQES_SWEEP_K=4"], "sampler_mx_k4"),
    ("sv_factored.hip", []),
    ("sv_q8.hip", []),
    ("sv_q8_gemm.hip", ["-mllvm", "-amdgpu-mfma-vgpr-form"]),
    ("sv_q8_fused.hip", ["-mllvm", "-amdgpu-mfma-vgpr-form", "-DNQES_FZ_K=1"], "sv_q8_fused_k1"),
    ("sv_q8_fused.hip", ["-mllvm", "-amdgpu-mfma-vgpr-form", "-DNQES_FZ_K=2"], "sv_q8_fused_k2"),
    ("sv_q8_fused.hip", ["-mllvm", "-amdgpu-mfma-vgpr-form", "-DNQES_FZ_K=4"], "sv_q8_fused_k4"),
    ("nqes_api.cpp", EXACT),
    ("noalloc.hip", EXACT),
    ("serve.cpp", EXACT),
    ("eig.cpp", EXACT),
]


def _compile(src, extra, obj=None):
    out = os.path.join(OBJDIR, (obj or os.path.basename(src)) + ".o")
    path = os.path.join(CSRC, src)
    lang = ["-x", "hip"] if src.endswith(".hip") else ["-x", "hip"]
    cmd = [HIPCC] + COMMON + extra + lang + ["-c", path, "-o", out]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return out


def build(jobs=8, verbose=False):
    os.makedirs(OBJDIR, exist_ok=True)
    os.makedirs(LIBDIR, exist_ok=True)
    os.makedirs(BINDIR, exist_ok=True)
    with ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        objs = list(ex.map(lambda se: _compile(*se), SOURCES))
    lib = os.path.join(LIBDIR, "libnqes.so")
    tmp = lib + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, lib)
    cli_src = os.path.join(CSRC, "cli.cpp")
    if os.path.exists(cli_src):
        exe = os.path.join(BINDIR, "rbm-matfree")
        cmd = ["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-Wall", "-I", os.path.join(REPO, "include"),
               cli_src, os.path.join(CSRC, "matv5.cpp"),
               "-L", LIBDIR, "-lnqes", "-lz", "-pthread", "-Wl,-rpath,$ORIGIN/../lib", "-o", exe + ".tmp"]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"cli build failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
        os.replace(exe + ".tmp", exe)
    if verbose:
        print("built", lib)
    return lib


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=8)
    args = ap.parse_args()
    build(args.jobs, verbose=True)
    sys.exit(0)
