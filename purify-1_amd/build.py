"""Builds the HIP library libpurify_amd.so for gfx950 (in-tree, next to the Python package).

    python purify-1_amd/build.py            # compile every .hip source, link the shared library
    python purify-1_amd/build.py --asm      # also dump the gfx950 assembly to _build/ for review

Each translation unit is compiled separately (in parallel) with hipcc, then linked.  Packed FP32
instructions are disabled for device code (-packed-fp32-ops).
"""
import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
OUT = os.path.join(HERE, "purify_amd", "libpurify_amd.so")
ARCH = os.environ.get("PURIFY_AMD_ARCH", "gfx950")

SOURCES = ["pf_fft.hip", "pf_grid.hip", "pf_wproj.hip", "pf_wmat.hip", "pf_wavelet.hip", "pf_op.hip", "pf_solver.hip",
           "pf_noalloc.hip", "pf_chirp.hip", "pf_build.hip", "pf_mem.hip"]

CFLAGS = [
    "-std=c++17", "-O3", "-fPIC", f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-result",
    "-I", os.path.join(ROOT, "include"), "-I", CSRC,
]


def hipcc():
    p = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    return p


def compile_one(src, asm=False):
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    cmd = [hipcc(), "-c", os.path.join(CSRC, src), "-o", obj] + CFLAGS
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stdout}\n{r.stderr}")
    if asm:
        s = os.path.join(BUILD, os.path.splitext(src)[0] + ".s")
        subprocess.run([hipcc(), "-S", "--cuda-device-only", os.path.join(CSRC, src), "-o", s] + CFLAGS,
                       check=True, capture_output=True)
    return obj, r.stderr


def build(asm=False, verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    srcs = [s for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    with cf.ThreadPoolExecutor(max_workers=min(len(srcs), os.cpu_count() or 4)) as ex:
        results = list(ex.map(lambda s: compile_one(s, asm), srcs))
    for _, err in results:
        if verbose and err.strip():
            print(err, file=sys.stderr)
    objs = [o for o, _ in results]
    tmp = OUT + ".tmp"
    cmd = [hipcc(), "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--asm", action="store_true")
    ap.add_argument("-v", action="store_true")
    a = ap.parse_args()
    print(build(asm=a.asm, verbose=a.v))
