"""Builds libfrehg_gpu.so (the gfx950 backend) in-tree with hipcc.

Usage: python -m frehg_amd.build  (or frehg_amd.build.build()).
Objects go to frehg_amd/_build/; the shared library to frehg_amd/libfrehg_gpu.so and the command
line to frehg_amd/frehg_run.
"""
import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libfrehg_gpu.so")
BUILD = os.path.join(HERE, "_build")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("FREHG_OFFLOAD_ARCH", "gfx950")

FLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-Wall",
         # packed FP32 instructions are disabled for all device code
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-I", os.path.join(HERE, "..", "include")]

SOURCES = ["krylov.hip", "re.hip", "swe.hip", "coupling.hip", "mlp.hip", "ensemble.hip", "solute.hip", "noalloc.hip",
           "probe.hip", "api.hip"]
# Host-only C++ above the C ABI (configuration front end), built with g++.
HOST_SOURCES = ["yaml.cpp", "config.cpp", "driver.cpp"]
CXX = os.environ.get("CXX", "g++")
HOST_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-Wall", "-Wextra", "-I", os.path.join(HERE, "..", "include")]
# The command line, linked against the library (found next to it at run time).
CLI = os.path.join(HERE, "frehg_run")

# Per-file flags. The MLP kernels keep their MFMA accumulators in ordinary VGPRs (the guess kernel
# otherwise reads every result back from accumulation registers: 36 extra moves per 16-cell tile).
FILE_FLAGS = {"mlp.hip": ["-mllvm", "-amdgpu-mfma-vgpr-form"]}


def _compile(src):
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC, *FLAGS, *FILE_FLAGS.get(src, []), "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stdout}\n{r.stderr}")
    return obj, r.stderr


def _compile_host(src):
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    cmd = [CXX, *HOST_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"{CXX} failed for {src}:\n{r.stdout}\n{r.stderr}")
    return obj, r.stderr


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    srcs = [s for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    with cf.ThreadPoolExecutor(max_workers=8) as ex:
        jobs = [ex.submit(_compile, s) for s in srcs] + [ex.submit(_compile_host, s) for s in HOST_SOURCES]
        results = [j.result() for j in jobs]
    if verbose:
        for _, err in results:
            if err.strip():
                print(err, file=sys.stderr)
    objs = [o for o, _ in results]
    tmp = OUT + ".tmp"
    cmd = [HIPCC, "-shared", f"--offload-arch={ARCH}", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, OUT)
    cmd = [CXX, *HOST_FLAGS, os.path.join(CSRC, "frehg_run.cpp"), "-o", CLI + ".tmp", "-L", HERE, "-lfrehg_gpu",
           "-Wl,-rpath,$ORIGIN"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"frehg_run link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(CLI + ".tmp", CLI)
    return OUT


if __name__ == "__main__":
    print(build(verbose=True))
