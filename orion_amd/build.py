"""Builds liborion_amd.so (HIP for gfx950 + host C++) in-tree, and the oracle.

    python -m orion_amd.build            # library + oracle (+ oracle/_ref if the reference is present)
    python -m orion_amd.build --sanitize # host C++ under ASan + UBSan, into orion_amd/_asan/

Compiler rules: every HIP and host translation unit is compiled with
-ffp-contract=off (no FMA contraction: the float operation sequence must equal
the CPU one), HIP code additionally with packed-FP32 instructions disabled.
"""
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
BUILD = PKG / "_build"
LIB = PKG / "liborion_amd.so"

HIP_SOURCES = ["runtime.hip", "path_kernels.hip", "sppm_kernels.hip", "batch.hip"]
CPP_SOURCES = ["api_host.cpp", "host_records.cpp", "ori_loader.cpp", "scenegen.cpp", "bvh.cpp", "png_io.cpp", "film_io.cpp"]

ARCH = os.environ.get("ORION_AMD_ARCH", "gfx950")
COMMON = ["-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-Wall", "-Wno-unused-function",
          "-Wno-unused-variable"]
HIP_FLAGS = COMMON + [f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang",
                      "-packed-fp32-ops", "-munsafe-fp-atomics"]


def _hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found (ROCm is required to build orion_amd)")


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"build step failed ({r.returncode}): {' '.join(cmd)}\n{r.stdout}")
    return r.stdout


def build_library(verbose=False):
    hipcc = _hipcc()
    BUILD.mkdir(exist_ok=True)
    jobs = []
    for s in HIP_SOURCES:
        jobs.append([hipcc, *HIP_FLAGS, "-c", str(CSRC / s), "-o", str(BUILD / (s + ".o"))])
    for s in CPP_SOURCES:
        jobs.append(["g++", *COMMON, "-O2", "-c", str(CSRC / s), "-o", str(BUILD / (s + ".o"))])
    with ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 4)) as ex:
        outs = list(ex.map(_run, jobs))
    if verbose:
        for o in outs:
            if o.strip():
                print(o)
    objs = [str(BUILD / (s + ".o")) for s in HIP_SOURCES + CPP_SOURCES]
    tmp = str(LIB) + ".tmp"
    _run([hipcc, "-shared", "-fPIC", f"--offload-arch={ARCH}", *objs, "-o", tmp, "-lpthread", "-lz"])
    os.replace(tmp, LIB)
    # the C++ host layer's driver (orion_amd/host), linked against the library
    drv = PKG / "orion_render"
    _run(["g++", "-O2", "-std=c++17", str(PKG / "host" / "render_ori.cpp"), "-o", str(drv) + ".tmp",
          f"-L{PKG}", "-lorion_amd", "-Wl,-rpath,$ORIGIN"])
    os.replace(str(drv) + ".tmp", drv)
    return LIB


def build_sanitized():
    """A copy of the library whose host C++ (scene loader, PNG reader, BVH
    builder, scene generators, film writers, record checks) is built with
    AddressSanitizer and UndefinedBehaviorSanitizer, into orion_amd/_asan/.
    The HIP translation units are built as usual (no GPU code is
    instrumented).  Load it with ORION_AMD_LIB=orion_amd/_asan/liborion_amd.so
    and gcc's sanitizer runtimes preloaded:

        LD_PRELOAD="$(gcc -print-file-name=libasan.so) $(gcc -print-file-name=libubsan.so)" \
        ASAN_OPTIONS=detect_leaks=0 ORION_AMD_LIB=orion_amd/_asan/liborion_amd.so \
        python -m pytest tests -m "not gpu"
    """
    hipcc = _hipcc()
    out = PKG / "_asan"
    objd = out / "obj"
    objd.mkdir(parents=True, exist_ok=True)
    san = ["-O1", "-g", "-fno-omit-frame-pointer", "-fsanitize=address,undefined", "-fno-sanitize-recover=undefined"]
    jobs = [[hipcc, *HIP_FLAGS, "-c", str(CSRC / s), "-o", str(objd / (s + ".o"))] for s in HIP_SOURCES]
    jobs += [["g++", *COMMON, *san, "-c", str(CSRC / s), "-o", str(objd / (s + ".o"))] for s in CPP_SOURCES]
    with ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 4)) as ex:
        list(ex.map(_run, jobs))
    objs = [str(objd / (s + ".o")) for s in HIP_SOURCES + CPP_SOURCES]
    lib = out / "liborion_amd.so"
    _run([hipcc, "-shared", "-fPIC", f"--offload-arch={ARCH}", *objs, "-o", str(lib), "-lpthread", "-lz"])
    return lib


def build_oracle():
    """Compiles the oracle (test infrastructure) and, where the reference tree
    exists, oracle/_ref from the reference's own sources."""
    _run(["make", "-s", "-C", str(ROOT / "oracle"), "oracle"])
    if Path("/root/reference/src/sampler/sobol.h").exists():
        _run(["make", "-s", "-j8", "-C", str(ROOT / "oracle"), "ref"])


def main():
    if "--sanitize" in sys.argv:
        print(f"built {build_sanitized()}")
        return
    build_library(verbose="-v" in sys.argv)
    build_oracle()
    print(f"built {LIB}")


if __name__ == "__main__":
    main()
