"""Build the native library liblitho_amd.so (HIP for gfx950) in the source tree.

Every translation unit under csrc/ is compiled by hipcc into _build/ and linked
with rocBLAS / rocSOLVER into c-version-lithography_amd/liblitho_amd.so.  Runs
without a GPU (hipcc cross-compiles).  Used by __graft_entry__.build().
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
BUILD = PKG.parent / "_build" / "litho_amd"
LIB = PKG / "liblitho_amd.so"
BENCH = PKG / "litho_bench"
CPP_TEST = PKG / "test_litho_cpp"
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

COMMON = ["-O3", "-std=c++17", "-fPIC", "-fvisibility=hidden", f"--offload-arch={ARCH}",
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-Wall", "-Wno-unused-function", "-Wno-unused-command-line-argument"]

SOURCES = ["socs_kernels.hip", "corr_mfma.hip", "generic_ops.hip", "opt_kernels.hip", "precompute.hip", "levelset.hip", "tiling.hip", "noalloc.hip", "binarize.hip", "epe.hip", "plan.cpp",
           "optics.cpp", "ctm_driver.cpp", "batch.cpp", "server.cpp", "lsm.cpp", "fullchip.cpp", "io.cpp",
           "litho_capi.cpp"]


def _compile(src: Path, obj: Path) -> tuple[str, int, str]:
    cmd = [HIPCC, *COMMON, "-x", "hip", "-c", str(src), "-o", str(obj), "-I", str(CSRC),
           "-I", str(PKG.parent / "include")]
    p = subprocess.run(cmd, capture_output=True, text=True)
    out = "\n".join(l for l in (p.stdout + p.stderr).splitlines() if "packed-fp32-ops" not in l)
    return src.name, p.returncode, out


def build(verbose: bool = False, jobs: int | None = None) -> Path:
    BUILD.mkdir(parents=True, exist_ok=True)
    objs = []
    jobs = jobs or min(len(SOURCES), max(1, (os.cpu_count() or 4)))
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        futs = []
        for s in SOURCES:
            obj = BUILD / (s + ".o")
            objs.append(obj)
            futs.append(ex.submit(_compile, CSRC / s, obj))
        failed = []
        for f in futs:
            name, rc, out = f.result()
            if verbose and out.strip():
                print(f"[{name}]\n{out}", file=sys.stderr)
            if rc != 0:
                failed.append((name, out))
    if failed:
        msg = "\n".join(f"--- {n}\n{o}" for n, o in failed)
        raise RuntimeError(f"hipcc failed:\n{msg}")
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-o", str(tmp), *map(str, objs),
           "-L/opt/rocm/lib", "-lrocsolver", "-lrocblas", "-Wl,-rpath,/opt/rocm/lib"]
    p = subprocess.run(cmd, capture_output=True, text=True)
    if p.returncode != 0:
        raise RuntimeError(f"link failed:\n{p.stdout}\n{p.stderr}")
    os.replace(tmp, LIB)
    # C++ caller of the C ABI, used for profiling runs without Python
    btmp = BENCH.with_suffix(".tmp")
    cmd = [HIPCC, "-O2", "-std=c++17", "-x", "hip", str(CSRC / "litho_bench_main.cpp"), "-o", str(btmp),
           f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
           "-Wno-unused-command-line-argument", "-L", str(PKG), "-llitho_amd", f"-Wl,-rpath,$ORIGIN",
           "-Wl,-rpath,/opt/rocm/lib"]
    p = subprocess.run(cmd, capture_output=True, text=True)
    if p.returncode != 0:
        raise RuntimeError(f"litho_bench link failed:\n{p.stdout}\n{p.stderr}")
    os.replace(btmp, BENCH)
    # C++ tests of the reference call surface (include/litho_cpp.hpp)
    ttmp = CPP_TEST.with_suffix(".tmp")
    cmd = [HIPCC, "-O2", "-std=c++17", "-x", "hip", str(PKG.parent / "tests" / "cpp" / "test_litho_cpp.cpp"), "-o",
           str(ttmp), f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
           "-Wno-unused-command-line-argument", "-I", str(PKG.parent / "include"), "-L", str(PKG), "-llitho_amd",
           f"-Wl,-rpath,$ORIGIN", "-Wl,-rpath,/opt/rocm/lib"]
    p = subprocess.run(cmd, capture_output=True, text=True)
    if p.returncode != 0:
        raise RuntimeError(f"test_litho_cpp build failed:\n{p.stdout}\n{p.stderr}")
    os.replace(ttmp, CPP_TEST)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
