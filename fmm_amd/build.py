"""Build libfmm_hip.so (gfx950) in-tree with hipcc, and the Fortran module
shim with flang when it is available.

    python -m fmm_amd.build            # library + Fortran shim
"""
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libfmm_hip.so")
# slowest first: the wide-order near-field units take minutes each (one kernel per order pair and damping)
SOURCES = ["near_wide_k7a.hip", "near_wide_k7b.hip", "near_wide_k6a.hip", "near_wide_k6b.hip", "near_wide_k5.hip",
           "near_wide_k34.hip", "expansions.hip", "expansions_mfma.hip",
           "near.hip", "m2l.hip", "selector.hip", "shift_mfma.hip", "m2l_gemm.hip", "m2l_block.hip", "tree.hip", "periodic.hip",
           "qmmm.hip", "prims.hip", "api.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLANG = os.environ.get("FLANG", "/opt/rocm/lib/llvm/bin/flang")
ARCH = "gfx950"
HIPFLAGS = ["-O3", "-std=c++17", f"--offload-arch={ARCH}", "-fPIC", "-munsafe-fp-atomics",
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            "-Wno-unused-result", "-I", os.path.join(ROOT, "include")]


def _compile(src):
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC, *HIPFLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stderr[-6000:]}")
    return obj


def build_library(verbose=True):
    os.makedirs(BUILD, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=min(len(SOURCES), os.cpu_count() or 4)) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr[-4000:]}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


def build_fortran(verbose=True):
    """Fortran module `fmm` (+ alias `octree`) over the C ABI, as a shared
    library plus .mod files under fmm_amd/fortran/_build."""
    if not os.path.exists(FLANG):
        if verbose:
            print("flang not found: Fortran shim skipped")
        return None
    src = os.path.join(HERE, "fortran", "fmm_hip.F90")
    out = os.path.join(HERE, "fortran", "_build")
    os.makedirs(out, exist_ok=True)
    obj = os.path.join(out, "fmm_hip.o")
    lib = os.path.join(HERE, "fortran", "libfmm_fortran.so")
    subprocess.run([FLANG, "-cpp", "-O2", "-fPIC", "-J", out, "-c", src, "-o", obj], check=True)
    subprocess.run([FLANG, "-shared", "-fPIC", "-o", lib, obj, f"-L{HERE}", "-lfmm_hip",
                    f"-Wl,-rpath,{HERE}"], check=True)
    drv = os.path.join(HERE, "fortran", "drive.x")
    subprocess.run([FLANG, "-cpp", "-O2", "-I", out, "-o", drv, os.path.join(HERE, "fortran", "drive.F90"),
                    f"-L{os.path.join(HERE, 'fortran')}", "-lfmm_fortran", f"-L{HERE}", "-lfmm_hip",
                    f"-Wl,-rpath,{os.path.join(HERE, 'fortran')}", f"-Wl,-rpath,{HERE}"], check=True)
    if verbose:
        print(f"built {lib} and {drv}")
    return lib


def build(verbose=True):
    build_library(verbose)
    if os.path.exists(os.path.join(HERE, "fortran", "fmm_hip.F90")):
        build_fortran(verbose)


if __name__ == "__main__":
    build()
    sys.exit(0)
