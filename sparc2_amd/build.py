"""Build libsparc_hip.so (the HIP kernels + C ABI) in-tree for gfx950, and
libsparc_h5.so (the reference's HDF5 file layouts, host code over the HDF5 C
library) when HDF5 is installed, and libsparc_fits.so (the forcing cube of
the PULSE = 0 driver, over CFITSIO) when CFITSIO is.

    python -m sparc2_amd.build [--jobs N]

Each .hip source is compiled to an object with hipcc, then linked into
sparc2_amd/libsparc_hip.so.  Packed FP32 instructions are disabled for all
device code.  HDF5 and CFITSIO are looked up under $SPARC_HDF5_DIR (default
/opt/conda).
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_obj")
LIB = os.path.join(HERE, "libsparc_hip.so")
H5LIB = os.path.join(HERE, "libsparc_h5.so")
FITSLIB = os.path.join(HERE, "libsparc_fits.so")
HDF5_DIR = os.environ.get("SPARC_HDF5_DIR", "/opt/conda")
SOURCES = ["api.hip", "stage.hip", "ops.hip", "damp.hip"]
ARCH = os.environ.get("SPARC_OFFLOAD_ARCH", "gfx950")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
CFLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result",
]


def _compile(src: str) -> str:
    obj = os.path.join(OBJ, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC, *CFLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stderr}")
    warn = "\n".join(l for l in r.stderr.splitlines()
                     if "packed-fp32-ops" not in l and l.strip())
    if warn:
        sys.stderr.write(warn + "\n")
    return obj


def build_h5(verbose: bool = True):
    """libsparc_h5.so from csrc/h5io.cpp with g++ (no device code); returns
    its path, or None when the HDF5 headers and libraries are absent."""
    inc, libdir = os.path.join(HDF5_DIR, "include"), os.path.join(HDF5_DIR, "lib")
    if not (os.path.exists(os.path.join(inc, "hdf5.h")) and
            os.path.exists(os.path.join(libdir, "libhdf5_hl.so"))):
        if verbose:
            print(f"HDF5 not found under {HDF5_DIR}: libsparc_h5.so not built")
        return None
    tmp = H5LIB + ".tmp"
    cmd = ["g++", "-O2", "-fPIC", "-shared", "-std=c++17", "-Wall", f"-I{inc}",
           os.path.join(CSRC, "h5io.cpp"), f"-L{libdir}", "-lhdf5_hl", "-lhdf5",
           f"-Wl,-rpath,{libdir}", "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"g++ failed for h5io.cpp:\n{r.stderr}")
    os.replace(tmp, H5LIB)
    if verbose:
        print(f"built {H5LIB}")
    return H5LIB


def build_fits(verbose: bool = True):
    """libsparc_fits.so from csrc/fitsio.cpp with g++ (the forcing cube of the
    PULSE = 0 driver); None when CFITSIO is absent."""
    inc, libdir = os.path.join(HDF5_DIR, "include"), os.path.join(HDF5_DIR, "lib")
    if not (os.path.exists(os.path.join(inc, "fitsio.h")) and
            os.path.exists(os.path.join(libdir, "libcfitsio.so"))):
        if verbose:
            print(f"CFITSIO not found under {HDF5_DIR}: libsparc_fits.so not built")
        return None
    tmp = FITSLIB + ".tmp"
    cmd = ["g++", "-O2", "-fPIC", "-shared", "-std=c++17", "-Wall", f"-I{inc}",
           os.path.join(CSRC, "fitsio.cpp"), f"-L{libdir}", "-lcfitsio",
           f"-Wl,-rpath,{libdir}", "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"g++ failed for fitsio.cpp:\n{r.stderr}")
    os.replace(tmp, FITSLIB)
    if verbose:
        print(f"built {FITSLIB}")
    return FITSLIB


FLANG = os.environ.get("FLANG", "/opt/rocm/lib/llvm/bin/flang")
ROOT = os.path.dirname(HERE)
SHIM_TEST = os.path.join(ROOT, "tests", "fortran", "_build", "drive_quiet")


def build_fortran_shim_test(verbose: bool = True):
    """The test program of the Fortran drop-in shim (fortran/sparc_gpu.f90
    over fortran/sparc_hip.f90, with the test's stand-in all_modules) linked
    against libsparc_hip.so; None when flang is absent."""
    if not os.path.exists(FLANG):
        if verbose:
            print(f"flang not found at {FLANG}: shim test program not built")
        return None
    out = os.path.dirname(SHIM_TEST)
    os.makedirs(out, exist_ok=True)
    srcs = [os.path.join(HERE, "fortran", "sparc_hip.f90"),
            os.path.join(ROOT, "tests", "fortran", "all_modules.f90"),
            os.path.join(HERE, "fortran", "sparc_gpu.f90"),
            os.path.join(ROOT, "tests", "fortran", "drive_quiet.f90")]
    objs = []
    for src in srcs:
        obj = os.path.join(out, os.path.splitext(os.path.basename(src))[0] + ".o")
        r = subprocess.run([FLANG, "-c", "-O2", "-fPIC", "-module-dir", out, src, "-o", obj],
                           capture_output=True, text=True, cwd=out)
        if r.returncode != 0:
            raise RuntimeError(f"flang failed for {src}:\n{r.stderr}")
        objs.append(obj)
    tmp = SHIM_TEST + ".tmp"
    r = subprocess.run([FLANG, *objs, f"-L{HERE}", "-lsparc_hip",
                        "-Wl,-rpath,$ORIGIN/../../../sparc2_amd", "-o", tmp],
                       capture_output=True, text=True, cwd=out)
    if r.returncode != 0:
        raise RuntimeError(f"flang link failed:\n{r.stderr}")
    os.replace(tmp, SHIM_TEST)
    if verbose:
        print(f"built {SHIM_TEST}")
    return SHIM_TEST


def build(jobs: int = 4, verbose: bool = True) -> str:
    os.makedirs(OBJ, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-lhipfft", "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    build_h5(verbose)
    build_fits(verbose)
    build_fortran_shim_test(verbose)
    return LIB


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=4)
    a = ap.parse_args()
    build(a.jobs)


if __name__ == "__main__":
    main()
