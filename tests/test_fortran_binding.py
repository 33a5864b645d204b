"""The Fortran ISO_C_BINDING module compiles and its bind(C) types match the
C structs of include/sparc_hip.h (checked by a C program built against the
header).  Skipped where no Fortran compiler is installed."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FLANG = os.environ.get("FLANG", "/opt/rocm/lib/llvm/bin/flang")
FLANG_RT = "/opt/rocm/lib/llvm/lib/clang/22/lib/x86_64-unknown-linux-gnu"


def test_fortran_types_match_the_c_header(tmp_path):
    if not (os.path.exists(FLANG) and shutil.which("gcc")):
        pytest.skip("flang or gcc missing")
    d = str(tmp_path)
    run = lambda *cmd: subprocess.run(list(cmd), cwd=d, check=True, capture_output=True, text=True)
    run(FLANG, "-c", "-fPIC", os.path.join(ROOT, "sparc2_amd", "fortran", "sparc_hip.f90"), "-o", "m.o")
    run(FLANG, "-c", "-fPIC", os.path.join(ROOT, "tests", "fortran", "fill_structs.f90"), "-o", "f.o")
    run("gcc", "-c", "-I", os.path.join(ROOT, "include"),
        os.path.join(ROOT, "tests", "fortran", "check_layout.c"), "-o", "c.o")
    lib_dir = os.path.join(ROOT, "sparc2_amd")
    if not os.path.exists(os.path.join(lib_dir, "libsparc_hip.so")):
        pytest.skip("libsparc_hip.so not built")
    run("gcc", "c.o", "f.o", "m.o", "-o", "check", f"-L{lib_dir}", "-lsparc_hip",
        f"-Wl,-rpath,{lib_dir}", f"-L{FLANG_RT}", "-lflang_rt.runtime", "-lm", "-lstdc++")
    out = run(os.path.join(d, "check"))
    assert out.stdout.startswith("OK"), out.stdout


def test_fortran_module_binds_every_entry_point():
    """Each C entry point of the header has a bind(C) interface in the
    Fortran module (sparc_hip_strerror is covered by sparc_error_message)."""
    import re
    hdr = open(os.path.join(ROOT, "include", "sparc_hip.h")).read()
    f90 = open(os.path.join(ROOT, "sparc2_amd", "fortran", "sparc_hip.f90")).read()
    declared = set(re.findall(r"\b(sparc_hip_\w+)\s*\(", hdr))
    bound = set(re.findall(r'bind\(C,\s*name="(sparc_hip_\w+)"\)', f90))
    assert bound <= declared, bound - declared
    assert declared - bound <= {"sparc_hip_strerror"}, declared - bound
    # ... and every interface is public (the module is private by default)
    pub = re.search(r"^\s*public ::(.*?)\n\s*\n", f90, re.S | re.M).group(1)
    public = set(re.findall(r"\b(sparc_\w+)", pub))
    assert (bound - {"sparc_hip_last_error"}) <= public, (bound - {"sparc_hip_last_error"}) - public
