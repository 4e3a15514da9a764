"""The C-ABI library loads and exports every entry point include/swm_gpu.h
declares (no device calls)."""
import os
import re

import pytest

from conftest import ROOT


def declared_symbols():
    with open(os.path.join(ROOT, "include", "swm_gpu.h")) as fh:
        text = fh.read()
    return sorted(set(re.findall(r"\b(swm_gpu_[a-z_]+)\s*\(", text)))


def test_header_declares_the_binding_list():
    from swm_amd import gpu
    assert declared_symbols() == sorted(gpu.EXPORTED)


def test_library_exports_all_symbols():
    import ctypes
    lib_path = os.path.join(ROOT, "swm_amd", "libswmgpu.so")
    if not os.path.exists(lib_path):
        pytest.fail("swm_amd/libswmgpu.so not built: run __graft_entry__.build()")
    lib = ctypes.CDLL(lib_path)
    for sym in declared_symbols():
        assert hasattr(lib, sym), sym


def test_binding_structs_match_header_layout():
    """ctypes mirrors of swm_desc / swm_tracer_desc have the C sizes (checked
    by compiling a tiny C program against the header)."""
    import ctypes
    import subprocess
    import tempfile
    from swm_amd import gpu
    src = ('#include <stdio.h>\n#include "swm_gpu.h"\nint main(void){printf("%zu %zu %zu\\n", sizeof(swm_desc),'
           ' sizeof(swm_tracer_desc), sizeof(swm_grid_desc)); return 0;}\n')
    with tempfile.TemporaryDirectory() as d:
        c = os.path.join(d, "t.c")
        with open(c, "w") as fh:
            fh.write(src)
        exe = os.path.join(d, "t")
        subprocess.run(["gcc", "-I" + os.path.join(ROOT, "include"), c, "-o", exe], check=True)
        out = subprocess.run([exe], capture_output=True, text=True, check=True).stdout.split()
    assert [int(x) for x in out] == [ctypes.sizeof(gpu.Desc), ctypes.sizeof(gpu.TracerDesc),
                                     ctypes.sizeof(gpu.GridDesc)]


def test_missing_library_fails_loudly(tmp_path):
    from swm_amd import gpu
    with pytest.raises(gpu.SwmGpuError):
        gpu._lib = None
        try:
            gpu.load_library(str(tmp_path / "absent.so"))
        finally:
            gpu._lib = None


def test_fortran_shim_compiles(tmp_path):
    """The ISO_C_BINDING module a maintainer adds to the reference build
    compiles with the image's Fortran compiler, and every bind(C) name it
    declares is exported by the library."""
    import shutil
    import subprocess
    fc = shutil.which("amdflang") or "/opt/rocm/bin/amdflang"
    if not os.path.exists(fc):
        pytest.skip("no Fortran compiler in this image")
    src = os.path.join(ROOT, "swm_amd", "fortran", "swm_gpu_shim.f90")
    r = subprocess.run([fc, "-c", src, "-o", str(tmp_path / "shim.o"), "-module-dir", str(tmp_path)],
                       capture_output=True, text=True, cwd=tmp_path)
    assert r.returncode == 0, r.stderr
    from swm_amd import gpu
    names = set(re.findall(r'bind\(C, name="([a-z_]+)"\)', open(src).read()))
    assert names and names <= set(gpu.EXPORTED)
