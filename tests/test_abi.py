"""The C-ABI boundary (include/mck.h) without a GPU: the library loads, exports every
declared entry point, and the ctypes mirror of the structs has the C layout."""
import ctypes as C
import os
import re
import shutil
import subprocess

import pytest

from mc_kernel_amd import abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "mck.h")
LIB = os.path.join(ROOT, "mc_kernel_amd", "libmck.so")

STRUCTS = {"mck_db_t": abi.MckDb, "mck_filter_t": abi.MckFilter,
           "mck_setup_t": abi.MckSetup, "mck_stats_t": abi.MckStats, "mck_tuning_t": abi.MckTuning}


def declared_functions():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"\b(mck_[a-z_0-9]+)\s*\(", text)))


def test_header_declares_exports():
    assert set(declared_functions()) == set(abi.EXPORTS)


@pytest.mark.skipif(not os.path.exists(LIB), reason="libmck.so not built")
def test_library_exports_every_declared_symbol():
    lib = C.CDLL(LIB)
    missing = [f for f in declared_functions() if not hasattr(lib, f)]
    assert not missing, missing


@pytest.mark.skipif(not os.path.exists(LIB), reason="libmck.so not built")
def test_last_error_without_gpu():
    lib = C.CDLL(LIB)
    lib.mck_last_error.restype = C.c_char_p
    assert isinstance(lib.mck_last_error(), bytes)


@pytest.mark.skipif(shutil.which("gcc") is None, reason="gcc not available")
def test_struct_layout_matches_c(tmp_path):
    lines = ["#include <stdio.h>", "#include <stddef.h>", '#include "%s"' % HEADER,
             "int main(void) {"]
    for cname, py in STRUCTS.items():
        lines.append('printf("%s size %%zu\\n", sizeof(%s));' % (cname, cname))
        for f in py._fields_:
            lines.append('printf("%s %s %%zu\\n", offsetof(%s, %s));'
                         % (cname, f[0], cname, f[0]))
    lines += ["return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c11", str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout
    got = {}
    for ln in out.splitlines():
        s, f, v = ln.split()
        got[(s, f)] = int(v)
    for cname, py in STRUCTS.items():
        assert got[(cname, "size")] == C.sizeof(py), cname
        for f in py._fields_:
            assert got[(cname, f[0])] == getattr(py, f[0]).offset, (cname, f[0])


def test_product_has_no_cpu_fallback(monkeypatch, tmp_path):
    """Without the built library the package raises; it never computes on the CPU."""
    import mc_kernel_amd as mk
    monkeypatch.setattr(mk, "_LIB", None, raising=False)
    monkeypatch.setattr(mk, "library_path", lambda: str(tmp_path / "missing.so"))
    with pytest.raises((RuntimeError, OSError)):
        mk.load_library()


def test_product_does_not_import_oracle():
    pkg = os.path.join(ROOT, "mc_kernel_amd")
    for dirpath, _, files in os.walk(pkg):
        for fn in files:
            if fn.endswith((".py", ".hip", ".cpp", ".h", ".c")):
                text = open(os.path.join(dirpath, fn), errors="replace").read()
                assert "import oracle" not in text and "from oracle" not in text, fn
                assert "liboracle" not in text or fn == "build.py", fn


FLANG = shutil.which("flang") or shutil.which("amdflang")


@pytest.mark.skipif(FLANG is None or not os.path.exists(LIB), reason="flang or libmck.so missing")
def test_fortran_binding_layout(tmp_path):
    """The iso_c_binding module (mc_kernel_amd/fortran/mck_gpu.f90) has the C layout of
    include/mck.h and links against the library (no GPU call)."""
    mod = os.path.join(ROOT, "mc_kernel_amd", "fortran", "mck_gpu.f90")
    prog = os.path.join(ROOT, "tests", "fortran", "layout_check.f90")
    subprocess.run([FLANG, "-c", mod, "-o", str(tmp_path / "mck_gpu.o"), "-module-dir", str(tmp_path)],
                   check=True, cwd=tmp_path)
    exe = tmp_path / "layout_check"
    subprocess.run([FLANG, prog, str(tmp_path / "mck_gpu.o"), "-I", str(tmp_path),
                    "-L" + os.path.dirname(LIB), "-lmck", "-Wl,-rpath," + os.path.dirname(LIB),
                    "-o", str(exe)], check=True, cwd=tmp_path)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout
    got = {}
    for ln in out.splitlines():
        parts = ln.split()
        if parts[0] == "last_error":
            continue
        got[(parts[0], parts[1])] = int(parts[2])
    for cname, py in STRUCTS.items():
        if (cname, "size") in got:
            assert got[(cname, "size")] == C.sizeof(py), cname
    for (cname, field), v in got.items():
        if field != "size":
            assert v == getattr(STRUCTS[cname], field).offset, (cname, field)


def test_python_binding_methods():
    """The Python binding keeps one method per entry point (no GPU call)."""
    import mc_kernel_amd as mk
    for m in ("slave_work", "stats", "locate", "load_strain", "integrate_model", "set_hetero",
              "integrate_hetero", "results", "seismograms", "base_kernels", "close"):
        assert callable(getattr(mk.MCKernel, m, None)), m
    assert callable(mk.load_het_rtpv)
