"""Merged database files (include/mck_dbio.h) on the CPU: the library exports what the
header declares; a database written in the merged layout (repack_database.py:151-187)
opens with every array, attribute and snapshot equal to what was written, memory-mapped
(contiguous merged_snapshots) or read; malformed files are rejected with a message."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from mc_kernel_amd import dbfile, synth

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "mck_dbio.h")
built = pytest.mark.skipif(not os.path.exists(dbfile.library_path()), reason="libmck_dbio.so not built")


def test_header_declares_exports():
    text = open(HEADER).read()
    assert set(re.findall(r"\b(mck_db[a-z_]*)\s*\(", text)) == set(dbfile.EXPORTS)


@built
def test_library_exports_every_declared_symbol():
    lib = C.CDLL(dbfile.library_path())
    assert all(hasattr(lib, f) for f in dbfile.EXPORTS)


def test_struct_layout_matches_c(tmp_path):
    import subprocess
    lines = ["#include <stdio.h>", "#include <stddef.h>", '#include "%s"' % HEADER, "int main(void) {",
             'printf("size %zu\\n", sizeof(mck_dbfile_t));']
    for f in dbfile.MckDbFile._fields_:
        lines.append('printf("%s %%zu\\n", offsetof(mck_dbfile_t, %s));' % (f[0], f[0]))
    lines += ["return 0;", "}"]
    src = tmp_path / "l.c"
    src.write_text("\n".join(lines))
    subprocess.run(["gcc", "-std=c11", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(tmp_path / "l")],
                   check=True)
    out = subprocess.run([str(tmp_path / "l")], check=True, capture_output=True, text=True).stdout.split("\n")
    got = dict(ln.split() for ln in out if ln)
    assert int(got["size"]) == C.sizeof(dbfile.MckDbFile)
    for f in dbfile.MckDbFile._fields_:
        assert int(got[f[0]]) == getattr(dbfile.MckDbFile, f[0]).offset, f[0]


@pytest.fixture(scope="module")
def written(tmp_path_factory):
    p = synth.make_problem("tiny", ntet=4)
    d = tmp_path_factory.mktemp("db")
    paths = {w: str(d / ("%s_merged_instaseis_db.nc4" % w)) for w in ("fwd", "bwd")}
    for w, path in paths.items():
        dbfile.write_problem_db(path, p, w)
    return p, paths


@built
@pytest.mark.parametrize("mapped", [True, False])
@pytest.mark.parametrize("which", ["fwd", "bwd"])
def test_round_trip(written, which, mapped):
    p, paths = written
    m = p.mesh
    u = p.fwd_u if which == "fwd" else p.bwd_u
    with dbfile.DbFile(paths[which], nvar_expected=10 if which == "fwd" else 5, map_snapshots=mapped) as f:
        assert f.mapped == (1 if mapped else 0)
        d = f.db
        assert (d.npol, d.nelem, d.npoints, d.nvar) == (4, m.nelem, m.npoints, 10 if which == "fwd" else 5)
        # stored as a 32-bit real, as the reference reads it (readfields.f90:651-655)
        assert d.amplitude == float(np.float32(1.0e20))
        assert (f.nt, f.dt, f.file_version) == (p.nt, p.dt, 3)
        assert f.planet_radius == p.setup.planet_radius
        assert f.source_shift == 15.0 and f.source_depth_km == p.src_depth
        for name, ref in (("mesh_s", m.mesh_s), ("mesh_z", m.mesh_z), ("mp_s", m.mp_s), ("mp_z", m.mp_z),
                          ("fem_mesh", m.fem), ("eltype", m.eltype), ("axis", m.axis), ("sem_mesh", m.sem)):
            np.testing.assert_array_equal(f.array(name), ref, err_msg=name)
        np.testing.assert_array_equal(f.array("snapshots"), np.asarray(u).reshape(m.nelem, -1, 5, 5, p.nt))
        np.testing.assert_array_equal(f.array("stf_dump"), p.stf.astype(np.float32).astype(np.float64))
        assert f.array("mesh_vp") is None


@built
def test_rejects_malformed(written, tmp_path):
    p, paths = written
    with pytest.raises(dbfile.DbIoError, match="cannot open"):
        dbfile.DbFile(str(tmp_path / "missing.nc4"))
    with pytest.raises(dbfile.DbIoError, match="5 variables, expected 10"):
        dbfile.DbFile(paths["bwd"], nvar_expected=10)
    bad = tmp_path / "not_hdf5.nc4"
    bad.write_bytes(b"CDF\x01" + b"\0" * 64)          # a NetCDF-3 header
    with pytest.raises(dbfile.DbIoError, match="cannot open"):
        dbfile.DbFile(str(bad))


@built
def test_fortran_binding_layout(tmp_path):
    """The iso_c_binding module mc_kernel_amd/fortran/mck_dbio.f90 has the C layout of
    mck_dbfile_t and links against libmck_dbio.so (a missing file fails with a message)."""
    import shutil
    import subprocess
    flang = shutil.which("flang") or shutil.which("amdflang")
    if flang is None or not os.path.exists(os.path.join(ROOT, "mc_kernel_amd", "libmck.so")):
        pytest.skip("flang or libmck.so missing")
    fdir = os.path.join(ROOT, "mc_kernel_amd", "fortran")
    for mod in ("mck_gpu.f90", "mck_dbio.f90"):
        subprocess.run([flang, "-c", os.path.join(fdir, mod), "-o", str(tmp_path / (mod[:-4] + ".o")),
                        "-module-dir", str(tmp_path)], check=True, cwd=tmp_path)
    exe = tmp_path / "dbio_layout_check"
    libdir = os.path.dirname(dbfile.library_path())
    subprocess.run([flang, os.path.join(ROOT, "tests", "fortran", "dbio_layout_check.f90"),
                    str(tmp_path / "mck_gpu.o"), str(tmp_path / "mck_dbio.o"), "-I", str(tmp_path),
                    "-L" + libdir, "-lmck_dbio", "-lmck", "-Wl,-rpath," + libdir, "-o", str(exe)],
                   check=True, cwd=tmp_path)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True, cwd=tmp_path).stdout
    got = {}
    for ln in out.splitlines():
        parts = ln.split(None, 2)
        got[parts[0]] = parts[1:]
    assert int(got["size"][0]) == C.sizeof(dbfile.MckDbFile)
    for field in ("nt", "dt", "file_version", "stf_dump", "mesh_eta", "mapped", "priv"):
        assert int(got[field][0]) == getattr(dbfile.MckDbFile, field).offset, field
    assert int(got["open"][0]) != 0 and "cannot open" in got["open"][1]
