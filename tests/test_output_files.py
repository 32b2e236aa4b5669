"""Output files of a kernel run (include/mck_output.h) on the CPU:

  * XDMF: the .nc datasets read back with HDF5's own h5dump equal what was written
    (points in single precision, grid 0-based, block ids, node_data / cell_data variables
    entry-major); the .xdmf is well-formed XML whose topology, geometry and one hyperslab
    attribute per entry point at those datasets (dump_data_xdmf, inversion_mesh.f90:
    2026-2265);
  * CSR: the binary records and counts equal a direct selection of |value| > threshold in
    kernel-major, point order (inversion_mesh.f90:1886-1953); the text form writes E15.8
    as Fortran does (known strings);
  * intermediate results: written and read back (master_queue.f90:586-659), completed =
    element_proc >= 1, dimension mismatches rejected;
  * the library exports every function include/mck_output.h declares.
"""
import os
import re
import shutil
import subprocess
import xml.etree.ElementTree as ET

import numpy as np
import pytest

from mc_kernel_amd import dbfile, output

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
H5DUMP = shutil.which("h5dump") or ("/opt/conda/bin/h5dump" if os.path.exists("/opt/conda/bin/h5dump") else None)
CONN = np.array([[1, 2, 3, 4], [2, 3, 5, 4], [3, 6, 5, 1]], np.int32)


def _h5(path, name, dtype, tmp):
    if H5DUMP is None:
        pytest.skip("h5dump not available")
    out = os.path.join(tmp, name.strip("/").replace("/", "_") + ".bin")
    subprocess.run([H5DUMP, "-d", name, "-b", "LE", "-o", out, path], check=True, capture_output=True)
    return np.fromfile(out, dtype)


def test_exports_declared():
    text = open(os.path.join(ROOT, "include", "mck_output.h")).read()
    assert set(re.findall(r"\bint (mck_[a-z_]+)\s*\(", text)) == set(output.EXPORTS)
    lib = dbfile.load_library()
    assert all(hasattr(lib, f) for f in output.EXPORTS)


def test_xdmf_files(tmp_path):
    rng = np.random.default_rng(1)
    verts = rng.random((6, 3)) * 6.371e6
    kern = rng.standard_normal((2, 6)).astype(np.float32)
    nit = rng.integers(1, 9, (2, 3)).astype(np.float32)
    base = str(tmp_path / "run_kernel")
    output.write_xdmf(base, "tet", verts, CONN, [("kernel", False, kern, ["K_x_a", "K_x_b"]),
                                                   ("niterations", True, nit, ["nit_a", "nit_b"])],
                      block_id=[1, 2, 1])
    nc = base + ".nc"
    np.testing.assert_array_equal(_h5(nc, "/points", np.float32, tmp_path), verts.astype(np.float32).ravel())
    np.testing.assert_array_equal(_h5(nc, "/grid", np.int32, tmp_path), (CONN - 1).ravel())
    np.testing.assert_array_equal(_h5(nc, "/block", np.int32, tmp_path), [1, 2, 1])
    np.testing.assert_array_equal(_h5(nc, "/node_data/kernel", np.float32, tmp_path), kern.ravel())
    np.testing.assert_array_equal(_h5(nc, "/cell_data/niterations", np.float32, tmp_path), nit.ravel())
    root = ET.parse(base + ".xdmf").getroot()
    dom = root.find("Domain")
    items = {d.get("Name"): d for d in dom.findall("DataItem")}
    assert items["grid"].get("Dimensions").split() == ["3", "4"]
    assert items["grid"].text.strip() == "run_kernel.nc:/grid"
    assert items["points"].get("Dimensions").split() == ["6", "3"]
    grid = dom.find("Grid").find("Grid")
    assert grid.find("Topology").get("TopologyType") == "Tetrahedron"
    assert int(grid.find("Topology").get("NumberOfElements")) == 3
    attrs = grid.findall("Attribute")
    assert [a.get("Name") for a in attrs] == ["blockid", "K_x_a", "K_x_b", "nit_a", "nit_b"]
    assert [a.get("Center") for a in attrs] == ["Cell", "Node", "Node", "Cell", "Cell"]
    for a, (j, n, path) in zip(attrs[1:], [(0, 6, "node_data/kernel"), (1, 6, "node_data/kernel"),
                                          (0, 3, "cell_data/niterations"), (1, 3, "cell_data/niterations")]):
        slab, data = a.find("DataItem").findall("DataItem")
        rows = [r.split() for r in slab.text.strip().splitlines()]
        assert rows == [[str(j), "0"], ["1", "1"], ["1", str(n)]]
        assert data.get("Dimensions").split() == ["2", str(n)]
        assert data.text.strip() == "run_kernel.nc:/" + path
    with pytest.raises(dbfile.DbIoError):
        output.write_xdmf(base, "prism", verts, CONN, [])


def test_csr_binary_and_text(tmp_path):
    rng = np.random.default_rng(2)
    vals = (rng.standard_normal((3, 50)) * 10.0 ** rng.integers(-6, 2, (3, 50))).astype(np.float32)
    thr = 1e-3
    base = str(tmp_path / "k")
    output.write_csr(base, vals, thr)
    mask = np.abs(vals.astype(np.float64)) > thr
    np.testing.assert_array_equal(np.fromfile(base + ".val", np.float32), vals[mask])
    np.testing.assert_array_equal(np.fromfile(base + ".ind", np.int32), np.nonzero(mask)[1] + 1)
    pnt = open(base + ".pnt").read().splitlines()
    assert [int(x) for x in pnt] == list(np.cumsum(mask.sum(1)))
    assert all(len(x) == 12 for x in pnt)
    # text form: E15.8 strings as Fortran writes them
    known = np.array([[1.0, -0.5, 123456.78, 1e-10, 0.0, 9.99999999]], np.float32)
    output.write_csr(base + "_a", known, -1.0, ascii=True)
    line = open(base + "_a.val").read().splitlines()[0]
    fields = [line[i:i + 16] for i in range(0, len(line), 16)]
    assert fields == ["  0.10000000E+01", " -0.50000000E+00", "  0.12345678E+06", "  0.10000000E-09",
                      "  0.00000000E+00", "  0.10000000E+02"]
    ind = open(base + "_a.ind").read().splitlines()[0]
    assert [int(x) for x in ind.split()] == [1, 2, 3, 4, 5, 6] and len(ind) == 6 * 16


def test_intermediate_round_trip(tmp_path):
    rng = np.random.default_rng(3)
    npt, ne, nk = 20, 15, 4
    arrays = dict(kernel=rng.standard_normal((nk, npt)).astype(np.float32),
                  variance=rng.random((nk, npt)).astype(np.float32),
                  niterations=rng.integers(0, 64, (nk, ne)).astype(np.int32),
                  element_proc=rng.integers(0, 3, ne).astype(np.int32),
                  bg_model=rng.random((12, npt)).astype(np.float32))
    path = str(tmp_path / "intermediate_results.nc")
    output.write_intermediate(path, **arrays)
    got = output.read_intermediate(path, npt, ne, nk, nmodel=12)
    for key, v in arrays.items():
        np.testing.assert_array_equal(got[key], v, err_msg=key)
    np.testing.assert_array_equal(got["completed"], arrays["element_proc"] >= 1)
    with pytest.raises(dbfile.DbIoError):
        output.read_intermediate(path, npt + 1, ne, nk)
