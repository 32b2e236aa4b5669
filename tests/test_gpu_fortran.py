"""A Fortran program drives the library end to end through the iso_c_binding module
(mc_kernel_amd/fortran/mck_gpu.f90), as the reference's do_slave would (slave.f90:94-219):
the tiny problem's set-up is written as raw arrays, the program (tests/fortran/slave_e2e,
built by the package build) reads them, fills mck_setup_t with c_loc of them, runs
mck_init, one task through mck_slave_work and mck_slave_work_sequential, and writes the
results.  Checks: the batched results equal the same task driven from Python bit for bit
and the oracle on the same stream positions to 1e-5; the sequential run walks the
reference's own Halton stream (element starts, iteration counts and end position equal to
the oracle's, kernels to 1e-5)."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

from mc_kernel_amd import abi, synth

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
EXE = os.path.join(HERE, "fortran", "slave_e2e")
TOL = 1e-5


def _arrays(p):
    """(field, address, bytes) of every non-null array of the set-up"""
    s = p.setup
    out = []
    for w, db in (("fwd", s.fwd), ("bwd", s.bwd)):
        sizes = {"mesh_s": 4 * db.npoints, "mesh_z": 4 * db.npoints, "mp_s": 4 * db.nelem, "mp_z": 4 * db.nelem,
                 "fem_mesh": 16 * db.nelem, "eltype": 4 * db.nelem, "axis": 4 * db.nelem,
                 "snapshots": 4 * db.nelem * db.nvar * (db.npol + 1) ** 2 * s.nt}
        for f, n in sizes.items():
            out.append(("%s.%s" % (w, f), getattr(db, f), n))
    nk, nr, nd = s.nkernel, s.nrec, s.model_ndepth
    sizes = {"stf_fwd": 8 * s.nt, "stf_src": 8 * s.nt, "rec_lat": 8 * nr, "rec_lon": 8 * nr, "seis": 8 * nr * s.nt,
             "filters": C.sizeof(abi.MckFilter) * s.nfilter, "kernel_receiver": 4 * nk, "kernel_filter": 4 * nk,
             "kernel_misfit": 4 * nk, "kernel_param": 4 * nk, "kernel_window": 16 * nk}
    for m in ("vp", "vs", "rho", "phi", "xi", "eta"):
        sizes["model_" + m] = 4 * nd
    for f, n in sizes.items():
        out.append((f, getattr(s, f), n))
    return [(f, C.cast(ptr, C.c_void_p).value, n) for f, ptr, n in out if C.cast(ptr, C.c_void_p).value]


def _write_problem(p, d):
    blob = abi.MckSetup.from_buffer_copy(p.setup)
    arrays = _arrays(p)
    for f, _, _ in arrays:                     # pointer fields travel as separate files
        if "." in f:
            w, g = f.split(".")
            setattr(getattr(blob, w), g, None)
        else:
            setattr(blob, f, None)
    with open(os.path.join(d, "setup.bin"), "wb") as fh:
        fh.write(bytes(blob))
    with open(os.path.join(d, "manifest.txt"), "w") as fh:
        for f, addr, n in arrays:
            fh.write("%s %d\n" % (f, n))
            with open(os.path.join(d, f + ".bin"), "wb") as g:
                g.write(C.string_at(addr, n))
    verts = np.ascontiguousarray(np.asarray(p.verts, np.float64).reshape(-1, 3))
    conn = np.ascontiguousarray(np.asarray(p.conn, np.int32))
    with open(os.path.join(d, "task.txt"), "w") as fh:
        fh.write("%d %d %d %d\n" % (len(verts), len(conn), conn.shape[1], 1))
    verts.tofile(os.path.join(d, "vertices.bin"))       # = Fortran vertices(3, nvert)
    conn.tofile(os.path.join(d, "conn.bin"))            # = Fortran conn(nvpe, nelems)
    return verts, conn


def _rel(g, r):
    n = np.linalg.norm(r)
    return np.linalg.norm(g - r) / n if n else np.linalg.norm(g)


def test_fortran_slave_end_to_end(tmp_path):
    if not os.path.exists(EXE):
        from mc_kernel_amd import build
        if build.build_fortran_e2e(force=False) is None:
            pytest.skip("no Fortran compiler")
    import mc_kernel_amd as mk
    from oracle import oracle
    p = synth.make_problem("tiny", ntet=10)
    verts, conn = _write_problem(p, str(tmp_path))
    print("fortran e2e: tiny, %d tets, %d kernels" % (len(conn), p.nkernel), flush=True)
    run = subprocess.run([EXE, str(tmp_path)], capture_output=True, text=True, timeout=300)
    assert run.returncode == 0, run.stdout + run.stderr
    assert "slave_e2e done" in run.stdout
    ne, nk = len(conn), p.nkernel
    rd = lambda f, t: np.fromfile(os.path.join(str(tmp_path), f), t)
    vals = rd("values.bin", np.float64).reshape(ne, nk)
    var = rd("variance.bin", np.float64).reshape(ne, nk)
    nit = rd("niterations.bin", np.int32).reshape(ne, nk)
    halton, evals = rd("halton.bin", np.int64)
    # the same task driven from Python: bit for bit
    k = mk.MCKernel(p.setup)
    try:
        ref_py = k.slave_work(verts, conn)
    finally:
        k.close()
    assert np.array_equal(vals, ref_py["values"]) and np.array_equal(var, ref_py["variance"])
    assert np.array_equal(nit, ref_py["niterations"])
    assert halton == ref_py["halton_step"] and evals == ref_py["stats"]["evaluations"]
    # the oracle on the batched form's stream positions
    P, mi = p.setup.npoints_per_step, p.setup.max_iter
    o = oracle.Oracle(p.setup)
    ref = o.slave_work(verts, conn, halton_offsets=np.arange(ne, dtype=np.int64) * (mi + 1) * P)
    assert np.array_equal(nit, ref["niterations"]) and evals == ref["evaluations"] > 0
    for j in range(nk):
        assert _rel(vals[:, j], ref["values"][:, j]) < TOL, j
    # the sequential run: the reference's own stream from position 0
    o2 = oracle.Oracle(p.setup)
    seq_ref = o2.slave_work(verts, conn)
    assert np.array_equal(rd("seq_offsets.bin", np.int64), seq_ref["offsets"])
    assert rd("seq_halton.bin", np.int64)[0] == seq_ref["halton_step"]
    assert np.array_equal(rd("seq_niterations.bin", np.int32).reshape(ne, nk), seq_ref["niterations"])
    sv = rd("seq_values.bin", np.float64).reshape(ne, nk)
    for j in range(nk):
        assert _rel(sv[:, j], seq_ref["values"][:, j]) < TOL, j
