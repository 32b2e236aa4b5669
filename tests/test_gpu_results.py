"""Result assembly on the GPU (mck_results_*, extract_receive_buffer master_queue.f90:
253-323 and finalize's conversions master_queue.f90:327-581) against the oracle (pinned
in tests/test_oracle_results.py), bit for bit: both add each point's contributions to
its FP64 accumulator one at a time in task order.

  * several random tasks with shared vertices, empty slots and recomputed elements,
    4 basis functions, model and heterogeneity integrals: accumulators and all six
    exported variables;
  * volumetric mode; an odd kernel count;
  * the results of the last slave_work / integrate_model taken from device memory give
    the same accumulators as the same results passed from the host; device tensors as
    inputs;
  * load / get round trip (resume), argument checks;
  * finalize's files from the device accumulators of a slave_work task, and a resume
    from the intermediate results written from them.
"""
import os

import numpy as np
import pytest

from mc_kernel_amd import abi, synth

pytestmark = pytest.mark.gpu

EXPORTS = [abi.MCK_RES_KERNEL, abi.MCK_RES_VARIANCE, abi.MCK_RES_ERROR, abi.MCK_RES_MODEL, abi.MCK_RES_NITER,
           abi.MCK_RES_HETERO]


def _kernel(odd=False):
    """The tiny problem (16 kernels), or with 3 kernels (odd column counts take the
    one-column paths of the accumulation kernels)."""
    import mc_kernel_amd as mk
    over = {}
    if odd:
        over = dict(receivers_deg=[30.0], kernels=synth._kernels(1, 1, ["CC"], ["vp", "vs"])
                    + synth._kernels(1, 1, ["AM"], ["vp"]))
    p = synth.make_problem("tiny", ntet=4, **over)
    return mk.MCKernel(p.setup), p


def _tasks(rng, ne, nk, nb, ntask=5, size=40):
    out = []
    for _ in range(ntask):
        t = (rng.permutation(ne)[:size] + 1).astype(np.int32)
        t[rng.random(size) < 0.1] = -1
        vals = rng.standard_normal((size, nb, nk)) * 10.0 ** rng.integers(-3, 3, (size, nb, 1))
        var = np.abs(rng.standard_normal((size, nb, nk)))
        nit = rng.integers(1, 64, (size, nk)).astype(np.int32)
        ct = rng.random(size)
        model = rng.standard_normal((size, nb, 12))
        het = rng.standard_normal((size, nb, 3)) * 1e-2
        out.append((t, vals, var, nit, ct, model, int(rng.integers(1, 9)), het))
    return out


def _compare(res, ref):
    got = res.get()
    np.testing.assert_array_equal(got["kernel"], ref.K)
    np.testing.assert_array_equal(got["variance"], ref.V)
    if ref.nm:
        np.testing.assert_array_equal(got["model"], ref.M)
    if ref.nh:
        np.testing.assert_array_equal(got["hetero"], ref.H)
    np.testing.assert_array_equal(got["niterations"], ref.niter)
    np.testing.assert_array_equal(got["computation_time"], ref.ctime)
    np.testing.assert_array_equal(got["element_proc"], ref.proc)
    for w in EXPORTS:
        if (w == abi.MCK_RES_MODEL and not ref.nm) or (w == abi.MCK_RES_HETERO and not ref.nh):
            continue
        np.testing.assert_array_equal(res.export(w), ref.export(w), err_msg=str(w))


@pytest.mark.parametrize("int_type,odd", [(0, False), (1, False), (0, True)])
def test_random_tasks_match_oracle(int_type, odd):
    from oracle import oracle
    print("results case int_type", int_type, "odd", odd, flush=True)
    k, p = _kernel(odd)
    try:
        rng = np.random.default_rng(11 + int_type)
        nv, ne, nk = 300, 500, k.nk
        nb = 4 if int_type == 0 else 1
        npt = nv if int_type == 0 else ne
        conn = np.array([rng.choice(nv, 4, replace=False) + 1 for _ in range(ne)], np.int32)
        c = conn if int_type == 0 else None
        res = k.results(int_type, npt, ne, nb, 12, c, nhet=3)
        ref = oracle.Results(int_type, npt, ne, nb, nk, 12, c, nhet=3)
        for t, vals, var, nit, ct, model, rank, het in _tasks(rng, ne, nk, nb):
            res.add(t, vals, var, nit, ct, model, rank, het)
            ref.add(t, vals, var, nit, ct, model, rank, het)
        _compare(res, ref)
    finally:
        k.close()


def test_device_resident_results():
    k, p = _kernel()
    try:
        out = k.slave_work(p.verts, p.conn, nbasis=4)
        model, _ = k.integrate_model(p.verts, p.conn, nbasis=4)
        ne = len(p.conn)
        nv = np.asarray(p.verts).reshape(-1, 3).shape[0]
        elems = np.arange(1, ne + 1, dtype=np.int32)
        a = k.results(0, nv, ne, 4, 12, p.conn)
        a.add(elems, rank=3)                              # last slave_work + integrate_model
        ga = a.get()
        b = k.results(0, nv, ne, 4, 12, p.conn)
        b.add(elems, out["values"], out["variance"], out["niterations"], out["computation_time"], model, rank=3)
        gb = b.get()
        for key in ga:
            np.testing.assert_array_equal(ga[key], gb[key], err_msg=key)
        assert np.abs(ga["kernel"]).max() > 0
        # device tensors as inputs
        import torch
        dev = [torch.from_numpy(np.ascontiguousarray(x)).cuda() for x in
               (out["values"], out["variance"], out["niterations"], out["computation_time"], model)]
        torch.cuda.synchronize()
        c = k.results(0, nv, ne, 4, 12, p.conn)
        c.add(elems, *dev, rank=3)
        gc = c.get()
        for key in ga:
            np.testing.assert_array_equal(gc[key], gb[key], err_msg=key)
    finally:
        k.close()


def test_load_get_and_checks():
    import mc_kernel_amd as mk
    k, p = _kernel()
    try:
        rng = np.random.default_rng(5)
        res = k.results(1, 50, 50, 1, 0)
        arrays = dict(kernel=rng.standard_normal((50, k.nk)), variance=rng.random((50, k.nk)),
                      niterations=rng.integers(0, 9, (50, k.nk)).astype(np.int32),
                      computation_time=rng.random(50), element_proc=rng.integers(1, 4, 50).astype(np.int32))
        res.load(**arrays)
        got = res.get()
        for key, v in arrays.items():
            np.testing.assert_array_equal(got[key], v, err_msg=key)
        # export straight into device memory
        import torch
        dev = torch.empty((k.nk, 50), dtype=torch.float32, device="cuda")
        torch.cuda.synchronize()
        mk._check(k.lib.mck_results_export(abi.MCK_RES_ERROR, dev.data_ptr()))
        np.testing.assert_array_equal(dev.cpu().numpy(), res.export(abi.MCK_RES_ERROR))
        with pytest.raises(mk.MckError):                  # element id out of range
            res.add([51], np.zeros((1, 1, k.nk)), np.zeros((1, 1, k.nk)), np.zeros((1, k.nk), np.int32))
        with pytest.raises(mk.MckError):                  # no slave_work of that size yet
            res.add(np.arange(1, 8, dtype=np.int32))
        with pytest.raises(mk.MckError):                  # no model integrals accumulated
            res.export(abi.MCK_RES_MODEL)
        with pytest.raises(mk.MckError):                  # onvertices needs the connectivity
            k.results(0, 10, 10, 4, 0)
    finally:
        k.close()


def test_finalize_files_and_resume(tmp_path):
    """slave_work -> device accumulators -> finalize's files; the intermediate results
    written from them resume into fresh accumulators (read_intermediate converts the
    single-precision file back to double, master_queue.f90:628-650)."""
    import shutil
    import subprocess
    from mc_kernel_amd import output
    k, p = _kernel()
    try:
        k.slave_work(p.verts, p.conn, nbasis=4)
        ne = len(p.conn)
        verts = np.asarray(p.verts, float).reshape(-1, 3)
        nv = verts.shape[0]
        res = k.results(0, nv, ne, 4, 0, p.conn)
        res.add(np.arange(1, ne + 1, dtype=np.int32), rank=1)
        names = ["k%d" % i for i in range(k.nk)]
        base = str(tmp_path / "run")
        var = output.finalize(res, base, names, "tet", verts, p.conn)
        assert [v[0] for v in var] == ["kernel", "error", "niterations", "computation_time", "element_proc", "volume"]
        assert os.path.exists(base + "_kernel.xdmf")
        h5dump = shutil.which("h5dump") or "/opt/conda/bin/h5dump"
        if os.path.exists(h5dump):
            out = str(tmp_path / "k.bin")
            subprocess.run([h5dump, "-d", "/node_data/kernel", "-b", "LE", "-o", out, base + "_kernel.nc"],
                           check=True, capture_output=True)
            np.testing.assert_array_equal(np.fromfile(out, np.float32), res.export(abi.MCK_RES_KERNEL).ravel())
        raw = res.get()
        path = str(tmp_path / "intermediate_results.nc")
        output.write_intermediate(path, res.export(abi.MCK_RES_KERNEL), res.export(abi.MCK_RES_VARIANCE),
                                  raw["niterations"].T, raw["element_proc"])
        back = output.read_intermediate(path, nv, ne, k.nk)
        assert back["completed"].all()
        res2 = k.results(0, nv, ne, 4, 0, p.conn)
        res2.load(kernel=back["kernel"].T.astype(np.float64), variance=back["variance"].T.astype(np.float64),
                  niterations=back["niterations"].T, element_proc=back["element_proc"])
        got = res2.get()
        np.testing.assert_array_equal(got["kernel"], raw["kernel"].astype(np.float32).astype(np.float64))
        np.testing.assert_array_equal(got["niterations"], raw["niterations"])
        np.testing.assert_array_equal(res2.export(abi.MCK_RES_KERNEL), res.export(abi.MCK_RES_KERNEL))
    finally:
        k.close()


def test_run_job_tasks_equal_one_task(tmp_path):
    """run_job over tasks of 3 elements gives the accumulators of one task holding every
    element (each element keeps its own stream range) and writes finalize's files; with
    the background-model integrals the model accumulators fill as well."""
    import mc_kernel_amd as mk
    from mc_kernel_amd.job import run_job
    p = synth.make_problem("tiny", ntet=8)
    k = mk.MCKernel(p.setup)
    try:
        a = run_job(k, p.verts, p.conn, nbasis=4, nelems_per_task=3, output_file=str(tmp_path / "job"))
        ga = a.get()
        k.halton_step = 0
        b = run_job(k, p.verts, p.conn, nbasis=4, nelems_per_task=len(p.conn))
        gb = b.get()
        c = run_job(k, p.verts, p.conn, nbasis=4, nelems_per_task=3, int_over_background=True)
        gc = c.get()
    finally:
        k.close()
    for key in ("kernel", "variance", "niterations", "element_proc"):
        np.testing.assert_array_equal(ga[key], gb[key], err_msg=key)
    assert np.abs(ga["kernel"]).max() > 0 and (ga["element_proc"] == 1).all()
    assert os.path.exists(str(tmp_path / "job_kernel.xdmf"))
    assert np.abs(gc["model"]).max() > 0


def test_device_resident_skipped_slot():
    """A task slot the slave skipped (first vertex id 0) holds zeros in device memory as
    in the host arrays, so both paths of mck_results_add agree for it."""
    k, p = _kernel()
    try:
        conn = np.array(p.conn, np.int32)
        conn[-1] = 0
        out = k.slave_work(p.verts, conn)
        ne = len(conn)
        elems = np.arange(1, ne + 1, dtype=np.int32)
        a = k.results(1, ne, ne, 1, 0)
        a.add(elems)
        b = k.results(1, ne, ne, 1, 0)
        b.add(elems, out["values"], out["variance"], out["niterations"], out["computation_time"])
        ga, gb = a.get(), b.get()
        for key in ("kernel", "variance", "niterations", "computation_time", "element_proc"):
            np.testing.assert_array_equal(ga[key], gb[key], err_msg=key)
        assert not ga["kernel"][-1].any() and np.abs(ga["kernel"][:-1]).max() > 0
    finally:
        k.close()


def test_run_job_voxels_with_model_and_hetero(tmp_path):
    """run_job over voxels with 8 basis functions (onvertices), the background-model and
    heterogeneity integrals and finalize's files: tasks of 3 voxels give the kernel
    accumulators of one task (model integrals off), and the model and heterogeneity
    accumulators fill when on."""
    import mc_kernel_amd as mk
    from mc_kernel_amd.job import run_job
    p = synth.make_problem("tiny", ntet=4)
    vv, vc = p.voxels(7, seed=2)
    v = np.asarray(vv, float).reshape(-1, 3)
    rng = np.random.default_rng(1)
    a_, b_ = v.min(0), v.max(0)
    xyz = a_ + (b_ - a_) * (-0.2 + 1.4 * rng.random((500, 3)))
    r = np.linalg.norm(xyz, axis=1)
    rtp = np.stack([r / 1000.0, 90.0 - np.degrees(np.arccos(xyz[:, 2] / r)),
                    np.degrees(np.arctan2(xyz[:, 1], xyz[:, 0]))], 1)
    dln = (rng.standard_normal((500, 3)) * 0.03).astype(np.float32)
    k = mk.MCKernel(p.setup)
    try:
        a = run_job(k, vv, vc, nbasis=8, nelems_per_task=3).get()
        k.halton_step = 0
        b = run_job(k, vv, vc, nbasis=8, nelems_per_task=len(vc)).get()
        c = run_job(k, vv, vc, nbasis=8, nelems_per_task=3, int_over_background=True, hetero=(rtp, dln),
                    output_file=str(tmp_path / "vox"))
        gc = c.get()
    finally:
        k.close()
    for key in ("kernel", "variance", "niterations"):
        np.testing.assert_array_equal(a[key], b[key], err_msg=key)
    assert np.abs(a["kernel"]).max() > 0
    assert np.abs(gc["model"]).max() > 0 and np.abs(gc["hetero"]).max() > 0
    assert os.path.exists(str(tmp_path / "vox_kernel.nc"))
