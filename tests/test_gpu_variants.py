"""Parity of the HIP path with the FP64 oracle over the source and strain variants the
reference supports: trace-only kernels, general moment tensors (all four simulations),
R and T receiver components, and the anisotropic base kernels (kernel.f90:514-570,
665-832).  Each case checks base kernels per point and a full pinned slave_work."""
import numpy as np
import pytest

from mc_kernel_amd import synth

pytestmark = pytest.mark.gpu

TOL = 1e-5
DC = [0.3e20, 1.0e20, -1.3e20, 0.5e20, -0.4e20, 0.7e20]   # general moment tensor

CASES = {
    "trace_vp": dict(kernels=synth._kernels(2, 2, ["CC", "AM"], ["vp", "lam"])),
    "general_mij_R": dict(mij=DC, component="R"),
    "general_mij_Z": dict(mij=DC, component="Z"),
    "explosion_T": dict(component="T"),
    "anisotropic": dict(kernels=synth._kernels(1, 2, ["CC"], ["vsh", "vsv", "vph", "vpv", "eta", "mu"]),
                        receivers_deg=[40.0]),
    # NO_DECONVOLVE_STF: the databases' time shifts instead of the STF deconvolution
    "no_deconvolution": dict(deconv_stf=0),
}


@pytest.fixture(scope="module", params=sorted(CASES))
def case(request):
    import mc_kernel_amd as mk
    from oracle import oracle
    p = synth.make_problem("tiny", ntet=10, **CASES[request.param])
    o = oracle.Oracle(p.setup)
    k = mk.MCKernel(p.setup)
    yield request.param, p, o, k
    k.close()


def _tet_around(c, a=60e3):
    c = np.asarray(c, float)
    return c + a * np.array([[1, 1, 1], [1, -1, -1], [-1, 1, -1], [-1, -1, 1]], float)


def test_masks_and_damping():
    """MASK_SOURCE_RECEIVER (slave.f90:466-482, 539-545) and DAMP_RADIUS_SOURCE_RECEIVER
    (dampen_field after every processed receiver, slave.f90:701-706): tets around the
    source and a receiver are added to the task."""
    import mc_kernel_amd as mk
    from oracle import oracle
    p = synth.make_problem("tiny", ntet=8)
    p.setup.mask_src_rec = 1
    p.setup.damp_radius = 2500e3
    src = synth.geo_xyz(p.src_lat, p.src_lon, synth.R_EARTH - p.src_depth * 1e3)
    rec = synth.geo_xyz(p.rec_lat[0], p.rec_lon[0], synth.R_EARTH - 30e3)
    verts = np.concatenate([np.asarray(p.verts, float).reshape(-1, 3), _tet_around(src), _tet_around(rec)])
    n0 = np.asarray(p.verts).reshape(-1, 3).shape[0]
    conn = np.concatenate([np.asarray(p.conn, np.int32),
                           np.arange(n0 + 1, n0 + 5, dtype=np.int32)[None],
                           np.arange(n0 + 5, n0 + 9, dtype=np.int32)[None]])
    o = oracle.Oracle(p.setup)
    ref = o.slave_work(verts, conn)
    assert not np.any(ref["values"][-2]) and not np.any(ref["niterations"][-2])   # source tet
    k = mk.MCKernel(p.setup)
    try:
        res = k.slave_work(verts, conn, halton_offsets=ref["offsets"])
    finally:
        k.close()
    assert np.array_equal(res["niterations"], ref["niterations"])
    assert res["stats"]["evaluations"] == ref["evaluations"]
    for j in range(p.nkernel):
        assert _rel(res["values"][:, j], ref["values"][:, j]) < TOL, j


def _rel(g, r):
    n = np.linalg.norm(r)
    return np.linalg.norm(g - r) / n if n else np.linalg.norm(g)


def test_base_kernels_variant(case):
    name, p, o, k = case
    x = p.sample_points(24, seed=41)
    g = k.base_kernels(x)
    r = o.base_kernels(x)
    for j in range(p.nkernel):
        for b in range(6):
            if not np.any(r[:, j, b]):
                assert not np.any(g[:, j, b]), (name, j, b)
                continue
            assert _rel(g[:, j, b], r[:, j, b]) < TOL, (name, j, b, _rel(g[:, j, b], r[:, j, b]))


def test_slave_work_variant(case):
    name, p, o, k = case
    ref = o.slave_work(p.verts, p.conn)
    res = k.slave_work(p.verts, p.conn, halton_offsets=ref["offsets"])
    assert np.array_equal(res["niterations"], ref["niterations"]), name
    assert res["stats"]["evaluations"] == ref["evaluations"], name
    for j in range(p.nkernel):
        assert _rel(res["values"][:, j], ref["values"][:, j]) < TOL, (name, j)


@pytest.mark.parametrize("nbasis", [1, 8])
def test_voxel_elements(nbasis):
    """Voxel inversion elements (random_voxel.f90): pseudo-random points in spherical
    blocks, block volume and (as written) centre for the travel-time mask; pinned to the
    oracle's stream positions.  With 8 basis functions every weight is 1."""
    import mc_kernel_amd as mk
    from oracle import oracle
    p = synth.make_problem("tiny", ntet=2, allowed_error=1e-30, allowed_relative_error=0.05)
    verts, conn = p.voxels(8, seed=11)
    o = oracle.Oracle(p.setup)
    ref = o.slave_work(verts, conn, nbasis=nbasis)
    assert ref["niterations"].max() > 2          # several MC steps
    k = mk.MCKernel(p.setup)
    try:
        res = k.slave_work(verts, conn, halton_offsets=ref["offsets"], nbasis=nbasis)
    finally:
        k.close()
    assert res["values"].shape == ref["values"].shape
    assert np.array_equal(res["niterations"], ref["niterations"])
    v, r = res["values"].reshape(len(conn), -1), ref["values"].reshape(len(conn), -1)
    assert np.abs(r).max() > 0
    err = np.linalg.norm(v - r) / np.linalg.norm(r)
    assert err < TOL, err
    if nbasis == 8:
        assert np.array_equal(res["values"][:, 0], res["values"][:, 7])


def test_pseudorandom_tetrahedra():
    """USE_PSEUDORANDOM_NUMBERS: the pseudo-random stream in tetrahedra, pinned to the
    oracle's stream positions."""
    import mc_kernel_amd as mk
    from oracle import oracle
    p = synth.make_problem("tiny", ntet=8)
    p.setup.quasirandom = 0
    o = oracle.Oracle(p.setup)
    ref = o.slave_work(p.verts, p.conn)
    k = mk.MCKernel(p.setup)
    try:
        res = k.slave_work(p.verts, p.conn, halton_offsets=ref["offsets"])
    finally:
        k.close()
    assert np.array_equal(res["niterations"], ref["niterations"])
    err = np.linalg.norm(res["values"] - ref["values"]) / np.linalg.norm(ref["values"])
    assert err < TOL, err


def test_c2_kernel_set():
    """The c2 kernel set: 50 receivers x (5 Gabor + Butterworth band pass) x CC/AM x vp/vs
    = 1200 kernels, 72 (window spectrum, base kernel) slots per receiver; reduced
    database and task so that the oracle runs in seconds."""
    import mc_kernel_amd as mk
    from oracle import oracle
    # receivers 10..34.5 deg: every kernel window (at most 1.8 x P time + 70 s) inside the
    # 1024-sample traces
    p = synth.make_problem("c2", nelem_target=2000, nt=1024, ntet=2, npoints_per_step=8, max_iter=2,
                           unmasked_frac=1.0, receivers_deg=[10.0 + 0.5 * i for i in range(50)])
    assert p.nkernel == 1200
    o = oracle.Oracle(p.setup)
    x = p.sample_points(4, seed=9)
    ref_b = o.base_kernels(x)
    ref = o.slave_work(p.verts, p.conn)
    k = mk.MCKernel(p.setup)
    try:
        g_b = k.base_kernels(x)
        res = k.slave_work(p.verts, p.conn, halton_offsets=ref["offsets"])
    finally:
        k.close()
    err_b = np.linalg.norm(g_b - ref_b) / np.linalg.norm(ref_b)
    assert err_b < TOL, err_b
    assert np.array_equal(res["niterations"], ref["niterations"])
    assert res["stats"]["evaluations"] == ref["evaluations"]
    err = np.linalg.norm(res["values"] - ref["values"]) / np.linalg.norm(ref["values"])
    assert err < TOL, err
