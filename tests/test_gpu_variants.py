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
