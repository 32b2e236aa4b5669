"""Heterogeneity integrals on the GPU (mck_set_hetero / mck_integrate_hetero,
integrate_3d_model slave.f90:874-924, load_het_rtpv / load_model_coeffs
heterogeneities.f90:63-188) against the oracle (pinned in tests/test_oracle_hetero.py,
its nearest point against the reference's own kd-tree): the grid search finds the same
nearest model points as the oracle's brute force, so the same stream positions give the
same number of MC steps per element and the same FP64 integrals, bit for bit, for
tetrahedra with 1 and 4 basis functions and voxels with 1 and 8; random clouds around
the elements and a regular (r, lat, lon) grid model; the integrals of the last call feed
the result accumulators from device memory; argument checks."""
import numpy as np
import pytest

from mc_kernel_amd import abi, synth

pytestmark = pytest.mark.gpu


def _rtp_of(xyz):
    xyz = np.asarray(xyz, float)
    r = np.linalg.norm(xyz, axis=1)
    lat = 90.0 - np.degrees(np.arccos(xyz[:, 2] / r))
    lon = np.degrees(np.arctan2(xyz[:, 1], xyz[:, 0]))
    return np.stack([r / 1000.0, lat, lon], 1)


def _model(kind, verts, seed):
    v = np.asarray(verts, float).reshape(-1, 3)
    rng = np.random.default_rng(seed)
    if kind == "cloud":
        a, b = v.min(0), v.max(0)
        rtp = _rtp_of(a + (b - a) * (-0.3 + 1.6 * rng.random((3000, 3))))
    else:
        rr = np.linalg.norm(v, axis=1)
        sph = _rtp_of(v)
        r = np.linspace(rr.min() / 1000 - 50, rr.max() / 1000 + 50, 12)
        lat = np.linspace(sph[:, 1].min() - 2, sph[:, 1].max() + 2, 25)
        lon = np.linspace(sph[:, 2].min() - 2, sph[:, 2].max() + 2, 25)
        rtp = np.stack(np.meshgrid(r, lat, lon, indexing="ij"), -1).reshape(-1, 3)
    dln = (rng.standard_normal((len(rtp), 3)) * 0.02).astype(np.float32)
    return rtp, dln


@pytest.mark.parametrize("elem,nbasis,kind", [("tet", 1, "cloud"), ("tet", 4, "cloud"), ("vox", 1, "cloud"),
                                              ("vox", 8, "grid"), ("tet", 1, "grid")])
def test_hetero_integrals_match_oracle(elem, nbasis, kind):
    import mc_kernel_amd as mk
    from oracle import oracle
    print("hetero case", elem, nbasis, kind, flush=True)
    p = synth.make_problem("tiny", ntet=6)
    verts, conn = (p.verts, p.conn) if elem == "tet" else p.voxels(4, seed=9)
    rtp, dln = _model(kind, verts, 3)
    ref, ns_ref, step_ref = oracle.integrate_hetero(rtp, dln, verts, conn, nbasis=nbasis, halton_step=0)
    k = mk.MCKernel(p.setup)
    try:
        k.set_hetero(rtp, dln)
        got, ns = k.integrate_hetero(verts, conn, nbasis=nbasis)
        step = k.halton_step
    finally:
        k.close()
    np.testing.assert_array_equal(ns, ns_ref)
    assert step == step_ref
    assert np.abs(ref).max() > 0
    np.testing.assert_array_equal(got, ref)


def test_hetero_into_results_and_checks():
    import mc_kernel_amd as mk
    p = synth.make_problem("tiny", ntet=6)
    rtp, dln = _model("cloud", p.verts, 4)
    k = mk.MCKernel(p.setup)
    try:
        with pytest.raises(mk.MckError):                  # no model set yet
            k.integrate_hetero(p.verts, p.conn)
        with pytest.raises(mk.MckError):                  # latitude outside [-90, 90]
            k.set_hetero([[6000.0, 95.0, 0.0]], [[0.0, 0.0, 0.0]])
        k.set_hetero(rtp, dln)
        het, _ = k.integrate_hetero(p.verts, p.conn, nbasis=4)
        out = k.slave_work(p.verts, p.conn, nbasis=4)
        ne = len(p.conn)
        nv = np.asarray(p.verts).reshape(-1, 3).shape[0]
        elems = np.arange(1, ne + 1, dtype=np.int32)
        a = k.results(0, nv, ne, 4, 0, p.conn, nhet=3)
        a.add(elems)                                      # slave_work and integrate_hetero from the device
        b = k.results(0, nv, ne, 4, 0, p.conn, nhet=3)
        b.add(elems, out["values"], out["variance"], out["niterations"], out["computation_time"], None, 1, het)
        ga, gb = a.get(), b.get()
        for key in ga:
            if ga[key] is not None:
                np.testing.assert_array_equal(ga[key], gb[key], err_msg=key)
        assert np.abs(ga["hetero"]).max() > 0
        np.testing.assert_array_equal(a.export(abi.MCK_RES_HETERO), ga["hetero"].T.astype(np.float32))
    finally:
        k.close()
