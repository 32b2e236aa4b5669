"""Seismograms by reciprocity (mck_setup_t.seis = NULL: mck_init computes them from the
backward database, load_seismogram / load_fw_points_rdbm, readfields.f90:1004-1048,
1469-1604) against the FP64 oracle (pinned in tests/test_oracle_seismogram.py):

  * the seismograms themselves, FP64 on both sides, to 1e-10 of their maximum, for Z, R
    and T receivers of an explosion and of a general moment tensor;
  * a whole pinned slave_work whose kernel windows and normalisations use them;
  * paged backward database: the same seismograms bit for bit;
  * a set-up with seismograms: mck_seismograms returns them unchanged.
"""
import numpy as np
import pytest

from mc_kernel_amd import synth

pytestmark = pytest.mark.gpu

DC = [0.3e20, 1.0e20, -1.3e20, 0.5e20, -0.4e20, 0.7e20]
CASES = {
    "explosion_Z": dict(component="Z"),
    "explosion_R": dict(component="R"),
    "general_Z": dict(mij=DC, component="Z"),
    "general_R": dict(mij=DC, component="R"),
    "general_T": dict(mij=DC, component="T"),
}


def _problem(**kw):
    p = synth.make_problem("tiny", ntet=8, receivers_deg=[25.0, 40.0, 70.0],
                           kernels=synth._kernels(3, 2, ["CC", "AM"], ["vp", "vs"]), **kw)
    p.setup.seis = None
    return p


@pytest.mark.parametrize("case", sorted(CASES))
def test_seismograms_match_oracle(case):
    import mc_kernel_amd as mk
    from oracle import oracle
    print("seismogram case", case, flush=True)
    p = _problem(**CASES[case])
    ref = oracle.Oracle(p.setup).seismograms()
    k = mk.MCKernel(p.setup)
    try:
        got = k.seismograms()
    finally:
        k.close()
    scale = np.abs(ref).max()
    assert scale > 0
    np.testing.assert_allclose(got, ref, rtol=0, atol=1e-10 * scale)


def test_slave_work_with_reciprocal_seismograms():
    import mc_kernel_amd as mk
    from oracle import oracle
    p = _problem(mij=DC, component="R")
    ref = oracle.Oracle(p.setup).slave_work(p.verts, p.conn)
    k = mk.MCKernel(p.setup)
    try:
        got = k.slave_work(p.verts, p.conn, halton_offsets=ref["offsets"])
    finally:
        k.close()
    np.testing.assert_array_equal(got["niterations"], ref["niterations"])
    for j in range(p.nkernel):
        r = ref["values"][:, j]
        n = np.linalg.norm(r)
        if n:
            assert np.linalg.norm(got["values"][:, j] - r) / n < 1e-5, j


def test_paged_database_same_seismograms():
    import mc_kernel_amd as mk
    out = []
    for page in (0, 37):
        p = _problem(mij=DC, component="Z")
        p.setup.db_page_elems = page
        k = mk.MCKernel(p.setup)
        try:
            out.append(k.seismograms())
        finally:
            k.close()
    np.testing.assert_array_equal(out[1], out[0])


def test_given_seismograms_are_returned():
    import mc_kernel_amd as mk
    p = synth.make_problem("tiny", ntet=2)
    k = mk.MCKernel(p.setup)
    try:
        np.testing.assert_array_equal(k.seismograms(), p.seis)
    finally:
        k.close()
