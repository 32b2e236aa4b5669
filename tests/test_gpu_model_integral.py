"""Model integrals (mck_integrate_model, integrate_1d_model slave.f90:815-869) on the GPU
against the FP64 oracle (pinned in tests/test_oracle_model_integral.py): the same stream
positions give the same number of MC steps per element and the same integrals, bit for
bit (FP64 on both sides, the same summation order), for tetrahedra with 1 and 4 basis functions and
voxels with 1 and 8; a constant model integrates to the parameters exactly; the stream
position advances past the task; empty slots (first vertex id <= 0) give zeros."""
import numpy as np
import pytest

from mc_kernel_amd import synth

pytestmark = pytest.mark.gpu


def _problem():
    return synth.make_problem("tiny", ntet=6)


@pytest.mark.parametrize("kind,nbasis", [("tet", 1), ("tet", 4), ("vox", 1), ("vox", 8)])
def test_model_integrals_match_oracle(kind, nbasis):
    import mc_kernel_amd as mk
    from oracle import oracle
    print("model integral case", kind, nbasis, flush=True)
    p = _problem()
    verts, conn = (p.verts, p.conn) if kind == "tet" else p.voxels(4, seed=9)
    ref, ns_ref, step_ref = oracle.Oracle(p.setup).integrate_model(verts, conn, nbasis=nbasis, halton_step=0)
    k = mk.MCKernel(p.setup)
    try:
        got, ns = k.integrate_model(verts, conn, nbasis=nbasis)
        step = k.halton_step
    finally:
        k.close()
    np.testing.assert_array_equal(ns, ns_ref)
    assert step == step_ref
    assert np.abs(ref).max() > 0
    # the same FP64 operations in the same order: bit for bit
    np.testing.assert_array_equal(got, ref)


def test_constant_model_and_empty_slot():
    import mc_kernel_amd as mk
    p = _problem()
    for i, val in enumerate((8000.0, 4500.0, 3300.0, 1.05, 0.95, 0.9)):
        p.model[i][:] = val
    conn = np.array(p.conn, np.int32)
    conn[-1] = 0                                      # an empty slot of a last task
    k = mk.MCKernel(p.setup)
    try:
        got, ns = k.integrate_model(p.verts, conn)
    finally:
        k.close()
    vp, vs, rho, phi, xi, eta = (np.float32(x) for x in (8000.0, 4500.0, 3300.0, 1.05, 0.95, 0.9))
    want = np.array([vp, vs, rho, vp, vp * np.sqrt(phi), vs, vs / np.sqrt(xi), eta, phi, xi,
                     rho * (vp * vp - np.float32(2) * (vs * vs)), rho * (vs * vs)], np.float32).astype(np.float64)
    np.testing.assert_array_equal(got[:-1, 0], np.tile(want, (len(conn) - 1, 1)))
    assert (ns[:-1] == 1).all()
    assert not got[-1].any() and ns[-1] == 0
