"""Pins of the oracle's result assembly (extract_receive_buffer, master_queue.f90:253-323;
finalize's conversions, master_queue.f90:327-581):

  * hand-summed tasks on a mesh of three tetrahedra sharing vertices: each vertex's K_x,
    Var and Bg_Model are the sums of the basis-function results that touch it, empty
    slots (-1) add nothing, a second task that recomputes an element reassigns its
    niterations, computation time and rank;
  * volumetric mode: the element's own row;
  * the same sums by an independent scatter formulation (numpy add.at over the
    connectivity) on random values, and task order does not matter for exactly
    representable values;
  * the error variable at its special cases (|K| > 1e-30 strictly, negative K, K = 0),
    rounding to single precision, and the entry-major layout of the outputs.
"""
import numpy as np
import pytest

from mc_kernel_amd import abi

# three tetrahedra on 6 vertices (1-based), vertex 3 in all of them
CONN = np.array([[1, 2, 3, 4], [2, 3, 5, 4], [3, 6, 5, 1]], np.int32)


def _res(int_type=0, nk=2, nmodel=0, nbasis=4, conn=CONN, npoints=6, nelements=3):
    from oracle import oracle
    return oracle.Results(int_type, npoints, nelements, nbasis, nk, nmodel, conn if int_type == 0 else None)


def test_hand_sums_onvertices():
    r = _res(nmodel=12)
    # task 1: elements 1 and 3, a trailing empty slot
    vals = np.zeros((3, 4, 2))
    var = np.zeros((3, 4, 2))
    mod = np.zeros((3, 4, 12))
    for iel in range(2):
        for b in range(4):
            vals[iel, b] = [10 * (iel + 1) + b, -(10 * (iel + 1) + b)]
            var[iel, b] = [iel + 1, 2 * (iel + 1)]
            mod[iel, b] = 100 * (iel + 1) + b
    vals[2] = 1e6                                  # must be ignored (empty slot)
    nit = np.array([[3, 4], [5, 6], [9, 9]], np.int32)
    r.add([1, 3, -1], vals, var, nit, computation_time=[0.5, 0.25, 9.0], model=mod, rank=2)
    K = np.zeros((6, 2))
    V = np.zeros((6, 2))
    M = np.zeros((6, 12))
    for slot, e in enumerate([1, 3]):
        for b in range(4):
            v = CONN[e - 1, b] - 1
            K[v] += vals[slot, b]
            V[v] += var[slot, b]
            M[v] += mod[slot, b]
    np.testing.assert_array_equal(r.K, K)
    np.testing.assert_array_equal(r.V, V)
    np.testing.assert_array_equal(r.M, M)
    # vertex 3 (index 2) gets basis 2 of element 1 (slot 0) and basis 0 of element 3 (slot 1)
    np.testing.assert_array_equal(r.K[2], [12 + 20, -(12 + 20)])
    # vertex 2 (index 1) only belongs to elements 1 and 2: element 1's basis 1
    np.testing.assert_array_equal(r.K[1], [11, -11])
    np.testing.assert_array_equal(r.niter, [[3, 4], [0, 0], [5, 6]])
    np.testing.assert_array_equal(r.ctime, [0.5, 0.0, 0.25])
    np.testing.assert_array_equal(r.proc, [2, 0, 2])
    # task 2 recomputes element 3: sums grow, niterations/time/rank are reassigned
    r.add([3], np.ones((1, 4, 2)), np.ones((1, 4, 2)), np.array([[7, 8]], np.int32), [1.5],
          model=np.ones((1, 4, 12)), rank=5)
    np.testing.assert_array_equal(r.K[2], [33, -31])
    np.testing.assert_array_equal(r.niter[2], [7, 8])
    assert r.ctime[2] == 1.5 and r.proc[2] == 5 and r.proc[0] == 2


def test_volumetric_rows():
    r = _res(int_type=1, nbasis=1, npoints=3)
    vals = np.arange(6, dtype=float).reshape(3, 1, 2)
    r.add([2, -1, 3], vals, vals * 2, np.ones((3, 2), np.int32))
    np.testing.assert_array_equal(r.K, [[0, 0], [0, 1], [4, 5]])
    np.testing.assert_array_equal(r.V, [[0, 0], [0, 2], [8, 10]])


def test_out_of_range_rejected():
    r = _res()
    with pytest.raises(RuntimeError):
        r.add([4], np.zeros((1, 4, 2)), np.zeros((1, 4, 2)), np.zeros((1, 2), np.int32))


def test_scatter_formulation_and_order():
    rng = np.random.default_rng(3)
    nv, ne, nk = 40, 60, 5
    conn = np.array([rng.choice(nv, 4, replace=False) + 1 for _ in range(ne)], np.int32)
    r = _res(nk=nk, conn=conn, npoints=nv, nelements=ne)
    tasks = [rng.permutation(ne)[:20] + 1 for _ in range(4)]
    K = np.zeros((nv, nk))
    for t in tasks:
        vals = rng.standard_normal((len(t), 4, nk))
        r.add(t, vals, np.abs(vals), np.ones((len(t), nk), np.int32))
        np.add.at(K, conn[t - 1].ravel() - 1, vals.reshape(-1, nk))
    np.testing.assert_allclose(r.K, K, rtol=1e-13, atol=1e-13)
    # integer values: every order gives the same sums exactly
    outs = []
    for perm in (tasks, tasks[::-1]):
        q = _res(nk=nk, conn=conn, npoints=nv, nelements=ne)
        for t in perm:
            vals = np.outer(t, np.arange(1, 4 * nk + 1)).reshape(len(t), 4, nk).astype(float)
            q.add(t, vals, vals, np.ones((len(t), nk), np.int32))
        outs.append(q.K.copy())
    np.testing.assert_array_equal(outs[0], outs[1])


def test_export_conversions():
    r = _res(nk=3, npoints=6)
    r.K[:] = 0.0
    r.K[0] = [4.0, -4.0, 0.0]
    r.V[0] = [16.0, 16.0, 5.0]
    r.K[1] = [1e-30, -2e-30, 1.0 + 2.0 ** -30]
    r.V[1] = [7.0, 8e-30, 0.0]
    err = r.export(abi.MCK_RES_ERROR)
    assert err.shape == (3, 6) and err.dtype == np.float32
    np.testing.assert_array_equal(err[:, 0], [2.0, 2.0, 1000.0])
    assert err[0, 1] == 1000.0                     # |K| = 1e-30 is not > 1e-30
    assert err[1, 1] == np.float32(2.0)            # sqrt(8e-30 / 2e-30)
    assert err[2, 1] == 0.0
    kern = r.export(abi.MCK_RES_KERNEL)
    assert kern[2, 1] == np.float32(1.0)           # rounded to single precision
    np.testing.assert_array_equal(kern[:, 0], [4.0, -4.0, 0.0])
    np.testing.assert_array_equal(r.export(abi.MCK_RES_VARIANCE)[:, 0], [16.0, 16.0, 5.0])
    r.niter[1] = [1, 2, 3]
    ni = r.export(abi.MCK_RES_NITER)
    assert ni.shape == (3, 3)
    np.testing.assert_array_equal(ni[:, 1], [1, 2, 3])
