"""Pins of the oracle's model integrals (integrate_1d_model, slave.f90:815-869): the MC
integral (volume 1, errors 1e-2) of the 12 background-model parameters (vp vs rho vph vpv
vsh vsv eta phi xi lam mu, background_model.f90:34-37, combined in fp32) per element and
basis function, 100 Halton points per step in tetrahedra, pseudo-random in voxels.

  * a constant model: every integral is the parameter (fp32 combination) exactly, after
    one step (zero variance) - for tetrahedra and voxels;
  * on-vertex basis functions with a constant model: integral b = c x the mean of the
    barycentric weight b over exactly the points drawn (the same Halton stream), and the
    integrator of b stops once its relative error is below 1e-2;
  * a model varying with radius: the integral is the volume average within the MC error
    (a fine deterministic quadrature of the same tetrahedron);
  * stream positions: element e starts at halton_step + e x 9999 x 100 and the step
    advances past the task.
"""
import numpy as np

from mc_kernel_amd import synth

NP = 100


def _problem(const=True, **kw):
    p = synth.make_problem("tiny", ntet=3, **kw)
    if const:
        for i, val in enumerate((8000.0, 4500.0, 3300.0, 1.05, 0.95, 0.9)):
            p.model[i][:] = val
    return p


def _expected_const():
    vp, vs, rho, phi, xi, eta = (np.float32(x) for x in (8000.0, 4500.0, 3300.0, 1.05, 0.95, 0.9))
    lam = rho * (vp * vp - np.float32(2) * (vs * vs))
    mu = rho * (vs * vs)
    vsv = vs / np.sqrt(xi)
    vpv = vp * np.sqrt(phi)
    return np.array([vp, vs, rho, vp, vpv, vs, vsv, eta, phi, xi, lam, mu], np.float32).astype(np.float64)


def test_constant_model_tetrahedra_and_voxels():
    from oracle import oracle
    p = _problem()
    o = oracle.Oracle(p.setup)
    m, nsteps, _ = o.integrate_model(p.verts, p.conn)
    np.testing.assert_array_equal(m[:, 0], np.tile(_expected_const(), (len(p.conn), 1)))
    assert (nsteps == 1).all()
    vv, vc = p.voxels(2, seed=4)
    m, nsteps, _ = o.integrate_model(vv, vc)
    np.testing.assert_array_equal(m[:, 0], np.tile(_expected_const(), (2, 1)))
    assert (nsteps == 1).all()


def test_vertex_basis_constant_model_is_mean_weight():
    from oracle import oracle
    p = _problem()
    o = oracle.Oracle(p.setup)
    m, nsteps, _ = o.integrate_model(p.verts, p.conn, nbasis=4, halton_step=0)
    c = _expected_const()
    verts = np.asarray(p.verts, float).reshape(-1, 3)
    for e in range(len(p.conn)):
        v = verts[np.asarray(p.conn[e]) - 1]
        pos = e * 9999 * NP
        n = int(nsteps[e]) * NP
        pts = oracle.tet_points(v, oracle.halton(n, step=pos))
        w, bad = oracle.tet_weights(v, pts)                  # [n][4]
        assert not bad
        np.testing.assert_allclose(w.sum(1), 1.0, atol=1e-12)
        for b in range(4):
            # each (basis, parameter) integrator stops at the first step where its
            # variance is below 1e-4 (absolute) or its relative error below 1e-2; its
            # integral is c times the mean weight over the points up to that step
            wb = w[:, b]
            for j in range(12):
                done = None
                for s in range(1, int(nsteps[e]) + 1):
                    k = s * NP
                    f = c[j] * wb[:k]
                    var = (np.mean(f * f) - np.mean(f) ** 2) / (k - 1)   # variance of the mean
                    if var < 1e-4 or np.sqrt(var) < 1e-2 * abs(np.mean(f)):
                        done = k
                        break
                assert done is not None, (b, j)
                np.testing.assert_allclose(m[e, b, j], c[j] * np.mean(wb[:done]), rtol=1e-12, err_msg=str((b, j)))
        assert m[e, :, 0].sum() != 0


def test_radial_model_is_volume_average():
    from oracle import oracle
    p = _problem(const=False)
    o = oracle.Oracle(p.setup)
    m, nsteps, _ = o.integrate_model(p.verts, p.conn)
    verts = np.asarray(p.verts, float).reshape(-1, 3)
    rng = np.random.default_rng(0)
    for e in range(len(p.conn)):
        v = verts[np.asarray(p.conn[e]) - 1]
        u = rng.random((200000, 3))
        pts = oracle.tet_points(v, u)
        r = np.linalg.norm(pts, axis=1)
        idx = np.clip((r / float(p.setup.model_dr) + 1).astype(int), 1, p.setup.model_ndepth) - 1
        vp_avg = p.model[0][idx].astype(np.float64).mean()
        assert abs(m[e, 0, 0] - vp_avg) < 4e-2 * abs(vp_avg)


def test_stream_positions():
    from oracle import oracle
    p = _problem(const=False)
    o = oracle.Oracle(p.setup)
    m1, n1, step = o.integrate_model(p.verts, p.conn, halton_step=1000)
    assert step == 1000 + len(p.conn) * 9999 * NP
    offs = 1000 + np.arange(len(p.conn), dtype=np.int64) * 9999 * NP
    m2, n2, step2 = o.integrate_model(p.verts, p.conn, halton_step=5, halton_offsets=offs)
    np.testing.assert_array_equal(m2, m1)
    np.testing.assert_array_equal(n2, n1)
    assert step2 == 5
