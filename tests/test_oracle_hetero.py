"""Pins of the oracle's heterogeneity integrals (integrate_3d_model, slave.f90:874-924;
load_het_rtpv / load_model_coeffs, heterogeneities.f90:63-188):

  * coordinates: (r, lat, lon) known answers, latitude outside [-90, 90] rejected;
  * the nearest model point equals the reference's own kd-tree (src/kdtree2.f90 built
    into oracle/_ref, build_hetero_kdtree's 3-D tree, kdtree2_n_nearest nn = 1) on random
    clouds, index and fp32 squared distance;
  * a constant model integrates to its values exactly after one step (dyadic values);
  * on-vertex basis functions with a constant model: c x the mean barycentric weight over
    exactly the points drawn, each integrator stopping at its own convergence step
    (allowed error 1e-2, relative 1e-1);
  * a two-point model (nearest = the side of a plane): the integral is the volume
    fraction on each side within the MC error (fine deterministic quadrature);
  * stream positions: element e starts at halton_step + e x 9999 x 10;
  * the heterogeneity file reader.
"""
import ctypes as C
import os

import numpy as np
import pytest

from mc_kernel_amd import synth

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_KD = os.path.join(ROOT, "oracle", "_ref", "libref_kdtree.so")
NP = 10


def _rtp_of(xyz):
    xyz = np.asarray(xyz, float)
    r = np.linalg.norm(xyz, axis=1)
    lat = 90.0 - np.degrees(np.arccos(xyz[:, 2] / r))
    lon = np.degrees(np.arctan2(xyz[:, 1], xyz[:, 0]))
    return np.stack([r / 1000.0, lat, lon], 1)


def test_coordinates():
    from oracle import oracle
    xyz = oracle.het_xyz([[6371.0, 0.0, 0.0], [6000.0, 90.0, 33.0], [3000.0, -90.0, 0.0], [5000.0, 0.0, 90.0]])
    # the reference's pi is truncated (3.1415926535898), so cos(90 deg) is -3.5e-15, not 0
    np.testing.assert_allclose(xyz[0], [6371000.0, 0.0, 0.0], atol=1e-7)
    assert xyz[0, 0] == np.float32(6371000.0) and xyz[0, 1] == 0.0
    np.testing.assert_array_equal(xyz[1], np.float32([0.0, 0.0, 6000000.0]))
    np.testing.assert_allclose(xyz[2], [0.0, 0.0, -3000000.0], atol=1e-3)
    np.testing.assert_allclose(xyz[3], [0.0, 5000000.0, 0.0], atol=1.0)
    with pytest.raises(ValueError):
        oracle.het_xyz([[6000.0, 90.5, 0.0]])


@pytest.mark.skipif(not os.path.exists(REF_KD), reason="oracle/_ref not built (needs the reference tree)")
def test_nearest_matches_reference_kdtree():
    from oracle import oracle
    L = C.CDLL(REF_KD)
    f32p = C.POINTER(C.c_float)
    L.ref_kd3_build.argtypes = [C.c_int, f32p]
    L.ref_kd3_nearest.argtypes = [f32p, C.POINTER(C.c_int), f32p]
    rng = np.random.default_rng(7)
    for n, spread in ((2000, 6.0e6), (300, 1.0e6)):
        pts = (rng.random((n, 3)) * 2 - 1).astype(np.float32) * np.float32(spread)
        L.ref_kd3_build(n, pts.ctypes.data_as(f32p))
        q = (rng.random((400, 3)) * 2.4 - 1.2).astype(np.float32) * np.float32(spread)
        got = oracle.het_nearest(pts, q)
        idx, dis = C.c_int(), C.c_float()
        for i in range(len(q)):
            qi = np.ascontiguousarray(q[i])
            L.ref_kd3_nearest(qi.ctypes.data_as(f32p), C.byref(idx), C.byref(dis))
            assert got[i] == idx.value - 1, i
            d = pts[got[i]] - qi
            mine = np.float32(np.float32(d[0] * d[0] + d[1] * d[1]) + d[2] * d[2])
            assert mine == np.float32(dis.value), i


def _cloud_around(p, n, seed, lo=-0.2, hi=1.2):
    v = np.asarray(p.verts, float).reshape(-1, 3)
    a, b = v.min(0), v.max(0)
    rng = np.random.default_rng(seed)
    pts = a + (b - a) * (lo + (hi - lo) * rng.random((n, 3)))
    return _rtp_of(pts)


def test_constant_model_exact():
    from oracle import oracle
    p = synth.make_problem("tiny", ntet=3)
    rtp = _cloud_around(p, 50, 1)
    dln = np.tile(np.float32([0.25, -0.125, 0.5]), (50, 1))
    h, ns, _ = oracle.integrate_hetero(rtp, dln, p.verts, p.conn)
    np.testing.assert_array_equal(h[:, 0], np.tile([0.25, -0.125, 0.5], (len(p.conn), 1)))
    assert (ns == 1).all()
    vv, vc = p.voxels(2, seed=3)
    rtpv = _cloud_around(type("Q", (), {"verts": vv})(), 50, 2)
    h, ns, _ = oracle.integrate_hetero(rtpv, dln, vv, vc)
    np.testing.assert_array_equal(h[:, 0], np.tile([0.25, -0.125, 0.5], (2, 1)))


def test_vertex_basis_constant_model_is_mean_weight():
    from oracle import oracle
    p = synth.make_problem("tiny", ntet=3)
    c = np.array([0.25, -0.125, 0.5])
    rtp = _cloud_around(p, 40, 5)
    h, ns, _ = oracle.integrate_hetero(rtp, np.tile(c.astype(np.float32), (40, 1)), p.verts, p.conn, nbasis=4)
    verts = np.asarray(p.verts, float).reshape(-1, 3)
    for e in range(len(p.conn)):
        v = verts[np.asarray(p.conn[e]) - 1]
        n = int(ns[e]) * NP
        pts = oracle.tet_points(v, oracle.halton(n, step=e * 9999 * NP))
        w, bad = oracle.tet_weights(v, pts)
        assert not bad
        for b in range(4):
            for j in range(3):
                done = None
                for s in range(1, int(ns[e]) + 1):
                    k = s * NP
                    f = c[j] * w[:k, b]
                    var = (np.mean(f * f) - np.mean(f) ** 2) / (k - 1)
                    if var < 1e-4 or np.sqrt(var) < 1e-1 * abs(np.mean(f)):
                        done = k
                        break
                assert done is not None
                np.testing.assert_allclose(h[e, b, j], c[j] * np.mean(w[:done, b]), rtol=1e-12)


def test_two_point_model_is_volume_fraction():
    from oracle import oracle
    p = synth.make_problem("tiny", ntet=3)
    verts = np.asarray(p.verts, float).reshape(-1, 3)
    rng = np.random.default_rng(0)
    for e in range(len(p.conn)):
        v = verts[np.asarray(p.conn[e]) - 1]
        cen = v.mean(0)
        nrm = np.cross(v[1] - v[0], v[2] - v[0])
        nrm /= np.linalg.norm(nrm)
        d = 0.05 * np.linalg.norm(v[1] - v[0])
        two = np.stack([cen + d * nrm, cen - d * nrm])
        h, ns, _ = oracle.integrate_hetero(_rtp_of(two), np.float32([[1, 0, 0], [0, 1, 0]]), v, [[1, 2, 3, 4]])
        pts = oracle.tet_points(v, rng.random((200000, 3)))
        xyz = oracle.het_xyz(_rtp_of(two)).astype(float)
        frac = np.mean(np.linalg.norm(pts - xyz[0], axis=1) < np.linalg.norm(pts - xyz[1], axis=1))
        # each (basis, parameter) integrator stops on its own, so the two need not sum to 1
        assert abs(h[0, 0, 0] - frac) < 0.15 and abs(h[0, 0, 1] - (1 - frac)) < 0.15
        assert h[0, 0, 2] == 0.0


def test_stream_positions():
    from oracle import oracle
    p = synth.make_problem("tiny", ntet=3)
    rtp = _cloud_around(p, 200, 9)
    dln = np.random.default_rng(1).standard_normal((200, 3)).astype(np.float32) * 0.01
    h1, n1, step = oracle.integrate_hetero(rtp, dln, p.verts, p.conn, halton_step=1000)
    assert step == 1000 + len(p.conn) * 9999 * NP
    offs = 1000 + np.arange(len(p.conn), dtype=np.int64) * 9999 * NP
    h2, n2, step2 = oracle.integrate_hetero(rtp, dln, p.verts, p.conn, halton_step=3, halton_offsets=offs)
    np.testing.assert_array_equal(h2, h1)
    np.testing.assert_array_equal(n2, n1)
    assert step2 == 3


def test_het_file_reader(tmp_path):
    import mc_kernel_amd as mk
    path = tmp_path / "het.dat"
    path.write_text("3\n6371.0 10.0 20.0 0.01 0.02 0.03\n5000 -45 170 -0.5 0 0.25\n4000.5 0 0 1 2 3\n")
    rtp, dln = mk.load_het_rtpv(str(path))
    np.testing.assert_array_equal(rtp, [[6371.0, 10.0, 20.0], [5000.0, -45.0, 170.0], [4000.5, 0.0, 0.0]])
    np.testing.assert_array_equal(dln, np.float32([[0.01, 0.02, 0.03], [-0.5, 0, 0.25], [1, 2, 3]]))
    path.write_text("1\n6371.0 91.0 20.0 0.01 0.02 0.03\n")
    with pytest.raises(ValueError):
        mk.load_het_rtpv(str(path))
