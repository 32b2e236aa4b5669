"""Pins of the oracle's reciprocity seismograms (load_seismogram / load_fw_points_rdbm,
readfields.f90:1004-1048, 1469-1604): the moment tensor rotated source xyz -> earth ->
receiver xyz -> receiver (s, phi, z), contracted with the full strain of the backward
database at the source point.

  * the two rotations added for it invert the pinned forward rotations;
  * an isotropic source is invariant under every rotation: Z = M0 x the strain trace
    at the source point, R = -M0 x the trace of the dipole strain without its azimuthal
    factor, T = 0;
  * a general moment tensor with the Z component: the double contraction M : eps is
    frame-invariant, so the seismogram equals M_earth : eps_earth with the earth-frame
    strain of the independent strain loader (load_bw_points);
  * a set-up without seismograms uses these (the kernels' windows and normalisation).
"""
import ctypes as C

import numpy as np
import pytest

from mc_kernel_amd import synth

DC = [0.3e20, 1.0e20, -1.3e20, 0.5e20, -0.4e20, 0.7e20]
DEG = np.pi / 180.0


def _oracle(seis_none=True, **kw):
    from oracle import oracle
    p = synth.make_problem("tiny", ntet=2, receivers_deg=[25.0, 40.0, 70.0], **kw)
    if seis_none:
        p.setup.seis = None
    return p, oracle.Oracle(p.setup)


def _src_xyz(p):
    return synth.geo_xyz(p.src_lat, p.src_lon, synth.R_EARTH - p.src_depth * 1e3)


def _phi_in_frame(p, r, xyz):
    from oracle import oracle
    s, ph, z = (np.zeros(1) for _ in range(3))
    x = np.ascontiguousarray(xyz, float).reshape(3)
    oracle.lib().ora_rotate_frame(1, oracle.P(x), p.rec_lon[r] * DEG, (90 - p.rec_lat[r]) * DEG,
                                  oracle.P(s), oracle.P(ph), oracle.P(z))
    return float(ph[0])


def test_inverse_rotations():
    from oracle import oracle
    rng = np.random.default_rng(3)
    for _ in range(5):
        t = rng.standard_normal(6)
        phi, theta = rng.uniform(0, 2 * np.pi), rng.uniform(0, np.pi)
        back = oracle.rotate_voigt_inv(oracle.rotate_voigt(t, phi, theta, to_earth=True), phi, theta, True)
        np.testing.assert_allclose(back, t, atol=1e-12)
        back = oracle.rotate_voigt_inv(oracle.rotate_voigt(t, phi), phi, earth_to_xyz_src=False)
        np.testing.assert_allclose(back, t, atol=1e-12)
        # the trace is invariant
        assert abs(oracle.rotate_voigt_inv(t, phi, theta)[:3].sum() - t[:3].sum()) < 1e-12


def test_isotropic_source_z_r_t():
    seis = {}
    for comp in ("Z", "R", "T"):
        p, o = _oracle(component=comp, kernels=synth._kernels(3, 1, ["CC"], ["vp"]))
        seis[comp] = o.seismograms()
        assert seis[comp].shape == (3, p.nt)
        if comp == "T":
            continue
        # trace of the strain at the source point in each receiver frame (trace-only
        # set-up: load_bw_points gives trace x azimuthal factor / amplitude)
        x = _src_xyz(p)[None]
        m0 = p.mij[0]
        for r in range(3):
            tr = o.load_strain(1 + r, x)[0, 0]
            if comp == "Z":
                want = m0 * tr
            else:
                want = -m0 * tr / np.cos(_phi_in_frame(p, r, x))   # a1 = cos(phi) for R
            np.testing.assert_allclose(seis[comp][r], want, rtol=0, atol=1e-10 * np.abs(want).max())
            assert np.abs(want).max() > 0
    assert np.abs(seis["T"]).max() <= 1e-12 * np.abs(seis["R"]).max()


def test_general_moment_tensor_z_is_frame_invariant():
    from oracle import oracle
    p, o = _oracle(mij=DC, component="Z")          # full strain kernel set (vp, vs)
    seis = o.seismograms()
    x = _src_xyz(p)[None]
    mv = np.array([DC[1], DC[2], DC[0], DC[4], DC[3], DC[5]])   # Voigt (tt, pp, rr, rp, rt, tp)
    m_earth = oracle.rotate_voigt(mv, p.src_lon * DEG, (90 - p.src_lat) * DEG, to_earth=True)
    w = np.array([1, 1, 1, 2, 2, 2.0])
    for r in range(3):
        eps = o.load_strain(1 + r, x)[0]             # earth frame, / amplitude
        want = (w[:, None] * m_earth[:, None] * eps).sum(0)
        np.testing.assert_allclose(seis[r], want, rtol=0, atol=1e-10 * np.abs(want).max())
        assert np.abs(want).max() > 0


def test_setup_without_seismograms_uses_reciprocity():
    from oracle import oracle
    p, o = _oracle()
    s = o.seismograms()
    x = p.sample_points(3, seed=2)
    b0 = o.base_kernels(x)
    p2, _ = _oracle(seis_none=False)
    p2.seis = np.ascontiguousarray(s)
    p2.setup.seis = p2.seis.ctypes.data_as(C.POINTER(C.c_double))
    b1 = oracle.Oracle(p2.setup).base_kernels(x)
    np.testing.assert_array_equal(b0, b1)
    assert not np.allclose(s, p.seis)               # not the synthetic seismograms
