"""Oracle pins for the functions that the reference's fixtures leave open.

  * dampen_field (readfields.f90:3184-3204): the reference's own known-answer test
    (test_readfields.f90:1189-1217): 1000 random points, r_src (0.5, 0.5, 0.5), r_max 0.2;
  * calc_basekernel's lambda and mu (kernel.f90:514-570), the only base kernels of the
    c0-c4 kernel sets: the trace-only branch equals the full-tensor branch on the trace,
    an isotropic forward strain gives K_mu = 2/3 K_lambda, both are symmetric in the two
    fields, and a unit strain gives -1 and -2;
  * add_stfs (filtering.f90:241-470): a signal convolved (in the time domain) with the
    database STF and filtered with the backward response is the filtered signal times the
    source magnitude; the forward response also applies the source STF and advances by the
    STF shift (a time-domain convolution and a circular shift); the stop at the highest
    frequency where its ratio exceeds 1e-2, and not below;
  * the slave_work loop (slave.f90:495-720): iteration counts, draws and evaluations
    counted by hand for integrals that never converge (the max_iter exit after the draw;
    kernels that need two base kernels count two iterations per step), for receivers all
    masked by the travel time (one step, one draw, zeros), and for a receiver inside the
    element (one iteration per step for its kernels).
"""
import numpy as np
import pytest

from mc_kernel_amd import abi, synth
from oracle import oracle


# --------------------------------------------------------------------------- dampen_field

def test_dampen_field_known_answer():
    """test_readfields.f90:1189-1217: points farther than r_max keep their values, the
    others are scaled by |r - r_src| / r_max (assert_comparable, 1e-5)."""
    rng = np.random.default_rng(1189)
    npoints = 1000
    field = rng.random((npoints, 2, 6))
    pts = rng.random((npoints, 3))
    r_src = np.array([0.5, 0.5, 0.5])
    got = oracle.dampen_field(field, pts, r_src, 0.2)
    dist = np.sqrt(((pts - r_src) ** 2).sum(axis=1))
    inside = dist <= 0.2
    assert 10 < inside.sum() < npoints
    np.testing.assert_allclose(got[~inside], field[~inside], rtol=1e-5)
    np.testing.assert_allclose(got[inside], field[inside] * (dist[inside] / 0.2)[:, None, None], rtol=1e-5)
    # r_max <= 0: no damping (readfields.f90:3194)
    np.testing.assert_array_equal(oracle.dampen_field(field, pts, r_src, 0.0), field)


# --------------------------------------------------------------------------- calc_basekernel

def _fields(rng, nd, nw=33):
    return rng.standard_normal((nd, nw)) + 1j * rng.standard_normal((nd, nw))


def test_basekernel_lambda_trace_branch_equals_full_branch():
    """The trace-only strain carries tr(e) in its one row; K_lambda from it equals
    K_lambda of the full tensors, -tr(e_f) tr(e_b)."""
    rng = np.random.default_rng(514)
    F, B = _fields(rng, 6), _fields(rng, 6)
    full = oracle.basekernel(0, True, F, B)
    trace = oracle.basekernel(0, False, F[:3].sum(0)[None], B[:3].sum(0)[None])
    np.testing.assert_allclose(trace, full, rtol=1e-12)
    # shear components do not enter K_lambda
    F2 = F.copy()
    F2[3:] = _fields(rng, 3)
    np.testing.assert_allclose(oracle.basekernel(0, True, F2, B), full, rtol=1e-12)


def test_basekernel_mu_isotropic_strain():
    """An isotropic forward strain a (1, 1, 1, 0, 0, 0) gives K_mu = -2 a tr(e_b) =
    2/3 K_lambda for any backward strain."""
    rng = np.random.default_rng(521)
    a = _fields(rng, 1)[0]
    F = np.zeros((6, a.size), complex)
    F[:3] = a
    B = _fields(rng, 6)
    lam = oracle.basekernel(0, True, F, B)
    mu = oracle.basekernel(1, True, F, B)
    np.testing.assert_allclose(mu, 2.0 / 3.0 * lam, rtol=1e-12)
    np.testing.assert_allclose(lam, -3 * a * B[:3].sum(0), rtol=1e-12)


def test_basekernel_symmetry_and_unit_strain():
    rng = np.random.default_rng(550)
    F, B = _fields(rng, 6), _fields(rng, 6)
    for ibk in (0, 1):
        np.testing.assert_allclose(oracle.basekernel(ibk, True, F, B), oracle.basekernel(ibk, True, B, F),
                                   rtol=1e-12)
    e = np.zeros((6, 4), complex)
    e[2] = 1.0
    np.testing.assert_array_equal(oracle.basekernel(0, True, e, e), -np.ones(4))
    np.testing.assert_array_equal(oracle.basekernel(1, True, e, e), -2 * np.ones(4))
    t = np.ones((1, 4), complex)
    np.testing.assert_array_equal(oracle.basekernel(1, False, t, t), -2 * np.ones(4))


# --------------------------------------------------------------------------- add_stfs

NT, DT = 512, 1.0
N = 2 * 512
NW = N // 2 + 1
DF = 1.0 / (N * DT)
T = np.arange(NT) * DT


def _gauss(t0, w):
    return np.exp(-(((T - t0) / w) ** 2))


def _signal():
    """a 40 s-period wavelet at 200 s, well inside the trace"""
    return np.exp(-(((T - 200.0) / 60.0) ** 2)) * np.cos(2 * np.pi * (T - 200.0) / 40.0)


def _apply(x, tf):
    """x (length <= N) zero padded, filtered with tf in the frequency domain"""
    pad = np.zeros(N)
    pad[:len(x)] = x
    return oracle.irfft(oracle.rfft(pad, DT) * tf, N, DF)


def _gabor():
    return oracle.filter_tf("Gabor", (40.0, 0.5, 0.0, 0.0), NW, DF)


@pytest.mark.parametrize("amplitude", [1.0, 3.5e19])
def test_add_stfs_backward_deconvolves_the_database_stf(amplitude):
    stf = _gauss(10.0, 1.0)
    tf = _gabor()
    rc, tff, tfb, msg = oracle.add_stfs(tf, stf, DT, amplitude)
    assert rc == 0, msg
    s = _signal()
    recorded = DT * np.convolve(s, stf)              # what the database holds for s
    got = _apply(recorded, tfb)
    want = amplitude * _apply(s, tf)
    assert np.linalg.norm(got - want) < 1e-9 * np.linalg.norm(want)
    # a unit earthquake STF at t = 0 and no shift: the forward response applies dt x
    # the unit impulse, i.e. the backward one times dt
    np.testing.assert_allclose(tff, tfb * DT, rtol=1e-12, atol=1e-300)


@pytest.mark.parametrize("shift_samples", [0, 25, -40])
def test_add_stfs_forward_applies_source_stf_and_shift(shift_samples):
    stf = _gauss(10.0, 1.0)
    src = _gauss(30.0, 3.0)
    tf = _gabor()
    amplitude = 2.0e18
    rc, tff, _, msg = oracle.add_stfs(tf, stf, DT, amplitude, stf_src=src, stf_shift=shift_samples * DT)
    assert rc == 0, msg
    s = _signal()
    got = _apply(DT * np.convolve(s, stf), tff)
    moved = np.zeros(N)
    v = DT * np.convolve(s, src)
    moved[:len(v)] = v
    moved = np.roll(moved, -shift_samples)          # x(t + shift), circularly
    want = amplitude * _apply(moved, tf)
    assert np.linalg.norm(got - want) < 1e-9 * np.linalg.norm(want)


def _ratio(tf, lp2):
    r = np.abs(tf * lp2)
    return r[-1] / r.max()


def test_add_stfs_stops_where_the_response_does_not_vanish():
    """A unit impulse as the database STF (|spectrum| = dt = 1): the responses are
    tf x amplitude x |lowpass|^2, the low pass being the 12th-order Butterworth at
    0.75 x the highest frequency; the reference stops when the response at the highest
    frequency exceeds 1e-2 x its maximum."""
    stf = np.zeros(NT)
    stf[5] = 1.0
    f = np.arange(NW) * DF
    lp = oracle.filter_tf("Butterw_LP", (1.0 / (f[-1] * 0.75), 0.0, 12.0, 0.0), NW, DF)
    lp2 = np.abs(lp) ** 2
    for a, stops in ((1.0, False), (5.0, True)):
        tf = np.exp(a * (f / f[-1]) ** 4).astype(complex)
        r = _ratio(tf, lp2)
        assert (r > 1e-2) == stops and abs(np.log10(r / 1e-2)) > 0.3, r
        rc, _, tfb, msg = oracle.add_stfs(tf, stf, DT, 1.0)
        assert rc == (-1 if stops else 0), (a, r, msg)
        if not stops:
            np.testing.assert_allclose(np.abs(tfb), np.abs(tf) * lp2, rtol=1e-12)
        else:
            assert "not vanishing" in msg
    # a Gabor filter on the synthetic databases' narrow STF passes
    assert oracle.add_stfs(_gabor(), _gauss(10.0, 1.0), DT)[0] == 0


# --------------------------------------------------------------------------- slave_work loop

PNAMES = {v: k for k, v in abi.PARAMS.items()}


def _loop_problem(**over):
    kw = dict(ntet=4, npoints_per_step=8)
    kw.update(over)
    p = synth.make_problem("tiny", **kw)
    return p


def _increments(p):
    """iterations a kernel counts per evaluated step: one per base kernel it needs
    (slave.f90:626-653, the count sits in the base-kernel loop)"""
    out = []
    for k in range(p.nkernel):
        rc, nb = oracle.param_needs(PNAMES[int(p.setup.kernel_param[k])])
        assert rc == 0
        out.append(int(nb.sum()))
    return np.array(out)


@pytest.mark.parametrize("max_iter", [6, 7])
def test_loop_never_converging(max_iter):
    """allowed errors 0: no kernel converges.  A step draws, exits if any kernel has
    exactly max_iter iterations, else evaluates every kernel (+1 per base kernel).  With
    vp kernels (+1) and vs kernels (+2): max_iter 6 stops after 3 steps (vs at 6), max_iter
    7 after 7 steps (vp at 7; vs passes 7 at 14).  Draws = steps + 1."""
    p = _loop_problem(max_iter=max_iter, allowed_error=0.0, allowed_relative_error=0.0)
    p.setup.vp_max = p.setup.vs_max = 1e12                 # no travel-time mask
    inc = _increments(p)
    assert set(inc) == {1, 2}
    steps = 1
    while not np.any(inc * steps == max_iter) and steps < 100:
        steps += 1
    assert steps == {6: 3, 7: 7}[max_iter]
    o = oracle.Oracle(p.setup)
    o.halton_step = 1000
    res = o.slave_work(p.verts, p.conn)
    P, ne = p.setup.npoints_per_step, len(p.conn)
    assert np.array_equal(res["niterations"], np.tile(inc * steps, (ne, 1)))
    assert np.array_equal(res["offsets"], 1000 + np.arange(ne) * (steps + 1) * P)
    assert res["halton_step"] == 1000 + ne * (steps + 1) * P
    assert res["evaluations"] == ne * steps * P * p.nkernel
    assert np.all(res["values"] != 0)


def test_loop_all_masked_converges_after_one_step():
    """Every receiver masked by the travel time (v_max 1 mm/s): each step gives zero base
    kernels and one iteration per kernel; after the first step (8 points > 2) the variance
    is 0 < (allowed_error V)^2, so every integral converges: one draw, zeros, no evaluation."""
    p = _loop_problem(max_iter=6, allowed_error=1e-10)
    p.setup.vp_max = p.setup.vs_max = 1e-3
    o = oracle.Oracle(p.setup)
    res = o.slave_work(p.verts, p.conn)
    ne, P = len(p.conn), p.setup.npoints_per_step
    assert np.array_equal(res["niterations"], np.ones((ne, p.nkernel), np.int32))
    assert not np.any(res["values"]) and not np.any(res["variance"])
    assert res["evaluations"] == 0
    assert np.array_equal(res["offsets"], np.arange(ne) * P)
    assert res["halton_step"] == ne * P


def test_loop_receiver_inside_element():
    """mask_src_rec with receiver 0 inside the element (slave.f90:539-551): receiver 0's
    kernels count one iteration per step, receiver 1's kernels one per base kernel; never
    converging with max_iter 6 gives 3 steps."""
    p = _loop_problem(max_iter=6, allowed_error=0.0, allowed_relative_error=0.0)
    p.setup.vp_max = p.setup.vs_max = 1e12
    p.setup.mask_src_rec = 1
    rec = synth.geo_xyz(p.rec_lat[0], p.rec_lon[0], synth.R_EARTH - 30e3)   # the tet holds the receiver
    a = 60e3
    tet = rec + a * np.array([[1, 1, 1], [1, -1, -1], [-1, 1, -1], [-1, -1, 1]], float)
    verts = np.concatenate([np.asarray(p.verts, float).reshape(-1, 3), tet])
    n0 = len(verts) - 4
    conn = np.concatenate([np.asarray(p.conn, np.int32)[:1], np.arange(n0 + 1, n0 + 5, dtype=np.int32)[None]])
    inc = _increments(p)
    rec_of = np.asarray(np.ctypeslib.as_array(p.setup.kernel_receiver, shape=(p.nkernel,)))
    o = oracle.Oracle(p.setup)
    res = o.slave_work(verts, conn)
    steps = 3
    assert np.array_equal(res["niterations"][0], inc * steps)
    want = np.where(rec_of == 0, steps, inc * steps)
    assert np.array_equal(res["niterations"][1], want)
    assert not np.any(res["values"][1][rec_of == 0])
    assert np.all(res["values"][1][rec_of == 1] != 0)
    P = p.setup.npoints_per_step
    assert res["evaluations"] == steps * P * (p.nkernel + int((rec_of == 1).sum()))
