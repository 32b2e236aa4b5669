"""Pins of the oracle's differentiable filter-convolve-window-integrate stage (the FP64
reference gradient of the GPU adjoint).

The forward must equal the oracle's own receiver loop (ora_base_kernels, a restatement
of slave.f90:586-640) on the same traces; the adjoint must equal central finite
differences of the forward in FP64.  The stage is linear in the fwd traces, linear or
quadratic in the bwd traces and quadratic in the filter responses, so central
differences carry no truncation error: they agree with the adjoint to rounding, and a
wrong sign, index, conjugate or transposed factor fails them."""
import numpy as np
import pytest

from mc_kernel_amd import synth

VARIANTS = {
    "full": {},
    "trace": dict(kernels=synth._kernels(2, 2, ["CC", "AM"], ["vp", "lam"])),
    "aniso": dict(kernels=synth._kernels(1, 2, ["CC"], ["vsh", "vsv", "vph", "vpv", "eta", "mu"]),
                  receivers_deg=[40.0]),
    "trapezoidal": dict(int_scheme="trapezoidal"),
    "deconv": dict(deconv_stf=1),
}


@pytest.fixture(scope="module", params=sorted(VARIANTS))
def stage(request):
    from oracle import oracle
    p = synth.make_problem("tiny", ntet=10, **{"deconv_stf": 0, **VARIANTS[request.param]})
    o = oracle.Oracle(p.setup)
    x = p.sample_points(2, seed=5)
    fw = o.load_strain(0, x)
    bw = np.stack([o.load_strain(1 + r, x) for r in range(o.nrec)], 1)
    return request.param, p, o, x, fw, bw


def test_forward_is_the_receiver_loop(stage):
    name, p, o, x, fw, bw = stage
    base = o.ad_forward(o.ad_transfer(), fw, bw)
    ref = o.base_kernels(x)
    assert np.abs(ref).max() > 0
    if name == "deconv":
        # TF = G x deconvolution factors is formed in another order than add_stfs
        assert np.abs(base - ref).max() <= 1e-12 * np.abs(ref).max()
    else:
        np.testing.assert_array_equal(base, ref)


def _directional(o, G, fw, bw, gb, which, d, h=1e-3):
    args = [G, fw, bw]
    plus, minus = list(args), list(args)
    plus[which] = args[which] + h * d
    minus[which] = args[which] - h * d
    lp = float((o.ad_forward(*plus) * gb).sum())
    lm = float((o.ad_forward(*minus) * gb).sum())
    return (lp - lm) / (2 * h)


def test_adjoint_matches_central_differences(stage):
    name, p, o, x, fw, bw = stage
    rng = np.random.default_rng(11)
    G = o.ad_transfer()
    gb = rng.standard_normal((fw.shape[0], o.nk, 6))
    gG, gfw, gbw = o.ad_backward(G, fw, bw, gb)
    for which in range(3):
        v = [G, fw, bw][which]
        if which == 0:
            d = (rng.standard_normal(v.shape) + 1j * rng.standard_normal(v.shape)) * np.abs(v).max()
            pred = float(np.sum(gG.real * d.real + gG.imag * d.imag))
        else:
            d = rng.standard_normal(v.shape) * np.abs(v).max()
            pred = float(np.sum([None, gfw, gbw][which] * d))
        fd = _directional(o, G, fw, bw, gb, which, d)
        assert abs(pred - fd) <= 1e-9 * abs(fd), (name, which, pred, fd)


def test_adjoint_single_entries(stage):
    """Single samples: the first and last trace samples (taper and cumsum ends) and the
    DC / Nyquist responses, whose imaginary parts enter only through G^2."""
    name, p, o, x, fw, bw = stage
    rng = np.random.default_rng(12)
    G = o.ad_transfer() * (1 + 0.3j)       # give DC and Nyquist imaginary parts
    gb = rng.standard_normal((fw.shape[0], o.nk, 6))
    gG, gfw, gbw = o.ad_backward(G, fw, bw, gb)
    nw, nt = G.shape[1], fw.shape[-1]
    f0 = 0                                  # a filter every variant's kernels use
    for q in (0, 1, nw // 3, nw - 1):
        for unit, part in ((1.0, "re"), (1j, "im")):
            d = np.zeros_like(G)
            d[f0, q] = unit * max(np.abs(G).max(), 1.0)
            fd = _directional(o, G, fw, bw, gb, 0, d)
            pred = (gG[f0, q].real if part == "re" else gG[f0, q].imag) * abs(d[f0, q])
            assert abs(pred - fd) <= 1e-8 * max(abs(fd), 1e-3 * np.abs(gG).max() * abs(d[f0, q])), (q, part)
    for t in (0, 1, nt // 2, nt - 2, nt - 1):
        for which, g, v in ((1, gfw, fw), (2, gbw, bw)):
            d = np.zeros_like(v)
            idx = (0, 0, t) if which == 1 else (0, 0, 0, t)
            d[idx] = np.abs(v).max()
            fd = _directional(o, G, fw, bw, gb, which, d)
            pred = g[idx] * d[idx]
            assert abs(pred - fd) <= 1e-8 * max(abs(fd), 1e-3 * np.abs(g).max() * d[idx]), (which, t)
