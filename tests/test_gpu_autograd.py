"""The filter-convolve-window-integrate stage behind torch.autograd (mck_ad_forward /
mck_ad_backward through mc_kernel_amd.autograd).

  * forward: equals the production path's base kernels (mck_base_kernels) and the FP64
    oracle's receiver loop at the same points, to 1e-5 of the L2 norm;
  * backward: the hand-written adjoint equals the oracle's FP64 adjoint (itself pinned by
    finite differences in tests/test_oracle_ad.py) to 1e-5 of the L2 norm, and central
    finite differences in FP64 of the oracle's forward (step 1e-3 relative) along random
    directions; the torch wiring is checked by finite differences of the GPU forward;
  * the c3 size (10,000 points, 10 receivers, 4 Gabor filters, nt 2048): Euler identities
    of the stage (linear in the fwd traces and, for lambda/mu kernels, in the bwd traces;
    quadratic in the filter responses) on the full gradients.
"""
import numpy as np
import pytest

from mc_kernel_amd import synth

pytestmark = pytest.mark.gpu

TOL = 1e-5

VARIANTS = {
    "full": {},
    "trace": dict(kernels=synth._kernels(2, 2, ["CC", "AM"], ["vp", "lam"])),
    "aniso": dict(kernels=synth._kernels(1, 2, ["CC"], ["vsh", "vsv", "vph", "vpv", "eta", "mu"]),
                  receivers_deg=[40.0]),
    "trapezoidal": dict(int_scheme="trapezoidal"),
    "deconv": dict(deconv_stf=1),
}


def rel(a, b):
    return float(np.linalg.norm(np.ravel(a - b)) / np.linalg.norm(np.ravel(b)))


@pytest.fixture(scope="module", params=sorted(VARIANTS))
def stage(request):
    import mc_kernel_amd as mk
    from oracle import oracle
    p = synth.make_problem("tiny", ntet=10, **{"deconv_stf": 0, **VARIANTS[request.param]})
    print("autograd case", request.param, flush=True)
    o = oracle.Oracle(p.setup)
    k = mk.MCKernel(p.setup)
    x = p.sample_points(12, seed=23)
    fw = k.load_strain(0, x)
    bw = np.stack([k.load_strain(1 + r, x) for r in range(k.R)], 1)
    yield request.param, p, o, k, x, fw, bw
    k.close()


def _tensors(fw, bw):
    import torch
    from mc_kernel_amd import autograd as ad
    tf = ad.transfer_functions().requires_grad_(True)
    f = torch.from_numpy(fw).cuda().requires_grad_(True)
    b = torch.from_numpy(bw).cuda().requires_grad_(True)
    return tf, f, b


def test_forward_parity(stage):
    from mc_kernel_amd import autograd as ad
    name, p, o, k, x, fw, bw = stage
    tf, f, b = _tensors(fw, bw)
    base = ad.filter_convolve_window(tf, f, b).detach().cpu().numpy()
    prod = k.base_kernels(x)
    ref = o.base_kernels(x)
    assert np.abs(ref).max() > 0
    assert rel(base, prod) < TOL, name
    assert rel(base, ref) < TOL, name
    # the oracle's stage on the GPU's own fp32 inputs
    G = tf.detach().cpu().numpy().astype(np.complex128)
    ref2 = o.ad_forward(G, fw.astype(np.float64), bw.astype(np.float64))
    assert rel(base, ref2) < TOL, name


def test_adjoint_vs_fp64_oracle(stage):
    from mc_kernel_amd import autograd as ad
    import torch
    name, p, o, k, x, fw, bw = stage
    tf, f, b = _tensors(fw, bw)
    base = ad.filter_convolve_window(tf, f, b)
    gb = np.random.default_rng(5).standard_normal(tuple(base.shape))
    (base * torch.from_numpy(gb).cuda()).sum().backward()
    G = tf.detach().cpu().numpy().astype(np.complex128)
    fw64, bw64 = fw.astype(np.float64), bw.astype(np.float64)
    gG, gfw, gbw = o.ad_backward(G, fw64, bw64, gb)
    assert rel(tf.grad.cpu().numpy(), gG) < TOL, name
    assert rel(f.grad.cpu().numpy(), gfw) < TOL, name
    assert rel(b.grad.cpu().numpy(), gbw) < TOL, name
    # central differences of the FP64 forward, step 1e-3 relative, along random directions
    rng = np.random.default_rng(6)
    args = [G, fw64, bw64]
    grads = [tf.grad.cpu().numpy(), f.grad.cpu().numpy(), b.grad.cpu().numpy()]
    for which in range(3):
        v = args[which]
        d = rng.standard_normal(v.shape) + (1j * rng.standard_normal(v.shape) if which == 0 else 0)
        h = 1e-3 * np.linalg.norm(v) / np.linalg.norm(d)
        plus, minus = list(args), list(args)
        plus[which] = v + h * d
        minus[which] = v - h * d
        fd = float(((o.ad_forward(*plus) - o.ad_forward(*minus)) * gb).sum()) / (2 * h)
        g = grads[which]
        pred = float(np.sum(g.real * np.real(d) + np.imag(g) * np.imag(d)))
        scale = float(np.linalg.norm(g) * np.linalg.norm(d))
        assert abs(pred - fd) <= TOL * scale, (name, which, pred, fd)


def test_autograd_wiring_finite_differences(stage):
    """loss.backward() through the torch Function agrees with central differences of the
    GPU forward itself (fp32 stage: tolerance 1e-3)."""
    import torch
    from mc_kernel_amd import autograd as ad
    name, p, o, k, x, fw, bw = stage
    tf, f, b = _tensors(fw, bw)
    gen = torch.Generator(device="cuda").manual_seed(7)
    base = ad.filter_convolve_window(tf, f, b)
    w = torch.randn(base.shape, generator=gen, device="cuda", dtype=torch.float64)
    (base * w).sum().backward()
    nrm = lambda a: float(torch.linalg.vector_norm(a.detach().to(torch.complex128 if a.is_complex() else torch.float64)))
    for t in (tf, f, b):
        if t.is_complex():
            d = torch.randn(t.shape, generator=gen, device="cuda", dtype=torch.complex64)
        else:
            # in-band directions: the traces of the neighbouring point (white noise would
            # sit mostly outside the filters, and its derivative would drown in the fp32
            # rounding of the forward values)
            d = t.detach().roll(1, 0)
        h = 1e-2 * nrm(t) / nrm(d)
        with torch.no_grad():
            args = [tf.detach(), f.detach(), b.detach()]
            i = [id(tf), id(f), id(b)].index(id(t))
            plus, minus = list(args), list(args)
            plus[i] = args[i] + h * d
            minus[i] = args[i] - h * d
            fd = float(((ad.filter_convolve_window(*plus) - ad.filter_convolve_window(*minus)) * w).sum()) / (2 * h)
            g = t.grad
            pred = float((g.real.double() * d.real.double()).sum()
                         + ((g.imag.double() * d.imag.double()).sum() if t.is_complex() else 0.0))
        assert abs(pred - fd) <= 1e-3 * abs(pred) + 1e-6 * nrm(g) * nrm(d), (name, i, pred, fd)


def test_c3_euler_identities():
    """c3: the c1 workload's stage (nt 2048, 10 receivers, 4 Gabor filters, CC and AM of
    vp and vs: 160 kernels) at 10,000 points, with the strain traces of the points.  With
    L = sum(w base):
      <dL/dfw, fw> = L        (every base kernel is linear in the fwd spectra)
      <dL/dbw, bw> = L        (lambda and mu are linear in the bwd spectra)
      Re <dL/dG, G> = 2 L     (TF_fwd TF_bwd = G^2 x constants)
    each to 1e-5 of the sums of absolute terms (the rounding scale of the products)."""
    import torch
    import mc_kernel_amd as mk
    from mc_kernel_amd import autograd as ad
    p = synth.make_problem("c1", nelem_target=2000, ntet=200)
    k = mk.MCKernel(p.setup)
    try:
        nd, nt, nw, nf, nr, nk = ad.dims()
        assert (nd, nt, nw, nf, nr, nk) == (6, 2048, 2049, 4, 10, 160)
        P = 10000
        x = p.sample_points(P, seed=31)
        print("c3 stage: loading strain traces", flush=True)
        fw = torch.from_numpy(k.load_strain(0, x)).cuda().requires_grad_(True)
        bw = torch.empty((P, nr, nd, nt), dtype=torch.float32, device="cuda")
        for r in range(nr):
            bw[:, r] = torch.from_numpy(k.load_strain(1 + r, x)).cuda()
        bw.requires_grad_(True)
        tf = ad.transfer_functions().requires_grad_(True)
        gen = torch.Generator(device="cuda").manual_seed(3)
        base = ad.filter_convolve_window(tf, fw, bw)
        w = torch.rand(base.shape, generator=gen, device="cuda", dtype=torch.float64)
        L = (base * w).sum()
        L.backward()
        torch.cuda.synchronize()
        print("c3 stage: forward and backward done", flush=True)
        Lv = float(L)
        scale = float((base.detach() * w).abs().sum())
        with torch.no_grad():
            tf_ = (fw.grad.double() * fw.double())
            lf, sf = float(tf_.sum()), float(tf_.abs().sum())
            lb = sb = 0.0
            for i in range(0, P, 1000):
                t = bw.grad[i:i + 1000].double() * bw[i:i + 1000].double()
                lb += float(t.sum())
                sb += float(t.abs().sum())
            t = (tf.grad.conj() * tf).real.double()
            lg, sg = float(t.sum()), float(t.abs().sum())
        print("c3 stage: L %.6e  <g,fw> %.6e  <g,bw> %.6e  Re<g,G>/2 %.6e" % (Lv, lf, lb, lg / 2), flush=True)
        assert abs(Lv) > 1e-3 * scale
        assert abs(lf - Lv) <= TOL * (sf + scale), (lf, Lv, sf, scale)
        assert abs(lb - Lv) <= TOL * (sb + scale), (lb, Lv, sb, scale)
        assert abs(lg - 2 * Lv) <= TOL * (sg + 2 * scale), (lg, Lv, sg, scale)
    finally:
        k.close()
