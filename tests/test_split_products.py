"""Error bounds of the split fp16 products that k_bwd_mfma uses for the contraction of the
element spectra with the pair weights (fp32 operands x = hi + lo, hi = fp16(x),
lo = fp16(x - hi), after a power-of-two scale that puts the largest magnitude in
[2^13, 2^14); x y ~ hi_x hi_y + hi_x lo_y + lo_x hi_y, each fp16 product exact in fp32).
The GPU parity tests hold the kernel to 1e-5 against the FP64 oracle; these check the
bounds its design rests on, with numpy's IEEE fp16 rounding (round to nearest even, as
v_cvt_f16_f32)."""
import numpy as np


def split(x):
    x = np.asarray(x, np.float32)
    hi = x.astype(np.float16)
    lo = (x - hi.astype(np.float32)).astype(np.float16)
    return hi, lo


def scale_exp(mx):
    """14 - e for mx = m 2^e, m in [0.5, 1): the scale that maps mx into [2^13, 2^14)."""
    return 14 - np.frexp(np.float32(mx))[1]


def test_scale_range():
    rng = np.random.default_rng(5)
    for mx in np.concatenate([10.0 ** rng.uniform(-30, 30, 2000), [1.0, 2.0 ** 13, 2.0 ** 14, 3e-38]]):
        mx = np.float32(mx)
        s = np.ldexp(np.float64(mx), int(scale_exp(mx)))
        assert 2.0 ** 13 <= s < 2.0 ** 14, (mx, s)


def scaled(rng, n, lo_exp):
    """n fp32 values of both signs with |x| in [2^lo_exp, 2^14)"""
    x = rng.choice([-1.0, 1.0], n) * 2.0 ** rng.uniform(lo_exp, 14, n)
    x = x.astype(np.float32)
    return x[np.abs(x) < 2.0 ** 14]


def test_split_keeps_22_bits():
    rng = np.random.default_rng(7)
    # where lo stays a normal fp16 number (|x| >= 2^-3 after the scale): 22 bits
    x = scaled(rng, 200000, -3)
    hi, lo = split(x)
    err = np.abs(x.astype(np.float64) - hi.astype(np.float64) - lo.astype(np.float64))
    assert np.all(err <= 2.0 ** -22 * np.abs(x.astype(np.float64)))
    # hi alone is fp16 (11 bits): the split is needed
    assert np.max(np.abs(x - hi.astype(np.float64)) / np.abs(x)) > 2.0 ** -13
    # below: an absolute error of at most half the smallest fp16 step (2^-25), i.e. 2^-38
    # of the largest magnitude of the scaled operand
    x = scaled(rng, 200000, -40)
    hi, lo = split(x)
    err = np.abs(x.astype(np.float64) - hi.astype(np.float64) - lo.astype(np.float64))
    assert np.all(err <= np.maximum(2.0 ** -22 * np.abs(x.astype(np.float64)), 2.0 ** -25))


def test_three_products_error():
    rng = np.random.default_rng(11)
    n = 100000
    x = scaled(rng, n, -3)[:90000]
    y = scaled(rng, n, -3)[:90000]
    xh, xl = split(x)
    yh, yl = split(y)
    f = lambda a: a.astype(np.float64)
    # every fp16 x fp16 product is exact in fp32 (22 significant bits)
    for a, b in ((xh, yh), (xh, yl), (xl, yh)):
        p = f(a) * f(b)
        assert np.array_equal(p.astype(np.float32).astype(np.float64), p)
    approx = f(xh) * f(yh) + f(xh) * f(yl) + f(xl) * f(yh)
    exact = f(x) * f(y)
    assert np.all(np.abs(approx - exact) <= 3.0 * 2.0 ** -22 * np.abs(exact))


def test_contraction_tile():
    """One 16-bin tile of the contraction, B(k, p) = sum_j U(k, j) W(j, p) over 50 traces
    with cancelling signs and rows 10^6 apart in magnitude: split products with fp32
    accumulation against FP64, each pair's column within 3e-6 of its norm (up to 1.0e-6
    here; fp32 products with fp32 accumulation give up to 2.4e-7 on the same tile)."""
    rng = np.random.default_rng(3)
    J = 50
    U = (rng.standard_normal((16, J)) * 10.0 ** rng.uniform(-12, -6, (16, 1))).astype(np.float32)
    W = (rng.standard_normal((J, 16)) * 10.0 ** rng.uniform(2, 6, (1, 16))).astype(np.float32)
    eu = scale_exp(np.max(np.abs(U)))
    ew = np.array([scale_exp(np.max(np.abs(W[:, p]))) for p in range(16)])
    Us = np.ldexp(U, eu).astype(np.float32)
    Ws = np.ldexp(W, ew[None, :]).astype(np.float32)
    uh, ul = split(Us)
    wh, wl = split(Ws)
    acc = np.zeros((16, 16), np.float32)
    for j in range(J):                                         # fp32 accumulation
        for a, b in ((uh, wh), (uh, wl), (ul, wh)):
            acc = (acc + np.outer(a[:, j].astype(np.float32), b[j, :].astype(np.float32))).astype(np.float32)
    got = np.ldexp(acc.astype(np.float64), -(eu + ew)[None, :])
    ref = U.astype(np.float64) @ W.astype(np.float64)
    for p in range(16):
        assert np.linalg.norm(got[:, p] - ref[:, p]) <= 3e-6 * np.linalg.norm(ref[:, p]), p
