"""Trace lengths that are not a multiple of the forward gather's 64-sample super-tiles (nt 1000)
or of 4 samples (nt 250), for a monopole source (4 carried components, 4 forward variables)
and a general moment tensor (6 components, all 10 variables: the gather's weights in LDS):
base kernels and a pinned slave_work against the FP64 oracle.  Bars as in the parity tests:
1e-5 relative to each kernel's L2 norm, iteration counts bit-exact."""
import numpy as np
import pytest

from mc_kernel_amd import synth

pytestmark = pytest.mark.gpu

TOL = 1e-5
DC = [0.3e20, 1.0e20, -1.3e20, 0.5e20, -0.4e20, 0.7e20]
CASES = {"nt250_general_R": dict(nt=250, mij=DC, component="R"), "nt1000_explosion_Z": dict(nt=1000)}


@pytest.mark.parametrize("case", sorted(CASES))
def test_odd_trace_lengths(case):
    import mc_kernel_amd as mk
    from oracle import oracle
    p = synth.make_problem("tiny", **CASES[case])
    print("trace lengths:", case, flush=True)
    o = oracle.Oracle(p.setup)
    x = p.sample_points(24, seed=31)
    ref_b = o.base_kernels(x)
    ref = o.slave_work(p.verts, p.conn)
    k = mk.MCKernel(p.setup)
    try:
        g = k.base_kernels(x)
        res = k.slave_work(p.verts, p.conn, halton_offsets=ref["offsets"])
    finally:
        k.close()
    for j in range(p.nkernel):
        for b in range(6):
            r = ref_b[:, j, b]
            if not np.any(r):
                assert not np.any(g[:, j, b]), (j, b)
                continue
            assert np.linalg.norm(g[:, j, b] - r) / np.linalg.norm(r) < TOL, (j, b)
    assert np.array_equal(res["niterations"], ref["niterations"])
    for j in range(p.nkernel):
        r = ref["values"][:, j]
        n = np.linalg.norm(r)
        if n > 0:
            assert np.linalg.norm(res["values"][:, j] - r) / n < TOL, j
