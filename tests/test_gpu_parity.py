"""Parity of the HIP path (through the C-ABI) with the FP64 CPU oracle.

Bars: element indices bit-exact; strain traces, base kernels and MC kernel values within
1e-5 relative to the L2 norm over points / voxels (fp32 wavefields and FFTs, fp64
accumulation of window integrals and MC sums).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TOL = 1e-5


@pytest.fixture(scope="module")
def gpu(tiny):
    import mc_kernel_amd as mk
    k = mk.MCKernel(tiny.setup)
    yield k
    k.close()


def _points(p, n, seed=3):
    return p.sample_points(n, seed=seed)


@pytest.mark.parametrize("frame", [0, 1, 2])
def test_locate_bitexact(tiny, tiny_oracle, gpu, frame):
    x = _points(tiny, 4000, seed=10 + frame)
    ge, gxi, geta, gtie = gpu.locate(frame, x)
    oe, oxi, oeta, otie = tiny_oracle.locate(frame, x)
    assert np.array_equal(gtie, otie)
    ok = otie == 0
    assert ok.mean() > 0.99
    assert np.array_equal(ge[ok], oe[ok]), "element indices differ"
    assert np.max(np.abs(gxi[ok] - oxi[ok])) < 1e-9
    assert np.max(np.abs(geta[ok] - oeta[ok])) < 1e-9


@pytest.mark.parametrize("frame", [0, 1])
def test_strain_traces(tiny, tiny_oracle, gpu, frame):
    x = _points(tiny, 48, seed=20 + frame)
    g = gpu.load_strain(frame, x).astype(np.float64)
    o = tiny_oracle.load_strain(frame, x)
    err = np.linalg.norm(g - o) / np.linalg.norm(o)
    assert err < TOL, err
    per = np.linalg.norm((g - o).reshape(len(x), -1), axis=1) / np.linalg.norm(o.reshape(len(x), -1), axis=1)
    assert per.max() < 10 * TOL, per.max()


def test_base_kernels(tiny, tiny_oracle, gpu):
    x = _points(tiny, 32, seed=30)
    g = gpu.base_kernels(x)
    o = tiny_oracle.base_kernels(x)
    for k in range(tiny.nkernel):
        for b in range(6):
            ref = o[:, k, b]
            if not np.any(ref):
                assert not np.any(g[:, k, b])
                continue
            err = np.linalg.norm(g[:, k, b] - ref) / np.linalg.norm(ref)
            assert err < TOL, (k, b, err)


def test_slave_work_pinned(tiny, gpu):
    from oracle import oracle
    o = oracle.Oracle(tiny.setup)
    ref = o.slave_work(tiny.verts, tiny.conn)
    res = gpu.slave_work(tiny.verts, tiny.conn, halton_offsets=ref["offsets"])
    assert np.array_equal(res["niterations"], ref["niterations"])
    assert res["halton_step"] == ref["halton_step"]
    assert res["stats"]["evaluations"] == ref["evaluations"]
    for k in range(tiny.nkernel):
        r = ref["values"][:, k]
        nrm = np.linalg.norm(r)
        if nrm == 0:
            assert not np.any(res["values"][:, k])
            continue
        err = np.linalg.norm(res["values"][:, k] - r) / nrm
        assert err < TOL, (k, err)


@pytest.mark.parametrize("start", [0, 12345])
def test_sequential_stream_is_the_reference_stream(tiny, gpu, start):
    """mck_slave_work_sequential walks the reference's own Halton stream (one position
    advanced by every draw, halton_sequence.f90:151, slave.f90:506-511) without positions
    computed elsewhere: the same element starts, iteration counts, evaluations and end
    position as the oracle's sequential loop; the batched form on those starts gives the
    same results bit for bit."""
    from oracle import oracle
    o = oracle.Oracle(tiny.setup)
    o.halton_step = start
    ref = o.slave_work(tiny.verts, tiny.conn)
    gpu.halton_step = start
    seq = gpu.slave_work(tiny.verts, tiny.conn, sequential=True)
    assert np.array_equal(seq["offsets"], ref["offsets"])
    assert seq["halton_step"] == ref["halton_step"] > start
    assert np.array_equal(seq["niterations"], ref["niterations"])
    assert seq["stats"]["evaluations"] == ref["evaluations"]
    for k in range(tiny.nkernel):
        r = ref["values"][:, k]
        nrm = np.linalg.norm(r)
        if nrm == 0:
            assert not np.any(seq["values"][:, k])
            continue
        assert np.linalg.norm(seq["values"][:, k] - r) / nrm < TOL, k
    b = gpu.slave_work(tiny.verts, tiny.conn, halton_offsets=seq["offsets"])
    assert np.array_equal(b["values"], seq["values"])
    assert np.array_equal(b["variance"], seq["variance"])
    assert np.array_equal(b["niterations"], seq["niterations"])
    assert b["halton_step"] == seq["halton_step"]


def test_plain_equals_batched(tiny, gpu):
    gpu.halton_step = 0
    a = gpu.slave_work(tiny.verts, tiny.conn)
    gpu.halton_step = 0
    b = gpu.slave_work(tiny.verts, tiny.conn, plain=True)
    assert np.array_equal(a["niterations"], b["niterations"])
    assert np.array_equal(a["values"], b["values"])
    assert np.array_equal(a["variance"], b["variance"])


def test_slave_work_vertex_basis(tiny, gpu):
    """Basis functions on the tetrahedra's vertices (nbasis 4): one integrator per basis
    function on the same points, values (ne, 4, nk)."""
    from oracle import oracle
    o = oracle.Oracle(tiny.setup)
    ref = o.slave_work(tiny.verts, tiny.conn, nbasis=4)
    res = gpu.slave_work(tiny.verts, tiny.conn, halton_offsets=ref["offsets"], nbasis=4)
    assert res["values"].shape == ref["values"].shape == (tiny.ntet, 4, tiny.nkernel)
    assert np.array_equal(res["niterations"], ref["niterations"])
    assert res["stats"]["evaluations"] == ref["evaluations"]
    for k in range(tiny.nkernel):
        r = ref["values"][:, :, k]
        nrm = np.linalg.norm(r)
        if nrm == 0:
            assert not np.any(res["values"][:, :, k])
            continue
        err = np.linalg.norm(res["values"][:, :, k] - r) / nrm
        assert err < TOL, (k, err)
    # the plain (reference-order) form agrees bit for bit with the batched one
    gpu.halton_step = 0
    a = gpu.slave_work(tiny.verts, tiny.conn, nbasis=4)
    gpu.halton_step = 0
    b = gpu.slave_work(tiny.verts, tiny.conn, plain=True, nbasis=4)
    assert np.array_equal(a["values"], b["values"]) and np.array_equal(a["niterations"], b["niterations"])
