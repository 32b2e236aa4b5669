"""The synthetic inputs the GPU parity tests rely on (no GPU):

* the database generator's samples are the wavelet sums they claim to be (numpy in
  float64 of the same wavelets, to float rounding);
* the kept oracle result of the whole c0 workload (tests/fixtures/c0_oracle.npz) is what
  the oracle computes now: a sample of its voxels is recomputed with the stored stream
  positions and compared bit for bit.
"""
import ctypes as C
import math
import os

import numpy as np

from mc_kernel_amd import abi, synth

HERE = os.path.dirname(os.path.abspath(__file__))


def _mix64(x):
    M = (1 << 64) - 1
    x = (x + 0x9e3779b97f4a7c15) & M
    x = ((x ^ (x >> 30)) * 0xbf58476d1ce4e5b9) & M
    x = ((x ^ (x >> 27)) * 0x94d049bb133111eb) & M
    return x ^ (x >> 31)


def _unif(seed, a, b, c):
    h = _mix64(seed ^ _mix64(a ^ _mix64(b ^ _mix64(c))))
    return (h >> 11) * (1.0 / 9007199254740992.0)


def test_generator_is_the_wavelet_sum():
    m = synth.Mesh(120)
    nt, dt, seed = 512, 2.0, 77
    u = synth.fill_db(m, 2, nt, dt, seed, nthreads=2)
    t = np.arange(nt) * dt
    rng = np.random.default_rng(0)
    for _ in range(20):
        e, v, j, i = int(rng.integers(m.nelem)), int(rng.integers(2)), int(rng.integers(5)), int(rng.integers(5))
        pt = int(m.sem[e, j, i])
        ref = np.zeros(nt)
        for w in range(8):
            u1, u2, u3 = (_unif(seed, pt, v, 4 * w + k) for k in range(3))
            amp = math.sqrt(-2.0 * math.log(1.0 - u1)) * math.cos(2.0 * math.pi * u2)
            tau = 100.0 + 800.0 * u3
            x = t - tau
            on = np.abs(x) <= 240.0
            ref[on] += amp * np.exp(-(x[on] / 40.0) ** 2) * np.cos(2.0 * np.pi * x[on] / 40.0)
        got = u[e, v, j, i].astype(np.float64)
        assert np.max(np.abs(got - ref)) <= 1e-6 * max(1.0, np.abs(ref).max()), (e, v, j, i)


def test_c0_fixture_current():
    from oracle import oracle
    ref = np.load(os.path.join(HERE, "fixtures", "c0_oracle.npz"))
    p = synth.make_problem("c0", fill=False)
    m = p.mesh
    fwd = np.empty((m.nelem, 10, 5, 5, p.nt), np.float32)
    bwd = np.empty((m.nelem, 5, 5, 5, p.nt), np.float32)
    p.setup.fwd.snapshots = abi.ptr(fwd, C.c_float)
    p.setup.bwd.snapshots = abi.ptr(bwd, C.c_float)
    lib = synth.synth_lib()
    lib.synth_lazy_create.argtypes = [abi.i32p, C.c_int32, C.c_int32, C.c_int32, C.c_double,
                                      C.c_uint64, C.c_uint64]
    lib.synth_lazy_create.restype = C.c_void_p
    user = lib.synth_lazy_create(abi.ptr(m.sem, C.c_int32), 10, 5, p.nt, p.dt, 1001, 1002)
    o = oracle.Oracle(p.setup)
    o.set_lazy_fill(C.cast(lib.synth_fill_element, C.c_void_p), user)
    live = np.nonzero(np.abs(ref["values"]).sum(1) > 0)[0]
    dead = np.nonzero(np.abs(ref["values"]).sum(1) == 0)[0]
    assert len(live) >= 3 and len(dead) >= 2
    sel = np.concatenate([live[:3], dead[:2]])
    r = o.slave_work(p.verts, p.conn[sel], halton_offsets=ref["offsets"][sel])
    assert np.array_equal(r["values"], ref["values"][sel])
    assert np.array_equal(r["niterations"], ref["niterations"][sel])
    lib.synth_lazy_free.argtypes = [C.c_void_p]
    del o
    lib.synth_lazy_free(user)
