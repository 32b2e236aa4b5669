"""Writes c0_oracle.npz: the FP64 oracle's slave_work of the whole c0 workload (2,000-quad
database at nt 1024, one receiver at 30 deg, 2 CC vp kernels, 500 tetrahedra, Halton
points from the start of the stream in task order), for the GPU parity test of that
workload (tests/test_gpu_configs.py).  The oracle run takes about two minutes on one core,
too long for every GPU test run, so its result is kept here; test_c0_fixture_current
re-derives a sample of it.

    python3 tests/fixtures/make_c0_oracle.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)


def main():
    from mc_kernel_amd import synth
    from oracle import oracle
    p = synth.make_problem("c0")
    o = oracle.Oracle(p.setup)
    r = o.slave_work(p.verts, p.conn)
    np.savez_compressed(os.path.join(HERE, "c0_oracle.npz"), values=r["values"], variance=r["variance"],
                        niterations=r["niterations"], offsets=r["offsets"], evaluations=r["evaluations"],
                        halton_step=r["halton_step"])
    print("c0: %d evaluations, %d voxels with nonzero kernels"
          % (r["evaluations"], int((np.abs(r["values"]).sum(1) > 0).sum())))


if __name__ == "__main__":
    main()
