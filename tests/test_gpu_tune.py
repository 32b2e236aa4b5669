"""The tuner (mck_tune, mck_set_tuning, MCK_TUNE at mck_init): it times the candidates of
the backward evaluation on the caller's points and measures the per-launch costs, applies
the per-launch choice (or a fixed candidate that beats it), and every candidate gives the
same results to the path's tolerance (the choice only moves time)."""
import numpy as np
import pytest

from mc_kernel_amd import synth

pytestmark = pytest.mark.gpu

TOL = 1e-5
CANDS = [(1, 1, 0), (1, 1, 1), (1, 2, 0), (1, 4, 0), (0, 1, 0)]


def _rel(g, r):
    n = np.linalg.norm(r)
    return np.linalg.norm(g - r) / n if n else np.linalg.norm(g)


def test_tune_applies_the_fastest_and_every_candidate_agrees():
    import mc_kernel_amd as mk
    from oracle import oracle
    p = synth.make_problem("tiny", ntet=10)
    ref = oracle.Oracle(p.setup).slave_work(p.verts, p.conn)
    k = mk.MCKernel(p.setup)
    try:
        t = k.tune(p.sample_points(256, seed=4), reps=2)
        assert t["ncandidates"] == 5
        cms = t["candidate_ms"]
        assert all(c > 0 for c in cms) and t["default_ms"] == cms[0]
        assert t["cost_group_us"] > 0 and t["cost_run_us"] > 0 and t["cost_pair_us"] > 0 and t["auto_ms"] > 0
        if t["bwd_path"] == -1:
            # the per-launch choice, kept unless a fixed candidate beats it by over 2 %
            assert t["best_ms"] == t["auto_ms"] <= 1.02 * min(cms)
        else:
            assert t["best_ms"] == min(cms) < t["auto_ms"] / 1.02
            assert (t["bwd_path"], t["mfma_waves"], t["mfma_xcd"]) == CANDS[int(np.argmin(cms))]
        tuned = k.slave_work(p.verts, p.conn, halton_offsets=ref["offsets"])
        assert np.array_equal(tuned["niterations"], ref["niterations"])
        for j in range(p.nkernel):
            assert _rel(tuned["values"][:, j], ref["values"][:, j]) < TOL, j
        for path, waves, xcd in CANDS + [(-1, 1, 0)]:
            k.set_tuning({"bwd_path": path, "mfma_waves": waves, "mfma_xcd": xcd})
            r = k.slave_work(p.verts, p.conn, halton_offsets=ref["offsets"])
            if path >= 0:
                assert r["stats"]["bwd_path"] == path
                assert (r["stats"]["pair_launches"] > 0) == (path == 0)
            assert np.array_equal(r["niterations"], ref["niterations"]), (path, waves, xcd)
            for j in range(p.nkernel):
                assert _rel(r["values"][:, j], ref["values"][:, j]) < TOL, (path, waves, xcd, j)
        with pytest.raises(mk.MckError):
            k.set_tuning({"bwd_path": 1, "mfma_waves": 3, "mfma_xcd": 0})
    finally:
        k.close()


def test_tune_at_init(monkeypatch):
    """MCK_TUNE=1: mck_init tunes itself on points spread through the mantle."""
    import mc_kernel_amd as mk
    monkeypatch.setenv("MCK_TUNE", "1")
    p = synth.make_problem("tiny", ntet=4)
    k = mk.MCKernel(p.setup)
    try:
        r = k.slave_work(p.verts, p.conn)
        assert r["stats"]["evaluations"] > 0
    finally:
        k.close()
