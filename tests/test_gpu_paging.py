"""Databases paged from host memory (mck_setup_t.db_page_elems > 0: the snapshots stay in
host memory and every launch copies the element pages its points touch into device page
buffers: db_device_pages - 2 pages keep a buffer of their own, the others rotate through
two) give the same results as resident databases, bit for bit:

  * slave_work (element-sorted launches: each page covers one range of the sorted points
    and pairs, untouched pages are not copied), for a monopole (explosion, Z) and a
    general moment tensor with an R receiver (all four forward simulations);
  * base kernels and strain traces (unsorted launches: every page covers all points and
    the kernels skip the points outside it);
  * a page larger than the database falls back to the resident copy;
  * databases opened from merged database files (include/mck_dbio.h), snapshots
    memory-mapped, resident or paged from the mapping.
"""
import numpy as np
import pytest

from mc_kernel_amd import synth

pytestmark = pytest.mark.gpu

DC = [0.3e20, 1.0e20, -1.3e20, 0.5e20, -0.4e20, 0.7e20]
CASES = {"explosion_Z": {}, "general_mij_R": dict(mij=DC, component="R")}
PAGE = 37          # elements per page: 13 pages of the tiny databases (462 elements)
TOL = 1e-5


def _with_kernel(page, kw, fn, device_pages=0, p=None):
    import mc_kernel_amd as mk
    p = p or synth.make_problem("tiny", ntet=12, **kw)
    p.setup.db_page_elems = page
    p.setup.db_device_pages = device_pages
    k = mk.MCKernel(p.setup)
    try:
        return p, fn(p, k)
    finally:
        k.close()


_ORACLE = {}


def _oracle(case):
    """The oracle's slave_work and base kernels of a case's task (kept per case)."""
    if case not in _ORACLE:
        from oracle import oracle
        p = synth.make_problem("tiny", ntet=12, **CASES[case])
        o = oracle.Oracle(p.setup)
        _ORACLE[case] = (o.slave_work(p.verts, p.conn), o.base_kernels(p.sample_points(40, seed=9)))
    return _ORACLE[case]


def _rel(g, r):
    n = np.linalg.norm(r)
    return np.linalg.norm(g - r) / n if n else np.linalg.norm(g)


def _task(p, k, offsets=None):
    res = k.slave_work(p.verts, p.conn, halton_offsets=offsets)
    x = p.sample_points(40, seed=9)
    traces = [k.load_strain(f, x) for f in (0, 1)]
    return res, k.base_kernels(x), traces


@pytest.mark.parametrize("device_pages", [2, 7, 15])
@pytest.mark.parametrize("case", sorted(CASES))
def test_paged_equals_resident(case, device_pages):
    """Paged = resident bit for bit, and both = the oracle on the oracle's stream
    positions (element ids and iteration counts exact, kernels to 1e-5).  2 and 7 buffers:
    element-sorted launches run page by page; 15 (every one of the 13 pages can keep a
    buffer): each launch runs once over all its items through the page table."""
    print("paging case", case, device_pages, flush=True)
    oref, obase = _oracle(case)
    task = lambda p, k: _task(p, k, oref["offsets"])
    p, (ref, base_ref, tr_ref) = _with_kernel(0, CASES[case], task)
    nelem = p.setup.fwd.nelem
    assert nelem > 4 * PAGE
    _, (got, base_got, tr_got) = _with_kernel(PAGE, CASES[case], task, device_pages)
    assert np.abs(ref["values"]).max() > 0
    for key in ("values", "variance", "niterations"):
        np.testing.assert_array_equal(got[key], ref[key], err_msg=key)
    assert got["halton_step"] == ref["halton_step"]
    np.testing.assert_array_equal(base_got, base_ref)
    for a, b in zip(tr_got, tr_ref):
        np.testing.assert_array_equal(a, b)
    # against the oracle
    assert np.array_equal(got["niterations"], oref["niterations"])
    assert got["stats"]["evaluations"] == oref["evaluations"]
    for j in range(p.nkernel):
        assert _rel(got["values"][:, j], oref["values"][:, j]) < TOL, j
        for b in range(6):
            assert _rel(base_got[:, j, b], obase[:, j, b]) < TOL, (j, b)
    # the task's launches copied pages, and never more than every page once per launch
    st, st0 = got["stats"], ref["stats"]
    assert st0["paged_bytes"] == 0
    assert st["paged_bytes"] > 0
    db_bytes = 4 * 25 * p.setup.nt * (p.setup.fwd.nvar * nelem + p.setup.bwd.nvar * p.setup.bwd.nelem)
    assert st["paged_bytes"] <= st["mc_steps"] * db_bytes
    assert st["evaluations"] == st0["evaluations"]


def test_page_larger_than_database_is_resident():
    p, ref = _with_kernel(0, {}, lambda p, k: k.slave_work(p.verts, p.conn))
    _, got = _with_kernel(10 ** 9, {}, lambda p, k: k.slave_work(p.verts, p.conn))
    np.testing.assert_array_equal(got["values"], ref["values"])
    assert got["stats"]["paged_bytes"] == 0


def test_fixed_page_buffers_are_copied_once(monkeypatch):
    """Unsorted launches (base kernels on the per-pair backward kernel) visit every page
    in order.  With 2 buffers every call copies every page of both databases; with 7,
    the first 5 pages used (0-4) keep their buffers and pages 5-12 rotate through the
    other two, so a second call copies exactly 5 pages per database less."""
    monkeypatch.setenv("MCK_BWD_PATH", "0")
    def two_calls(p, k):
        x = p.sample_points(40, seed=9)
        k.base_kernels(x)
        s1 = k.stats()["paged_bytes"]
        b = k.base_kernels(x)
        return b, k.stats()["paged_bytes"] - s1
    p, (b2, d2) = _with_kernel(PAGE, {}, two_calls, 2)
    _, (b7, d7) = _with_kernel(PAGE, {}, two_calls, 7)
    np.testing.assert_array_equal(b7, b2)
    f, b = p.setup.fwd, p.setup.bwd
    assert f.nelem > 12 * PAGE and b.nelem > 12 * PAGE
    elem_bytes = lambda db: db.nvar * 25 * p.setup.nt * 4
    full = f.nelem * elem_bytes(f) + b.nelem * elem_bytes(b)
    assert d2 == full
    assert d7 == full - 5 * PAGE * (elem_bytes(f) + elem_bytes(b))


@pytest.mark.parametrize("page", [0, PAGE])
def test_slave_work_from_database_files(tmp_path, page):
    """Databases written as merged database files and opened with mck_db_open (snapshots
    memory-mapped from the file) give the same slave_work as the in-memory databases,
    resident and paged from the mapping.  The file stores the source magnitude as a
    32-bit real (as the reference reads it), so the in-memory set-up uses that value."""
    import mc_kernel_amd as mk
    from mc_kernel_amd import dbfile
    p = synth.make_problem("tiny", ntet=12)
    amp = float(np.float32(1.0e20))
    p.setup.fwd.amplitude = amp
    p.setup.bwd.amplitude = amp
    k = mk.MCKernel(p.setup)
    try:
        ref = k.slave_work(p.verts, p.conn)
    finally:
        k.close()
    paths = {w: str(tmp_path / ("%s.nc4" % w)) for w in ("fwd", "bwd")}
    for w, path in paths.items():
        dbfile.write_problem_db(path, p, w)
    with dbfile.DbFile(paths["fwd"], 10) as ff, dbfile.DbFile(paths["bwd"], 5) as fb:
        assert ff.mapped and fb.mapped
        p.setup.fwd = ff.db
        p.setup.bwd = fb.db
        p.setup.db_page_elems = page
        k = mk.MCKernel(p.setup)
        try:
            got = k.slave_work(p.verts, p.conn)
        finally:
            k.close()
        from oracle import oracle
        p.setup.db_page_elems = 0
        o = oracle.Oracle(p.setup)
        oref = o.slave_work(p.verts, p.conn, halton_offsets=np.arange(len(p.conn), dtype=np.int64)
                            * (p.setup.max_iter + 1) * p.setup.npoints_per_step)
    assert np.abs(ref["values"]).max() > 0
    for key in ("values", "variance", "niterations"):
        np.testing.assert_array_equal(got[key], ref[key], err_msg=key)
    assert (got["stats"]["paged_bytes"] > 0) == (page > 0)
    # the oracle on the databases read back from the files, on the task's stream positions
    assert np.array_equal(got["niterations"], oref["niterations"])
    for j in range(p.nkernel):
        assert _rel(got["values"][:, j], oref["values"][:, j]) < TOL, j
