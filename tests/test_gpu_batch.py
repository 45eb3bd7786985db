"""GPU tests of the batched multi-solve (mfs_batch_*): every solve of a
batch gives the same bits as the same solve run alone, also when solves
join and leave the batch between steps."""
import numpy as np
import pytest

import minifs_amd as mfs

pytestmark = pytest.mark.gpu

FIELDS = ("rho", "u", "v", "w", "p", "res", "dt", "dq", "dqdx", "dqdy", "dqdz")


def _case(n, mach, alpha, seed, cells="tet"):
    p = mfs.make_params(mach=mach, alpha=alpha, cfl=10.0)
    g = mfs.Grid.box_bump(n, seed=seed, cells=cells).decouple().preprocess(p)
    return g, p


def _alone(g, p, iters, precision):
    s = mfs.Solver(g, p, precision=precision)
    s.step(iters)
    out = {nm: s.field(nm) for nm in FIELDS}
    h = s.history()
    s.close()
    return out, h


def _same(member, ref, href):
    for nm in FIELDS:
        np.testing.assert_array_equal(member.field(nm), ref[nm], err_msg=nm)
    np.testing.assert_array_equal(member.history(), href)


@pytest.mark.parametrize("precision", [mfs.MFS_FP64, mfs.MFS_MIXED])
def test_batch_members_equal_standalone(precision):
    print(f"[batch] precision={precision}", flush=True)
    cases = [_case(6, 0.5, 0.0, 1), _case(9, 0.8, 1.25, 2), _case(4, 0.3, -3.0, 3)]
    b = mfs.Batch(precision)
    members = [b.add(g, p) for g, p in cases]
    assert len(b) == 3
    b.step(4)
    b.step(3)
    for m, (g, p) in zip(members, cases):
        ref, href = _alone(g, p, 7, precision)
        _same(m, ref, href)
    b.close()


@pytest.mark.parametrize("precision,mode,cells", [(mfs.MFS_FP64, 2, "tet"), (mfs.MFS_MIXED, 4, "tet"),
                                                  (mfs.MFS_MIXED, 2, "prism"), (mfs.MFS_FP64, 4, "hex")])
def test_batch_gradient_on_second_stream_equals_standalone(precision, mode, cells):
    """A batch whose gradient runs on a second stream (after the update, or
    after the next iteration's forward sweep), with a solve joining and one
    leaving between steps: every member equals the same solve alone."""
    print(f"[batch] gradient stream {mode} precision={precision} {cells}", flush=True)
    a, c, d = _case(7, 0.6, 1.0, 31, cells), _case(5, 0.7, 2.0, 32, cells), _case(6, 0.45, -1.0, 33, cells)
    b = mfs.Batch(precision, gradient_stream=mode)
    ma = b.add(*a)
    mc = b.add(*c)
    b.step(3)
    md = b.add(*d)
    b.step(2)
    b.remove(mc)
    b.step(4)
    for m, case, iters in ((ma, a, 9), (md, d, 6)):
        ref, href = _alone(*case, iters, precision)
        _same(m, ref, href)
    b.close()


def test_batch_join_and_leave():
    a, c, d = _case(7, 0.6, 1.0, 11), _case(5, 0.7, 2.0, 12), _case(8, 0.45, -1.0, 13)
    b = mfs.Batch(mfs.MFS_FP64)
    ma = b.add(*a)
    mc = b.add(*c)
    b.step(3)
    md = b.add(*d)      # joins after 3 steps
    b.step(2)
    b.remove(mc)        # leaves after 5 steps
    b.step(4)
    for m, case, iters in ((ma, a, 9), (md, d, 6)):
        ref, href = _alone(*case, iters, mfs.MFS_FP64)
        _same(m, ref, href)
    assert len(b) == 2
    b.close()


def test_batch_polyhedral_family():
    cases = [_case(4, 0.5, 0.0, 21, "hex"), _case(3, 0.8, 1.0, 22, "prism")]
    b = mfs.Batch(mfs.MFS_FP64)
    members = [b.add(g, p) for g, p in cases]
    b.step(5)
    for m, (g, p) in zip(members, cases):
        ref, href = _alone(g, p, 5, mfs.MFS_FP64)
        _same(m, ref, href)
    # a tetrahedral mesh cannot join a batch of 6-face cells
    g, p = _case(3, 0.5, 0.0, 23)
    with pytest.raises(mfs.MfsError):
        b.add(g, p)
    b.close()


@pytest.mark.parametrize("precision", [mfs.MFS_FP64, mfs.MFS_MIXED])
def test_batch_ghost_refresh_equals_standalone(precision):
    """SetGhostVariables after each update (mfs_ctx_set_ghost_refresh) runs
    for the batch members that ask for it, with the same bits as alone."""
    print(f"[batch] ghost refresh precision={precision}", flush=True)
    cases = [_case(6, 0.5, 2.0, 5), _case(7, 0.7, -1.0, 6)]
    refs = []
    for g, p in cases:
        s = mfs.Solver(g, p, precision=precision)
        s.ghost_refresh(True)
        s.step(7)
        refs.append(({nm: s.field(nm) for nm in FIELDS}, s.history()))
        s.close()
    b = mfs.Batch(precision)
    members = [b.add(g, p) for g, p in cases]
    for m in members:
        m.ghost_refresh(True)
    b.step(7)
    b.sync()
    for m, (ref, href) in zip(members, refs):
        _same(m, ref, href)
    b.close()
    for g, _ in cases:
        g.close()


@pytest.mark.parametrize("precision", [mfs.MFS_FP64, mfs.MFS_MIXED])
def test_attach_continues_standalone_and_threaded_contexts(precision):
    """mfs_batch_attach: a context stepped alone and then attached continues
    bit for bit, and one built on another thread while the batch steps joins
    like mfs_batch_add."""
    import concurrent.futures as cf

    print(f"[batch] attach precision={precision}", flush=True)
    a, c, t = _case(7, 0.6, 1.0, 31), _case(6, 0.75, 2.5, 32), _case(9, 0.4, -2.0, 33)
    b = mfs.Batch(precision)
    mc = b.add(*c)
    s = mfs.Solver(*a, precision=precision)
    s.step(3)
    with cf.ThreadPoolExecutor(1) as pool:
        fut = pool.submit(mfs.Solver, *t, precision=precision)
        b.step(2)
        st = fut.result()
    ma = b.attach(s)
    assert s._h is None
    mt = b.attach(st)
    assert len(b) == 3
    b.step(4)
    for m, case, iters in ((ma, a, 7), (mc, c, 6), (mt, t, 4)):
        ref, href = _alone(*case, iters, precision)
        _same(m, ref, href)
    b.close()


def test_attach_rejects_another_precision():
    g, p = _case(4, 0.5, 0.0, 41)
    b = mfs.Batch(mfs.MFS_MIXED)
    s = mfs.Solver(g, p, precision=mfs.MFS_FP64)
    with pytest.raises(mfs.MfsError):
        b.attach(s)
    s.step(2)  # still the caller's context
    assert len(s.history()) == 2
    s.close()
    assert len(b) == 0
    b.close()


@pytest.mark.parametrize("precision", [mfs.MFS_FP64, mfs.MFS_MIXED])
def test_batch_with_symmetry_and_interface_patches(tmp_path, precision):
    """Solves whose patches are remapped to SYMM and INTERFACE (their
    symmetry-plane node projection runs per solve in the batch) next to a
    plain one: the same bits as each alone."""
    print(f"[batch] symm/interface precision={precision}", flush=True)
    cases = []
    for k, bc_text in enumerate(("2\n1 3\n2 4\n", "2\n1 4\n2 10\n", None)):
        g0 = mfs.Grid.box_bump(6 + k, seed=40 + k).decouple()
        p = mfs.make_params(mach=0.6, alpha=1.5, cfl=10.0)
        if bc_text is None:
            cases.append((g0.preprocess(p), p))
            continue
        mesh, bc = str(tmp_path / f"m{k}.mfl"), str(tmp_path / f"m{k}.bc")
        g0.save(mesh, bc)
        g0.close()
        open(bc, "w").write(bc_text)
        cases.append((mfs.Grid.load(mesh, bc).preprocess(p), p))
    b = mfs.Batch(precision)
    members = [b.add(g, p) for g, p in cases]
    b.step(6)
    for m, (g, p) in zip(members, cases):
        ref, href = _alone(g, p, 6, precision)
        _same(m, ref, href)
    b.close()


@pytest.mark.parametrize("precision", [mfs.MFS_FP64, mfs.MFS_MIXED])
def test_batch_of_large_solves_equals_standalone(precision):
    """Solves of 94k, 162k, 384k and 584k cells (thousands of chunks each,
    long interleaved item lists, many XCD groups per solve): the same bits
    as alone after a few fused iterations.  Alone, the 584k-cell solve
    (2,282 cell blocks) reduces its residual norms in two launches, the
    threads' folds on many CUs and then the tree; in the batch one workgroup
    does both: the history records are the same bits."""
    print(f"[batch] large solves precision={precision}", flush=True)
    cases = [_case(25, 0.5, 0.0, 61), _case(30, 0.8, 1.25, 62), _case(40, 0.4, -2.0, 63), _case(46, 0.7, 2.0, 64)]
    b = mfs.Batch(precision)
    members = [b.add(g, p) for g, p in cases]
    b.step(3)
    b.step(2)
    for m, (g, p) in zip(members, cases):
        ref, href = _alone(g, p, 5, precision)
        _same(m, ref, href)
    b.close()
    for g, _ in cases:
        g.close()
