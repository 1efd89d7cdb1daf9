"""Cahn-Hilliard checkpoint / restart and snapshot output on the GPU.

Bar: a solver restored from a checkpoint continues bit for bit as the
original (and as an uninterrupted run); snapshots load with numpy.load and
hold the current field exactly; a served request resumed from both time levels
finishes identical to the uninterrupted request."""
import numpy as np
import pytest

import oracle
from custen_amd import analysis, api, synth
from tests.helpers import bits_equal, mismatches

pytestmark = pytest.mark.gpu


def ic(n, r):
    return synth.uniform(n * n, 5000 + r, -0.1, 0.1).reshape(n, n)


@pytest.mark.parametrize("n", [64, 130])
def test_restore_continues_bit_for_bit(gpu, n, tmp_path):
    c0 = ic(n, 3)
    path = str(tmp_path / "ch.ckpt")
    a = api.CHSolver(api.ch_params(n), c0)
    a.run(7)
    a.checkpoint(path)
    assert a.steps_done == 7
    a.run(9)
    ca, coa = a.field()
    a.close()
    b = api.CHSolver.restore(path)
    assert b.n == n and b.steps_done == 7
    b.run(9)
    assert b.steps_done == 16
    cb, cob = b.field()
    b.close()
    assert bits_equal(ca, cb), mismatches(ca, cb)
    assert bits_equal(coa, cob)
    # and the uninterrupted oracle
    rc, rco = c0.ravel().copy(), c0.ravel().copy()
    oracle.CH(n).step(rc, rco, steps=16)
    assert bits_equal(cb, rc.reshape(n, n))
    # checkpoint size: header + two fields
    assert (tmp_path / "ch.ckpt").stat().st_size == 56 + 2 * n * n * 8


def test_restore_then_plain_mode(gpu, tmp_path):
    """A restored solver also continues identically on the reference-shaped path."""
    n = 64
    c0 = ic(n, 4)
    a = api.CHSolver(api.ch_params(n), c0)
    a.run(3, mode=1)
    a.checkpoint(str(tmp_path / "x.ckpt"))
    a.run(4, mode=1)
    ca, _ = a.field()
    a.close()
    b = api.CHSolver.restore(str(tmp_path / "x.ckpt"))
    b.run(4, mode=0)
    cb, _ = b.field()
    b.close()
    assert bits_equal(ca, cb)


def test_snapshot_is_numpy_readable(gpu, tmp_path):
    n = 96
    s = api.CHSolver(api.ch_params(n), ic(n, 5))
    s.run(5)
    p = api.ch_params(n)
    t = 5 * p.dtOverDx * p.lx / n
    path = tmp_path / analysis.snapshot_name(t)
    s.snapshot(str(path))
    c, _ = s.field()
    s.close()
    snap = np.load(path, allow_pickle=False)
    assert snap.dtype == np.float64 and snap.shape == (n, n)
    assert bits_equal(snap, c)
    (t_loaded, field), = analysis.load_series(str(tmp_path))
    assert abs(t_loaded - t) < 1e-9 and bits_equal(field, c)
    stats = analysis.coarsening(field)
    assert np.isfinite(stats["s"]) and np.isfinite(stats["inv_k1"])


def test_server_resume_matches_uninterrupted(gpu, tmp_path):
    n = 64
    c0 = ic(n, 6)
    a = api.CHSolver(api.ch_params(n), c0)
    a.run(4)
    c4, co4 = a.field()
    a.run(6)
    c10, _ = a.field()
    a.close()
    srv = api.CHServer(budget_bytes=4 * api.ch_footprint(n), max_slots=2)
    other = srv.submit(api.ch_params(48), 8, ic(48, 9))
    rid = srv.submit(api.ch_params(n), 6, c4, c_old=co4)
    srv.run_until_idle()
    got = srv.result(rid)
    assert srv.query(other)[0] == 2
    srv.close()
    assert bits_equal(got, c10), mismatches(got, c10)


def test_field_levels_every_residue(gpu, tmp_path):
    """The solver keeps its time levels in three buffers used in turn, and a
    run of k steps fuses each step's final phase with the next step's rhs
    except the last.  Runs of 1 to 5 steps in a row reach every buffer
    residue and every run shape; after each run both levels and a
    checkpoint written then equal the oracle."""
    n = 96
    c0 = ic(n, 11)
    s = api.CHSolver(api.ch_params(n), c0)
    rc, rco = c0.ravel().copy(), c0.ravel().copy()
    ref = oracle.CH(n)
    done = 0
    for k in (1, 2, 3, 1, 4, 5, 2):
        s.run(k)
        ref.step(rc, rco, steps=k)
        done += k
        c, co = s.field()
        assert mismatches(c, rc.reshape(n, n)) == 0, (done, k)
        assert mismatches(co, rco.reshape(n, n)) == 0, (done, k)
        path = str(tmp_path / f"ch{done}.ckpt")
        s.checkpoint(path)
        r = api.CHSolver.restore(path)
        cr, cor = r.field()
        r.close()
        assert bits_equal(cr, c) and bits_equal(cor, co), done
    s.close()


def test_server_short_requests(gpu):
    """Requests of 1 to 4 steps share a lane, so its batches are of every
    length from 1 (a lone rhs and final) up; each result equals the oracle."""
    sizes = [48, 48, 48, 48, 32]
    steps = [1, 2, 3, 4, 1]
    ics = [ic(n, 60 + r) for r, n in enumerate(sizes)]
    srv = api.CHServer(budget_bytes=8 * api.ch_footprint(48), max_slots=8)
    ids = [srv.submit(api.ch_params(n), k, c) for n, k, c in zip(sizes, steps, ics)]
    srv.run_until_idle()
    for r, (n, k, c) in enumerate(zip(sizes, steps, ics)):
        rc, rco = c.ravel().copy(), c.ravel().copy()
        oracle.CH(n).step(rc, rco, steps=k)
        assert mismatches(srv.result(ids[r]), rc.reshape(n, n)) == 0, (n, k)
    srv.close()
