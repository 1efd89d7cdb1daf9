"""Cahn-Hilliard ADI on the GPU (single solver and continuous-batching server)
against the oracle's restatement of serialCahnADI.c.

Bar: bit-exact fields after every checked step (the kernels perform the
oracle's operations in the same order, unfused), mass conserved to 1e-12,
and each served request bit-identical alone and in a mix."""
import numpy as np
import pytest

import oracle
from custen_amd import api, synth
from tests.helpers import bits_equal, mismatches

pytestmark = pytest.mark.gpu


def ic(n, r):
    return synth.uniform(n * n, 5000 + r, -0.1, 0.1).reshape(n, n)


@pytest.mark.parametrize("n,steps", [(16, 3), (64, 10), (130, 4), (256, 5), (512, 2)])
@pytest.mark.parametrize("mode", [0, 1])
def test_solver_matches_oracle(gpu, n, steps, mode):
    c0 = ic(n, n)
    s = api.CHSolver(api.ch_params(n), c0)
    s.run(steps, mode=mode)
    c, co = s.field()
    s.close()
    ref = oracle.CH(n)
    rc, rco = c0.ravel().copy(), c0.ravel().copy()
    ref.step(rc, rco, steps=steps)
    assert mismatches(c, rc.reshape(n, n)) == 0
    assert mismatches(co, rco.reshape(n, n)) == 0


@pytest.mark.parametrize("n,steps", [(1024, 3), (2048, 2)])
@pytest.mark.parametrize("mode", [0, 1])
def test_solver_matches_oracle_config_sizes(gpu, n, steps, mode):
    """The two largest grid sizes of the served workload, both step modes."""
    test_solver_matches_oracle(gpu, n, steps, mode)


@pytest.mark.parametrize("n,gamma", [(130, 1e-5), (130, 6.3e-4), (130, 4.24e-3), (200, 1e-4),
                                     (600, 1e-6), (200, 8.7155907323626728e-05)])
def test_solver_matches_oracle_coefficients(gpu, n, gamma):
    """Coefficients whose factors settle early or never.  The column solve
    takes the rows of the factors' constant runs as uniform constants; with
    these coefficients the runs start at rows 8, 19, 127 (none: the last row
    alone), 16, 12 and 16, so a sweep has one step that reads the factors and
    the rest constant, or every step reading the factors.  The last pivot is
    one that the two-operation division check refuses, so its constant steps
    take the three-correction division."""
    c0 = ic(n, 3)
    s = api.CHSolver(api.ch_params(n, gamma=gamma), c0)
    s.run(3)
    c, co = s.field()
    s.close()
    ref = oracle.CH(n, gamma=gamma)
    rc, rco = c0.ravel().copy(), c0.ravel().copy()
    ref.step(rc, rco, steps=3)
    assert mismatches(c, rc.reshape(n, n)) == 0
    assert mismatches(co, rco.reshape(n, n)) == 0


@pytest.mark.parametrize("n", [256, 2048])
def test_mass_conservation_200_steps(gpu, n):
    c0 = ic(n, 1)
    s = api.CHSolver(api.ch_params(n), c0)
    s.run(200)
    c, _ = s.field()
    s.close()
    assert abs(c.mean() - c0.mean()) <= 1e-12
    assert np.all(np.isfinite(c))


def test_server_alone_vs_mix(gpu):
    sizes = [32, 64, 16, 128, 64, 48]
    steps = [7, 5, 9, 4, 6, 8]
    ics = [ic(n, r) for r, n in enumerate(sizes)]
    srv = api.CHServer(budget_bytes=3 * api.ch_footprint(128), max_slots=3)
    ids = [srv.submit(api.ch_params(n), k, c) for n, k, c in zip(sizes, steps, ics)]
    srv.run_until_idle()
    mix = [srv.result(i) for i in ids]
    for i in ids:
        assert srv.query(i)[0] == 2
    srv.close()
    for r, (n, k, c) in enumerate(zip(sizes, steps, ics)):
        alone = api.CHServer(budget_bytes=api.ch_footprint(n), max_slots=1)
        rid = alone.submit(api.ch_params(n), k, c)
        alone.run_until_idle()
        assert bits_equal(alone.result(rid), mix[r])
        alone.close()
        ref = oracle.CH(n)
        rc, rco = c.ravel().copy(), c.ravel().copy()
        ref.step(rc, rco, steps=k)
        assert mismatches(mix[r], rc.reshape(n, n)) == 0


def test_server_results_moved_to_pageable(gpu, monkeypatch):
    """With no room for results in pinned memory every finished result moves
    to pageable memory after its copy; the results are unchanged and can be
    read more than once."""
    monkeypatch.setenv("CUSTEN_CH_PINNED_RESULT_BYTES", "0")
    sizes = [32, 48, 32, 64]
    steps = [3, 5, 4, 2]
    ics = [ic(n, 70 + r) for r, n in enumerate(sizes)]
    srv = api.CHServer(budget_bytes=4 * api.ch_footprint(64), max_slots=4)
    ids = [srv.submit(api.ch_params(n), k, c) for n, k, c in zip(sizes, steps, ics)]
    srv.run_until_idle()
    for r, (n, k, c) in enumerate(zip(sizes, steps, ics)):
        first = srv.result(ids[r])
        assert bits_equal(first, srv.result(ids[r]))
        ref = oracle.CH(n)
        rc, rco = c.ravel().copy(), c.ravel().copy()
        ref.step(rc, rco, steps=k)
        assert mismatches(first, rc.reshape(n, n)) == 0
    srv.close()


def test_server_alone_vs_mix_config_sizes(gpu):
    """A mix that includes the served workload's 1024 and 2048 grids: each
    result is the same alone and in the mix, and equals the oracle."""
    sizes = [2048, 64, 1024, 256, 16, 512]
    steps = [2, 5, 3, 4, 6, 2]
    ics = [ic(n, 40 + r) for r, n in enumerate(sizes)]
    budget = sum(api.ch_footprint(n) for n in sizes)
    srv = api.CHServer(budget_bytes=budget, max_slots=4)
    ids = [srv.submit(api.ch_params(n), k, c) for n, k, c in zip(sizes, steps, ics)]
    srv.run_until_idle()
    mix = [srv.result(i) for i in ids]
    srv.close()
    for r, (n, k, c) in enumerate(zip(sizes, steps, ics)):
        print(f"alone vs mix: n={n} steps={k}", flush=True)
        alone = api.CHServer(budget_bytes=api.ch_footprint(n), max_slots=1)
        rid = alone.submit(api.ch_params(n), k, c)
        alone.run_until_idle()
        assert bits_equal(alone.result(rid), mix[r])
        alone.close()
        ref = oracle.CH(n)
        rc, rco = c.ravel().copy(), c.ravel().copy()
        ref.step(rc, rco, steps=k)
        assert mismatches(mix[r], rc.reshape(n, n)) == 0


def test_server_admission_budget(gpu):
    fp = api.ch_footprint(64)
    srv = api.CHServer(budget_bytes=fp + fp // 2, max_slots=8)  # room for one 64^2 request
    ids = [srv.submit(api.ch_params(64), 3, ic(64, r)) for r in range(3)]
    left = srv.step()
    states = [srv.query(i)[0] for i in ids]
    assert states.count(1) == 1 and states.count(0) == 2 and left == 3
    srv.run_until_idle()
    assert all(srv.query(i)[0] == 2 for i in ids)
    with pytest.raises(api.CustenError) as ei:
        srv.submit(api.ch_params(256), 3, ic(256, 9))
    assert ei.value.code == 5
    srv.close()


def test_modes_interleave(gpu):
    """Fused (mode 0) and reference-shaped (mode 1) steps can alternate on one
    simulation; the result equals the oracle after the total step count."""
    n = 96
    c0 = ic(n, 3)
    s = api.CHSolver(api.ch_params(n), c0)
    s.run(3, mode=0)
    s.run(2, mode=1)
    s.run(2, mode=0)
    c, co = s.field()
    s.close()
    ref = oracle.CH(n)
    rc, rco = c0.ravel().copy(), c0.ravel().copy()
    ref.step(rc, rco, steps=7)
    assert mismatches(c, rc.reshape(n, n)) == 0
    assert mismatches(co, rco.reshape(n, n)) == 0
