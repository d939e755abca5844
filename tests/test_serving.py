"""Request loop: admission by device-memory footprint (CPU, with a stand-in
solver) and, on the GPU, that every request's result is the same alone and
in a concurrent mix."""
import threading
import time

import numpy as np
import pytest

from optimet_amd.problem import drude_cluster
from optimet_amd.serving import Request, Server, footprint, summarize


def _req(rid, arrival, P=3, nmax=3, nl=2, seed=1):
    c = drude_cluster(P, nmax, 700e-9, 40, 80, 500, 10, seed=seed + rid)
    return Request(rid, arrival, c, np.linspace(500e-9, 800e-9, nl))


class _Fake:
    """Stand-in solver: sleeps, records how much footprint is held at once."""

    def __init__(self, server, dt=0.05):
        self.server, self.dt = server, dt
        self.lock = threading.Lock()
        self.running = 0
        self.peak_running = 0
        self.peak_held = 0
        self.order = []

    def __call__(self, slot, req, batch):
        with self.lock:
            self.running += 1
            self.order.append(req.id)
            self.peak_running = max(self.peak_running, self.running)
            self.peak_held = max(self.peak_held, self.server._held)
        time.sleep(self.dt)
        with self.lock:
            self.running -= 1
        return {"slot": slot, "batch": batch}


def _server(slots, budget):
    s = Server(slots=slots, budget=budget, solve_fn=lambda *a: None)
    fake = _Fake(s)
    s._solve_fn = fake
    return s, fake


def test_admission_respects_budget_and_order():
    reqs = [_req(i, 0.0) for i in range(6)]
    one = footprint(reqs[0].cluster, 2, "lu32ir")
    s, fake = _server(slots=4, budget=int(2.5 * one))  # room for two requests at a time
    res = s.run(reqs, realtime=False)
    assert all(r.error is None for r in res)
    assert fake.peak_held <= s.budget
    assert fake.peak_running == 2
    assert fake.order == sorted(fake.order)  # first come, first served


def test_slots_bound_concurrency_and_arrivals_are_honoured():
    reqs = [_req(i, 0.02 * i) for i in range(5)]
    s, fake = _server(slots=2, budget=1 << 50)
    res = s.run(reqs, realtime=True)
    assert fake.peak_running == 2
    for r in res:
        assert r.admitted >= r.arrival - 1e-3 and r.finished > r.admitted
    summ = summarize(res, {r.id: 2 for r in reqs})
    assert summ["completed"] == 5 and summ["p99_s"] >= summ["p50_s"] > 0


def test_oversize_request_is_rejected_not_queued_forever():
    big = _req(0, 0.0, P=40, nmax=8)
    small = _req(1, 0.0)
    s, fake = _server(slots=2, budget=footprint(small.cluster, 2, "lu32ir") + 1)
    res = s.run([big, small], realtime=False)
    assert res[0].error is not None and "exceeds" in res[0].error
    assert res[1].error is None


def test_batch_shrinks_to_fit():
    req = _req(0, 0.0, nl=16)
    per = footprint(req.cluster, 1, "lu32ir")
    s = Server(slots=1, budget=5 * per, solve_fn=lambda *a: None)
    b, nbytes = s.plan(req)
    assert b == 4 and nbytes <= s.budget


def test_admitted_footprint_covers_the_one_at_a_time_fallback():
    """A sweep that redoes a system alone (omx_solve_sweep's partial-pivoting
    fallback) first releases the batch's matrix buffers, so the request peaks
    at max(footprint(batch), footprint(1)): the footprint plan() admits it
    with covers that peak for every batch size it can choose, down to a
    budget that fits one system only."""
    req = _req(0, 0.0, P=6, nmax=5, nl=16)
    one = footprint(req.cluster, 1, "lu32ir")
    for budget in (one, 3 * one, 16 * one, 64 * one):
        s = Server(slots=1, budget=budget, solve_fn=lambda *a: None)
        b, nbytes = s.plan(req)
        assert nbytes <= budget
        assert nbytes >= max(footprint(req.cluster, b, "lu32ir"), one)
    for kind in ("lu64", "lu32ir"):
        for b in (1, 2, 7, 32):
            assert footprint(req.cluster, b, kind) >= footprint(req.cluster, 1, kind)


@pytest.mark.gpu
def test_results_same_alone_and_in_the_mix():
    """Each request solved alone (one slot, one at a time) and in a mix of
    concurrent requests (three slots, all arriving together) gives the same
    coefficients and cross sections to 1e-10 (the same kernels and
    iterations run; only their interleaving on the device differs)."""
    reqs = [_req(0, 0.0, P=4, nmax=4, nl=3), _req(1, 0.0, P=6, nmax=3, nl=2), _req(2, 0.0, P=3, nmax=5, nl=4),
            _req(3, 0.0, P=5, nmax=4, nl=1)]
    alone = {}
    srv = Server(slots=1, budget=8 << 30, kind="lu32ir")
    try:
        for r in reqs:
            (res,) = srv.run([r], realtime=False)
            assert res.error is None, res.error
            alone[r.id] = res.out
    finally:
        srv.close()
    srv = Server(slots=3, budget=8 << 30, kind="lu32ir")
    try:
        mixed = srv.run(reqs, realtime=False)
    finally:
        srv.close()
    for res in mixed:
        assert res.error is None, res.error
        a = alone[res.id]
        for key in ("x_sca", "x_sca_sh"):
            d = np.linalg.norm(res.out[key] - a[key]) / np.linalg.norm(a[key])
            assert d < 1e-10, (res.id, key, d)
        assert np.allclose(res.out["xs"], a["xs"], rtol=1e-10, atol=0)


def test_small_request_backfills_while_head_waits_for_memory():
    big = [_req(i, 0.0, P=6, nmax=4) for i in range(2)]
    small = _req(2, 0.0, P=2, nmax=2, nl=1)
    fb = footprint(big[0].cluster, 2, "lu32ir")
    s, fake = _server(slots=3, budget=int(1.5 * fb))  # one big request at a time
    fake.dt = 0.1
    res = s.run(big + [small], realtime=False)
    assert all(r.error is None for r in res)
    # the small request started before the second big one
    assert res[2].admitted < res[1].admitted


def test_result_cache_answers_repeated_wavelengths():
    """A wavelength of a cluster solved before is answered from the cache:
    only the new wavelengths reach the solver, results come back in request
    order, and the hit rate counts per wavelength.  A cluster that differs in
    any input (here one permittivity) is a different key."""
    s = Server(slots=1, budget=1 << 40, solve_fn=lambda *a: None)
    s._solve_fn = None
    seen = []

    def device(slot, req, batch, lams):
        seen.append(list(lams))
        n = lams.size
        return {"x_sca": np.outer(lams, np.arange(3.0)) + 1j, "xs": np.tile(lams[:, None], (1, 5)),
                "iterations": np.zeros((n, 2), np.int32)}

    s._device_solve = device
    a = _req(0, 0.0, nl=3)
    b = Request(1, 0.0, a.cluster, np.array([a.wavelengths[2], 900e-9, a.wavelengths[0]]))
    res = s.run([a, b], realtime=False)
    assert all(r.error is None for r in res), [r.error for r in res]
    assert seen == [list(a.wavelengths), [900e-9]]
    np.testing.assert_array_equal(res[1].out["xs"][:, 0], b.wavelengths)
    np.testing.assert_array_equal(res[1].out["x_sca"][2], res[0].out["x_sca"][0])
    st = s.cache.stats()
    assert (st["hits"], st["misses"]) == (2, 4)
    c = _req(0, 0.0, nl=3)
    c.cluster.eps_r = c.cluster.eps_r.copy()
    c.cluster.eps_r[0] += 1e-12
    s.run([Request(2, 0.0, c.cluster, a.wavelengths)], realtime=False)
    assert seen[-1] == list(a.wavelengths)


def test_small_requests_are_grouped_oldest_first():
    """With grouping on, small requests that have arrived go into one job of
    at most `group` systems (one per request and wavelength), oldest first; a
    request above the size limit and one not yet arrived are not taken in."""
    reqs = [_req(i, 0.0, nl=2) for i in range(5)] + [_req(5, 0.0, P=40, nmax=8, nl=1), _req(6, 0.3, nl=1)]
    s, fake = _server(slots=2, budget=1 << 50)
    s.group = 6
    jobs = []

    def group_fn(slot, grp):
        jobs.append([r.id for r in grp])
        time.sleep(0.05)
        return {r.id: {"group": len(grp)} for r in grp}

    s._group_fn = group_fn
    res = s.run(reqs, realtime=True)
    assert all(r.error is None for r in res), [r.error for r in res]
    assert jobs[0] == [0, 1, 2]  # 6 systems: three requests of two wavelengths
    assert sorted(i for j in jobs for i in j) == [0, 1, 2, 3, 4]
    assert all(len(j) >= 2 for j in jobs)
    assert fake.order == [5, 6]  # the large one and the late single one run alone
    assert res[3].out["group"] == 2 and res[0].admitted <= res[3].admitted
    assert s.group_footprint([reqs[0], reqs[1]]) >= footprint(reqs[0].cluster, 4, "lu32ir")


def test_group_respects_the_budget():
    reqs = [_req(i, 0.0, nl=1) for i in range(6)]
    s, fake = _server(slots=3, budget=1 << 50)
    s.group = 8
    s.budget = s.group_footprint(reqs[:3])  # room for three systems per job
    sizes = []
    s._group_fn = lambda slot, grp: (sizes.append(len(grp)), {r.id: {} for r in grp})[1]
    res = s.run(reqs, realtime=False)
    assert all(r.error is None for r in res)
    assert max(sizes) <= 3 and sum(sizes) + len(fake.order) == 6


@pytest.mark.gpu
def test_grouped_requests_match_their_sweeps():
    """Small requests served in one group job (omx_solve_group over every
    request and wavelength) give the coefficients and cross sections of the
    same requests served one at a time by batched sweeps, to 1e-10."""
    reqs = [_req(0, 0.0, P=4, nmax=4, nl=2), _req(1, 0.0, P=6, nmax=3, nl=1), _req(2, 0.0, P=3, nmax=5, nl=3),
            _req(3, 0.0, P=2, nmax=6, nl=2)]
    print("grouped requests: N =", [r.cluster.N for r in reqs], flush=True)
    alone = {}
    srv = Server(slots=1, budget=8 << 30, kind="lu32ir")
    try:
        for r in reqs:
            (res,) = srv.run([r], realtime=False)
            assert res.error is None, res.error
            alone[r.id] = res.out
    finally:
        srv.close()
    srv = Server(slots=2, budget=8 << 30, kind="lu32ir", group=16)
    try:
        grouped = srv.run(reqs, realtime=False)
        assert srv.groups_run == 1
    finally:
        srv.close()
    for res in grouped:
        assert res.error is None, res.error
        assert res.batch == 8  # every request's wavelengths in one job
        a = alone[res.id]
        for key in ("x_sca", "x_sca_sh"):
            d = np.linalg.norm(res.out[key] - a[key]) / np.linalg.norm(a[key])
            assert d < 1e-10, (res.id, key, d)
        assert np.allclose(res.out["xs"], a["xs"], rtol=1e-10, atol=0)
