"""Request loop over the MI355X backend: continuous admission of cluster
solves by device-memory footprint.

A *request* is one cluster solved at a list of wavelengths, at ω and 2ω (the
reference's `Simulation::scan_wavelengths`, Simulation.cpp:604-685, one
request per input file).

The server runs `slots` workers. Each worker owns one device context, with
its own HIP streams and buffers, so the work of concurrent requests overlaps
on the GPU.

Admission is first-come-first-served. A request starts as soon as a slot is
free and its HBM footprint fits in what the admitted requests leave of the
budget: requests that arrive while others run are admitted as slots and
memory free up. While the oldest request waits for memory, a later one whose
footprint is at most an eighth of the budget may start if it fits.

Inside a request the wavelengths are solved as batched sweeps
(`Solver.sweep`, i.e. `omx_solve_sweep`). The T-matrices, the assembled S and V
and their factorisations stay resident from the ω solve to the 2ω solve. A
worker trims its matrix buffers after each request, so the footprint it
holds is the one it was admitted with. A system that the batch cannot solve
is redone alone with partial pivoting; the sweep releases the batch's matrix
buffers before that, so the request then holds one system's footprint, which
is at most the batch's.

The scheduler is plain Python over the C ABI (ctypes releases the GIL during
the calls); `solve_fn` can be replaced for CPU tests of the admission logic.
"""
import collections
import threading
import time

import numpy as np

from . import solver as _solver


def footprint(cluster, batch, kind):
    """Device bytes of a batch of `batch` systems of this cluster: the FP64
    matrix (16 N^2) and, per solver kind, its FP32 copy (8 N^2) or FP64
    factors (16 N^2), for S and (with the second harmonic) V, plus the
    per-system vectors, panel inverses and the two Krylov bases (restart 40)."""
    Np = -(-cluster.N // 64) * 64
    per_matrix = Np * Np * (24 if kind == "lu32ir" else 32)
    mats = 2 if cluster.sh else 1
    vectors = Np * (16 * 16 + 2 * 16 * 2 * 64 + 2 * 16 * 41 + 16 * 32)
    return batch * (mats * per_matrix + vectors)


class Request:
    def __init__(self, rid, arrival, cluster, wavelengths):
        self.id = rid
        self.arrival = float(arrival)
        self.cluster = cluster
        self.wavelengths = np.asarray(wavelengths, dtype=np.float64)


class Result:
    def __init__(self, req):
        self.id = req.id
        self.arrival = req.arrival
        self.admitted = None
        self.finished = None
        self.batch = None
        self.bytes = 0
        self.out = None
        self.error = None

    @property
    def latency(self):
        return None if self.finished is None else self.finished - self.arrival


class Server:
    """slots: concurrent requests (one device context each); budget: device
    bytes the admitted requests may hold (default: 60% of the free memory at
    start); batch_cap: most wavelengths per batched sweep."""

    def __init__(self, slots=3, budget=None, kind="lu32ir", device=0, batch_cap=32, solve_fn=None):
        self.kind = kind
        self.slots = slots
        self.batch_cap = batch_cap
        self._solve_fn = solve_fn
        self._contexts = []
        if solve_fn is None:
            self._contexts = [_solver.Context(device) for _ in range(slots)]
            if budget is None:
                free, _, _ = self._contexts[0].mem_info()
                budget = int(0.6 * free)
        self.budget = int(budget if budget is not None else 1 << 40)
        self._lock = threading.Condition()
        self._on_done = None
        self._held = 0
        self._free_slots = list(range(slots))

    def close(self):
        for c in self._contexts:
            c.close()
        self._contexts = []

    def plan(self, req):
        """Sweep batch size and footprint of a request: the largest batch
        (<= batch_cap, <= its wavelengths) whose footprint fits the budget."""
        b = max(1, min(self.batch_cap, req.wavelengths.size))
        while b > 1 and footprint(req.cluster, b, self.kind) > self.budget:
            b //= 2
        return b, footprint(req.cluster, b, self.kind)

    def _solve(self, slot, req, batch):
        if self._solve_fn is not None:
            return self._solve_fn(slot, req, batch)
        ctx = self._contexts[slot]
        s = _solver.Solver(req.cluster, kind=self.kind, context=ctx)
        try:
            return s.sweep(req.wavelengths, batch=batch, outputs=("x_sca", "x_sca_sh"))
        finally:
            ctx.trim()

    def _worker(self, slot, req, res):
        try:
            res.out = self._solve(slot, req, res.batch)
        except Exception as e:  # reported per request; the server keeps going
            res.error = repr(e)
        res.finished = time.perf_counter() - self._t0
        if self._on_done is not None:
            self._on_done(res)
        with self._lock:
            self._held -= res.bytes
            self._free_slots.append(slot)
            self._lock.notify_all()

    def run(self, requests, realtime=True, deadline=None, on_done=None):
        """Serves the requests (arrival times in seconds from the start;
        realtime=False admits each one as soon as it fits; requests not
        admitted by `deadline` seconds are reported as not served;
        on_done(result) is called from the worker as each request ends).
        Returns the results in request order."""
        self._on_done = on_done
        reqs = sorted(requests, key=lambda r: r.arrival)
        results = {r.id: Result(r) for r in reqs}
        pending = collections.deque(reqs)
        threads = []
        self._t0 = time.perf_counter()
        with self._lock:
            while pending:
                req = pending[0]
                now = time.perf_counter() - self._t0
                if deadline is not None and now > deadline:
                    for r in pending:
                        results[r.id].error = "not admitted before the deadline"
                    pending.clear()
                    break
                if realtime and req.arrival > now:
                    self._lock.wait(timeout=req.arrival - now)
                    continue
                batch, nbytes = self.plan(req)
                if nbytes > self.budget:
                    res = results[req.id]
                    res.error = f"footprint {nbytes} B exceeds the budget {self.budget} B"
                    res.admitted = res.finished = now
                    pending.popleft()
                    continue
                if not self._free_slots:
                    self._lock.wait(timeout=0.05)
                    continue
                if self._held + nbytes > self.budget:
                    # the head waits for memory: a small request that has
                    # arrived and fits now may start beside the running ones
                    # (at most budget / 8, so the head is not starved)
                    small = None
                    for cand in list(pending)[1:]:
                        if realtime and cand.arrival > now:
                            break
                        b2, n2 = self.plan(cand)
                        if n2 <= self.budget // 8 and self._held + n2 <= self.budget:
                            small = (cand, b2, n2)
                            break
                    if small is None:
                        self._lock.wait(timeout=0.05)
                        continue
                    req, batch, nbytes = small
                    pending.remove(req)
                else:
                    pending.popleft()
                slot = self._free_slots.pop()
                self._held += nbytes
                res = results[req.id]
                res.admitted, res.batch, res.bytes = now, batch, nbytes
                th = threading.Thread(target=self._worker, args=(slot, req, res), daemon=True)
                threads.append(th)
                th.start()
        for th in threads:
            th.join()
        return [results[r.id] for r in requests]


def summarize(results, wavelengths_per_request):
    """p50/p99 latency (s), throughput in wavelength-solves (omega + 2 omega)
    per second over the busy span, errors."""
    done = [r for r in results if r.error is None]
    lat = [r.latency for r in done]
    ended = [r.finished for r in results if r.finished is not None]
    span = max(ended) - min(r.arrival for r in results) if ended else 0.0
    solves = sum(wavelengths_per_request[r.id] for r in done)
    return {"requests": len(results), "completed": len(done), "errors": len(results) - len(done),
            "p50_s": float(np.percentile(lat, 50)) if lat else None,
            "p99_s": float(np.percentile(lat, 99)) if lat else None,
            "solves": solves, "span_s": span, "solves_per_s": solves / span if span > 0 else None}
