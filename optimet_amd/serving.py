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

With kind "auto" a request of N >= 16384 unknowns is solved wavelength by
wavelength with the automatic solver choice (the mixed Krylov within the
mixed LU's cost, else the mixed LU on the same matrices), one system at a
time; smaller requests use batched sweeps with the automatic choice (V's
batched mixed LU and, from N = 2048 with the second harmonic, S by lockstep
FP64 GMRES beside it).

With grouping on (`group` > 0), small requests (N <= GROUP_N, at most
GROUP_LAMBDAS wavelengths, N within GROUP_RATIO of the oldest one's) that
have arrived are gathered, oldest first, into
one job of at most `group` systems (one per request and wavelength): each
system is assembled in a member context of the slot's pool, and the members
are solved together by `omx_solve_group` (batched LU and lockstep refinement
over the identity-bordered systems, in the slot's context). Systems of
different requests, clusters and sizes thus share one batched launch. The
job's footprint is the members' own buffers plus the batch.

Results are cached by content: a wavelength of a cluster (geometry,
materials, excitation) that an earlier request solved is answered from the
cache without device work, and the server counts hits and misses.

The scheduler is plain Python over the C ABI (ctypes releases the GIL during
the calls); `solve_fn` can be replaced for CPU tests of the admission logic.
"""
import collections
import hashlib
import threading
import time

import numpy as np

from . import solver as _solver


KRYLOV_N = 16384  # "auto": requests from this order on are solved one wavelength at a time
GROUP_N = 4096     # with grouping on: requests up to this order ...
GROUP_LAMBDAS = 8  # ... and this many wavelengths may be solved in one group,
GROUP_RATIO = 1.5  # with orders within this factor of the oldest one's (padding costs flops)


def footprint(cluster, batch, kind):
    """Device bytes of a batch of `batch` systems of this cluster: the FP64
    matrix (16 N^2) and, per solver kind, its FP32 copy (8 N^2) or FP64
    factors (16 N^2), for S and (with the second harmonic) V, plus the
    per-system vectors, panel inverses and the two Krylov bases (restart 40);
    "auto" is the mixed LU's footprint (the Krylov path holds the same
    matrices plus a 31-vector inner basis)."""
    Np = -(-cluster.N // 64) * 64
    per_matrix = Np * Np * (32 if kind == "lu64" else 24)
    if kind == "auto" and cluster.N >= KRYLOV_N:
        batch = 1
        per_matrix += Np * 16 * 32 // (2 if cluster.sh else 1)
    mats = 2 if cluster.sh else 1
    vectors = Np * (16 * 16 + 2 * 16 * 2 * 64 + 2 * 16 * 41 + 16 * 32)
    return batch * (mats * per_matrix + vectors)


class Request:
    def __init__(self, rid, arrival, cluster, wavelengths):
        self.id = rid
        self.arrival = float(arrival)
        self.cluster = cluster
        self.wavelengths = np.asarray(wavelengths, dtype=np.float64)


def cluster_key(c):
    """Content hash of everything a solve depends on except the wavelength."""
    h = hashlib.sha1()
    for a in (c.xyz, c.radius, c.eps_r, c.eps_r_sh, c.mu_r, c.ksippp, c.ksiparppar, c.gamma):
        h.update(np.ascontiguousarray(a).tobytes())
    h.update(repr((c.nmax, bool(c.sh), c.bg_eps_r, c.bg_mu_r, c.theta_k, c.phi_k, c.etheta, c.ephi,
                   c.meta.get("drude"))).encode())
    return h.hexdigest()


class ResultCache:
    """Per-wavelength results (x_sca, x_sca_sh, cross sections, iterations)
    of solved clusters, keyed by content, least recently used dropped past
    `max_bytes` of host memory."""

    def __init__(self, max_bytes=4 << 30):
        self.max_bytes = max_bytes
        self.bytes = 0
        self.hits = 0
        self.misses = 0
        self._d = collections.OrderedDict()
        self._lock = threading.Lock()

    def get(self, key):
        with self._lock:
            v = self._d.get(key)
            if v is None:
                self.misses += 1
                return None
            self._d.move_to_end(key)
            self.hits += 1
            return v

    def put(self, key, value):
        size = sum(a.nbytes for a in value.values() if isinstance(a, np.ndarray))
        with self._lock:
            if key in self._d:
                return
            self._d[key] = value
            self.bytes += size
            while self.bytes > self.max_bytes and self._d:
                _, old = self._d.popitem(last=False)
                self.bytes -= sum(a.nbytes for a in old.values() if isinstance(a, np.ndarray))

    def stats(self):
        n = self.hits + self.misses
        return {"hits": self.hits, "misses": self.misses, "hit_rate": self.hits / n if n else 0.0,
                "entries": len(self._d), "bytes": self.bytes}


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

    def __init__(self, slots=3, budget=None, kind="lu32ir", device=0, batch_cap=32, solve_fn=None,
                 cache_bytes=4 << 30, profile=False, group=0, group_fn=None):
        self.kind = kind
        self.group = 0 if profile else int(group)  # most systems per group job (0: no grouping)
        self._group_fn = group_fn  # stand-in for a group job (CPU tests)
        self._pools = {}  # slot -> member contexts of its group jobs
        self.device = device
        self.groups_run = 0
        self.cache = ResultCache(cache_bytes)
        self.profile = profile  # per-launch HIP events; each request's counters kept in prof_log
        self.prof_log = []
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
        for pool in self._pools.values():
            for c in pool:
                c.close()
        self._contexts = []
        self._pools = {}

    def groupable(self, req):
        return (self.group > 0 and self.kind in ("lu32ir", "lu64", "auto") and req.cluster.N <= GROUP_N
                and req.wavelengths.size <= GROUP_LAMBDAS)

    def group_footprint(self, reqs):
        """Members' own buffers (one system each) plus the batch at the
        largest member's padded order."""
        kind = "lu64" if self.kind == "lu64" else "lu32ir"
        members = [r.cluster for r in reqs for _ in range(r.wavelengths.size)]
        big = max(members, key=lambda c: c.N)
        return sum(footprint(c, 1, kind) for c in members) + footprint(big, len(members), kind)

    def _solve_group(self, slot, reqs):
        """One group job: every (request, wavelength) not in the cache as a
        member system, solved together; returns each request's result in the
        sweep's layout."""
        if self._group_fn is not None:
            return self._group_fn(slot, reqs)
        keys = {r.id: cluster_key(r.cluster) for r in reqs}
        cached = {(r.id, float(l)): self.cache.get((keys[r.id], float(l))) for r in reqs for l in r.wavelengths}
        todo = [(r, float(l)) for r in reqs for l in r.wavelengths if cached[(r.id, float(l))] is None]
        fresh = {}
        if todo:
            pool = self._pools.setdefault(slot, [])
            while len(pool) < len(todo):
                pool.append(_solver.Context(self.device))
            kind = "lu64" if self.kind == "lu64" else "lu32ir"
            leader = self._contexts[slot]
            try:
                mem = []
                for k, (r, lam) in enumerate(todo):
                    sv = _solver.Solver(r.cluster.with_wavelength(lam), kind=kind, context=pool[k])
                    sv.update()
                    mem.append(sv)
                outs = _solver.solve_group(mem, leader=leader)
                for (r, lam), sv, o in zip(todo, mem, outs):
                    cs = sv.cross_sections()
                    st = sv.stats()
                    v = {"x_sca": o["x_sca"],
                         "xs": np.array([cs["cext"], cs["csca"], cs["cabs"], cs["csca_sh"], cs["cabs_sh"]]),
                         "iterations": np.array([st["iterations"], st["iterations_sh"]], dtype=np.int32)}
                    if r.cluster.sh:
                        v["x_sca_sh"] = o["x_sca_sh"]
                    fresh[(r.id, lam)] = v
                    self.cache.put((keys[r.id], lam), v)
            finally:
                for c in pool[:len(todo)]:
                    c.trim()
                leader.trim()
            self.groups_run += 1
        out = {}
        for r in reqs:
            per = [fresh.get((r.id, float(l))) or cached[(r.id, float(l))] for l in r.wavelengths]
            out[r.id] = {k: np.stack([p[k] for p in per]) for k in per[0]}
        return out

    def plan(self, req):
        """Sweep batch size and footprint of a request: the largest batch
        (<= batch_cap, <= its wavelengths) whose footprint fits the budget."""
        b = max(1, min(self.batch_cap, req.wavelengths.size))
        while b > 1 and footprint(req.cluster, b, self.kind) > self.budget:
            b //= 2
        return b, footprint(req.cluster, b, self.kind)

    def _solve(self, slot, req, batch):
        """The request's wavelengths not in the cache, solved; then every
        wavelength's result from the cache.  Returns the sweep's layout:
        x_sca, x_sca_sh (n_lambda, N), xs (n_lambda, 5), iterations."""
        if self._solve_fn is not None:  # a stand-in for the whole request (CPU tests of the admission rules)
            return self._solve_fn(slot, req, batch)
        key = cluster_key(req.cluster)
        lams = req.wavelengths
        cached = [self.cache.get((key, float(l))) for l in lams]
        todo = np.array([l for l, v in zip(lams, cached) if v is None])
        fresh = {}
        if todo.size:
            out = self._device_solve(slot, req, batch, todo)
            for i, l in enumerate(todo):
                fresh[float(l)] = {k: v[i] for k, v in out.items()}
                self.cache.put((key, float(l)), fresh[float(l)])
        per = [fresh[float(l)] if v is None else v for l, v in zip(lams, cached)]
        return {k: np.stack([p[k] for p in per]) for k in per[0]}

    def _device_solve(self, slot, req, batch, lams):
        ctx = self._contexts[slot]
        c = req.cluster
        if self.profile:
            ctx.set_profile(True)  # also resets the context's counters
        try:
            if self.kind == "auto" and c.N >= KRYLOV_N:  # one wavelength at a time, automatic choice
                s = _solver.Solver(c.with_wavelength(lams[0]), kind="auto", context=ctx)
                outs = {"x_sca": [], "x_sca_sh": [], "xs": [], "iterations": []}
                for lam in lams:
                    s.set_cluster(c.with_wavelength(lam))
                    s.update()
                    o = s.solve()
                    cs = s.cross_sections()
                    st = s.stats()
                    outs["x_sca"].append(o["x_sca"])
                    if c.sh:
                        outs["x_sca_sh"].append(o["x_sca_sh"])
                    outs["xs"].append([cs["cext"], cs["csca"], cs["cabs"], cs["csca_sh"], cs["cabs_sh"]])
                    outs["iterations"].append([st["iterations"], st["iterations_sh"]])
                return {k: np.array(v) for k, v in outs.items() if len(v)}
            s = _solver.Solver(c, kind=self.kind, context=ctx)
            return s.sweep(lams, batch=batch, outputs=("x_sca", "x_sca_sh"))
        finally:
            if self.profile:
                self.prof_log.append((req.id, ctx.stats()))
            ctx.trim()

    def _group_worker(self, slot, reqs, ress, nbytes):
        try:
            outs = self._solve_group(slot, reqs)
            for r, res in zip(reqs, ress):
                res.out = outs[r.id]
        except Exception as e:  # reported per request; the server keeps going
            for res in ress:
                res.error = repr(e)
        t = time.perf_counter() - self._t0
        for res in ress:
            res.finished = t
            if self._on_done is not None:
                self._on_done(res)
        with self._lock:
            self._held -= nbytes
            self._free_slots.append(slot)
            self._lock.notify_all()

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
                if self.groupable(req):
                    # the head and the small requests that have arrived after
                    # it, oldest first, up to `group` systems and the budget
                    grp, nsys = [req], req.wavelengths.size
                    for cand in list(pending)[1:]:
                        if realtime and cand.arrival > now:
                            break
                        if (self.groupable(cand) and nsys + cand.wavelengths.size <= self.group
                                and max(cand.cluster.N, req.cluster.N) <= GROUP_RATIO * min(cand.cluster.N,
                                                                                           req.cluster.N)
                                and self.group_footprint(grp + [cand]) <= self.budget):
                            grp.append(cand)
                            nsys += cand.wavelengths.size
                    if len(grp) > 1:
                        gbytes = self.group_footprint(grp)
                        if gbytes <= self.budget:
                            if not self._free_slots or self._held + gbytes > self.budget:
                                self._lock.wait(timeout=0.05)
                                continue
                            for r in grp:
                                pending.remove(r)
                            slot = self._free_slots.pop()
                            self._held += gbytes
                            ress = [results[r.id] for r in grp]
                            for res in ress:
                                res.admitted, res.batch, res.bytes = now, nsys, 0
                            ress[0].bytes = gbytes
                            th = threading.Thread(target=self._group_worker, args=(slot, grp, ress, gbytes),
                                                  daemon=True)
                            threads.append(th)
                            th.start()
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
