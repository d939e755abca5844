#!/usr/bin/env python3
"""Cluster solves/sec on one MI355X: mixed precision vs FP64 LU.

Workload: the 64-sphere Drude cluster (radii U[40,80] nm in a 1 µm cube,
10 nm minimum gap, seed 2000), nmax = 8, λ = 800 nm, fundamental and second
harmonic: N = 10240 complex unknowns per frequency.  One step is one full
cluster solve as OPTIMET's solver does it per wavelength
(PreconditionedMatrixSolver.h:45-100): plane-wave coefficients, T-matrices,
Q, S and V assembly, the ω solve, the second-harmonic source, the 2ω solve and
the cross sections.  Inputs are synthetic (seeded geometry, Drude
permittivity); the sphere data enters device memory inside the step, as it
does per wavelength in the reference's scan loop.

The headline value is the mixed-precision path (complex-FP32 LU on MFMA +
FP64 GMRES-IR); the FP64 LU path runs beside it, and the JSON line reports
both, their p50/p99 latency and the coefficient error of the mixed path
against the FP64 one.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FP32_MFMA_TFLOPS = 157.3   # MI355X dense FP32 matrix peak
BF16_MFMA_TFLOPS = 2500.0  # MI355X dense bf16 matrix peak
# the split-bf16 update issues six bf16 products per real FP32-accurate product
SPLIT_TFLOPS = BF16_MFMA_TFLOPS / 6
FP64_MFMA_TFLOPS = 78.6    # MI355X dense FP64 matrix peak
HBM_GBS = 8000.0
GEMM_KERNEL = {"lu32ir": "gemm_split<128,128,128,4,2,16,8,true,true> (bf16x3 split, complex FP32, double-buffered split planes, nontemporal C; next-pair 128 columns on gemm_tile<float2,32,32>)",
               "lu64": "gemm_tile<double2,128,64,128,4,2,16> (next-pair 128 columns on gemm_tile<double2,32,32>)"}
TRAFFIC_FILE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles",
                            "traffic_gemm_split_config2.txt")


def traffic_per_launch():
    """HBM bytes per launch of the FP32 bulk update from the committed
    rocprofv3 --pmc summary (FETCH_SIZE + WRITE_SIZE, copy-calibrated)."""
    try:
        vals = {}
        with open(TRAFFIC_FILE) as f:
            for line in f:
                if "=" in line and not line.startswith("#"):
                    k, v = line.split("=", 1)
                    vals[k.strip()] = float(v.split()[0])
        return {"bytes_per_split_launch": vals["fetch_bytes_per_launch"] + vals["write_bytes_per_launch"],
                "algorithmic_bytes_per_split_launch": vals["algorithmic_bytes_per_launch"]}
    except (OSError, KeyError, ValueError):
        return None


def roofline(stats, kind, peak):
    """Bulk trailing update of the LU: algorithmic flop (8 m n k per launch)
    over the HIP-event time of its launches on the stream they run on."""
    t = stats["t_lu_gemm_ms"] * 1e-3
    n = stats["lu_gemm_launches"]
    if t <= 0 or n <= 0:
        return None
    achieved = stats["lu_gemm_flop"] / t / 1e12
    return {"bound": "mfma", "kernel": GEMM_KERNEL[kind], "achieved": achieved, "peak": peak, "unit": "TFLOP/s",
            "frac": achieved / peak, "launches": n, "avg_launch_us": 1e6 * t / n,
            "flop_per_launch": stats["lu_gemm_flop"] / n}


def stage_rooflines(c, st, mixed_copy):
    """The other stages of a step against their bounds, from the profiled
    step's HIP-event times.  Assembly (objects, incident field, Q and the
    pair kernel for S and V): algorithmic bytes = the matrices it stores,
    16 N^2 per FP64 matrix plus 8 N^2 per complex-FP32 copy on the mixed path
    (the pair kernel is bound by instruction issue, not by these stores: see
    the counter profile).  Refinement matvecs: 16 Np^2 bytes per product."""
    mats = 2 if c.sh else 1
    store = mats * (16.0 + (8.0 if mixed_copy else 0.0)) * c.N * c.N
    t = st["t_assemble_ms"] * 1e-3
    out = {}
    if t > 0:
        gbs = store / t / 1e9
        out["assembly"] = {"bound": "hbm", "kernel": "pair_kernel (S and V) with objects, incident and origin kernels",
                           "achieved": gbs, "peak": HBM_GBS, "unit": "GB/s", "frac": gbs / HBM_GBS,
                           "bytes": store, "ms": st["t_assemble_ms"]}
    mv = matvec_roofline(st, "gemv_part_kernel (FP64 matvec of the refinement) + gemv_reduce_kernel")
    if mv is not None:
        out["refinement_matvec"] = mv
    return out


def pct(v, q):
    return float(np.percentile(np.asarray(v), q)) if v else float("nan")


def run_path(cluster, kind, steps, warmup, profile, pivoting="block", graphs=False):
    """Times `steps` unprofiled steps (the value and latencies), then, with
    `profile`, runs one more step with per-launch HIP events on (which
    serialises the omega and 2 omega factorisations) for the stage and GEMM
    figures."""
    from optimet_amd import solver

    ctx = solver.Context()
    s = solver.Solver(cluster, kind=kind, context=ctx, pivoting=pivoting)
    ctx.set_graphs(graphs)
    for _ in range(warmup):
        s.update()
        s.solve()
        s.cross_sections()
    lat = []
    out = cs = None
    t_start = time.perf_counter()
    for _ in range(steps):
        t0 = time.perf_counter()
        s.update()
        out = s.solve()
        cs = s.cross_sections()
        lat.append(time.perf_counter() - t0)
    total = time.perf_counter() - t_start
    st = s.stats()
    if profile:
        ctx.set_profile(True)
        s.update()
        s.solve()
        s.cross_sections()
        st = s.stats()
        ctx.set_profile(False)
    ctx.close()
    return {"lat": lat, "total": total, "out": out, "cs": cs, "stats": st}


def run_sweep(cluster, lams, kind, steps, warmup, batch, graphs=False):
    """Batched wavelength sweeps (omx_solve_sweep): every step solves all
    wavelengths at omega and 2 omega."""
    from optimet_amd import solver

    ctx = solver.Context()
    s = solver.Solver(cluster, kind=kind, context=ctx)
    ctx.set_graphs(graphs)
    for _ in range(warmup):
        s.sweep(lams, batch=batch, outputs=())
    lat = []
    out = None
    t_start = time.perf_counter()
    for _ in range(steps):
        t0 = time.perf_counter()
        out = s.sweep(lams, batch=batch, outputs=("x_sca", "x_sca_sh"))
        lat.append(time.perf_counter() - t0)
    total = time.perf_counter() - t_start
    st = s.stats()
    # one more sweep with HIP events around the batched FP64 matvecs of the
    # refinement (the roofline of the line)
    ctx.set_profile(True)
    s.sweep(lams, batch=batch, outputs=())
    prof = s.stats()
    ctx.set_profile(False)
    ctx.close()
    return {"lat": lat, "total": total, "out": out, "stats": st, "prof": prof}


def matvec_roofline(st, kernel="gemv_part_kernel (batched FP64 matvec of the refinement) + gemv_reduce_kernel"):
    """The Krylov or refinement matvecs: their algorithmic bytes (16 Np^2 per
    FP64 product, 8 Np^2 per FP32-matrix product, per system) over the
    HIP-event time of their launches on the stream they run on."""
    if st["t_matvec_ms"] <= 0 or st["matvec_calls"] <= 0:
        return None
    gbs = st["matvec_bytes"] / (st["t_matvec_ms"] * 1e-3) / 1e9
    return {"bound": "hbm", "kernel": kernel,
            "achieved": gbs, "peak": HBM_GBS, "unit": "GB/s", "frac": gbs / HBM_GBS, "traffic": None,
            "launches": st["matvec_calls"], "avg_launch_us": 1e3 * st["t_matvec_ms"] / st["matvec_calls"],
            "bytes_per_launch": st["matvec_bytes"] / st["matvec_calls"]}


def tuned(Np, nb, fp64):
    """The sweep autotuner's choice for a shape (see omx_set_autotune)."""
    from optimet_amd import solver
    return solver.tune_query(Np, nb, fp64)


def sweep_main(args):
    """Config 3: one 16-sphere Drude cluster, nmax 10 (N = 3840), 256
    wavelengths in [400, 900] nm, omega and 2 omega, batched."""
    from optimet_amd.problem import config3_cluster, config3_wavelengths

    c = config3_cluster()
    lams = config3_wavelengths(256)
    graphs = args.graphs == "on"
    # the value: the automatic choice (S by lockstep FP64 GMRES beside V's
    # batched mixed LU); beside it the mixed LU for both and the FP64 LU
    mixed = run_sweep(c, lams, args.sweep_kind, args.steps, args.warmup, args.batch, graphs=graphs)
    mixed_lu = run_sweep(c, lams, "lu32ir", args.steps, args.warmup, args.batch, graphs=graphs) \
        if args.sweep_kind != "lu32ir" else None
    fp64 = None if args.no_fp64 else run_sweep(c, lams, "lu64", args.steps, args.warmup, args.batch, graphs=graphs)
    nl = lams.size
    value = args.steps * nl / mixed["total"]
    ms = mixed["stats"]
    line = {
        "metric": "cluster solves/sec + p99 latency, mixed-precision vs FP64 LU at rel err <=1e-9",
        "value": value, "unit": "solves/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": 1e3 * mixed["total"] / args.steps, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None,
        "dtype": ("c128 GMRES (omega, lockstep) beside c64 LU (FP32-accurate bf16x3 MFMA products) + c128 GMRES-IR "
                  "(2 omega)") if args.sweep_kind == "auto" else "c64 LU (FP32-accurate bf16x3 MFMA products) + c128 GMRES-IR",
        "data": "synthetic (seeded Drude cluster)",
        "config": {"workload": c.meta["name"], "P": c.nobj, "nmax": c.nmax, "N": c.N, "wavelengths": nl,
                   "frequencies": "omega+2omega", "solver": args.sweep_kind, "batch": ms["sweep_batch"],
                   "graphs": args.graphs},
        "step": "one sweep of all wavelengths (a solve = one wavelength at omega and 2 omega)",
        "p50_ms_per_sweep": 1e3 * pct(mixed["lat"], 50), "p99_ms_per_sweep": 1e3 * pct(mixed["lat"], 99),
        "max_iterations": [ms["iterations"], ms["iterations_sh"]], "max_residual": [ms["residual"], ms["residual_sh"]],
        "roofline": matvec_roofline(mixed["prof"]),
        "autotune": {"lu32ir": tuned(ms["Np"], ms["sweep_batch"], False),
                     "lu64": tuned(ms["Np"], ms["sweep_batch"], True) if fp64 is not None else None},
        "systems_redone_alone": ms["pivot_fallbacks"],
    }
    if mixed_lu is not None:
        ls = mixed_lu["stats"]
        line["mixed_lu"] = {"solves_per_s": args.steps * nl / mixed_lu["total"],
                            "p50_ms_per_sweep": 1e3 * pct(mixed_lu["lat"], 50),
                            "max_iterations": [ls["iterations"], ls["iterations_sh"]],
                            "roofline": matvec_roofline(mixed_lu["prof"])}
    if fp64 is not None:
        fs = fp64["stats"]
        rel = lambda a, b: float(np.max(np.linalg.norm(a - b, axis=1) / np.linalg.norm(b, axis=1)))
        line["fp64_lu"] = {"solves_per_s": args.steps * nl / fp64["total"],
                           "p50_ms_per_sweep": 1e3 * pct(fp64["lat"], 50),
                           "max_iterations": [fs["iterations"], fs["iterations_sh"]]}
        line["speedup_vs_fp64"] = value / line["fp64_lu"]["solves_per_s"]
        if mixed_lu is not None:
            line["mixed_lu"]["speedup_vs_fp64"] = line["mixed_lu"]["solves_per_s"] / line["fp64_lu"]["solves_per_s"]
        mo, fo = mixed["out"], fp64["out"]
        line["max_rel_err_vs_fp64"] = {"x_sca": rel(mo["x_sca"], fo["x_sca"]),
                                       "x_sca_sh": rel(mo["x_sca_sh"], fo["x_sca_sh"]),
                                       "cext": float(np.max(np.abs(mo["xs"][:, 0] - fo["xs"][:, 0])
                                                            / np.abs(fo["xs"][:, 0])))}
    print(json.dumps(line), flush=True)


def config4_main(args):
    """Config 4: 512 Drude spheres, nmax 6 (N = 49152, a 38.7 GB FP64 matrix),
    fundamental frequency.  Legs: the default solver choice ("auto": the mixed
    Krylov within the mixed LU's cost, falling back to the mixed LU), FP64
    GMRES, the mixed Krylov, the mixed LU + GMRES-IR, the FP64 LU (the
    comparison path of the metric) and the ACA-compressed operator with the
    reference's GMRES settings (tol 1e-6).  The Krylov paths run to the same
    final tolerance, 1e-10."""
    from optimet_amd import solver
    from optimet_amd.problem import config4

    c = config4()
    res = {}
    legs = args.legs.split(",")
    for kind in legs:
        ctx = solver.Context()
        if kind == "aca":  # <ACA compression="yes"/> with the reference's GMRES settings
            s = solver.Solver(c, kind=kind, context=ctx)
        else:
            s = solver.Solver(c, kind=kind, context=ctx, tol=1e-10 if kind.startswith("gmres") else 1e-16, maxit=800)
        ctx.set_profile(True)
        t0 = time.perf_counter()
        s.update()
        t1 = time.perf_counter()
        aca = s.aca_info()[1] if kind == "aca" else None
        converged = True
        try:
            out = s.solve()
            cs = s.cross_sections()
        except solver.OmxError as e:  # the unpreconditioned GMRES may stop above its tolerance
            if e.code != -6:
                raise
            converged, out, cs = False, None, {"cext": None}
        t2 = time.perf_counter()
        st = s.stats()
        ctx.close()
        res[kind] = {"solve_s": t2 - t1, "update_s": t1 - t0, "converged": converged, "iterations": st["iterations"],
                     "residual": st["residual"], "matvec_calls": st["matvec_calls"],
                     "matvec_GBs": (st["matvec_bytes"] / (st["t_matvec_ms"] * 1e-3) / 1e9)
                     if st["t_matvec_ms"] > 0 else None,
                     "fallbacks": st["pivot_fallbacks"],
                     "out": out, "cext": cs["cext"], "stats": st, "aca": aca}
        if kind in ("lu32ir", "lu64"):
            res[kind]["roofline"] = roofline(st, kind, SPLIT_TFLOPS if kind == "lu32ir" else FP64_MFMA_TFLOPS)
    keys = ("update_s", "solve_s", "converged", "iterations", "residual", "matvec_calls", "matvec_GBs", "cext")
    if not all(k in res for k in ("auto", "lu32ir", "lu64")):  # a partial run (profiling): the legs' figures only
        print(json.dumps({k: {f: v[f] for f in keys + ("aca",)} for k, v in res.items()}), flush=True)
        return
    total = {k: v["update_s"] + v["solve_s"] for k, v in res.items()}
    ref = res["lu64"]["out"]["x_sca"]
    rel_x = {k: float(np.linalg.norm(v["out"]["x_sca"] - ref) / np.linalg.norm(ref))
             for k, v in res.items() if v["out"] is not None}
    au = res["auto"]
    picked = "gmres32ir" if au["fallbacks"] == 0 and c.N >= 16384 else "lu32ir"
    line = {
        "metric": "cluster solves/sec + p99 latency, mixed-precision vs FP64 LU at rel err <=1e-9",
        "value": 1.0 / total["auto"], "unit": "solves/s", "n_gpus": 1, "steps": 1, "warmup": 0,
        "ms_per_step": 1e3 * total["auto"], "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "c128 flexible GMRES over complex-FP32 GMRES cycles (FP32 matrix, FP64 vectors)"
        if picked == "gmres32ir" else "c64 LU (FP32-accurate bf16x3 MFMA products) + c128 GMRES-IR",
        "data": "synthetic (seeded Drude cluster)",
        "config": {"workload": c.meta["name"], "P": c.nobj, "nmax": c.nmax, "N": c.N, "frequencies": "omega",
                   "solver": "auto", "picked": picked,
                   "rule": "N >= 16384: the mixed Krylov within a budget of the mixed LU's cost (0.0143 N FP32 "
                           "products), else or when it does not converge there, the mixed LU + GMRES-IR"},
        "roofline": matvec_roofline(au["stats"], "gemv_part_kernel<float2> (inner products on the FP32 copy) and "
                                    "gemv_part_kernel<cplx> (outer FP64 products), with gemv_reduce_kernel")
        if picked == "gmres32ir" else res["lu32ir"].get("roofline"),
        "speedup_vs_fp64": total["lu64"] / total["auto"],
        "speedup_mixed_lu_vs_fp64_lu": total["lu64"] / total["lu32ir"],
        "rel_err_x_sca_vs_fp64_lu": rel_x,
        "legs": {k: {f: v[f] for f in keys} for k, v in res.items() if k != "aca"},
        "lu_rooflines": {k: res[k].get("roofline") for k in ("lu32ir", "lu64")},
    }
    if "aca" in res:
        a = res["aca"]
        ai = a["aca"]
        line["aca"] = {f: a[f] for f in keys}
        line["aca"].update({"compressed_GB": ai["bytes"] / 1e9, "dense_GB": ai["dense_bytes"] / 1e9,
                            "fraction_of_dense": ai["bytes"] / ai["dense_bytes"],
                            "reference_size_MB": ai["reference_mb"], "compressed_blocks": int(ai["compressed_blocks"]),
                            "dense_blocks": int(ai["dense_blocks"]), "mean_rank": ai["mean_rank"],
                            "max_rank": int(ai["max_rank"]), "device_GB": a["stats"]["device_bytes"] / 1e9,
                            "rel_diff_x_sca_vs_fp64_lu": rel_x.get("aca")})
    print(json.dumps(line), flush=True)


def group_main(args):
    """Continuous batching of small uneven requests (omx_solve_group): 64
    seeded Drude clusters (P in 2..16, nmax in 4..7, one wavelength each in
    400-900 nm, omega and 2 omega), all arriving at once.  Served by the
    request loop with --slots concurrent slots, without grouping (every
    request a one-system sweep in its slot) and with groups of up to --group
    systems; beside them the plain API, one cluster after another
    (omx_update + omx_solve) and in groups of similar padded order.  A solve
    is one cluster at omega and 2 omega."""
    from optimet_amd import solver
    from optimet_amd.problem import drude_cluster
    from optimet_amd.serving import Request, Server, summarize

    rng = np.random.default_rng(5000)
    cl = []
    for i in range(64):
        P, nmax = int(rng.integers(2, 17)), int(rng.integers(4, 8))
        lam = float(rng.uniform(400e-9, 900e-9))
        cl.append(drude_cluster(P, nmax, lam, 40, 80, 600 + 40 * P, 10, seed=5000 + i))
    reqs = [Request(i, 0.0, c, np.array([c.wavelength])) for i, c in enumerate(cl)]
    G = max(2, args.group)

    def served(group):
        srv = Server(slots=args.slots, kind="lu32ir", group=group)
        try:
            for _ in range(max(1, args.warmup)):  # tuning, first-touch allocations, member pools
                srv.cache = type(srv.cache)(srv.cache.max_bytes)
                srv.run(reqs, realtime=False)
            srv.cache = type(srv.cache)(srv.cache.max_bytes)
            res = srv.run(reqs, realtime=False)
            groups = srv.groups_run
        finally:
            srv.close()
        assert all(r.error is None for r in res), [r.error for r in res if r.error][:3]
        summ = summarize(res, {r.id: 1 for r in reqs})
        return summ, {r.id: r.out for r in res}, groups

    order = sorted(range(len(cl)), key=lambda i: cl[i].N)

    def api_one_by_one():
        ctx = solver.Context()
        t0 = time.perf_counter()
        for i in order:
            s = solver.Solver(cl[i], kind="lu32ir", context=ctx)
            s.update()
            s.solve()
            s.cross_sections()
        t = time.perf_counter() - t0
        ctx.close()
        return t

    def api_grouped(pool, leader):
        t0 = time.perf_counter()
        for g0 in range(0, len(order), G):
            mem = [solver.Solver(cl[i], kind="lu32ir", context=pool[k]) for k, i in enumerate(order[g0:g0 + G])]
            for s in mem:
                s.update()
            solver.solve_group(mem, leader=leader)
            for s in mem:
                s.cross_sections()
        return time.perf_counter() - t0

    plain, plain_out, _ = served(0)
    grouped, grouped_out, ngroups = served(G)
    err = max(float(np.linalg.norm(grouped_out[i]["x_sca"] - plain_out[i]["x_sca"])
                    / np.linalg.norm(plain_out[i]["x_sca"])) for i in plain_out)
    cerr = max(float(np.max(np.abs(grouped_out[i]["xs"][:, 0] - plain_out[i]["xs"][:, 0])
                            / np.abs(plain_out[i]["xs"][:, 0]))) for i in plain_out)
    pool = [solver.Context() for _ in range(G)]
    leader = solver.Context()
    api_one_by_one()
    api_grouped(pool, leader)
    t1 = api_one_by_one()
    t2 = api_grouped(pool, leader)
    for c in pool + [leader]:
        c.close()
    line = {"metric": "cluster solves/sec, uneven single-wavelength requests batched across requests",
            "value": grouped["solves_per_s"], "unit": "solves/s", "n_gpus": 1, "higher_is_better": True,
            "vs_baseline": None, "dtype": "c64 LU (FP32-accurate bf16x3 MFMA products) + c128 GMRES-IR",
            "data": "synthetic (seeded Drude clusters)",
            "config": {"workload": "64 requests of one wavelength, P 2-16, nmax 4-7, N %d-%d, omega+2omega, %d slots, "
                       "groups of up to %d systems" % (min(c.N for c in cl), max(c.N for c in cl), args.slots, G)},
            "served_grouped": {k: grouped[k] for k in ("solves_per_s", "p50_s", "p99_s", "span_s")},
            "served_plain": {k: plain[k] for k in ("solves_per_s", "p50_s", "p99_s", "span_s")},
            "groups_run": ngroups, "speedup_vs_plain_serving": grouped["solves_per_s"] / plain["solves_per_s"],
            "api_one_by_one_solves_per_s": len(cl) / t1, "api_grouped_solves_per_s": len(cl) / t2,
            "max_rel_diff_x_sca": err, "max_rel_diff_cext": cerr}
    print(json.dumps(line), flush=True)


def config5_main(args):
    """Config 5: the seeded Poisson request stream (P ~ U{2..128}, nmax ~
    U{4..10}, 1-64 wavelengths, Drude or constant permittivity, omega and
    2 omega) served in real time by the request loop; a bounded sample: the
    first `--requests` requests (optionally, admission closed after
    `--deadline` s).  The mixed path serves with the automatic choice per
    request (batched mixed-LU sweeps; the mixed Krylov from N = 16384 on);
    the comparison path with batched FP64-LU sweeps.  The roofline comes from
    a profiled replay of the first `--roofline-requests` requests."""
    from optimet_amd.problem import serving_requests
    from optimet_amd.serving import Request, Server, summarize

    stream = serving_requests(n=args.requests)
    reqs = [Request(r["id"], r["arrival"], r["cluster"], r["wavelengths"]) for r in stream]
    nl = {r.id: r.wavelengths.size for r in reqs}
    out = {}
    for kind in (("auto",) if args.no_fp64 else ("auto", "lu64")):
        srv = Server(slots=args.slots, kind=kind, group=args.group)

        def progress(r, kind=kind):
            print(f"[{kind}] request {r.id} done at {r.finished:.1f} s (latency {r.latency:.1f} s"
                  f"{', error' if r.error else ''})", file=sys.stderr, flush=True)

        try:
            res = srv.run(reqs, realtime=True, deadline=args.deadline, on_done=progress)
            cache = srv.cache.stats()
        finally:
            srv.close()
        summ = summarize(res, nl)
        summ["error_kinds"] = sorted({r.error.split("(")[0] for r in res if r.error})
        summ["cache"] = cache
        summ["groups_run"] = srv.groups_run
        summ["latency_s"] = {r.id: r.latency for r in res if r.error is None}
        out[kind] = (summ, res)
    # roofline: the dominant kernel over a profiled replay of the first requests
    srv = Server(slots=args.slots, kind="auto", profile=True)
    try:
        srv.run(reqs[:args.roofline_requests], realtime=False)
        log = [st for _, st in srv.prof_log]
    finally:
        srv.close()
    gemm_ms = sum(st["t_lu_gemm_ms"] for st in log)
    gemm_flop = sum(st["lu_gemm_flop"] for st in log)
    mv_ms = sum(st["t_matvec_ms"] for st in log)
    mv_bytes = sum(st["matvec_bytes"] for st in log)
    if gemm_ms >= mv_ms and gemm_ms > 0:
        a = gemm_flop / (gemm_ms * 1e-3) / 1e12
        roof = {"bound": "mfma", "kernel": GEMM_KERNEL["lu32ir"] + ", batched over each request's wavelengths",
                "achieved": a, "peak": SPLIT_TFLOPS, "unit": "TFLOP/s", "frac": a / SPLIT_TFLOPS, "traffic": None,
                "launches": sum(st["lu_gemm_launches"] for st in log)}
    elif mv_ms > 0:
        a = mv_bytes / (mv_ms * 1e-3) / 1e9
        roof = {"bound": "hbm", "kernel": "gemv_part_kernel (refinement and Krylov products)", "achieved": a,
                "peak": HBM_GBS, "unit": "GB/s", "frac": a / HBM_GBS, "traffic": None}
    else:
        roof = None
    if roof is not None:
        roof["sample"] = (f"profiled replay of the first {args.roofline_requests} requests: {mv_ms:.0f} ms of "
                          "HIP-event time in matvecs (the Krylov and refinement products; the batched sweeps' bulk "
                          "LU updates are not timed per launch)" if gemm_ms == 0 else
                          f"profiled replay of the first {args.roofline_requests} requests: LU bulk updates "
                          f"{gemm_ms:.0f} ms, matvecs {mv_ms:.0f} ms of HIP-event time")
    m = out["auto"][0]
    line = {
        "metric": "cluster solves/sec + p99 latency, mixed-precision vs FP64 LU at rel err <=1e-9",
        "value": m["solves_per_s"], "unit": "solves/s", "n_gpus": 1, "steps": 1, "warmup": 0,
        "ms_per_step": 1e3 * m["span_s"], "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "c64 LU (FP32-accurate bf16x3 MFMA products) + c128 GMRES-IR; c128 GMRES over c64 GMRES cycles "
                 "from N = 16384", "data": "synthetic (seeded request stream)",
        "config": {"workload": "config5-serving-stream", "requests": args.requests, "slots": args.slots,
                   "deadline_s": args.deadline, "rate_hz": 50.0, "solver": "auto", "group": args.group},
        "step": "the served sample; a solve = one wavelength of a request at omega and 2 omega",
        "p50_latency_s": m["p50_s"], "p99_latency_s": m["p99_s"], "mixed": m, "roofline": roof,
    }
    if "lu64" in out:
        f = out["lu64"][0]
        line["fp64_lu"] = f
        line["speedup_vs_fp64"] = (m["solves_per_s"] / f["solves_per_s"]) if f["solves_per_s"] else None
    print(json.dumps(line), flush=True)


def host_facts():
    """Host the CPU baseline ran on: core counts, affinity, CPU model, threads, BLAS."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    blas = None
    try:
        cfg = np.show_config(mode="dicts")
        blas = cfg.get("Build Dependencies", {}).get("blas", {}).get("name")
    except Exception:
        pass
    return {"nproc": os.cpu_count(), "sched_getaffinity": len(os.sched_getaffinity(0)), "cpu_model": model,
            "omp_num_threads": os.environ.get("OMP_NUM_THREADS"), "blas": blas}


def cpu_baseline(cluster):
    """The oracle (the C++ FP64 restatement of the reference's algorithm) on
    the GPU box's host, two lines:
    (i) all cores, measured in full: one config-2 solve as the reference does
        it per wavelength -- Q, S and V assembled by the oracle with OpenMP over
        block rows, both dense solves by LAPACK zgesv (numpy), the SH source
        (OpenMP over objects) and the cross sections; the CG tables are
        per-nmax setup, timed apart;
    (ii) one core, reference-faithful (the reference is single-threaded):
        pair blocks, a column-pivoted QR solve and one object's SH source
        timed on a sample and extrapolated by O(P^2), O(N^3), O(P)."""
    import oracle

    facts = host_facts()
    threads = int(facts["omp_num_threads"] or facts["sched_getaffinity"] or 1)
    t0 = time.perf_counter()
    tables = oracle.cg_tables(cluster.nmax)
    t_tables = time.perf_counter() - t0
    stages = {}
    t0 = time.perf_counter()
    Q = oracle.source_vector(cluster)
    S = oracle.scattering_matrix(cluster)
    stages["assembly_S_Q_s"] = time.perf_counter() - t0
    t1 = time.perf_counter()
    x = np.linalg.solve(S, Q)
    del S
    stages["solve_omega_s"] = time.perf_counter() - t1
    t1 = time.perf_counter()
    iaux = np.concatenate([oracle.object_vector(cluster, j, oracle.OBJ_IAUX) for j in range(cluster.nobj)])
    K, _ = oracle.source_sh(cluster, tables, x * iaux)
    stages["sh_source_s"] = time.perf_counter() - t1
    t1 = time.perf_counter()
    V = oracle.scattering_matrix(cluster, sh=True)
    stages["assembly_V_s"] = time.perf_counter() - t1
    t1 = time.perf_counter()
    xsh = np.linalg.solve(V, K)
    del V
    stages["solve_2omega_s"] = time.perf_counter() - t1
    t1 = time.perf_counter()
    oracle.cross_sections(cluster, x, xsh)
    stages["cross_sections_s"] = time.perf_counter() - t1
    total = time.perf_counter() - t0
    # (ii) one core, reference-faithful, sampled and extrapolated
    npairs, nsolve = 1024, 2048
    t_pair, t_solve, t_sh = oracle.time_sample(cluster, npairs, nsolve, 1)
    P, N = cluster.nobj, cluster.N
    est = t_pair * (2 * P * (P - 1) + 2 * P) + 2 * t_solve * (N / nsolve) ** 3 + t_sh * P
    return {"value": 1.0 / total, "unit": "solves/s", "cores": threads, "kind": "port",
            "sample": f"one full solve of the same workload (P={P}, N={N}, omega+2omega) on {threads} threads: oracle "
                      "assembly (OpenMP), LAPACK zgesv, SH source, cross sections; CG tables (per-nmax setup) "
                      f"{t_tables:.1f} s apart",
            "stages_s": stages, "host": facts,
            "single_core_reference_faithful": {
                "value": 1.0 / est, "unit": "solves/s", "cores": 1, "kind": "port", "extrapolated": True,
                "sample": f"{npairs} pair couplings+blocks, one {nsolve}x{nsolve} column-pivoted QR solve and one "
                          f"object's SH source, extrapolated to P={P}, N={N}, omega+2omega (est. {est:.0f} s per solve)"}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-fp64", action="store_true", help="skip the FP64 LU comparison path")
    ap.add_argument("--pivoting", choices=("block", "partial"), default="block",
                    help="config2: pivot search of the mixed path's factorisations")
    ap.add_argument("--legs", default="auto,gmres64,gmres32ir,lu32ir,lu64,aca",
                    help="config4: solver paths to run (a subset prints their figures only)")
    ap.add_argument("--group", type=int, default=16, help="group: clusters per omx_solve_group")
    ap.add_argument("--sweep-kind", choices=("auto", "lu32ir"), default="auto",
                    help="config3: solver kind of the mixed sweep")
    ap.add_argument("--workload", choices=("config2", "config3", "config4", "config5", "group"), default="config2",
                    help="config2: one 64-sphere cluster (headline); config3: 256-wavelength batched sweep; "
                         "config4: 512 spheres, N = 49152, GMRES vs mixed LU; config5: request stream")
    ap.add_argument("--requests", type=int, default=8, help="config5: requests of the stream served")
    ap.add_argument("--slots", type=int, default=6, help="config5: concurrent requests")
    ap.add_argument("--roofline-requests", type=int, default=4,
                    help="config5: requests of the profiled replay the roofline comes from")
    ap.add_argument("--deadline", type=float, default=None, help="config5: admission closes after this many s")
    ap.add_argument("--batch", type=int, default=0, help="systems per batch of the config3 sweep (0: automatic)")
    ap.add_argument("--graphs", choices=("on", "off"), default="off",
                    help="config2/config3: factorisations captured in HIP graphs and replayed (omx_set_graphs)")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only copy a MIB-sized device buffer (calibration of rocprofv3 FETCH_SIZE/WRITE_SIZE)")
    args = ap.parse_args()
    if args.copy_probe:
        import torch

        n = args.copy_probe * (1 << 20) // 4
        a = torch.ones(n, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes_read": 4 * n, "copy_probe_bytes_written": 4 * n}), flush=True)
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported: this build runs on one GPU "
              "(the path would shard by block rows of S and whole systems per GPU)", file=sys.stderr)
        sys.exit(2)

    if args.workload == "config3":
        sweep_main(args)
        return
    if args.workload == "config4":
        config4_main(args)
        return
    if args.workload == "config5":
        config5_main(args)
        return
    if args.workload == "group":
        group_main(args)
        return

    from optimet_amd.problem import config2

    c = config2()
    graphs = args.graphs == "on"
    # the value: the automatic solver choice (S by FP64 GMRES beside V's mixed
    # LU, S's own mixed LU if GMRES falls short); beside it the mixed LU for
    # both systems and the FP64 LU of the metric's comparison
    best = run_path(c, "auto", args.steps, args.warmup, True, pivoting=args.pivoting, graphs=graphs)
    mixed = run_path(c, "lu32ir", args.steps, args.warmup, True, pivoting=args.pivoting, graphs=graphs)
    fp64 = None if args.no_fp64 else run_path(c, "lu64", args.steps, args.warmup, True, graphs=graphs)

    ms = best["stats"]
    value = args.steps / best["total"]
    roof = roofline(ms, "lu32ir", SPLIT_TFLOPS)
    if roof is not None:
        roof["peak_basis"] = "bf16 dense 2500 TFLOP/s / 6 bf16 products per FP32-accurate real product"
        tr = traffic_per_launch()
        roof["traffic"] = tr["bytes_per_split_launch"] if tr else None
        if tr:
            roof["traffic_note"] = ("HBM bytes per gemm_split launch (the wide part of a pair's bulk update; "
                                    f"algorithmic {tr['algorithmic_bytes_per_split_launch']:.3g} B), rocprofv3 "
                                    "FETCH_SIZE + WRITE_SIZE, copy-calibrated")
    line = {
        "metric": "cluster solves/sec + p99 latency, mixed-precision vs FP64 LU at rel err <=1e-9",
        "value": value,
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * best["total"] / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c128 GMRES (omega) beside c64 LU (FP32-accurate bf16x3 MFMA products) + c128 GMRES-IR (2 omega)",
        "data": "synthetic (seeded Drude cluster)",
        "config": {"workload": c.meta["name"], "P": c.nobj, "nmax": c.nmax, "N": c.N,
                   "wavelength_nm": c.wavelength * 1e9, "frequencies": "omega+2omega", "solver": "auto",
                   "graphs": args.graphs},
        "solver_choice": {"omega": "FP64 GMRES" if ms["pivot_fallbacks"] == 0 else "mixed LU (GMRES fell short)",
                          "2omega": "mixed LU + FP64 GMRES-IR"},
        "p50_ms": 1e3 * pct(best["lat"], 50),
        "p99_ms": 1e3 * pct(best["lat"], 99),
        "stages_ms": {k: ms[k] for k in ("t_update_ms", "t_assemble_ms", "t_solve_ms", "t_sh_source_ms")},
        "iterations": [ms["iterations"], ms["iterations_sh"]],
        "residual": [ms["residual"], ms["residual_sh"]],
        "lu_factor_ms_total": ms["t_factor_total_ms"],
        "lu_panel_ms_total": ms["t_lu_panel_ms"],
        "roofline": roof,
        "stage_rooflines": stage_rooflines(c, ms, mixed_copy=True),
        "mixed_lu": {"solves_per_s": args.steps / mixed["total"], "p50_ms": 1e3 * pct(mixed["lat"], 50),
                     "p99_ms": 1e3 * pct(mixed["lat"], 99),
                     "iterations": [mixed["stats"]["iterations"], mixed["stats"]["iterations_sh"]],
                     "roofline": roofline(mixed["stats"], "lu32ir", SPLIT_TFLOPS),
                     "lu_factor_ms_total": mixed["stats"]["t_factor_total_ms"]},
        "autotune": {"lu32ir": tuned(ms["Np"], 1, False),
                     "lu64": tuned(ms["Np"], 1, True) if fp64 is not None else None},
    }
    if fp64 is not None:
        fs = fp64["stats"]
        rel = lambda a, b: float(np.linalg.norm(a - b) / np.linalg.norm(b))
        line["fp64_lu"] = {"solves_per_s": args.steps / fp64["total"], "p50_ms": 1e3 * pct(fp64["lat"], 50),
                           "p99_ms": 1e3 * pct(fp64["lat"], 99),
                           "roofline": roofline(fs, "lu64", FP64_MFMA_TFLOPS),
                           "iterations": [fs["iterations"], fs["iterations_sh"]],
                           "lu_factor_ms_total": fs["t_factor_total_ms"], "lu_panel_ms_total": fs["t_lu_panel_ms"]}
        line["speedup_vs_fp64"] = value / line["fp64_lu"]["solves_per_s"]
        line["mixed_lu"]["speedup_vs_fp64"] = line["mixed_lu"]["solves_per_s"] / line["fp64_lu"]["solves_per_s"]
        fo = fp64["out"]

        def errs(run):
            mo = run["out"]
            return {"x_sca": rel(mo["x_sca"], fo["x_sca"]), "x_int": rel(mo["x_int"], fo["x_int"]),
                    "x_sca_sh": rel(mo["x_sca_sh"], fo["x_sca_sh"]),
                    "cext": abs(run["cs"]["cext"] - fp64["cs"]["cext"]) / abs(fp64["cs"]["cext"]),
                    "csca_sh": abs(run["cs"]["csca_sh"] - fp64["cs"]["csca_sh"]) / abs(fp64["cs"]["csca_sh"])}
        line["rel_err_vs_fp64"] = errs(best)
        line["mixed_lu"]["rel_err_vs_fp64"] = errs(mixed)
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(c)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
