"""The sweep autotuner: a tuned shape is cached and read back, tuning the
same shape again picks the same products and triangular solves (or the
candidates it hesitates between lie within the 5% margin), and a sweep gives
the same answer with the tuner on and off."""
import numpy as np
import pytest

from optimet_amd.problem import drude_cluster

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def sweep(c, lams, batch, autotune):
    from optimet_amd import solver
    s = solver.Solver(c, kind="lu32ir")
    s.ctx.set_autotune(autotune)
    out = s.sweep(lams, batch=batch, outputs=("x_sca", "x_sca_sh"))
    Np = s.stats()["Np"]
    s.ctx.close()
    return out, Np


def test_tuning_is_cached_and_reproducible():
    from optimet_amd import solver
    c = drude_cluster(12, 8, 700e-9, 40, 80, 900, 10, seed=21)  # N = 1920
    lams = np.linspace(480e-9, 860e-9, 8)
    solver.tune_reset()
    print(f"tuning: N {c.N}, batch 8", flush=True)
    on, Np = sweep(c, lams, 8, True)
    first = solver.tune_query(Np, 8)
    assert first is not None and first["trsv"] in (0, 1, 2)
    print("choice", first, flush=True)
    again, _ = sweep(c, lams, 8, True)  # cached: no second tuning
    assert solver.tune_query(Np, 8) == first
    solver.tune_reset()
    assert solver.tune_query(Np, 8) is None
    sweep(c, lams, 8, True)
    second = solver.tune_query(Np, 8)
    print("retuned", second, flush=True)
    for key, idx in (("split", (0, 1)), ("trsv", (2, 3, 4))):
        if second[key] != first[key]:  # a hesitation: the two picks were within the margin of each other
            t = second["times_ms"]
            pick = {"split": lambda v: 0 if v else 1, "trsv": lambda v: 2 + v}[key]
            a, b = t[pick(first[key])], t[pick(second[key])]
            assert abs(a - b) <= 0.12 * max(a, b), (key, first, second)
    off, _ = sweep(c, lams, 8, False)
    for k in ("x_sca", "x_sca_sh"):
        assert rel(on[k], off[k]) < 1e-10, k
        assert rel(again[k], on[k]) < 1e-12, k
