"""Writes the golden fixtures of tests/golden/ with the CPU oracle.

The reference holds no expected outputs for this path (its regression data
lives in an absent submodule), so these fixtures pin the oracle's own output
on two seeded cases: the GPU parity tests and the oracle pins compare
against them.  Run from the repository root:  python -m tests.golden.make_golden
"""
import os

import numpy as np

from optimet_amd.problem import config1, drude_cluster

HERE = os.path.dirname(os.path.abspath(__file__))

CASES = {
    # 3 dielectric spheres on the x axis, nmax 6, fundamental only (288 unknowns)
    "three_spheres_nmax6": config1,
    # 4 seeded Drude spheres, nmax 4, fundamental and second harmonic (192 unknowns)
    "drude4_nmax4_sh": lambda: drude_cluster(4, 4, 800e-9, 40, 80, 400, 10, seed=7),
}


def main():
    import oracle

    for name, build in CASES.items():
        c = build()
        out = oracle.solve_all(c, method=0)
        keep = {k: np.asarray(v) for k, v in out.items() if k in ("x_sca", "x_sca_sh", "cext", "csca", "cabs",
                                                                  "csca_sh")}
        np.savez(os.path.join(HERE, name + ".npz"), **keep)
        print(name, {k: v.shape for k, v in keep.items()})


if __name__ == "__main__":
    main()
