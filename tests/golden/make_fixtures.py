"""Write the regression fixtures of tests/golden/ (run from the repo root).

* grid30_oracle.npz: a seeded 30-bus synthetic grid, 4 load scenarios and the
  FP64 oracle NR solution of each at tolerance 1e-12 (inputs + outputs).

The reference's own Jacobian/RHS fixtures case118/ and case1354pegase/ are
copied (gzipped) from custom_linear_solver/tests/datasets/power/<case>/.
"""

from __future__ import annotations

import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT / "gpu-powerflow_amd"))
sys.path.insert(0, str(ROOT))

from cupf import synthetic as syn  # noqa: E402
from oracle import pyoracle as po  # noqa: E402


def main() -> None:
    g = syn.make_grid(30, 45, 5, seed=7)
    f = syn.load_factors(4, g.n_bus, 0.7, 1.3, seed=8)
    S = g.sbus_scaled(f)
    V0 = g.flat_start(4)
    V, it, nm, st = po.nr_solve_batch(g.n_bus, g.indptr, g.indices, g.data, S, V0, g.pv, g.pq,
                                      tol=1e-12, max_iter=30)
    assert (st == 0).all() and (nm <= 1e-12).all()
    np.savez_compressed(Path(__file__).with_name("grid30_oracle.npz"), indptr=g.indptr,
                        indices=g.indices, data=g.data, pv=g.pv, pq=g.pq, sbus=S, v0=V0, V=V,
                        iterations=it, norm=nm)
    print("wrote grid30_oracle.npz", it, nm)


if __name__ == "__main__":
    main()
