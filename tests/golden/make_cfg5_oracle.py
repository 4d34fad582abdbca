"""Writes cfg5_oracle.json: FP64 oracle frequencies for the large combined case.

Square lattice (a = 1), eps_bg = 12, two eps = 1 rods: r = 0.2 a at (0, 0) and
r = 0.12 a at (0.5, 0.5), analytic smoothing.
  * 256 x 256 grid, 32 bands, TM and TE, one-k-point solves (the case is
    batched over k-points, each solve from the deterministic start) at
    k = (0.2, 0), (0.5, 0.25) and M = (0.5, 0.5);
  * 512 x 512 grid, TM, 16 bands, at M.
LOBPCG tol 1e-7, every band converged (exit rule 1), max_iter 400.
The oracle (oracle/) is the CPU restatement of the reference path; these runs
take tens of minutes of CPU, so the GPU parity test reads them from here.
Run from the repository root:  python tests/golden/make_cfg5_oracle.py
"""
import concurrent.futures as cf
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
from oracle import oracle as O  # noqa: E402

ATOMS = [(0.0, 0.0, 0.2, 1.0), (0.5, 0.5, 0.12, 1.0)]
CASES = [(256, 32, pol, k) for pol in ("TM", "TE") for k in ((0.2, 0.0), (0.5, 0.25), (0.5, 0.5))]
CASES.append((512, 16, "TM", (0.5, 0.5)))


def run(case):
    n, nb, pol, k = case
    t = time.time()
    job = O.Job(lattice="square", eps_bg=12.0, atoms=ATOMS, nx=n, ny=n, pol=pol, n_bands=nb, tol=1e-7,
                max_iter=400, exit_rule=1, kpath=np.array([k], dtype=np.float64))
    om, it, cv, tot = O.bandstructure(job)
    print(f"{n} {nb} {pol} {k}: {time.time() - t:.0f} s, {int(it[0])} iterations", flush=True)
    return {"grid": n, "n_bands": nb, "pol": pol, "k": list(k), "omegas": om[0].tolist(),
            "iterations": int(it[0]), "converged": int(cv[0])}


if __name__ == "__main__":
    with cf.ThreadPoolExecutor(max_workers=min(len(CASES), os.cpu_count() or 1)) as ex:
        res = list(ex.map(run, CASES))
    out = {"lattice": "square", "eps_bg": 12.0, "atoms": ATOMS, "tol": 1e-7, "exit_rule": 1, "max_iter": 400,
           "smoothing": True, "cases": res}
    with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "cfg5_oracle.json"), "w") as f:
        json.dump(out, f, indent=1)
