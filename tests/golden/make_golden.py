"""Writes tests/golden/tiny_dsgdpp.json: a 6 x 8 problem, rank 3, d = 2
(2 x 4 blocks), 3 WOR/COWOR epochs with Nzsl_L2(0.5), produced by the CPU
oracle (oracle/dsgdpp_oracle.cpp).  It records schedules, task seeds, every
visit order, per-epoch losses and the factors after epoch 1 as hex floats, so
later changes to the oracle or the GPU path are checked against fixed values.

Run:  python tests/golden/make_golden.py
"""
import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
from oracle import pyoracle as po  # noqa: E402


def problem():
    m, n, r = 6, 8, 3
    cells = [(0, 0), (0, 3), (0, 6), (1, 1), (1, 2), (1, 7), (2, 0), (2, 4), (2, 5), (3, 1),
             (3, 3), (3, 6), (4, 2), (4, 4), (4, 7), (5, 0), (5, 5), (5, 6), (1, 4), (3, 0)]
    i = np.array([c[0] for c in cells], dtype=np.int32)
    j = np.array([c[1] for c in cells], dtype=np.int32)
    x = np.array([((a * 7 + b * 3) % 5) - 2.0 + 0.25 * (a - b) for a, b in cells])
    W = np.array([[((a * r + z) % 7) / 10.0 - 0.3 for z in range(r)] for a in range(m)])
    H = np.array([[((b * 5 + z * 3) % 9) / 12.0 - 0.35 for b in range(n)] for z in range(r)])
    return m, n, r, i, j, x, W, H


def main():
    m, n, r, i, j, x, W, H = problem()
    o = po.Oracle(update="Nzsl_L2(0.5)", loss="Nzsl_L2(0.5)", rank=r, tasks_per_rank=2, seed=2024)
    o.load_coo(0, m, n, i, j, x)
    o.set_factors(W, H)
    o.record(True)
    out = {"m": m, "n": n, "rank": r, "d": 2, "seed": 2024, "update": "Nzsl_L2(0.5)",
           "i": i.tolist(), "j": j.tolist(), "x": [v.hex() for v in x],
           "W0": [v.hex() for v in W.ravel()], "H0": [v.hex() for v in H.ravel()],
           "eps": [0.05, 0.05, 0.025], "epochs": []}
    out["loss0"] = o.loss(0).hex()
    for e, eps in enumerate(out["eps"]):
        o.epoch(eps)
        ep = {"schedule": o.schedule().tolist(), "seeds": [str(int(s)) for s in o.task_seeds()],
              "perms": [[o.permutation(t, s).tolist() for s in range(4)] for t in range(2)],
              "loss": o.loss(0).hex()}
        if e == 0:
            Wo, Ho = o.get_factors()
            ep["W"] = [v.hex() for v in Wo.ravel()]
            ep["H"] = [v.hex() for v in Ho.ravel()]
        out["epochs"].append(ep)
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tiny_dsgdpp.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=1)
    print("wrote", path)


if __name__ == "__main__":
    main()
