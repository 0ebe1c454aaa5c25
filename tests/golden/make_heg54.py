"""Writes heg54_oracle.json: the CPU oracle's full solve of the electron gas
with 54 electrons, r_s = 1.0, n^2 <= 36 (1850 s.p. states; reference CLI
--electronGas -r 1.0 -s 32 -p 4 -t 1e-10), the headline workload.

The GPU tests hold the FP64 GPU solve to it iteration by iteration (same
count, same stop reason) and the mixed-precision solve to its converged
energy per particle; bench.py reports the mixed solve's error against it.
The oracle follows the reference loop (src/main.cpp:491-682): T2 = 0,
mixing 0.5, E_prev = 100, continue while |dE| > tol.

    python tests/golden/make_heg54.py          # about 20 s on 8 cores
"""
from __future__ import annotations

import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle.pyoracle import Oracle, set_threads  # noqa: E402

ARGS = (1.0, 32, 4)
TOL = 1e-10
SEED = 20261204
N_SAMPLES = 4096


def main():
    set_threads(os.cpu_count() or 1)
    t0 = time.time()
    o = Oracle.electron_gas(*ARGS)
    mbpt2 = o.mbpt2()
    e, stop, _ = o.solve(TOL)
    t2 = o.t2()
    idx = np.sort(np.random.default_rng(SEED).choice(o.nt2, N_SAMPLES, replace=False))
    de = np.diff(np.concatenate([[100.0], e]))
    out = {
        "workload": "electron gas N=54, r_s=1.0, n^2<=36",
        "reference_cli": "--electronGas -r 1.0 -s 32 -p 4 -t 1e-10",
        "args": list(ARGS),
        "nsp": o.nsp, "nh": o.nh, "nt2": o.nt2,
        "eref": o.eref(), "mbpt2": mbpt2,
        "tol": TOL, "stop": stop, "energies": [float(x) for x in e],
        # |dE| of the last two iterations: the stop rule's margin on either side of tol
        "last_abs_dE": [float(abs(de[-2])), float(abs(de[-1]))],
        "t2_sum_sq": float(np.dot(t2, t2)),
        "t2_max_abs": float(np.max(np.abs(t2))),
        "t2_sample_seed": SEED,
        "t2_sample_index": idx.tolist(),
        "t2_sample_value": t2[idx].tolist(),
        "oracle_seconds": time.time() - t0,
    }
    with open(os.path.join(HERE, "heg54_oracle.json"), "w") as f:
        json.dump(out, f, indent=1)
    print(f"{len(e)} iterations, stop {stop}, E {e[-1]!r}, {out['oracle_seconds']:.0f} s")


if __name__ == "__main__":
    main()
