"""Writes heg66_oracle.json: the CPU oracle's solve of the electron gas with
66 electrons, r_s = 0.5, n^2 <= 52 (3150 s.p. states; reference CLI
--electronGas -r 0.5 -s 46 -p 5).  At this size the GPU tests cannot run the
oracle themselves (about 8 s per FP64 iteration on 8 cores), so they compare
against this file.

The oracle runs in its matrix-free mode (V^pppp, 54 GB in the reference's
layout, rebuilt per channel inside the ladder, calc.cpp:233-262); its results
do not depend on the mode (tests/test_oracle_pins.py checks that at N=14).

    python tests/golden/make_heg66.py          # a few minutes on 8 cores
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

ARGS = (0.5, 46, 5)
TOL = 1e-10
SEED = 20261203
N_SAMPLES = 4096


def main():
    set_threads(os.cpu_count() or 1)
    t0 = time.time()
    o = Oracle.electron_gas(*ARGS, matrix_free=True)
    mbpt2 = o.mbpt2()
    e, stop, _ = o.solve(TOL)
    t2 = o.t2()
    idx = np.sort(np.random.default_rng(SEED).choice(o.nt2, N_SAMPLES, replace=False))
    out = {
        "workload": "electron gas N=66, r_s=0.5, n^2<=52",
        "reference_cli": "--electronGas -r 0.5 -s 46 -p 5",
        "args": list(ARGS),
        "nsp": o.nsp, "nh": o.nh, "nt2": o.nt2,
        "eref": o.eref(), "mbpt2": mbpt2,
        "tol": TOL, "stop": stop, "energies": [float(x) for x in e],
        "t2_sum_sq": float(np.dot(t2, t2)),
        "t2_sample_seed": SEED,
        "t2_sample_index": idx.tolist(),
        "t2_sample_value": t2[idx].tolist(),
        "oracle_seconds": time.time() - t0,
    }
    with open(os.path.join(HERE, "heg66_oracle.json"), "w") as f:
        json.dump(out, f, indent=1)
    print(f"{len(e)} iterations, stop {stop}, E {e[-1]!r}, {out['oracle_seconds']:.0f} s")


if __name__ == "__main__":
    main()
