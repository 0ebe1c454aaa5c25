"""The headline workload end to end against the oracle: the electron gas with
54 electrons, r_s = 1.0, n^2 <= 36 (1850 s.p. states, 2,925,948 amplitudes;
reference CLI --electronGas -r 1.0 -s 32 -p 4 -t 1e-10).

The oracle's full FP64 solve is committed as tests/golden/heg54_oracle.json
(written by tests/golden/make_heg54.py).  Bars:
  * FP64 GPU solve (stored V and matrix-free): MBPT2 and every iteration's
    energy within 1e-10 relative, the same iteration count and stop reason
    (the stop rule's margin, |dE| of the last two iterations against tol, is
    5-10% of tol, far above the 1e-14 relative differences of the energies),
    the amplitudes within 1e-10 x max|t| at 4096 seeded positions and their
    sum of squares within 1e-10 relative;
  * mixed-precision solves (FP32 contractions, FP64 sweeps): converged
    energy within 1e-7 Ha per particle of the oracle's converged energy, for
    the switch points 1e-6 and 1e-8 and for FP32 iterations only.
"""
import json
import os

import numpy as np
import pytest

from coupled_cluster_doubles_amd import FP32, FP64, Context, Model

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
E_REL = 1e-10
A_REL = 1e-10
PER_PARTICLE = 1e-7


@pytest.fixture(scope="module")
def golden():
    with open(os.path.join(HERE, "golden", "heg54_oracle.json")) as f:
        return json.load(f)


@pytest.fixture(scope="module")
def model(golden):
    m = Model.electron_gas(*golden["args"])
    assert (m.n_sp, m.n_holes, m.n_t2) == (golden["nsp"], golden["nh"], golden["nt2"])
    return m


@pytest.mark.parametrize("free", [False, True])
def test_heg54_fp64_solve_matches_oracle(require_gpu, golden, model, free):
    print(f"heg54 FP64 solve, matrix_free={free}", flush=True)
    assert abs(model.eref() - golden["eref"]) <= 1e-14 * abs(golden["eref"])
    ctx = Context(model, FP64, matrix_free=free)
    assert abs(ctx.mbpt2() - golden["mbpt2"]) <= E_REL * abs(golden["mbpt2"])
    e_g, stop_g, n_g = ctx.solve(golden["tol"])
    e_o = np.array(golden["energies"])
    assert stop_g == golden["stop"] == "converged"
    assert n_g == len(e_o)
    assert np.max(np.abs(e_g - e_o) / np.abs(e_o)) < E_REL
    assert e_g[0] == pytest.approx(0.5 * golden["mbpt2"], rel=1e-12)
    t = ctx.get_t2()
    idx = np.array(golden["t2_sample_index"])
    ref = np.array(golden["t2_sample_value"])
    assert np.max(np.abs(t[idx] - ref)) <= A_REL * golden["t2_max_abs"]
    assert abs(np.dot(t, t) - golden["t2_sum_sq"]) <= A_REL * golden["t2_sum_sq"]
    assert abs(np.max(np.abs(t)) - golden["t2_max_abs"]) <= A_REL * golden["t2_max_abs"]


@pytest.mark.parametrize("switch_tol", [1e-6, 1e-8, 1e-10])
def test_heg54_mixed_solve_within_1e7_per_particle_of_oracle(require_gpu, golden, model, switch_tol):
    """The mixed solve's converged energy against the oracle's (not against
    the GPU's own FP64 solve); switch_tol = tol runs FP32 iterations only."""
    print(f"heg54 mixed solve, switch_tol={switch_tol}", flush=True)
    ctx = Context(model, FP32, matrix_free=True)
    e, stop, n, n32 = ctx.solve_mixed(golden["tol"], switch_tol)
    assert stop == "converged" and n32 > 0
    if switch_tol > golden["tol"]:
        assert n > n32
    err = abs(e[-1] - golden["energies"][-1]) / golden["nh"]
    assert err <= PER_PARTICLE, err
