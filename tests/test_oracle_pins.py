"""Pins of the CPU oracle (oracle/ccd_oracle.c) before it is trusted.

The reference ships no tests or fixtures for this path and cannot be built
here (it needs <cblas.h>, which the image lacks), so the oracle is pinned by
what the mathematics fixes:

* pairing closed forms: E_ref = 2 sum_{p<=4}(p-1) xi - 2g, and
  MBPT2 = sum_{p<=4<q} (g^2/4) / (2(p-q) xi - g);
* an independent dense spin-orbital CCD (oracle/dense_ccd.py), iteration by
  iteration, on the pairing model and on three tiny electron-gas bases;
* each diagram on random amplitudes against its dense textbook form, so a
  dropped term, wrong sign or transposed operand fails on its own;
* the electron-gas channel histograms (N=14, N=54) counted independently;
* protocol edge cases: first iterate = MBPT2/2, max-iteration halt, NaN exit,
  T2 antisymmetry.
"""
import numpy as np
import pytest

from oracle import dense_ccd as dc
from oracle.pyoracle import Oracle


def _rel(a, b):
    return abs(a - b) / max(abs(b), 1e-300)


# ---------------------------------------------------------------- closed forms

def test_pairing_closed_forms():
    o = Oracle.pairing(8, 8, 0.5, 1.0)
    g, xi = 0.5, 1.0
    eref = 2 * sum((p - 1) * xi for p in range(1, 5)) - 4 * g / 2
    mbpt2 = sum((g * g / 4) / (2 * (p - q) * xi - g) for p in range(1, 5) for q in range(5, 9))
    assert o.eref() == pytest.approx(eref, abs=1e-14)
    assert o.eref() == 11.0
    assert _rel(o.mbpt2(), mbpt2) < 1e-14
    assert _rel(o.mbpt2(), -0.14320318401454299) < 1e-14


@pytest.mark.parametrize("g", [0.2, 0.8])
def test_pairing_mbpt2_other_g(g):
    o = Oracle.pairing(8, 8, g, 1.0)
    mbpt2 = sum((g * g / 4) / (2 * (p - q) - g) for p in range(1, 5) for q in range(5, 9))
    assert _rel(o.mbpt2(), mbpt2) < 1e-13
    assert o.eref() == pytest.approx(12 - 2 * g, abs=1e-13)


# ---------------------------------------------------------------- full iteration vs dense CCD

def test_pairing_iterations_match_dense():
    o = Oracle.pairing(8, 8, 0.5, 1.0)
    e_o, stop, _ = o.solve(1e-12)
    s = dc.pairing(8, 8, 0.5, 1.0)
    e_d, stop_d = s.solve(1e-12)
    assert stop == stop_d == "converged"
    assert len(e_o) == len(e_d)
    assert np.max(np.abs(e_o - np.array(e_d))) < 1e-13
    assert e_o[0] == pytest.approx(0.5 * o.mbpt2(), rel=1e-14)


@pytest.mark.parametrize("n_shells", [3, 4, 5])
def test_electron_gas_tiny_match_dense(n_shells):
    o = Oracle.electron_gas(1.0, n_shells, 2)
    h = dc.electron_gas(1.0, n_shells, 2)
    assert o.nsp == h.nsp and o.nh == h.nh == 14
    assert h.fock_offdiag == 0.0
    assert _rel(o.eref(), h.eref) < 1e-14
    assert _rel(o.mbpt2(), h.mbpt2()) < 1e-13
    e_o, stop, _ = o.solve(1e-10)
    e_d, stop_d = h.solve(1e-10)
    assert stop == stop_d == "converged" and len(e_o) == len(e_d)
    assert np.max(np.abs(e_o - np.array(e_d)) / np.abs(e_d)) < 1e-12


# ---------------------------------------------------------------- diagrams on random amplitudes

def _index_maps(o, s):
    """Oracle s.p. index -> dense s.p. index, via quantum numbers."""
    qn = o.qn()
    if isinstance(s.labels[0][0], tuple):
        key = {lab: k for k, lab in enumerate(s.labels)}
        return np.array([key[((int(q[0]), int(q[1]), int(q[2])), int(q[3]))] for q in qn])
    key = {lab: k for k, lab in enumerate(s.labels)}
    return np.array([key[(int(q[0]), int(q[1]))] for q in qn])


def _setup(kind):
    if kind == "pairing":
        return Oracle.pairing(8, 8, 0.5, 1.0), dc.pairing(8, 8, 0.5, 1.0)
    return Oracle.electron_gas(1.0, 3, 2), dc.electron_gas(1.0, 3, 2)


@pytest.mark.parametrize("kind", ["pairing", "heg"])
def test_diagrams_match_dense(kind):
    o, s = _setup(kind)
    rng = np.random.default_rng(7)
    mp = _index_maps(o, s)
    nh = o.nh
    lab = o.t2_labels()
    A, B, I, J = (mp[lab[:, k]] for k in range(4))
    A, B = A - nh, B - nh
    td = rng.uniform(-1, 1, s.v_ppHH.shape)
    # only symmetry-allowed elements exist in the blocked layout
    mask = np.zeros(td.shape, dtype=bool)
    mask[A, B, I, J] = True
    td = np.where(mask, td, 0.0)
    t = td[A, B, I, J]
    # ladder (calcT2Diff_pphh)
    d = o.ladder(t)
    dd = s.ladder(td)[A, B, I, J]
    assert np.max(np.abs(d - dd)) <= 1e-12 * np.max(np.abs(dd))
    # ring (calcT2Diff_phhp) through the ph remap
    tph = o.scatter_ph(t)
    r = o.ring(tph)
    lph = o.tph_labels()                    # a i j b
    a2, i2, j2, b2 = (mp[lph[:, k]] for k in range(4))
    rd = s.ring_unpermuted(td)[a2 - nh, i2, j2, b2 - nh]
    assert np.max(np.abs(r - rd)) <= 1e-12 * np.max(np.abs(rd))
    # the remap is a bijection onto the phhp layout
    assert np.array_equal(np.sort(tph), np.sort(t))
    # one-body intermediates
    chi, psi = o.onebody(t)
    chid, psid = s.onebody(td)
    assert np.allclose(chi, chid[mp[:nh]], rtol=0, atol=1e-12 * np.max(np.abs(chid)))
    assert np.allclose(psi, psid[mp[nh:] - nh], rtol=0, atol=1e-12 * np.max(np.abs(psid)))


@pytest.mark.parametrize("kind", ["pairing", "heg"])
def test_full_step_matches_dense(kind):
    o, s = _setup(kind)
    rng = np.random.default_rng(11)
    mp = _index_maps(o, s)
    nh = o.nh
    lab = o.t2_labels()
    A, B, I, J = (mp[lab[:, k]] for k in range(4))
    A, B = A - nh, B - nh
    mask = np.zeros(s.v_ppHH.shape, dtype=bool)
    mask[A, B, I, J] = True
    if kind == "pairing":
        # The reference keeps only the diagonal of the one-body intermediates
        # (calc.cpp:16-17, 79-80).  For the pairing model that holds for
        # amplitudes with the pair structure of V (both pairs in one level),
        # which every CCD iterate has; a generic S_z-conserving T does not.
        mask &= s.v_ppHH != 0
    td = np.where(mask, rng.uniform(-0.1, 0.1, s.v_ppHH.shape), 0.0)
    td = td - td.transpose(1, 0, 2, 3)
    td = td - td.transpose(0, 1, 3, 2)
    t = td[A, B, I, J]
    t_new, e = o.step(t, 0.5)
    td_new = 0.5 * s.residual(td) / s.denom + 0.5 * td
    ref = td_new[A, B, I, J]
    assert np.max(np.abs(t_new - ref)) <= 1e-12 * np.max(np.abs(ref))
    assert _rel(e, s.energy(td_new)) < 1e-12


# ---------------------------------------------------------------- channel histograms

@pytest.mark.parametrize("args,expect", [
    ((1.0, 19, 2), dict(nsp=778, nt2=88548, active=75, maxpp=764, maxhh=14, medhh=2, mod=1902)),
    ((1.0, 32, 4), dict(nsp=1850, nt2=2925948, active=375, maxpp=1796, maxhh=54, medhh=6, mod=5298)),
])
def test_electron_gas_channel_histogram(args, expect):
    o = Oracle.electron_gas(*args, load=False)
    d = o.dims()
    act = (d["pp"] * d["hh"]) > 0
    assert o.nsp == expect["nsp"]
    assert o.nt2 == expect["nt2"]
    assert int(act.sum()) == expect["active"]
    assert int(d["pp"].max()) == expect["maxpp"]
    assert int(d["hh"].max()) == expect["maxhh"]
    assert int(np.median(d["hh"][act])) == expect["medhh"]
    assert int(((d["mph"] * d["mhp"]) > 0).sum()) == expect["mod"]
    assert int((d["mph"] * d["mhp"]).sum()) == o.nt2


# ---------------------------------------------------------------- protocol edge cases

def test_max_iterations_halt():
    o = Oracle.pairing(8, 8, 0.5, 1.0)
    e, stop, _ = o.solve(-1.0)            # |dE| > -1 always holds
    assert stop == "max_iterations" and len(e) == 101
    e, stop, _ = o.solve(0.0)             # the pairing iterate reaches an exact fixed point
    assert stop == "converged" and len(e) < 101


def test_nan_exits_loop():
    o = Oracle.pairing(8, 8, 0.0, 0.0)     # every denominator 0, V = 0: 0/0
    e, stop, _ = o.solve(1e-8)
    assert stop == "nan" and len(e) == 1 and np.isnan(e[0])


def test_tolerance_above_one_runs_nothing():
    o = Oracle.pairing(8, 8, 0.5, 1.0)
    e, stop, _ = o.solve(2.0)             # dE starts at 1.0 (main.cpp:492)
    assert len(e) == 0 and stop == "converged"


def test_amplitudes_stay_antisymmetric():
    o = Oracle.electron_gas(1.0, 4, 2)
    o.solve(1e-6)
    t = o.t2()
    lab = o.t2_labels()
    key = {tuple(r): k for k, r in enumerate(lab.tolist())}
    sw = np.array([key[(b, a, i, j)] for a, b, i, j in lab.tolist()])
    assert np.max(np.abs(t + t[sw])) <= 1e-14 * np.max(np.abs(t))


def test_inf_matter_published_energy():
    """The reference's only published result (src/README.md:8-10): pure neutron
    matter, Minnesota potential, density 0.08 fm^-3, 14 neutrons, tolerance
    1e-8, E_CCD/A = 10.155078 MeV.  The README's positional "3 shells" is the
    modern -s 4 (momenta with n^2 <= 3, 54 states; -s 3 gives 10.166175)."""
    o = Oracle.inf_matter(0.08, 4, 2)
    assert o.nsp == 54 and o.nh == 14
    e, stop, _ = o.solve(1e-8)
    assert stop == "converged"
    e_per_a = (o.eref() + e[-1]) / 14
    assert abs(e_per_a - 10.155078) < 5e-7
    assert e[0] == pytest.approx(0.5 * o.mbpt2(), rel=1e-13)



def test_oracle_matrix_free_ladder_is_identical():
    """The oracle's -m 1 ladder (V^pppp rebuilt per channel, calc.cpp:233-262)
    gives the stored-V results bit for bit; the N=66 golden file relies on it."""
    a = Oracle.electron_gas(1.0, 19, 2)
    b = Oracle.electron_gas(1.0, 19, 2, matrix_free=True)
    t = np.random.default_rng(5).uniform(-0.05, 0.05, a.nt2)
    assert np.array_equal(a.ladder(t), b.ladder(t))
    e1, s1, _ = a.solve(1e-10)
    e2, s2, _ = b.solve(1e-10)
    assert s1 == s2 == "converged" and np.array_equal(e1, e2)


def test_heg66_golden_file_is_consistent():
    """The committed N=66 oracle solve: protocol invariants that hold for any
    correct run (first iterate = MBPT2/2, the stop rule at the last step)."""
    import json
    import os
    with open(os.path.join(os.path.dirname(__file__), "golden", "heg66_oracle.json")) as f:
        g = json.load(f)
    e = np.array(g["energies"])
    assert g["nsp"] == 3150 and g["nh"] == 66 and g["nt2"] == 7765428
    assert e[0] == pytest.approx(0.5 * g["mbpt2"], rel=1e-13)
    assert abs(e[-1] - e[-2]) <= g["tol"] < np.min(np.abs(np.diff(e[:-1])))
    o = Oracle.electron_gas(*g["args"], load=False)
    assert o.nsp == g["nsp"] and o.nh == g["nh"] and o.nt2 == g["nt2"]
    assert o.eref() == g["eref"]


def test_heg54_golden_file_is_what_the_oracle_computes():
    """The committed N=54 oracle solve (the headline workload): protocol
    invariants, and its MBPT2 and first two iterations recomputed here by the
    oracle (OpenMP sums over channels: equal to 1e-13 relative)."""
    import json
    import os
    with open(os.path.join(os.path.dirname(__file__), "golden", "heg54_oracle.json")) as f:
        g = json.load(f)
    e = np.array(g["energies"])
    assert g["nsp"] == 1850 and g["nh"] == 54 and g["nt2"] == 2925948
    assert g["stop"] == "converged" and len(e) == 33
    assert e[0] == pytest.approx(0.5 * g["mbpt2"], rel=1e-13)
    assert abs(e[-1] - e[-2]) <= g["tol"] < np.min(np.abs(np.diff(e[:-1])))
    o = Oracle.electron_gas(*g["args"])
    assert o.nsp == g["nsp"] and o.nh == g["nh"] and o.nt2 == g["nt2"]
    assert o.eref() == g["eref"]
    assert _rel(o.mbpt2(), g["mbpt2"]) < 1e-13
    t = np.zeros(o.nt2)
    for k in range(2):
        t, ek = o.step(t)
        assert _rel(ek, e[k]) < 1e-13
