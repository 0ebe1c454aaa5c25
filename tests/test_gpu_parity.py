"""Parity of the HIP path (through the C ABI) with the CPU oracle.

Tolerances (FP64 path): energies within 1e-10 relative of the oracle, the
bar the reference comparison is held to; amplitude arrays within
1e-12 * max|reference| (the only differences are summation order and FMA
contraction).  The ph remap is index work and must be exact.
"""
import numpy as np
import pytest

from coupled_cluster_doubles_amd import FP64, Context, Model
from oracle.pyoracle import Oracle

pytestmark = pytest.mark.gpu

E_REL = 1e-10
A_REL = 1e-12


def _rel(a, b):
    return abs(a - b) / max(abs(b), 1e-300)


def _pair(kind, args):
    if kind == "pairing":
        return Model.pairing(*args), Oracle.pairing(*args)
    if kind == "infmatter":
        return Model.inf_matter(*args), Oracle.inf_matter(*args)
    return Model.electron_gas(*args), Oracle.electron_gas(*args)


def test_pairing_solve(require_gpu):
    m, o = _pair("pairing", (8, 8, 0.5, 1.0))
    ctx = Context(m, FP64)
    assert _rel(ctx.mbpt2(), o.mbpt2()) < 1e-13
    e_g, stop_g, n_g = ctx.solve(1e-12)
    e_o, stop_o, _ = o.solve(1e-12)
    assert stop_g == stop_o == "converged"
    assert n_g == len(e_o)
    assert np.max(np.abs(e_g - e_o) / np.abs(e_o)) < E_REL
    assert e_g[0] == pytest.approx(0.5 * o.mbpt2(), rel=1e-13)


def test_electron_gas_n14_solve(require_gpu):
    m, o = _pair("heg", (1.0, 19, 2))
    ctx = Context(m, FP64)
    assert _rel(ctx.mbpt2(), o.mbpt2()) < E_REL
    e_g, stop_g, n_g = ctx.solve(1e-10)
    e_o, stop_o, _ = o.solve(1e-10)
    assert stop_g == stop_o == "converged"
    assert n_g == len(e_o)
    k = n_g
    assert np.max(np.abs(e_g[:k] - e_o[:k]) / np.abs(e_o[:k])) < E_REL
    assert _rel(e_g[-1], e_o[-1]) < E_REL
    t_g, t_o = ctx.get_t2(), o.t2()
    assert np.max(np.abs(t_g - t_o)) <= 1e-9 * np.max(np.abs(t_o))


def test_inf_matter_solve_and_published_energy(require_gpu):
    """Pure neutron matter, Minnesota potential, density 0.08 fm^-3, 14
    neutrons in 54 states, tolerance 1e-8: E_CCD/A = 10.155078 MeV
    (src/README.md:8-10)."""
    m, o = _pair("infmatter", (0.08, 4, 2))
    ctx = Context(m, FP64)
    assert _rel(ctx.mbpt2(), o.mbpt2()) < E_REL
    e_g, stop_g, n_g = ctx.solve(1e-8)
    e_o, stop_o, _ = o.solve(1e-8)
    assert stop_g == stop_o == "converged"
    assert n_g == len(e_o)
    assert np.max(np.abs(e_g - e_o) / np.abs(e_o)) < E_REL
    assert abs((m.eref() + e_g[-1]) / 14 - 10.155078) < 5e-7


@pytest.mark.parametrize("kind,args", [
    ("pairing", (8, 8, 0.5, 1.0)),
    ("heg", (1.0, 4, 2)),
    ("heg", (1.0, 19, 2)),
    ("infmatter", (0.08, 4, 2)),
    ("pairing", (14, 14, 0.4, 1.0)),      # hh = 98: 7 column tiles
    ("pairing", (16, 16, 0.6, 1.0)),      # hh = 128: 8 column tiles
])
def test_diagrams_on_random_amplitudes(require_gpu, kind, args):
    m, o = _pair(kind, args)
    ctx = Context(m, FP64)
    rng = np.random.default_rng(3)
    t = rng.uniform(-0.05, 0.05, m.n_t2)
    lab = m.t2_labels()
    key = {tuple(r): k for k, r in enumerate(lab.tolist())}
    sw_ab = np.array([key[(b, a, i, j)] for a, b, i, j in lab.tolist()])
    sw_ij = np.array([key[(a, b, j, i)] for a, b, i, j in lab.tolist()])
    ta = t - t[sw_ab]
    ta = ta - ta[sw_ij]
    # ladder: calcT2Diff_pphh (the GPU ladder works on the antisymmetric part,
    # which is all the reference's iterates ever are)
    d_g, d_o = ctx.ladder(ta), o.ladder(ta)
    assert np.max(np.abs(d_g - d_o)) <= A_REL * np.max(np.abs(d_o))
    # ph remap: exact
    tph_g, tph_o = ctx.remap_ph(t), o.scatter_ph(t)
    assert np.array_equal(tph_g, tph_o)
    # ring: calcT2Diff_phhp
    r_g, r_o = ctx.ring(tph_o), o.ring(tph_o)
    assert np.max(np.abs(r_g - r_o)) <= A_REL * np.max(np.abs(r_o))
    # one-body intermediates
    chi_g, psi_g = ctx.onebody(t)
    chi_o, psi_o = o.onebody(t)
    assert np.max(np.abs(chi_g - chi_o)) <= A_REL * np.max(np.abs(chi_o))
    assert np.max(np.abs(psi_g - psi_o)) <= A_REL * np.max(np.abs(psi_o))
    # a full step from antisymmetric amplitudes
    tn_g, e_g = ctx.step(ta)
    tn_o, e_o = o.step(ta)
    assert np.max(np.abs(tn_g - tn_o)) <= A_REL * np.max(np.abs(tn_o))
    assert _rel(e_g, e_o) < E_REL


@pytest.mark.parametrize("kind,args", [
    ("pairing", (2, 2, 0.5, 1.0)),       # one pair in two levels: 6 amplitudes
    ("pairing", (3, 2, 0.3, 1.0)),       # odd level count
    ("pairing", (5, 6, 0.4, 1.0)),       # more holes than particles
    ("pairing", (10, 4, 0.5, 1.0)),      # more particles than holes
    ("heg", (1.0, 2, 1)),                # 2 electrons in 14 states
    ("heg", (2.0, 3, 1)),                # 2 electrons in 38 states, r_s = 2
])
def test_small_and_ragged_problems(require_gpu, kind, args):
    """Full solves on the smallest and most uneven bases: the reference's
    protocol (energies, iteration count, stop reason) in both V modes, and
    the mixed solve within 1e-7 per particle."""
    from coupled_cluster_doubles_amd import FP32
    m, o = _pair(kind, args)
    e_o, stop_o, _ = o.solve(1e-12)
    for free in (False, True):
        ctx = Context(m, FP64, matrix_free=free)
        assert _rel(ctx.mbpt2(), o.mbpt2()) < E_REL
        e_g, stop_g, n_g = ctx.solve(1e-12)
        assert stop_g == stop_o == "converged"
        assert n_g == len(e_o)
        k = n_g
        assert np.max(np.abs(e_g[:k] - e_o[:k]) / np.abs(e_o[:k])) < E_REL
    e_m, stop_m, _, _ = Context(m, FP32).solve_mixed(1e-12, 1e-8)
    assert stop_m == "converged" and abs(e_m[-1] - e_o[-1]) / m.n_holes <= 1e-7


@pytest.mark.parametrize("kind,args", [
    ("heg", (1.0, 1, 1)),                # every state a hole: no amplitudes
    ("pairing", (4, 0, 0.5, 1.0)),       # no particles: no holes, no amplitudes
])
def test_empty_problems(require_gpu, kind, args):
    """No amplitudes at all: MBPT2 = 0, and the loop runs the reference's two
    iterations (E = 0 after E_prev = 100, then dE = 0) and stops converged."""
    m, o = _pair(kind, args)
    assert m.n_t2 == 0
    e_o, stop_o, _ = o.solve(1e-10)
    ctx = Context(m, FP64)
    assert ctx.mbpt2() == 0.0
    e_g, stop_g, n_g = ctx.solve(1e-10)
    assert (stop_g, n_g) == (stop_o, len(e_o)) == ("converged", 2)
    assert np.all(e_g == 0.0)
    assert ctx.get_t2().size == 0


def test_host_v_matches_generated_v(require_gpu):
    m = Model.electron_gas(1.0, 6, 2)
    c_gen = Context(m, FP64)
    c_host = Context(m, FP64, host_v=True)
    assert c_gen.mbpt2() == c_host.mbpt2()
    t = np.random.default_rng(5).uniform(-0.05, 0.05, m.n_t2)
    a, ea = c_gen.step(t)
    b, eb = c_host.step(t)
    assert np.array_equal(a, b) and ea == eb


def test_inf_matter_generated_v_matches_host_v(require_gpu):
    """The device Minnesota generator against the host chains (which equal the
    oracle's bit for bit): the device exp may differ in the last bit."""
    m = Model.inf_matter(0.08, 4, 2)
    c_gen = Context(m, FP64)
    c_host = Context(m, FP64, host_v=True)
    assert abs(c_gen.mbpt2() - c_host.mbpt2()) <= 1e-14 * abs(c_host.mbpt2())
    t = np.random.default_rng(5).uniform(-0.05, 0.05, m.n_t2)
    a, ea = c_gen.step(t)
    b, eb = c_host.step(t)
    assert np.max(np.abs(a - b)) <= 1e-13 * np.max(np.abs(b)) and abs(ea - eb) <= 1e-13 * abs(eb)


def test_headline_n54_matches_oracle_two_iterations(require_gpu):
    m, o = _pair("heg", (1.0, 32, 4))
    ctx = Context(m, FP64)
    mb = ctx.mbpt2()
    assert _rel(mb, o.mbpt2()) < E_REL
    e0 = ctx.iterate()
    assert _rel(e0, 0.5 * mb) < 1e-12            # zero start, mixing 1/2
    e1 = ctx.iterate()
    t_o = np.zeros(o.nt2)
    t_o, eo0 = o.step(t_o)
    t_o, eo1 = o.step(t_o)
    assert _rel(e0, eo0) < E_REL and _rel(e1, eo1) < E_REL
    t_g = ctx.get_t2()
    assert np.max(np.abs(t_g - t_o)) <= A_REL * np.max(np.abs(t_o))


def test_headline_n54_amplitudes_stay_antisymmetric(require_gpu):
    m = Model.electron_gas(1.0, 32, 4)
    ctx = Context(m, FP64)
    for _ in range(4):
        ctx.iterate()
    t = ctx.get_t2()
    lab = m.t2_labels()
    # (b,a) partner of each element via a dictionary on (a,b,i,j) keys
    n = m.n_sp
    keys = ((lab[:, 0].astype(np.int64) * n + lab[:, 1]) * n + lab[:, 2]) * n + lab[:, 3]
    swk = ((lab[:, 1].astype(np.int64) * n + lab[:, 0]) * n + lab[:, 2]) * n + lab[:, 3]
    order = np.argsort(keys)
    pos = order[np.searchsorted(keys[order], swk)]
    assert np.max(np.abs(t + t[pos])) <= 1e-13 * np.max(np.abs(t))


@pytest.mark.parametrize("kind,args", [
    ("pairing", (8, 8, 0.5, 1.0)),       # generic per-element generator
    ("heg", (1.0, 19, 2)),               # electron-gas table generator
    ("heg", (1.0, 32, 4)),               # headline shape
    ("infmatter", (0.08, 4, 2)),         # Minnesota generator
])
def test_matrix_free_is_bit_identical_to_stored(require_gpu, kind, args):
    """-m 1 twin: V^pppp rebuilt inside the ladder gives the stored-V iterates
    bit for bit (same elements, same summation order)."""
    m = _pair(kind, args)[0]
    stored = Context(m, FP64)
    free = Context(m, FP64, matrix_free=True)
    assert free.device_bytes < stored.device_bytes or kind == "pairing"
    for _ in range(3):
        assert free.iterate() == stored.iterate()
    assert np.array_equal(free.get_t2(), stored.get_t2())
    t = np.random.default_rng(11).uniform(-0.05, 0.05, m.n_t2)
    assert np.array_equal(free.ladder(t), stored.ladder(t))


def test_matrix_free_solve_matches_oracle(require_gpu):
    m, o = _pair("heg", (1.0, 19, 2))
    ctx = Context(m, FP64, matrix_free=True)
    e_g, stop_g, n_g = ctx.solve(1e-10)
    e_o, stop_o, _ = o.solve(1e-10)
    assert stop_g == stop_o == "converged"
    assert n_g == len(e_o)
    k = n_g
    assert np.max(np.abs(e_g[:k] - e_o[:k]) / np.abs(e_o[:k])) < E_REL


# ---- mixed precision (FP32 contractions, FP64 update and energy)

F32_REL = 2e-5          # FP32 contraction vs FP64 oracle, relative to the largest element


def test_fp32_diagrams_close_to_oracle(require_gpu):
    from coupled_cluster_doubles_amd import FP32
    m, o = _pair("heg", (1.0, 19, 2))
    ctx = Context(m, FP32)
    rng = np.random.default_rng(7)
    t = rng.uniform(-0.05, 0.05, m.n_t2)
    lab = m.t2_labels()
    key = {tuple(r): k for k, r in enumerate(lab.tolist())}
    sw_ab = np.array([key[(b, a, i, j)] for a, b, i, j in lab.tolist()])
    sw_ij = np.array([key[(a, b, j, i)] for a, b, i, j in lab.tolist()])
    ta = t - t[sw_ab]
    ta = ta - ta[sw_ij]
    d_g, d_o = ctx.ladder(ta), o.ladder(ta)
    assert np.max(np.abs(d_g - d_o)) <= F32_REL * np.max(np.abs(d_o))
    tph = o.scatter_ph(t)
    r_g, r_o = ctx.ring(tph), o.ring(tph)
    assert np.max(np.abs(r_g - r_o)) <= F32_REL * np.max(np.abs(r_o))
    # the same context in FP64 is the FP64 path
    ctx.set_precision(FP64)
    assert np.array_equal(ctx.ladder(ta), Context(m, FP64).ladder(ta))


@pytest.mark.parametrize("kind,args", [
    ("pairing", (12, 12, 0.5, 1.0)),      # mod channels of 3 and 5 column tiles
    ("pairing", (10, 10, 0.5, 1.0)),      # 2 and 4 tiles
    ("pairing", (16, 8, 0.5, 1.0)),       # 3 and 6 tiles
    ("heg", (1.0, 10, 3)),                # 1 and 2 tiles, many channels
    ("heg", (1.0, 32, 4)),                # headline shape: 1 to 4 tiles
])
def test_ring_every_tile_class(require_gpu, kind, args):
    """calcT2Diff_phhp on random T^ph for every mod-channel width the
    LDS-staged (1-5 tiles) and per-wave (6-8 tiles) ring kernels take, in
    FP64 (1e-12) and FP32 (2e-5, relative to the largest element), and with
    the staged ring on BF16x3 split products (2e-4; channels of 6-8 tiles
    stay FP32)."""
    from coupled_cluster_doubles_amd import FP32
    m, o = _pair(kind, args)
    tph = np.random.default_rng(5).uniform(-0.05, 0.05, m.n_t2)
    r_o = o.ring(tph)
    r64 = Context(m, FP64).ring(tph)
    assert np.max(np.abs(r64 - r_o)) <= A_REL * np.max(np.abs(r_o))
    r32 = Context(m, FP32).ring(tph)
    assert np.max(np.abs(r32 - r_o)) <= F32_REL * np.max(np.abs(r_o))
    ctx = Context(m, FP32)
    p = ctx.ladder_plan()
    p["ring_fp32"] = "bf16x3"
    ctx.set_ladder_plan(p)
    rb = ctx.ring(tph)
    err = np.max(np.abs(rb - r_o)) / np.max(np.abs(r_o))
    print(f"{kind}{args} BF16x3 ring: max |R - R_oracle| / max|R_oracle| = {err:.2e}", flush=True)
    assert err <= 2e-4
    # the FP64 iterations of the same context keep the FP64 ring
    ctx.set_precision(FP64)
    assert np.array_equal(ctx.ring(tph), r64)


@pytest.mark.parametrize("free", [False, True])
def test_precision_switch_continues_the_same_iteration(require_gpu, free):
    """FP32 iterations keep T^ph in FP32; after switching a context to FP64
    it must iterate exactly as an FP64 context started from the same
    amplitudes (T^ph rebuilt in the new type, nothing stale)."""
    from coupled_cluster_doubles_amd import FP32
    m = Model.electron_gas(1.0, 19, 2)
    a = Context(m, FP32, matrix_free=free)
    a.iterate_n(3)
    t3 = a.get_t2()
    a.set_precision(FP64)
    ea = a.iterate_n(2)
    b = Context(m, FP64, matrix_free=free)
    b.set_t2(t3)
    eb = b.iterate_n(2)
    assert np.array_equal(ea, eb)
    assert np.array_equal(a.get_t2(), b.get_t2())
    # and back: FP32 again continues as an FP32 context from those amplitudes
    a.set_precision(FP32)
    ea32 = a.iterate_n(2)
    c = Context(m, FP32, matrix_free=free)
    c.set_t2(b.get_t2())
    assert np.array_equal(ea32, c.iterate_n(2))


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_entry_points_leave_the_iteration_state_alone(require_gpu, prec):
    """ccd_ladder, ccd_ring and ccd_remap_ph compute a caller's term in the
    context's work arrays (the antisymmetrised rows, T^ph); the next
    iterations must still start from the context's own amplitudes."""
    from coupled_cluster_doubles_amd import FP32
    p = FP32 if prec == "fp32" else FP64
    m = Model.electron_gas(1.0, 19, 2)
    a = Context(m, p, matrix_free=True)
    a.iterate_n(2)
    t = a.get_t2()
    b = Context(m, p, matrix_free=True)
    b.set_t2(t)
    rng = np.random.default_rng(11)
    a.ladder(rng.uniform(-0.05, 0.05, m.n_t2))
    a.ring(rng.uniform(-0.05, 0.05, m.n_t2))
    a.remap_ph(rng.uniform(-0.05, 0.05, m.n_t2))
    assert np.array_equal(a.get_t2(), t)
    assert np.array_equal(a.iterate_n(2), b.iterate_n(2))
    assert np.array_equal(a.get_t2(), b.get_t2())


@pytest.mark.parametrize("free", [False, True])
def test_fp32_stored_and_generated_v_agree_bitwise(require_gpu, free):
    from coupled_cluster_doubles_amd import FP32
    m = Model.electron_gas(1.0, 19, 2)
    a = Context(m, FP32)
    b = Context(m, FP32, matrix_free=free)
    for _ in range(3):
        assert a.iterate() == b.iterate()
    assert np.array_equal(a.get_t2(), b.get_t2())


@pytest.mark.parametrize("args,n_part", [((1.0, 19, 2), 14), ((1.0, 32, 4), 54)])
def test_mixed_solve_within_1e7_per_particle(require_gpu, args, n_part):
    """FP32 contractions then FP64 sweeps: converged energy within 1e-7 Ha per
    particle of the FP64 solve (in practice within the tolerance itself)."""
    from coupled_cluster_doubles_amd import FP32
    m = Model.electron_gas(*args)
    e64, stop64, n64 = Context(m, FP64, matrix_free=True).solve(1e-10)
    e, stop, n, n32 = Context(m, FP32, matrix_free=True).solve_mixed(1e-10, 1e-6)
    assert stop == stop64 == "converged"
    assert n32 > 0 and n > n32
    assert abs(e[-1] - e64[-1]) / n_part <= 1e-7


# ---- device-side loop control

def _host_loop(ctx, tol, max_iters=100):
    """The reference protocol run from the host, one synchronous step at a time."""
    ctx.reset()
    diff, prev, out = 1.0, 100.0, []
    while abs(diff) > tol:
        if len(out) > max_iters:
            return np.array(out), "max_iterations"
        e = ctx.iterate()
        diff, prev = e - prev, e
        out.append(e)
    return np.array(out), ("nan" if np.isnan(diff) else "converged")


@pytest.mark.parametrize("tol", [1e-12, 0.0, -1.0])
def test_device_loop_matches_host_protocol(require_gpu, tol):
    m = Model.pairing(8, 8, 0.5, 1.0)
    ctx = Context(m, FP64)
    e_d, stop_d, n_d = ctx.solve(tol)
    e_h, stop_h = _host_loop(ctx, tol)
    assert stop_d == stop_h and n_d == len(e_h)
    assert np.array_equal(e_d, e_h)
    if tol < 0:
        assert n_d == 101
    # the amplitudes left behind are those of the last real iteration
    t_d = ctx.get_t2()
    ctx.solve(tol)
    assert np.array_equal(ctx.get_t2(), t_d)


def test_iterate_n_matches_single_steps(require_gpu):
    m = Model.electron_gas(1.0, 19, 2)
    a, b = Context(m, FP64, matrix_free=True), Context(m, FP64, matrix_free=True)
    e_n = a.iterate_n(7)
    e_1 = np.array([b.iterate() for _ in range(7)])
    assert np.array_equal(e_n, e_1)
    assert np.array_equal(a.get_t2(), b.get_t2())


@pytest.mark.parametrize("kind,args,tol,prec", [
    ("pairing", (8, 8, 0.5, 1.0), 1e-12, "fp64"),
    ("heg", (1.0, 19, 2), 1e-10, "fp64"),
    ("heg", (1.0, 19, 2), 1e-10, "mixed"),
])
def test_iterating_on_after_a_solve_continues_from_its_amplitudes(require_gpu, kind, args, tol, prec):
    """The last iteration of a device loop stops the loop while its T remap
    and the psi part of its reduction may still be running; the next
    iterations must start from the solve's amplitudes exactly as a fresh
    context given them does (T^ph, chi and psi rebuilt from them)."""
    from coupled_cluster_doubles_amd import FP32
    m = _pair(kind, args)[0]
    p = FP32 if prec == "mixed" else FP64
    for free in (False, True):
        a = Context(m, p, matrix_free=free)
        if prec == "mixed":
            a.solve_mixed(tol, 1e-6)     # leaves the context in FP32, its own setting
        else:
            a.solve(tol)
        t = a.get_t2()
        ea = a.iterate_n(2)
        b = Context(m, p, matrix_free=free)
        b.set_t2(t)
        assert np.array_equal(ea, b.iterate_n(2))
        assert np.array_equal(a.get_t2(), b.get_t2())
