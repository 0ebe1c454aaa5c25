"""The pppp-ladder plan (ccd_get/set_ladder_plan, ccd_tune_ladder).

Blocks per compute unit, the tile assignment and the launch partition only
move work between blocks and launches (every tile writes its own outputs),
so they must give the same iterates bit for bit.  The kernel kinds change
the summation (16x16x4 MFMA instead of 4x4x1 for class 1) or the arithmetic
(BF16x3 split products, about 2^-16 relative per product): they are held to
the oracle -- FP32 kinds within 2e-5 of the largest element of the ladder,
BF16x3 within 2e-4 -- and a mixed solve with them to 1e-7 Ha per particle
of the oracle's converged energy.  The symmetric class-1 kind (bf16x3sym)
builds each element of V^pppp once for both of its rows, and bf16x3w is its
kernel without the symmetry: both held to the oracle like BF16x3, also
through whole iterations.  Stored and generated V stay bit-identical under
every kind.
"""
import numpy as np
import pytest

from coupled_cluster_doubles_amd import (FP32, FP64, TUNE_BF16X3, TUNE_KINDS, TUNE_RING_BF16X3, CCDError, Context,
                                         Model)
from oracle.pyoracle import Oracle

pytestmark = pytest.mark.gpu

F32_REL = 2e-5
BF16X3_REL = 2e-4


def _model(kind, args):
    return Model.pairing(*args) if kind == "pairing" else Model.electron_gas(*args)


def _oracle(kind, args):
    return Oracle.pairing(*args) if kind == "pairing" else Oracle.electron_gas(*args)


def _plan(ctx, **kw):
    p = ctx.ladder_plan()
    for k, v in kw.items():
        p[k] = v if isinstance(v, str) else list(v)
    ctx.set_ladder_plan(p)
    return p


def _antisym(m, t):
    lab = m.t2_labels()
    key = {tuple(r): k for k, r in enumerate(lab.tolist())}
    sw_ab = np.array([key[(b, a, i, j)] for a, b, i, j in lab.tolist()])
    sw_ij = np.array([key[(a, b, j, i)] for a, b, i, j in lab.tolist()])
    ta = t - t[sw_ab]
    return ta - ta[sw_ij]


def test_default_plan_and_validation(require_gpu):
    m = Model.electron_gas(1.0, 19, 2)
    free = Context(m, FP32, matrix_free=True).ladder_plan()
    assert free["kind_fp32"] == ["default"] * 8 and free["kind_fp64"] == ["default"] * 8
    assert free["blocks_per_cu"] == [3] * 8 and free["assign"] == ["model"] * 8
    assert free["partition"] == "split"
    ctx = Context(m, FP64)
    assert ctx.ladder_plan()["blocks_per_cu"] == [0] * 8
    p = ctx.ladder_plan()
    for bad in ({"kind_fp64": ["bf16x3"] * 8}, {"blocks_per_cu": [17] * 8}, {"blocks_per_cu": [-1] * 8},
                {"kind_fp64": ["bf16x3sym"] + ["default"] * 7},
                {"kind_fp32": ["bf16x3", "bf16x3sym"] + ["bf16x3"] * 6},
                {"kind_fp64": ["bf16x3w"] + ["default"] * 7},
                {"kind_fp32": ["bf16x3", "bf16x3w"] + ["bf16x3"] * 6}):
        q = dict(p, **bad)
        with pytest.raises(CCDError):
            ctx.set_ladder_plan(q)
    assert ctx.ladder_plan() == p   # a rejected plan leaves the old one


@pytest.mark.parametrize("kind,args,prec,free", [
    ("heg", (1.0, 19, 2), FP32, True),
    ("heg", (1.0, 19, 2), FP64, True),
    ("heg", (1.0, 10, 3), FP32, False),
    ("pairing", (16, 16, 0.5, 1.0), FP64, False),   # hh up to 128: classes 5-8
    ("pairing", (16, 16, 0.5, 1.0), FP32, False),
    ("heg", (1.0, 19, 2), "sym", True),
    ("heg", (1.0, 10, 3), "sym", False),
    ("heg", (1.0, 19, 2), "w", True),
])
def test_scheduling_fields_change_no_result(require_gpu, kind, args, prec, free):
    m = _model(kind, args)
    k1 = {"sym": "bf16x3sym", "w": "bf16x3w"}.get(prec)   # FP32 iterations, class 1 on a wave-k kernel
    if k1:
        prec = FP32
    plans = [
        {},
        {"blocks_per_cu": [1] * 8, "assign": ["channel"] * 8},
        {"blocks_per_cu": [6] * 8, "assign": ["longest"] * 8, "partition": "one"},
        {"blocks_per_cu": [0, 2, 4, 0, 1, 3, 0, 2], "assign": ["longest", "channel"] * 4, "partition": "one"},
        {"blocks_per_cu": [2, 3, 1, 6, 4, 2, 3, 1], "assign": ["xcd", "runs", "xcd", "model"] * 2},
    ]
    ref = None
    for q, kw in enumerate(plans):
        print(f"{kind}{args} prec={prec} free={free} plan {q}: {kw}", flush=True)
        ctx = Context(m, prec, matrix_free=free)
        _plan(ctx, **kw, **({"kind_fp32": [k1] + ["bf16x3"] * 7} if k1 else {}))
        got = (ctx.iterate_n(4), ctx.get_t2())
        ctx.close()
        if ref is None:
            ref = got
            continue
        assert np.array_equal(got[0], ref[0]) and np.array_equal(got[1], ref[1]), q


@pytest.mark.parametrize("kind,args", [
    ("heg", (1.0, 19, 2)),
    ("heg", (1.0, 10, 3)),
    ("pairing", (16, 16, 0.5, 1.0)),
])
def test_kernel_kinds_close_to_oracle(require_gpu, kind, args):
    m, o = _model(kind, args), _oracle(kind, args)
    ta = _antisym(m, np.random.default_rng(11).uniform(-0.05, 0.05, m.n_t2))
    d_o = o.ladder(ta)
    scale = np.max(np.abs(d_o))
    for kinds, tol in ((["mfma16"] * 8, F32_REL), (["bf16x3"] * 8, BF16X3_REL),
                       (["default", "bf16x3", "mfma16", "bf16x3", "default", "bf16x3", "mfma16", "bf16x3"], BF16X3_REL),
                       (["bf16x3sym"] + ["default"] * 7, BF16X3_REL), (["bf16x3w"] + ["default"] * 7, BF16X3_REL)):
        ctx = Context(m, FP32)
        _plan(ctx, kind_fp32=kinds)
        err = np.max(np.abs(ctx.ladder(ta) - d_o)) / scale
        print(f"{kind}{args} kinds={kinds}: max |D - D_oracle| / max|D_oracle| = {err:.2e}", flush=True)
        assert err <= tol, (kinds, err)
        # the FP64 iterations of the same context keep the FP64 kernel
        ctx.set_precision(FP64)
        assert np.array_equal(ctx.ladder(ta), Context(m, FP64).ladder(ta))


def _antisym_fast(m, t):
    """(1 - P_ab)(1 - P_ij) t over the pphh labels (vectorised _antisym)."""
    lab = m.t2_labels().astype(np.int64)
    n = int(lab.max()) + 1
    key = lambda a, b, i, j: ((a * n + b) * n + i) * n + j
    k0 = key(lab[:, 0], lab[:, 1], lab[:, 2], lab[:, 3])
    order = np.argsort(k0)
    def where(k):
        pos = order[np.searchsorted(k0, k, sorter=order)]
        assert np.array_equal(k0[pos], k)
        return pos
    sw_ab = where(key(lab[:, 1], lab[:, 0], lab[:, 2], lab[:, 3]))
    sw_ij = where(key(lab[:, 0], lab[:, 1], lab[:, 3], lab[:, 2]))
    ta = t - t[sw_ab]
    return ta - ta[sw_ij]


def test_ladder_kinds_at_n54_against_oracle(require_gpu):
    """The production workload (electron gas N=54, matrix-free V): the pppp
    ladder of every FP32 kernel kind against the oracle's, within the kind's
    tolerance relative to the largest element, and the FP64 ladder within
    1e-12.  Classes 1-4 of the plan (one to four 16-column tiles, hh up to
    54) occur here."""
    args = (1.0, 32, 4)
    m, o = Model.electron_gas(*args), Oracle.electron_gas(*args, matrix_free=True)
    ta = _antisym_fast(m, np.random.default_rng(54).uniform(-0.05, 0.05, m.n_t2))
    d_o = o.ladder(ta)
    o.close()
    scale = np.max(np.abs(d_o))
    ctx = Context(m, FP64, matrix_free=True)
    err = np.max(np.abs(ctx.ladder(ta) - d_o)) / scale
    print(f"N=54 FP64: max |D - D_oracle| / max|D_oracle| = {err:.2e}", flush=True)
    assert err <= 1e-12
    ctx.close()
    for kinds, tol in ((["default"] * 8, F32_REL), (["mfma16"] * 8, F32_REL), (["bf16x3"] * 8, BF16X3_REL),
                       (["bf16x3sym"] + ["bf16x3"] * 7, BF16X3_REL), (["bf16x3w"] + ["bf16x3"] * 7, BF16X3_REL)):
        ctx = Context(m, FP32, matrix_free=True)
        _plan(ctx, kind_fp32=kinds)
        err = np.max(np.abs(ctx.ladder(ta) - d_o)) / scale
        print(f"N=54 kinds={kinds[0]}: max |D - D_oracle| / max|D_oracle| = {err:.2e}", flush=True)
        assert err <= tol, (kinds, err)
        ctx.close()


@pytest.mark.parametrize("kinds", [["bf16x3"] * 8, ["mfma16"] * 8, ["bf16x3sym"] + ["bf16x3"] * 7,
                                   ["bf16x3w"] + ["bf16x3"] * 7])
def test_stored_and_generated_v_agree_bitwise_under_every_kind(require_gpu, kinds):
    m = Model.electron_gas(1.0, 19, 2)
    a, b = Context(m, FP32), Context(m, FP32, matrix_free=True)
    _plan(a, kind_fp32=kinds)
    _plan(b, kind_fp32=kinds)
    assert np.array_equal(a.iterate_n(3), b.iterate_n(3))
    assert np.array_equal(a.get_t2(), b.get_t2())


@pytest.mark.parametrize("kinds,ring", [(["bf16x3"] * 8, "fp32"), (["bf16x3sym"] + ["bf16x3"] * 7, "fp32"),
                                        (["bf16x3w"] + ["bf16x3"] * 7, "fp32"), (["bf16x3"] * 8, "bf16x3"),
                                        (["bf16x3w"] + ["bf16x3"] * 7, "bf16x3")])
def test_mixed_solve_with_bf16x3_products_within_1e7_per_particle(require_gpu, kinds, ring):
    """FP32 iterations on BF16x3 products (the ladder, with or without the
    ring), then FP64 sweeps: the converged energy within 1e-7 Ha per particle
    of the oracle's (N=14 electrons)."""
    args = (1.0, 19, 2)
    m, o = Model.electron_gas(*args), Oracle.electron_gas(*args)
    e_o, stop_o, _ = o.solve(1e-10)
    ctx = Context(m, FP32, matrix_free=True)
    _plan(ctx, kind_fp32=kinds, ring_fp32=ring)
    e, stop, n, n32 = ctx.solve_mixed(1e-10, 1e-6)
    print(f"{kinds[0]} ring {ring} mixed: {n32} + {n - n32} iterations, |dE|/N = {abs(e[-1] - e_o[-1]) / 14:.2e}",
          flush=True)
    assert stop == stop_o == "converged" and n32 > 0
    assert abs(e[-1] - e_o[-1]) / 14 <= 1e-7


def test_symmetric_kernel_at_n66_matches_bf16x3(require_gpu):
    """Electron gas N=66 (matrix-free, r_s=0.5, 3,150 states): class-1
    channels of up to 1,154 packed pairs, so of two k segments (the symmetric
    kernel's tiles of the second segment only clear their slab).  Its ladder
    against the BF16x3 kernel's, both within BF16X3_REL of the exact one
    (held to the oracle at N=54), so within twice that of each other; and
    three whole FP32 iterations on each within 1e-6 of each other (relative
    energy)."""
    m = Model.electron_gas(0.5, 46, 5)
    ta = _antisym_fast(m, np.random.default_rng(66).uniform(-0.05, 0.05, m.n_t2))
    got = {}
    for kinds in (["bf16x3"] * 8, ["bf16x3sym"] + ["bf16x3"] * 7):
        ctx = Context(m, FP32, matrix_free=True)
        _plan(ctx, kind_fp32=kinds)
        got[kinds[0]] = (ctx.ladder(ta), ctx.iterate_n(3))
        ctx.close()
    (d0, e0), (d1, e1) = got["bf16x3"], got["bf16x3sym"]
    err = np.max(np.abs(d1 - d0)) / np.max(np.abs(d0))
    print(f"N=66 ladder bf16x3sym vs bf16x3: {err:.2e}; energies {e0} {e1}", flush=True)
    assert err <= 2 * BF16X3_REL
    assert np.max(np.abs(e1 - e0) / np.abs(e0)) <= 1e-6


def test_tuner_keeps_results_or_tolerance(require_gpu):
    """Without kind flags the tuner only moves work: the tuned context's
    solve equals an untuned one bit for bit.  With them it may pick other
    kernels: the mixed solve then stays within 1e-7 Ha per particle.  Either
    way it leaves T2 = 0 and a valid plan whose chosen timings are recorded."""
    m = Model.electron_gas(1.0, 19, 2)
    ref = Context(m, FP32, matrix_free=True).solve_mixed(1e-10, 1e-6)
    ctx = Context(m, FP32, matrix_free=True)
    p = ctx.tune_ladder(0)
    assert not ctx.get_t2().any()
    assert p["kind_fp32"] == ["default"] * 8
    assert p["us_fp32"][0] > 0 and all(v > 0 for v in p["us_partition"].values())
    got = ctx.solve_mixed(1e-10, 1e-6)
    assert np.array_equal(got[0], ref[0]) and got[1:] == ref[1:]
    p = ctx.tune_ladder(TUNE_KINDS | TUNE_BF16X3 | TUNE_RING_BF16X3)
    print(f"tuned plan: {p}", flush=True)
    assert p["kind_fp64"] == ["default"] * 8
    e, stop, n, n32 = ctx.solve_mixed(1e-10, 1e-6)
    assert stop == "converged"
    assert abs(e[-1] - ref[0][-1]) / 14 <= 1e-7
