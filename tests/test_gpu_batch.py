"""Batches of independent solves in one context (ccd_batch_*).

Each solve of a batch must give exactly what a context of its own gives:
the same MBPT2, the same energy in every iteration, the same stop reason
and iteration count, and the same final amplitudes, bit for bit -- also
when the solves stop at different iterations (a stopped solve keeps its
amplitudes while the others go on).  Single contexts are held to the
oracle in test_gpu_parity.py; here a few solves are also checked against
the oracle directly.  The 64-solve seeded sweep (electron gas N = 14/38/54
and pairing models) runs in full.
"""
import numpy as np
import pytest

from coupled_cluster_doubles_amd import FP32, FP64, Batch, Context, Model
from coupled_cluster_doubles_amd.sweep import sweep_configs, sweep_models
from oracle.pyoracle import Oracle

pytestmark = pytest.mark.gpu


def _small():
    cfg = [("pairing", (8, 8, 0.5, 1.0)), ("heg", (1.0, 19, 2)), ("pairing", (12, 12, 0.3, 1.0)),
           ("heg", (3.0, 4, 2)), ("heg", (0.5, 19, 2)), ("infmatter", (0.08, 4, 2))]
    make = {"pairing": Model.pairing, "heg": Model.electron_gas, "infmatter": Model.inf_matter}
    return cfg, [make[k](*a) for k, a in cfg]


def _same_as_single(models, results, batch, mbpt2=None, tol=1e-10):
    for q, m in enumerate(models):
        c = Context(m, FP64)
        if mbpt2 is not None:
            assert mbpt2[q] == c.mbpt2(), q
        e, stop, n = c.solve(tol)
        eb, stop_b, n_b = results[q]
        assert (stop_b, n_b) == (stop, n), q
        assert np.array_equal(eb, e, equal_nan=True), q     # a diverging solve ends in NaN
        assert np.array_equal(batch.t2(q), c.get_t2(), equal_nan=True), q
        c.close()


def test_batch_solves_match_single_contexts_bitwise(require_gpu):
    cfg, models = _small()
    b = Batch(models, FP64)
    assert b.n_problems == len(models)
    assert b.n_t2 == sum(m.n_t2 for m in models)
    mb = b.mbpt2()
    res = b.solve(1e-10)
    counts = {r[2] for r in res}
    assert len(counts) > 1            # the solves stop at different iterations
    _same_as_single(models, res, b, mb)
    # a second solve from T2 = 0 repeats the first
    res2 = b.solve(1e-10)
    for r, r2 in zip(res, res2):
        assert r[1:] == r2[1:] and np.array_equal(r[0], r2[0])


def test_batch_against_oracle(require_gpu):
    cfg, models = _small()
    b = Batch(models, FP64)
    res = b.solve(1e-10)
    for q, (kind, args) in enumerate(cfg[:3]):
        o = Oracle.pairing(*args) if kind == "pairing" else Oracle.electron_gas(*args)
        e_o, stop_o, _ = o.solve(1e-10)
        e_g, stop_g, n_g = res[q]
        assert stop_g == stop_o == "converged"
        assert n_g == len(e_o)
        assert np.max(np.abs(e_g - e_o) / np.abs(e_o)) < 1e-10
        o.close()


def test_batch_modes_and_iterate_n(require_gpu):
    _, models = _small()
    stored = Batch(models, FP64)
    free = Batch(models, FP64, matrix_free=True)
    es, ef = stored.iterate_n(4), free.iterate_n(4)
    assert es.shape == (len(models), 4)
    assert np.array_equal(es, ef)
    assert np.array_equal(stored.get_t2(), free.get_t2())
    for q, m in enumerate(models):
        assert np.array_equal(Context(m, FP64).iterate_n(4), es[q]), q


def test_batch_rejects_single_solve_calls(require_gpu):
    from coupled_cluster_doubles_amd import CCDError
    _, models = _small()
    b = Batch(models[:2], FP64)
    with pytest.raises(CCDError, match="batch"):
        b.iterate()
    with pytest.raises(CCDError, match="batch"):
        Context.solve(b, 1e-10)


def test_batch_mixed_precision_within_1e7_per_particle(require_gpu):
    """A solve that finishes its FP32 phase early waits, frozen, for the
    others; its iterates are still exactly those of a context of its own."""
    _, models = _small()
    e64 = [r[0][-1] for r in Batch(models, FP64).solve(1e-10)]
    res = Batch(models, FP32).solve_mixed(1e-10, 1e-6)
    for q, m in enumerate(models):
        e, stop, n, n32 = res[q]
        assert stop == "converged" and 0 < n32 <= n
        assert abs(e[-1] - e64[q]) / m.n_holes <= 1e-7
        e1, stop1, n1, n321 = Context(m, FP32).solve_mixed(1e-10, 1e-6)
        assert (stop1, n1, n321) == (stop, n, n32) and np.array_equal(e1, e), q


def test_batch_matrix_free_mixed_matches_single_contexts_bitwise(require_gpu):
    """Matrix-free mixed batch: the electron-gas solves generate V from their
    own tables, the others through the generic generator, and the channels
    of solves that have stopped are skipped; every solve still equals a
    matrix-free context of its own bit for bit."""
    _, models = _small()
    b = Batch(models, FP32, matrix_free=True)
    res = b.solve_mixed(1e-10, 1e-6)
    assert len({r[2] for r in res}) > 1          # the solves stop at different iterations
    for q, m in enumerate(models):
        c = Context(m, FP32, matrix_free=True)
        e, stop, n, n32 = c.solve_mixed(1e-10, 1e-6)
        eb, stop_b, n_b, n32_b = res[q]
        assert (stop_b, n_b, n32_b) == (stop, n, n32), q
        assert np.array_equal(eb, e), q
        assert np.array_equal(b.t2(q), c.get_t2()), q
        c.close()
    # and the FP64 iterations of the batch continue like the contexts'
    en = b.iterate_n(3)
    for q, m in enumerate(models):
        c = Context(m, FP32, matrix_free=True)
        c.solve_mixed(1e-10, 1e-6)
        assert np.array_equal(en[q], c.iterate_n(3)), q
        c.close()


def test_sweep64_batch(require_gpu):
    """The seeded 64-solve sweep as one batch: FP64 results bit-identical to
    64 separate contexts, mixed precision within 1e-7 per particle with V
    stored and matrix-free."""
    cfg = sweep_configs()
    models = sweep_models()
    b = Batch(models, FP64)
    mb = b.mbpt2()
    res = b.solve(1e-10)
    _same_as_single(models, res, b, mb)
    e64 = [r[0][-1] for r in res]
    stops = [r[1] for r in res]
    b.close()
    print("sweep stop reasons:", {k: stops.count(k) for k in set(stops)})
    for free in (False, True):
        mixed = Batch(models, FP32, matrix_free=free).solve_mixed(1e-10, 1e-6)
        for q, m in enumerate(models):
            e, stop, n, n32 = mixed[q]
            if stops[q] == "converged":
                assert stop == "converged", (free, q, cfg[q])
                assert abs(e[-1] - e64[q]) / m.n_holes <= 1e-7, (free, q, cfg[q])
            else:                              # the FP64 solve diverges: so does the mixed one
                assert stop != "converged", (free, q, cfg[q])


def test_batch_iterate_n_keeps_going_past_a_diverging_solve(require_gpu):
    """Electron gas N=38 at r_s = 4.72 diverges under the reference's Jacobi
    iteration (NaN after 18 iterations); in a batch it stops there and the
    other solve is unaffected."""
    div = Model.electron_gas(4.722872261101834, 25, 3)
    ok = Model.pairing(10, 10, 0.45, 1.0)
    en = Batch([div, ok], FP64).iterate_n(25)
    e_div, stop, n = Context(div, FP64).solve(0.0, max_iters=30)
    assert stop == "nan" and n < 25
    assert np.array_equal(en[0, :n], e_div, equal_nan=True)
    assert np.all(np.isnan(en[0, n:]))
    assert np.array_equal(en[1], Context(ok, FP64).iterate_n(25))


@pytest.mark.parametrize("k1", ["bf16x3w", "bf16x3sym"])
def test_batch_wave_k_ladder_kinds_match_single_contexts_bitwise(require_gpu, k1):
    """The class-1 wave-k ladder kernels (bf16x3w, and the symmetric kind with
    its partial sums read by the update) in a matrix-free mixed batch: each
    electron-gas solve generates V from its own table, the others through
    the generic generator, stopped solves are skipped, and every solve equals
    a context of its own on the same kernel kinds bit for bit."""
    _, models = _small()
    kinds = [k1] + ["bf16x3"] * 7

    def pinned(ctx):
        p = ctx.ladder_plan()
        p["kind_fp32"] = kinds
        ctx.set_ladder_plan(p)
        return ctx

    b = pinned(Batch(models, FP32, matrix_free=True))
    res = b.solve_mixed(1e-10, 1e-6)
    assert len({r[2] for r in res}) > 1          # the solves stop at different iterations
    for q, m in enumerate(models):
        c = pinned(Context(m, FP32, matrix_free=True))
        e, stop, n, n32 = c.solve_mixed(1e-10, 1e-6)
        eb, stop_b, n_b, n32_b = res[q]
        print(f"{k1} solve {q}: {stop_b} after {n_b} ({n32_b} FP32) iterations", flush=True)
        assert (stop_b, n_b, n32_b) == (stop, n, n32), q
        assert np.array_equal(eb, e), q
        assert np.array_equal(b.t2(q), c.get_t2()), q
        c.close()
