"""Every iteration schedule gives the same results bit for bit: four streams
or one, launched directly, or one stream replayed from a captured HIP graph
of four iterations.  Each graph is replayed many times inside one solve, and the
solves alternate with direct (eager) work on the same context -- per-diagram
entry points, amplitude downloads, iterate_n calls that mix graph replays
with single launches -- and every result is compared with the four-stream
direct schedule's."""
import numpy as np
import pytest

from coupled_cluster_doubles_amd import FP32, FP64, Batch, Context, Model

pytestmark = pytest.mark.gpu

SCHEDS = ("streams", "serial", "serial_graph")


def _run(ctx, m, tol, mixed):
    out = {}
    if mixed:
        e, stop, n, n32 = ctx.solve_mixed(tol, 1e-6)
        out["n32"] = n32
    else:
        e, stop, n = ctx.solve(tol)
    out.update(e=e, stop=stop, n=n, t=ctx.get_t2())
    # eager work between graph replays, then iterate_n across batch borders
    out["ladder"] = ctx.ladder(np.random.default_rng(3).uniform(-0.05, 0.05, m.n_t2))
    out["it6"] = ctx.iterate_n(6)
    out["it4"] = ctx.iterate_n(4)
    out["t_end"] = ctx.get_t2()
    # and a second solve from T2 = 0, replaying the same graphs again
    out["e2"] = (ctx.solve_mixed(tol, 1e-6) if mixed else ctx.solve(tol))[0]
    return out


@pytest.mark.parametrize("kind,args,prec,free,tol", [
    ("pairing", (8, 8, 0.5, 1.0), "fp64", False, 1e-12),
    ("heg", (1.0, 19, 2), "fp64", True, 1e-10),
    ("heg", (1.0, 19, 2), "mixed", True, 1e-10),
    ("heg", (1.0, 10, 3), "mixed", False, 1e-10),
])
def test_every_schedule_gives_the_same_results(require_gpu, kind, args, prec, free, tol):
    m = Model.pairing(*args) if kind == "pairing" else Model.electron_gas(*args)
    p = FP32 if prec == "mixed" else FP64
    ref = None
    for sched in SCHEDS:
        print(f"{kind}{args} {prec} free={free} schedule={sched}", flush=True)
        ctx = Context(m, p, matrix_free=free)
        ctx.set_schedule(sched)
        assert ctx.schedule()[0] == sched
        got = _run(ctx, m, tol, prec == "mixed")
        ctx.close()
        if ref is None:
            ref = got
            continue
        for k, v in ref.items():
            if isinstance(v, np.ndarray):
                assert np.array_equal(got[k], v), (sched, k)
            else:
                assert got[k] == v, (sched, k)


def test_tuner_picks_a_schedule_and_batches_follow_it(require_gpu):
    """The tuner times every schedule and keeps the fastest (amplitudes back
    at zero); a batch of solves gives the same per-solve results under every
    schedule."""
    ctx = Context(Model.pairing(8, 8, 0.5, 1.0), FP64)
    assert ctx.schedule()[0] == "streams"
    ctx.set_schedule("auto")
    assert not ctx.get_t2().any()
    name, us = ctx.schedule()
    assert name in SCHEDS and all(v > 0 for v in us.values())
    assert us[name] == min(us.values())
    models = [Model.pairing(8, 8, 0.5, 1.0), Model.electron_gas(1.0, 6, 2), Model.pairing(6, 6, 0.3, 1.0)]
    ref = None
    for sched in SCHEDS:
        b = Batch(models, FP32)
        b.set_schedule(sched)
        res = b.solve_mixed(1e-10, 1e-6)
        t = b.get_t2()
        b.close()
        if ref is None:
            ref = (res, t)
            continue
        for (e, s, n, n32), (e0, s0, n0, n320) in zip(res, ref[0]):
            assert np.array_equal(e, e0) and (s, n, n32) == (s0, n0, n320), sched
        assert np.array_equal(t, ref[1]), sched
