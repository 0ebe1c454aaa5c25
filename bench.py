"""Benchmark: CCD iterations/s on the electron gas, N=54, r_s=1.0, n^2<=36.

    python bench.py [--gpus 1] [--steps K] [--warmup W]
                    [--precision mixed|fp64|fp32] [--mode free|stored]
                    [--no-cpu-baseline] [--config heg54|heg66|heg14|pairing|sweep64]

A step is one full Jacobi iteration of the CCD amplitude equations over the
whole problem (both ladders, ring, one-body terms, remaps, update, energy),
with the amplitudes and interaction resident in HBM.

--precision mixed (default): the steps are FP32-contraction iterations (the
bulk of a mixed solve); afterwards a full mixed solve (FP32 iterations while
|dE| > --switch-tol, then FP64 sweeps to |dE| <= 1e-10) and a full FP64 solve
are timed from T2 = 0 and their energies compared per particle; the mixed
solve is also timed with the switch at 1e-6 and with no FP64 sweeps
(switch_sweep).  fp64 / fp32: steps in that arithmetic only.

--mode free (the default for the electron gas) is the reference's
matrix-free mode (-m 1): V^pppp is rebuilt inside the pppp ladder every
iteration instead of being read from HBM; the iterates are bit-identical to
--mode stored (the default for the pairing model and the sweep).

--config sweep64: the seeded sweep of 64 independent solves (electron gas
N=14/38/54 and pairing models, coupled_cluster_doubles_amd/sweep.py) as one
batch; a step advances every solve by one iteration.  Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

CONFIGS = {
    # reference CLI: --electronGas -r 1.0 -s 32 -p 4 -t 1e-10
    "heg54": dict(kind="heg", args=(1.0, 32, 4),
                  name="electron gas N=54, r_s=1.0, n^2<=36 (1850 s.p. states), CCD iteration"),
    # reference CLI: --electronGas -r 0.5 -s 46 -p 5
    "heg66": dict(kind="heg", args=(0.5, 46, 5),
                  name="electron gas N=66, r_s=0.5, n^2<=52 (3150 s.p. states), CCD iteration"),
    # reference CLI: --electronGas -r 1.0 -s 19 -p 2 -t 1e-10
    "heg14": dict(kind="heg", args=(1.0, 19, 2),
                  name="electron gas N=14, r_s=1.0, n^2<=20 (778 s.p. states), CCD iteration"),
    # reference CLI: --pairing -l 8 -n 8 -g 0.5 -x 1.0 -t 1e-12
    "pairing": dict(kind="pairing", args=(8, 8, 0.5, 1.0),
                    name="pairing model, 8 levels, 8 particles, g=0.5, CCD iteration"),
    "sweep64": dict(kind="sweep", args=(),
                    name="64 seeded solves in one batch (electron gas N=14/38/54, r_s in [0.5,5]; "
                         "pairing 8-16 levels, g in [0.2,0.8]), one iteration of every solve"),
}
HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6
FP32_PEAK_TFLOPS = 157.3


def make_model(cfg):
    from coupled_cluster_doubles_amd import Model
    if cfg["kind"] == "heg":
        return Model.electron_gas(*cfg["args"])
    return Model.pairing(*cfg["args"])


def make_context(cfg, prec, free, models=None):
    from coupled_cluster_doubles_amd import Batch, Context
    if cfg["kind"] == "sweep":
        return Batch(models, prec, matrix_free=free)
    return Context(models[0], prec, matrix_free=free)


def packed_ladder_flops(models):
    """MFMA-side flops of one pass of the packed pppp ladder: per direct
    channel 2 npk^2 hh (V t over pairs c<d) + 2 npk hh^2 (T X), npk = pairs a<b."""
    f = 0.0
    for model in models:
        d = model.dims()
        pairs = model.pairs("pp")
        o = 0
        for pp, hh in zip(d["pp_dim"], d["hh_dim"]):
            blk = pairs[o:o + pp]
            o += pp
            if pp == 0 or hh == 0:
                continue
            npk = int(np.count_nonzero(blk[:, 0] < blk[:, 1]))
            f += 2.0 * npk * npk * hh + 2.0 * npk * hh * hh
    return f


def cpu_baseline(kind, args, iters, threads, matrix_free=False):
    """The FP64 CPU oracle (plain C + OpenMP) timed on this host: iterations/s."""
    from oracle.pyoracle import Oracle, set_threads
    set_threads(threads)
    if kind == "heg":
        o = Oracle.electron_gas(*args, matrix_free=matrix_free)
    else:
        o = Oracle.pairing(*args)
    t = np.zeros(o.nt2)
    t, _ = o.step(t)                      # first iteration untimed (page-in)
    t0 = time.perf_counter()
    for _ in range(iters):
        t, e = o.step(t)
    dt = time.perf_counter() - t0
    o.close()
    return iters / dt, e


def cpu_baseline_sweep(iters, threads):
    """Seconds per sweep iteration on the oracle, from one representative solve
    per basis (an iteration's cost depends on the basis, not on r_s or g),
    weighted by the basis' count in the sweep."""
    from coupled_cluster_doubles_amd.sweep import sweep_configs
    count = {}
    for kind, args in sweep_configs():
        key = (kind, args[1:] if kind == "heg" else args[:2])
        count.setdefault(key, [0, (kind, args)])[0] += 1
    sec = 0.0
    for key, (n, (kind, args)) in sorted(count.items()):
        rate, _ = cpu_baseline(kind, args, iters, threads)
        sec += n / rate
    return 1.0 / sec


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=30)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--config", default="heg54", choices=sorted(CONFIGS))
    ap.add_argument("--precision", default="mixed", choices=["mixed", "fp64", "fp32"])
    ap.add_argument("--switch-tol", type=float, default=1e-8,
                    help="mixed solve: |dE| below which FP64 sweeps replace FP32 iterations")
    ap.add_argument("--mode", default=None, choices=["stored", "free"],
                    help="stored: V^pppp read from HBM; free: rebuilt in the ladder (reference -m 1); "
                         "default free for the electron gas, stored otherwise")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-iters", type=int, default=3)
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="first copy this many MiB device-to-device (a kernel of known traffic, "
                         "to calibrate hardware byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build runs the CCD iteration on a single "
              "MI355X (multi-GPU channel sharding is a design note, not built).", file=sys.stderr)
        sys.exit(2)

    from coupled_cluster_doubles_amd import FP32, FP64, Context

    if args.copy_probe_mib > 0:
        import torch
        n = args.copy_probe_mib * (1 << 20) // 4
        src = torch.ones(n, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    cfg = CONFIGS[args.config]
    prec = FP64 if args.precision == "fp64" else FP32
    if args.mode is None:
        args.mode = "free" if cfg["kind"] == "heg" else "stored"
    free = args.mode == "free"
    sweep = cfg["kind"] == "sweep"
    if sweep:
        from coupled_cluster_doubles_amd.sweep import sweep_models
        models = sweep_models()
    else:
        models = [make_model(cfg)]
    t_setup = time.perf_counter()
    ctx = make_context(cfg, prec, free, models)
    setup_s = time.perf_counter() - t_setup
    mbpt2 = ctx.mbpt2()
    work = ctx.work()

    # the steps are queued back to back (no host round trip between them);
    # the call returns after the last one has finished on the device
    ctx.iterate_n(args.warmup)
    t0 = time.perf_counter()
    e = ctx.iterate_n(args.steps)
    dt = time.perf_counter() - t0
    e = e[:, -1].tolist() if sweep else float(e[-1])
    ms_step = 1e3 * dt / args.steps
    iters_per_s = args.steps / dt

    # per-phase device time from HIP events on the context's stream
    phases = ctx.time_phases(max(3, min(args.steps, 10)))
    ladder_ms = float(phases[2])
    lb = work["ladder_bytes"]
    gemm_tflops = work["flops"] / (ms_step * 1e-3) / 1e12
    if args.mode == "stored":
        achieved = lb / (ladder_ms * 1e-3) / 1e9
        roof = {"bound": "hbm", "kernel": "k_ladder_v", "achieved": achieved, "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "algorithmic_bytes": lb}
    else:
        lf = packed_ladder_flops(models)
        peak = FP64_PEAK_TFLOPS if prec == FP64 else FP32_PEAK_TFLOPS
        achieved = lf / (ladder_ms * 1e-3) / 1e12
        roof = {"bound": "mfma", "kernel": "k_ladder_v", "achieved": achieved, "peak": peak,
                "unit": "TFLOP/s", "frac": achieved / peak, "algorithmic_flops": lf,
                "algorithmic_bytes": lb}

    traffic = None
    tfile = os.path.join(ROOT, "profiles",
                         f"ladder_d_traffic_{args.config}_{args.precision}_{args.mode}.json")
    if os.path.exists(tfile):
        with open(tfile) as f:
            traffic = json.load(f).get("bytes_per_launch")

    mixed = None
    if args.precision == "mixed" and not sweep:
        # full solves from T2 = 0, timed on the host clock around the call
        model = models[0]
        n_part = model.n_holes
        t0 = time.perf_counter()
        e_m, stop_m, n_m, n32 = ctx.solve_mixed(1e-10, args.switch_tol)
        t_m = time.perf_counter() - t0
        ctx64 = Context(model, FP64, matrix_free=free)
        ctx64.solve(1e-10)                       # warm
        t0 = time.perf_counter()
        e_d, stop_d, n_d = ctx64.solve(1e-10)
        t_d = time.perf_counter() - t0
        mixed = {"tol": 1e-10, "switch_tol": args.switch_tol,
                 "mixed_iters": n_m, "mixed_fp32_iters": n32, "mixed_stop": stop_m,
                 "mixed_solve_s": t_m, "mixed_energy": float(e_m[-1]),
                 "fp64_iters": n_d, "fp64_stop": stop_d, "fp64_solve_s": t_d,
                 "fp64_energy": float(e_d[-1]),
                 "abs_dE_per_particle": abs(float(e_m[-1]) - float(e_d[-1])) / n_part,
                 "speedup_vs_fp64_solve": t_d / t_m}
        # the same solve with other switch points (1e-10 = tol: FP32 iterations only)
        sw_runs = []
        for sw in (1e-6, 1e-8, 1e-10):
            ctx.solve_mixed(1e-10, sw)                 # warm
            t0 = time.perf_counter()
            e_s, stop_s, n_s, n32_s = ctx.solve_mixed(1e-10, sw)
            t_s = time.perf_counter() - t0
            sw_runs.append({"switch_tol": sw, "iters": n_s, "fp32_iters": n32_s, "stop": stop_s,
                          "solve_s": t_s, "speedup_vs_fp64_solve": t_d / t_s,
                          "abs_dE_per_particle": abs(float(e_s[-1]) - float(e_d[-1])) / n_part})
        mixed["switch_sweep"] = sw_runs
        ctx64.close()
    elif args.precision == "mixed":
        # every solve of the sweep from T2 = 0, mixed and FP64, one batch each
        t0 = time.perf_counter()
        rm = ctx.solve_mixed(1e-10, args.switch_tol)
        t_m = time.perf_counter() - t0
        ctx.close()
        ctx64 = make_context(cfg, FP64, free, models)
        t0 = time.perf_counter()
        rd = ctx64.solve(1e-10)
        t_d = time.perf_counter() - t0
        conv = [q for q in range(len(models)) if rd[q][1] == "converged"]
        mixed = {"tol": 1e-10, "switch_tol": args.switch_tol,
                 "fp64_stops": {k: sum(r[1] == k for r in rd) for k in ("converged", "max_iterations", "nan")},
                 "mixed_converged_where_fp64_did": all(rm[q][1] == "converged" for q in conv),
                 "max_abs_dE_per_particle": max(abs(float(rm[q][0][-1]) - float(rd[q][0][-1]))
                                                / models[q].n_holes for q in conv),
                 "batch_iters_mixed": max(r[2] for r in rm), "batch_iters_fp64": max(r[2] for r in rd),
                 "mixed_solve_s": t_m, "fp64_solve_s": t_d, "speedup_vs_fp64_solve": t_d / t_m}
        ctx64.close()

    cpu = None
    if not args.no_cpu_baseline:
        threads = min(os.cpu_count() or 1, 16)
        # V^pppp in the reference's layout is 54 GB at N=66: the oracle then
        # rebuilds it per channel (the reference's -m 1), else it reads it
        cpu_free = args.config == "heg66"
        if sweep:
            v = cpu_baseline_sweep(args.cpu_iters, threads)
            sample = (f"{args.cpu_iters} FP64 iterations by the C oracle of one solve per basis "
                      f"of the sweep (8 bases), weighted by the basis counts")
        else:
            v, _ = cpu_baseline(cfg["kind"], cfg["args"], args.cpu_iters, threads, cpu_free)
            sample = (f"{args.cpu_iters} FP64 iterations of the same workload by the C oracle "
                      f"(OpenMP over channels, naive dgemm, V^pppp "
                      f"{'rebuilt per channel' if cpu_free else 'stored'}), after one untimed iteration")
        cpu = {"value": v, "unit": "iter/s", "cores": threads, "kind": "port", "sample": sample}

    out = {
        "metric": "CCD iter/s + channel-GEMM TFLOP/s, HEG N=54; mixed vs FP64 at |dE|/N<=1e-7",
        "value": iters_per_s,
        "unit": "iter/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if prec == FP64 else "f32+f64",
        "mixed_vs_fp64": mixed,
        "data": "synthetic: the reference's closed-form basis and Coulomb interaction, T2 from zero",
        "config": {"workload": cfg["name"], "precision": args.precision, "mode": args.mode,
                   "n_t2": ctx.n_t2, "device_bytes": ctx.device_bytes},
        "gemm_tflops": gemm_tflops,
        "gemm_flops_per_iter": work["flops"],
        "mbpt2": mbpt2.tolist() if sweep else mbpt2,
        "corr_after_steps": e,
        "phase_ms": {"ring": phases[0], "ladder_x": phases[1], "ladder_d": phases[2],
                     "update": phases[3], "reductions": phases[4], "iteration": phases[5]},
        "setup_s": setup_s,
        "roofline": dict(roof, traffic=traffic),
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
