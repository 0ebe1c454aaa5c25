"""Benchmark: CCD iterations/s on the electron gas, N=54, r_s=1.0, n^2<=36.

    python bench.py [--gpus 1] [--steps K] [--warmup W]
                    [--precision mixed|fp64|fp32] [--mode free|stored]
                    [--no-cpu-baseline] [--config heg54|heg66|heg14|pairing|sweep64]

A step is one full Jacobi iteration of the CCD amplitude equations over the
whole problem (both ladders, ring, one-body terms, remaps, update, energy),
with the amplitudes and interaction resident in HBM.

--precision mixed (default): the steps are FP32-contraction iterations (the
bulk of a mixed solve); afterwards a full mixed solve (FP32 iterations while
|dE| > --switch-tol, then FP64 sweeps to |dE| <= 1e-10) and full FP64 solves
(the same V mode, and the stored-V "plain" FP64 path) are timed from T2 = 0;
their energies are compared per particle with each other and, where the
oracle's solve of the workload is committed (tests/golden), with the
oracle's.  The mixed solve is also timed with the switch at 1e-6 and with no
FP64 sweeps (switch_sweep).  fp64 / fp32: steps in that arithmetic only.

--mode free (the default for the electron gas) is the reference's
matrix-free mode (-m 1): V^pppp is rebuilt inside the pppp ladder every
iteration instead of being read from HBM; the iterates are bit-identical to
--mode stored (the default for the pairing model and the sweep).

Before timing, every context tunes its iteration schedule and its pppp-ladder
plan on the workload (ccd_set_schedule(AUTO), ccd_tune_ladder): FP32
iterations may then run some ladder classes on BF16x3 split products where
those are faster (--no-ladder-tune keeps the default plan); the plan and an
A/B of the kernel kinds per class are reported.

Per-kernel figures come from ccd_time_kernels (HIP events around each launch
group, the groups in order on the context's stream) and the algorithmic work
of each group (ccd_kernel_work); `roofline` is the group with the largest
single-stream time.

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
    "heg54": dict(kind="heg", args=(1.0, 32, 4), golden="heg54_oracle.json",
                  name="electron gas N=54, r_s=1.0, n^2<=36 (1850 s.p. states), CCD iteration"),
    # reference CLI: --electronGas -r 0.5 -s 46 -p 5
    "heg66": dict(kind="heg", args=(0.5, 46, 5), golden="heg66_oracle.json",
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
# MI355X published peaks
HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6
FP32_PEAK_TFLOPS = 157.3
N_CU = 256
CLOCK_GHZ = 2.4            # peak engine clock; one wave64 VALU instruction per CU per clock

LADDER_GROUPS = ("ladder_tx", "ladder_v1", "ladder_v2", "ladder_v34", "ladder_v58")
KERNEL_OF_GROUP = {
    "ring": "k_ring_staged / k_ring", "remap_r": "k_transpose (R to pphh)",
    "ladder_x": "k_ladder_x", "ladder_tx": "k_ladder_tx",
    "ladder_v1": "k_ladder_v<{ct},1,{src}>", "ladder_v2": "k_ladder_v<{ct},2,{src}>",
    "ladder_v34": "k_ladder_v2<{ct},3,4,{src}>", "ladder_v58": "k_ladder_v2<{ct},5..8,{src}>",
    "update": "k_update<true,{ct}>", "remap_t": "k_transpose<double,{ct}> (T to T^ph)",
    "reduce": "k_reduce",
}
# the size classes of the pppp ladder each launch group times
CLASSES_OF_GROUP = {"ladder_v1": (1,), "ladder_v2": (2,), "ladder_v34": (3, 4), "ladder_v58": (5, 6, 7, 8)}
KIND_LABEL = {"default": "{dflt}", "mfma16": "16x16x4 MFMA", "bf16x3": "BF16x3 split products, 16x16x32 BF16 MFMA",
              "bf16x3sym": "BF16x3 split products on the symmetry of V^pppp, 16x16x16 and 16x16x32 BF16 MFMA",
              "bf16x3w": "BF16x3 split products, waves split k, four rows per lane, 16x16x16 BF16 MFMA"}
BF16X3_KINDS = ("bf16x3", "bf16x3sym", "bf16x3w")
BF16_PEAK_TFLOPS = 2500.0   # dense BF16 MFMA; a BF16x3 product costs three


def ladder_label(g, ct, src, plan):
    """k_ladder_grp instance and kernel kind of a ladder launch group."""
    classes = CLASSES_OF_GROUP[g]
    kinds = plan["kind_fp32" if ct == "float" else "kind_fp64"]
    parts = []
    for n in classes:
        dflt = "one-row-per-lane 4x4x1 MFMA" if (ct == "float" and n == 1) else "16x16x4 MFMA"
        parts.append(f"class {n}: " + KIND_LABEL[kinds[n - 1]].format(dflt=dflt))
    return f"k_ladder_grp<{ct},{src}> (" + "; ".join(parts) + ")"


def group_kind(g, ct, plan):
    kinds = {plan["kind_fp32" if ct == "float" else "kind_fp64"][n - 1] for n in CLASSES_OF_GROUP[g]}
    return kinds.pop() if len(kinds) == 1 else "mixed"


def make_model(cfg):
    from coupled_cluster_doubles_amd import Model
    if cfg["kind"] == "heg":
        return Model.electron_gas(*cfg["args"])
    return Model.pairing(*cfg["args"])


def make_context(cfg, prec, free, models):
    from coupled_cluster_doubles_amd import Batch, Context
    if cfg["kind"] == "sweep":
        return Batch(models, prec, matrix_free=free)
    return Context(models[0], prec, matrix_free=free)


def cpu_threads():
    """The host threads the CPU baseline uses: the process's CPU share
    (OMP_NUM_THREADS where the host sets it), else every visible CPU."""
    env = os.environ.get("OMP_NUM_THREADS", "")
    return int(env) if env.isdigit() and int(env) > 0 else (os.cpu_count() or 1)


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


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


REPS = 5   # back-to-back launches of each group per timed iteration


def kernel_report(ctx, iters, ct, src):
    """Single-stream time per launch, algorithmic work and roofline fraction
    of every launch group of one iteration in the context's current
    arithmetic; and the sum of the groups' times."""
    t = ctx.time_kernels(iters, REPS)
    w = ctx.kernel_work()
    plan = ctx.ladder_plan()
    peak_tf = FP64_PEAK_TFLOPS if ct == "double" else FP32_PEAK_TFLOPS
    out = {}
    for g, ms in t.items():
        if g == "iteration":
            continue
        sec = ms * 1e-3
        lab = ladder_label(g, ct, src, plan) if g in CLASSES_OF_GROUP else KERNEL_OF_GROUP[g].format(ct=ct, src=src)
        e = {"kernel": lab, "ms": ms, "flops": w[g]["flops"], "bytes": w[g]["bytes"]}
        if g in CLASSES_OF_GROUP:
            e["kind"] = group_kind(g, ct, plan)
        if g == "ring" and ct == "float" and plan.get("ring_fp32") == "bf16x3":
            e["kernel"] += " (staged ring: BF16x3 split products, 16x16x16 BF16 MFMA)"
            e["kind"] = "bf16x3"
        if sec > 0:
            e["tflops"] = w[g]["flops"] / sec / 1e12
            e["gbs"] = w[g]["bytes"] / sec / 1e9
            e["frac_mfma"] = e["tflops"] / peak_tf
            e["frac_hbm"] = e["gbs"] / HBM_PEAK_GBS
        if w[g]["elements"] > 0:
            e["generated_elements"] = w[g]["elements"]
            e["elements_per_s"] = w[g]["elements"] / sec if sec > 0 else None
        out[g] = e
    return out, sum(v["ms"] for v in out.values())


def load_json(path):
    if os.path.exists(path):
        with open(path) as f:
            return json.load(f)
    return None


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
    ap.add_argument("--no-solves", action="store_true", help="skip the full solves (profiling runs)")
    ap.add_argument("--cpu-iters", type=int, default=3)
    ap.add_argument("--no-ladder-tune", action="store_true",
                    help="keep the default pppp-ladder plan instead of tuning it on the workload "
                         "(FP32 iterations may then use BF16x3 split products where they are faster)")
    ap.add_argument("--plan-json", default=None,
                    help="use this pppp-ladder plan (a JSON object shaped like the line's ladder_plan) "
                         "instead of tuning one (profiling runs: the same kernels as the tuned run)")
    ap.add_argument("--no-kind-ab", action="store_true", help="skip the A/B of the ladder kernel kinds")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="first copy this many MiB device-to-device (a kernel of known traffic, "
                         "to calibrate hardware byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build runs the CCD iteration on a single "
              "MI355X (multi-GPU channel sharding is a design note, not built).", file=sys.stderr)
        sys.exit(2)

    from coupled_cluster_doubles_amd import FP32, FP64, TUNE_BF16X3, TUNE_KINDS, Context

    def tune(c, p):
        """The schedule and the ladder plan tuned on the context's problem (the
        ladder in the arithmetic p; FP32 may pick other kernel kinds).  The
        sweep does not offer BF16x3 products to the tuner: its slowest
        near-divergent solves contract so slowly that the BF16x3 rounding
        (about 2^-16 per product) can hold |dE| above the switch tolerance
        in the FP32 phase."""
        c.set_schedule("auto")
        if not args.no_ladder_tune:
            flags = 0 if p == FP64 else TUNE_KINDS if sweep else TUNE_KINDS | TUNE_BF16X3
            c.tune_ladder(flags)
            c.set_schedule("auto")

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
    ct = "double" if prec == FP64 else "float"
    if args.mode is None:
        args.mode = "free" if cfg["kind"] == "heg" else "stored"
    free = args.mode == "free"
    src = "LV_GAS" if (free and cfg["kind"] == "heg") else "LV_GENERIC" if free else "LV_STORED"
    sweep = cfg["kind"] == "sweep"
    if sweep:
        from coupled_cluster_doubles_amd.sweep import sweep_models
        models = sweep_models()
    else:
        models = [make_model(cfg)]
    t_setup = time.perf_counter()
    ctx = make_context(cfg, prec, free, models)
    setup_s = time.perf_counter() - t_setup
    t_tune = time.perf_counter()
    if args.plan_json:
        with open(args.plan_json) as f:
            ctx.set_ladder_plan(json.load(f))
        ctx.set_schedule("auto")
    else:
        tune(ctx, prec)           # time the schedules and ladder plans on this problem, keep the fastest
    tune_s = time.perf_counter() - t_tune
    plan = ctx.ladder_plan()
    mbpt2 = ctx.mbpt2()
    work = ctx.work()
    sched, sched_us = ctx.schedule()

    # the steps are queued back to back (no host round trip between them);
    # the call returns after the last one has finished on the device
    ctx.iterate_n(args.warmup)
    t0 = time.perf_counter()
    e = ctx.iterate_n(args.steps)
    dt = time.perf_counter() - t0
    e = e[:, -1].tolist() if sweep else float(e[-1])
    ms_step = 1e3 * dt / args.steps
    iters_per_s = args.steps / dt

    # per launch group, single stream, in the steps' arithmetic
    n_timed = max(3, min(args.steps, 10))
    kern, iter_ms_serial = kernel_report(ctx, n_timed, ct, src)
    kern64 = None
    if prec == FP32:     # the same context's FP64 iterations
        ctx.set_precision(FP64)
        kern64, iter64_ms_serial = kernel_report(ctx, n_timed, "double", src)
        ctx.set_precision(FP32)
    # A/B of the FP32 ladder kernel kinds: every class on one kind, the other
    # plan fields as tuned, single stream, microseconds per launch of each group
    kind_ab = None
    if prec == FP32 and not args.no_kind_ab:
        kind_ab = {}
        for kind in ("default", "mfma16", "bf16x3", "bf16x3sym", "bf16x3w"):
            # the wave-k kernels are for class 1: the others on BF16x3
            q = dict(plan, kind_fp32=[kind] * 8 if kind not in ("bf16x3sym", "bf16x3w") else [kind] + ["bf16x3"] * 7)
            ctx.set_ladder_plan(q)
            tk = ctx.time_kernels(n_timed, REPS)
            kind_ab[kind] = {g: tk[g] * 1e3 for g in CLASSES_OF_GROUP}
        # and the staged ring's two kernels (microseconds per launch of the ring group)
        for ring in ("fp32", "bf16x3"):
            ctx.set_ladder_plan(dict(plan, ring_fp32=ring))
            kind_ab[f"ring_{ring}"] = {"ring": ctx.time_kernels(n_timed, REPS)["ring"] * 1e3}
        ctx.set_ladder_plan(plan)

    # executed flops of an iteration: the products on packed pairs
    executed_flops = sum(v["flops"] for v in kern.values())
    out_eff = {"gemm_tflops_effective": work["flops"] / (ms_step * 1e-3) / 1e12,
               "gemm_tflops_effective_note": "the reference's ordered-pair flops (gemm_flops_per_iter) "
                                             "over the step time; the device runs the pppp ladder on "
                                             "pairs a<b, a quarter of those flops",
               "gemm_tflops_executed": executed_flops / (ms_step * 1e-3) / 1e12,
               "gemm_flops_per_iter": work["flops"], "executed_flops_per_iter": executed_flops}

    # roofline of the dominant launch group, measured on the stream it runs on
    dom = max(kern, key=lambda g: kern[g]["ms"])
    # the committed counter summary of this workload: per class-1 kernel kind
    # (profiles/ladder_v1_<config>_<precision>_<mode>_<kind>.json, with the
    # calibrated traffic of every launch group); a ladder group's own figures
    # only where its kernel kind is the profiled one
    k1 = plan["kind_fp32" if ct == "float" else "kind_fp64"][0]
    prof_name = f"ladder_v1_{args.config}_{args.precision}_{args.mode}_{k1}.json"
    prof_all = load_json(os.path.join(ROOT, "profiles", prof_name))
    prof = None
    if prof_all:
        if dom == "ladder_v1":
            prof = prof_all
        elif not dom.startswith("ladder_v") and prof_all.get("traffic_per_group", {}).get(dom):
            prof = {"bytes_per_launch": prof_all["traffic_per_group"][dom]}

    def roof_of(g, rep, peak_tf):
        r = rep[g]
        if g.startswith("ladder_v") and free:
            if r.get("kind") in BF16X3_KINDS:   # three BF16 products per FP32 product
                peak_tf = BF16_PEAK_TFLOPS / 3
            d = {"bound": "mfma", "kernel": r["kernel"], "achieved": r["tflops"], "peak": peak_tf,
                 "unit": "TFLOP/s", "frac": r["tflops"] / peak_tf, "algorithmic_flops": r["flops"],
                 "algorithmic_bytes": r["bytes"], "launch_ms": r["ms"]}
            if r.get("kind") in BF16X3_KINDS:
                d["peak_note"] = "dense BF16 MFMA peak / 3 (each product is three BF16 products)"
        elif g in ("ring", "ladder_x", "ladder_tx") or g.startswith("ladder_v"):
            d = {"bound": "mfma" if g == "ring" else "hbm", "kernel": r["kernel"], "launch_ms": r["ms"],
                 "algorithmic_flops": r["flops"], "algorithmic_bytes": r["bytes"]}
            if d["bound"] == "mfma":
                d.update(achieved=r["tflops"], peak=peak_tf, unit="TFLOP/s", frac=r["tflops"] / peak_tf)
            else:
                d.update(achieved=r["gbs"], peak=HBM_PEAK_GBS, unit="GB/s", frac=r["gbs"] / HBM_PEAK_GBS)
        else:
            d = {"bound": "hbm", "kernel": r["kernel"], "achieved": r["gbs"], "peak": HBM_PEAK_GBS,
                 "unit": "GB/s", "frac": r["gbs"] / HBM_PEAK_GBS, "algorithmic_bytes": r["bytes"],
                 "launch_ms": r["ms"]}
        return d

    peak = FP64_PEAK_TFLOPS if prec == FP64 else FP32_PEAK_TFLOPS
    roof = roof_of(dom, kern, peak)
    roof["traffic"] = prof.get("bytes_per_launch") if prof else None
    roof["traffic_source"] = (f"profiles/{prof_name} "
                              "(rocprofv3 FETCH_SIZE / WRITE_SIZE, calibrated)") if prof else None
    if prof and prof.get("valu_insts_per_element") and kern[dom].get("generated_elements"):
        # VALU-issue roof of the generated ladder: every generated element costs
        # ipe VALU lane-instructions; a CU issues one wave64 instruction per clock
        ipe = prof["valu_insts_per_element"]
        peak_el = N_CU * CLOCK_GHZ * 1e9 * 64 / ipe
        ach = kern[dom]["elements_per_s"]
        roof["valu_roof"] = {"unit": "elements/s", "achieved": ach, "peak": peak_el, "frac": ach / peak_el,
                             "valu_insts_per_element": ipe, "insts_source": prof.get("valu_source")}

    ladder_ms = sum(kern[g]["ms"] for g in LADDER_GROUPS)
    ladder_flops = sum(kern[g]["flops"] for g in LADDER_GROUPS)
    phase = {"groups": list(LADDER_GROUPS), "ms": ladder_ms, "algorithmic_flops": ladder_flops,
             "achieved": ladder_flops / (ladder_ms * 1e-3) / 1e12, "peak": peak, "unit": "TFLOP/s"}
    phase["frac"] = phase["achieved"] / peak
    fp64 = None
    if kern64 is not None:
        r64 = roof_of(dom, kern64, FP64_PEAK_TFLOPS)
        l64 = sum(kern64[g]["ms"] for g in LADDER_GROUPS)
        f64 = sum(kern64[g]["flops"] for g in LADDER_GROUPS)
        fp64 = {"dominant": r64, "ladder_phase": {"ms": l64, "achieved": f64 / (l64 * 1e-3) / 1e12,
                                                 "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                                                 "frac": f64 / (l64 * 1e-3) / 1e12 / FP64_PEAK_TFLOPS},
                "iteration_ms_single_stream": iter64_ms_serial}
    hbm = {g: {"kernel": kern[g]["kernel"], "ms": kern[g]["ms"], "algorithmic_GBs": kern[g].get("gbs"),
               "frac": kern[g].get("frac_hbm")} for g in ("remap_t", "remap_r", "update")}
    if not sweep:
        # the remaps' index tables: a source and a destination address (int32)
        # and a permutation entry (uint16) per element
        idx = 10.0 * ctx.n_t2
        for g in ("remap_t", "remap_r"):
            if kern[g]["ms"] > 0:
                hbm[g]["kernel_GBs_with_index_tables"] = (kern[g]["bytes"] + idx) / (kern[g]["ms"] * 1e-3) / 1e9

    mixed = None
    gold = load_json(os.path.join(ROOT, "tests", "golden", cfg["golden"])) if cfg.get("golden") else None
    if args.precision == "mixed" and not sweep and not args.no_solves:
        # full solves from T2 = 0, each run once untimed, then timed on the
        # host clock around the call
        model = models[0]
        n_part = model.n_holes

        def timed(f):
            f()
            t0 = time.perf_counter()
            r = f()
            return r, time.perf_counter() - t0

        (e_m, stop_m, n_m, n32), t_m = timed(lambda: ctx.solve_mixed(1e-10, args.switch_tol))
        ctx64 = Context(model, FP64, matrix_free=free)
        tune(ctx64, FP64)
        (e_d, stop_d, n_d), t_d = timed(lambda: ctx64.solve(1e-10))
        ctx64.close()
        ctx_st = Context(model, FP64, matrix_free=False)     # the plain FP64 GPU path: V^pppp stored
        tune(ctx_st, FP64)
        (e_s, stop_s, n_s), t_s = timed(lambda: ctx_st.solve(1e-10))
        ctx_st.close()
        mixed = {"tol": 1e-10, "switch_tol": args.switch_tol,
                 "mixed_iters": n_m, "mixed_fp32_iters": n32, "mixed_stop": stop_m,
                 "mixed_solve_s": t_m, "mixed_energy": float(e_m[-1]),
                 "fp64_iters": n_d, "fp64_stop": stop_d, "fp64_solve_s": t_d,
                 "fp64_energy": float(e_d[-1]), "fp64_mode": args.mode,
                 "fp64_stored_iters": n_s, "fp64_stored_stop": stop_s, "fp64_stored_solve_s": t_s,
                 "abs_dE_per_particle": abs(float(e_m[-1]) - float(e_d[-1])) / n_part,
                 "speedup_vs_fp64_solve": t_d / t_m, "speedup_vs_fp64_stored_solve": t_s / t_m}
        if gold:
            e_o = gold["energies"][-1]
            mixed["oracle_energy"] = e_o
            mixed["oracle_iters"] = len(gold["energies"])
            mixed["abs_dE_per_particle_vs_oracle"] = abs(float(e_m[-1]) - e_o) / n_part
            mixed["fp64_abs_dE_per_particle_vs_oracle"] = abs(float(e_d[-1]) - e_o) / n_part
            mixed["fp64_stored_abs_dE_per_particle_vs_oracle"] = abs(float(e_s[-1]) - e_o) / n_part
        # the same solve with other switch points (1e-10 = tol: FP32 iterations only)
        sw_runs = []
        for sw in (1e-6, 1e-8, 1e-10):
            (e_w, stop_w, n_w, n32_w), t_w = timed(lambda: ctx.solve_mixed(1e-10, sw))
            r = {"switch_tol": sw, "iters": n_w, "fp32_iters": n32_w, "stop": stop_w,
                 "solve_s": t_w, "speedup_vs_fp64_solve": t_d / t_w,
                 "abs_dE_per_particle": abs(float(e_w[-1]) - float(e_d[-1])) / n_part}
            if gold:
                r["abs_dE_per_particle_vs_oracle"] = abs(float(e_w[-1]) - gold["energies"][-1]) / n_part
            sw_runs.append(r)
        mixed["switch_sweep"] = sw_runs
    elif args.precision == "mixed" and not args.no_solves:
        # every solve of the sweep from T2 = 0, mixed and FP64, one batch each
        t0 = time.perf_counter()
        rm = ctx.solve_mixed(1e-10, args.switch_tol)
        t_m = time.perf_counter() - t0
        ctx.close()
        ctx64 = make_context(cfg, FP64, free, models)
        tune(ctx64, FP64)
        t0 = time.perf_counter()
        rd = ctx64.solve(1e-10)
        t_d = time.perf_counter() - t0
        conv = [q for q in range(len(models)) if rd[q][1] == "converged"]
        cfgs = __import__("coupled_cluster_doubles_amd.sweep", fromlist=["sweep_configs"]).sweep_configs()
        mixed = {"tol": 1e-10, "switch_tol": args.switch_tol,
                 "fp64_iters_per_solve": [r[2] for r in rd],
                 "not_converged": [{"solve": q, "config": list(cfgs[q]), "fp64": rd[q][1], "iters": rd[q][2]}
                                   for q in range(len(models)) if rd[q][1] != "converged"],
                 "fp64_stops": {k: sum(r[1] == k for r in rd) for k in ("converged", "max_iterations", "nan")},
                 "mixed_converged_where_fp64_did": all(rm[q][1] == "converged" for q in conv),
                 "max_abs_dE_per_particle": max(abs(float(rm[q][0][-1]) - float(rd[q][0][-1]))
                                                / models[q].n_holes for q in conv),
                 "batch_iters_mixed": max(r[2] for r in rm), "batch_iters_fp64": max(r[2] for r in rd),
                 "mixed_solve_s": t_m, "fp64_solve_s": t_d, "speedup_vs_fp64_solve": t_d / t_m}
        ctx64.close()

    cpu = None
    if not args.no_cpu_baseline:
        threads = cpu_threads()
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
                      f"(OpenMP over channels, V^pppp "
                      f"{'rebuilt per channel' if cpu_free else 'stored'}), after one untimed iteration")
        cpu = {"value": v, "unit": "iter/s", "cores": threads, "kind": "port", "sample": sample,
               "cpu_model": cpu_model(), "host_cpus_visible": os.cpu_count(),
               "threads_rule": "OMP_NUM_THREADS (the process's CPU share) where set, else every visible CPU",
               "blas": "none: the oracle's own row-major triple-loop dgemm (oracle/ccd_oracle.c)"}

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
        "dtype": "f64" if prec == FP64 else ("f32+bf16x3+f64" if set(BF16X3_KINDS) & set(plan["kind_fp32"]) or plan.get("ring_fp32") == "bf16x3"
                                             else "f32+f64"),
        "mixed_vs_fp64": mixed,
        "data": "synthetic: the reference's closed-form basis and Coulomb interaction, T2 from zero",
        "config": {"workload": cfg["name"], "precision": args.precision, "mode": args.mode,
                   "n_t2": ctx.n_t2, "device_bytes": ctx.device_bytes},
        **out_eff,
        "mbpt2": mbpt2.tolist() if sweep else mbpt2,
        "corr_after_steps": e,
        "iteration_ms_single_stream": iter_ms_serial,
        "kernels": kern,
        "setup_s": setup_s,
        "schedule": {"chosen": sched, "tuner_us_per_iter": sched_us},
        "ladder_plan": plan,
        "ladder_kind_ab_us": kind_ab,
        "tune_s": tune_s,
        "roofline": roof,
        "roofline_ladder_phase": phase,
        "fp64": fp64,
        "hbm": hbm,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
