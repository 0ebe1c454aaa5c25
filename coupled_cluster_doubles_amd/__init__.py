"""MI355X-native coupled-cluster-doubles iteration.

Python mirror of the reference's interface for the CCD hot path
(justinlietz/coupled_cluster_doubles): a basis model (PairingSPBasis,
ElectronGasSPBasis, src/parse.cpp:231-244) and a device context that
provides MBPT2Corr, the diagram functions of src/calc.cpp and the
iteration loop of src/main.cpp:491-682.  All compute runs in
libccd_amd.so on the GPU; there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import CCDError, LadderPlan, Problem, check, lib

FP64 = 0
FP32 = 1
MATRIX_FREE = 0x100   # OR into the precision: V^pppp rebuilt inside the ladder (reference -m 1)
STOP_REASONS = {0: "converged", 1: "max_iterations", 2: "nan"}
# iteration schedules (CCD_SCHED_*): how the same kernels are launched
SCHEDULES = {"auto": 0, "streams": 1, "serial": 2, "serial_graph": 3}
# pppp-ladder plan vocabulary (CCD_LK_*, CCD_LA_*, CCD_LP_*, CCD_TUNE_*)
LADDER_KINDS = {"default": 0, "mfma16": 1, "bf16x3": 2, "bf16x3sym": 3, "bf16x3w": 4}
RING_KINDS = {"fp32": 0, "bf16x3": 1}
LADDER_ASSIGN = {"model": 0, "channel": 1, "longest": 2, "runs": 3, "xcd": 4}
LADDER_PARTITIONS = {"split": 0, "one": 1}
TUNE_KINDS = 1
TUNE_BF16X3 = 2
TUNE_RING_BF16X3 = 4
# launch groups of an iteration, in CCD_KT_* order (include/ccd_amd.h)
KERNEL_GROUPS = ("ring", "remap_r", "ladder_x", "ladder_tx", "ladder_v1", "ladder_v2", "ladder_v34",
                 "ladder_v58", "update", "remap_t", "reduce")

__all__ = ["Model", "Context", "Batch", "CCDError", "FP64", "FP32", "MATRIX_FREE", "STOP_REASONS",
           "TUNE_KINDS", "TUNE_BF16X3", "TUNE_RING_BF16X3"]


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


class Model:
    """Host problem model: s.p. basis, channel blocking and interaction."""

    def __init__(self, handle):
        self.h = handle
        n_sp, n_h, n_ch, n_t2 = C.c_int32(), C.c_int32(), C.c_int32(), C.c_int64()
        check(lib().ccd_model_sizes(self.h, C.byref(n_sp), C.byref(n_h), C.byref(n_ch),
                                    C.byref(n_t2)), "ccd_model_sizes")
        self.n_sp, self.n_holes = n_sp.value, n_h.value
        self.n_chan, self.n_t2 = n_ch.value, n_t2.value

    @classmethod
    def pairing(cls, n_pair_states=8, n_particles=8, g=0.5, xi=1.0):
        h = C.c_void_p()
        check(lib().ccd_model_pairing(n_pair_states, n_particles, g, xi, C.byref(h)),
              "ccd_model_pairing")
        return cls(h)

    @classmethod
    def electron_gas(cls, r_s=1.0, n_shells=4, n_particle_shells=2):
        h = C.c_void_p()
        check(lib().ccd_model_electron_gas(r_s, n_shells, n_particle_shells, C.byref(h)),
              "ccd_model_electron_gas")
        return cls(h)

    @classmethod
    def inf_matter(cls, density=0.08, n_shells=4, n_particle_shells=2, tz_max=1):
        """Pure neutron matter with the Minnesota potential (V generated on the device)."""
        h = C.c_void_p()
        check(lib().ccd_model_inf_matter(density, tz_max, n_shells, n_particle_shells, C.byref(h)),
              "ccd_model_inf_matter")
        return cls(h)

    def eref(self) -> float:
        e = C.c_double()
        check(lib().ccd_model_eref(self.h, C.byref(e)), "ccd_model_eref")
        return e.value

    def problem(self, host_v: bool = False):
        p = C.POINTER(Problem)()
        f = lib().ccd_model_problem_host_v if host_v else lib().ccd_model_problem
        check(f(self.h, C.byref(p)), "ccd_model_problem")
        return p

    def eps(self) -> np.ndarray:
        p = self.problem().contents
        return np.ctypeslib.as_array(C.cast(p.eps, C.POINTER(C.c_double)), (self.n_sp,)).copy()

    def dims(self):
        p = self.problem().contents
        out = {}
        for k in ("pp_dim", "hh_dim", "ph_dim", "hp_dim"):
            n = p.n_chan if k in ("pp_dim", "hh_dim") else p.n_mod
            out[k] = np.ctypeslib.as_array(C.cast(getattr(p, k), C.POINTER(C.c_int32)), (n,)).copy()
        return out

    def pairs(self, kind: str) -> np.ndarray:
        """Concatenated ordered pairs of one class ('pp', 'hh', 'ph', 'hp')."""
        p = self.problem().contents
        dims = self.dims()[kind + "_dim"]
        n = int(dims.sum())
        if n == 0:
            return np.zeros((0, 2), dtype=np.int32)
        arr = np.ctypeslib.as_array(C.cast(getattr(p, kind + "_pairs"), C.POINTER(C.c_int32)), (2 * n,))
        return arr.reshape(n, 2).copy()

    def t2_labels(self) -> np.ndarray:
        """(a, b, i, j) of every pphh element."""
        d = self.dims()
        pp, hh = self.pairs("pp"), self.pairs("hh")
        out = np.empty((self.n_t2, 4), dtype=np.int32)
        e = op = oh = 0
        for P, H in zip(d["pp_dim"], d["hh_dim"]):
            if P and H:
                blk = out[e:e + P * H].reshape(P, H, 4)
                blk[:, :, 0:2] = pp[op:op + P][:, None, :]
                blk[:, :, 2:4] = hh[oh:oh + H][None, :, :]
            e += P * H
            op += P
            oh += H
        return out

    def close(self):
        if self.h:
            lib().ccd_model_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Context:
    """Device-resident CCD state for one problem on the current HIP device."""

    def __init__(self, model: Model, precision: int = FP64, host_v: bool = False,
                 matrix_free: bool = False):
        self.model = model
        self.h = C.c_void_p()
        mode = precision | (MATRIX_FREE if matrix_free else 0)
        check(lib().ccd_ctx_create(model.problem(host_v), mode, C.byref(self.h)),
              "ccd_ctx_create")
        n, b = C.c_int64(), C.c_int64()
        check(lib().ccd_ctx_info(self.h, C.byref(n), C.byref(b)), "ccd_ctx_info")
        self.n_t2, self.device_bytes = n.value, b.value
        self.n_holes = model.n_holes
        self.n_sp = model.n_sp

    def close(self):
        if self.h:
            lib().ccd_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def mbpt2(self) -> float:
        e = C.c_double()
        check(lib().ccd_mbpt2(self.h, C.byref(e)), "ccd_mbpt2")
        return e.value

    def mbpt2_noalloc(self, d_corr: int, stream: int = 0):
        """ccd_mbpt2_noalloc: MBPT2 into the device double at address d_corr,
        queued on `stream` (a hipStream_t handle; 0 = the context's stream),
        with no allocation and no synchronisation."""
        check(lib().ccd_mbpt2_noalloc(self.h, C.c_void_p(d_corr), C.c_void_p(stream or None)),
              "ccd_mbpt2_noalloc")

    def reset(self):
        check(lib().ccd_reset(self.h), "ccd_reset")

    def iterate(self, mixing: float = 0.5) -> float:
        e = C.c_double()
        check(lib().ccd_iterate(self.h, mixing, C.byref(e)), "ccd_iterate")
        return e.value

    def iterate_n(self, n: int, mixing: float = 0.5) -> np.ndarray:
        """n iterations with no host round trip in between; their energies."""
        en = np.zeros(max(n, 1))
        check(lib().ccd_iterate_n(self.h, n, mixing, _ptr(en)), "ccd_iterate_n")
        return en[:n].copy()

    def solve(self, tol: float, max_iters: int = 100, mixing: float = 0.5, cap: int = 256):
        en = np.zeros(cap)
        n, stop = C.c_int(), C.c_int()
        check(lib().ccd_solve(self.h, tol, max_iters, mixing, _ptr(en), cap, C.byref(n),
                              C.byref(stop)), "ccd_solve")
        return en[:min(n.value, cap)].copy(), STOP_REASONS[stop.value], n.value

    def solve_mixed(self, tol: float, switch_tol: float, max_iters: int = 100, mixing: float = 0.5,
                    cap: int = 256):
        """FP32 contractions while |dE| > switch_tol, then FP64 sweeps to |dE| <= tol
        (contexts created with FP32).  Returns (energies, stop, n_iter, n_fp32)."""
        en = np.zeros(cap)
        n, n32, stop = C.c_int(), C.c_int(), C.c_int()
        check(lib().ccd_solve_mixed(self.h, tol, switch_tol, max_iters, mixing, _ptr(en), cap,
                                    C.byref(n), C.byref(n32), C.byref(stop)), "ccd_solve_mixed")
        return en[:min(n.value, cap)].copy(), STOP_REASONS[stop.value], n.value, n32.value

    def set_schedule(self, sched: str):
        """'auto' (time the others on this problem and keep the fastest;
        resets T2 to zero), 'streams', 'serial' or 'serial_graph'."""
        check(lib().ccd_set_schedule(self.h, SCHEDULES[sched]), "ccd_set_schedule")

    def schedule(self):
        """(schedule in use, {schedule: tuner's microseconds per iteration})."""
        k, us = C.c_int(), np.zeros(3)
        check(lib().ccd_get_schedule(self.h, C.byref(k), _ptr(us)), "ccd_get_schedule")
        names = {v: n for n, v in SCHEDULES.items()}
        return names[k.value], {names[q + 1]: float(us[q]) for q in range(3)}

    def ladder_plan(self) -> dict:
        """The pppp-ladder plan: per size class 1..8 (16(n-1) < hh <= 16n) the
        kernel kind for FP64 / FP32 iterations, blocks per compute unit (0: one
        per tile) and tile assignment; the launch partition; the ring's kernel
        in FP32 iterations; the tuner's microseconds per class launch and per
        iteration of each partition."""
        p = LadderPlan()
        check(lib().ccd_get_ladder_plan(self.h, C.byref(p)), "ccd_get_ladder_plan")
        kinds = {v: k for k, v in LADDER_KINDS.items()}
        assign = {v: k for k, v in LADDER_ASSIGN.items()}
        parts = {v: k for k, v in LADDER_PARTITIONS.items()}
        return {
            "kind_fp64": [kinds[p.kind[0][n]] for n in range(8)],
            "kind_fp32": [kinds[p.kind[1][n]] for n in range(8)],
            "blocks_per_cu": [int(p.blocks_per_cu[n]) for n in range(8)],
            "assign": [assign[p.assign[n]] for n in range(8)],
            "partition": parts[p.partition],
            "us_fp64": [float(p.us[0][n]) for n in range(8)],
            "us_fp32": [float(p.us[1][n]) for n in range(8)],
            "us_partition": {"split": float(p.us_partition[0]), "one": float(p.us_partition[1])},
            "ring_fp32": {v: k for k, v in RING_KINDS.items()}[p.ring_fp32],
        }

    def set_ladder_plan(self, plan: dict):
        """Set the plan from a dict shaped like ladder_plan() (the us fields are ignored)."""
        p = LadderPlan()
        for n in range(8):
            p.kind[0][n] = LADDER_KINDS[plan["kind_fp64"][n]]
            p.kind[1][n] = LADDER_KINDS[plan["kind_fp32"][n]]
            p.blocks_per_cu[n] = plan["blocks_per_cu"][n]
            p.assign[n] = LADDER_ASSIGN[plan["assign"][n]]
        p.partition = LADDER_PARTITIONS[plan["partition"]]
        p.ring_fp32 = RING_KINDS[plan.get("ring_fp32", "fp32")]
        check(lib().ccd_set_ladder_plan(self.h, C.byref(p)), "ccd_set_ladder_plan")

    def tune_ladder(self, flags: int = 0) -> dict:
        """Tune the ladder plan for the current arithmetic (TUNE_KINDS,
        TUNE_BF16X3, TUNE_RING_BF16X3 widen the kernels tried); resets T2 to
        zero."""
        check(lib().ccd_tune_ladder(self.h, flags), "ccd_tune_ladder")
        return self.ladder_plan()

    def set_precision(self, precision: int):
        """Arithmetic of the contractions of the following iterations (FP64 or FP32)."""
        check(lib().ccd_set_precision(self.h, precision), "ccd_set_precision")

    def get_t2(self) -> np.ndarray:
        t = np.empty(self.n_t2)
        check(lib().ccd_get_t2(self.h, _ptr(t)), "ccd_get_t2")
        return t

    def set_t2(self, t: np.ndarray):
        t = np.ascontiguousarray(t, dtype=np.float64)
        assert t.size == self.n_t2
        check(lib().ccd_set_t2(self.h, _ptr(t)), "ccd_set_t2")

    def _io(self, fn, name, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        assert x.size == self.n_t2
        y = np.empty(self.n_t2)
        check(fn(self.h, _ptr(x), _ptr(y)), name)
        return y

    def ladder(self, t):
        return self._io(lib().ccd_ladder, "ccd_ladder", t)

    def ring(self, tph):
        return self._io(lib().ccd_ring, "ccd_ring", tph)

    def remap_ph(self, t):
        return self._io(lib().ccd_remap_ph, "ccd_remap_ph", t)

    def onebody(self, t):
        t = np.ascontiguousarray(t, dtype=np.float64)
        chi = np.empty(self.n_holes)
        psi = np.empty(self.n_sp - self.n_holes)
        check(lib().ccd_onebody(self.h, _ptr(t), _ptr(chi), _ptr(psi)), "ccd_onebody")
        return chi, psi

    def step(self, t_old, mixing: float = 0.5):
        t_old = np.ascontiguousarray(t_old, dtype=np.float64)
        t_new = np.empty(self.n_t2)
        e = C.c_double()
        check(lib().ccd_step(self.h, _ptr(t_old), _ptr(t_new), mixing, C.byref(e)), "ccd_step")
        return t_new, e.value

    def time_phases(self, iters: int, mixing: float = 0.5) -> np.ndarray:
        ms = np.zeros(6)
        check(lib().ccd_time_phases(self.h, iters, mixing, _ptr(ms)), "ccd_time_phases")
        return ms

    def time_kernels(self, iters: int, reps: int = 1, mixing: float = 0.5) -> dict:
        """Average device time (ms) of one launch of each launch group of an
        iteration, the groups run in order on one stream, each `reps` times
        back to back (ccd_time_kernels)."""
        ms = np.zeros(len(KERNEL_GROUPS) + 1)
        check(lib().ccd_time_kernels(self.h, iters, reps, mixing, _ptr(ms), ms.size), "ccd_time_kernels")
        out = dict(zip(KERNEL_GROUPS, ms[:-1].tolist()))
        out["iteration"] = float(ms[-1])
        return out

    def kernel_work(self) -> dict:
        """Algorithmic flops, bytes and generated V elements of each launch
        group of one iteration, in the arithmetic of the next iterations
        (ccd_kernel_work)."""
        n = len(KERNEL_GROUPS)
        f, b, e = np.zeros(n), np.zeros(n), np.zeros(n)
        check(lib().ccd_kernel_work(self.h, _ptr(f), _ptr(b), _ptr(e), n), "ccd_kernel_work")
        return {g: {"flops": float(f[k]), "bytes": float(b[k]), "elements": float(e[k])}
                for k, g in enumerate(KERNEL_GROUPS)}

    def work(self):
        f, fp, b = C.c_double(), C.c_double(), C.c_double()
        check(lib().ccd_work(self.h, C.byref(f), C.byref(fp), C.byref(b)), "ccd_work")
        return {"flops": f.value, "flops_pppp": fp.value, "ladder_bytes": b.value}


class Batch(Context):
    """Many independent solves in one device context: every kernel of an
    iteration covers all of them in one launch, each solve keeps its own
    stop rule and energy log, and each one's results are bit-identical to a
    Context of its own.  No counterpart in the reference (one solve per
    ccd.exe process)."""

    def __init__(self, models, precision: int = FP64, host_v: bool = False,
                 matrix_free: bool = False):
        self.models = list(models)
        self.model = None
        self.h = C.c_void_p()
        n = len(self.models)
        probs = (C.c_void_p * max(n, 1))(*[C.cast(m.problem(host_v), C.c_void_p) for m in self.models])
        mode = precision | (MATRIX_FREE if matrix_free else 0)
        check(lib().ccd_batch_create(probs, n, mode, C.byref(self.h)), "ccd_batch_create")
        nt, b = C.c_int64(), C.c_int64()
        check(lib().ccd_ctx_info(self.h, C.byref(nt), C.byref(b)), "ccd_ctx_info")
        self.n_t2, self.device_bytes = nt.value, b.value
        k = C.c_int()
        self.t2_offset = np.zeros(n + 1, dtype=np.int64)
        check(lib().ccd_batch_info(self.h, C.byref(k), _ptr(self.t2_offset)), "ccd_batch_info")
        self.n_problems = k.value
        self.n_holes = sum(m.n_holes for m in self.models)
        self.n_sp = sum(m.n_sp for m in self.models)

    def mbpt2(self) -> np.ndarray:
        e = np.zeros(self.n_problems)
        check(lib().ccd_batch_mbpt2(self.h, _ptr(e)), "ccd_batch_mbpt2")
        return e

    def iterate_n(self, n: int, mixing: float = 0.5) -> np.ndarray:
        """n iterations of every solve; energies [n_problems, n]."""
        en = np.zeros(max(n, 1) * self.n_problems)
        check(lib().ccd_batch_iterate_n(self.h, n, mixing, _ptr(en)), "ccd_batch_iterate_n")
        return en[:n * self.n_problems].reshape(self.n_problems, n).copy()

    def solve(self, tol: float, max_iters: int = 100, mixing: float = 0.5, cap: int = 256):
        """Every solve from T2 = 0: a list of (energies, stop, n_iter)."""
        P = self.n_problems
        en = np.zeros(cap * P)
        n, stop = np.zeros(P, dtype=np.int32), np.zeros(P, dtype=np.int32)
        check(lib().ccd_batch_solve(self.h, tol, max_iters, mixing, _ptr(en), cap, _ptr(n), _ptr(stop)),
              "ccd_batch_solve")
        return [(en[q * cap:q * cap + min(int(n[q]), cap)].copy(), STOP_REASONS[int(stop[q])], int(n[q]))
                for q in range(P)]

    def solve_mixed(self, tol: float, switch_tol: float, max_iters: int = 100, mixing: float = 0.5,
                    cap: int = 256):
        """Every solve in mixed precision: a list of (energies, stop, n_iter, n_fp32)."""
        P = self.n_problems
        en = np.zeros(cap * P)
        n, n32, stop = (np.zeros(P, dtype=np.int32) for _ in range(3))
        check(lib().ccd_batch_solve_mixed(self.h, tol, switch_tol, max_iters, mixing, _ptr(en), cap,
                                          _ptr(n), _ptr(n32), _ptr(stop)), "ccd_batch_solve_mixed")
        return [(en[q * cap:q * cap + min(int(n[q]), cap)].copy(), STOP_REASONS[int(stop[q])], int(n[q]),
                 int(n32[q])) for q in range(P)]

    def t2(self, q: int) -> np.ndarray:
        """Current amplitudes of solve q (its own pphh layout)."""
        return self.get_t2()[self.t2_offset[q]:self.t2_offset[q + 1]].copy()

