"""ctypes wrapper of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: used by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py to check and time the MI355X path.  The product
package never imports it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liboracle.so")
_lib = None


def build() -> str:
    """Compile liboracle.so with the committed Makefile."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            raise RuntimeError(f"{_LIB} is not built; run `make -C oracle`")
        L = C.CDLL(_LIB)
        vp, sz, dp, ip = C.c_void_p, C.c_size_t, C.POINTER(C.c_double), C.c_void_p
        L.orc_create_pairing.restype = vp
        L.orc_create_pairing.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double]
        L.orc_create_heg.restype = vp
        L.orc_create_heg.argtypes = [C.c_double, C.c_int, C.c_int]
        L.orc_create_infmatter.restype = vp
        L.orc_create_infmatter.argtypes = [C.c_double, C.c_int, C.c_int]
        L.orc_destroy.argtypes = [vp]
        L.orc_load.argtypes = [vp]
        L.orc_set_matrix_free.argtypes = [vp, C.c_int]
        L.orc_set_threads.argtypes = [C.c_int]
        for f in ("orc_nsp", "orc_nholes", "orc_nchannels", "orc_nt2"):
            getattr(L, f).restype = sz
            getattr(L, f).argtypes = [vp]
        L.orc_eref.restype = C.c_double
        L.orc_eref.argtypes = [vp]
        L.orc_mbpt2.restype = C.c_double
        L.orc_mbpt2.argtypes = [vp]
        L.orc_get_eps.argtypes = [vp, ip]
        L.orc_get_qn.argtypes = [vp, ip]
        L.orc_get_dims.argtypes = [vp, ip, ip, ip, ip]
        L.orc_t2_labels.argtypes = [vp, ip, ip, ip, ip]
        L.orc_tph_labels.argtypes = [vp, ip, ip, ip, ip]
        L.orc_get_vhhpp.argtypes = [vp, ip]
        L.orc_ladder.argtypes = [vp, ip, ip]
        L.orc_ring.argtypes = [vp, ip, ip]
        L.orc_scatter_ph.argtypes = [vp, ip, ip]
        L.orc_onebody.argtypes = [vp, ip, ip, ip]
        L.orc_step.restype = C.c_double
        L.orc_step.argtypes = [vp, ip, ip, C.c_double]
        L.orc_solve.restype = C.c_int
        L.orc_solve.argtypes = [vp, C.c_double, ip, C.c_int, C.POINTER(C.c_int), dp]
        L.orc_get_t2.argtypes = [vp, ip]
        _lib = L
    return _lib


def _p(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


STOP = {0: "converged", 1: "max_iterations", 2: "nan"}


class Oracle:
    """One basis + its channel-blocked chains, reference layout."""

    def __init__(self, handle, load=True, matrix_free=False):
        self.L = lib()
        self.h = handle
        if matrix_free:
            self.L.orc_set_matrix_free(self.h, 1)
        if load and self.L.orc_load(self.h) != 0:
            raise MemoryError("oracle: allocation failed")

    @classmethod
    def pairing(cls, n_levels=8, n_particles=8, g=0.5, xi=1.0, load=True):
        return cls(lib().orc_create_pairing(n_levels, n_particles, g, xi), load)

    @classmethod
    def electron_gas(cls, r_s=1.0, n_shells=4, n_particle_shells=2, load=True, matrix_free=False):
        return cls(lib().orc_create_heg(r_s, n_shells, n_particle_shells), load, matrix_free)

    @classmethod
    def inf_matter(cls, density=0.08, n_shells=4, n_particle_shells=2, load=True):
        """Pure neutron matter, Minnesota potential (reference --infMatter, tzMax 1)."""
        return cls(lib().orc_create_infmatter(density, n_shells, n_particle_shells), load)

    def close(self):
        if self.h:
            self.L.orc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- sizes and setup data
    @property
    def nsp(self):
        return int(self.L.orc_nsp(self.h))

    @property
    def nh(self):
        return int(self.L.orc_nholes(self.h))

    @property
    def nchannels(self):
        return int(self.L.orc_nchannels(self.h))

    @property
    def nt2(self):
        return int(self.L.orc_nt2(self.h))

    def eref(self):
        return float(self.L.orc_eref(self.h))

    def mbpt2(self):
        return float(self.L.orc_mbpt2(self.h))

    def eps(self):
        e = np.empty(self.nsp)
        self.L.orc_get_eps(self.h, _p(e))
        return e

    def qn(self):
        q = np.empty((self.nsp, 5), dtype=np.int32)
        self.L.orc_get_qn(self.h, _p(q))
        return q

    def dims(self):
        n = self.nchannels
        arrs = [np.empty(n, dtype=np.int64) for _ in range(4)]
        self.L.orc_get_dims(self.h, *[_p(a) for a in arrs])
        return dict(zip(("pp", "hh", "mph", "mhp"), arrs))

    def t2_labels(self):
        n = self.nt2
        arrs = [np.empty(n, dtype=np.int32) for _ in range(4)]
        self.L.orc_t2_labels(self.h, *[_p(a) for a in arrs])
        return np.stack(arrs, axis=1)            # columns a b i j

    def tph_labels(self):
        n = self.nt2
        arrs = [np.empty(n, dtype=np.int32) for _ in range(4)]
        self.L.orc_tph_labels(self.h, *[_p(a) for a in arrs])
        return np.stack(arrs, axis=1)            # columns a i j b

    def vhhpp(self):
        v = np.empty(self.nt2)
        self.L.orc_get_vhhpp(self.h, _p(v))
        return v

    # -- diagrams
    def ladder(self, t):
        t = np.ascontiguousarray(t, dtype=np.float64)
        d = np.empty_like(t)
        self.L.orc_ladder(self.h, _p(t), _p(d))
        return d

    def ring(self, tph):
        tph = np.ascontiguousarray(tph, dtype=np.float64)
        r = np.empty_like(tph)
        self.L.orc_ring(self.h, _p(tph), _p(r))
        return r

    def scatter_ph(self, t):
        t = np.ascontiguousarray(t, dtype=np.float64)
        out = np.empty_like(t)
        self.L.orc_scatter_ph(self.h, _p(t), _p(out))
        return out

    def onebody(self, t):
        t = np.ascontiguousarray(t, dtype=np.float64)
        chi = np.empty(self.nh)
        psi = np.empty(self.nsp - self.nh)
        self.L.orc_onebody(self.h, _p(t), _p(chi), _p(psi))
        return chi, psi

    def step(self, t_old, mixing=0.5):
        t_old = np.ascontiguousarray(t_old, dtype=np.float64)
        t_new = np.empty_like(t_old)
        e = self.L.orc_step(self.h, _p(t_old), _p(t_new), mixing)
        return t_new, float(e)

    def solve(self, tol, cap=128):
        en = np.zeros(cap)
        stop = C.c_int(0)
        secs = C.c_double(0.0)
        n = self.L.orc_solve(self.h, tol, _p(en), cap, C.byref(stop), C.byref(secs))
        return en[:min(n, cap)].copy(), STOP[stop.value], float(secs.value)

    def t2(self):
        t = np.empty(self.nt2)
        self.L.orc_get_t2(self.h, _p(t))
        return t


def set_threads(n: int):
    lib().orc_set_threads(int(n))
