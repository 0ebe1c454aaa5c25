"""Dense spin-orbital CCD in numpy -- an independent check of the oracle.

TEST INFRASTRUCTURE ONLY.  Nothing in the product imports this module.

It shares no code and no data layout with the channel-blocked oracle
(oracle/ccd_oracle.c) or with the GPU path: the single-particle basis is
enumerated by brute force, the antisymmetrised matrix elements are built as
dense 4-index arrays straight from the Hamiltonian, and the amplitude
equations are the textbook spin-orbital CCD equations written with
numpy.einsum.  Only the physical constants of the electron gas and the
iteration protocol are taken from the reference, because they define the
numbers the reference prints:

* constants: src/ElectronGasSPBasis.cpp:178-190 (Hartree/Angstrom units);
* protocol: src/main.cpp:491-519 (T2 = 0, mixing 0.5, E_prev = 100,
  loop while |dE| > tol, halt after iteration 100).

Textbook CCD (e.g. Shavitt & Bartlett, Many-Body Methods in Chemistry and
Physics, ch. 9; Crawford & Schaefer, Rev. Comp. Chem. 14 (2000)), canonical
HF orbitals so the Fock matrix is diagonal:

  D_abij t_ij^ab = <ab||ij> + 1/2 <ab||cd> t_ij^cd + 1/2 <kl||ij> t_kl^ab
                 + P(ij)P(ab) <kb||cj> t_ik^ac + 1/4 <kl||cd> t_ij^cd t_kl^ab
                 + P(ij) <kl||cd> t_ik^ac t_jl^bd
                 - 1/2 P(ij) <kl||cd> t_ik^dc t_lj^ab
                 - 1/2 P(ab) <kl||cd> t_lk^ac t_ij^db
"""
from __future__ import annotations

import math

import numpy as np


def _sum_of_three_squares(n: int) -> bool:
    """True if n = x^2 + y^2 + z^2 for integers x, y, z (brute force)."""
    r = int(math.isqrt(n)) + 1
    for x in range(r):
        for y in range(x, r):
            rest = n - x * x - y * y
            if rest < 0:
                break
            z = int(math.isqrt(rest))
            if z * z == rest:
                return True
    return False


def shell_energy(n_shells: int) -> int:
    """Largest n^2 when n_shells distinct values of n^2 (0 included) are kept."""
    vals = [0]
    k = 0
    while len(vals) < n_shells:
        k += 1
        if _sum_of_three_squares(k):
            vals.append(k)
    return vals[-1]


class DenseSystem:
    """A closed-shell system as dense spin-orbital arrays, holes first."""

    def __init__(self, eps0, vfun, nh, labels):
        self.eps0 = np.asarray(eps0, dtype=np.float64)
        self.nsp = len(self.eps0)
        self.nh = nh
        self.labels = labels            # per s.p. state, a hashable quantum-number tuple
        self._v = vfun                   # vfun(P, Q, R, S) -> dense block over index lists
        h = np.arange(nh)
        p = np.arange(nh, self.nsp)
        self.h, self.p = h, p
        self.v_hhhh = vfun(h, h, h, h)
        self.v_hhpp = vfun(h, h, p, p)
        self.v_pppp = vfun(p, p, p, p)
        self.v_hpph = vfun(h, p, p, h)
        self.v_ppHH = vfun(p, p, h, h)
        # normal-ordered (Fock) diagonal: f_p = e_p + sum_j <pj||pj>
        allp = np.arange(self.nsp)
        vpjpj = vfun(allp, h, allp, h)            # [p, j, q, k]
        f = np.einsum("pjqj->pq", vpjpj, optimize=True)
        off = f - np.diag(np.diag(f))
        self.fock_offdiag = float(np.abs(off).max()) if off.size else 0.0
        self.f = self.eps0 + np.diag(f)
        self.eref = float(self.eps0[:nh].sum() + 0.5 * np.einsum("ijij->", self.v_hhhh))
        fh, fp = self.f[:nh], self.f[nh:]
        self.denom = (fh[None, None, :, None] + fh[None, None, None, :]
                      - fp[:, None, None, None] - fp[None, :, None, None])

    def mbpt2(self) -> float:
        t = self.v_ppHH / self.denom
        return float(0.25 * np.einsum("ijab,abij->", self.v_hhpp, t))

    def energy(self, t) -> float:
        return float(0.25 * np.einsum("ijab,abij->", self.v_hhpp, t))

    # ---- pieces, so the oracle's diagrams can be checked one by one ----
    def ladder(self, t):
        """1/2 <ab||cd> t_ij^cd + 1/2 t_kl^ab (<kl||ij> + 1/2 <kl||cd> t_ij^cd)."""
        x = self.v_hhhh + 0.5 * np.einsum("klcd,cdij->klij", self.v_hhpp, t, optimize=True)
        return (0.5 * np.einsum("abcd,cdij->abij", self.v_pppp, t, optimize=True)
                + 0.5 * np.einsum("abkl,klij->abij", t, x, optimize=True))

    def ring_unpermuted(self, t):
        """f(a,i,j,b) = sum_kc t_ik^ac (<kb||cj> + 1/2 sum_ld <kl||cd> t_lj^db)."""
        x = self.v_hpph + 0.5 * np.einsum("klcd,dblj->kbcj", self.v_hhpp, t, optimize=True)
        return np.einsum("acik,kbcj->aijb", t, x, optimize=True)

    def onebody(self, t):
        """chi_i = sum_jab <ij||ab> t_ij^ab, psi_a = sum_bij <ij||ab> t_ij^ab."""
        w = np.einsum("ijab,abij->abij", self.v_hhpp, t, optimize=True)
        return w.sum(axis=(0, 1, 3)), w.sum(axis=(1, 2, 3))

    def residual(self, t):
        r = self.v_ppHH.copy()
        r += self.ladder(t)
        f = self.ring_unpermuted(t)                       # [a, i, j, b]
        ring = np.einsum("aijb->abij", f, optimize=True)
        r += (ring - ring.transpose(0, 1, 3, 2) - ring.transpose(1, 0, 2, 3)
              + ring.transpose(1, 0, 3, 2))
        # full (not assumed diagonal) one-body intermediates
        xh = np.einsum("klcd,dcik->li", self.v_hhpp, t, optimize=True)       # X_li
        th = -0.5 * np.einsum("li,ablj->abij", xh, t, optimize=True)
        r += th - th.transpose(0, 1, 3, 2)
        xp = np.einsum("klcd,aclk->da", self.v_hhpp, t, optimize=True)       # X_da
        tp = -0.5 * np.einsum("da,dbij->abij", xp, t, optimize=True)
        r += tp - tp.transpose(1, 0, 2, 3)
        return r

    def solve(self, tol, mixing=0.5, max_iters=100):
        """Reference protocol; returns (energies per iteration, stop reason)."""
        t = np.zeros_like(self.v_ppHH)
        energies = []
        ediff, prev = 1.0, 100.0
        it = 0
        stop = "converged"
        while abs(ediff) > tol:
            if it > max_iters:
                stop = "max_iterations"
                break
            tn = self.residual(t) / self.denom
            t = mixing * tn + (1.0 - mixing) * t
            e = self.energy(t)
            ediff, prev = e - prev, e
            energies.append(e)
            it += 1
        self.t = t
        return energies, stop


def pairing(n_levels: int, n_particles: int, g: float, xi: float) -> DenseSystem:
    """Pairing Hamiltonian H = xi sum_p (p-1) n_p - g/2 sum_pq P+_p P_q.

    Two spin orbitals (level p, spin +/-) per level, lowest n_particles/2
    levels doubly occupied.  The antisymmetrised element follows from the
    pair operator: <p+ p-||q+ q-> = -g/2, with the sign of each swap.
    """
    lev = np.repeat(np.arange(1, n_levels + 1), 2)
    spin = np.tile([1, -1], n_levels)
    eps0 = (lev - 1.0) * xi

    def vfun(P, Q, R, S):
        lp, lq, lr, ls = (lev[x] for x in (P, Q, R, S))
        sp, sq, sr, ss = (spin[x] for x in (P, Q, R, S))
        bra = (lp[:, None] == lq[None, :]) & (sp[:, None] == -sq[None, :])
        ket = (lr[:, None] == ls[None, :]) & (sr[:, None] == -ss[None, :])
        same = sp[:, None, None, None] == sr[None, None, :, None]
        val = np.where(same, -0.5 * g, 0.5 * g)
        return np.where(bra[:, :, None, None] & ket[None, None, :, :], val, 0.0)

    labels = list(zip(lev.tolist(), spin.tolist()))
    return DenseSystem(eps0, vfun, n_particles, labels)


# Electron-gas constants of src/ElectronGasSPBasis.cpp:178-190 (Hartree, Angstrom).
def heg_constants(r_s: float, n_particles: int):
    hbarc = 0.1973269788 * 10000 / 27.21138505
    massc2 = 0.5109989461 * 1000000 / 27.21138505
    alpha = 1.0 / 137.035999139
    r_bohr = hbarc / (massc2 * alpha)
    length = (4.0 / 3.0 * math.pi * n_particles) ** (1.0 / 3.0) * r_s * r_bohr
    kin = hbarc * hbarc / (2.0 * massc2) * 4 * math.pi ** 2 / length ** 2
    return length, alpha * hbarc, kin


def electron_gas(r_s: float, n_shells: int, n_particle_shells: int) -> DenseSystem:
    """Plane waves in a periodic cube, n^2 <= shell_energy(n_shells), two spins;
    holes fill n^2 <= shell_energy(n_particle_shells).  Coulomb interaction
    4 pi e^2 / (L^3 q^2) with the q = 0 term dropped."""
    emax, ef = shell_energy(n_shells), shell_energy(n_particle_shells)
    r = int(math.isqrt(emax)) + 1
    ks = [(x, y, z) for x in range(-r, r + 1) for y in range(-r, r + 1)
          for z in range(-r, r + 1) if x * x + y * y + z * z <= emax]
    states = [(k, s) for k in ks for s in (-1, 1)]
    states.sort(key=lambda st: (st[0][0] ** 2 + st[0][1] ** 2 + st[0][2] ** 2 > ef,))
    nh = sum(1 for k, s in states if k[0] ** 2 + k[1] ** 2 + k[2] ** 2 <= ef)
    n = np.array([k for k, s in states], dtype=np.int64)
    sz = np.array([s for k, s in states], dtype=np.int64)
    length, e2, kin = heg_constants(r_s, nh)
    eps0 = kin * (n * n).sum(axis=1)

    def coul(dq):                       # dq [..., 3] integer momentum transfer
        q2 = (dq * dq).sum(axis=-1) * (2 * math.pi / length) ** 2
        with np.errstate(divide="ignore"):
            val = e2 / length ** 3 * 4 * math.pi / q2
        return np.where((dq == 0).all(axis=-1), 0.0, val)

    def vfun(P, Q, R, S):
        nP, nQ, nR, nS = (n[x] for x in (P, Q, R, S))
        sP, sQ, sR, sS = (sz[x] for x in (P, Q, R, S))
        ktot = (nP[:, None, None, None, :] + nQ[None, :, None, None, :]
                - nR[None, None, :, None, :] - nS[None, None, None, :, :])
        cons = (ktot == 0).all(axis=-1)
        d1 = nP[:, None, None, None, :] - nR[None, None, :, None, :]
        d2 = nP[:, None, None, None, :] - nS[None, None, None, :, :]
        s1 = (sP[:, None, None, None] == sR[None, None, :, None]) & (sQ[None, :, None, None] == sS[None, None, None, :])
        s2 = (sP[:, None, None, None] == sS[None, None, None, :]) & (sQ[None, :, None, None] == sR[None, None, :, None])
        d1 = np.broadcast_to(d1, cons.shape + (3,))
        d2 = np.broadcast_to(d2, cons.shape + (3,))
        val = np.where(s1, coul(d1), 0.0) - np.where(s2, coul(d2), 0.0)
        return np.where(cons, val, 0.0)

    labels = [(tuple(k), s) for k, s in states]
    return DenseSystem(eps0, vfun, nh, labels)
