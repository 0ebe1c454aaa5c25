"""The seeded parameter sweep: 64 independent CCD solves, half electron gas
and half pairing model, run as one batch (one launch per kernel per
iteration).

Electron gas: N in {14, 38, 54} with its basis n^2 <= {20, 28, 36}
(reference flags -p {2, 3, 4} and -s {19, 25, 32}), r_s uniform in
[0.5, 5.0].  Pairing: 8-16 doubly degenerate levels at half filling (an even
number of levels, so the particles pair up), spacing 1, g uniform in
[0.2, 0.8].  Solves alternate electron gas / pairing; the draws come from
numpy.random.default_rng(seed) in that order.
"""
from __future__ import annotations

import numpy as np

SEED = 20261203
# (particles, -s shell index, -p Fermi shells) of the electron-gas bases
HEG_BASES = ((14, 19, 2), (38, 25, 3), (54, 32, 4))


def sweep_configs(seed: int = SEED, n: int = 64):
    """[(kind, args)] with kind 'heg' (r_s, n_shells, n_particle_shells) or
    'pairing' (n_levels, n_particles, g, xi)."""
    rng = np.random.default_rng(seed)
    out = []
    for q in range(n):
        if q % 2 == 0:
            _, s, p = HEG_BASES[int(rng.integers(len(HEG_BASES)))]
            out.append(("heg", (float(rng.uniform(0.5, 5.0)), s, p)))
        else:
            levels = 8 + 2 * int(rng.integers(5))
            out.append(("pairing", (levels, levels, float(rng.uniform(0.2, 0.8)), 1.0)))
    return out


def sweep_models(seed: int = SEED, n: int = 64):
    from . import Model
    return [Model.electron_gas(*a) if k == "heg" else Model.pairing(*a) for k, a in sweep_configs(seed, n)]
