"""Synthetic workloads of the benchmark configurations (instance lists only;
vectors are generated on the device from per-instance seeds)."""
from __future__ import annotations

import numpy as np


def batch_1d_instances(n: int = 512, seed: int = 4, L_range=(8, 12), U_max: float = 12.0):
    """n independent 1D periodic Hubbard rings: L ~ U{8..12}, N_up and N_dn ~
    U{1..L/2}, U ~ U[0, 12], t = 1.  Returns a list of (L, N_up, N_dn, U)."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        L = int(rng.integers(L_range[0], L_range[1] + 1))
        nu = int(rng.integers(1, L // 2 + 1))
        nd = int(rng.integers(1, L // 2 + 1))
        U = float(rng.uniform(0.0, U_max))
        out.append((L, nu, nd, U))
    return out
