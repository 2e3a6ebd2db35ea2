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


SERVE_LATTICES = [(L,) for L in range(8, 15)] + [(3, 3), (4, 3)]


def serving_requests(n: int = 200, rate: float = 20.0, seed: int = 5, *, ground_steps: int = 150,
                     kpm_moments: int = 256, kpm_nvec: int = 4, kpm_bound_steps: int = 40,
                     krylov_dim: int = 30, dt: float = 0.05, t_max: float = 5.0,
                     lattices=None, mix=(0.7, 0.2, 0.1)):
    """Poisson request stream at `rate` requests/s: 70 % ground state (FP64
    Lanczos, 150 steps), 20 % KPM density of states (256 moments, FP32 H.v,
    4 random-phase vectors, bounds from 40 FP64 Lanczos steps), 10 % Krylov
    exp(-iHt) (t ~ U[0, 5], dt = 0.05, Krylov dimension 30).  Lattice uniform
    over the periodic 1D chains L = 8..14 and the 3 x 3 and 4 x 3 squares;
    N_up, N_dn ~ U{1..L/2}; U ~ U[0, 12]; t = 1; per-request seeds.  Returns
    Request objects with arrival times (seconds from the start)."""
    from . import REQ_EVOLVE, REQ_GROUND, REQ_KPM, Request
    rng = np.random.default_rng(seed)
    lattices = SERVE_LATTICES if lattices is None else lattices
    t = 0.0
    out = []
    for i in range(n):
        t += float(rng.exponential(1.0 / rate))
        sizes = lattices[int(rng.integers(len(lattices)))]
        L = int(np.prod(sizes))
        nu = int(rng.integers(1, L // 2 + 1))
        nd = int(rng.integers(1, L // 2 + 1))
        U = float(rng.uniform(0.0, 12.0))
        x = float(rng.uniform())
        common = dict(sizes=tuple(sizes), n_up=nu, n_dn=nd, t=1.0, u=U, seed=1000 + i, arrival=t)
        if x < mix[0]:
            out.append(Request(REQ_GROUND, steps=ground_steps, **common))
        elif x < mix[0] + mix[1]:
            out.append(Request(REQ_KPM, steps=kpm_bound_steps, moments=kpm_moments, nvec=kpm_nvec,
                               **common))
        else:
            out.append(Request(REQ_EVOLVE, steps=krylov_dim, t_final=float(rng.uniform(0.0, t_max)),
                               dt=dt, **common))
    return out
