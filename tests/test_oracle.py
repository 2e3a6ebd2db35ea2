"""Pins the oracle (oracle/hubbard_oracle.c) before it is trusted as the checker.

Each oracle function is checked against something that can catch a plausible
mistake in it:
  basis order            golden vectors written by the compiled reference (146 sectors)
  CSC matrix build       golden matrices from the compiled reference (28 models: 1D/2D/3D,
                         merged double bonds, self bonds, non-integer t/U), Hermiticity,
                         U = 0 free-fermion closed forms and t = 0 energies beyond N = 3
  product-order H.v      the CSC build through the DFS permutation, bit for bit
  Lanczos                scipy eigsh, ports of test_lanczos.cpp known answers
  tridiagonal/Jacobi     Toeplitz closed form, numpy
  Krylov expm            scipy expm, port of test_evolve_state.cpp
  KPM moments            exact spectral sum from a dense eigendecomposition
  Jackson kernel         port of test_kpm.cpp properties, closed form
"""
import hashlib
import json
import os

import numpy as np
import pytest
import scipy.linalg as sla
import scipy.sparse.linalg as spla

from oracle import pyoracle as po

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def _manifest():
    with open(os.path.join(GOLD, "manifest.json")) as f:
        return json.load(f)


def csc_from_dense(M):
    import scipy.sparse as sp
    A = sp.csc_matrix(np.asarray(M, np.complex128))
    A.sort_indices()
    return po.CSC(A.shape[0], A.indptr.astype(np.int64), A.indices.astype(np.int64),
                  A.data.astype(np.complex128))


def free_fermion_e0(sizes, n_up, n_dn, t=1.0):
    """Sum of the lowest single-particle levels of -t * (periodic nearest-neighbour
    hopping, merged bonds included), per spin."""
    L = int(np.prod(sizes))
    h = np.zeros((L, L))
    for f, to in po.hubbard_bonds(sizes):
        h[f, to] += -t
        h[to, f] += -t
    eps = np.sort(np.linalg.eigvalsh(h))
    return eps[:n_up].sum() + eps[:n_dn].sum()


# ------------------------------------------------------------------ basis

def test_choose_matches_math():
    import math
    for n in range(0, 33):
        for m in range(0, n + 2):
            assert po.choose(n, m) == (math.comb(n, m) if m <= n else 0)


def test_basis_order_matches_reference_golden():
    man = _manifest()
    stored = np.load(os.path.join(GOLD, "ref_basis.npz"))
    assert len(man["basis"]) == 146
    for e in man["basis"]:
        up, dn = po.basis_states(e["L"], e["n_up"], e["n_dn"])
        st = np.stack([up, dn], axis=1)
        assert len(st) == e["dim"]
        assert hashlib.sha256(np.ascontiguousarray(st, np.uint64).tobytes()).hexdigest() == e["sha256"], e["name"]
        if e.get("stored"):
            assert np.array_equal(st, stored[e["name"]])


def test_basis_small_example_order():
    # L = 2, one up, one down: (up0 dn0), (up0 dn1), (up1 dn1), (dn0 up1)
    up, dn = po.basis_states(2, 1, 1)
    assert list(zip(up.tolist(), dn.tolist())) == [(1, 1), (1, 2), (2, 2), (2, 1)]


def test_basis_size_formula_and_membership():
    for L, a, b in [(8, 4, 4), (6, 3, 1), (10, 5, 5), (9, 0, 4)]:
        up, dn = po.basis_states(L, a, b)
        assert len(up) == po.choose(L, a) * po.choose(L, b)
        assert all(bin(int(u)).count("1") == a for u in up[:200])
        assert len(set(zip(up.tolist(), dn.tolist()))) == len(up)


# ------------------------------------------------------------------ matrix

def test_csc_matches_reference_golden():
    man = _manifest()
    H = np.load(os.path.join(GOLD, "ref_hamiltonian.npz"))
    assert len(man["hamiltonian"]) == 28
    for e in man["hamiltonian"]:
        n = e["name"]
        A = po.build_csc(e["dims"], e["n_up"], e["n_dn"], e["t"], e["U"])
        M = A.to_scipy().toarray()
        R = np.zeros_like(M)
        ij = H[n + "__ij"]
        R[ij[:, 0], ij[:, 1]] = H[n + "__val"]
        assert np.array_equal(M, R), n


@pytest.mark.parametrize("sizes,nu,nd,U", [([8], 4, 4, 4.0), ([3, 3], 3, 3, 8.0), ([2, 2, 2], 3, 2, 5.0)])
def test_csc_hermitian_beyond_reference_range(sizes, nu, nd, U):
    A = po.build_csc(sizes, nu, nd, 1.0, U).to_scipy()
    assert abs(A - A.getH()).max() == 0


@pytest.mark.parametrize("sizes,nu,nd", [([8], 4, 4), ([10], 5, 5), ([3, 3], 4, 3), ([4, 2], 3, 3),
                                         ([2, 2, 2], 2, 3), ([4, 4], 2, 2)])
def test_u0_ground_state_is_free_fermion_sum(sizes, nu, nd):
    A = po.build_csc(sizes, nu, nd, 1.0, 0.0).to_scipy().real
    e0 = spla.eigsh(A, k=1, which="SA", tol=1e-12)[0][0]
    assert abs(e0 - free_fermion_e0(sizes, nu, nd)) < 1e-9


def test_u0_closed_form_c1():
    # 1D L = 8, 4 + 4, U = 0: E0 = 2 * sum of the lowest four -2 cos(2 pi k / 8)
    A = po.build_csc([8], 4, 4, 1.0, 0.0).to_scipy().real
    e0 = spla.eigsh(A, k=1, which="SA", tol=1e-13)[0][0]
    assert abs(e0 - (-9.656854249492380)) < 1e-10


@pytest.mark.parametrize("sizes,nu,nd,U", [([6], 3, 3, 4.0), ([6], 4, 3, 2.5), ([3, 3], 5, 5, 8.0)])
def test_t0_energy_counts_double_occupancy(sizes, nu, nd, U):
    A = po.build_csc(sizes, nu, nd, 0.0, U).to_scipy()
    L = int(np.prod(sizes))
    d = A.diagonal().real
    assert A.nnz == np.count_nonzero(d)
    assert abs(d.min() - U * max(0, nu + nd - L)) < 1e-12


# --------------------------------------------------------- product order

@pytest.mark.parametrize("sizes,nu,nd,t,U", [([8], 4, 4, 1.0, 4.0), ([3, 3], 4, 3, 1.0, 8.0),
                                             ([4, 2], 3, 2, 0.5, 3.0), ([2, 2, 2], 3, 3, 1.0, 4.0),
                                             ([3, 1], 2, 1, 1.0, 2.0), ([10], 3, 2, 1.0, 6.0),
                                             ([2, 3], 3, 3, -1.0, 5.0)])
def test_product_order_oracle_matches_csc_bitwise(sizes, nu, nd, t, U):
    A = po.build_csc(sizes, nu, nd, t, U)
    P = po.ProductOperator(sizes, nu, nd, t, U)
    perm = po.product_permutation(sizes, nu, nd)
    rng = np.random.default_rng(7)
    x_ref = (rng.integers(-4, 5, A.D) + 1j * rng.integers(-4, 5, A.D)).astype(np.complex128)
    y_ref = A.matvec(x_ref)
    x_dev = np.zeros_like(x_ref)
    x_dev[perm] = x_ref
    y_dev = P.matvec(x_dev)
    assert np.array_equal(y_dev[perm], y_ref)


# ------------------------------------------------------------------ Lanczos

def test_lanczos_oracle_energy_matches_eigsh_c1():
    A = po.build_csc([8], 4, 4, 1.0, 4.0)
    rng = np.random.default_rng(1)
    b = rng.standard_normal(A.D) + 1j * rng.standard_normal(A.D)
    al, be, bn = po.lanczos_csc(A, b, 200)
    w = po.tridiag_eigvals(al, be)
    e_ref = spla.eigsh(A.to_scipy().real, k=1, which="SA", tol=1e-14)[0][0]
    assert abs(w[0] - e_ref) / abs(e_ref) < 1e-10
    assert abs(bn - np.linalg.norm(b)) < 1e-12


def test_lanczos_negated_mirrors_alphas():
    A = po.build_csc([6], 3, 2, 1.0, 3.0)
    rng = np.random.default_rng(2)
    b = rng.standard_normal(A.D) + 1j * rng.standard_normal(A.D)
    al, be, _ = po.lanczos_csc(A, b, 30)
    aln, ben, _ = po.lanczos_csc(A, b, 30, negate=True)
    assert np.array_equal(aln, -al) and np.array_equal(ben, be)


def test_lanczos_diagonal_known_answer():
    # test_lanczos.cpp: DiagonalOperator {1, 3, -2}: max 3, min -2
    A = csc_from_dense(np.diag([1.0, 3.0, -2.0]))
    b = np.array([0.3, 0.5, 0.7], np.complex128)
    val, x = po.find_min_eigenpair_csc(A, b, 3)
    assert abs(val + 2.0) < 1e-6
    Ax = A.matvec(x)
    assert abs(np.vdot(x, Ax).real / np.vdot(x, x).real - val) < 1e-6


def test_lanczos_chain_known_answer():
    # test_lanczos.cpp:221-236: 80-site chain, diag 2, hopping -1, 50 steps
    n = 80
    M = 2.0 * np.eye(n) - np.eye(n, k=1) - np.eye(n, k=-1)
    A = csc_from_dense(-M)  # max eigenpair of M = -(min of -M)
    b = np.ones(n, np.complex128) + 0.1 * np.arange(n)
    val, _ = po.find_min_eigenpair_csc(A, b, 50)
    exact_max = 2.0 - 2.0 * np.cos(n * np.pi / (n + 1.0))
    assert abs(-val - exact_max) < 2e-2


def test_lanczos_breakdown_and_zero_norm():
    A = csc_from_dense(np.diag([1.0, 2.0, 3.0, 4.0]))
    b = np.array([1.0, 0, 0, 0], np.complex128)  # eigenvector: breakdown after one step
    al, be, bn = po.lanczos_csc(A, b, 4)
    assert len(al) == 1 and len(be) == 0 and al[0] == 1.0
    with pytest.raises(RuntimeError):
        po.lanczos_csc(A, np.zeros(4, np.complex128), 4)


# ------------------------------------------------------- small eigensolvers

def test_tridiag_eigvals_toeplitz_closed_form():
    n, a, b = 40, 1.5, -0.7
    w = po.tridiag_eigvals(np.full(n, a), np.full(n - 1, b))
    exact = np.sort(a + 2 * b * np.cos(np.arange(1, n + 1) * np.pi / (n + 1)))
    assert np.abs(w - exact).max() < 1e-12


def test_tridiag_and_jacobi_match_numpy():
    rng = np.random.default_rng(3)
    a = rng.standard_normal(25)
    b = rng.standard_normal(24)
    T = np.diag(a) + np.diag(b, 1) + np.diag(b, -1)
    ref = np.linalg.eigvalsh(T)
    assert np.abs(po.tridiag_eigvals(a, b) - ref).max() < 1e-12
    w, V = po.jacobi_eig(T)
    assert np.abs(w - ref).max() < 1e-12
    assert np.abs(T @ V - V * w).max() < 1e-11


# ------------------------------------------------------------------ Krylov

def test_time_evolve_matches_expm_full_krylov():
    A = po.build_csc([4], 2, 2, 1.0, 3.0)
    H = A.to_scipy().toarray()
    rng = np.random.default_rng(4)
    psi = rng.standard_normal(A.D) + 1j * rng.standard_normal(A.D)
    out = po.time_evolve_csc(A, psi, 0.37, k=0)
    ref = sla.expm(-1j * 0.37 * H) @ psi
    assert np.linalg.norm(out - ref) / np.linalg.norm(ref) < 1e-10


def test_time_evolve_two_level_known_answer():
    # test_evolve_state.cpp: H = diag(1, 2), psi0 = (1, i), t = 0.3 (three steps of 0.1)
    A = csc_from_dense(np.diag([1.0, 2.0]))
    psi = np.array([1.0, 1j])
    for _ in range(3):
        psi = po.time_evolve_csc(A, psi, 0.1, k=2)
    ref = np.array([np.exp(-0.3j), 1j * np.exp(-0.6j)])
    assert np.linalg.norm(psi - ref) / np.linalg.norm(ref) < 1e-8


# ------------------------------------------------------------------ KPM

def test_kpm_moments_match_spectral_sum():
    A = po.build_csc([6], 2, 2, 1.0, 4.0)
    H = A.to_scipy().toarray().real
    lam, Q = np.linalg.eigh(H)
    a, b = (lam[-1] - lam[0]) / 2 * 1.05, (lam[-1] + lam[0]) / 2
    rng = np.random.default_rng(5)
    R, M = 3, 40
    vecs = np.exp(2j * np.pi * rng.random((R, A.D)))
    mu = po.kpm_moments_csc(A, a, b, M, vecs)
    x = (lam - b) / a
    w = (np.abs(Q.T @ vecs.T) ** 2).sum(axis=1)
    ref = np.array([(w * np.cos(n * np.arccos(x))).sum() for n in range(M)]) / (R * A.D)
    assert np.abs(mu - ref).max() < 1e-12
    assert abs(mu[0] - 1.0) < 1e-14


def test_product_order_kpm_and_evolution_match_spectral_forms():
    """The product-order KPM moments and stepped Krylov evolution (used at sizes
    where no CSC fits) against the spectral sum and expm of the dense H, through
    the DFS permutation: 2D 3 x 2, 2 + 2, U = 3."""
    sizes, nu, nd, t, U = [3, 2], 2, 2, 1.0, 3.0
    A = po.build_csc(sizes, nu, nd, t, U)
    P = po.ProductOperator(sizes, nu, nd, t, U)
    perm = po.product_permutation(sizes, nu, nd)
    H = A.to_scipy().toarray()
    lam, Q = np.linalg.eigh(H)
    a, b = (lam[-1] - lam[0]) / 2 * 1.05, (lam[-1] + lam[0]) / 2
    rng = np.random.default_rng(9)
    R, M = 2, 48
    vecs_ref = np.exp(2j * np.pi * rng.random((R, A.D)))
    vecs_dev = np.zeros_like(vecs_ref)
    vecs_dev[:, perm] = vecs_ref
    mu = P.kpm_moments(a, b, M, vecs_dev)
    x = (lam - b) / a
    w = (np.abs(Q.conj().T @ vecs_ref.T) ** 2).sum(axis=1)
    ref = np.array([(w * np.cos(n * np.arccos(x))).sum() for n in range(M)]) / (R * A.D)
    assert np.abs(mu - ref).max() < 1e-12
    psi_ref = rng.standard_normal(A.D) + 1j * rng.standard_normal(A.D)
    psi_dev = np.zeros_like(psi_ref)
    psi_dev[perm] = psi_ref
    out = P.time_evolve_steps(psi_dev, 0.0, 0.33, 0.1, k=30)  # 0.1, 0.1, 0.1, 0.03
    exact = sla.expm(-1j * 0.33 * H) @ psi_ref
    assert np.linalg.norm(out[perm] - exact) / np.linalg.norm(exact) < 1e-10
    assert np.array_equal(P.time_evolve_steps(psi_dev, 0.0, 0.0, 0.1), psi_dev)


def test_jackson_kernel_properties_and_closed_form():
    g = po.jackson_kernel(50)
    assert abs(g[0] - 1.0) < 1e-10
    assert np.all(g >= -1e-10) and np.all(g <= 1 + 1e-10)
    assert np.all(np.diff(g) <= 1e-10)
    N = 50
    n = np.arange(N + 1)
    ref = ((N + 1 - n) * np.cos(np.pi * n / (N + 1)) + np.sin(np.pi * n / (N + 1)) / np.tan(np.pi / (N + 1))) / (N + 1)
    assert np.abs(g - ref).max() < 1e-13


# ------------------------------------------------------------------ correlation

def test_opposite_spin_correlation_matches_reference_golden():
    """G_updn(r) of the compiled reference (every displacement, 1D/2D/3D, its
    exact range N <= 3): the oracle's diagonal equals it exactly."""
    gold = np.load(os.path.join(GOLD, "ref_correlation.npz"))
    for c in _manifest()["correlation"]:
        d = po.opposite_spin_correlation_diag(c["dims"], c["n_up"], c["n_dn"])
        up, dn = po.basis_states(int(np.prod(c["dims"])), c["n_up"], c["n_dn"])
        st = gold[c["name"] + "__states"]
        assert np.array_equal(st[:, 0], np.asarray(up, np.uint64)), c["name"]
        assert np.array_equal(st[:, 1], np.asarray(dn, np.uint64)), c["name"]
        assert np.array_equal(d, gold[c["name"] + "__diag"]), c["name"]


@pytest.mark.parametrize("sizes,nu,nd", [([8], 4, 3), ([3, 3], 4, 4), ([2, 2, 2], 3, 2)])
def test_opposite_spin_correlation_sum_rule(sizes, nu, nd):
    """sum_r G(r) = N_up N_dn / N for every state (sum_r n_{i+r,dn} = N_dn)."""
    N = int(np.prod(sizes))
    d = po.opposite_spin_correlation_diag(sizes, nu, nd)
    assert np.allclose(d.sum(axis=0), nu * nd / N, rtol=0, atol=1e-14)
    x = np.random.default_rng(1).normal(size=d.shape[1])
    g = po.opposite_spin_correlation(sizes, nu, nd, x)
    assert abs(g.sum() - nu * nd / N) < 1e-12


# ------------------------------------------------------------ momentum sectors

def test_momentum_basis_matches_reference_golden():
    """Fixed-momentum bases of the compiled reference (every K of ten 1D/2D/3D
    lattices): the oracle's filtered DFS order equals them, and the sectors of
    one lattice partition the real-space basis."""
    gold = np.load(os.path.join(GOLD, "ref_momentum.npz"))
    totals = {}
    for c in _manifest()["momentum"]:
        if c["kind"] != "basis":
            continue
        up, dn = po.momentum_sector_states(c["dims"], c["n_up"], c["n_dn"], c["K"])
        st = gold[c["name"] + "__states"]
        assert np.array_equal(st[:, 0], np.asarray(up, np.uint64)), c["name"]
        assert np.array_equal(st[:, 1], np.asarray(dn, np.uint64)), c["name"]
        key = (tuple(c["dims"]), c["n_up"], c["n_dn"])
        totals[key] = totals.get(key, 0) + len(up)
    for (dims, nu, nd), tot in totals.items():
        N = int(np.prod(dims))
        assert tot == po.choose(N, nu) * po.choose(N, nd)


def test_momentum_hamiltonian_matches_reference_golden():
    """The oracle's factorized operator (rho_{-q,dn} then rho_{q,up}) against the
    reference's symbolic HubbardModelMomentum matrix, every element (1e-12)."""
    gold = np.load(os.path.join(GOLD, "ref_momentum.npz"))
    n = 0
    for c in _manifest()["momentum"]:
        if c["kind"] != "hamiltonian":
            continue
        H = po.momentum_hamiltonian(c["dims"], c["n_up"], c["n_dn"], c["K"], c["t"], c["U"])
        G = np.zeros_like(H)
        ij, val = gold[c["name"] + "__ij"], gold[c["name"] + "__val"]
        G[ij[:, 0], ij[:, 1]] = val
        assert H.shape == (c["dim"], c["dim"]), c["name"]
        assert np.abs(H - G).max() < 1e-12, c["name"]
        n += 1
    assert n == 11


@pytest.mark.parametrize("sizes,nu,nd,U", [([6], 3, 2, 4.0), ([4], 2, 2, 6.0),
                                           ([3, 2], 2, 2, 8.0), ([2, 2, 2], 2, 2, 4.0),
                                           ([5], 2, 3, 3.0)])
def test_momentum_sectors_reproduce_real_space_spectrum(sizes, nu, nd, U):
    """The momentum-space operator is the real-space Hubbard model in another
    basis: the union of the spectra of all K sectors equals the real-space
    spectrum (pins the interaction's signs beyond the reference's exact range)."""
    N = int(np.prod(sizes))
    A = po.build_csc(sizes, nu, nd, 1.0, U).to_scipy().toarray()
    w_real = np.linalg.eigvalsh(A)
    w_k = np.concatenate([np.linalg.eigvalsh(po.momentum_hamiltonian(sizes, nu, nd, K, 1.0, U))
                          for K in range(N)])
    assert len(w_k) == len(w_real)
    assert np.abs(np.sort(w_k) - w_real).max() < 1e-10


# ------------------------------------------------------------ optical conductivity

def test_current_operator_matches_reference_golden():
    """J_d(Q) between every sector pair (K, K + Q) against the compiled reference's
    symbolic HubbardModelMomentum::current on the full basis, restricted to the
    sectors.  1e-12 absolute: where L_d = 2, v_d(k) = 2t pi sin(pi) ~ 1e-15, which
    CurrentOperatorFixedK keeps (|v| >= 1e-15, the oracle's rule) and the symbolic
    path zeroes (|coefficient| < 1000 eps); every other entry agrees exactly."""
    gold = np.load(os.path.join(GOLD, "ref_conductivity.npz"))
    n = 0
    for c in _manifest()["conductivity"]:
        if c["kind"] != "current":
            continue
        dims, nu, nd = c["dims"], c["n_up"], c["n_dn"]
        N = int(np.prod(dims))
        st = gold[c["name"] + "__states"]
        full = {(int(u), int(d)): i for i, (u, d) in enumerate(st)}
        G = np.zeros((len(st), len(st)), dtype=np.complex128)
        ij = gold[c["name"] + "__ij"]
        G[ij[:, 0], ij[:, 1]] = gold[c["name"] + "__val"]
        Qc = c["Q"] + [0] * (3 - len(dims))
        dd = list(dims) + [1] * (3 - len(dims))
        Q = Qc[0] + dd[0] * Qc[1] + dd[0] * dd[1] * Qc[2]
        covered = 0
        for K in range(N):
            J = po.current_operator(dims, nu, nd, K, Q, c["direction"], c["t"])
            su, sd = po.momentum_sector_states(dims, nu, nd, K)
            tu, td = po.momentum_sector_states(dims, nu, nd, po.momentum_add(dims, K, Q))
            rows = [full[s] for s in zip(tu, td)]
            cols = [full[s] for s in zip(su, sd)]
            assert np.abs(J - G[np.ix_(rows, cols)]).max() <= 1e-12, (c["name"], K)
            covered += len(cols)
        assert covered == len(st)
        n += 1
    assert n == 7


def test_optical_conductivity_matches_reference_golden():
    gold = np.load(os.path.join(GOLD, "ref_conductivity.npz"))
    n = 0
    for c in _manifest()["conductivity"]:
        if c["kind"] != "sigma":
            continue
        w, sg, sr = po.optical_conductivity(gold[c["name"] + "__corr"], c["dt"], c["beta"],
                                            c["volume"])
        assert np.abs(w - gold[c["name"] + "__omega"]).max() <= 1e-12 * w.max()
        ref = gold[c["name"] + "__sigma"]
        assert np.abs(sg - ref).max() <= 1e-12 * np.abs(ref).max()
        assert abs(sr - c["sum_rule"]) <= 1e-12 * max(abs(c["sum_rule"]), 1e-300)
        n += 1
    assert n == 4


def test_optical_conductivity_of_a_constant_correlator_is_a_geometric_sum():
    """C(t) = 1: the damped trapezoid sum is dt [(1 - r^n)/(1 - r) - (1 + r^(n-1))/2]
    with r = exp((i omega - eta) dt), in closed form."""
    n, dt, beta, vol = 64, 0.1, 3.0, 5.0
    w, sg, _ = po.optical_conductivity(np.ones(n), dt, beta, vol)
    eta = 4.0 / ((n - 1) * dt)
    r = np.exp((1j * w - eta) * dt)
    integral = dt * ((1 - r ** n) / (1 - r) - (1 + r ** (n - 1)) / 2)
    pref = np.where(w < 1e-12, beta / vol, -np.expm1(-beta * w) / np.where(w < 1e-12, 1, w) / vol)
    assert np.abs(sg - pref * integral).max() <= 1e-10 * np.abs(sg).max()


# ------------------------------------------------------------ generic expressions

def _golden_terms(gold, name):
    coef, nops, ops = gold[name + "__coef"], gold[name + "__nops"], gold[name + "__ops"]
    terms, k = [], 0
    for c, n in zip(coef, nops):
        terms.append((complex(c), [tuple(int(v) for v in ops[k + i]) for i in range(int(n))]))
        k += int(n)
    return terms


def test_expression_matrix_matches_reference_golden():
    """compute_matrix_elements of ten generic expressions (random Hermitian one-
    and two-body terms with spin flips, odd and non-conserving monomials,
    constants, merged, cancelling and sub-tolerance terms; real-space and
    fixed-momentum bases) against the compiled reference, and the basis order
    the driver printed against the oracle's enumeration."""
    gold = np.load(os.path.join(GOLD, "ref_expression.npz"))
    n = 0
    for c in _manifest()["expression"]:
        if c["kind"] not in ("expr", "mexpr", "nexpr"):
            continue
        dims, nu, nd = c["dims"], c["n_up"], c["n_dn"]
        L = int(np.prod(dims))
        st = gold[c["name"] + "__states"]
        if c["kind"] == "nexpr":  # with_fixed_particle_number(L, N = n_up)
            up, dn = po.fixed_n_states(L, nu)
        elif c["kind"] == "expr":
            up, dn = po.basis_states(L, nu, nd)
        else:
            Kc = c["K"] + [0] * (3 - len(dims))
            dd = dims + [1] * (3 - len(dims))
            up, dn = po.momentum_sector_states(dims, nu, nd,
                                               Kc[0] + dd[0] * (Kc[1] + dd[1] * Kc[2]))
        assert np.array_equal(np.stack([up, dn], 1).astype(np.uint64), st), c["name"]
        M = po.expression_matrix(L, st, _golden_terms(gold, c["name"]))
        R = np.zeros_like(M)
        ij = gold[c["name"] + "__ij"]
        R[ij[:, 0], ij[:, 1]] = gold[c["name"] + "__val"]
        assert np.count_nonzero(M) == len(ij), c["name"]
        assert np.abs(M - R).max() <= 1e-14 * max(1.0, np.abs(R).max()), c["name"]
        n += 1
    assert n == 14


def test_cereal_bytes_match_reference_golden():
    """The byte layout of arma_cereal.h's save through cereal::BinaryOutputArchive,
    against archives the vendored cereal wrote (empty, small and 37 x 37)."""
    gold = np.load(os.path.join(GOLD, "ref_expression.npz"))
    n = 0
    for c in _manifest()["expression"]:
        if c["kind"] != "cereal":
            continue
        b = po.cereal_sparse_bytes(c["n_rows"], c["n_cols"], gold[c["name"] + "__rows"],
                                   gold[c["name"] + "__cols"], gold[c["name"] + "__vals"])
        assert b == gold[c["name"] + "__bytes"].tobytes(), c["name"]
        assert len(b) == 28 + 32 * c["nnz"]
        n += 1
    assert n == 3


def test_expression_one_body_commutator_identity():
    """[c+_a c_b, c+_b c_a] = n_a - n_b within a fixed (N_up, N_dn) sector (same
    spin, so the products stay in the sector): pins the fermionic signs beyond
    the reference's 7-operator range (L = 7, 3 + 3, every pair)."""
    L, nu, nd = 7, 3, 3
    up, dn = po.basis_states(L, nu, nd)
    st = list(zip(up, dn))
    for s in (0, 1):
        for a in range(L):
            for b in range(L):
                if a == b:
                    continue
                A = po.expression_matrix(L, st, [(1.0, [(0, s, a), (1, s, b)])])
                B = po.expression_matrix(L, st, [(1.0, [(0, s, b), (1, s, a)])])
                N = po.expression_matrix(L, st, [(1.0, [(0, s, a), (1, s, a)]),
                                                 (-1.0, [(0, s, b), (1, s, b)])])
                assert np.array_equal(A @ B - B @ A, N)


def test_expression_two_body_reduces_to_one_body_products():
    """c+_a c+_b c_c c_d = (c+_a c_d)(c+_b c_c) - delta_bd c+_a c_c (anticommutation,
    all one spin), on L = 6, 3 + 2 for random index sets including coincident ones."""
    L, nu, nd = 6, 3, 2
    up, dn = po.basis_states(L, nu, nd)
    st = list(zip(up, dn))
    rng = np.random.default_rng(3)
    for _ in range(40):
        s = int(rng.integers(2))
        a, b, c, d = (int(v) for v in rng.integers(L, size=4))
        T = po.expression_matrix(L, st, [(1.0, [(0, s, a), (0, s, b), (1, s, c), (1, s, d)])])
        P = (po.expression_matrix(L, st, [(1.0, [(0, s, a), (1, s, d)])])
             @ po.expression_matrix(L, st, [(1.0, [(0, s, b), (1, s, c)])]))
        if b == d:
            P = P - po.expression_matrix(L, st, [(1.0, [(0, s, a), (1, s, c)])])
        assert np.array_equal(T, P), (s, a, b, c, d)


@pytest.mark.parametrize("sizes,nu,nd,t,U", [([6], 3, 3, 1.0, 4.0), ([3, 2], 3, 2, 0.7, 6.0),
                                             ([2, 2, 2], 2, 2, 1.0, 3.0), ([2], 1, 1, 1.0, 2.0)])
def test_hubbard_expression_equals_the_hubbard_oracle(sizes, nu, nd, t, U):
    """The Hubbard Hamiltonian written as an expression (bond loops of
    hubbard_model.h, merged bonds at size 2) through expression_matrix equals the
    oracle's Hubbard CSC (a separate construction), beyond the reference's range."""
    import manybody_amd.expression as ex
    L = int(np.prod(sizes))
    up, dn = po.basis_states(L, nu, nd)
    e = ex.hubbard_hamiltonian(t, U, sizes)
    terms = [(c, [(o >> 15, (o >> 14) & 1, o & 0x3FFF) for o in k]) for k, c in e.terms.items()]
    M = po.expression_matrix(L, list(zip(up, dn)), terms)
    R = po.build_csc(sizes, nu, nd, t, U).to_scipy().toarray()
    assert np.abs(M - R).max() <= 1e-13


# ------------------------------------------------------------ spectral bounds

def test_power_method_known_answer():
    """Port of test_linear_operator.cpp:128-135: diag(3, -1, 0.5), seed (1, 1, 1),
    40 iterations -> 3 within 1e-3."""
    d = np.array([3.0, -1.0, 0.5])
    assert abs(po.power_method(lambda v: d * v, np.ones(3), 40) - 3.0) < 1e-3
    assert po.power_method(lambda v: d * v, np.zeros(3), 40) == 0.0
    assert po.power_method(lambda v: d * v, np.ones(3), 0) == 0.0


def test_estimate_bounds_known_answer():
    """Port of test_linear_operator.cpp:137-144: diag(-2, 0.5, 4), 40 iterations,
    padding 0.2 -> (-2, 4) within 1e-3 (the reference seeds with randu; here a
    seeded U[0,1) complex vector)."""
    d = np.array([-2.0, 0.5, 4.0])
    rng = np.random.default_rng(3)
    seed = rng.random(3) + 1j * rng.random(3)
    a, b = po.estimate_bounds(lambda v: d * v, seed, 40, 0.2)
    assert abs(a + 2.0) < 1e-3 and abs(b - 4.0) < 1e-3


def test_power_method_unconjugated_dot_converges_for_real_symmetric():
    """With a complex seed the estimate Re(v.w)/Re(v.v) (no conjugation) still
    converges to the dominant eigenvalue of a real symmetric matrix; bounds
    match eigvalsh for a spectrum with clear gaps."""
    rng = np.random.default_rng(5)
    Q, _ = np.linalg.qr(rng.standard_normal((30, 30)))
    lam = np.concatenate([[-9.0], np.linspace(-3, 3, 28), [7.0]])
    A = (Q * lam) @ Q.T
    seed = rng.standard_normal(30) + 1j * rng.standard_normal(30)
    assert abs(po.power_method(lambda v: A @ v, seed, 200) + 9.0) < 1e-8
    a, b = po.estimate_bounds(lambda v: A @ v, seed, 400, 0.1)
    assert abs(a + 9.0) < 1e-6 and abs(b - 7.0) < 1e-6


# ------------------------------------------------------------ KPM projector

def test_projector_moments_and_lorentz_known_answers():
    """Ports of test_kpm.cpp:30-58: Lorentz kernel g_0 = 1 and 0 <= g <= 1;
    projector moment sum rules mu_0(0) = 1/2, mu_0(-1) = 0, mu_0(1) = 1."""
    g = po.lorentz_kernel(100)
    assert abs(g[0] - 1.0) < 1e-10 and np.all(g >= -1e-10) and np.all(g <= 1 + 1e-10)
    assert abs(po.projector_moments(50, 0.0)[0] - 0.5) < 1e-10
    assert abs(po.projector_moments(50, -1.0)[0]) < 1e-10
    assert abs(po.projector_moments(50, 1.0)[0] - 1.0) < 1e-10


def test_projector_moments_are_chebyshev_coefficients_of_the_step():
    """mu_m = (2 - delta_m0) / pi * int_{-1}^{eps} T_m(x) / sqrt(1 - x^2) dx, by
    quadrature in theta = arccos x (closed form pinned independently)."""
    from scipy.integrate import quad
    for eps in (-0.7, 0.0, 0.35, 0.9):
        mu = po.projector_moments(12, eps)
        for m in range(13):
            val = quad(lambda th: np.cos(m * th), np.arccos(eps), np.pi)[0]
            assert abs(mu[m] - (2.0 - (m == 0)) / np.pi * val) < 1e-12


def test_kpm_projector_known_answers():
    """Ports of test_kpm.cpp:74-146 and :188-204 (with the dense projector's exact
    bounds passed as E_min, E_max): order 200 reproduces the exact projector
    (Frobenius < 1e-2), is nearly idempotent (< 0.1), counts 3 states below 0."""
    H = np.diag([-2.0, -0.5, 0.5, 2.0])
    H[0, 1] = H[1, 0] = H[2, 3] = H[3, 2] = 0.1
    w, V = np.linalg.eigh(H)
    P_exact = (V[:, w < 0] @ V[:, w < 0].T)
    P = po.kpm_projector_matrix(H, 200, w.min(), w.max())
    assert np.linalg.norm(P - P_exact) < 1e-2
    H = np.diag([-1.5, -0.5, 0.5, 1.5])
    P = po.kpm_projector_matrix(H, 200, -1.5, 1.5)
    assert np.linalg.norm(P @ P - P) < 0.1
    H = np.diag([-2.0, -1.0, -0.5, 0.5, 1.0, 2.0])
    P = po.kpm_projector_matrix(H, 200, -2.0, 2.0)
    assert abs(np.trace(P).real - 3.0) < 0.1


def test_fixed_n_basis_matches_reference_golden():
    """with_fixed_particle_number (basis.cpp:70-89, 240-265): the oracle's DFS
    against the compiled reference for every L <= 6, N <= min(2L, 7); the size
    is C(2L, N)."""
    import math
    gold = np.load(os.path.join(GOLD, "ref_expression.npz"))
    n = 0
    for c in _manifest()["expression"]:
        if c["kind"] != "nbasis":
            continue
        up, dn = po.fixed_n_states(c["L"], c["N"])
        st = np.stack([up, dn], 1).astype(np.uint64).reshape(-1, 2)
        assert np.array_equal(st, gold[c["name"] + "__states"]), c["name"]
        assert len(up) == math.comb(2 * c["L"], c["N"])
        n += 1
    assert n == 39


def _rect_case(gold, c):
    coef, nops, ops = gold[c["name"] + "__coef"], gold[c["name"] + "__nops"], gold[c["name"] + "__ops"]
    terms, k = [], 0
    for cc, m in zip(coef, nops):
        terms.append((complex(cc), [tuple(int(v) for v in ops[k + i]) for i in range(int(m))]))
        k += int(m)
    spec = c["row"]
    L = spec[1] if spec[0] in "sn" else int(np.prod(spec[2:2 + spec[1]]))
    return L, terms


def test_rectangular_and_vector_elements_match_reference_golden():
    """rectangular_matrix / vector_elements (matrix_elements.h:31-51, 110-167)
    against the compiled reference on eleven cases: creators, annihilators and
    pair creation between sectors of different particle numbers, the square
    case, a J(Q)-shaped operator between momentum sectors, bases of different
    kinds, and vector elements on real-space, fixed-N and momentum bases."""
    gold = np.load(os.path.join(GOLD, "ref_expression.npz"))
    n = 0
    for c in _manifest()["expression"]:
        if c["kind"] not in ("rexpr", "vexpr"):
            continue
        L, terms = _rect_case(gold, c)
        rows = [tuple(int(v) for v in r) for r in gold[c["name"] + "__rows"]]
        if c["kind"] == "rexpr":
            cols = [tuple(int(v) for v in r) for r in gold[c["name"] + "__cols"]]
            M = po.rectangular_matrix(L, rows, cols, terms)
        else:
            M = po.vector_elements(L, rows, terms)[:, None]
        G = np.zeros_like(M)
        ij = gold[c["name"] + "__ij"]
        G[ij[:, 0], ij[:, 1]] = gold[c["name"] + "__val"]
        assert np.abs(M - G).max(initial=0) <= 1e-14 * max(1.0, np.abs(G).max(initial=0)), c["name"]
        assert np.count_nonzero(M) == c["nnz"] > 0, c["name"]
        n += 1
    assert n == 11


def test_rectangular_adjoint_and_pair_creation_identities():
    """<i|A|j> between bases B1 <- B2 equals conj(<j|A^dagger|i>) between B2 <- B1
    (random creators, annihilators and spin flips, L = 5, (2,1) <- (1,1) and
    fixed N = 3 <- 2); c+_{a s} c+_{b s'}|0> is one basis state of norm 1 for
    every pair of distinct modes, with the sign of the reordering."""
    rng = np.random.default_rng(5)
    L = 5
    terms = []
    for _ in range(12):
        nc = int(rng.integers(1, 3))
        ops = [(0, int(rng.integers(2)), int(rng.integers(L))) for _ in range(nc)]
        ops += [(1, int(rng.integers(2)), int(rng.integers(L))) for _ in range(nc - 1)]
        terms.append((complex(rng.standard_normal(), rng.standard_normal()), ops))
    adj = [(c.conjugate(), [(1 - t, s, o) for t, s, o in reversed(ops)]) for c, ops in terms]
    for b1, b2 in (((2, 1), (1, 1)), ((1, 2), (1, 1))):
        s1 = list(zip(*po.basis_states(L, *b1)))
        s2 = list(zip(*po.basis_states(L, *b2)))
        A = po.rectangular_matrix(L, s1, s2, terms)
        B = po.rectangular_matrix(L, s2, s1, adj)
        assert np.count_nonzero(A) > 10 and np.abs(A - B.conj().T).max() < 1e-14
    f3 = list(zip(*po.fixed_n_states(L, 3)))
    f2 = list(zip(*po.fixed_n_states(L, 2)))
    A = po.rectangular_matrix(L, f3, f2, terms)
    B = po.rectangular_matrix(L, f2, f3, adj)
    assert np.count_nonzero(A) > 10 and np.abs(A - B.conj().T).max() < 1e-14
    idx = {st: i for i, st in enumerate(f2)}
    for m1 in range(2 * L):
        for m2 in range(2 * L):
            if m1 == m2:
                continue
            (s1, o1), (s2, o2) = divmod(m1, L), divmod(m2, L)
            v = po.vector_elements(L, f2, [(1.0, [(0, s1, o1), (0, s2, o2)])])
            up = (1 << o1 if s1 == 0 else 0) | (1 << o2 if s2 == 0 else 0)
            dn = (1 << o1 if s1 == 1 else 0) | (1 << o2 if s2 == 1 else 0)
            want = np.zeros(len(f2))
            want[idx[(up, dn)]] = 1.0 if m1 < m2 else -1.0  # canonical order: ascending modes
            assert np.array_equal(v, want), (m1, m2)
