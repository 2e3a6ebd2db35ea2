"""GPU parity of the Krylov layer (Lanczos, lowest eigenpair, Krylov exp(-iHt),
KPM moments) against the oracle, plus closed forms at full 4 x 4 size.

Tolerances (FP64): Lanczos alpha/beta 1e-12 relative over 40 steps (rounding
grows along the recurrence), energies 1e-10 relative, evolved states 1e-10
relative, FP64 KPM moments 1e-11 absolute.  FP32 KPM moments against the FP64
run: 5e-4 absolute (stated, not hidden).
"""
import numpy as np
import pytest
import scipy.sparse.linalg as spla

import manybody_amd as mb
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu


def make_op(sizes, nu, nd, t, U):
    L = int(np.prod(sizes))
    basis = mb.FermionBasis.sector(L, nu, nd)
    if len(sizes) == 1:
        model = mb.HubbardModel(t, U, sizes[0])
    else:
        model = mb.HubbardModel2D(t, U, *sizes)
    return model.operator(basis)


def test_lanczos_alpha_beta_match_oracle_c1(cuda):
    sizes, nu, nd, t, U = [8], 4, 4, 1.0, 4.0
    op = make_op(sizes, nu, nd, t, U)
    P = po.ProductOperator(sizes, nu, nd, t, U)
    b = op.gaussian(seed=2024)
    dec = mb.lanczos_pass_one(op, b, 40)
    al, be, bn = P.lanczos(mb.from_planar(b), 40)
    assert dec.steps_taken == len(al) == 40 and len(dec.betas) == 39
    assert abs(dec.b_norm - bn) / bn < 1e-14
    assert np.abs(np.array(dec.alphas) - al).max() / np.abs(al).max() < 1e-12
    assert np.abs(np.array(dec.betas) - be).max() / np.abs(be).max() < 1e-12


def test_ground_state_c1_matches_oracle_and_eigsh(cuda):
    """1D L = 8, U = 4, 4 + 4 (D = 4,900), 200-step FP64 Lanczos."""
    sizes, nu, nd, t, U = [8], 4, 4, 1.0, 4.0
    op = make_op(sizes, nu, nd, t, U)
    b = op.gaussian(seed=1)
    ep = mb.find_min_eigenpair(op, 200, b)
    A = po.build_csc(sizes, nu, nd, t, U)
    e_ref = spla.eigsh(A.to_scipy().real, k=1, which="SA", tol=1e-14)[0][0]
    assert abs(ep.value - e_ref) / abs(e_ref) < 1e-10
    # oracle's own 200-step Lanczos on the same start vector (reference order)
    b_ref = op.to_reference_order(b).cpu().numpy()
    e_orc, x_orc = po.find_min_eigenpair_csc(A, b_ref, 200)
    assert abs(ep.value - e_orc) / abs(e_orc) < 1e-10
    # eigenvector from pass two: residual and overlap with the oracle's
    x = op.to_reference_order(ep.vector).cpu().numpy()
    r = np.linalg.norm(A.matvec(x) - ep.value * x) / np.linalg.norm(x)
    assert r < 1e-6
    ov = abs(np.vdot(x, x_orc)) / (np.linalg.norm(x) * np.linalg.norm(x_orc))
    assert abs(ov - 1.0) < 1e-8


def test_ground_state_u0_closed_form_c1(cuda):
    op = make_op([8], 4, 4, 1.0, 0.0)
    ep = mb.find_min_eigenpair(op, 200, op.gaussian(seed=3), want_vector=False)
    assert abs(ep.value - (-9.656854249492380)) < 1e-10 * 9.66


def test_ground_state_2d_4x4_u0_closed_form_full_size(cuda):
    """4 x 4 periodic, U = 0, 7 + 7 (D = 130,873,600): E0 = -24 exactly."""
    op = make_op([4, 4], 7, 7, 1.0, 0.0)
    ep = mb.find_min_eigenpair(op, 150, op.gaussian(seed=4), want_vector=False)
    assert abs(ep.value + 24.0) / 24.0 < 1e-10


def test_lanczos_first_steps_c2_match_oracle(cuda):
    """4 x 4, U = 8, 7 + 7: the first 4 alpha/beta against the CPU oracle."""
    sizes, nu, nd, t, U = [4, 4], 7, 7, 1.0, 8.0
    op = make_op(sizes, nu, nd, t, U)
    b = op.gaussian(seed=5)
    dec = mb.lanczos_pass_one(op, b, 4)
    P = po.ProductOperator(sizes, nu, nd, t, U, threads=16)
    al, be, bn = P.lanczos(mb.from_planar(b), 4)
    assert np.abs(np.array(dec.alphas) - al).max() / np.abs(al).max() < 1e-12
    assert np.abs(np.array(dec.betas) - be).max() / np.abs(be).max() < 1e-12


def test_lanczos_breakdown_and_zero_norm(cuda):
    import torch
    op = make_op([4], 0, 4, 1.0, 2.0)  # D = 1: breakdown after the first step
    b = op.gaussian(seed=6)
    dec = mb.lanczos_pass_one(op, b, 5)
    assert dec.steps_taken == 1 and dec.betas == []
    op2 = make_op([6], 3, 3, 1.0, 2.0)
    with pytest.raises(RuntimeError, match="zero norm"):
        mb.lanczos_pass_one(op2, torch.zeros((2, op2.dimension()), dtype=torch.float64,
                                             device=cuda), 5)
    dec0 = mb.lanczos_pass_one(op2, op2.gaussian(1), 0)
    assert dec0.steps_taken == 0 and dec0.b_norm > 0


def test_time_evolve_matches_oracle(cuda):
    sizes, nu, nd, t, U = [6], 3, 2, 1.0, 3.0
    op = make_op(sizes, nu, nd, t, U)
    A = po.build_csc(sizes, nu, nd, t, U)
    psi = op.gaussian(seed=7)
    out = op.to_reference_order(mb.time_evolve_state(op, psi, 0.45, mb.EvolutionOptions(30)))
    ref = po.time_evolve_csc(A, op.to_reference_order(psi).cpu().numpy(), 0.45, 30)
    out = out.cpu().numpy()
    assert np.linalg.norm(out - ref) / np.linalg.norm(ref) < 1e-10


def test_time_evolve_steps_conserves_norm_and_matches_single_steps(cuda):
    import torch
    op = make_op([10], 5, 4, 1.0, 4.0)
    psi = op.gaussian(seed=8)
    n0 = torch.linalg.vector_norm(psi).item()
    stepped = mb.time_evolve_state_steps(op, psi, 0.0, 0.3, 0.1)
    manual = psi
    for _ in range(3):
        manual = mb.time_evolve_state(op, manual, 0.1)
    assert abs(torch.linalg.vector_norm(stepped).item() - n0) / n0 < 1e-10
    assert torch.linalg.vector_norm(stepped - manual).item() / n0 < 1e-12


def test_kpm_fp64_moments_match_oracle(cuda):
    sizes, nu, nd, t, U = [8], 3, 3, 1.0, 4.0
    op = make_op(sizes, nu, nd, t, U)
    A = po.build_csc(sizes, nu, nd, t, U)
    ev = np.linalg.eigvalsh(A.to_scipy().toarray().real)
    a, b = (ev[-1] - ev[0]) / 2 * 1.02, (ev[-1] + ev[0]) / 2  # spectrum inside [-1, 1]
    M, nvec, seed = 64, 4, 99
    mu = np.array(mb.kpm_dos_moments(op, a, b, M, nvec, seed, fp32=False))
    vecs = op.random_phase(seed, nvec, fp32=False)  # (2 nvec, D)
    vecs_ref = np.stack([op.to_reference_order(vecs[2 * r:2 * r + 2].contiguous()).cpu().numpy()
                         for r in range(nvec)])
    ref = po.kpm_moments_csc(A, a, b, M, vecs_ref)
    assert np.abs(mu - ref).max() < 1e-11
    assert abs(mu[0] - 1.0) < 1e-14


def test_kpm_fp32_error_vs_fp64_is_small(cuda):
    op = make_op([12], 6, 6, 1.0, 4.0)
    # the spectrum of the L = 12 ring, 6 + 6, U = 4 is [-6.92, 30.92] (scipy eigsh)
    a, b = 19.5, 12.0
    mu64 = np.array(mb.kpm_dos_moments(op, a, b, 256, 4, 5, fp32=False))
    mu32 = np.array(mb.kpm_dos_moments(op, a, b, 256, 4, 5, fp32=True))
    err = np.abs(mu64 - mu32).max()
    print(f"KPM FP32 vs FP64 max |d mu| (L=12, 6+6, M=256, R=4): {err:.3e}")
    assert err < 5e-4


# ------------------------------------------------------------------ reorthogonalised Lanczos

def test_reorthogonalized_lanczos_ritz_values_are_eigenvalues(cuda):
    """Full reorthogonalisation run to the dimension (D = 90, strongly degenerate
    spectrum): it stops by breakdown once the reachable space is exhausted,
    every Ritz value is an eigenvalue of the dense H (no spurious values) and
    every distinct eigenvalue is found.  (In floating point the Krylov space of
    an exactly degenerate H does not close after #distinct steps: rounding seeds
    the other directions of each degenerate eigenspace, so degenerate
    eigenvalues may appear more than once, each time correctly.)"""
    sizes, nu, nd, t, U = [6], 2, 1, 1.0, 3.0  # D = 15 * 6 = 90
    op = make_op(sizes, nu, nd, t, U)
    A = po.build_csc(sizes, nu, nd, t, U)
    ev = np.linalg.eigvalsh(A.to_scipy().toarray().real)
    dec = mb.lanczos_pass_one(op, op.gaussian(5), 90, reorthogonalize=True)
    assert 40 <= dec.steps_taken <= 90
    w, _ = mb.tridiag_eig(dec.alphas, dec.betas, vectors=False)
    assert max(np.abs(ev - x).min() for x in w) < 1e-9
    assert max(np.abs(w - e).min() for e in ev) < 1e-9


def test_reorthogonalized_ground_state_c1_no_ghosts(cuda):
    """C1 (D = 4,900), 200 steps: same E0 as eigsh; with reorthogonalisation no
    more Ritz values sit at E0 than its multiplicity (plain Lanczos repeats
    converged values as ghosts); the eigenvector built from the stored vectors
    has a small residual and is the pass-two eigenvector up to a phase."""
    sizes, nu, nd, t, U = [8], 4, 4, 1.0, 4.0
    op = make_op(sizes, nu, nd, t, U)
    b = op.gaussian(seed=1)
    A = po.build_csc(sizes, nu, nd, t, U)
    low = np.sort(spla.eigsh(A.to_scipy().real, k=6, which="SA", tol=1e-14)[0])
    e_ref = low[0]
    mult = int(np.sum(np.abs(low - e_ref) < 1e-8))
    dec = mb.lanczos_pass_one(op, b, 200, reorthogonalize=True)
    w, _ = mb.tridiag_eig(dec.alphas, dec.betas, vectors=False)
    assert abs(w[0] - e_ref) / abs(e_ref) < 1e-10
    assert 1 <= int(np.sum(np.abs(w - e_ref) < 1e-6)) <= mult
    dec_plain = mb.lanczos_pass_one(op, b, 200)
    w_plain, _ = mb.tridiag_eig(dec_plain.alphas, dec_plain.betas, vectors=False)
    print("E0 multiplicity", mult, "; Ritz values within 1e-6 of E0: reorthogonalised",
          int(np.sum(np.abs(w - e_ref) < 1e-6)), "plain", int(np.sum(np.abs(w_plain - e_ref) < 1e-6)))
    ep = mb.find_min_eigenpair(op, 200, b, reorthogonalize=True)
    assert abs(ep.value - e_ref) / abs(e_ref) < 1e-10
    x = op.to_reference_order(ep.vector).cpu().numpy()
    r = np.linalg.norm(A.matvec(x) - ep.value * x) / np.linalg.norm(x)
    assert r < 1e-8
    x2 = op.to_reference_order(mb.find_min_eigenpair(op, 200, b).vector).cpu().numpy()
    ov = abs(np.vdot(x, x2)) / (np.linalg.norm(x) * np.linalg.norm(x2))
    assert abs(ov - 1.0) < 1e-8
    assert abs(np.linalg.norm(x) - dec.b_norm) / dec.b_norm < 1e-10  # |x| = |b| (lanczos.h:283-290)


# ------------------------------------------------------------------ FP32 Krylov time evolution

def test_time_evolve_fp32_error_vs_fp64_is_stated(cuda):
    """exp(-iHt) with FP32 Krylov vectors and H.v (state, alpha/beta reductions and
    the exponential of T_k in FP64): 1D ring L = 12, 6 + 6, U = 4 (D = 853,776),
    t = 0 -> 1 in 20 steps of 0.05, Krylov dimension 30.  Stated bar: relative
    state error 1e-4 against the FP64 run, norm kept to 1e-5."""
    import torch
    op = make_op([12], 6, 6, 1.0, 4.0)
    psi = op.gaussian(9)
    psi = psi / torch.linalg.vector_norm(psi)
    opts = mb.EvolutionOptions(30)
    out64 = mb.time_evolve_state_steps(op, psi, 0.0, 1.0, 0.05, opts)
    out32 = mb.time_evolve_state_steps(op, psi, 0.0, 1.0, 0.05, opts, fp32=True)
    err = (torch.linalg.vector_norm(out32 - out64) / torch.linalg.vector_norm(out64)).item()
    drift = abs(torch.linalg.vector_norm(out32).item() - 1.0)
    print(f"FP32 Krylov evolution: relative state error vs FP64 {err:.3e}, norm drift {drift:.3e}")
    assert err < 1e-4
    assert drift < 1e-5
    # the FP64 path itself against the oracle on a small sector
    sizes, nu, nd, t, U = [8], 3, 2, 1.0, 4.0
    op8 = make_op(sizes, nu, nd, t, U)
    A = po.build_csc(sizes, nu, nd, t, U)
    b = op8.gaussian(4)
    b_ref = op8.to_reference_order(b).cpu().numpy()
    ref = po.time_evolve_csc(A, b_ref, 0.7, 30)
    got32 = op8.to_reference_order(mb.time_evolve_state(op8, b, 0.7, opts, fp32=True)).cpu().numpy()
    assert np.linalg.norm(got32 - ref) / np.linalg.norm(ref) < 1e-5


# ------------------------------------------------------------ spectral bounds

@pytest.mark.parametrize("kind", ["real", "momentum", "expression"])
def test_power_method_and_bounds_match_oracle(cuda, kind):
    """power_method / estimate_bounds (linear_operator_utils.h:18-78) on the
    device against the oracle's restatement with the same seed and iteration
    counts (1e-9 relative: the same iterates up to rounding)."""
    import torch
    if kind == "real":
        basis = mb.FermionBasis.sector(8, 4, 3)
        op = mb.HubbardModel(1.0, 4.0, 8).operator(basis)
        A = po.build_csc([8], 4, 3, 1.0, 4.0)
        mat = A.to_scipy()
    elif kind == "momentum":
        basis = mb.FermionBasis.with_fixed_particle_number_spin_momentum(9, 5, 1, [3, 3], [1, 0])
        op = mb.HubbardMomentumFactorizedFixedK(basis, [3, 3], 1.0, 6.0, [1, 0])
        mat = po.momentum_hamiltonian([3, 3], 3, 2, 1, 1.0, 6.0)
    else:
        basis = mb.FermionBasis.sector(7, 3, 3)
        e = mb.HubbardModel(1.0, 3.0, 7).hamiltonian() + mb.hopping(0.4j, 0, 3, mb.Spin.Down)
        op = mb.FermionExpressionOperator(basis, e)
        up, dn = po.basis_states(7, 3, 3)
        terms = [(c, [(o >> 15, (o >> 14) & 1, o & 0x3FFF) for o in k]) for k, c in e.terms.items()]
        mat = po.expression_matrix(7, list(zip(up, dn)), terms)
    seed = op.gaussian(seed=12)
    seed_ref = op.to_reference_order(seed).cpu().numpy()
    mv = (lambda v: mat @ v)
    got = mb.power_method(op, seed, 30)
    want = po.power_method(mv, seed_ref, 30)
    assert abs(got - want) <= 1e-9 * abs(want)
    b = mb.estimate_bounds(op, seed, 30, 0.1)
    wa, wb = po.estimate_bounds(mv, seed_ref, 30, 0.1)
    assert abs(b.alpha - wa) <= 1e-9 * abs(wa) and abs(b.beta - wb) <= 1e-9 * abs(wb)
    # the seed is not modified
    assert torch.equal(seed, op.gaussian(seed=12))


def diag_operator(d):
    """diag(d) on the device: one up electron on len(d) orbitals (basis order =
    orbital order) with sum_i d_i n_i, through the expression compiler."""
    e = mb.FermionExpression()
    for i, v in enumerate(d):
        e += float(v) * mb.density(mb.Spin.Up, i)
    return mb.FermionExpressionOperator(mb.FermionBasis.sector(len(d), 1, 0), e)


def test_reference_linear_operator_known_answers_on_device(cuda):
    """Ports of test_linear_operator.cpp:128-160 run through the device path:
    power_method on diag(3, -1, 0.5) from (1, 1, 1) -> 3; estimate_bounds on
    diag(-2, 0.5, 4) -> (-2, 4) (1e-3); Exp with k = D on diag(-2, 0.5, 1.5)
    applied to (1, -2, 0.5) -> exp(d) v (1e-6)."""
    import torch
    op = diag_operator([3.0, -1.0, 0.5])
    ones = mb.to_planar(np.ones(3, np.complex128))
    assert abs(mb.power_method(op, ones, 40) - 3.0) < 1e-3
    op = diag_operator([-2.0, 0.5, 4.0])
    b = mb.estimate_bounds(op, op.gaussian(seed=4), 40, 0.2)
    assert abs(b.alpha + 2.0) < 1e-3 and abs(b.beta - 4.0) < 1e-3
    d = np.array([-2.0, 0.5, 1.5])
    op = diag_operator(d)
    v = np.array([1.0, -2.0, 0.5], np.complex128)
    out = mb.from_planar(mb.Exp(op, krylov_steps=3).apply(mb.to_planar(v)))
    assert np.abs(out - np.exp(d) * v).max() < 1e-6


def test_exp_matches_expm_on_hubbard(cuda):
    import scipy.linalg as sla
    op = mb.HubbardModel(0.3, 0.5, 6).operator(mb.FermionBasis.sector(6, 3, 2))
    A = po.build_csc([6], 3, 2, 0.3, 0.5).to_scipy().toarray()
    v = op.gaussian(seed=8)
    got = op.to_reference_order(mb.Exp(op, 30).apply(v)).cpu().numpy()
    want = sla.expm(A) @ op.to_reference_order(v).cpu().numpy()
    assert np.linalg.norm(got - want) <= 1e-10 * np.linalg.norm(want)
    z = op.empty().zero_()
    assert np.array_equal(mb.Exp(op).apply(z).cpu().numpy(), z.cpu().numpy())  # |v| = 0


def matrix_operator(A):
    """A real symmetric n x n matrix (n <= 32) as a device operator: one up
    electron on n orbitals with sum_ij A_ij c+_i c_j (one particle: no fermionic
    signs, basis order = orbital order), through the expression compiler."""
    A = np.asarray(A, float)
    e = mb.FermionExpression()
    for i in range(A.shape[0]):
        for j in range(A.shape[1]):
            if A[i, j] != 0.0:
                e += A[i, j] * mb.one_body(mb.Spin.Up, i, mb.Spin.Up, j)
    return mb.FermionExpressionOperator(mb.FermionBasis.sector(A.shape[0], 1, 0), e)


def tridiag(al, be):
    m = len(al)
    T = np.diag(al)
    for i in range(m - 1):
        T[i, i + 1] = T[i + 1, i] = be[i]
    return T


def e1_solve(al, be):
    e1 = np.zeros(len(al))
    e1[0] = 1.0
    return np.linalg.solve(tridiag(al, be), e1)


def dev(op, v):
    import torch
    return op.to_device_order(torch.from_numpy(np.asarray(v, np.complex128)).to("cuda"))


def host(op, x):
    return op.to_reference_order(x).cpu().numpy()


def chain(n, d, h):
    return np.diag(np.full(n, d)) + np.diag(np.full(n - 1, h), 1) + np.diag(np.full(n - 1, h), -1)


def test_reference_lanczos_known_answers_on_device(cuda):
    """Ports of test_lanczos.cpp:30-236 through the device path."""
    import scipy.linalg as sla
    # lanczos_solve_matches_direct_solution (:30-52)
    A = chain(5, 4.0, -1.0)
    op = matrix_operator(A)
    b = np.linspace(1.0, 5.0, 5)
    x = host(op, mb.solve(op, dev(op, b), 5, e1_solve))
    assert np.linalg.norm(x - np.linalg.solve(A, b)) / np.linalg.norm(np.linalg.solve(A, b)) < 1e-8
    # lanczos_{max,min}_eigenpair_matches_diagonal (:54-70)
    op = matrix_operator(np.diag([1.0, 3.0, -2.0]))
    seed = op.gaussian(seed=1)
    mx = mb.find_max_eigenpair(op, 3, seed)
    assert abs(mx.value - 3.0) < 1e-6
    v = host(op, mx.vector)
    w = np.diag([1.0, 3.0, -2.0]) @ v
    assert abs(np.sum(v * w).real / np.sum(v * v).real - mx.value) < 1e-6  # arma::dot
    assert abs(mb.find_min_eigenpair(op, 3, seed).value + 2.0) < 1e-6
    # lanczos_exp_solver_matches_direct_expm (:104-134)
    A = chain(4, 2.0, -1.0)
    op = matrix_operator(A)
    b = np.array([1.0, 0.0, 0.0, 0.0])
    x = host(op, mb.solve(op, dev(op, b), 4, lambda al, be: sla.expm(tridiag(al, be))[:, 0]))
    assert np.linalg.norm(x - sla.expm(A) @ b) < 1e-8
    # lanczos_inverse_approximation_improves_with_k (:136-164)
    A = np.diag([10.0, 8.0, 6.0, 5.0]) + np.diag([1.0] * 3, 1) + np.diag([1.0] * 3, -1)
    op = matrix_operator(A)
    b = np.ones(4)
    exact = np.linalg.solve(A, b)
    err = [np.linalg.norm(host(op, mb.solve(op, dev(op, b), k, e1_solve)) - exact)
           / np.linalg.norm(exact) for k in (2, 3, 4)]
    assert err[1] <= err[0] and err[2] <= err[1] and err[2] < 1e-10
    # lanczos_decomposition_tracks_steps_and_norm (:166-186)
    op = matrix_operator(chain(6, 4.0, 1.0))
    d = mb.lanczos_pass_one(op, dev(op, np.ones(6)), 5)
    assert 0 < d.steps_taken <= 5
    assert len(d.alphas) == d.steps_taken and len(d.betas) + 1 == d.steps_taken
    assert abs(d.b_norm - np.sqrt(6.0)) < 1e-12 and all(v > 0 for v in d.betas)
    # lanczos_max_eigenpair_matches_nearest_neighbor_chain (:221-236) at n = 32
    # (the basis holds at most 32 orbitals; the reference uses n = 80, k = 50)
    n, k = 32, 20
    op = matrix_operator(chain(n, 2.0, -1.0))
    ep = mb.find_max_eigenpair(op, k, op.gaussian(seed=2), want_vector=False)
    assert abs(ep.value - (2.0 - 2.0 * np.cos(n * np.pi / (n + 1)))) < 2e-2


def test_solve_generic_callback_and_errors(cuda):
    """solve with a complex-valued solver (exp(-i t T) e1) equals time evolution;
    a solver returning the wrong length is an error."""
    op = mb.HubbardModel(1.0, 4.0, 6).operator(mb.FermionBasis.sector(6, 3, 2))
    psi = op.gaussian(seed=5)
    import scipy.linalg as sla
    x = mb.solve(op, psi, 30, lambda al, be: sla.expm(-0.3j * tridiag(al, be))[:, 0])
    y = mb.time_evolve_state(op, psi, 0.3, mb.EvolutionOptions(30))
    assert (x - y).abs().max().item() <= 1e-12 * y.abs().max().item()
    with pytest.raises(RuntimeError):
        mb.solve(op, psi, 10, lambda al, be: np.zeros(3))


def test_reference_evolve_state_known_answers_on_device(cuda):
    """Ports of test_evolve_state.cpp:43-121 (2 x 2 diagonal Hamiltonians as
    one-particle expressions, k = 2): stepped exp(-iHt) and its callback times,
    imaginary time, against the exact exponentials (1e-8)."""
    import torch
    H = np.diag([1.0, 2.0])
    op = matrix_operator(H)
    psi0 = np.array([1.0, 1j])
    out = mb.time_evolve_state_steps(op, dev(op, psi0), 0.0, 0.3, 0.1, mb.EvolutionOptions(2))
    want = np.exp(-1j * 0.3 * np.diag(H)) * psi0
    assert np.linalg.norm(host(op, out) - want) / np.linalg.norm(want) < 1e-8
    H = np.diag([0.5, 1.5])
    op = matrix_operator(H)
    psi0 = np.array([0.3 - 0.4j, 0.2 + 0.1j])
    times, norms = [], []

    def cb(t, state):
        times.append(t)
        norms.append(torch.linalg.vector_norm(state).item())

    out = mb.time_evolve_state_steps(op, dev(op, psi0), 0.0, 0.3, 0.1, mb.EvolutionOptions(2),
                                     callback=cb)
    assert len(times) == 4 and abs(times[0]) < 1e-12 and abs(times[-1] - 0.3) < 1e-12
    assert max(abs(n - np.linalg.norm(psi0)) for n in norms) < 1e-12
    want = np.exp(-1j * 0.3 * np.diag(H)) * psi0
    assert np.linalg.norm(host(op, out) - want) / np.linalg.norm(want) < 1e-8
    with pytest.raises(RuntimeError):  # a failing callback aborts
        mb.time_evolve_state_steps(op, dev(op, psi0), 0.0, 0.3, 0.1, mb.EvolutionOptions(2),
                                   callback=lambda t, s: 1 / 0)
    H = np.diag([0.2, 1.1])
    op = matrix_operator(H)
    psi0 = np.array([0.4 - 0.1j, -0.2 + 0.3j])
    out = mb.imaginary_time_evolve_state(op, dev(op, psi0), 0.7, mb.EvolutionOptions(2))
    want = np.exp(-0.7 * np.diag(H)) * psi0
    assert np.linalg.norm(host(op, out) - want) / np.linalg.norm(want) < 1e-8


def test_kpm_projector_on_device_matches_oracle(cuda):
    """SparseKPMProjector (kpm.h:184-241) applied on the device: ports of
    test_kpm.cpp:74-121 through one-particle matrix operators, and a Hubbard
    sector against the oracle's vector recurrence (1e-11 relative)."""
    H = np.diag([-2.0, -0.5, 0.5, 2.0])
    H[0, 1] = H[1, 0] = H[2, 3] = H[3, 2] = 0.1
    w, V = np.linalg.eigh(H)
    P_exact = V[:, w < 0] @ V[:, w < 0].T
    op = matrix_operator(H)
    proj = mb.SparseKPMProjector(op, 200, w.min(), w.max(), 0.0)
    cols = [host(op, proj.apply(dev(op, np.eye(4)[:, j]))) for j in range(4)]
    assert np.linalg.norm(np.stack(cols, 1) - P_exact) < 1e-2
    A = chain(6, 0.0, -1.0)
    op = matrix_operator(A)
    v = np.random.default_rng(1).standard_normal(6)
    got = host(op, mb.SparseKPMProjector(op, 100, -2.0, 2.0, 0.0).apply(dev(op, v)))
    want = po.kpm_projector_matrix(A, 100, -2.0, 2.0, 0.0) @ v
    assert np.linalg.norm(got - want) < 1e-10
    # a Hubbard sector: E_F in the middle of the spectrum
    hop = mb.HubbardModel(1.0, 4.0, 8).operator(mb.FermionBasis.sector(8, 4, 3))
    Ac = po.build_csc([8], 4, 3, 1.0, 4.0).to_scipy()
    x = hop.gaussian(seed=3)
    got = host(hop, mb.SparseKPMProjector(hop, 64, -9.0, 22.0, 3.0, 0.05).apply(x))
    want = po.kpm_projector_apply(lambda u: Ac @ u, host(hop, x), 64, -9.0, 22.0, 3.0, 0.05)
    assert np.linalg.norm(got - want) <= 1e-11 * np.linalg.norm(want)


def test_kpm_projector_from_operator_matches_oracle(cuda):
    """KPMProjector(H, M, E_fermi, padding) (kpm.h:67-180) on a device operator:
    the bounds from Lanczos equal the exact extreme eigenvalues (k >= D), the
    built matrix equals the oracle's projector matrix with those bounds, and at
    order 200 with E_fermi between the 3rd and 4th levels it is close to the
    projector on the three lowest states (trace ~ 3, idempotent), as
    test_kpm.cpp:100-146 checks for its dense matrices."""
    sizes, nu, nd = [4], 2, 1
    A = po.build_csc(sizes, nu, nd, 1.0, 4.0).to_scipy().toarray()
    w = np.linalg.eigvalsh(A)
    op = mb.HubbardModel(1.0, 4.0, 4).operator(mb.FermionBasis.sector(4, nu, nd))
    ef = 0.5 * (w[2] + w[3])
    P = mb.KPMProjector(op, 200, ef, 0.01)
    r = P.rescaling()
    assert abs(r.a - (w[-1] - w[0]) / 2 * 1.01) < 1e-10 and abs(r.b - (w[-1] + w[0]) / 2) < 1e-10
    Pm = P.build_matrix()
    want = po.kpm_projector_matrix(A, 200, P.e_min, P.e_max, ef, 0.01)
    assert np.abs(Pm - want).max() < 1e-12
    assert abs(np.trace(Pm).real - 3.0) < 0.1
    assert np.abs(Pm @ Pm - Pm).max() < 0.05
    with pytest.raises(RuntimeError):
        mb.KPMProjector(op, 50, w[-1] + 100.0)


def test_kpm_moments_on_any_operator(cuda):
    """KPM DOS moments through the generic path (any operator, FP64): the
    compiled Hubbard expression gives the native real-space moments (same
    phases, 1e-12); a fixed-momentum sector gives the oracle's moments for the
    same vectors (1e-11); Shifted(H, 1) with b + 1 gives H's moments.  FP32
    runs on the Hubbard kernels only: the expression (Hubbard form) gives the
    native FP32 moments to FP32 rounding, the composite is refused."""
    sizes, nu, nd, t, U = [8], 3, 3, 1.0, 4.0
    native = make_op(sizes, nu, nd, t, U)
    A = po.build_csc(sizes, nu, nd, t, U)
    ev = np.linalg.eigvalsh(A.to_scipy().toarray().real)
    a, b = (ev[-1] - ev[0]) / 2 * 1.02, (ev[-1] + ev[0]) / 2
    M, nvec, seed = 64, 3, 7
    mu_n = np.array(mb.kpm_dos_moments(native, a, b, M, nvec, seed, fp32=False))
    expr = mb.FermionExpressionOperator(native.basis, mb.HubbardModel(t, U, 8).hamiltonian())
    mu_e = np.array(mb.kpm_dos_moments(expr, a, b, M, nvec, seed, fp32=False))
    assert np.abs(mu_e - mu_n).max() < 1e-12
    mu_s = np.array(mb.kpm_dos_moments(mb.Shifted(native, 1.0), a, b + 1.0, M, nvec, seed,
                                       fp32=False))
    assert np.abs(mu_s - mu_n).max() < 1e-12
    mu32_e = np.array(mb.kpm_dos_moments(expr, a, b, M, nvec, seed, fp32=True))
    mu32_n = np.array(mb.kpm_dos_moments(native, a, b, M, nvec, seed, fp32=True))
    assert np.abs(mu32_e - mu32_n).max() < 1e-5
    assert np.abs(mu32_n - mu_n).max() < 1e-4
    with pytest.raises(ValueError):
        mb.kpm_dos_moments(mb.Shifted(native, 1.0), a, b + 1.0, M, nvec, seed, fp32=True)
    # a fixed-momentum sector against the oracle with the same phase vectors
    import scipy.sparse as sps
    msz, K = [6], 1
    basis = mb.FermionBasis.with_fixed_particle_number_spin_momentum(6, 5, 1, msz, [K])
    H = mb.HubbardMomentumFactorizedFixedK(basis, msz, 1.0, 4.0, [K])
    Hd = po.momentum_hamiltonian(msz, 3, 2, K, 1.0, 4.0)
    w = np.linalg.eigvalsh(Hd)
    a, b = (w[-1] - w[0]) / 2 * 1.05, (w[-1] + w[0]) / 2
    mu = np.array(mb.kpm_dos_moments(H, a, b, 40, 2, 3, fp32=False))
    vecs = H.random_phase(3, 2, fp32=False)
    vref = np.stack([H.to_reference_order(vecs[2 * r:2 * r + 2].contiguous()).cpu().numpy()
                     for r in range(2)])
    sp = sps.csc_matrix(Hd)
    csc = po.CSC(sp.shape[0], sp.indptr.astype(np.int64), sp.indices.astype(np.int64),
                 sp.data.astype(np.complex128))
    ref = po.kpm_moments_csc(csc, a, b, 40, vref)
    assert np.abs(mu - ref).max() < 1e-11 and abs(mu[0] - 1.0) < 1e-14


@pytest.mark.parametrize("fp32", [False, True])
def test_time_evolve_steps_graph_replays_equal_eager_steps(cuda, monkeypatch, fp32):
    """time_evolve_state_steps captures the device part of a Krylov step once
    (norm, normalisation, k Lanczos steps into the stored basis) and replays it
    at every step, with the host's exp(-i tau T) e1 and the combine in between;
    MANYBODY_AMD_KRYLOV_GRAPH=0 runs every step eagerly.  The same kernels run in
    the same order, so every intermediate state (seen through the callback) and
    the final one are bit-identical, over 12 steps (the last one shorter) on a
    real-space sector, a 2D sector, a momentum sector and a zero state."""
    import torch
    cases = [make_op([10], 5, 4, 1.0, 4.0), make_op([3, 2], 2, 2, 1.0, 6.0)]
    if not fp32:  # momentum sectors run in FP64 only
        kb = mb.FermionBasis.with_fixed_particle_number_spin_momentum(8, 6, 0, [8], [2])
        cases.append(mb.HubbardMomentumFactorizedFixedK(kb, [8], 1.0, 4.0, [2]))
    for op in cases:
        psi = op.gaussian(seed=41)
        for start in (psi, torch.zeros_like(psi)):
            runs = {}
            for mode in ("1", "0"):
                monkeypatch.setenv("MANYBODY_AMD_KRYLOV_GRAPH", mode)
                seen = []
                out = mb.time_evolve_state_steps(op, start, 0.0, 0.53, 0.05, mb.EvolutionOptions(20),
                                                 fp32=fp32, callback=lambda t, s: seen.append((t, s.clone())))
                runs[mode] = (out, seen)
            (o1, s1), (o0, s0) = runs["1"], runs["0"]
            assert len(s1) == len(s0) == 12
            for (ta, a), (tb, b) in zip(s1, s0):
                assert ta == tb and torch.equal(a, b)
            assert torch.equal(o1, o0)
            if start is psi:
                n0 = torch.linalg.vector_norm(psi).item()
                assert abs(torch.linalg.vector_norm(o1).item() - n0) / n0 < (1e-5 if fp32 else 1e-10)
            else:
                assert torch.count_nonzero(o1).item() == 0


def test_time_evolve_steps_small_sector_kernel_matches_general_path_and_oracle(cuda, monkeypatch):
    """FP64 stepping on real-space sectors of at most 2048 states runs each
    step's Lanczos in one workgroup (MANYBODY_AMD_KRYLOV_SMALL=0: the general
    path).  Its sums run in another order, so: the states agree to 1e-12
    relative after 20 steps, and match the dense exp(-iHt) of the oracle's
    matrix to 1e-9, on 1D and 2D sectors, one where k exceeds the dimension
    (the run breaks down at the invariant subspace), and a zero state."""
    import scipy.linalg as sla
    import torch
    for sizes, nu, nd in [([8], 2, 3), ([3, 3], 1, 2), ([4], 1, 1)]:
        op = make_op(sizes, nu, nd, 1.0, 4.0)
        assert op.dimension() <= 2048
        psi = op.gaussian(seed=5)
        outs = {}
        for mode in ("1", "0"):
            monkeypatch.setenv("MANYBODY_AMD_KRYLOV_SMALL", mode)
            outs[mode] = mb.time_evolve_state_steps(op, psi, 0.0, 1.0, 0.05, mb.EvolutionOptions(30))
        o1, o0 = outs["1"], outs["0"]
        assert torch.linalg.vector_norm(o1 - o0).item() <= 1e-12 * torch.linalg.vector_norm(o0).item()
        A = po.build_csc(sizes, nu, nd, 1.0, 4.0).to_scipy().toarray()
        x = op.to_reference_order(psi).cpu().numpy()
        want = sla.expm(-1j * 1.0 * A) @ x
        got = op.to_reference_order(o1).cpu().numpy()
        assert np.linalg.norm(got - want) <= 1e-9 * np.linalg.norm(want), sizes
        monkeypatch.setenv("MANYBODY_AMD_KRYLOV_SMALL", "1")
        z = mb.time_evolve_state_steps(op, torch.zeros_like(psi), 0.0, 0.2, 0.05, mb.EvolutionOptions(30))
        assert torch.count_nonzero(z).item() == 0
