"""Operator algebra (linear_operator.h:93-281) on device operators: Negated,
Scaled, Shifted, NegShift, Sum, Difference, Composed and Identity.

The unit cases port test_linear_operator.cpp:9-127, whose DiagonalOperator is
here the expression sum_i d_i n_{up,i} on the 3-orbital sector with one up
particle (reference order: orbitals 0, 1, 2, so the operator is diag(d)).  The
Krylov cases check what the algebra fixes: spectra of c H, H + c, -H, H^2 and
a sum of model parts, and the phase that a shift adds to time evolution.
"""
import numpy as np
import pytest
import torch

import manybody_amd as mb
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu


def diag_op(d):
    basis = mb.FermionBasis.sector(3, 1, 0)
    e = mb.density(mb.Spin.Up, 0) * 0.0
    for i, v in enumerate(d):
        e += mb.density(mb.Spin.Up, i) * float(v)
    op = mb.FermionExpressionOperator(basis, e)
    op._keep = basis
    return op


def apply(op, v, cuda):
    x = torch.tensor(np.asarray(v, np.complex128), device=cuda)
    return op.apply(x).cpu().numpy()


def test_reference_algebra_unit_cases(cuda):
    """test_linear_operator.cpp:9-127, values as the reference checks them."""
    op = diag_op([1.0, 2.0, -1.0])
    v = [2.0, -1.0, 3.0]
    neg = -op
    assert np.array_equal(apply(neg, v, cuda), [-2.0, 2.0, 3.0]) and neg.dimension() == 3
    sh = mb.Shifted(op, 1.5)
    assert np.array_equal(apply(sh, v, cuda), [5.0, -3.5, 1.5]) and sh.dimension() == 3
    ns = mb.NegShift(op, 1.5)
    assert np.array_equal(apply(ns, v, cuda), [1.0, 0.5, 7.5]) and ns.dimension() == 3
    op = diag_op([1.0, -2.0, 0.5])
    v = [2.0, 3.0, -4.0]
    assert np.array_equal(apply(op * 2.0, v, cuda), [4.0, -12.0, -4.0])
    assert np.array_equal(apply(2.0 * op, v, cuda), [4.0, -12.0, -4.0])
    a, b = diag_op([1.0, 2.0, 3.0]), diag_op([-1.0, 4.0, 0.5])
    v = [2.0, -1.0, 3.0]
    assert np.array_equal(apply(a + b, v, cuda), [0.0, -6.0, 10.5]) and (a + b).dimension() == 3
    assert np.array_equal(apply(a - b, v, cuda), [4.0, 2.0, 7.5])
    a, b = diag_op([2.0, 0.5, -1.0]), diag_op([3.0, -2.0, 4.0])
    assert np.array_equal(apply(a * b, [1.0, -1.0, 2.0], cuda), [6.0, 1.0, -8.0])
    ident = mb.Identity(mb.FermionBasis.sector(3, 1, 0))
    assert np.array_equal(apply(ident, [1.0, -2.0, 5.0], cuda), [1.0, -2.0, 5.0])
    assert ident.dimension() == 3


def test_complex_scalars_and_nesting(cuda):
    """Complex scale and shift, and a nested composite, against numpy."""
    d = np.array([0.5, -1.25, 2.0])
    a, b = diag_op(d), diag_op([1.0, 3.0, -0.5])
    c = 0.75 - 1.5j
    v = np.array([1.0 + 2.0j, -0.5j, 3.0 - 1.0j])
    e = (mb.Shifted(a * c, 0.25j) - b) * a
    want = (c * d + 0.25j - np.array([1.0, 3.0, -0.5])) * (d * v)
    assert np.abs(apply(e, v, cuda) - want).max() < 1e-15


def test_krylov_on_composites(cuda):
    """Lanczos, eigenpairs and time evolution through composites: 1D ring L = 8,
    4 + 4, U = 4 (D = 4,900)."""
    basis = mb.FermionBasis.sector(8, 4, 4)
    model = mb.HubbardModel(1.0, 4.0, 8)
    H = model.operator(basis)
    b = H.gaussian(seed=3)
    e0 = mb.find_min_eigenpair(H, 200, b, want_vector=False).value
    emax = mb.find_max_eigenpair(H, 200, b, want_vector=False).value
    tol = 1e-10 * abs(e0)
    assert abs(mb.find_min_eigenpair(mb.Shifted(H, 3.0), 200, b, False).value - (e0 + 3.0)) < tol
    assert abs(mb.find_min_eigenpair(2.0 * H, 200, b, False).value - 2.0 * e0) < 2 * tol
    assert abs(mb.find_min_eigenpair(-H, 200, b, False).value + emax) < tol
    assert abs(mb.find_min_eigenpair(mb.NegShift(H, 1.0), 200, b, False).value
               - (1.0 - emax)) < tol
    assert abs(mb.find_min_eigenpair(H - 2.0 * mb.Identity(basis), 200, b, False).value
               - (e0 - 2.0)) < tol
    # kinetic (native real-space kernel) + interaction (compiled expression)
    kin = mb.HubbardModel(1.0, 0.0, 8).operator(basis)
    inter = mb.FermionExpressionOperator(
        basis, sum((4.0 * mb.density(mb.Spin.Up, i) * mb.density(mb.Spin.Down, i)
                    for i in range(8)), mb.density(mb.Spin.Up, 0) * 0.0))
    assert abs(mb.find_min_eigenpair(kin + inter, 200, b, False).value - e0) < tol
    # a shift only adds the phase exp(-i c t)
    psi = H.gaussian(seed=9)
    y0 = mb.time_evolve_state(H, psi, 0.7)  # device order: planes (re, im)
    y1 = mb.time_evolve_state(mb.Shifted(H, 0.4), psi, 0.7)
    z0 = y0[0] + 1j * y0[1]
    z1 = y1[0] + 1j * y1[1]
    assert torch.abs(z1 - z0 * np.exp(-0.4j * 0.7)).max().item() < 1e-10


def test_composed_square_spectrum(cuda):
    """min eig of H H = min lambda^2 over the spectrum (ring L = 6, 3 + 3, U = 3,
    dense spectrum from the oracle's CSC)."""
    basis = mb.FermionBasis.sector(6, 3, 3)
    H = mb.HubbardModel(1.0, 3.0, 6).operator(basis)
    lam = np.linalg.eigvalsh(po.build_csc([6], 3, 3, 1.0, 3.0).to_scipy().toarray())
    got = mb.find_min_eigenpair(H * H, 400, H.gaussian(seed=4), want_vector=False).value
    assert abs(got - np.min(lam ** 2)) < 1e-9 * max(1.0, np.max(lam ** 2))


@pytest.mark.parametrize("kind", ["momentum", "fixed_n"])
def test_sum_in_reference_order_on_grouped_bases(cuda, kind):
    """(A + B) and (A B) applied in reference order equal the operands applied one
    after another, on a fixed-momentum and a fixed-N basis."""
    if kind == "momentum":
        basis = mb.FermionBasis.with_fixed_particle_number_spin_momentum(9, 5, 1, [3, 3], [1, 2])
        A = mb.HubbardMomentumFactorizedFixedK(basis, [3, 3], 1.0, 4.0, [1, 2])
        B = mb.FermionExpressionOperator(basis, mb.HubbardModelMomentum(0.3, 1.0, [3, 3]).hamiltonian())
    else:
        basis = mb.FermionBasis.with_fixed_particle_number(6, 5)
        A = mb.HubbardModel(1.0, 4.0, 6).operator(basis)
        B = mb.FermionExpressionOperator(basis, mb.one_body(mb.Spin.Up, 0, mb.Spin.Down, 1)
                                         + mb.one_body(mb.Spin.Down, 1, mb.Spin.Up, 0))
    rng = np.random.default_rng(2)
    D = A.dimension()
    x = torch.tensor(rng.standard_normal(D) + 1j * rng.standard_normal(D), device=cuda)
    ax, bx = A.apply(x), B.apply(x)
    s = (A + B).apply(x)
    assert torch.abs(s - (ax + bx)).max().item() <= 1e-14 * torch.abs(s).max().item()
    c = (A * B).apply(x)
    want = A.apply(bx)
    assert torch.abs(c - want).max().item() <= 1e-14 * torch.abs(want).max().item()


def test_composite_rejects_mismatched_operands(cuda):
    a = mb.HubbardModel(1.0, 4.0, 6).operator(mb.FermionBasis.sector(6, 3, 3))
    b = mb.HubbardModel(1.0, 4.0, 6).operator(mb.FermionBasis.sector(6, 2, 4))
    with pytest.raises(Exception):
        a + b
    c = mb.HubbardModel(1.0, 4.0, 6).operator(mb.FermionBasis.sector(6, 3, 3))
    assert (a - c).dimension() == a.dimension()


def small_op(d, hop=0.0):
    """sum_i d_i n_{up,i} (+ hop (c+_0 c_1 + h.c.)) on the len(d)-orbital sector with
    one up particle: the matrix diag(d) with H(0, 1) = H(1, 0) = hop."""
    basis = mb.FermionBasis.sector(len(d), 1, 0)
    e = mb.density(mb.Spin.Up, 0) * 0.0
    for i, v in enumerate(d):
        e += mb.density(mb.Spin.Up, i) * float(v)
    if hop:
        e += mb.hopping(hop, 0, 1, mb.Spin.Up)
    op = mb.FermionExpressionOperator(basis, e)
    op._keep = basis
    return op


def dense_of(op, cuda):
    D = op.dimension()
    return np.stack([apply(op, np.eye(D)[j], cuda) for j in range(D)], axis=1)


def test_reference_rescaling_operator_cases(cuda):
    """test_rescaling.cpp:46-132 on device operators: estimate_rescaling puts the
    spectrum inside (-1, 1) and, without padding, maps diag(0, 1, 2) onto
    [-1, 1] exactly; rescale_hamiltonian keeps the eigenvectors and equals
    (H - b I) / a."""
    H = small_op([-2.0, -1.0, 1.0, 3.0])
    r = mb.estimate_rescaling(H)
    w = np.linalg.eigvalsh(dense_of(mb.rescale_hamiltonian(H, r), cuda))
    assert w.min() > -1.0 and w.max() < 1.0
    H = small_op([0.0, 1.0, 2.0])
    r0 = mb.estimate_rescaling(H, 0.0)
    assert abs(r0.a - 1.0) < 1e-10 and abs(r0.b - 1.0) < 1e-10
    r1 = mb.estimate_rescaling(H, 0.1)
    assert r1.a > r0.a and abs(r1.b - 1.0) < 1e-10
    H = small_op([-1.0, 0.0, 2.0], hop=0.5)
    M = dense_of(H, cuda)
    assert abs(M[0, 1] - 0.5) < 1e-15 and abs(M[1, 0] - 0.5) < 1e-15
    _, V0 = np.linalg.eigh(M)
    _, V1 = np.linalg.eigh(dense_of(mb.rescale_hamiltonian(H, mb.estimate_rescaling(H)), cuda))
    for i in range(3):
        assert abs(abs(np.vdot(V0[:, i], V1[:, i])) - 1.0) < 1e-10
    H = small_op([-2.0, -1.0, 1.0, 3.0], hop=0.5)
    r = mb.Rescaling.from_bounds(-2.5, 3.5)
    want = (dense_of(H, cuda) - r.b * np.eye(4)) / r.a
    assert np.linalg.norm(dense_of(mb.rescale_hamiltonian(H, r), cuda) - want) < 1e-14
    with pytest.raises(ValueError):
        mb.rescale_hamiltonian(H, mb.Rescaling(0.0, 1.0))


def test_operator_overloads_reject_non_operators(cuda):
    """A + c and A - c are not defined (the reference shifts with Shifted); numpy
    scalars scale."""
    H = small_op([1.0, 2.0, 3.0])
    with pytest.raises(TypeError):
        H + 1.0
    with pytest.raises(TypeError):
        H - 1.0
    with pytest.raises(TypeError):
        H * "x"
    assert np.array_equal(apply(np.float64(2.0) * H, [1.0, 1.0, 1.0], cuda), [2.0, 4.0, 6.0])
