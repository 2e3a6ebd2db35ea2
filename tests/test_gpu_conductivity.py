"""Optical conductivity on the GPU: the current operator J_d(Q) between momentum
sectors (current_operator_fixed_k.h), imaginary-time evolution
(evolve_state.h:184-197) and the stochastic current-current correlator of
hubbard_momentum_optical_conductivity_factorized.cpp:226-283.

J and J^dagger are checked against the oracle's dense J (itself pinned to the
compiled reference's CurrentOperatorFixedK) at 1e-13 relative; exp(-tau H)
against scipy's expm of the oracle's dense H_K at 1e-10; the correlator against
its spectral form computed from eigendecompositions of the dense H_K and
H_{K+Q} with the same seeded start vectors, at 1e-9 relative to C(0).

Vectors of sector K with momentum components kcoords(K) are device-order planes
(2, D); the operators' to_device_order / to_reference_order convert.
"""
import numpy as np
import pytest
import torch

import manybody_amd as mb
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu


def kcoords(sizes, K):
    c = [K % sizes[0]]
    if len(sizes) > 1:
        c.append((K // sizes[0]) % sizes[1])
    if len(sizes) > 2:
        c.append(K // (sizes[0] * sizes[1]))
    return c


def sector(sizes, nu, nd, K, t=1.0, U=4.0):
    N = int(np.prod(sizes))
    kc = kcoords(sizes, K)
    b = mb.FermionBasis.with_fixed_particle_number_spin_momentum(N, nu + nd, nu - nd, sizes, kc)
    return b, mb.HubbardMomentumFactorizedFixedK(b, sizes, t, U, kc)


def ref(op, x):
    return op.to_reference_order(x).cpu().numpy()


def dev(op, z, cuda):
    return op.to_device_order(torch.from_numpy(np.ascontiguousarray(z)).to(cuda))


# (sizes, n_up, n_dn, Q, direction, t); every K of the lattice is tried
CASES = [([6], 3, 2, 2, 0, 1.0), ([5], 2, 2, 1, 0, 0.7), ([3, 2], 2, 2, 3, 1, 1.0),
         ([4, 2], 3, 2, 1, 0, 1.0), ([2, 2, 2], 2, 2, 5, 2, 1.0), ([4], 2, 2, 0, 0, 1.0),
         ([3, 3], 2, 2, 0, 1, 1.3), ([4, 4], 2, 1, 1, 0, 1.0), ([4, 4], 2, 2, 5, 1, 1.0)]


@pytest.mark.parametrize("sizes,nu,nd,Q,d,t", CASES)
def test_current_operator_matches_oracle(cuda, sizes, nu, nd, Q, d, t):
    N = int(np.prod(sizes))
    rng = np.random.default_rng(N * 100 + Q * 10 + d)
    for K in range(N):
        KQ = po.momentum_add(sizes, K, Q)
        bk, hk = sector(sizes, nu, nd, K)
        bq, hq = sector(sizes, nu, nd, KQ)
        J = mb.CurrentOperatorFixedK(bk, bq, sizes, t, Q, d)
        Jd = po.current_operator(sizes, nu, nd, K, Q, d, t)
        assert Jd.shape == (J.target_dimension(), J.source_dimension())
        if 0 in Jd.shape:
            continue
        x = rng.standard_normal(Jd.shape[1]) + 1j * rng.standard_normal(Jd.shape[1])
        y = rng.standard_normal(Jd.shape[0]) + 1j * rng.standard_normal(Jd.shape[0])
        want = Jd @ x
        got = ref(hq, J.apply(dev(hk, x, cuda)))
        scale = max(np.abs(want).max(), 1.0)
        assert np.abs(got - want).max() <= 1e-13 * scale, (sizes, K)
        want = Jd.conj().T @ y
        got = ref(hk, J.adjoint_apply(dev(hq, y, cuda)))
        scale = max(np.abs(want).max(), 1.0)
        assert np.abs(got - want).max() <= 1e-13 * scale, (sizes, K)


def test_current_operator_adjoint_identity_on_4x4_7p7(cuda):
    """<y|J x> = <J^dagger y|x> on the 4 x 4, 7 + 7 sectors K = (0,0) -> (1,0)
    (D about 8.2 million each): a size-independent property of the pair."""
    bk, hk = sector([4, 4], 7, 7, 0, 1.0, 8.0)
    bq, hq = sector([4, 4], 7, 7, 1, 1.0, 8.0)
    J = mb.CurrentOperatorFixedK(bk, bq, [4, 4], 1.0, 1, 0)
    x, y = hk.gaussian(seed=3), hq.gaussian(seed=4)
    jx, jty = J.apply(x), J.adjoint_apply(y)
    cx = torch.complex(jx[0], jx[1])
    cy = torch.complex(y[0], y[1])
    lhs = torch.vdot(cy, cx).item()
    rhs = torch.vdot(torch.complex(jty[0], jty[1]), torch.complex(x[0], x[1])).item()
    assert abs(lhs - rhs) <= 1e-11 * abs(lhs)
    # J conserves nothing else: the image is not trivially zero
    assert torch.linalg.vector_norm(jx).item() > 0.1 * torch.linalg.vector_norm(x).item()


@pytest.mark.parametrize("sizes,nu,nd,K,tau,k", [([6], 3, 2, 1, 0.7, 60), ([3, 2], 2, 2, 3, 1.5, 0),
                                                 ([3, 3], 3, 2, 4, 0.2, 30)])
def test_imaginary_time_evolution_matches_expm(cuda, sizes, nu, nd, K, tau, k):
    import scipy.linalg as sla
    _, hk = sector(sizes, nu, nd, K, 1.0, 6.0)
    Hk = po.momentum_hamiltonian(sizes, nu, nd, K, 1.0, 6.0)
    psi = hk.gaussian(seed=11)
    out = mb.imaginary_time_evolve_state(hk, psi, tau, mb.EvolutionOptions(k))
    want = sla.expm(-tau * Hk) @ ref(hk, psi)
    assert np.linalg.norm(ref(hk, out) - want) <= 1e-10 * np.linalg.norm(want)


def spectral_correlator(Hk, Hq, Jd, starts, beta, dt, steps):
    """C(i dt) = avg_n <v_b| e^{i H_K t} J^dagger e^{-i H_{K+Q} t} J |v_b> / Z from
    eigendecompositions (no time stepping)."""
    ek, uk = np.linalg.eigh(Hk)
    eq, uq = np.linalg.eigh(Hq)
    Jq = uq.conj().T @ Jd @ uk          # J in the eigenbases: <b|J|a>
    C = np.zeros(steps, dtype=complex)
    for v in starts:
        v = v / np.linalg.norm(v)
        a = np.exp(-0.5 * beta * ek) * (uk.conj().T @ v)   # v_b in the H_K eigenbasis
        Z = np.vdot(a, a).real
        ja = Jq @ a
        for i in range(steps):
            tt = i * dt
            phi1 = np.exp(-1j * ek * tt) * a
            phi2 = np.exp(-1j * eq * tt) * ja
            C[i] += np.vdot(phi1, Jq.conj().T @ phi2) / Z
    return C / len(starts)


@pytest.mark.parametrize("sizes,nu,nd,K,Q,d", [([6], 3, 2, 1, 1, 0), ([3, 2], 2, 2, 0, 3, 0),
                                               ([4], 2, 2, 0, 0, 0), ([3, 3], 2, 2, 1, 4, 1)])
def test_optical_correlator_matches_spectral_form(cuda, sizes, nu, nd, K, Q, d):
    t, U, beta, dt, steps, samples, seed = 1.0, 4.0, 2.0, 0.1, 12, 3, 77
    KQ = po.momentum_add(sizes, K, Q)
    bk, hk = sector(sizes, nu, nd, K, t, U)
    bq, hq = sector(sizes, nu, nd, KQ, t, U)
    J = mb.CurrentOperatorFixedK(bk, bq, sizes, t, Q, d)
    C = mb.optical_correlator(hk, hq, J, samples, beta, dt, steps, seed, mb.EvolutionOptions(40))
    starts = [ref(hk, hk.gaussian(seed=seed + n)) for n in range(samples)]
    want = spectral_correlator(po.momentum_hamiltonian(sizes, nu, nd, K, t, U),
                               po.momentum_hamiltonian(sizes, nu, nd, KQ, t, U),
                               po.current_operator(sizes, nu, nd, K, Q, d, t),
                               starts, beta, dt, steps)
    assert C.shape == (steps,)
    # (a side of length 2 has v_d = 2 t pi sin(pi k_d) = 0: no such case here)
    assert abs(C[0].imag) <= 1e-12 * abs(C[0]) and C[0].real > 0  # <J v_b|J v_b> / Z
    assert np.abs(C - want).max() <= 1e-9 * abs(want[0])
    # sigma(omega) of the GPU correlator, the host transform against the oracle's
    vol = float(np.prod(sizes))
    r = mb.compute_optical_conductivity(C, dt, beta, vol)
    w, sg, sr = po.optical_conductivity(C, dt, beta, vol)
    assert np.abs(r.sigma - sg).max() <= 1e-12 * np.abs(sg).max()
    assert abs(r.sum_rule - sr) <= 1e-12 * abs(sr)


def test_current_operator_rejects_invalid_arguments(cuda):
    bk, _ = sector([6], 3, 2, 1)
    bq, _ = sector([6], 3, 2, 3)
    with pytest.raises(ValueError):  # target momentum is not K + Q
        mb.CurrentOperatorFixedK(bk, bq, [6], 1.0, 1, 0)
    with pytest.raises(ValueError):  # direction beyond the lattice dimensionality
        mb.CurrentOperatorFixedK(bk, bq, [6], 1.0, 2, 1)
    with pytest.raises(ValueError):  # real-space bases
        real = mb.FermionBasis.sector(6, 3, 2)
        mb.CurrentOperatorFixedK(real, real, [6], 1.0, 0, 0)
    with pytest.raises(ValueError):  # lattice does not match the bases
        mb.CurrentOperatorFixedK(bk, bq, [3, 2], 1.0, 2, 0)


def test_correlator_with_expression_hamiltonians(cuda):
    """The correlator accepts any operator on the current operator's bases: the
    momentum-space Hubbard model compiled as an expression gives the factorized
    operator's correlator (1e-10 of C(0))."""
    sizes, nu, nd, K, Q = [6], 3, 2, 1, 1
    bk, hk = sector(sizes, nu, nd, K, 1.0, 4.0)
    bq, hq = sector(sizes, nu, nd, po.momentum_add(sizes, K, Q), 1.0, 4.0)
    J = mb.CurrentOperatorFixedK(bk, bq, sizes, 1.0, Q, 0)
    e = mb.HubbardModelMomentum(1.0, 4.0, sizes).hamiltonian()
    ek = mb.FermionExpressionOperator(bk, e)
    eq = mb.FermionExpressionOperator(bq, e)
    C0 = mb.optical_correlator(hk, hq, J, 2, 1.5, 0.1, 8, 11)
    C1 = mb.optical_correlator(ek, eq, J, 2, 1.5, 0.1, 8, 11)
    assert np.abs(C1 - C0).max() <= 1e-10 * abs(C0[0])
    with pytest.raises(ValueError):  # operators on other bases
        mb.optical_correlator(hq, hk, J, 1, 1.0, 0.1, 2, 1)


def _mbasis(sizes, n, K):
    return mb.FermionBasis.with_fixed_particle_number_spin_momentum(
        int(np.prod(sizes)), n, 0, sizes, K)


def _kflat(sizes, K):
    return K[0] + (sizes[0] * K[1] if len(K) > 1 else 0)


def _check_against_rectangular(cuda, sizes, K, Q, direction, n, rng, t=1.0):
    """CurrentOperatorFixedK (device order) against compute_rectangular_matrix_
    elements of the symbolic current, both directions, absolute 1e-10 as the
    reference checks (test_current_operator_fixed_k.cpp:48-172)."""
    KQ = [(k + q) % s for k, q, s in zip(K, Q, sizes)]
    bk, bq = _mbasis(sizes, n, K), _mbasis(sizes, n, KQ)
    if bk.size() == 0 or bq.size() == 0:
        return False
    model = mb.HubbardModelMomentum(t, 4.0, sizes)
    Jd = mb.compute_rectangular_matrix_elements(bq, bk, model.current(Q, direction)).to_dense()
    J = mb.CurrentOperatorFixedK(bk, bq, sizes, t, _kflat(sizes, Q), direction)
    hk = mb.HubbardMomentumFactorizedFixedK(bk, sizes, t, 4.0, K)
    hq = mb.HubbardMomentumFactorizedFixedK(bq, sizes, t, 4.0, KQ)
    v = rng.standard_normal(bk.size()) + 1j * rng.standard_normal(bk.size())
    assert np.abs(ref(hq, J.apply(dev(hk, v, cuda))) - Jd @ v).max() < 1e-10
    u = rng.standard_normal(bq.size()) + 1j * rng.standard_normal(bq.size())
    assert np.abs(ref(hk, J.adjoint_apply(dev(hq, u, cuda))) - Jd.conj().T @ u).max() < 1e-10
    return True


def test_reference_current_operator_cases_1d(cuda):
    """test_current_operator_fixed_k.cpp:26-129: the 4-site ring with 2 particles
    at S_z = 0, every K and Q, the operator and its adjoint against the
    rectangular matrix of HubbardModelMomentum::current."""
    rng = np.random.default_rng(11)
    bk, bq = _mbasis([4], 2, [0]), _mbasis([4], 2, [1])
    J = mb.CurrentOperatorFixedK(bk, bq, [4], 1.0, 1, 0)
    assert J.source_dimension() == bk.size() and J.target_dimension() == bq.size()
    done = sum(_check_against_rectangular(cuda, [4], [K], [Q], 0, 2, rng)
               for K in range(4) for Q in range(4))
    assert done > 8


def test_reference_current_operator_cases_2d_and_q0(cuda):
    """test_current_operator_fixed_k.cpp:131-241: 3x3 with 2 particles, Q = (1, 1),
    both directions; Q = 0 is diagonal (here: J(0) = sum_k v(k) n_k, the
    rectangular matrix of one sector with itself); the adjoint identity
    <u|J v> = <J^dagger u|v> at Q = 2 on the ring."""
    rng = np.random.default_rng(12)
    for d in (0, 1):
        assert _check_against_rectangular(cuda, [3, 3], [1, 0], [1, 1], d, 2, rng)
    b = _mbasis([4], 2, [1])
    Jd = mb.compute_rectangular_matrix_elements(
        b, b, mb.HubbardModelMomentum(1.0, 4.0, [4]).current([0], 0)).to_dense()
    assert np.abs(Jd - np.diag(np.diag(Jd))).max() < 1e-10
    assert _check_against_rectangular(cuda, [4], [1], [0], 0, 2, rng)
    bk, bq = _mbasis([4], 2, [1]), _mbasis([4], 2, [3])
    J = mb.CurrentOperatorFixedK(bk, bq, [4], 1.0, 2, 0)
    hk = mb.HubbardMomentumFactorizedFixedK(bk, [4], 1.0, 4.0, [1])
    hq = mb.HubbardMomentumFactorizedFixedK(bq, [4], 1.0, 4.0, [3])
    v = rng.standard_normal(bk.size()) + 1j * rng.standard_normal(bk.size())
    u = rng.standard_normal(bq.size()) + 1j * rng.standard_normal(bq.size())
    jv = ref(hq, J.apply(dev(hk, v, cuda)))
    jtu = ref(hk, J.adjoint_apply(dev(hq, u, cuda)))
    assert abs(np.vdot(u, jv) - np.vdot(jtu, v)) < 1e-10


def test_reference_current_velocity_values(cuda):
    """test_current_operator_fixed_k.cpp:276-300: the symbolic current's
    coefficients are v_d(k) = 2t (2 pi / L) sin(2 pi k / L), k = 0 and k = L/2
    skipped (v = 0, or below 1000 eps after sin(pi) rounding)."""
    t, L = 1.3, 6
    e = mb.HubbardModelMomentum(t, 2.0, [L]).current([1], 0)
    for k in range(L):
        v = 2 * t * (2 * np.pi / L) * np.sin(2 * np.pi * k / L)
        key = (((k + 1) % L), 0x8000 | k)
        if abs(v) < 1e-12:
            assert key not in e.terms
        else:
            assert abs(e.terms[key] - v) < 1e-15
