"""GPU observables of a state: the opposite-spin correlation G_updn(r) for every
displacement (hubbard_model.h:40-53, 105-120, 184-200) and the Rayleigh
quotient / residual pair.

Small sectors: against the oracle (which is pinned to the reference's own
correlation outputs) at 1e-12 relative.  Full 4 x 4 7+7 size: identities that
hold for every state -- the sum rule sum_r G(r) = N_up N_dn / N, and
E(t, U) - E(t, 0) = U N G(0) with the energies from the independent H.v path.
"""
import ctypes as C

import numpy as np
import pytest
import scipy.sparse.linalg as spla
import torch

import manybody_amd as mb
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

MODELS = {1: mb.HubbardModel, 2: mb.HubbardModel2D, 3: mb.HubbardModel3D}


def make_op(sizes, nu, nd, t, U):
    basis = mb.FermionBasis.sector(int(np.prod(sizes)), nu, nd)
    return MODELS[len(sizes)](t, U, *sizes).operator(basis)


def random_ref(D, seed):
    rng = np.random.default_rng(seed)
    return rng.standard_normal(D) + 1j * rng.standard_normal(D)


@pytest.mark.parametrize("sizes,nu,nd", [([6], 3, 3), ([8], 3, 5), ([3, 2], 2, 3),
                                         ([4, 3], 3, 2), ([2, 2, 2], 2, 2)])
def test_correlation_matches_oracle(cuda, sizes, nu, nd):
    op = make_op(sizes, nu, nd, 1.0, 4.0)
    x_ref = random_ref(op.dimension(), seed=len(sizes) * 100 + nu)
    x = op.to_device_order(torch.from_numpy(x_ref).to(cuda))
    g = op.opposite_spin_correlation(x)
    g_orc = po.opposite_spin_correlation(sizes, nu, nd, x_ref)
    assert g.shape == g_orc.shape == (int(np.prod(sizes)),)
    assert np.abs(g - g_orc).max() <= 1e-12 * np.abs(g_orc).max()


def test_ground_state_correlation_and_residual(cuda):
    """1D L = 8, U = 4, 4 + 4: exact ground state (oracle matrix, eigsh) in,
    correlation against the oracle, energy against eigsh, residual ~ 0; then a
    random state against the oracle's Rayleigh quotient and residual."""
    sizes, nu, nd, t, U = [8], 4, 4, 1.0, 4.0
    op = make_op(sizes, nu, nd, t, U)
    A = po.build_csc(sizes, nu, nd, t, U)
    w, v = spla.eigsh(A.to_scipy().real, k=1, which="SA", tol=1e-14)
    psi = v[:, 0].astype(np.complex128)
    x = op.to_device_order(torch.from_numpy(psi).to(cuda))
    g = op.opposite_spin_correlation(x)
    g_orc = po.opposite_spin_correlation(sizes, nu, nd, psi)
    assert np.abs(g - g_orc).max() <= 1e-12 * np.abs(g_orc).max()
    e, r = op.energy_residual(x)
    assert abs(e - w[0]) <= 1e-12 * abs(w[0])
    assert r < 1e-9
    # random state: E = <x|A|x>/<x|x>, r = |Ax - E x| / |x|
    y_ref = random_ref(op.dimension(), seed=5)
    y = op.to_device_order(torch.from_numpy(y_ref).to(cuda))
    e, r = op.energy_residual(y)
    Ay = A.matvec(y_ref)
    e_orc = (np.vdot(y_ref, Ay) / np.vdot(y_ref, y_ref)).real
    r_orc = np.linalg.norm(Ay - e_orc * y_ref) / np.linalg.norm(y_ref)
    assert abs(e - e_orc) <= 1e-12 * abs(e_orc)
    assert abs(r - r_orc) <= 1e-12 * r_orc


def test_lanczos_eigenpair_residual(cuda):
    """Residual of find_min_eigenpair's vector on the L = 10 ring, 5 + 5
    (D = 63,504; odd N per spin: closed shell, non-degenerate ground state),
    and a symmetry of its correlation: the state is invariant under spin flip
    (N_up = N_dn), so G(r) = G(-r) must hold for the computed vector."""
    sizes, nu, nd = [10], 5, 5
    op = make_op(sizes, nu, nd, 1.0, 4.0)
    ep = mb.find_min_eigenpair(op, 300, op.gaussian(seed=11))
    e, r = op.energy_residual(ep.vector)
    assert abs(e - ep.value) <= 1e-10 * abs(ep.value)
    assert r < 1e-6
    g = op.opposite_spin_correlation(ep.vector)
    L = sizes[0]
    neg = [(-i) % L for i in range(L)]
    assert np.abs(g - g[neg]).max() < 1e-7
    assert abs(g.sum() - nu * nd / L) < 1e-12


def test_full_size_identities(cuda):
    """4 x 4, 7 + 7 (D = 130,873,600): sum rule for a random state, and
    E(t, U) - E(t, 0) = U N G(0) for the same state: the energies come from the
    H.v path of two models, G(0) from the correlation path."""
    basis = mb.FermionBasis.sector(16, 7, 7)
    op8 = mb.HubbardModel2D(1.0, 8.0, 4, 4).operator(basis)
    op0 = mb.HubbardModel2D(1.0, 0.0, 4, 4).operator(basis)
    x = op8.gaussian(seed=3)
    g = op8.opposite_spin_correlation(x)
    assert abs(g.sum() - 49 / 16) <= 1e-12 * 49 / 16
    e8, _ = op8.energy_residual(x)
    e0, _ = op0.energy_residual(x)
    assert abs((e8 - e0) - 8.0 * 16 * g[0]) <= 1e-12 * (abs(e8) + abs(e0))
    del x
    torch.cuda.empty_cache()


def test_bond_list_model_has_no_correlation(cuda):
    """A model made from a bond list has no lattice displacements: MB_ERR_INVALID."""
    lib = mb.lib()
    basis = mb.FermionBasis.sector(4, 2, 2)
    frm = (C.c_int * 3)(0, 1, 2)
    to = (C.c_int * 3)(1, 2, 3)
    h = C.c_void_p()
    assert lib.mb_hubbard_create_bonds(basis._h, 3, frm, to, 1.0, 2.0, C.byref(h)) == 0
    op = mb.HubbardOperator(h, basis, None)  # owns h from here on
    x = op.gaussian(seed=1)
    with pytest.raises(ValueError):
        op.opposite_spin_correlation(x)
    e, r = op.energy_residual(x)
    assert np.isfinite(e) and r > 0
