"""Fixed-momentum sectors on the GPU: the factorized momentum-space Hubbard
operator (hubbard_momentum_factorized_fixed_k.h) against the oracle, and
against the real-space model through the spectrum.

The H.v tests run each kernel (per-row and persistent, see KERNELS).
Tolerances: H.v 1e-12 relative to the largest output entry (different summation
order than the oracle's dense matrix); ground energies 1e-10 relative (Lanczos,
fully converged on these sizes) and 1e-8 at 4 x 4, 7 + 7 (300 steps on both
sides); norm and energy under exp(-iHt) 1e-10.
"""
import ctypes as C

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


def make_mop(sizes, nu, nd, K, t, U):
    N = int(np.prod(sizes))
    kc = kcoords(sizes, K)
    basis = mb.FermionBasis.with_fixed_particle_number_spin_momentum(N, nu + nd, nu - nd, sizes, kc)
    return mb.HubbardMomentumFactorizedFixedK(basis, sizes, t, U, kc)


CASES = [([6], 3, 2, 1.0, 4.0), ([5], 2, 2, 0.7, 3.0), ([3, 2], 2, 2, 1.0, 8.0),
         ([2, 2, 2], 2, 2, 1.0, 4.0), ([4, 2], 3, 2, 1.0, 6.0), ([4, 4], 2, 1, 1.0, 8.0)]


# H.v kernels: the per-row kernel, and the persistent one (taken by default from
# 2^20 states on) forced on small sectors
KERNELS = ["rows", "lock"]


@pytest.fixture(params=KERNELS)
def mom_kernel(request, monkeypatch):
    monkeypatch.setenv("MANYBODY_AMD_MOM_KERNEL", request.param)
    return request.param


@pytest.mark.parametrize("sizes,nu,nd,t,U", CASES)
def test_momentum_hv_matches_oracle(cuda, mom_kernel, sizes, nu, nd, t, U):
    N = int(np.prod(sizes))
    rng = np.random.default_rng(N * 10 + nu)
    for K in range(N):
        op = make_mop(sizes, nu, nd, K, t, U)
        D = op.dimension()
        Hk = po.momentum_hamiltonian(sizes, nu, nd, K, t, U)
        assert Hk.shape == (D, D)
        if D == 0:
            continue
        x_ref = rng.standard_normal(D) + 1j * rng.standard_normal(D)
        x = op.to_device_order(torch.from_numpy(x_ref).to(cuda))
        y = op.to_reference_order(op.hv(x)).cpu().numpy()
        want = Hk @ x_ref
        assert np.abs(y - want).max() <= 1e-12 * max(np.abs(want).max(), 1.0), (sizes, K)
        # the reference-order entry point (gather, H.v, scatter in one call)
        yr = torch.empty(D, dtype=torch.complex128, device=cuda)
        work = torch.empty(4 * D, dtype=torch.float64, device=cuda)
        mb._check(mb.lib().mb_apply_ref_c128(op._h, mb._ptr(torch.from_numpy(x_ref).to(cuda)),
                                             mb._ptr(yr), mb._ptr(work), mb._stream()))
        assert np.abs(yr.cpu().numpy() - want).max() <= 1e-12 * max(np.abs(want).max(), 1.0)


@pytest.mark.parametrize("sizes,nu,nd,U", [([8], 4, 4, 4.0), ([3, 3], 3, 2, 6.0),
                                           ([2, 2, 2], 3, 2, 4.0)])
def test_momentum_ground_state_is_the_real_space_one(cuda, mom_kernel, sizes, nu, nd, U):
    """min over K of the sector ground energies = the real-space ground energy."""
    N = int(np.prod(sizes))
    basis = mb.FermionBasis.sector(N, nu, nd)
    real = (mb.HubbardModel(1.0, U, sizes[0]) if len(sizes) == 1 else
            mb.HubbardModel2D(1.0, U, *sizes) if len(sizes) == 2 else
            mb.HubbardModel3D(1.0, U, *sizes)).operator(basis)
    e_real = mb.find_min_eigenpair(real, 200, real.gaussian(seed=1), want_vector=False).value
    e_k = []
    for K in range(N):
        op = make_mop(sizes, nu, nd, K, 1.0, U)
        if op.dimension() == 0:
            continue
        k = min(200, op.dimension())
        ep = mb.find_min_eigenpair(op, k, op.gaussian(seed=2 + K))
        e, r = op.energy_residual(ep.vector)
        assert abs(e - ep.value) <= 1e-9 * abs(ep.value) + 1e-12
        e_k.append(ep.value)
    assert abs(min(e_k) - e_real) <= 1e-10 * abs(e_real)


def test_momentum_time_evolution_conserves_norm_and_energy(cuda):
    op = make_mop([3, 3], 3, 2, 4, 1.0, 6.0)
    psi = op.gaussian(seed=9)
    e0, _ = op.energy_residual(psi)
    n0 = psi.norm().item()
    out = mb.time_evolve_state_steps(op, psi, 0.0, 0.5, 0.05, mb.EvolutionOptions(30))
    e1, _ = op.energy_residual(out)
    assert abs(out.norm().item() - n0) <= 1e-10 * n0
    assert abs(e1 - e0) <= 1e-10 * abs(e0)
    # the same evolution from the oracle's dense matrix (scipy expm)
    import scipy.linalg as sla
    Hk = po.momentum_hamiltonian([3, 3], 3, 2, 4, 1.0, 6.0)
    want = sla.expm(-0.5j * Hk) @ op.to_reference_order(psi).cpu().numpy()
    got = op.to_reference_order(out).cpu().numpy()
    assert np.linalg.norm(got - want) <= 1e-9 * np.linalg.norm(want)


def test_momentum_operator_rejects_real_space_only_calls(cuda):
    op = make_mop([6], 3, 2, 0, 1.0, 4.0)
    x = op.gaussian(seed=3)
    with pytest.raises(ValueError):
        op.hv(x.float())
    with pytest.raises(ValueError):  # FP32 KPM needs a real-space Hubbard operator
        mb.kpm_dos_moments(op, 30.0, 0.0, 8, 1, seed=1, fp32=True)
    with pytest.raises(ValueError):
        op.opposite_spin_correlation(x)
    basis = mb.FermionBasis.with_fixed_particle_number_spin_momentum(6, 5, 1, [6], [0])
    with pytest.raises(ValueError):
        mb.HubbardModel(1.0, 4.0, 6).operator(basis)
    with pytest.raises(ValueError):  # wrong target momentum for the basis
        mb.HubbardMomentumFactorizedFixedK(basis, [6], 1.0, 4.0, [1])


def test_momentum_4x4_7p7_ground_state(cuda):
    """4 x 4, 7 + 7 at U = 8: the lowest sector energy over the momenta matches the
    real-space 300-step Lanczos energy on the full D = 130,873,600.  The square
    lattice's point group maps every K to one of (0,0), (1,0), (1,1), (2,0),
    (2,1), (2,2) with the same spectrum, so those six sectors suffice."""
    basis = mb.FermionBasis.sector(16, 7, 7)
    real = mb.HubbardModel2D(1.0, 8.0, 4, 4).operator(basis)
    dec = mb.lanczos_pass_one(real, real.gaussian(seed=20260101), 300)
    e_real = mb.tridiag_eig(dec.alphas, dec.betas, vectors=False)[0][0]
    del real, basis
    torch.cuda.empty_cache()
    e_k = []
    for kx, ky in [(0, 0), (1, 0), (1, 1), (2, 0), (2, 1), (2, 2)]:
        op = make_mop([4, 4], 7, 7, kx + 4 * ky, 1.0, 8.0)
        assert 7_500_000 < op.dimension() < 8_800_000  # about D / 16
        d = mb.lanczos_pass_one(op, op.gaussian(seed=kx + 4 * ky), 300)
        e_k.append(mb.tridiag_eig(d.alphas, d.betas, vectors=False)[0][0])
    assert abs(min(e_k) - e_real) <= 1e-8 * abs(e_real)


def test_persistent_kernel_equals_per_row_kernel_at_4x4_7p7(cuda, monkeypatch):
    """4 x 4, 7 + 7 at U = 8, sectors K = (0,0) and (2,1) (D_K about 8.2 M): the
    persistent kernel, the default at this size, against the per-row kernel.
    H.v to 1e-14 of the largest entry (the two sum the same terms in different
    orders); 20 Lanczos coefficients to 1e-12 relative."""
    for K in (0, 2 + 4 * 1):
        op = make_mop([4, 4], 7, 7, K, 1.0, 8.0)
        x = op.gaussian(seed=77 + K)
        out = {}
        for kind in ("rows", "lock"):
            monkeypatch.setenv("MANYBODY_AMD_MOM_KERNEL", kind)
            y = op.hv(x)
            dec = mb.lanczos_pass_one(op, x, 20)
            out[kind] = (y, np.array(dec.alphas), np.array(dec.betas))
            print(f"K={K} kernel {kind}: |Hx| = {y.abs().max().item():.6f}", flush=True)
        y0, a0, b0 = out["rows"]
        scale = y0.abs().max().item()
        for key, (y, a, b) in out.items():
            assert (y - y0).abs().max().item() <= 1e-14 * scale, (K, key)
            assert np.abs(a - a0).max() <= 1e-12 * np.abs(a0).max(), (K, key)
            assert np.abs(b - b0).max() <= 1e-12 * np.abs(b0).max(), (K, key)
        del op, x, out
        torch.cuda.empty_cache()
