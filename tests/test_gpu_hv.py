"""GPU parity of the matrix-free H.v kernels against the oracle.

Integer-valued inputs (entries in [-4, 4]) make every sum exact in FP64 (and in
FP32 for these sizes), so the device result must equal the oracle's bit for
bit: that pins indices, signs, bonds and the diagonal.  Gaussian inputs are
checked to a relative 1e-14 (FP64; the summation order differs from the
oracle's).  The full 4 x 4, 7 + 7 sector (D = 130,873,600) is checked bit for
bit too.
"""
import numpy as np
import pytest

import manybody_amd as mb
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

CASES = [
    ("1d_L8_4_4", [8], 4, 4, 1.0, 4.0),
    ("1d_L10_5_5", [10], 5, 5, 1.0, 6.0),
    ("1d_L12_6_5", [12], 6, 5, 1.0, 3.0),
    ("1d_L7_0_3", [7], 0, 3, 1.0, 2.0),
    ("2d_2x2_2_1", [2, 2], 2, 1, 1.0, 4.0),
    ("2d_3x3_4_3", [3, 3], 4, 3, 1.0, 8.0),
    ("2d_4x3_3_3", [4, 3], 3, 3, 0.5, 2.5),
    ("2d_3x1_2_1_selfbond", [3, 1], 2, 1, 1.0, 2.0),
    ("2d_4x4_3_2", [4, 4], 3, 2, 1.0, 8.0),
    ("3d_2x2x2_3_3", [2, 2, 2], 3, 3, 1.0, 4.0),
    ("2d_4x2_3_3_two_magnitudes", [4, 2], 3, 3, 1.0, 4.0),
    ("2d_2x3_2_2_two_magnitudes", [2, 3], 2, 2, 1.0, 2.5),
]

# amplitudes that are not exact binary fractions: parity at a relative 1e-14
FP_CASES = [
    ("2d_2x3_2_2_t07", [2, 3], 2, 2, 0.7, 2.3),
    ("1d_L10_5_4_t03", [10], 5, 4, -0.3, 1.7),
]


def model_of(sizes, t, U):
    if len(sizes) == 1:
        return mb.HubbardModel(t, U, sizes[0])
    if len(sizes) == 2:
        return mb.HubbardModel2D(t, U, *sizes)
    return mb.HubbardModel3D(t, U, *sizes)


def make_op(sizes, nu, nd, t, U):
    L = int(np.prod(sizes))
    basis = mb.FermionBasis.sector(L, nu, nd)
    return model_of(sizes, t, U).operator(basis)


# H.v paths: the two passes (up hops, then the row pass) with the up hops from
# the block-staged kernel (the default where its tables exist) or from the
# per-row strip kernel, and the one-pass row kernel that forms the up hops
# itself (the default for sectors whose plane group fits the Infinity Cache);
# the bit-exact tests run all three
@pytest.fixture(params=[("0", None), ("0", "strip"), ("1", None)],
                ids=["two_pass", "two_pass_strip", "one_pass"])
def hv_path(request, monkeypatch):
    fused, up = request.param
    monkeypatch.setenv("MANYBODY_AMD_FUSED", fused)
    if up is None:
        monkeypatch.delenv("MANYBODY_AMD_UP_KERNEL", raising=False)
    else:
        monkeypatch.setenv("MANYBODY_AMD_UP_KERNEL", up)
    return request.param


def int_vec(rng, n):
    return (rng.integers(-4, 5, n) + 1j * rng.integers(-4, 5, n)).astype(np.complex128)


@pytest.mark.parametrize("name,sizes,nu,nd,t,U", CASES, ids=[c[0] for c in CASES])
def test_hv_c128_bitexact_vs_product_oracle(cuda, hv_path, name, sizes, nu, nd, t, U):
    import torch
    op = make_op(sizes, nu, nd, t, U)
    P = po.ProductOperator(sizes, nu, nd, t, U)
    assert (op.d_up, op.d_dn) == (P.Dup, P.Ddn)
    x = int_vec(np.random.default_rng(1), P.D)
    y = mb.from_planar(op.hv(mb.to_planar(x)))
    assert np.array_equal(y, P.matvec(x))


@pytest.mark.parametrize("name,sizes,nu,nd,t,U", CASES[:6], ids=[c[0] for c in CASES[:6]])
def test_apply_reference_order_bitexact_vs_csc_oracle(cuda, hv_path, name, sizes, nu, nd, t, U):
    """The whole boundary: reference order in, gather, H.v, scatter, reference order out."""
    import torch
    op = make_op(sizes, nu, nd, t, U)
    A = po.build_csc(sizes, nu, nd, t, U)
    x = int_vec(np.random.default_rng(2), A.D)
    y = op.apply(torch.from_numpy(x).to(cuda)).cpu().numpy()
    assert np.array_equal(y, A.matvec(x))


# edge sectors: one site (self bonds on every axis), 32 sites (bit 31 of the
# strings), a full species, the vacuum, a 3D lattice with a size-1 axis
EDGE_CASES = [
    ("1d_L1_1_1_selfbond", [1], 1, 1, 1.0, 3.0),
    ("1d_L32_1_1_bit31", [32], 1, 1, 1.0, 4.0),
    ("1d_L32_2_1_bit31", [32], 2, 1, 0.5, 4.0),
    ("1d_L6_6_3_full_up", [6], 6, 3, 1.0, 4.0),
    ("1d_L5_0_0_vacuum", [5], 0, 0, 1.0, 4.0),
    ("3d_2x2x1_2_2", [2, 2, 1], 2, 2, 1.0, 4.0),
    ("3d_1x1x1_1_0", [1, 1, 1], 1, 0, 1.0, 4.0),
]


@pytest.mark.parametrize("name,sizes,nu,nd,t,U", EDGE_CASES, ids=[c[0] for c in EDGE_CASES])
def test_hv_edge_sectors_bitexact(cuda, hv_path, name, sizes, nu, nd, t, U):
    """Integer inputs, bit for bit against the product oracle, and through the
    reference-order boundary against the CSC oracle where it applies (L <= 31)."""
    import torch
    op = make_op(sizes, nu, nd, t, U)
    P = po.ProductOperator(sizes, nu, nd, t, U)
    assert (op.d_up, op.d_dn) == (P.Dup, P.Ddn)
    x = int_vec(np.random.default_rng(3), P.D)
    assert np.array_equal(mb.from_planar(op.hv(mb.to_planar(x))), P.matvec(x))
    if int(np.prod(sizes)) <= 31:
        A = po.build_csc(sizes, nu, nd, t, U)
        y = op.apply(torch.from_numpy(x).to(cuda)).cpu().numpy()
        assert np.array_equal(y, A.matvec(x))
    # the Krylov layer on the sector: a ground energy (D = 1 breaks down at once)
    ep = mb.find_min_eigenpair(op, min(60, P.D), op.gaussian(seed=2), want_vector=False)
    if P.D <= 64:
        A = po.build_csc(sizes, nu, nd, t, U).to_scipy().toarray()
        assert abs(ep.value - np.linalg.eigvalsh(A)[0]) <= 1e-10 * max(1.0, abs(ep.value))


def test_permutation_matches_oracle_and_roundtrips(cuda):
    import torch
    sizes, nu, nd = [3, 3], 4, 3
    op = make_op(sizes, nu, nd, 1.0, 1.0)
    perm = po.product_permutation(sizes, nu, nd)
    x_ref = np.arange(op.dimension()).astype(np.complex128) * (1 + 2j)
    xd = op.to_device_order(torch.from_numpy(x_ref).to(cuda))
    expect = np.zeros_like(x_ref)
    expect[perm] = x_ref
    assert np.array_equal(mb.from_planar(xd), expect)
    back = op.to_reference_order(xd).cpu().numpy()
    assert np.array_equal(back, x_ref)


@pytest.mark.parametrize("name,sizes,nu,nd,t,U", CASES[:4], ids=[c[0] for c in CASES[:4]])
def test_hv_c64_bitexact_vs_product_oracle(cuda, hv_path, name, sizes, nu, nd, t, U):
    import torch
    op = make_op(sizes, nu, nd, t, U)
    P = po.ProductOperator(sizes, nu, nd, t, U)
    x = int_vec(np.random.default_rng(3), P.D)
    y = mb.from_planar(op.hv(mb.to_planar(x, dtype=torch.float32)))
    assert np.array_equal(y, P.matvec(x))


def test_hv_c64_multi_matches_single(cuda):
    import torch
    sizes, nu, nd = [10], 5, 4
    op = make_op(sizes, nu, nd, 1.0, 4.0)
    P = po.ProductOperator(sizes, nu, nd, 1.0, 4.0)
    rng = np.random.default_rng(4)
    for nvec in (2, 4, 8, 16):
        X = np.stack([int_vec(rng, P.D) for _ in range(nvec)])  # (nvec, D)
        Y = mb.from_planar(op.hv(mb.to_planar(X, dtype=torch.float32)))
        for r in range(nvec):
            assert np.array_equal(Y[r], P.matvec(np.ascontiguousarray(X[r]))), (nvec, r)


GAUSS_CASES = CASES[:3] + FP_CASES


@pytest.mark.parametrize("name,sizes,nu,nd,t,U", GAUSS_CASES, ids=[c[0] for c in GAUSS_CASES])
def test_hv_c128_gaussian_tolerance(cuda, name, sizes, nu, nd, t, U):
    import torch
    op = make_op(sizes, nu, nd, t, U)
    P = po.ProductOperator(sizes, nu, nd, t, U)
    x_t = op.gaussian(seed=123)
    x = mb.from_planar(x_t)
    y = mb.from_planar(op.hv(x_t))
    ref = P.matvec(x)
    assert np.linalg.norm(y - ref) / np.linalg.norm(ref) < 1e-14


def test_gaussian_fill_is_seeded_per_reference_index(cuda):
    import torch
    op = make_op([10], 5, 5, 1.0, 4.0)
    a = op.gaussian(7)
    b = op.gaussian(7)
    c = op.gaussian(8)
    assert torch.equal(a, b) and not torch.equal(a, c)
    v = mb.from_planar(a)
    assert abs(v.real.mean()) < 0.01 and abs(v.imag.mean()) < 0.01
    assert abs(v.real.var() - 1) < 0.01 and abs(v.imag.var() - 1) < 0.01


def test_hv_c2_full_size_bitexact(cuda):
    """4 x 4 periodic, U = 8, 7 + 7: D = 130,873,600, integer inputs, bit for bit."""
    import torch
    sizes, nu, nd, t, U = [4, 4], 7, 7, 1.0, 8.0
    op = make_op(sizes, nu, nd, t, U)
    assert op.dimension() == 130873600
    P = po.ProductOperator(sizes, nu, nd, t, U, threads=16)
    x = int_vec(np.random.default_rng(5), P.D)
    xd = mb.to_planar(x)
    y = mb.from_planar(op.hv(xd))
    del xd
    ref = P.matvec(x)
    assert np.array_equal(y, ref)


def test_hv_c2_hermitian(cuda):
    """<x|H y> = <H x|y> on seeded Gaussian vectors at full size (FP64)."""
    import torch
    op = make_op([4, 4], 7, 7, 1.0, 8.0)
    def c(t):
        return torch.complex(t[0], t[1])
    x = op.gaussian(1)
    y = op.gaussian(2)
    lhs = torch.vdot(c(x), c(op.hv(y))).item()
    rhs = torch.vdot(c(op.hv(x)), c(y)).item()
    assert abs(lhs - rhs) / abs(lhs) < 1e-12


@pytest.mark.parametrize("name,sizes,nu,nd,t,U", CASES[:5], ids=[c[0] for c in CASES[:5]])
def test_generic_path_matches_two_pass(cuda, name, sizes, nu, nd, t, U, monkeypatch):
    """The one-kernel fallback path (rows too long for LDS) gives the same H.v."""
    op = make_op(sizes, nu, nd, t, U)
    P = po.ProductOperator(sizes, nu, nd, t, U)
    x = int_vec(np.random.default_rng(9), P.D)
    monkeypatch.setenv("MANYBODY_AMD_HV_PATH", "generic")
    y = mb.from_planar(op.hv(mb.to_planar(x)))
    monkeypatch.delenv("MANYBODY_AMD_HV_PATH")
    assert np.array_equal(y, P.matvec(x))
