"""The row kernels of the two-pass H.v on sectors with D_dn = 11,440 (4 x 4, 7 down
electrons, 1 or 2 up): the persistent hop-count-sorted kernel (the default at
this row length) against the per-row compact kernel (selected with
MANYBODY_AMD_ROW_KERNEL=compact) and the oracle, through every epilogue: plain
H.v, Lanczos (pass one), reconstruction (the eigenvector pass) and Chebyshev
(KPM moments, FP64 and FP32).

Tolerances: integer-valued inputs are exact in FP64, so H.v is compared bit
for bit with the oracle; Gaussian inputs 1e-14 relative between the kernels
(they add a string's hops in a different order); Lanczos alpha/beta 1e-12
relative over 30 steps; eigenvalue 1e-12 relative, eigenvector 1e-9 relative;
FP64 KPM moments 1e-12 absolute (mu_0 = 1 to 1e-12), FP32 KPM moments 1e-5
absolute between the kernels and 1e-4 against FP64 (mu_0 = 1 to 1e-6).
The last test compares the one-pass H.v (the row kernel forms its row's up
hops itself) with the two passes on 1D, 2D and 3D sectors.
"""
import contextlib
import os

import numpy as np
import pytest
import torch

import manybody_amd as mb
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

T_HOP, U = 1.0, 8.0


@pytest.fixture(autouse=True)
def two_passes(monkeypatch):
    """These tests compare the row kernels of the two-pass H.v, so the one-pass
    kernel (the default on sectors that fit the Infinity Cache) is switched off;
    test_one_pass_hv_equals_two_passes switches it on for its own comparison."""
    monkeypatch.setenv("MANYBODY_AMD_FUSED", "0")


@contextlib.contextmanager
def compact_rows():
    old = os.environ.get("MANYBODY_AMD_ROW_KERNEL")
    os.environ["MANYBODY_AMD_ROW_KERNEL"] = "compact"
    try:
        yield
    finally:
        if old is None:
            del os.environ["MANYBODY_AMD_ROW_KERNEL"]
        else:
            os.environ["MANYBODY_AMD_ROW_KERNEL"] = old


def make_op(nu, nd=7):
    basis = mb.FermionBasis.sector(16, nu, nd)
    return mb.HubbardModel2D(T_HOP, U, 4, 4).operator(basis)


def test_hv_integer_inputs_bit_exact_vs_oracle(cuda):
    nu, nd = 1, 7
    op = make_op(nu, nd)
    D = op.dimension()
    assert (op.d_up, op.d_dn) == (16, 11440)
    rng = np.random.default_rng(7)
    x_ref = (rng.integers(-4, 5, D) + 1j * rng.integers(-4, 5, D)).astype(np.complex128)
    x = op.to_device_order(torch.from_numpy(x_ref).to(cuda))
    y = op.to_reference_order(op.hv(x)).cpu().numpy()
    A = po.build_csc([4, 4], nu, nd, T_HOP, U)
    assert np.array_equal(y, A.matvec(x_ref))


@contextlib.contextmanager
def no_fold():
    old = os.environ.get("MANYBODY_AMD_LANCZOS_FOLD")
    os.environ["MANYBODY_AMD_LANCZOS_FOLD"] = "0"
    try:
        yield
    finally:
        if old is None:
            del os.environ["MANYBODY_AMD_LANCZOS_FOLD"]
        else:
            os.environ["MANYBODY_AMD_LANCZOS_FOLD"] = old


@pytest.mark.parametrize("nu", [1, 2])
def test_row_first_lanczos_matches_up_then_row(cuda, nu):
    """Lanczos pass one with the row pass first (the previous step's axpy done
    while the row is staged, no separate axpy pass; the default on these
    sectors) against the up-then-row step with its axpy pass
    (MANYBODY_AMD_LANCZOS_FOLD=0) and against the oracle: alpha/beta to 1e-12
    relative over 40 steps (the two sum alpha and beta in different orders),
    the lowest Ritz value of 120 steps to 1e-12, and the eigenvector through
    pass two."""
    op = make_op(nu)
    b = op.gaussian(seed=41 + nu)
    d_f = mb.lanczos_pass_one(op, b, 40)
    with no_fold():
        d_a = mb.lanczos_pass_one(op, b, 40)
    assert d_f.steps_taken == d_a.steps_taken == 40
    assert np.abs(np.array(d_f.alphas) - d_a.alphas).max() <= 1e-12 * np.abs(d_a.alphas).max()
    assert np.abs(np.array(d_f.betas) - d_a.betas).max() <= 1e-12 * np.abs(d_a.betas).max()
    P = po.ProductOperator([4, 4], nu, 7, T_HOP, U)
    al, be, _ = P.lanczos(mb.from_planar(b), 6)
    assert np.abs(np.array(d_f.alphas[:6]) - al).max() <= 1e-12 * np.abs(al).max()
    assert np.abs(np.array(d_f.betas[:5]) - be).max() <= 1e-12 * np.abs(be).max()
    ep_f = mb.find_min_eigenpair(op, 120, b)
    with no_fold():
        ep_a = mb.find_min_eigenpair(op, 120, b)
    assert abs(ep_f.value - ep_a.value) <= 1e-12 * abs(ep_a.value)
    assert (ep_f.vector - ep_a.vector).norm().item() <= 1e-9 * ep_a.vector.norm().item()
    e, r = op.energy_residual(ep_f.vector)
    assert abs(e - ep_f.value) <= 1e-10 * abs(ep_f.value)


@pytest.mark.parametrize("nu", [1, 2])
def test_hv_sorted_matches_compact(cuda, nu):
    op = make_op(nu)
    x = op.gaussian(seed=11 + nu)
    y_s = op.hv(x)
    with compact_rows():
        y_c = op.hv(x)
    scale = y_c.abs().max().item()
    assert (y_s - y_c).abs().max().item() <= 1e-14 * scale
    # FP32, two complex vectors (four planes)
    xf = torch.cat([x, op.gaussian(seed=99)]).float()
    yf_s = op.hv(xf)
    with compact_rows():
        yf_c = op.hv(xf)
    assert (yf_s - yf_c).abs().max().item() <= 1e-5 * yf_c.abs().max().item()


def test_lanczos_sorted_matches_compact(cuda):
    op = make_op(2)
    b = op.gaussian(seed=3)
    d_s = mb.lanczos_pass_one(op, b, 30)
    with compact_rows():
        d_c = mb.lanczos_pass_one(op, b, 30)
    assert d_s.steps_taken == d_c.steps_taken == 30
    a_s, a_c = np.array(d_s.alphas), np.array(d_c.alphas)
    b_s, b_c = np.array(d_s.betas), np.array(d_c.betas)
    assert np.abs(a_s - a_c).max() <= 1e-12 * np.abs(a_c).max()
    assert np.abs(b_s - b_c).max() <= 1e-12 * np.abs(b_c).max()


def test_eigenpair_sorted_matches_compact(cuda):
    """find_min_eigenpair: pass one (Lanczos epilogue) and the eigenvector pass
    (reconstruction epilogue)."""
    op = make_op(1)
    b = op.gaussian(seed=5)
    ep_s = mb.find_min_eigenpair(op, 80, b)
    with compact_rows():
        ep_c = mb.find_min_eigenpair(op, 80, b)
    assert abs(ep_s.value - ep_c.value) <= 1e-12 * abs(ep_c.value)
    diff = (ep_s.vector - ep_c.vector).norm().item()
    assert diff <= 1e-9 * ep_c.vector.norm().item()
    e, r = op.energy_residual(ep_s.vector)
    assert abs(e - ep_s.value) <= 1e-10 * abs(ep_s.value)


@pytest.mark.parametrize("fp32", [False, True])
def test_kpm_sorted_matches_compact(cuda, fp32):
    """Chebyshev epilogue; the spectrum of 4 x 4 (2 up, 7 down) at U = 8 lies in
    [-36, 52], inside [b - a, b + a] = [-38, 54]."""
    op = make_op(2)
    a, bb = 46.0, 8.0
    mu_s = np.array(mb.kpm_dos_moments(op, a, bb, 64, 2, seed=21, fp32=fp32))
    with compact_rows():
        mu_c = np.array(mb.kpm_dos_moments(op, a, bb, 64, 2, seed=21, fp32=fp32))
    # mu_0 = <r|r> / <r|r>: exact in FP64, to FP32 rounding of the vectors in FP32
    assert abs(mu_s[0] - 1.0) < (1e-6 if fp32 else 1e-12)
    tol = 1e-5 if fp32 else 1e-12
    assert np.abs(mu_s - mu_c).max() <= tol
    if fp32:
        mu_64 = np.array(mb.kpm_dos_moments(op, a, bb, 64, 2, seed=21, fp32=False))
        assert np.abs(mu_s - mu_64).max() <= 1e-4


# ------------------------------------------------------------ multi-plane row kernel

@contextlib.contextmanager
def row_planes(n):
    old = os.environ.get("MANYBODY_AMD_ROW_PLANES")
    os.environ["MANYBODY_AMD_ROW_PLANES"] = str(n)
    try:
        yield
    finally:
        if old is None:
            del os.environ["MANYBODY_AMD_ROW_PLANES"]
        else:
            os.environ["MANYBODY_AMD_ROW_PLANES"] = old


@pytest.mark.parametrize("L,nu,nd", [(14, 7, 7), (12, 6, 6), (10, 5, 4)])
def test_multi_plane_row_kernel_equals_one_plane_kernel(cuda, L, nu, nd):
    """The row pass over several planes at once (two FP32 planes by default;
    2, 4 or 8 with MANYBODY_AMD_ROW_PLANES) against the one-plane compact
    kernel (MANYBODY_AMD_ROW_PLANES=1): the same
    sums in the same order (the compiler may contract multiply-adds
    differently), so H.v agrees to 1e-15 relative, Lanczos coefficients to
    1e-13, FP64 KPM moments to 1e-13 and FP32 moments to 1e-6; integer inputs
    bit for bit against the oracle."""
    op = mb.HubbardModel(T_HOP, 4.0, L).operator(mb.FermionBasis.sector(L, nu, nd))
    x = op.gaussian(17)
    w, _ = mb.tridiag_eig(*[getattr(mb.lanczos_pass_one(op, x, 60), f) for f in ("alphas", "betas")],
                          vectors=False)
    a, c = (w[-1] - w[0]) / 2 * 1.05, (w[-1] + w[0]) / 2  # spectrum inside [-1, 1]
    with row_planes(1):
        y1 = op.hv(x)
        dec1 = mb.lanczos_pass_one(op, x, 12)
        mu1 = np.array(mb.kpm_dos_moments(op, a, c, 32, 8, 3, fp32=True))
        mu1d = np.array(mb.kpm_dos_moments(op, a, c, 32, 4, 3, fp32=False))
    for cap in (2, 4, 8):
        with row_planes(cap):
            y = op.hv(x)
            dec = mb.lanczos_pass_one(op, x, 12)
        assert torch.abs(y - y1).max().item() <= 1e-15 * torch.abs(y1).max().item()
        assert np.abs(np.array(dec.alphas) - dec1.alphas).max() <= 1e-13 * np.abs(dec1.alphas).max()
        assert np.abs(np.array(dec.betas) - dec1.betas).max() <= 1e-13 * np.abs(dec1.betas).max()
        with row_planes(cap):
            mu = np.array(mb.kpm_dos_moments(op, a, c, 32, 8, 3, fp32=True))
            mud = np.array(mb.kpm_dos_moments(op, a, c, 32, 4, 3, fp32=False))
        assert np.abs(mu - mu1).max() <= 1e-6, cap
        assert np.abs(mud - mu1d).max() <= 1e-13, cap
    # integer inputs: bit for bit against the oracle (FP64: the default kernel)
    rng = np.random.default_rng(11)
    D = op.dimension()
    z = (rng.integers(-4, 5, D) + 1j * rng.integers(-4, 5, D)).astype(np.complex128)
    P = po.ProductOperator([L], nu, nd, T_HOP, 4.0)
    yz = op.hv(torch.from_numpy(np.stack([z.real, z.imag])).to(cuda)).cpu().numpy()
    assert np.array_equal(yz[0] + 1j * yz[1], P.matvec(z))
    # and the multi-vector FP32 H.v on integer inputs (the default two planes
    # per workgroup, then eight)
    zz = torch.from_numpy(np.stack([z.real, z.imag] * 4).astype(np.float32)).to(cuda)
    ref = P.matvec(z).astype(np.complex64)
    for cap in (None, 8):
        with (row_planes(cap) if cap else contextlib.nullcontext()):
            yy = op.hv(zz).cpu().numpy()
        for r in range(4):
            assert np.array_equal(yy[2 * r] + 1j * yy[2 * r + 1], ref), (cap, r)


@contextlib.contextmanager
def one_pass(mode):
    """MANYBODY_AMD_FUSED: "0" the two passes (strip, then row), "1" the one-pass
    kernel wherever the tables allow it."""
    old = os.environ.get("MANYBODY_AMD_FUSED")
    os.environ["MANYBODY_AMD_FUSED"] = mode
    try:
        yield
    finally:
        if old is None:
            del os.environ["MANYBODY_AMD_FUSED"]
        else:
            os.environ["MANYBODY_AMD_FUSED"] = old


@pytest.mark.parametrize("sizes,nu,nd", [([14], 7, 7), ([12], 6, 6), ([10], 5, 4), ([8], 4, 4),
                                         ([4, 3], 6, 6), ([3, 3], 4, 4), ([2, 2, 3], 5, 5)])
def test_one_pass_hv_equals_two_passes(cuda, sizes, nu, nd):
    """The one-pass H.v (each row kernel forms its row's up hops from the source
    rows, MANYBODY_AMD_FUSED=1) against the two passes (=0).  Both sum the up
    hops in the same order and add the down hops and the diagonal the same way;
    the compiler may contract multiply-adds differently (and in FP64 the two
    passes use the one-plane row kernel), so: FP64 H.v to 1e-15 relative, FP32
    H.v over four and eight planes to 1e-6, 12 Lanczos coefficients to 1e-13,
    FP64 KPM moments to 1e-13 and FP32 moments to 1e-6.  Integer inputs: bit
    for bit against the oracle."""
    N = int(np.prod(sizes))
    model = (mb.HubbardModel(T_HOP, 4.0, sizes[0]) if len(sizes) == 1 else
             mb.HubbardModel2D(T_HOP, 4.0, *sizes) if len(sizes) == 2 else
             mb.HubbardModel3D(T_HOP, 4.0, *sizes))
    op = model.operator(mb.FermionBasis.sector(N, nu, nd))
    x = op.gaussian(23)
    xf = torch.cat([x, op.gaussian(29)]).float()
    x8 = torch.cat([xf, xf.flip(0)])
    w, _ = mb.tridiag_eig(*[getattr(mb.lanczos_pass_one(op, x, 40), f) for f in ("alphas", "betas")],
                          vectors=False)
    a, c = (w[-1] - w[0]) / 2 * 1.05, (w[-1] + w[0]) / 2
    out = {}
    for mode in ("0", "1"):
        with one_pass(mode):
            dec = mb.lanczos_pass_one(op, x, 12)
            out[mode] = (op.hv(x), op.hv(xf), op.hv(x8), np.array(dec.alphas), np.array(dec.betas),
                         np.array(mb.kpm_dos_moments(op, a, c, 32, 8, 5, fp32=True)),
                         np.array(mb.kpm_dos_moments(op, a, c, 32, 4, 5, fp32=False)))
        print(f"{sizes} {nu}+{nd} D={op.dimension()} mode {mode}: |Hx| = "
              f"{out[mode][0].abs().max().item():.6f}", flush=True)
    tol = (1e-15, 1e-6, 1e-6, 1e-13, 1e-13)
    for r0, r1, tl in zip(out["0"][:5], out["1"][:5], tol):
        if isinstance(r0, torch.Tensor):
            assert (r1 - r0).abs().max().item() <= tl * r0.abs().max().item()
        else:
            assert np.abs(r1 - r0).max() <= tl * np.abs(r0).max()
    assert np.abs(out["1"][5] - out["0"][5]).max() <= 1e-6
    assert np.abs(out["1"][6] - out["0"][6]).max() <= 1e-13
    # integer inputs against the oracle, one pass
    rng = np.random.default_rng(31)
    D = op.dimension()
    z = (rng.integers(-4, 5, D) + 1j * rng.integers(-4, 5, D)).astype(np.complex128)
    P = po.ProductOperator(sizes, nu, nd, T_HOP, 4.0)
    with one_pass("1"):
        yz = op.hv(torch.from_numpy(np.stack([z.real, z.imag])).to(cuda)).cpu().numpy()
    assert np.array_equal(yz[0] + 1j * yz[1], P.matvec(z))


@pytest.mark.parametrize("sizes", [[24], [6, 4], [8, 3]])
def test_sorted_row_kernel_on_other_lattices(cuda, sizes):
    """Sectors of 24 sites with 4 down electrons (D_dn = 10,626) also take the
    sorted row kernel (one hop magnitude: no lattice side of length 2), with
    fewer hops per string (at most 8 on the ring, 16 on the tori) and so other
    sign patterns and bank-pair colourings of the sorted table: integer inputs
    bit for bit against the oracle, Gaussian inputs against the per-row compact
    kernel to 1e-14."""
    nu, nd = 1, 4
    L = int(np.prod(sizes))
    model = (mb.HubbardModel(T_HOP, U, L) if len(sizes) == 1 else
             mb.HubbardModel2D(T_HOP, U, *sizes))
    op = model.operator(mb.FermionBasis.sector(L, nu, nd))
    assert (op.d_up, op.d_dn) == (24, 10626)
    D = op.dimension()
    rng = np.random.default_rng(L + len(sizes))
    x_ref = (rng.integers(-4, 5, D) + 1j * rng.integers(-4, 5, D)).astype(np.complex128)
    x = op.to_device_order(torch.from_numpy(x_ref).to(cuda))
    y = op.to_reference_order(op.hv(x)).cpu().numpy()
    A = po.build_csc(sizes, nu, nd, T_HOP, U)
    assert np.array_equal(y, A.matvec(x_ref))
    g = op.gaussian(seed=5)
    y_s = op.hv(g)
    with compact_rows():
        y_c = op.hv(g)
    assert (y_s - y_c).abs().max().item() <= 1e-14 * y_c.abs().max().item()
