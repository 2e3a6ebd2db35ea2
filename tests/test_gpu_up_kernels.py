"""The up-hop pass of the two-pass H.v: the block-staged kernel (strings that
agree in their high bits share a block of rows staged in LDS; the default
where its tables exist) against the per-row strip kernel
(MANYBODY_AMD_UP_KERNEL=strip).

Both add a row's hops in the same order (the order of the row's entries, hops
that leave the block first) with the same arithmetic, so the results must be
equal bit for bit on Gaussian inputs too, in FP64 and FP32, through every
epilogue of the row pass.  Integer inputs are also checked bit for bit against
the oracle.  The one-pass kernel is switched off (MANYBODY_AMD_FUSED=0) so the
two-pass path runs on every sector.
"""
import contextlib
import os

import numpy as np
import pytest
import torch

import manybody_amd as mb
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def two_passes(monkeypatch):
    monkeypatch.setenv("MANYBODY_AMD_FUSED", "0")
    monkeypatch.delenv("MANYBODY_AMD_UP_KERNEL", raising=False)


@contextlib.contextmanager
def strip_kernel():
    old = os.environ.get("MANYBODY_AMD_UP_KERNEL")
    os.environ["MANYBODY_AMD_UP_KERNEL"] = "strip"
    try:
        yield
    finally:
        if old is None:
            del os.environ["MANYBODY_AMD_UP_KERNEL"]
        else:
            os.environ["MANYBODY_AMD_UP_KERNEL"] = old


def make_op(sizes, nu, nd, t, U):
    L = int(np.prod(sizes))
    basis = mb.FermionBasis.sector(L, nu, nd)
    if len(sizes) == 1:
        model = mb.HubbardModel(t, U, sizes[0])
    elif len(sizes) == 2:
        model = mb.HubbardModel2D(t, U, *sizes)
    else:
        model = mb.HubbardModel3D(t, U, *sizes)
    return model.operator(basis)


SECTORS = [
    ("1d_L8_4_4", [8], 4, 4, 1.0, 4.0),
    ("1d_L12_6_6", [12], 6, 6, 1.0, 4.0),
    ("1d_L14_7_7", [14], 7, 7, 1.0, 4.0),
    ("2d_4x3_3_4", [4, 3], 3, 4, 1.0, 6.0),
    ("2d_4x4_5_4", [4, 4], 5, 4, 1.0, 8.0),
    ("2d_4x4_7_2_t07", [4, 4], 7, 2, 0.7, 2.3),
    ("3d_2x2x3_4_3", [2, 2, 3], 4, 3, 1.0, 4.0),
]


@pytest.mark.parametrize("name,sizes,nu,nd,t,U", SECTORS, ids=[s[0] for s in SECTORS])
def test_block_kernel_equals_strip_kernel(cuda, name, sizes, nu, nd, t, U):
    op = make_op(sizes, nu, nd, t, U)
    x = op.gaussian(seed=31)
    y_b = op.hv(x)
    with strip_kernel():
        y_s = op.hv(x)
    assert torch.equal(y_b, y_s)
    # FP32, four complex vectors at once
    X = torch.cat([op.gaussian(seed=40 + k) for k in range(4)]).to(torch.float32)  # (8, D)
    Y_b = op.hv(X)
    with strip_kernel():
        Y_s = op.hv(X)
    assert torch.equal(Y_b, Y_s)


@pytest.mark.parametrize("name,sizes,nu,nd,t,U", SECTORS[:5], ids=[s[0] for s in SECTORS[:5]])
def test_block_kernel_integer_inputs_vs_oracle(cuda, name, sizes, nu, nd, t, U):
    op = make_op(sizes, nu, nd, t, U)
    P = po.ProductOperator(sizes, nu, nd, t, U)
    rng = np.random.default_rng(17)
    x = (rng.integers(-4, 5, P.D) + 1j * rng.integers(-4, 5, P.D)).astype(np.complex128)
    y = mb.from_planar(op.hv(mb.to_planar(x)))
    assert np.array_equal(y, P.matvec(x))


def test_block_kernel_epilogues_equal_strip_kernel(cuda):
    """Lanczos pass one, the eigenvector (reconstruction epilogue) and FP64 KPM
    moments on the 4 x 4, 5 + 4 sector: the same numbers with either up kernel."""
    op = make_op([4, 4], 5, 4, 1.0, 8.0)
    b = op.gaussian(seed=5)
    dec_b = mb.lanczos_pass_one(op, b, 30)
    ep_b = mb.find_min_eigenpair(op, 60, b)
    mu_b = mb.kpm_dos_moments(op, 40.0, 5.0, 32, 2, 3, fp32=False)
    with strip_kernel():
        dec_s = mb.lanczos_pass_one(op, b, 30)
        ep_s = mb.find_min_eigenpair(op, 60, b)
        mu_s = mb.kpm_dos_moments(op, 40.0, 5.0, 32, 2, 3, fp32=False)
    assert list(dec_b.alphas) == list(dec_s.alphas)
    assert list(dec_b.betas) == list(dec_s.betas)
    assert ep_b.value == ep_s.value
    assert torch.equal(ep_b.vector, ep_s.vector)
    assert np.array_equal(np.asarray(mu_b), np.asarray(mu_s))


def test_block_kernel_full_c2_sector_equals_strip_kernel(cuda):
    """4 x 4 periodic, U = 8, 7 + 7 (D = 130,873,600): Gaussian input, both up
    kernels, bit for bit."""
    op = make_op([4, 4], 7, 7, 1.0, 8.0)
    x = op.gaussian(seed=2)
    y_b = op.hv(x)
    with strip_kernel():
        y_s = op.hv(x)
    assert torch.equal(y_b, y_s)
