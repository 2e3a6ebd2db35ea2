"""mb_lanczos_pass_one_noalloc (no device allocation, static per-GPU scratch)
against mb_lanczos_pass_one, through the C ABI.

Both run the same kernels with the same fixed-order reductions, so the bar is
bit-for-bit equality of alpha, beta, the step count and |b|.  Two calls in a
row share the scratch array: the first on the largest L = 12 sector
(6 + 6, D = 853,776, 41 MB of work vectors), the second on a smaller one
(L = 10, 4 + 5, D = 52,920), then the first again.  A run larger than the
array is refused with MB_ERR_INVALID before any device work, and the entry
point still works afterwards; operators other than real-space Hubbard models
are refused.
"""
import ctypes as C

import numpy as np
import pytest

import manybody_amd as mb

pytestmark = pytest.mark.gpu


def ring(L, nu, nd, U):
    return mb.HubbardModel(1.0, U, L).operator(mb.FermionBasis.sector(L, nu, nd))


def call(name, op, b, k):
    al = (C.c_double * max(k, 1))()
    be = (C.c_double * max(k, 1))()
    steps = C.c_int()
    bn = C.c_double()
    st = getattr(mb.lib(), name)(op._h, mb._ptr(b), k, al, be, C.byref(steps), C.byref(bn),
                                 mb._stream())
    m = steps.value
    return st, np.array(al[:m]), np.array(be[:max(m - 1, 0)]), m, bn.value


def same(op, b, k, label):
    print(f"no-alloc Lanczos: {label}, D = {op.dimension()}, k = {k}", flush=True)
    s0, a0, b0, m0, n0 = call("mb_lanczos_pass_one", op, b, k)
    s1, a1, b1, m1, n1 = call("mb_lanczos_pass_one_noalloc", op, b, k)
    assert s0 == 0 and s1 == 0, (s0, s1, mb.lib().mb_last_error().decode())
    assert m1 == m0 == k
    assert n1 == n0
    assert np.array_equal(a1, a0)
    assert np.array_equal(b1, b0)
    return a1


def test_noalloc_lanczos_equals_allocating_entry_point(cuda):
    big = ring(12, 6, 6, 4.0)
    small = ring(10, 4, 5, 7.5)
    bb, bs = big.gaussian(3), small.gaussian(4)
    a_first = same(big, bb, 60, "first call, L=12 6+6")
    same(small, bs, 40, "second call, smaller input L=10 4+5")
    a_again = same(big, bb, 60, "third call, L=12 6+6 again")
    assert np.array_equal(a_first, a_again)
    # and through the Python wrapper
    dec = mb.lanczos_pass_one_noalloc(small, bs, 40)
    ref = mb.lanczos_pass_one(small, bs, 40)
    assert dec.alphas == ref.alphas and dec.betas == ref.betas


def test_noalloc_lanczos_refuses_oversize_and_other_operators(cuda):
    op14 = ring(14, 7, 7, 4.0)  # D = 11,778,624: 565 MB of work vectors
    print("no-alloc Lanczos: oversize L=14 7+7 (refused on the host)", flush=True)
    st, *_ = call("mb_lanczos_pass_one_noalloc", op14, op14.gaussian(1), 10)
    assert st == mb.MB_ERR_INVALID
    assert "scratch" in mb.lib().mb_last_error().decode()
    op = ring(8, 4, 4, 4.0)
    neg = mb.Negated(op)
    print("no-alloc Lanczos: composite operator (refused on the host)", flush=True)
    st, *_ = call("mb_lanczos_pass_one_noalloc", neg, op.gaussian(2), 10)
    assert st == mb.MB_ERR_INVALID
    assert "real-space" in mb.lib().mb_last_error().decode()
    same(op, op.gaussian(2), 30, "after the refusals, L=8 4+4")
