"""Generic FermionExpressions compiled for the GPU (compute_matrix_elements,
matrix_elements.h:74-105, as a device operator and a device-built sparse
export).

Checked against: the compiled reference's matrices for ten expressions (golden,
1e-13 relative on H.v, entries exactly located and 1e-14 in value); the
oracle's op-by-op expression_matrix on larger random expressions (1e-13
relative); the native Hubbard operators for the Hubbard expressions, real space
up to the full 4 x 4, 7 + 7 sector and momentum space (1e-13 relative); ground
energies through Lanczos (1e-10 relative); the cereal archive against the
oracle's byte layout of the exported entries.
"""
import numpy as np
import pytest
import torch

import manybody_amd as mb
from manybody_amd.expression import FermionExpression
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu


def _gold():
    import os
    return np.load(os.path.join(os.path.dirname(__file__), "golden", "ref_expression.npz"))


def _manifest():
    import json
    import os
    with open(os.path.join(os.path.dirname(__file__), "golden", "manifest.json")) as f:
        return json.load(f)["expression"]


def expr_from_terms(terms):
    e = FermionExpression()
    for c, ops in terms:
        key = tuple((0x8000 if t else 0) | (0x4000 if s else 0) | o for t, s, o in ops)
        e += FermionExpression({key: c})
    return e


def oracle_terms(e):
    return [(c, [(o >> 15, (o >> 14) & 1, o & 0x3FFF) for o in k]) for k, c in e.terms.items()]


def kflat(dims, K):
    Kc = list(K) + [0] * (3 - len(dims))
    dd = list(dims) + [1] * (3 - len(dims))
    return Kc[0] + dd[0] * (Kc[1] + dd[1] * Kc[2])


def basis_for(dims, nu, nd, K=None, fixed_n=False):
    L = int(np.prod(dims))
    if fixed_n:  # with_fixed_particle_number(L, N = nu)
        return mb.FermionBasis.with_fixed_particle_number(L, nu)
    if K is None:
        return mb.FermionBasis.sector(L, nu, nd)
    return mb.FermionBasis.with_fixed_particle_number_spin_momentum(L, nu + nd, nu - nd, dims, K)


def apply_ref(op, x_ref, cuda):
    return op.apply(torch.from_numpy(np.ascontiguousarray(x_ref)).to(cuda)).cpu().numpy()


def test_expression_operator_matches_reference_golden(cuda):
    gold = _gold()
    rng = np.random.default_rng(1)
    n = 0
    for c in _manifest():
        if c["kind"] not in ("expr", "mexpr", "nexpr"):
            continue
        name = c["name"]
        coef, nops, ops = gold[name + "__coef"], gold[name + "__nops"], gold[name + "__ops"]
        terms, k = [], 0
        for cc, m in zip(coef, nops):
            terms.append((complex(cc), [tuple(int(v) for v in ops[k + i]) for i in range(int(m))]))
            k += int(m)
        basis = basis_for(c["dims"], c["n_up"], c["n_dn"], c["K"], c["kind"] == "nexpr")
        op = mb.FermionExpressionOperator(basis, expr_from_terms(terms))
        D = op.dimension()
        assert D == c["dim"]
        R = np.zeros((D, D), np.complex128)
        ij = gold[name + "__ij"]
        R[ij[:, 0], ij[:, 1]] = gold[name + "__val"]
        x = rng.standard_normal(D) + 1j * rng.standard_normal(D)
        want = R @ x
        got = apply_ref(op, x, cuda)
        assert np.abs(got - want).max() <= 1e-13 * max(1.0, np.abs(want).max()), name
        sp = mb.compute_matrix_elements(op)
        r, cl, v = sp.entries()
        order = np.lexsort((ij[:, 0], ij[:, 1]))  # column-major, rows ascending
        assert sp.nnz == len(ij), name
        assert np.array_equal(r.astype(np.int64), ij[order, 0]), name
        assert np.array_equal(cl.astype(np.int64), ij[order, 1]), name
        assert np.abs(v - gold[name + "__val"][order]).max(initial=0) <= 1e-14 * max(
            1.0, np.abs(v).max(initial=0)), name
        n += 1
    assert n == 14


def random_expression(rng, L, n_one, n_two, n_odd=0, n_three=0):
    """Hermitian one-, two- and three-body terms (spin flips included) and extra
    non-Hermitian odd monomials."""
    e = FermionExpression()

    def op(t):
        return (0x8000 if t else 0) | (0x4000 if rng.integers(2) else 0) | int(rng.integers(L))

    for nb, cnt in ((1, n_one), (2, n_two), (3, n_three)):
        for _ in range(cnt):
            key = tuple(op(0) for _ in range(nb)) + tuple(op(1) for _ in range(nb))
            t = FermionExpression({key: complex(rng.standard_normal(), rng.standard_normal())})
            e += t + t.adjoint()
    for _ in range(n_odd):
        key = tuple(op(int(rng.integers(2))) for _ in range(int(rng.integers(1, 6))))
        e += FermionExpression({key: complex(rng.standard_normal(), rng.standard_normal())})
    return e


@pytest.mark.parametrize("dims,nu,nd,K", [([8], 3, 3, None), ([3, 3], 2, 3, None),
                                          ([2, 2, 2], 3, 2, None), ([9], 1, 4, None),
                                          ([6], 3, 2, [2]), ([3, 3], 2, 2, [1, 2]),
                                          ([2, 2, 2], 2, 2, [1, 0, 1])])
def test_expression_operator_matches_oracle_beyond_reference_range(cuda, dims, nu, nd, K):
    rng = np.random.default_rng(sum(dims) * 7 + nu)
    L = int(np.prod(dims))
    e = random_expression(rng, L, 12, 10, n_odd=6, n_three=3)
    if K is not None:  # add momentum-conserving terms so the sector has off-diagonal work
        hm = mb.HubbardModelMomentum(0.8, 3.0, dims).hamiltonian()
        e += hm
    basis = basis_for(dims, nu, nd, K)
    op = mb.FermionExpressionOperator(basis, e)
    if K is None:
        up, dn = po.basis_states(L, nu, nd)
    else:
        up, dn = po.momentum_sector_states(dims, nu, nd, kflat(dims, K))
    st = list(zip(up, dn))
    M = po.expression_matrix(L, st, oracle_terms(e))
    x = rng.standard_normal(len(st)) + 1j * rng.standard_normal(len(st))
    want = M @ x
    got = apply_ref(op, x, cuda)
    assert np.abs(got - want).max() <= 1e-13 * np.abs(want).max()
    # the exported matrix: same nonzero pattern, values to rounding
    sp = mb.compute_matrix_elements(op)
    r, c, v = sp.entries()
    rr, cc, vv = po.sparse_column_major(M)
    assert np.array_equal(r, rr) and np.array_equal(c, cc)
    assert np.abs(v - vv).max() <= 1e-13 * np.abs(vv).max()


@pytest.mark.parametrize("dims,nu,nd,U", [([10], 5, 4, 4.0), ([3, 4], 3, 3, 8.0),
                                          ([2, 3, 2], 3, 2, 5.0), ([2, 2], 2, 1, 3.0)])
def test_hubbard_expression_equals_native_operator(cuda, dims, nu, nd, U):
    L = int(np.prod(dims))
    basis = mb.FermionBasis.sector(L, nu, nd)
    model = (mb.HubbardModel(1.0, U, dims[0]) if len(dims) == 1 else
             mb.HubbardModel2D(1.0, U, *dims) if len(dims) == 2 else
             mb.HubbardModel3D(1.0, U, *dims))
    native = model.operator(basis)
    op = mb.FermionExpressionOperator(basis, model.hamiltonian())
    x = native.gaussian(seed=5)
    y0, y1 = native.hv(x), op.hv(x)
    assert torch.abs(y1 - y0).max().item() <= 1e-13 * torch.abs(y0).max().item()
    e0 = mb.find_min_eigenpair(native, 120, x, want_vector=False).value
    e1 = mb.find_min_eigenpair(op, 120, x, want_vector=False).value
    assert abs(e1 - e0) <= 1e-10 * abs(e0)


@pytest.mark.parametrize("dims,nu,nd", [([3, 4], 3, 3), ([10], 5, 4)])
def test_expression_strip_and_one_pass_paths_agree(cuda, monkeypatch, dims, nu, nd):
    """Real-space sectors run the up-only terms in a strip pass; the generic-path
    switch selects the one-pass kernel instead.  Both give the same H.v and the
    same Lanczos coefficients (fused epilogue on both)."""
    rng = np.random.default_rng(4)
    L = int(np.prod(dims))
    e = random_expression(rng, L, 10, 8, n_odd=3) + mb.HubbardModel(1.0, 4.0, L).hamiltonian()
    op = mb.FermionExpressionOperator(mb.FermionBasis.sector(L, nu, nd), e)
    x = op.gaussian(seed=6)
    y_strip = op.hv(x)
    d_strip = mb.lanczos_pass_one(op, x, 40)
    monkeypatch.setenv("MANYBODY_AMD_HV_PATH", "generic")
    y_one = op.hv(x)
    d_one = mb.lanczos_pass_one(op, x, 40)
    monkeypatch.delenv("MANYBODY_AMD_HV_PATH")
    assert torch.abs(y_strip - y_one).max().item() <= 1e-13 * torch.abs(y_one).max().item()
    assert np.allclose(d_strip.alphas, d_one.alphas, rtol=1e-12, atol=1e-12)
    assert np.allclose(d_strip.betas, d_one.betas, rtol=1e-12, atol=1e-12)


def test_hubbard_expression_on_4x4_7p7_equals_native_operator(cuda):
    """The full 4 x 4, 7 + 7 sector (D = 130,873,600): the compiled expression of
    HubbardModel2D.hamiltonian() gives the native two-pass H.v."""
    basis = mb.FermionBasis.sector(16, 7, 7)
    model = mb.HubbardModel2D(1.0, 8.0, 4, 4)
    op = mb.FermionExpressionOperator(basis, model.hamiltonian())
    x = op.gaussian(seed=20260101)
    y1 = op.hv(x)
    del op
    native = model.operator(basis)
    y0 = native.hv(x)
    err = torch.abs(y1 - y0).max().item()
    assert err <= 1e-13 * torch.abs(y0).max().item()


@pytest.mark.parametrize("dims,nu,nd,K", [([4, 4], 3, 3, [1, 2]), ([6], 3, 3, [0]),
                                          ([2, 2, 2], 3, 2, [1, 1, 0])])
def test_momentum_hubbard_expression_equals_factorized_operator(cuda, dims, nu, nd, K):
    L = int(np.prod(dims))
    basis = basis_for(dims, nu, nd, K)
    fact = mb.HubbardMomentumFactorizedFixedK(basis, dims, 1.0, 6.0, K)
    op = mb.FermionExpressionOperator(basis, mb.HubbardModelMomentum(1.0, 6.0, dims).hamiltonian())
    x = fact.gaussian(seed=3)
    y0, y1 = fact.hv(x), op.hv(x)
    assert torch.abs(y1 - y0).max().item() <= 1e-13 * torch.abs(y0).max().item()
    assert L == int(np.prod(dims))


def test_expression_export_cereal_roundtrip(cuda, tmp_path):
    rng = np.random.default_rng(9)
    basis = mb.FermionBasis.sector(7, 3, 2)
    e = random_expression(rng, 7, 8, 6, n_odd=2)
    sp = mb.compute_matrix_elements(basis, e)
    path = tmp_path / "h.bin"
    sp.save(str(path))
    r, c, v = sp.entries()
    assert path.read_bytes() == po.cereal_sparse_bytes(sp.shape[0], sp.shape[1], r, c, v)
    back = mb.SparseMatrix.load(str(path))
    r2, c2, v2 = back.entries()
    assert np.array_equal(r, r2) and np.array_equal(c, c2) and np.array_equal(v, v2)
    assert sp.to_scipy().shape == (basis.size(), basis.size())


def test_expression_operator_rejects_invalid_use(cuda):
    basis = mb.FermionBasis.sector(4, 2, 1)
    with pytest.raises(ValueError):  # orbital beyond the basis
        mb.FermionExpressionOperator(basis, mb.density(0, 5))
    # n_0up n_1dn is not of Hubbard form: the generic expression kernels, FP64 only
    op = mb.FermionExpressionOperator(basis, mb.hopping(1.0, 0, 1, 0) + mb.density(1, 2) +
                                      mb.density_density(0, 0, 1, 1))
    x = op.gaussian(seed=1)
    with pytest.raises(ValueError):
        op.hv(x.float())
    with pytest.raises(ValueError):  # FP32 KPM needs the Hubbard tables
        mb.kpm_dos_moments(op, 10.0, 0.0, 8, 1, seed=1, fp32=True)
    mu = mb.kpm_dos_moments(op, 10.0, 0.0, 8, 1, seed=1)  # the default runs FP64 here
    assert abs(mu[0] - 1.0) < 1e-14
    with pytest.raises(ValueError):
        op.opposite_spin_correlation(x)
    empty = mb.FermionExpressionOperator(basis, FermionExpression())
    assert torch.count_nonzero(empty.hv(x)).item() == 0
    assert mb.compute_matrix_elements(empty).nnz == 0


def test_hubbard_on_fixed_n_basis_reference_example(cuda):
    """The reference's exact-diagonalisation example (examples/hubbard_exact_
    diagonalization.cpp): HubbardModel(1, 4, 4) on with_fixed_particle_number(4, 2)
    and on the spin-0 sector; the full-basis spectrum is the union of the three
    spin sectors' spectra (oracle CSC), and the model's operator handle (built by
    mb_hubbard_create) equals the compiled expression."""
    model = mb.HubbardModel(1.0, 4.0, 4)
    basis = mb.FermionBasis.with_fixed_particle_number(4, 2)
    op = model.operator(basis)
    assert op.dimension() == 28
    H = mb.compute_matrix_elements(basis, model.hamiltonian()).to_dense()
    w = np.linalg.eigvalsh(H)
    sectors = np.concatenate([np.linalg.eigvalsh(po.build_csc([4], nu, 2 - nu, 1.0, 4.0)
                                                 .to_scipy().toarray()) for nu in range(3)])
    assert np.abs(np.sort(sectors) - w).max() < 1e-12
    x = op.gaussian(seed=3)
    y0 = op.hv(x)
    y1 = mb.FermionExpressionOperator(basis, model.hamiltonian()).hv(x)
    assert torch.abs(y1 - y0).max().item() <= 1e-14 * torch.abs(y0).max().item()
    assert torch.abs(op.to_device_order(op.apply(op.to_reference_order(x))) - y0).max().item() < 1e-12


@pytest.mark.parametrize("L,N", [(6, 5), (5, 4), (7, 3)])
def test_spin_flip_expressions_on_fixed_n_match_oracle(cuda, L, N):
    """Expressions with spin flips (terms moving particles between spins) on a
    fixed-N basis against the oracle, beyond the reference's range."""
    rng = np.random.default_rng(L * 10 + N)
    e = random_expression(rng, L, 12, 8, n_odd=4) + mb.HubbardModel(1.0, 3.0, L).hamiltonian()
    for i in range(L - 1):  # transverse field and spin exchange
        e += 0.4 * (mb.one_body(mb.Spin.Up, i, mb.Spin.Down, i)
                    + mb.one_body(mb.Spin.Down, i, mb.Spin.Up, i))
    basis = mb.FermionBasis.with_fixed_particle_number(L, N)
    op = mb.FermionExpressionOperator(basis, e)
    up, dn = po.fixed_n_states(L, N)
    M = po.expression_matrix(L, list(zip(up, dn)), oracle_terms(e))
    x = rng.standard_normal(len(up)) + 1j * rng.standard_normal(len(up))
    want = M @ x
    got = apply_ref(op, x, cuda)
    assert np.abs(got - want).max() <= 1e-13 * np.abs(want).max()
    r, c, v = mb.compute_matrix_elements(op).entries()
    rr, cc, vv = po.sparse_column_major(M)
    assert np.array_equal(r, rr) and np.array_equal(c, cc)
    assert np.abs(v - vv).max() <= 1e-13 * np.abs(vv).max()


def test_fixed_n_ground_energy_is_the_sector_minimum(cuda):
    """Lanczos on with_fixed_particle_number(8, 8) (D = 12,870) for the ring
    Hubbard model gives the lowest ground energy over the nine spin sectors."""
    model = mb.HubbardModel(1.0, 4.0, 8)
    op = model.operator(mb.FermionBasis.with_fixed_particle_number(8, 8))
    e_full = mb.find_min_eigenpair(op, 200, op.gaussian(seed=1), want_vector=False).value
    e_sec = []
    for nu in range(0, 9):
        s = model.operator(mb.FermionBasis.sector(8, nu, 8 - nu))
        e_sec.append(mb.find_min_eigenpair(s, min(200, s.dimension()), s.gaussian(seed=2),
                                           want_vector=False).value)
    assert abs(e_full - min(e_sec)) <= 1e-9 * abs(min(e_sec))


def test_matrix_elements_density_on_single_orbital(cuda):
    """Port of test_matrix_elements.cpp:54-68: n_up0 on with_fixed_particle_number
    (1, 1) is diag(1, 0)."""
    basis = mb.FermionBasis.with_fixed_particle_number(1, 1)
    M = mb.compute_matrix_elements(basis, mb.density(mb.Spin.Up, 0)).to_dense()
    assert np.array_equal(M, np.array([[1, 0], [0, 0]], np.complex128))


def _basis_from_spec(spec):
    """Basis of a driver spec: ["s", L, nup, ndn], ["n", L, N] or
    ["m", nd, d.., nup, ndn, K..]."""
    if spec[0] == "s":
        return mb.FermionBasis.sector(spec[1], spec[2], spec[3])
    if spec[0] == "n":
        return mb.FermionBasis.with_fixed_particle_number(spec[1], spec[2])
    nd = spec[1]
    dims = spec[2:2 + nd]
    nu, ndn = spec[2 + nd], spec[3 + nd]
    K = spec[4 + nd:4 + 2 * nd]
    return mb.FermionBasis.with_fixed_particle_number_spin_momentum(
        int(np.prod(dims)), nu + ndn, nu - ndn, dims, K)


def test_rectangular_and_vector_elements_match_reference_golden(cuda):
    """compute_rectangular_matrix_elements / compute_vector_elements through the
    device against the compiled reference: creators, annihilators and pair
    creation between sectors of different particle numbers, the square case, a
    J(Q) between momentum sectors, bases of different kinds."""
    gold = _gold()
    n = 0
    for c in _manifest():
        if c["kind"] not in ("rexpr", "vexpr"):
            continue
        name = c["name"]
        coef, nops, ops = gold[name + "__coef"], gold[name + "__nops"], gold[name + "__ops"]
        terms, k = [], 0
        for cc, m in zip(coef, nops):
            terms.append((complex(cc), [tuple(int(v) for v in ops[k + i]) for i in range(int(m))]))
            k += int(m)
        e = expr_from_terms(terms)
        rb = _basis_from_spec(c["row"])
        assert [rb.state(i) for i in range(rb.size())] == \
            [tuple(int(v) for v in r) for r in gold[name + "__rows"]], name
        ij, val = gold[name + "__ij"], gold[name + "__val"]
        if c["kind"] == "vexpr":
            v = mb.compute_vector_elements(rb, e)
            want = np.zeros(rb.size(), np.complex128)
            want[ij[:, 0]] = val
            assert np.abs(v - want).max() <= 1e-14 * max(1.0, np.abs(want).max()), name
        else:
            cb = _basis_from_spec(c["col"])
            sp = mb.compute_rectangular_matrix_elements(rb, cb, e)
            assert sp.shape == (rb.size(), cb.size())
            r, cl, v = sp.entries()
            order = np.lexsort((ij[:, 0], ij[:, 1]))
            assert sp.nnz == len(ij), name
            assert np.array_equal(r.astype(np.int64), ij[order, 0]), name
            assert np.array_equal(cl.astype(np.int64), ij[order, 1]), name
            assert np.abs(v - val[order]).max() <= 1e-14 * max(1.0, np.abs(v).max()), name
        n += 1
    assert n == 11


@pytest.mark.parametrize("row,col", [(("n", 6, 4), ("n", 6, 3)), (("s", 6, 3, 2), ("s", 6, 2, 2)),
                                     (("s", 5, 2, 2), ("n", 5, 3)), (("n", 7, 2), ("s", 7, 1, 1))])
def test_rectangular_matrix_elements_match_oracle_beyond_reference_range(cuda, row, col):
    """Random creators, c+ c+ c, c+ c and spin flips between two bases against
    the oracle, and the adjoint relation <i|A|j> = conj(<j|A^dagger|i>) on the
    device."""
    L = row[1]
    rng = np.random.default_rng(sum(row[1:]) * 13 + sum(col[1:]))
    e = random_expression(rng, L, 6, 4, n_odd=10)
    for _ in range(6):
        e += FermionExpression({(int(rng.integers(2)) * 0x4000 | int(rng.integers(L)),):
                                complex(rng.standard_normal(), rng.standard_normal())})
    rb, cb = _basis_from_spec(list(row)), _basis_from_spec(list(col))
    rs = [rb.state(i) for i in range(rb.size())]
    cs = [cb.state(i) for i in range(cb.size())]
    M = po.rectangular_matrix(L, rs, cs, oracle_terms(e))
    sp = mb.compute_rectangular_matrix_elements(rb, cb, e)
    r, c, v = sp.entries()
    rr, cc, vv = po.sparse_column_major(M)
    assert len(vv) > 0
    assert np.array_equal(r, rr) and np.array_equal(c, cc)
    assert np.abs(v - vv).max() <= 1e-13 * np.abs(vv).max()
    adj = mb.compute_rectangular_matrix_elements(cb, rb, e.adjoint()).to_dense()
    assert np.abs(adj.conj().T - sp.to_dense()).max() <= 1e-13 * np.abs(vv).max()


@pytest.mark.parametrize("sizes,nu,nd,K,Q,d", [([4], 2, 2, [1], [1], 0), ([3, 3], 2, 1, [1, 2], [2, 0], 1)])
def test_rectangular_current_between_momentum_sectors(cuda, sizes, nu, nd, K, Q, d):
    """J_d(Q) = sum_{k, s} v_d(k) c+_{k+Q s} c_{k s} from sector K to K + Q as
    compute_rectangular_matrix_elements does it in the reference's example
    (hubbard_momentum_optical_conductivity.cpp:243), against the oracle's current
    operator (pinned to the reference's symbolic current)."""
    N = int(np.prod(sizes))
    flat = lambda v: v[0] + (sizes[0] * v[1] if len(v) > 1 else 0)  # noqa: E731
    Kf, Qf = flat(K), flat(Q)
    terms = []
    for k in range(N):
        vk = po.velocity(sizes, 1.0, d, k)
        if abs(vk) < 1e-15:
            continue
        for s in (0, 1):
            terms.append((vk, [(0, s, po.momentum_add(sizes, k, Qf)), (1, s, k)]))
    e = expr_from_terms(terms)
    KQ = [(a + b) % n for a, b, n in zip(K, Q, sizes)]
    src = mb.FermionBasis.with_fixed_particle_number_spin_momentum(N, nu + nd, nu - nd, sizes, K)
    tgt = mb.FermionBasis.with_fixed_particle_number_spin_momentum(N, nu + nd, nu - nd, sizes, KQ)
    got = mb.compute_rectangular_matrix_elements(tgt, src, e).to_dense()
    want = po.current_operator(sizes, nu, nd, Kf, Qf, d, 1.0)
    assert np.count_nonzero(want) > 0
    assert np.abs(got - want).max() <= 1e-13 * np.abs(want).max()


def hubbard_form_expression(rng, L, bonds, U, mags, one_way=0):
    """Real hops on the bonds with amplitudes from `mags` magnitudes (random
    signs), spin-dependent on-site energies, one U on every site, and
    `one_way` one-directional hops (a non-Hermitian term; still of Hubbard
    form: the kernels gather along each entry's own direction)."""
    amps = rng.choice(np.asarray(mags, dtype=float), size=len(bonds)) * rng.choice([-1.0, 1.0], len(bonds))
    terms = []
    for (i, j), t in zip(bonds, amps):
        for s in (0, 1):
            terms.append((complex(t), [(0, s, i), (1, s, j)]))
            terms.append((complex(t), [(0, s, j), (1, s, i)]))
    for _ in range(one_way):
        i, j = (int(v) for v in rng.choice(L, 2, replace=False))
        s = int(rng.integers(2))
        terms.append((complex(mags[0]), [(0, s, i), (1, s, j)]))
    for i in range(L):
        for s in (0, 1):
            terms.append((complex(rng.standard_normal()), [(0, s, i), (1, s, i)]))
        terms.append((complex(U), [(0, 0, i), (1, 0, i), (0, 1, i), (1, 1, i)]))
    return expr_from_terms(terms)


def ring_bonds(L):
    return [(i, (i + 1) % L) for i in range(L)]


def grid_bonds(dims):
    L = int(np.prod(dims))
    out = []
    for s in range(L):
        c = np.unravel_index(s, dims[::-1])[::-1]
        for ax in range(len(dims)):
            n = list(c)
            n[ax] = (n[ax] + 1) % dims[ax]
            t = int(np.ravel_multi_index(n[::-1], dims[::-1]))
            if t != s:
                out.append((s, t))
    return out


@pytest.mark.parametrize("fused", ["1", "0"])
@pytest.mark.parametrize("dims,nu,nd,mags,one_way", [([8], 3, 3, [1.0], 0), ([3, 3], 2, 3, [0.7, 1.3], 0),
                                                     ([2, 2, 2], 3, 2, [1.0], 3), ([9], 1, 4, [0.5, 1.0, 2.0], 2)])
def test_hubbard_form_expression_runs_on_hubbard_kernels(cuda, monkeypatch, fused, dims, nu, nd, mags,
                                                         one_way):
    """An expression of Hubbard form (real hops per species, real on-site energies,
    one U on every site) runs on the Hubbard tables and kernels built from its own
    hop lists: H.v equals the oracle's matrix, the generic expression kernels
    (MANYBODY_AMD_EXPR_NATIVE=0) and, in FP32, the FP64 result to FP32 rounding;
    Lanczos coefficients equal the generic path's.  Several magnitudes take the
    packed (ELL) row kernel and the strip up pass; one-directional hops check the
    gather direction."""
    monkeypatch.setenv("MANYBODY_AMD_FUSED", fused)
    rng = np.random.default_rng(sum(dims) * 11 + nu + len(mags))
    L = int(np.prod(dims))
    e = hubbard_form_expression(rng, L, grid_bonds(dims) if len(dims) > 1 else ring_bonds(L), 4.0,
                                mags, one_way)
    basis = basis_for(dims, nu, nd)
    op = mb.FermionExpressionOperator(basis, e)
    up, dn = po.basis_states(L, nu, nd)
    M = po.expression_matrix(L, list(zip(up, dn)), oracle_terms(e))
    x = rng.standard_normal(M.shape[0]) + 1j * rng.standard_normal(M.shape[0])
    want = M @ x
    assert np.abs(apply_ref(op, x, cuda) - want).max() <= 1e-13 * np.abs(want).max()
    xd = op.to_device_order(torch.from_numpy(x).to(cuda))
    y = op.hv(xd)
    y32 = op.hv(xd.float()).double()  # FP32 exists only on the Hubbard kernels
    assert torch.abs(y32 - y).max().item() <= 2e-6 * torch.abs(y).max().item()
    d = mb.lanczos_pass_one(op, xd, 30)
    monkeypatch.setenv("MANYBODY_AMD_EXPR_NATIVE", "0")
    gen = mb.FermionExpressionOperator(basis, e)
    monkeypatch.delenv("MANYBODY_AMD_EXPR_NATIVE")
    yg = gen.hv(xd)
    assert torch.abs(y - yg).max().item() <= 1e-13 * torch.abs(yg).max().item()
    dg = mb.lanczos_pass_one(gen, xd, 30)
    assert d.steps_taken == dg.steps_taken
    assert np.allclose(d.alphas, dg.alphas, rtol=1e-11, atol=1e-12)
    assert np.allclose(d.betas, dg.betas, rtol=1e-11, atol=1e-12)
    with pytest.raises(ValueError, match="Hubbard form"):
        gen.hv(xd.float())
    if one_way == 0:  # Hermitian: KPM moments, FP32 on the Hubbard kernels against FP64
        a = float(np.abs(M).sum(axis=1).max()) * 1.01
        mu32 = mb.kpm_dos_moments(op, a, 0.0, 32, 2, seed=3, fp32=True)
        mu64 = mb.kpm_dos_moments(gen, a, 0.0, 32, 2, seed=3)
        assert abs(mu64[0] - 1.0) < 1e-13
        assert np.abs(np.asarray(mu32) - mu64).max() <= 1e-4


@pytest.mark.parametrize("extra", ["nonuniform_u", "neighbour_v", "complex_hop", "pair_hop"])
def test_expressions_outside_hubbard_form_keep_the_expression_kernels(cuda, extra):
    """Terms that leave Hubbard form (a site with another U, n_up n_dn on two
    sites, a complex hop, a pair hop) keep the generic expression kernels: FP32
    H.v is refused, and FP64 H.v still equals the oracle."""
    rng = np.random.default_rng(3)
    L, nu, nd = 8, 3, 3
    e = hubbard_form_expression(rng, L, ring_bonds(L), 4.0, [1.0])
    extra_terms = {
        "nonuniform_u": [(0.5 + 0j, [(0, 0, 2), (1, 0, 2), (0, 1, 2), (1, 1, 2)])],
        "neighbour_v": [(0.5 + 0j, [(0, 0, 2), (1, 0, 2), (0, 1, 3), (1, 1, 3)])],
        "complex_hop": [(0.3j, [(0, 0, 1), (1, 0, 4)]), (-0.3j, [(0, 0, 4), (1, 0, 1)])],
        "pair_hop": [(0.4 + 0j, [(0, 0, 1), (0, 1, 1), (1, 1, 2), (1, 0, 2)]),
                     (0.4 + 0j, [(0, 0, 2), (0, 1, 2), (1, 1, 1), (1, 0, 1)])],
    }[extra]
    e += expr_from_terms(extra_terms)
    basis = basis_for([L], nu, nd)
    op = mb.FermionExpressionOperator(basis, e)
    xd = op.gaussian(seed=9)
    with pytest.raises(ValueError, match="Hubbard form"):
        op.hv(xd.float())
    up, dn = po.basis_states(L, nu, nd)
    M = po.expression_matrix(L, list(zip(up, dn)), oracle_terms(e))
    x = rng.standard_normal(M.shape[0]) + 1j * rng.standard_normal(M.shape[0])
    want = M @ x
    assert np.abs(apply_ref(op, x, cuda) - want).max() <= 1e-13 * np.abs(want).max()


def test_hubbard_expression_on_4x4_7p7_lanczos_and_fp32(cuda):
    """The 4 x 4, 7 + 7 Hubbard expression runs the native kernels (block up pass,
    sorted row pass, row-first Lanczos step): 40 Lanczos coefficients equal the
    native operator's to 1e-12, and FP32 H.v equals the native FP32 H.v to FP32
    rounding (the same tables but for the order of each string's hops)."""
    basis = mb.FermionBasis.sector(16, 7, 7)
    model = mb.HubbardModel2D(1.0, 8.0, 4, 4)
    op = mb.FermionExpressionOperator(basis, model.hamiltonian())
    x = op.gaussian(seed=77)
    d1 = mb.lanczos_pass_one(op, x, 40)
    y1 = op.hv(x.float())
    del op
    native = model.operator(basis)
    d0 = mb.lanczos_pass_one(native, x, 40)
    y0 = native.hv(x.float())
    assert d1.steps_taken == d0.steps_taken == 40
    assert np.abs(np.array(d1.alphas) - d0.alphas).max() <= 1e-12 * np.abs(d0.alphas).max()
    assert np.abs(np.array(d1.betas) - d0.betas).max() <= 1e-12 * np.abs(d0.betas).max()
    assert torch.abs(y1 - y0).max().item() <= 1e-5 * torch.abs(y0).max().item()
