"""Write the golden fixtures in tests/golden/ from the compiled reference.

Runs oracle/_ref/ref_driver (the reference's own algebra core, compiled from the
reference sources by `make -C oracle ref`) on small sectors and stores what it
prints: the basis states in reference index order, and the nonzero matrix
elements of HubbardModel / HubbardModel2D / HubbardModel3D.  Only sectors the
reference represents exactly are used: N_up + N_dn <= 7 for the basis, <= 3 for
the full Hamiltonian, <= 5 with U = 0 (7-operator key, fermion/term.h:11-19,
dropped long products, expression_base.h:184-186).

Run where the reference tree exists (it is not part of this repository):
    python tests/golden/make_golden.py
Outputs: tests/golden/ref_basis.npz, tests/golden/ref_hamiltonian.npz,
tests/golden/ref_correlation.npz (diagonal of the opposite-spin correlation
operator G_updn(r) of hubbard_model.h, every displacement r, N_up + N_dn <= 3),
tests/golden/ref_momentum.npz (fixed-momentum basis state lists for every K of
ten lattices, and the momentum-space Hamiltonian HubbardModelMomentum on eleven
sectors inside its exact range), tests/golden/ref_conductivity.npz (the
current operator J_d(Q) on seven lattices and compute_optical_conductivity on
four seeded correlators), tests/golden/ref_expression.npz (compute_matrix_elements
of generic FermionExpressions read by the driver: random Hermitian one- and
two-body terms with spin flips, odd and non-conserving monomials, constants,
cancelling and sub-tolerance terms, on real-space and fixed-momentum bases; and
cereal binary archives of sparse matrices from the vendored cereal),
tests/golden/manifest.json.
"""
from __future__ import annotations

import hashlib
import json
import os
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
DRIVER = os.path.join(REPO, "oracle", "_ref", "ref_driver")


def run(args):
    out = subprocess.run([DRIVER] + [str(a) for a in args], check=True, capture_output=True,
                         text=True).stdout
    states, elems = [], []
    for line in out.splitlines():
        p = line.split()
        if p[0] == "S":
            states.append((int(p[1]), int(p[2])))
        elif p[0] == "E":
            elems.append((int(p[1]), int(p[2]), float(p[3]), float(p[4])))
    return np.array(states, np.uint64).reshape(-1, 2), elems


def digest(states: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(states, np.uint64).tobytes()).hexdigest()


def _random_expression(rng, L, n_one, n_two, hermitian=True, odd=0):
    """Terms [(coeff, [(type, spin, orbital), ...])]: one-body c+_{a s} c_{b s'} and
    two-body c+ c+ c c with random modes (spin flips included) and complex
    coefficients, each followed by its adjoint when hermitian; `odd` extra
    monomials of 1 to 3 random operators."""
    def op(t):
        return (t, int(rng.integers(2)), int(rng.integers(L)))

    def adj(ops):
        return [(1 - t, s, o) for t, s, o in reversed(ops)]

    terms = []
    for n_ops, count in ((2, n_one), (4, n_two)):
        for _ in range(count):
            ops = [op(0) for _ in range(n_ops // 2)] + [op(1) for _ in range(n_ops // 2)]
            c = complex(rng.standard_normal(), rng.standard_normal())
            terms.append((c, ops))
            if hermitian:
                terms.append((c.conjugate(), adj(ops)))
    for _ in range(odd):
        ops = [op(int(rng.integers(2))) for _ in range(int(rng.integers(1, 4)))]
        terms.append((complex(rng.standard_normal(), rng.standard_normal()), ops))
    return terms


def _momentum_expression(rng, dims, n_one, n_two):
    """Momentum-conserving terms on the lattice `dims` (orbital = flat momentum):
    c+_{k s} c_{k s'} (spin flips included) and c+_{k1+q s} c+_{k2-q s'} c_{k2 s'} c_{k1 s}
    with complex coefficients and their adjoints, plus two non-conserving terms."""
    sys.path.insert(0, REPO)
    from oracle import pyoracle as po
    N = int(np.prod(dims))
    terms = []

    def add(c, ops):
        terms.append((c, ops))
        terms.append((c.conjugate(), [(1 - t, s, o) for t, s, o in reversed(ops)]))

    for _ in range(n_one):
        k = int(rng.integers(N))
        add(complex(rng.standard_normal(), rng.standard_normal()),
            [(0, int(rng.integers(2)), k), (1, int(rng.integers(2)), k)])
    for _ in range(n_two):
        k1, k2, q = (int(v) for v in rng.integers(N, size=3))
        s1, s2 = int(rng.integers(2)), int(rng.integers(2))
        add(complex(rng.standard_normal(), rng.standard_normal()),
            [(0, s1, po.momentum_add(dims, k1, q)), (0, s2, po.momentum_add(dims, k2, q, -1)),
             (1, s2, k2), (1, s1, k1)])
    terms.append((0.4, [(0, 0, 1 % N), (1, 0, 0)]))
    terms.append((0.2j, [(0, 1, 0), (1, 1, 0), (0, 0, 1 % N), (1, 1, 1 % N)]))
    return terms


def _hubbard_terms(L, t, U):
    """HubbardModel(t, U, L).hamiltonian() as raw terms, in the reference's loop
    order (hubbard_model.h:13-35), for the fixed-N example basis."""
    terms = []
    for i in range(L):
        j = (i + 1) % L
        for s in (0, 1):
            terms.append((-t, [(0, s, i), (1, s, j)]))
            terms.append((-t, [(0, s, j), (1, s, i)]))
    for i in range(L):
        terms.append((U, [(0, 0, i), (1, 0, i), (0, 1, i), (1, 1, i)]))
    return terms


def _special_terms(L):
    """Monomials whose action needs care: 1 - n as c c+, c+ c c+ c, a repeated
    creator (zero), a constant, a coefficient below 1000 eps, a key added twice
    (merged) and a pair that cancels (erased)."""
    return [(0.75, [(1, 0, 1), (0, 0, 1)]),
            (-1.25 + 0.5j, [(0, 0, 0), (1, 0, 0), (0, 1, L - 1), (1, 1, 0)]),
            (2.0, [(0, 1, 2), (0, 1, 2), (1, 1, 1), (1, 1, 0)]),
            (0.3 - 0.1j, []),
            (1e-14, [(0, 0, 1), (1, 0, 0)]),
            (0.5, [(0, 1, 1), (1, 1, 0)]), (0.25j, [(0, 1, 1), (1, 1, 0)]),
            (1.5, [(0, 0, 2), (1, 1, 2)]), (-1.5, [(0, 0, 2), (1, 1, 2)]),
            (0.6, [(0, 1, 0), (1, 0, 1), (0, 0, 1), (1, 1, 2)])]


def _terms_text(terms):
    return "".join(f"{complex(c).real:.17g} {complex(c).imag:.17g} {len(ops)} "
                   + " ".join(f"{t} {s} {o}" for t, s, o in ops) + "\n" for c, ops in terms)


def expression_cases():
    """compute_matrix_elements of generic expressions through the driver, and
    cereal archives of three sparse matrices."""
    rng = np.random.default_rng(77)
    arrays, manifest = {}, []
    cases = [("expr", [4], 2, 1, None, _random_expression(rng, 4, 6, 4)),
             ("expr", [5], 1, 1, None, _random_expression(rng, 5, 5, 3)),
             ("expr", [3], 2, 2, None, _random_expression(rng, 3, 4, 0, odd=4)),
             ("expr", [4], 2, 1, None, _special_terms(4)),
             ("expr", [6], 1, 2, None, _random_expression(rng, 6, 6, 3, hermitian=False, odd=3)),
             ("mexpr", [4], 2, 1, [1], _momentum_expression(rng, [4], 4, 6)),
             ("mexpr", [2, 2], 1, 2, [1, 0], _momentum_expression(rng, [2, 2], 4, 6)),
             ("mexpr", [5], 1, 2, [2], _momentum_expression(rng, [5], 5, 8)),
             ("mexpr", [7], 2, 1, [3], _momentum_expression(rng, [7], 6, 16)),
             ("mexpr", [2, 2], 2, 1, [0, 1], _special_terms(4)),
             ("nexpr", [4], 2, None, None, _hubbard_terms(4, 1.0, 4.0)),
             ("nexpr", [4], 3, None, None, _random_expression(rng, 4, 6, 4, odd=2)),
             ("nexpr", [5], 2, None, None, _random_expression(rng, 5, 8, 5)),
             ("nexpr", [3], 3, None, None, _special_terms(3))]
    for k, (mode, dims, nu, nd, K, terms) in enumerate(cases):
        args = [mode] + ([dims[0], nu, nd] if mode == "expr" else [dims[0], nu] if mode == "nexpr"
                         else [len(dims)] + dims + [nu, nd] + K)
        out = subprocess.run([DRIVER] + [str(a) for a in args], input=_terms_text(terms),
                             check=True, capture_output=True, text=True).stdout
        st = np.array([(int(l.split()[1]), int(l.split()[2])) for l in out.splitlines()
                       if l[0] == "S"], np.uint64).reshape(-1, 2)
        el = np.array([[float(x) for x in l.split()[1:]] for l in out.splitlines()
                       if l[0] == "E"], np.float64).reshape(-1, 4)
        name = f"{mode}{k}_{'x'.join(map(str, dims))}_{nu}_{nd}"  # nexpr: nu = N
        arrays[name + "__states"] = st
        arrays[name + "__ij"] = el[:, :2].astype(np.int64)
        arrays[name + "__val"] = el[:, 2] + 1j * el[:, 3]
        arrays[name + "__coef"] = np.array([complex(c) for c, _ in terms], np.complex128)
        arrays[name + "__nops"] = np.array([len(o) for _, o in terms], np.int64)
        arrays[name + "__ops"] = np.array([v for _, o in terms for op in o for v in op],
                                          np.int64).reshape(-1, 3)
        manifest.append({"name": name, "kind": mode, "dims": dims, "n_up": nu, "n_dn": nd,
                         "K": K, "dim": int(len(st)), "nnz": int(len(el))})
    for L in range(1, 7):
        for N in range(0, min(2 * L, 7) + 1):
            st, _ = run(["nbasis", L, N])
            name = f"nbasis_{L}_{N}"
            arrays[name + "__states"] = st
            manifest.append({"name": name, "kind": "nbasis", "L": L, "N": N, "dim": int(len(st))})
    for k, (nr, nc, nnz) in enumerate([(0, 0, 0), (4, 5, 3), (37, 37, 200)]):
        r = rng.integers(0, max(nr, 1), nnz)
        c = rng.integers(0, max(nc, 1), nnz)
        key = sorted(set(zip(c.tolist(), r.tolist())))   # column-major, unique
        v = rng.standard_normal(len(key)) + 1j * rng.standard_normal(len(key))
        txt = f"{nr} {nc}\n" + "".join(f"{rr} {cc} {x.real:.17g} {x.imag:.17g}\n"
                                       for (cc, rr), x in zip(key, v))
        path = os.path.join(HERE, "_cereal.bin")
        subprocess.run([DRIVER, "cereal", path], input=txt, check=True, text=True)
        with open(path, "rb") as f:
            blob = f.read()
        os.remove(path)
        name = f"cereal{k}_{nr}x{nc}"
        arrays[name + "__rows"] = np.array([rr for _, rr in key], np.uint64)
        arrays[name + "__cols"] = np.array([cc for cc, _ in key], np.uint64)
        arrays[name + "__vals"] = v
        arrays[name + "__bytes"] = np.frombuffer(blob, np.uint8)
        manifest.append({"name": name, "kind": "cereal", "n_rows": nr, "n_cols": nc,
                         "nnz": len(key)})
    rect_arrays, rect_manifest = rectangular_cases()
    arrays.update(rect_arrays)
    return arrays, manifest + rect_manifest


def _signature_terms(rng, L, signatures, count):
    """`count` terms per signature; a signature lists (type, spin) per operator,
    spin None for a random one; orbitals random, complex coefficients."""
    terms = []
    for sig in signatures:
        for _ in range(count):
            ops = [(t, int(rng.integers(2)) if sp is None else sp, int(rng.integers(L)))
                   for t, sp in sig]
            terms.append((complex(rng.standard_normal(), rng.standard_normal()), ops))
    return terms


def _momentum_current_terms(rng, dims, Q):
    """sum_{k, s} v_k c+_{k+Q s} c_{k s} with random complex v_k: a J(Q)-shaped
    operator from sector K to K + Q (orbital = flat momentum)."""
    sys.path.insert(0, REPO)
    from oracle import pyoracle as po
    N = int(np.prod(dims))
    return [(complex(rng.standard_normal(), rng.standard_normal()),
             [(0, s, po.momentum_add(dims, k, Q)), (1, s, k)]) for k in range(N) for s in (0, 1)]


def rectangular_cases():
    """compute_rectangular_matrix_elements between two bases (row <- column) and
    compute_vector_elements through the driver: particle-number changing
    operators (creators, annihilators, pair creation), the square case, a J(Q)
    between momentum sectors and bases of different kinds."""
    rng = np.random.default_rng(78)
    C, A = 0, 1
    cases = [
        ("rexpr", ["s", 4, 2, 1], ["s", 4, 1, 1],
         _signature_terms(rng, 4, [[(C, 0)], [(C, 0), (C, 1), (A, 1)], [(C, 1)], [(C, 0), (A, 0)]], 3)),
        ("rexpr", ["n", 4, 3], ["n", 4, 2],
         _signature_terms(rng, 4, [[(C, None)], [(C, None), (C, None), (A, None)]], 4)),
        ("rexpr", ["s", 4, 2, 2], ["s", 4, 1, 1],
         [(complex(rng.standard_normal(), 0.3), [(C, 0, i), (C, 1, i)]) for i in range(4)]
         + _signature_terms(rng, 4, [[(C, 1), (C, 0)]], 4)),
        ("rexpr", ["s", 3, 1, 0], ["s", 3, 1, 1],
         _signature_terms(rng, 3, [[(A, 1)], [(C, 0), (A, 0), (A, 1)]], 3)),
        ("rexpr", ["s", 4, 1, 1], ["s", 4, 1, 1], _random_expression(rng, 4, 5, 3)),
        ("rexpr", ["m", 1, 4, 1, 1, 1], ["m", 1, 4, 1, 1, 0], _momentum_current_terms(rng, [4], 1)),
        ("rexpr", ["m", 2, 2, 2, 1, 1, 1, 0], ["s", 4, 1, 1], _random_expression(rng, 4, 6, 2)),
        ("rexpr", ["s", 2, 2, 1], ["n", 2, 2],
         _signature_terms(rng, 2, [[(C, None)], [(C, None), (C, None), (A, None)]], 3)),
        ("vexpr", ["s", 3, 1, 1], None,
         _signature_terms(rng, 3, [[(C, 0), (C, 1)], [(C, 1), (C, 0)], [(C, 0)], []], 3)
         + [(0.5, [(C, 0, 0), (C, 1, 1), (A, 1, 1), (C, 1, 1)])]),
        ("vexpr", ["n", 3, 2], None, _signature_terms(rng, 3, [[(C, None), (C, None)]], 8)),
        ("vexpr", ["m", 1, 4, 1, 1, 1], None,
         [(complex(rng.standard_normal(), rng.standard_normal()), [(C, s0, a), (C, 1 - s0, b)])
          for s0 in (0, 1) for a in range(4) for b in range(4)]),
    ]
    arrays, manifest = {}, []
    for k, (mode, row, col, terms) in enumerate(cases):
        args = [mode] + row + (col or [])
        out = subprocess.run([DRIVER] + [str(a) for a in args], input=_terms_text(terms),
                             check=True, capture_output=True, text=True).stdout
        rows = np.array([(int(l.split()[1]), int(l.split()[2])) for l in out.splitlines()
                         if l[0] == "S"], np.uint64).reshape(-1, 2)
        cols = np.array([(int(l.split()[1]), int(l.split()[2])) for l in out.splitlines()
                         if l[0] == "C"], np.uint64).reshape(-1, 2)
        el = np.array([[float(x) for x in l.split()[1:]] for l in out.splitlines()
                       if l[0] == "E"], np.float64).reshape(-1, 4)
        name = f"{mode}{k}"
        arrays[name + "__rows"] = rows
        arrays[name + "__cols"] = cols
        arrays[name + "__ij"] = el[:, :2].astype(np.int64)
        arrays[name + "__val"] = el[:, 2] + 1j * el[:, 3]
        arrays[name + "__coef"] = np.array([complex(c) for c, _ in terms], np.complex128)
        arrays[name + "__nops"] = np.array([len(o) for _, o in terms], np.int64)
        arrays[name + "__ops"] = np.array([v for _, o in terms for op in o for v in op],
                                          np.int64).reshape(-1, 3)
        manifest.append({"name": name, "kind": mode, "row": row, "col": col,
                         "n_rows": int(len(rows)), "n_cols": int(len(cols)) if col else 1,
                         "nnz": int(len(el))})
    return arrays, manifest


def main():
    if not os.path.exists(DRIVER):
        sys.exit("build the reference checker first: make -C oracle ref")
    basis_arrays, basis_manifest = {}, []
    # every sector with L <= 7 and N <= 7: full lists for L <= 5, digests beyond
    for L in range(1, 8):
        for nu in range(0, L + 1):
            for nd in range(0, L + 1):
                if nu + nd > 7 or nu + nd == 0:
                    continue
                st, _ = run(["basis", L, nu, nd])
                name = f"b_{L}_{nu}_{nd}"
                entry = {"name": name, "L": L, "n_up": nu, "n_dn": nd, "dim": int(len(st)),
                         "sha256": digest(st)}
                if L <= 5:
                    basis_arrays[name] = st
                    entry["stored"] = True
                basis_manifest.append(entry)

    ham_cases = [
        # (kind, dims, n_up, n_dn, t, U)
        ("h1d", [2], 1, 1, 1.0, 4.0),
        ("h1d", [3], 1, 1, 1.0, 4.0),
        ("h1d", [4], 1, 1, 1.0, 4.0),
        ("h1d", [4], 2, 1, 1.0, 4.0),
        ("h1d", [5], 2, 1, 0.7, 2.5),
        ("h1d", [6], 1, 2, 1.0, 8.0),
        ("h1d", [8], 2, 1, 1.0, 4.0),
        ("h1d", [8], 3, 0, 1.0, 4.0),
        ("h1d", [8], 0, 3, -0.3, 1.0),
        ("h1d", [7], 3, 2, 1.0, 0.0),
        ("h1d", [8], 3, 2, 1.25, 0.0),
        ("h1d", [1], 1, 1, 1.0, 3.0),
        ("h2d", [2, 2], 1, 1, 1.0, 4.0),
        ("h2d", [2, 2], 2, 1, 1.0, 4.0),
        ("h2d", [3, 3], 1, 1, 1.0, 8.0),
        ("h2d", [3, 3], 2, 1, 1.0, 8.0),
        ("h2d", [4, 4], 2, 1, 1.0, 8.0),
        ("h2d", [4, 4], 1, 1, 0.5, 3.0),
        ("h2d", [2, 3], 2, 1, 1.0, 2.0),
        ("h2d", [3, 2], 1, 2, 1.0, 2.0),
        ("h2d", [3, 1], 1, 1, 1.0, 2.0),
        ("h2d", [1, 3], 2, 1, 1.0, 2.0),
        ("h2d", [4, 3], 1, 2, 1.0, 6.0),
        ("h2d", [3, 3], 3, 2, 1.0, 0.0),
        ("h2d", [4, 2], 2, 2, 1.0, 0.0),
        ("h3d", [2, 2, 2], 1, 1, 1.0, 4.0),
        ("h3d", [2, 2, 2], 2, 1, 1.0, 4.0),
        ("h3d", [3, 2, 2], 1, 1, 1.0, 5.0),
    ]
    ham_arrays, ham_manifest = {}, []
    for kind, dims, nu, nd, t, U in ham_cases:
        st, el = run([kind] + dims + [nu, nd, t, U])
        name = f"{kind}_{'x'.join(map(str, dims))}_{nu}_{nd}_t{t}_U{U}"
        el = np.array(el, np.float64).reshape(-1, 4)
        ham_arrays[name + "__states"] = st
        ham_arrays[name + "__ij"] = el[:, :2].astype(np.int64)
        ham_arrays[name + "__val"] = el[:, 2] + 1j * el[:, 3]
        ham_manifest.append({"name": name, "kind": kind, "dims": dims, "n_up": nu, "n_dn": nd,
                             "t": t, "U": U, "dim": int(len(st)), "nnz": int(len(el))})

    # G_updn(r): every displacement of each lattice; the operator is diagonal in
    # the occupation basis, so its diagonal (reference index order) is stored
    corr_cases = [("g1d", [4], 2, 1), ("g1d", [5], 1, 2), ("g1d", [6], 2, 1), ("g1d", [7], 1, 1),
                  ("g2d", [2, 2], 1, 1), ("g2d", [3, 2], 2, 1), ("g2d", [2, 3], 1, 2),
                  ("g3d", [2, 2, 2], 1, 1)]
    corr_arrays, corr_manifest = {}, []
    for kind, dims, nu, nd in corr_cases:
        N = int(np.prod(dims))
        diags = []
        for r in range(N):
            rr = [r % dims[0]] + ([(r // dims[0]) % dims[1]] if len(dims) > 1 else []) + \
                 ([r // (dims[0] * dims[1])] if len(dims) > 2 else [])
            st, el = run([kind] + dims + [nu, nd] + rr)
            D = len(st)
            g = np.zeros(D)
            for i, j, re, im in el:
                if i != j or im != 0.0:
                    raise SystemExit(f"{kind} {dims} r={rr}: off-diagonal or complex element")
                g[int(i)] = re
            diags.append(g)
        name = f"{kind}_{'x'.join(map(str, dims))}_{nu}_{nd}"
        corr_arrays[name + "__states"] = st
        corr_arrays[name + "__diag"] = np.array(diags)  # [r, reference index]
        corr_manifest.append({"name": name, "kind": kind, "dims": dims, "n_up": nu, "n_dn": nd,
                              "dim": int(len(st)), "displacements": N})

    # fixed total momentum K: basis state lists (every K of each lattice) and the
    # momentum-space Hamiltonian (HubbardModelMomentum) inside its exact range
    mom_arrays, mom_manifest = {}, []
    mbasis_cases = [([4], 2, 1), ([5], 2, 2), ([6], 3, 2), ([7], 2, 3), ([2, 2], 2, 1),
                    ([3, 2], 2, 2), ([2, 3], 3, 1), ([3, 3], 2, 2), ([4, 2], 3, 2),
                    ([2, 2, 2], 2, 2)]
    for dims, nu, nd in mbasis_cases:
        N = int(np.prod(dims))
        for K in range(N):
            kc = [K % dims[0]] + ([(K // dims[0]) % dims[1]] if len(dims) > 1 else []) + \
                 ([K // (dims[0] * dims[1])] if len(dims) > 2 else [])
            st, _ = run(["mbasis", len(dims)] + dims + [nu, nd] + kc)
            name = f"mb_{'x'.join(map(str, dims))}_{nu}_{nd}_K{K}"
            mom_arrays[name + "__states"] = st
            mom_manifest.append({"name": name, "kind": "basis", "dims": dims, "n_up": nu,
                                 "n_dn": nd, "K": K, "dim": int(len(st))})
    mh_cases = [([2], 1, 1, 0, 0.0, 3.0), ([4], 1, 1, 1, 1.0, 4.0), ([4], 2, 1, 2, 1.0, 4.0),
                ([5], 1, 2, 3, 0.7, 2.5), ([6], 2, 1, 0, 1.0, 8.0), ([4], 2, 2, 1, 1.5, 0.0),
                ([6], 3, 2, 4, 1.0, 0.0), ([2, 2], 2, 1, 1, 1.0, 4.0), ([3, 2], 1, 2, 5, 1.0, 6.0),
                ([3, 3], 2, 1, 4, 1.0, 8.0), ([2, 2, 2], 1, 2, 3, 1.0, 4.0)]
    for dims, nu, nd, K, t, U in mh_cases:
        kc = [K % dims[0]] + ([(K // dims[0]) % dims[1]] if len(dims) > 1 else []) + \
             ([K // (dims[0] * dims[1])] if len(dims) > 2 else [])
        st, el = run(["mh", len(dims)] + dims + [nu, nd] + kc + [t, U])
        name = f"mh_{'x'.join(map(str, dims))}_{nu}_{nd}_K{K}_t{t}_U{U}"
        el = np.array(el, np.float64).reshape(-1, 4)
        mom_arrays[name + "__states"] = st
        mom_arrays[name + "__ij"] = el[:, :2].astype(np.int64)
        mom_arrays[name + "__val"] = el[:, 2] + 1j * el[:, 3]
        mom_manifest.append({"name": name, "kind": "hamiltonian", "dims": dims, "n_up": nu,
                             "n_dn": nd, "K": K, "t": t, "U": U, "dim": int(len(st)),
                             "nnz": int(len(el))})

    # current operator J_d(Q) of HubbardModelMomentum on the full basis (two
    # operators per term: exact up to 5 particles), and compute_optical_conductivity
    # on seeded correlators
    cond_arrays, cond_manifest = {}, []
    j_cases = [([4], 1, 1, [1], 0, 1.0), ([5], 2, 1, [2], 0, 0.7), ([6], 2, 2, [1], 0, 1.0),
               ([3, 2], 2, 1, [1, 1], 0, 1.0), ([3, 2], 1, 2, [0, 1], 1, 1.3),
               ([2, 2, 2], 2, 1, [1, 0, 1], 2, 1.0), ([4], 2, 2, [0], 0, 1.0)]
    for dims, nu, nd, Q, dr, t in j_cases:
        st, el = run(["mj", len(dims)] + dims + [nu, nd] + Q + [dr, t])
        name = f"mj_{'x'.join(map(str, dims))}_{nu}_{nd}_Q{'-'.join(map(str, Q))}_d{dr}_t{t}"
        el = np.array(el, np.float64).reshape(-1, 4)
        cond_arrays[name + "__states"] = st
        cond_arrays[name + "__ij"] = el[:, :2].astype(np.int64)
        cond_arrays[name + "__val"] = el[:, 2] + 1j * el[:, 3]
        cond_manifest.append({"name": name, "kind": "current", "dims": dims, "n_up": nu,
                              "n_dn": nd, "Q": Q, "direction": dr, "t": t, "dim": int(len(st))})
    rng = np.random.default_rng(20260101)
    for n, dt, beta, vol in [(2, 0.5, 1.0, 1.0), (40, 0.1, 2.0, 16.0), (200, 0.05, 10.0, 8.0),
                             (101, 0.2, 0.0, 12.0)]:
        c = rng.standard_normal(n) + 1j * rng.standard_normal(n)
        txt = "".join(f"{v.real:.17g} {v.imag:.17g}\n" for v in c)
        out = subprocess.run([DRIVER, "sigma", str(dt), str(beta), str(vol)], input=txt,
                             check=True, capture_output=True, text=True).stdout
        W = np.array([[float(x) for x in l.split()[1:]] for l in out.splitlines() if l[0] == "W"])
        R = [float(l.split()[1]) for l in out.splitlines() if l[0] == "R"][0]
        name = f"sigma_n{n}_dt{dt}_b{beta}_v{vol}"
        cond_arrays[name + "__corr"] = c
        cond_arrays[name + "__omega"] = W[:, 0]
        cond_arrays[name + "__sigma"] = W[:, 1] + 1j * W[:, 2]
        cond_manifest.append({"name": name, "kind": "sigma", "n": n, "dt": dt, "beta": beta,
                              "volume": vol, "sum_rule": R})
    np.savez_compressed(os.path.join(HERE, "ref_conductivity.npz"), **cond_arrays)

    expr_arrays, expr_manifest = expression_cases()
    np.savez_compressed(os.path.join(HERE, "ref_expression.npz"), **expr_arrays)

    np.savez_compressed(os.path.join(HERE, "ref_momentum.npz"), **mom_arrays)
    np.savez_compressed(os.path.join(HERE, "ref_basis.npz"), **basis_arrays)
    np.savez_compressed(os.path.join(HERE, "ref_correlation.npz"), **corr_arrays)
    np.savez_compressed(os.path.join(HERE, "ref_hamiltonian.npz"), **ham_arrays)
    with open(os.path.join(HERE, "manifest.json"), "w") as f:
        json.dump({"generator": "tests/golden/make_golden.py",
                   "source": "reference algebra core compiled by oracle/Makefile (target ref)",
                   "basis": basis_manifest, "hamiltonian": ham_manifest,
                   "correlation": corr_manifest, "momentum": mom_manifest,
                   "conductivity": cond_manifest, "expression": expr_manifest}, f, indent=1)
    print(f"{len(basis_manifest)} basis sectors, {len(ham_manifest)} Hamiltonians, "
          f"{len(corr_manifest)} correlation operators")


if __name__ == "__main__":
    main()
