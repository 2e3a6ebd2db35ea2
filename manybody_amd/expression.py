"""Fermionic expressions (the reference's FermionExpression algebra,
src/algebra/fermion/expression.h, term.h, expression_base.h) as the host-side
input of the expression compiler, and the sparse matrices it exports.

An expression is a map from operator strings to complex coefficients.  An
operator is stored as the reference's OperatorStorage (operator.h:14-40): bit
15 set for an annihilator, bit 14 set for spin down, bits 0-13 the orbital.  A
string o_1 o_2 ... o_n is the product with o_n acting first.  Adding merges
equal strings and drops coefficients below 1000 eps (ExpressionBase::add_to,
expression_base.h:85-97); products concatenate strings.  Unlike the reference's
fixed 7-operator key (fermion/term.h:11-19), strings of any length are kept.
"""
from __future__ import annotations

import ctypes as C
from typing import Dict, Iterable, Tuple

import numpy as np

_TOL = 1000.0 * np.finfo(np.float64).eps

ANNIHILATION = 0x8000
DOWN = 0x4000


class Spin:
    Up = 0
    Down = 1


def _op(annihilate: bool, spin: int, orbital: int) -> int:
    if not 0 <= orbital <= 0x3FFF:
        raise ValueError("orbital out of range")
    return (ANNIHILATION if annihilate else 0) | (DOWN if spin else 0) | int(orbital)


def _adjoint_op(o: int) -> int:
    return o ^ ANNIHILATION


class FermionExpression:
    """Sum of complex coefficients times operator strings (expression_base.h)."""

    def __init__(self, terms: Dict[Tuple[int, ...], complex] = None):
        self.terms: Dict[Tuple[int, ...], complex] = {}
        for k, c in (terms or {}).items():
            self._add(tuple(k), complex(c))

    # ExpressionBase::add_to
    def _add(self, key: Tuple[int, ...], c: complex):
        if abs(c) ** 2 < _TOL * _TOL:
            return
        if key in self.terms:
            s = self.terms[key] + c
            if abs(s) ** 2 < _TOL * _TOL:
                del self.terms[key]
            else:
                self.terms[key] = s
        else:
            self.terms[key] = c

    @staticmethod
    def scalar(c: complex) -> "FermionExpression":
        return FermionExpression({(): c})

    def __len__(self):
        return len(self.terms)

    def size(self) -> int:
        return len(self.terms)

    def copy(self) -> "FermionExpression":
        e = FermionExpression()
        e.terms = dict(self.terms)
        return e

    def __iadd__(self, other):
        if not isinstance(other, FermionExpression):
            other = FermionExpression.scalar(other)
        for k, c in other.terms.items():
            self._add(k, c)
        return self

    def __add__(self, other):
        r = self.copy()
        r += other
        return r

    __radd__ = __add__

    def __neg__(self):
        return self * -1.0

    def __isub__(self, other):
        self += -other if isinstance(other, FermionExpression) else -complex(other)
        return self

    def __sub__(self, other):
        r = self.copy()
        r -= other
        return r

    def __mul__(self, other):
        if isinstance(other, FermionExpression):
            r = FermionExpression()
            for k1, c1 in self.terms.items():
                for k2, c2 in other.terms.items():
                    r._add(k1 + k2, c1 * c2)
            return r
        v = complex(other)
        r = FermionExpression()
        if abs(v) ** 2 < _TOL * _TOL:  # ExpressionBase::scale clears
            return r
        for k, c in self.terms.items():
            r.terms[k] = c * v
        return r

    def __rmul__(self, other):
        if isinstance(other, FermionExpression):
            return other * self
        return self * other

    def adjoint(self) -> "FermionExpression":
        """adjoint (expression_base.h:361-368): reversed strings of adjoint operators,
        conjugated coefficients."""
        r = FermionExpression()
        for k, c in self.terms.items():
            r._add(tuple(_adjoint_op(o) for o in reversed(k)), c.conjugate())
        return r

    def arrays(self):
        """(coeffs (n, 2) float64, offsets (n + 1) int64, ops uint16) for the C ABI."""
        keys = list(self.terms.keys())
        coeffs = np.zeros((len(keys), 2), np.float64)
        offsets = np.zeros(len(keys) + 1, np.int64)
        ops = []
        for t, k in enumerate(keys):
            c = self.terms[k]
            coeffs[t] = (c.real, c.imag)
            ops.extend(k)
            offsets[t + 1] = len(ops)
        return coeffs, offsets, np.array(ops if ops else [0], np.uint16)

    def __repr__(self):
        return f"FermionExpression({len(self.terms)} terms)"


def _mono(c: complex, *ops: int) -> FermionExpression:
    return FermionExpression({tuple(ops): c})


# fermion/term.h:21-53
def creation(spin: int, orbital: int) -> FermionExpression:
    return _mono(1.0, _op(False, spin, orbital))


def annihilation(spin: int, orbital: int) -> FermionExpression:
    return _mono(1.0, _op(True, spin, orbital))


def one_body(s1: int, o1: int, s2: int, o2: int) -> FermionExpression:
    return _mono(1.0, _op(False, s1, o1), _op(True, s2, o2))


def two_body(s1, o1, s2, o2, s3, o3, s4, o4) -> FermionExpression:
    return _mono(1.0, _op(False, s1, o1), _op(False, s2, o2), _op(True, s3, o3),
                 _op(True, s4, o4))


def density(spin: int, orbital: int) -> FermionExpression:
    return one_body(spin, orbital, spin, orbital)


def density_density(s1: int, i: int, s2: int, j: int) -> FermionExpression:
    return density(s1, i) * density(s2, j)


def hopping(coeff, frm: int, to: int = None, spin: int = None) -> FermionExpression:
    """hopping(coeff, from, to, spin) = coeff c+_from c_to + conj(coeff) c+_to c_from
    (expression_base.h:395-403); hopping(from, to, spin) with coeff 1."""
    if spin is None:  # hopping(from, to, spin)
        coeff, frm, to, spin = 1.0, coeff, frm, to
    c = complex(coeff)
    r = _mono(c, _op(False, spin, frm), _op(True, spin, to))
    r += _mono(c.conjugate(), _op(False, spin, to), _op(True, spin, frm))
    return r


# ------------------------------------------------------------ model expressions

def hubbard_interaction(u: float, n_sites: int) -> FermionExpression:
    e = FermionExpression()
    for i in range(n_sites):
        e += complex(u) * density_density(Spin.Up, i, Spin.Down, i)
    return e


def hubbard_hamiltonian(t: float, u: float, sizes) -> FermionExpression:
    """HubbardModel{,2D,3D}::hamiltonian() = kinetic() + interaction()."""
    sizes = list(sizes)
    e = hubbard_kinetic(t, sizes)
    e += hubbard_interaction(u, int(np.prod(sizes)))
    return e


def hubbard_kinetic(t: float, sizes) -> FermionExpression:
    """-t sum over bonds and spins of hopping(site, next) in the reference's loop
    order (hubbard_model.h:13-22, 69-84, 141-160): x outermost, site index
    x + Lx y + Lx Ly z, neighbours (x+1, y, z), (x, y+1, z), (x, y, z+1)."""
    sizes = [int(v) for v in sizes]
    e = FermionExpression()
    c = complex(-t)
    if len(sizes) == 1:
        L = sizes[0]
        for i in range(L):  # hubbard_model.h:13-22
            e += c * hopping(i, (i + 1) % L, Spin.Up)
            e += c * hopping(i, (i + 1) % L, Spin.Down)
        return e
    Lx, Ly = sizes[0], sizes[1]
    Lz = sizes[2] if len(sizes) == 3 else 1
    for x in range(Lx):
        for y in range(Ly):
            for z in range(Lz):
                s = x + Lx * (y + Ly * z)
                nbrs = [((x + 1) % Lx) + Lx * (y + Ly * z), x + Lx * (((y + 1) % Ly) + Ly * z)]
                if len(sizes) == 3:
                    nbrs.append(x + Lx * (y + Ly * ((z + 1) % Lz)))
                for nb in nbrs:
                    e += c * hopping(s, nb, Spin.Up)
                    e += c * hopping(s, nb, Spin.Down)
    return e


def hubbard_momentum_current(t: float, sizes, Q, direction: int) -> FermionExpression:
    """HubbardModelMomentum::current(Q, direction) (hubbard_model_momentum.h:52-91):
    sum_k v_d(k) (c+_{k+Q,up} c_{k,up} + c+_{k+Q,dn} c_{k,dn}) with the group velocity
    v_d(k) = 2t (2 pi / L_d) sin(2 pi k_d / L_d) (:38-48); terms with v exactly 0
    are skipped; k = kx + Lx ky + Lx Ly kz."""
    sizes = [int(v) for v in sizes]
    Q = [int(v) for v in Q]
    if len(Q) != len(sizes):
        raise ValueError("Momentum transfer Q must have the same dimension as the system.")
    if direction < 0 or direction >= len(sizes):
        raise ValueError("Direction index out of bounds.")
    d3 = sizes + [1] * (3 - len(sizes))
    q3 = Q + [0] * (3 - len(Q))
    N = int(np.prod(sizes))
    e = FermionExpression()
    for k in range(N):
        kc = [k % d3[0], (k // d3[0]) % d3[1], k // (d3[0] * d3[1])]
        L = sizes[direction]
        v = 2.0 * t * (2.0 * np.pi / L) * np.sin(2.0 * np.pi * kc[direction] / L)
        if v == 0.0:
            continue
        kq = [(kc[i] + q3[i]) % d3[i] for i in range(3)]
        kqf = kq[0] + d3[0] * (kq[1] + d3[1] * kq[2])
        e += _mono(complex(v), _op(False, 0, kqf), _op(True, 0, k))
        e += _mono(complex(v), _op(False, 1, kqf), _op(True, 1, k))
    return e


def hubbard_momentum_hamiltonian(t: float, u: float, sizes) -> FermionExpression:
    """HubbardModelMomentum::hamiltonian() (hubbard_model_momentum.h:119-175):
    sum_k eps(k) (n_k,up + n_k,dn) + (U/N) sum_{k1,k2,q}
    c+_{k1+q,up} c+_{k2-q,dn} c_{k2,dn} c_{k1,up}; k = kx + Lx ky + Lx Ly kz,
    eps(k) = -2t sum_d cos(2 pi k_d / L_d)."""
    sizes = [int(v) for v in sizes]
    d3 = sizes + [1] * (3 - len(sizes))
    N = int(np.prod(sizes))

    def coords(k):
        return [k % d3[0], (k // d3[0]) % d3[1], k // (d3[0] * d3[1])]

    def flat(c):
        return c[0] + d3[0] * (c[1] + d3[1] * c[2])

    e = FermionExpression()
    for k in range(N):
        kc = coords(k)
        eps = 0.0
        for dd in range(len(sizes)):
            eps += -2.0 * t * np.cos(2.0 * np.pi * kc[dd] / sizes[dd])
        e += _mono(complex(eps), _op(False, 0, k), _op(True, 0, k))
        e += _mono(complex(eps), _op(False, 1, k), _op(True, 1, k))
    uc = complex(u / N)
    for k1 in range(N):
        for k2 in range(N):
            for q in range(N):
                a, b, c = coords(k1), coords(k2), coords(q)
                k1q = flat([(a[i] + c[i]) % d3[i] for i in range(3)])
                k2q = flat([(b[i] + d3[i] - c[i]) % d3[i] for i in range(3)])
                e += _mono(uc, _op(False, 0, k1q), _op(False, 1, k2q), _op(True, 1, k2),
                           _op(True, 0, k1))
    return e


# ------------------------------------------------------------ sparse matrices

class SparseMatrix:
    """arma::sp_cx_mat as the reference stores it: nonzeros column-major with rows
    ascending; save / load in the cereal binary format (arma_cereal.h:9-30)."""

    def __init__(self, handle):
        from . import lib, _check
        self._h = handle
        self._lib = lib()  # kept for __del__, which may run at interpreter shutdown
        r, c, n = C.c_uint64(), C.c_uint64(), C.c_int64()
        _check(lib().mb_sparse_shape(handle, C.byref(r), C.byref(c), C.byref(n)))
        self.shape = (int(r.value), int(c.value))
        self.nnz = int(n.value)

    def __del__(self):
        h, L = getattr(self, "_h", None), getattr(self, "_lib", None)
        if h and L is not None:
            L.mb_sparse_destroy(h)
            self._h = None

    def entries(self):
        """(rows uint64, cols uint64, values complex128) in storage order."""
        from . import lib, _check
        n = max(self.nnz, 1)
        rows = np.zeros(n, np.uint64)
        cols = np.zeros(n, np.uint64)
        vals = np.zeros(2 * n, np.float64)
        _check(lib().mb_sparse_entries(self._h, rows.ctypes.data_as(C.c_void_p),
                                       cols.ctypes.data_as(C.c_void_p),
                                       vals.ctypes.data_as(C.c_void_p)))
        return rows[:self.nnz], cols[:self.nnz], (vals[0::2] + 1j * vals[1::2])[:self.nnz]

    def to_scipy(self):
        import scipy.sparse as sp
        r, c, v = self.entries()
        return sp.csc_matrix((v, (r.astype(np.int64), c.astype(np.int64))), shape=self.shape)

    def to_dense(self):
        M = np.zeros(self.shape, np.complex128)
        r, c, v = self.entries()
        M[r.astype(np.int64), c.astype(np.int64)] = v
        return M

    def save(self, path: str):
        from . import lib, _check
        _check(lib().mb_sparse_save_cereal(self._h, str(path).encode()))

    @staticmethod
    def load(path: str) -> "SparseMatrix":
        from . import lib, _check
        h = C.c_void_p()
        _check(lib().mb_sparse_load_cereal(str(path).encode(), C.byref(h)))
        return SparseMatrix(h)

    @staticmethod
    def from_entries(shape, rows, cols, vals) -> "SparseMatrix":
        """From entries; they are sorted column-major here (duplicates rejected)."""
        from . import lib, _check
        rows = np.asarray(rows, np.uint64)
        cols = np.asarray(cols, np.uint64)
        vals = np.asarray(vals, np.complex128)
        order = np.lexsort((rows, cols))
        rows, cols, vals = rows[order], cols[order], vals[order]
        v = np.ascontiguousarray(np.stack([vals.real, vals.imag], 1).reshape(-1))
        h = C.c_void_p()
        _check(lib().mb_sparse_create(int(shape[0]), int(shape[1]), len(vals),
                                      np.ascontiguousarray(rows).ctypes.data_as(C.c_void_p),
                                      np.ascontiguousarray(cols).ctypes.data_as(C.c_void_p),
                                      v.ctypes.data_as(C.c_void_p), C.byref(h)))
        return SparseMatrix(h)
