"""numpy front end of the oracle (oracle/hubbard_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker or the timed CPU baseline,
never by the product package.

Every function mirrors one reference routine; the citations are on the C side.
Complex vectors are numpy complex128 arrays in reference index order unless a
function says "product order".
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")
_lib = None

_i64p = np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_u64p = np.ctypeslib.ndpointer(np.uint64, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_c128p = np.ctypeslib.ndpointer(np.complex128, flags="C_CONTIGUOUS")


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            raise RuntimeError(f"oracle library missing: {_LIB_PATH} (run make -C oracle)")
        L = C.CDLL(_LIB_PATH)
        L.orc_choose.restype = C.c_uint64
        L.orc_choose.argtypes = [C.c_uint64, C.c_uint64]
        L.orc_basis_enumerate.restype = C.c_int64
        L.orc_basis_enumerate.argtypes = [C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_int64]
        L.orc_hubbard_bonds.restype = C.c_int
        L.orc_hubbard_bonds.argtypes = [C.c_int, _i32p, _i32p, _i32p, C.c_int]
        L.orc_build_csc.restype = C.c_int64
        L.orc_build_csc.argtypes = [C.c_int, _u64p, _u64p, C.c_int64, C.c_int, _i32p, _i32p,
                                    C.c_double, C.c_double, _i64p, C.c_void_p, C.c_void_p,
                                    C.c_int64, C.c_int]
        L.orc_set_progress.restype = None
        L.orc_set_progress.argtypes = [C.c_int]
        L.orc_set_threads.restype = None
        L.orc_set_threads.argtypes = [C.c_int]
        L.orc_spmv_csc.restype = None
        L.orc_spmv_csc.argtypes = [C.c_int64, _i64p, _i64p, _c128p, _c128p, _c128p]
        L.orc_species_table.restype = C.c_int64
        L.orc_species_table.argtypes = [C.c_int, C.c_int, C.c_int, _i32p, _i32p, C.c_double,
                                        _u64p, _i64p, _i32p, _f64p, _f64p, C.c_int64]
        prod_args = [C.c_int64, C.c_int64, _u64p, _u64p, _i64p, _i32p, _f64p, _f64p,
                     _i64p, _i32p, _f64p, _f64p, _f64p]
        L.orc_hv_product.restype = None
        L.orc_hv_product.argtypes = prod_args + [_c128p, _c128p, C.c_int]
        L.orc_diag_table.restype = None
        L.orc_diag_table.argtypes = [C.c_double, C.c_int, _f64p]
        L.orc_lanczos_csc.restype = C.c_int
        L.orc_lanczos_csc.argtypes = [C.c_int64, _i64p, _i64p, _c128p, C.c_int, _c128p, C.c_int,
                                      _f64p, _f64p, C.POINTER(C.c_double)]
        L.orc_lanczos_product.restype = C.c_int
        L.orc_lanczos_product.argtypes = prod_args + [_c128p, C.c_int, _f64p, _f64p,
                                                      C.POINTER(C.c_double), C.c_int]
        L.orc_tridiag_eigvals.restype = None
        L.orc_tridiag_eigvals.argtypes = [C.c_int, _f64p, _f64p, _f64p]
        L.orc_jacobi_eig.restype = None
        L.orc_jacobi_eig.argtypes = [C.c_int, _f64p, _f64p, _f64p]
        L.orc_lanczos_pass_two_csc.restype = None
        L.orc_lanczos_pass_two_csc.argtypes = [C.c_int64, _i64p, _i64p, _c128p, C.c_int, _c128p,
                                               C.c_int, _f64p, _f64p, C.c_double, _f64p, _c128p]
        L.orc_time_evolve_csc.restype = None
        L.orc_time_evolve_csc.argtypes = [C.c_int64, _i64p, _i64p, _c128p, _c128p, C.c_double,
                                          C.c_int, _c128p]
        L.orc_kpm_moments_csc.restype = None
        L.orc_kpm_moments_csc.argtypes = [C.c_int64, _i64p, _i64p, _c128p, C.c_double, C.c_double,
                                          C.c_int, C.c_int, _c128p, _f64p]
        L.orc_time_evolve_product.restype = None
        L.orc_time_evolve_product.argtypes = prod_args + [_c128p, C.c_double, C.c_int, _c128p,
                                                          C.c_int]
        L.orc_kpm_moments_product.restype = None
        L.orc_kpm_moments_product.argtypes = prod_args + [C.c_double, C.c_double, C.c_int,
                                                          C.c_int, _c128p, _f64p, C.c_int]
        L.orc_jackson_kernel.restype = None
        L.orc_jackson_kernel.argtypes = [C.c_int, _f64p]
        _lib = L
    return _lib


def choose(n: int, m: int) -> int:
    return int(lib().orc_choose(n, m))


def basis_states(n_sites: int, n_up: int, n_dn: int):
    """(up_bits, dn_bits) arrays in reference index order (basis.cpp:267-302)."""
    D = lib().orc_basis_enumerate(n_sites, n_up, n_dn, None, None, 0)
    if D < 0:
        raise ValueError("invalid sector")
    up = np.zeros(D, np.uint64)
    dn = np.zeros(D, np.uint64)
    got = lib().orc_basis_enumerate(n_sites, n_up, n_dn, up.ctypes.data, dn.ctypes.data, D)
    assert got == D
    return up, dn


def hubbard_bonds(sizes) -> np.ndarray:
    """Bond list (from, to) of HubbardModel{,2D,3D} in loop order."""
    sz = np.asarray(sizes, np.int32)
    cap = 3 * int(np.prod(sz)) + 3
    f = np.zeros(cap, np.int32)
    t = np.zeros(cap, np.int32)
    n = lib().orc_hubbard_bonds(len(sz), sz, f, t, cap)
    if n < 0:
        raise ValueError("bad lattice")
    return np.stack([f[:n], t[:n]], axis=1)


@dataclass
class CSC:
    D: int
    col_ptr: np.ndarray
    row_idx: np.ndarray
    val: np.ndarray  # complex128

    def matvec(self, x: np.ndarray) -> np.ndarray:
        y = np.zeros(self.D, np.complex128)
        lib().orc_spmv_csc(self.D, self.col_ptr, self.row_idx, self.val,
                           np.ascontiguousarray(x, np.complex128), y)
        return y

    def to_scipy(self):
        import scipy.sparse as sp
        return sp.csc_matrix((self.val, self.row_idx, self.col_ptr), shape=(self.D, self.D))


def build_csc(sizes, n_up: int, n_dn: int, t: float, U: float, threads: int = 0,
              states=None) -> CSC:
    """compute_matrix_elements<sp_cx_mat>(basis, HubbardModel*.hamiltonian())."""
    L = int(np.prod(sizes))
    up, dn = states if states is not None else basis_states(L, n_up, n_dn)
    bonds = hubbard_bonds(sizes)
    f = np.ascontiguousarray(bonds[:, 0])
    to = np.ascontiguousarray(bonds[:, 1])
    D = len(up)
    cp = np.zeros(D + 1, np.int64)
    nnz = lib().orc_build_csc(L, up, dn, D, len(f), f, to, t, U, cp, None, None, 0, threads)
    if nnz < 0:
        raise RuntimeError("CSC build failed")
    ri = np.zeros(max(nnz, 1), np.int64)
    val = np.zeros(max(nnz, 1), np.complex128)
    got = lib().orc_build_csc(L, up, dn, D, len(f), f, to, t, U, cp, ri.ctypes.data,
                              val.ctypes.data, nnz, threads)
    assert got == nnz
    return CSC(D, cp, ri[:nnz].copy(), val[:nnz].copy())


@dataclass
class Species:
    strings: np.ndarray
    row_ptr: np.ndarray
    cols: np.ndarray
    vals: np.ndarray
    onsite: np.ndarray


def species_table(sizes, n: int, t: float) -> Species:
    L = int(np.prod(sizes))
    bonds = hubbard_bonds(sizes)
    f = np.ascontiguousarray(bonds[:, 0])
    to = np.ascontiguousarray(bonds[:, 1])
    Ds = choose(L, n)
    cap = Ds * (2 * len(f) + 1)
    strings = np.zeros(Ds, np.uint64)
    rp = np.zeros(Ds + 1, np.int64)
    cols = np.zeros(max(cap, 1), np.int32)
    vals = np.zeros(max(cap, 1), np.float64)
    onsite = np.zeros(Ds, np.float64)
    nnz = lib().orc_species_table(L, n, len(f), f, to, t, strings, rp, cols, vals, onsite, cap)
    if nnz < 0:
        raise RuntimeError("species table failed")
    return Species(strings, rp, cols[:nnz].copy(), vals[:nnz].copy(), onsite)


class ProductOperator:
    """Matrix-free H in product order x[u * D_dn + d] (u, d = ascending bitstring ranks)."""

    def __init__(self, sizes, n_up: int, n_dn: int, t: float, U: float, threads: int = 0):
        L = int(np.prod(sizes))
        self.up = species_table(sizes, n_up, t)
        self.dn = species_table(sizes, n_dn, t)
        self.diagU = np.zeros(L + 1, np.float64)
        lib().orc_diag_table(U, L, self.diagU)
        self.Dup = len(self.up.strings)
        self.Ddn = len(self.dn.strings)
        self.D = self.Dup * self.Ddn
        self.threads = threads

    def _args(self):
        u, d = self.up, self.dn
        return (self.Dup, self.Ddn, u.strings, d.strings, u.row_ptr, u.cols, u.vals, u.onsite,
                d.row_ptr, d.cols, d.vals, d.onsite, self.diagU)

    def matvec(self, x: np.ndarray) -> np.ndarray:
        y = np.zeros(self.D, np.complex128)
        lib().orc_hv_product(*self._args(), np.ascontiguousarray(x, np.complex128), y,
                             self.threads)
        return y

    def lanczos(self, b: np.ndarray, k: int):
        al = np.zeros(max(k, 1))
        be = np.zeros(max(k, 1))
        bn = C.c_double(0.0)
        m = lib().orc_lanczos_product(*self._args(), np.ascontiguousarray(b, np.complex128), k,
                                      al, be, C.byref(bn), self.threads)
        if m < 0:
            raise RuntimeError("Input vector has zero norm")
        return al[:m].copy(), be[:max(m - 1, 0)].copy(), bn.value

    def time_evolve(self, psi: np.ndarray, t: float, k: int = 30) -> np.ndarray:
        """exp(-i t H) psi with k Krylov steps (evolve_state.h:120-178), product order."""
        out = np.zeros(self.D, np.complex128)
        lib().orc_time_evolve_product(*self._args(), np.ascontiguousarray(psi, np.complex128),
                                      t, k, out, self.threads)
        return out

    def time_evolve_steps(self, psi: np.ndarray, t0: float, t1: float, dt: float,
                          k: int = 30) -> np.ndarray:
        """time_evolve_state_steps (evolve_state.h:204-231): steps of dt from t0,
        the last one clipped to t1."""
        if dt == 0.0 or t1 == t0:
            return psi
        t = t0
        while t < t1:
            step = min(dt, t1 - t)
            psi = self.time_evolve(psi, step, k)
            t += step
        return psi

    def kpm_moments(self, a: float, b: float, M: int, vecs: np.ndarray) -> np.ndarray:
        """mu_n = (1 / (R D)) sum_r Re <r|T_n((H - b) / a)|r>, n < M, product order."""
        vecs = np.ascontiguousarray(np.atleast_2d(vecs), np.complex128)
        mu = np.zeros(M)
        lib().orc_kpm_moments_product(*self._args(), a, b, M, vecs.shape[0], vecs, mu,
                                      self.threads)
        return mu


def product_permutation(sizes, n_up: int, n_dn: int, ref_states=None) -> np.ndarray:
    """perm[ref_index] = product index u * D_dn + d, from the DFS enumeration
    and sorted single-species strings (independent of any closed-form rank)."""
    L = int(np.prod(sizes))
    up, dn = ref_states if ref_states is not None else basis_states(L, n_up, n_dn)
    su = species_table(sizes, n_up, 1.0).strings
    sd = species_table(sizes, n_dn, 1.0).strings
    u = np.searchsorted(su, up)
    d = np.searchsorted(sd, dn)
    return (u.astype(np.int64) * len(sd) + d).astype(np.int64)


def lanczos_csc(A: CSC, b: np.ndarray, k: int, negate: bool = False):
    al = np.zeros(max(k, 1))
    be = np.zeros(max(k, 1))
    bn = C.c_double(0.0)
    m = lib().orc_lanczos_csc(A.D, A.col_ptr, A.row_idx, A.val, int(negate),
                              np.ascontiguousarray(b, np.complex128), k, al, be, C.byref(bn))
    if m < 0:
        raise RuntimeError("Input vector has zero norm")
    return al[:m].copy(), be[:max(m - 1, 0)].copy(), bn.value


def tridiag_eigvals(a, b) -> np.ndarray:
    a = np.ascontiguousarray(a, np.float64)
    m = len(a)
    bb = np.ascontiguousarray(np.resize(np.asarray(b, np.float64), max(m - 1, 1)), np.float64)
    if m > 1:
        bb[: m - 1] = np.asarray(b, np.float64)[: m - 1]
    w = np.zeros(m)
    lib().orc_tridiag_eigvals(m, a, bb, w)
    return w


def jacobi_eig(A: np.ndarray):
    A = np.ascontiguousarray(A, np.float64)
    m = A.shape[0]
    w = np.zeros(m)
    V = np.zeros((m, m))
    lib().orc_jacobi_eig(m, A, w, V)
    return w, V


def find_min_eigenpair_csc(A: CSC, b: np.ndarray, k: int):
    """find_min_eigenpair (lanczos.h:283-290) with an explicit start vector:
    Lanczos on -H, top eigenpair of T, pass two on -H."""
    al, be, bn = lanczos_csc(A, b, k, negate=True)
    m = len(al)
    T = np.diag(al) + np.diag(be[: m - 1], 1) + np.diag(be[: m - 1], -1)
    w, V = jacobi_eig(T)
    y = np.ascontiguousarray(V[:, m - 1] * bn)
    x = np.zeros(A.D, np.complex128)
    lib().orc_lanczos_pass_two_csc(A.D, A.col_ptr, A.row_idx, A.val, 1,
                                   np.ascontiguousarray(b, np.complex128), m, al,
                                   np.ascontiguousarray(np.resize(be, max(m, 1))), bn, y, x)
    return -w[m - 1], x


def time_evolve_csc(A: CSC, psi: np.ndarray, t: float, k: int = 30) -> np.ndarray:
    out = np.zeros(A.D, np.complex128)
    lib().orc_time_evolve_csc(A.D, A.col_ptr, A.row_idx, A.val,
                              np.ascontiguousarray(psi, np.complex128), t, k, out)
    return out


def kpm_moments_csc(A: CSC, a: float, b: float, M: int, vecs: np.ndarray) -> np.ndarray:
    vecs = np.ascontiguousarray(vecs, np.complex128)
    R = vecs.shape[0]
    mu = np.zeros(M)
    lib().orc_kpm_moments_csc(A.D, A.col_ptr, A.row_idx, A.val, a, b, M, R, vecs, mu)
    return mu


def jackson_kernel(N: int) -> np.ndarray:
    g = np.zeros(N + 1)
    lib().orc_jackson_kernel(N, g)
    return g


def lattice_shift_sites(sizes, r: int) -> np.ndarray:
    """site i -> site i + r on the periodic lattice (site = x + Lx y + Lx Ly z,
    index.h:82-103; r is the flattened displacement rx + Lx ry + Lx Ly rz)."""
    dims = list(sizes) + [1] * (3 - len(sizes))
    Lx, Ly, Lz = dims
    rx, ry, rz = r % Lx, (r // Lx) % Ly, r // (Lx * Ly)
    out = np.zeros(Lx * Ly * Lz, np.int64)
    for z in range(Lz):
        for y in range(Ly):
            for x in range(Lx):
                out[x + Lx * y + Lx * Ly * z] = ((x + rx) % Lx) + Lx * ((y + ry) % Ly) + \
                    Lx * Ly * ((z + rz) % Lz)
    return out


def opposite_spin_correlation_diag(sizes, n_up: int, n_dn: int) -> np.ndarray:
    """Diagonal of G_updn(r) = (1/N) sum_i n_{i,up} n_{i+r,dn}
    (hubbard_model.h:40-53, 105-120, 184-200) for every displacement r, in
    reference index order: array [N, D]."""
    N = int(np.prod(sizes))
    up, dn = basis_states(N, n_up, n_dn)
    up = np.asarray(up, np.uint64)
    dn = np.asarray(dn, np.uint64)
    out = np.zeros((N, len(up)))
    for r in range(N):
        m = lattice_shift_sites(sizes, r)
        for i in range(N):
            out[r] += ((up >> np.uint64(i)) & np.uint64(1)).astype(np.float64) * \
                ((dn >> np.uint64(m[i])) & np.uint64(1)).astype(np.float64)
    return out / N


def opposite_spin_correlation(sizes, n_up: int, n_dn: int, x_ref: np.ndarray) -> np.ndarray:
    """<x|G_updn(r)|x> / <x|x> for every displacement r."""
    p = np.abs(np.asarray(x_ref)) ** 2
    return opposite_spin_correlation_diag(sizes, n_up, n_dn) @ p / p.sum()


# ------------------------------------------------------------ momentum sectors
# Fixed total momentum K (basis.cpp:128-185, 304-356; the DFS of the real-space
# basis with a filter at the leaves, so the sector index is the rank among the
# K-states in the real-space reference order) and the factorized momentum-space
# Hubbard operator of hubbard_momentum_factorized_fixed_k.h:92-117:
#   H = t sum_{k,s} eps(k) n_{k,s} + (U/N) sum_q rho_{q,up} rho_{-q,dn},
#   eps(k) = -2 sum_d cos(2 pi k_d / L_d), rho_{q,s} = sum_k c+_{k+q,s} c_{k,s}.
# Orbital k is the momentum with coordinates k = k_x + L_x k_y + L_x L_y k_z.

def _dims3(sizes):
    return list(sizes) + [1] * (3 - len(sizes))


def momentum_add(sizes, k: int, q: int, sign: int = 1) -> int:
    """Flat index of k + sign * q (component-wise, periodic)."""
    Lx, Ly, Lz = _dims3(sizes)
    kc = (k % Lx, (k // Lx) % Ly, k // (Lx * Ly))
    qc = (q % Lx, (q // Lx) % Ly, q // (Lx * Ly))
    c = [(kc[d] + sign * qc[d]) % (Lx, Ly, Lz)[d] for d in range(3)]
    return c[0] + Lx * c[1] + Lx * Ly * c[2]


def string_momentum(sizes, bits: int) -> int:
    """Total momentum (flat) of an occupation string."""
    tot = 0
    k = 0
    while bits:
        if bits & 1:
            tot = momentum_add(sizes, tot, k)
        bits >>= 1
        k += 1
    return tot


def dispersion(sizes, k: int) -> float:
    """eps(k) = -2 sum_d cos(2 pi k_d / L_d) (hubbard_momentum_factorized_fixed_k.h:201-215)."""
    Lx, Ly, Lz = _dims3(sizes)
    kc = (k % Lx, (k // Lx) % Ly, k // (Lx * Ly))
    e = 0.0
    for d, L in enumerate(list(sizes)):
        e += -2.0 * np.cos(2.0 * np.pi * kc[d] / L)
    return e


def momentum_sector_states(sizes, n_up: int, n_dn: int, K: int):
    """(up, dn) bit strings of the sector, in reference index order."""
    N = int(np.prod(sizes))
    up, dn = basis_states(N, n_up, n_dn)
    keep = [i for i in range(len(up))
            if momentum_add(sizes, string_momentum(sizes, int(up[i])),
                            string_momentum(sizes, int(dn[i]))) == K]
    return [int(up[i]) for i in keep], [int(dn[i]) for i in keep]


def _move(bits: int, k: int, kq: int):
    """c+_{kq} c_k on one species' string: (new bits, sign) or None if zero.
    sign = (-1)^(occupied orbitals strictly between k and kq)."""
    if not (bits >> k) & 1:
        return None
    if kq == k:
        return bits, 1
    if (bits >> kq) & 1:
        return None
    lo, hi = min(k, kq), max(k, kq)
    between = bin(bits & (((1 << hi) - 1) ^ ((1 << (lo + 1)) - 1))).count("1")
    return (bits ^ (1 << k)) | (1 << kq), (-1) ** between


def momentum_hamiltonian(sizes, n_up: int, n_dn: int, K: int, t: float, U: float):
    """Dense matrix of the fixed-K factorized operator on the sector (reference order):
    column j = H |j>, rho_{-q,dn} applied first, then rho_{q,up}."""
    N = int(np.prod(sizes))
    ups, dns = momentum_sector_states(sizes, n_up, n_dn, K)
    idx = {(u, d): i for i, (u, d) in enumerate(zip(ups, dns))}
    D = len(ups)
    H = np.zeros((D, D), dtype=np.complex128)
    eps = [dispersion(sizes, k) for k in range(N)]
    for j, (u, d) in enumerate(zip(ups, dns)):
        e = 0.0
        for k in range(N):
            if (u >> k) & 1:
                e += eps[k]
        for k in range(N):
            if (d >> k) & 1:
                e += eps[k]
        H[j, j] += t * e
        if abs(U) <= 1e-15:
            continue
        for q in range(N):
            mq = momentum_add(sizes, 0, q, -1)
            for kd in range(N):
                rd = _move(d, kd, momentum_add(sizes, kd, mq))
                if rd is None:
                    continue
                for ku in range(N):
                    ru = _move(u, ku, momentum_add(sizes, ku, q))
                    if ru is None:
                        continue
                    i = idx.get((ru[0], rd[0]))
                    if i is not None:
                        H[i, j] += U / N * ru[1] * rd[1]
    return H


# ------------------------------------------------------------ optical conductivity
# Current operator J_d(Q) = sum_{k,s} v_d(k) c+_{k+Q,s} c_{k,s}, sector K -> K + Q,
# v_d(k) = 2 t (2 pi / L_d) sin(2 pi k_d / L_d) (current_operator_fixed_k.h:14-293),
# and the damped Fourier transform of the current-current correlator
# (optical_conductivity.cpp:8-80).

def velocity(sizes, t: float, direction: int, k: int) -> float:
    Lx, Ly, Lz = _dims3(sizes)
    kc = (k % Lx, (k // Lx) % Ly, k // (Lx * Ly))
    L = list(sizes)[direction]
    return 2.0 * t * (2.0 * np.pi / L) * np.sin(2.0 * np.pi * kc[direction] / L)


def current_operator(sizes, n_up: int, n_dn: int, K: int, Q: int, direction: int, t: float):
    """Dense J_d(Q) from sector K (columns) to sector K + Q (rows), reference order."""
    N = int(np.prod(sizes))
    src_u, src_d = momentum_sector_states(sizes, n_up, n_dn, K)
    KQ = momentum_add(sizes, K, Q)
    tgt_u, tgt_d = momentum_sector_states(sizes, n_up, n_dn, KQ)
    idx = {(u, d): i for i, (u, d) in enumerate(zip(tgt_u, tgt_d))}
    J = np.zeros((len(tgt_u), len(src_u)), dtype=np.complex128)
    v = [velocity(sizes, t, direction, k) for k in range(N)]
    for j, (u, d) in enumerate(zip(src_u, src_d)):
        for spin in (0, 1):
            bits = u if spin == 0 else d
            for k in range(N):
                if abs(v[k]) < 1e-15:
                    continue
                r = _move(bits, k, momentum_add(sizes, k, Q))
                if r is None:
                    continue
                key = (r[0], d) if spin == 0 else (u, r[0])
                i = idx.get(key)
                if i is not None:
                    J[i, j] += r[1] * v[k]
    return J


def optical_conductivity(correlation, dt: float, beta: float, volume: float):
    """(frequencies, sigma, sum_rule) of optical_conductivity.cpp:8-80: trapezoid
    weights, damping exp(-eta t) with eta = 4 / T, omega from 0 to pi/dt in steps
    of eta / 4, prefactor (1 - exp(-beta omega)) / (omega V) (beta / V at 0)."""
    c = np.asarray(correlation, dtype=np.complex128)
    n = len(c)
    if n < 2 or not dt > 0 or not volume > 0 or not beta >= 0:
        raise ValueError("invalid arguments")
    T = (n - 1) * dt
    eta = 4.0 / T
    dw = eta / 4.0
    nf = int(np.floor((np.pi / dt) / dw)) + 1
    w = np.full(n, dt)
    w[0] *= 0.5
    w[-1] *= 0.5
    series = c * w * np.exp(-eta * dt) ** np.arange(n)
    freqs = np.arange(nf) * dw
    sigma = np.zeros(nf, dtype=np.complex128)
    for i, om in enumerate(freqs):
        integral = np.sum(series * np.exp(1j * om * dt * np.arange(n)))
        pref = beta / volume if abs(om) < 1e-12 else -np.expm1(-beta * om) / (om * volume)
        sigma[i] = pref * integral
    wts = np.ones(nf)
    wts[0] = wts[-1] = 0.5
    return freqs, sigma, float(np.sum(sigma.real * wts) * dw)


# ------------------------------------------------------------ generic expressions

_TOL = 1000.0 * np.finfo(np.float64).eps   # tolerances::tolerance<double>()


def _add_to(target: dict, key, c: complex):
    """ExpressionBase::add_to (expression_base.h:85-97): drop |c| < 1000 eps,
    merge by key, erase a merged sum below 1000 eps."""
    if abs(c) ** 2 < _TOL * _TOL:
        return
    if key in target:
        target[key] += c
        if abs(target[key]) ** 2 < _TOL * _TOL:
            del target[key]
    else:
        target[key] = c


def expression_terms(terms):
    """[(coeff, [(type, spin, orbital), ...]), ...] -> {ops tuple: coeff} merged in
    the order given (FermionExpression += FermionExpression(FermionMonomial))."""
    out: dict = {}
    for c, ops in terms:
        _add_to(out, tuple(tuple(int(v) for v in op) for op in ops), complex(c))
    return out


def apply_monomial(ops, L: int, up: int, dn: int):
    """The product ops[0] ops[1] ... ops[-1] applied to the state
    c+_{up i1} ... c+_{up ia} c+_{dn j1} ... c+_{dn jb} |0> (ascending, up first: the
    reference's state key order, basis.cpp:273-276) one operator at a time, last
    first: c_m / c+_m on mode m = spin L + orbital gives (-1)^(occupied modes below
    m).  Returns (up, dn, sign) or None."""
    s = up | (dn << L)
    sign = 1
    for typ, spin, orb in reversed(ops):
        m = spin * L + orb
        occ = (s >> m) & 1
        if typ == 1 and not occ or typ == 0 and occ:
            return None
        if bin(s & ((1 << m) - 1)).count("1") & 1:
            sign = -sign
        s ^= 1 << m
    return s & ((1 << L) - 1), s >> L, sign


def expression_matrix(L: int, states, terms):
    """Dense <i|A|j> of compute_matrix_elements (matrix_elements.h:74-105) over
    `states` [(up, dn), ...] in basis order: each column from A|j>, outputs outside
    the basis dropped, contributions to one output merged by add_to."""
    expr = expression_terms(terms)
    idx = {(int(u), int(d)): i for i, (u, d) in enumerate(states)}
    D = len(states)
    M = np.zeros((D, D), dtype=np.complex128)
    for j, (u, d) in enumerate(states):
        col: dict = {}
        for ops, c in expr.items():
            r = apply_monomial(ops, L, int(u), int(d))
            if r is None:
                continue
            i = idx.get((r[0], r[1]))
            if i is not None:
                _add_to(col, i, c * r[2])
        for i, v in col.items():
            M[i, j] = v
    return M


def rectangular_matrix(L: int, row_states, col_states, terms):
    """Dense <i|A|j> of compute_rectangular_matrix_elements (matrix_elements.h:
    110-167): i over `row_states`, j over `col_states` (each [(up, dn), ...] in
    basis order); A|j> canonicalised and merged by add_to, outputs outside the
    row basis dropped (fermion state normalisation is 1)."""
    expr = expression_terms(terms)
    idx = {(int(u), int(d)): i for i, (u, d) in enumerate(row_states)}
    M = np.zeros((len(row_states), len(col_states)), dtype=np.complex128)
    for j, (u, d) in enumerate(col_states):
        col: dict = {}
        for ops, c in expr.items():
            r = apply_monomial(ops, L, int(u), int(d))
            if r is None:
                continue
            i = idx.get((r[0], r[1]))
            if i is not None:
                _add_to(col, i, c * r[2])
        for i, v in col.items():
            M[i, j] = v
    return M


def vector_elements(L: int, states, terms):
    """compute_vector_elements (matrix_elements.h:31-51): element i is the constant
    term of canonicalize(adjoint(<i|) A), i.e. <i|A|0> with |0> the vacuum."""
    return rectangular_matrix(L, states, [(0, 0)], terms)[:, 0]


def sparse_column_major(M):
    """(rows, cols, values) of the nonzeros of M in arma::SpMat iteration order:
    column by column, rows ascending."""
    cols, rows = np.nonzero(M.T)
    return rows.astype(np.uint64), cols.astype(np.uint64), M[rows, cols]


def cereal_sparse_bytes(n_rows: int, n_cols: int, rows, cols, vals) -> bytes:
    """The cereal BinaryOutputArchive bytes of arma_cereal.h's save()
    (src/utils/arma_cereal.h:9-16): the class version (uint32 0, written by cereal
    for a versioned save the first time the type is seen), n_rows, n_cols,
    n_nonzero (arma::uword, 64-bit), then per nonzero row, col (uint64) and the
    value as real, imag (float64), little endian."""
    import struct
    out = [struct.pack("<I", 0), struct.pack("<QQQ", n_rows, n_cols, len(vals))]
    for r, c, v in zip(rows, cols, vals):
        v = complex(v)
        out.append(struct.pack("<QQdd", int(r), int(c), v.real, v.imag))
    return b"".join(out)


# ------------------------------------------------------------ spectral bounds

def power_method(matvec, v, iterations: int) -> float:
    """power_method (linear_operator_utils.h:18-50): v normalised, then per
    iteration w = A v, estimate Re(sum v w) / Re(sum v v) with arma::dot (no
    conjugation), v = w / |w|; 0 for a zero vector or a zero denominator."""
    if iterations == 0:
        return 0.0
    v = np.asarray(v, np.complex128)
    n = np.linalg.norm(v)
    if n == 0.0:
        return 0.0
    v = v / n
    eig = 0.0
    for _ in range(iterations):
        w = matvec(v)
        den = np.sum(v * v).real
        if den == 0.0:
            return 0.0
        eig = np.sum(v * w).real / den
        n = np.linalg.norm(w)
        if n == 0.0:
            return 0.0
        v = w / n
    return float(eig)


def estimate_bounds(matvec, seed, power_iterations: int, spectral_padding: float):
    """estimate_bounds (linear_operator_utils.h:52-78): (alpha, beta) with
    alpha <= beta, from power iterations on A, A + shift and shift - A."""
    dominant = abs(power_method(matvec, seed, power_iterations))
    if dominant == 0.0:
        return 0.0, 0.0
    shift = dominant * (1.0 + spectral_padding) + np.finfo(np.float64).eps
    beta = power_method(lambda v: matvec(v) + shift * v, seed, power_iterations) - shift
    alpha = shift - power_method(lambda v: -matvec(v) + shift * v, seed, power_iterations)
    return (alpha, beta) if alpha <= beta else (beta, alpha)


# ------------------------------------------------------------ KPM projector

def projector_moments(M: int, epsilon: float) -> np.ndarray:
    """kpm::projector_moments (kpm.h:46-59): mu_0 = 1 - acos(eps)/pi,
    mu_m = -2 sin(m acos eps) / (m pi)."""
    ac = np.arccos(epsilon)
    m = np.arange(1, M + 1)
    return np.concatenate([[1.0 - ac / np.pi], -2.0 * np.sin(m * ac) / (m * np.pi)])


def lorentz_kernel(N: int, lam: float = 4.0) -> np.ndarray:
    """kpm::lorentz_kernel (kpm.h:32-43): sinh(lam (1 - n/N)) / sinh(lam)."""
    n = np.arange(N + 1)
    return np.sinh(lam * (1.0 - n / N)) / np.sinh(lam)


def kpm_projector_matrix(A, M: int, e_min: float, e_max: float, e_fermi: float = 0.0,
                         padding: float = 0.01) -> np.ndarray:
    """The matrix of SparseKPMProjector (kpm.h:184-241) for a dense A: the
    Chebyshev recurrence on matrices, damped by the Jackson kernel."""
    A = np.asarray(A, np.complex128)
    a = (e_max - e_min) / 2.0 * (1.0 + padding)
    b = (e_max + e_min) / 2.0
    Hs = (A - b * np.eye(len(A))) / a
    c = jackson_kernel(M) * projector_moments(M, (e_fermi - b) / a)
    Tp, P = np.eye(len(A), dtype=np.complex128), c[0] * np.eye(len(A), dtype=np.complex128)
    if M == 0:
        return P
    Tc = Hs.copy()
    P = P + c[1] * Tc
    for m in range(2, M + 1):
        Tn = 2.0 * Hs @ Tc - Tp
        P = P + c[m] * Tn
        Tp, Tc = Tc, Tn
    return P


def kpm_projector_apply(matvec, v, M: int, e_min: float, e_max: float, e_fermi: float = 0.0,
                        padding: float = 0.01) -> np.ndarray:
    """SparseKPMProjector::apply (kpm.h:219-233): the vector recurrence."""
    a = (e_max - e_min) / 2.0 * (1.0 + padding)
    b = (e_max + e_min) / 2.0
    c = jackson_kernel(M) * projector_moments(M, (e_fermi - b) / a)
    v = np.asarray(v, np.complex128)
    res = c[0] * v
    if M == 0:
        return res
    Tp, Tc = v, (matvec(v) - b * v) / a
    res = res + c[1] * Tc
    for m in range(2, M + 1):
        Tn = 2.0 * (matvec(Tc) - b * Tc) / a - Tp
        res = res + c[m] * Tn
        Tp, Tc = Tc, Tn
    return res


# ------------------------------------------------------------ fixed-N basis

def fixed_n_states(n_sites: int, n_particles: int):
    """Basis::with_fixed_particle_number (basis.cpp:70-89, 240-265): depth-first
    over modes (spin, orbital); at each level the up candidates from the current
    orbital on, then the down ones, a mode allowed when its orbital is above the
    last one's or equal with a higher spin; a leaf at N modes.  Returns
    (up_bits, dn_bits) lists in index order."""
    ups, dns = [], []

    def rec(cur, first):
        if len(cur) == n_particles:
            ups.append(sum(1 << o for s, o in cur if s == 0))
            dns.append(sum(1 << o for s, o in cur if s == 1))
            return
        for s in (0, 1):
            for i in range(first, n_sites):
                if not cur or cur[-1][1] < i or (cur[-1][1] == i and s > cur[-1][0]):
                    cur.append((s, i))
                    rec(cur, i)
                    cur.pop()

    rec([], 0)
    return ups, dns
