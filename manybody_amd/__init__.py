"""manybody_amd — MI355X-native Hubbard H.v path (Python mirror of the C ABI).

The product is the HIP library ``libmanybody_hip.so`` in this directory and its C
ABI (``include/manybody_hip.h``).  This module binds it with ctypes and mirrors
the reference's names (keyehzy/manybody):

=============================================  =====================================
reference                                      here
=============================================  =====================================
FermionBasis::with_fixed_particle_number_and_   FermionBasis.with_fixed_particle_
spin (src/algebra/basis.cpp:91-126)            number_and_spin
HubbardModel / 2D / 3D                          HubbardModel, HubbardModel2D,
(src/algebra/fermion/model/hubbard_model.h)    HubbardModel3D
SparseComplexMatrixOperator::apply              HubbardOperator.apply (reference
(src/numerics/linear_operator.h:81-91)          order) / .hv (device order)
lanczos_pass_one (src/numerics/lanczos.h:130)   lanczos_pass_one
find_min_eigenpair (lanczos.h:283-290)          find_min_eigenpair
time_evolve_state[_steps] (evolve_state.h)      time_evolve_state[_steps]
kpm::jackson_kernel (kpm.h:19-30)               jackson_kernel
rescaling::from_bounds (rescaling.h:33-38)      rescaling_from_bounds
=============================================  =====================================

Vectors are torch tensors on the GPU (torch is used for device memory and the
stream only).  Errors raise the Python analogue of the reference's exception:
std::runtime_error -> RuntimeError, std::invalid_argument -> ValueError,
std::out_of_range -> IndexError.  There is no CPU fallback: without the built
library every call raises.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, field
from typing import List, Optional

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libmanybody_hip.so")
HEADER_PATH = os.path.join(os.path.dirname(_HERE), "include", "manybody_hip.h")

MB_OK, MB_ERR_INVALID, MB_ERR_RUNTIME, MB_ERR_RANGE, MB_ERR_DEVICE, MB_ERR_NOMEM = range(6)

_lib = None

_vp = C.c_void_p
_dp = C.POINTER(C.c_double)
_SIGS = {
    "mb_last_error": (C.c_char_p, []),
    "mb_version": (C.c_char_p, []),
    "mb_basis_create": (C.c_int, [C.c_int, C.c_int, C.c_int, C.POINTER(_vp)]),
    "mb_basis_destroy": (None, [_vp]),
    "mb_basis_size": (C.c_int64, [_vp]),
    "mb_momentum_basis_create": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int, _vp, _vp,
                                           C.POINTER(_vp)]),
    "mb_momentum_hubbard_create": (C.c_int, [_vp, C.c_int, _vp, C.c_double, C.c_double, _vp,
                                             C.POINTER(_vp)]),
    "mb_basis_state": (C.c_int, [_vp, C.c_int64, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]),
    "mb_basis_index_of": (C.c_int, [_vp, C.c_uint64, C.c_uint64, C.POINTER(C.c_int64)]),
    "mb_basis_product_index": (C.c_int, [_vp, C.c_int64, C.c_int64, _vp]),
    "mb_hubbard_create": (C.c_int, [_vp, C.c_int, _vp, C.c_double, C.c_double, C.POINTER(_vp)]),
    "mb_hubbard_create_bonds": (C.c_int, [_vp, C.c_int, _vp, _vp, C.c_double, C.c_double,
                                          C.POINTER(_vp)]),
    "mb_hubbard_destroy": (None, [_vp]),
    "mb_hubbard_dimension": (C.c_int64, [_vp]),
    "mb_hubbard_shape": (C.c_int, [_vp, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]),
    "mb_hv_c128": (C.c_int, [_vp, _vp, _vp, _vp]),
    "mb_hv_c64": (C.c_int, [_vp, _vp, _vp, _vp]),
    "mb_hv_c64_multi": (C.c_int, [_vp, C.c_int, _vp, _vp, _vp]),
    "mb_to_device_order_c128": (C.c_int, [_vp, _vp, _vp, _vp]),
    "mb_to_reference_order_c128": (C.c_int, [_vp, _vp, _vp, _vp]),
    "mb_apply_ref_c128": (C.c_int, [_vp, _vp, _vp, _vp, _vp]),
    "mb_fill_gaussian_c128": (C.c_int, [_vp, C.c_uint64, _vp, _vp]),
    "mb_fill_phase_c64_multi": (C.c_int, [_vp, C.c_uint64, C.c_int, _vp, _vp]),
    "mb_fill_phase_c128_multi": (C.c_int, [_vp, C.c_uint64, C.c_int, _vp, _vp]),
    "mb_lanczos_pass_one": (C.c_int, [_vp, _vp, C.c_int, _dp, _dp, C.POINTER(C.c_int), _dp, _vp]),
    "mb_lanczos_pass_one_noalloc": (C.c_int, [_vp, _vp, C.c_int, _dp, _dp, C.POINTER(C.c_int),
                                              _dp, _vp]),
    "mb_find_min_eigenpair": (C.c_int, [_vp, _vp, C.c_int, _dp, _vp, _vp]),
    "mb_time_evolve": (C.c_int, [_vp, _vp, C.c_double, C.c_int, _vp, _vp]),
    "mb_time_evolve_steps": (C.c_int, [_vp, _vp, C.c_double, C.c_double, C.c_double, C.c_int,
                                       _vp, _vp]),
    "mb_tridiag_eig": (C.c_int, [C.c_int, _vp, _vp, _vp, _vp]),
    "mb_kpm_dos_moments": (C.c_int, [_vp, C.c_double, C.c_double, C.c_int, C.c_int, C.c_uint64,
                                     C.c_int, _dp, _vp]),
    "mb_jackson_kernel": (C.c_int, [C.c_int, _dp]),
    "mb_batch_create": (C.c_int, [C.c_int, _vp, C.POINTER(_vp)]),
    "mb_batch_destroy": (None, [_vp]),
    "mb_batch_footprint": (C.c_int64, [_vp]),
    "mb_batch_fill_gaussian": (C.c_int, [_vp, _vp, _vp]),
    "mb_batch_set_start": (C.c_int, [_vp, C.c_int, _vp, _vp]),
    "mb_batch_set_mapping": (C.c_int, [_vp, C.c_int, C.c_int]),
    "mb_batch_autotune": (C.c_int, [_vp, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int), _vp]),
    "mb_batch_set_persistent": (C.c_int, [_vp, C.c_int64]),
    "mb_batch_lanczos": (C.c_int, [_vp, C.c_int, C.c_int, _vp, _vp, _vp, _vp, _vp]),
    "mb_batch_ground_energies": (C.c_int, [_vp, C.c_int, C.c_int, _vp, _vp]),
    "mb_lanczos_pass_one_ex": (C.c_int, [_vp, _vp, C.c_int, C.c_int, _vp, _vp, C.POINTER(C.c_int),
                                         C.POINTER(C.c_double), _vp]),
    "mb_find_min_eigenpair_ex": (C.c_int, [_vp, _vp, C.c_int, C.c_int, C.POINTER(C.c_double), _vp,
                                           _vp]),
    "mb_time_evolve_ex": (C.c_int, [_vp, _vp, C.c_double, C.c_int, C.c_int, _vp, _vp]),
    "mb_time_evolve_steps_ex": (C.c_int, [_vp, _vp, C.c_double, C.c_double, C.c_double, C.c_int,
                                          C.c_int, _vp, _vp]),
    "mb_opposite_spin_correlation": (C.c_int, [_vp, _vp, _vp, _vp]),
    "mb_exp_apply": (C.c_int, [_vp, _vp, C.c_int, _vp, _vp]),
    "mb_projector_moments": (C.c_int, [C.c_int, C.c_double, _dp]),
    "mb_lorentz_kernel": (C.c_int, [C.c_int, C.c_double, _dp]),
    "mb_kpm_projector_apply": (C.c_int, [_vp, C.c_int, C.c_double, C.c_double, C.c_double,
                                         C.c_double, _vp, _vp, _vp]),
    "mb_time_evolve_steps_cb": (C.c_int, [_vp, _vp, C.c_double, C.c_double, C.c_double, C.c_int,
                                          C.c_int, _vp, _vp, _vp, _vp]),
    "mb_lanczos_solve": (C.c_int, [_vp, _vp, C.c_int, _vp, _vp, _vp, _vp]),
    "mb_find_max_eigenpair": (C.c_int, [_vp, _vp, C.c_int, C.POINTER(C.c_double), _vp, _vp]),
    "mb_power_method": (C.c_int, [_vp, _vp, C.c_int, C.POINTER(C.c_double), _vp]),
    "mb_estimate_bounds": (C.c_int, [_vp, _vp, C.c_int, C.c_double, C.POINTER(C.c_double),
                                     C.POINTER(C.c_double), _vp]),
    "mb_fixed_n_basis_create": (C.c_int, [C.c_int, C.c_int, C.POINTER(_vp)]),
    "mb_expression_create": (C.c_int, [_vp, C.c_int64, _vp, _vp, _vp, C.POINTER(_vp)]),
    "mb_compute_matrix_elements": (C.c_int, [_vp, C.POINTER(_vp), _vp]),
    "mb_compute_rectangular_matrix_elements": (C.c_int, [_vp, _vp, C.c_int64, _vp, _vp, _vp,
                                                         C.POINTER(_vp), _vp]),
    "mb_compute_vector_elements": (C.c_int, [_vp, C.c_int64, _vp, _vp, _vp, _vp, _vp]),
    "mb_operator_negated": (C.c_int, [_vp, C.POINTER(_vp)]),
    "mb_operator_scaled": (C.c_int, [_vp, C.c_double, C.c_double, C.POINTER(_vp)]),
    "mb_operator_shifted": (C.c_int, [_vp, C.c_double, C.c_double, C.POINTER(_vp)]),
    "mb_operator_neg_shifted": (C.c_int, [_vp, C.c_double, C.c_double, C.POINTER(_vp)]),
    "mb_operator_sum": (C.c_int, [_vp, _vp, C.POINTER(_vp)]),
    "mb_operator_difference": (C.c_int, [_vp, _vp, C.POINTER(_vp)]),
    "mb_operator_composed": (C.c_int, [_vp, _vp, C.POINTER(_vp)]),
    "mb_operator_identity": (C.c_int, [_vp, C.POINTER(_vp)]),
    "mb_operator_rescaled": (C.c_int, [_vp, C.c_double, C.c_double, C.POINTER(_vp)]),
    "mb_sparse_create": (C.c_int, [C.c_uint64, C.c_uint64, C.c_int64, _vp, _vp, _vp,
                                   C.POINTER(_vp)]),
    "mb_sparse_destroy": (None, [_vp]),
    "mb_sparse_shape": (C.c_int, [_vp, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64),
                                  C.POINTER(C.c_int64)]),
    "mb_sparse_entries": (C.c_int, [_vp, _vp, _vp, _vp]),
    "mb_sparse_save_cereal": (C.c_int, [_vp, C.c_char_p]),
    "mb_sparse_load_cereal": (C.c_int, [C.c_char_p, C.POINTER(_vp)]),
    "mb_current_create": (C.c_int, [_vp, _vp, C.c_int, _vp, C.c_double, C.c_int64, C.c_int,
                                    C.POINTER(_vp)]),
    "mb_current_destroy": (None, [_vp]),
    "mb_current_apply": (C.c_int, [_vp, _vp, _vp, _vp]),
    "mb_current_adjoint_apply": (C.c_int, [_vp, _vp, _vp, _vp]),
    "mb_imaginary_time_evolve": (C.c_int, [_vp, _vp, C.c_double, C.c_int, _vp, _vp]),
    "mb_optical_correlator": (C.c_int, [_vp, _vp, _vp, C.c_int, C.c_double, C.c_double, C.c_int,
                                        C.c_int, C.c_uint64, _dp, _vp]),
    "mb_optical_conductivity": (C.c_int, [_dp, C.c_int, C.c_double, C.c_double, C.c_double, _dp,
                                          _dp, C.c_int, C.POINTER(C.c_int), _dp]),
    "mb_energy_residual": (C.c_int, [_vp, _vp, C.POINTER(C.c_double), C.POINTER(C.c_double), _vp]),
    "mb_server_create": (C.c_int, [C.c_int, C.c_int64, C.c_int64, C.POINTER(_vp)]),
    "mb_server_destroy": (None, [_vp]),
    "mb_server_submit": (C.c_int, [_vp, _vp, C.POINTER(C.c_int64)]),
    "mb_server_wait": (C.c_int, [_vp, C.c_int64, _vp, _vp, C.c_int]),
    "mb_server_drain": (C.c_int, [_vp]),
    "mb_server_now": (C.c_double, [_vp]),
    "mb_request_footprint": (C.c_int64, [_vp]),
    "mb_rescaling_from_bounds": (C.c_int, [C.c_double, C.c_double, C.c_double, _dp, _dp]),
}


def lib():
    """Load libmanybody_hip.so (raises if it has not been built).

    torch carries its own HIP runtime, and it has to bring the device up before
    this library's does (loading the library registers its device code with the
    HIP runtime), or torch later finds no device and the library's allocations
    fail: so on a machine with a GPU (/dev/kfd present) torch is initialised
    first.  Without a GPU the library loads alone: host-only users (bases,
    index lookups, the KPM and rescaling helpers, the tridiagonal solver) never
    import torch."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"HIP library not built: {LIB_PATH} (run __graft_entry__.build())")
        if os.path.exists("/dev/kfd"):
            try:
                import torch
                torch.cuda.is_available()
            except ImportError:
                pass
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _check(status: int):
    if status == MB_OK:
        return
    msg = lib().mb_last_error().decode()
    if status == MB_ERR_INVALID:
        raise ValueError(msg)
    if status == MB_ERR_RANGE:
        raise IndexError(msg)
    if status == MB_ERR_NOMEM:
        raise MemoryError(msg)
    raise RuntimeError(msg)


def _stream():
    import torch
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def _ptr(t) -> C.c_void_p:
    if not t.is_cuda:
        raise ValueError("expected a GPU tensor")
    if not t.is_contiguous():
        raise ValueError("expected a contiguous tensor")
    return C.c_void_p(t.data_ptr())


# ------------------------------------------------------------------ basis

class FermionBasis:
    """FermionBasis restricted to fixed particle number and spin projection."""

    def __init__(self, handle, orbitals: int, n_up: int, n_dn: int):
        self._h = handle
        self.orbitals = orbitals
        self.n_up = n_up
        self.n_dn = n_dn
        self.particles = (n_up + n_dn) if n_up is not None else None

    @staticmethod
    def with_fixed_particle_number_and_spin(orbitals: int, particles: int,
                                            spin_projection: int) -> "FermionBasis":
        h = C.c_void_p()
        _check(lib().mb_basis_create(orbitals, particles, spin_projection, C.byref(h)))
        up = (particles + spin_projection) // 2
        return FermionBasis(h, orbitals, up, particles - up)

    @staticmethod
    def with_fixed_particle_number(orbitals: int, particles: int) -> "FermionBasis":
        """Basis::with_fixed_particle_number (basis.cpp:70-89, 240-265): every state
        of `particles` particles, any N_up, in the reference's order."""
        h = C.c_void_p()
        _check(lib().mb_fixed_n_basis_create(int(orbitals), int(particles), C.byref(h)))
        b = FermionBasis(h, orbitals, None, None)
        b.particles = int(particles)
        return b

    @staticmethod
    def sector(orbitals: int, n_up: int, n_dn: int) -> "FermionBasis":
        return FermionBasis.with_fixed_particle_number_and_spin(orbitals, n_up + n_dn, n_up - n_dn)

    @staticmethod
    def with_fixed_particle_number_spin_momentum(orbitals: int, particles: int,
                                                 spin_projection: int, lattice_size,
                                                 momentum) -> "FermionBasis":
        """Basis::with_fixed_particle_number_spin_momentum (basis.cpp:128-185): the
        states of total momentum `momentum` on the periodic lattice `lattice_size`
        (orbital k = the momentum k_x + L_x k_y + L_x L_y k_z), in the reference's
        order (its real-space order restricted to the sector)."""
        sizes = tuple(int(v) for v in lattice_size)
        K = tuple(int(v) for v in momentum)
        h = C.c_void_p()
        _check(lib().mb_momentum_basis_create(orbitals, particles, spin_projection, len(sizes),
                                              (C.c_int * len(sizes))(*sizes),
                                              (C.c_int * len(K))(*K), C.byref(h)))
        up = (particles + spin_projection) // 2
        b = FermionBasis(h, orbitals, up, particles - up)
        b.lattice_size, b.momentum = sizes, K
        return b

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.mb_basis_destroy(self._h)
            self._h = None

    def size(self) -> int:
        return int(lib().mb_basis_size(self._h))

    __len__ = size

    def state(self, index: int):
        """(up_bits, dn_bits) of basis.set[index]."""
        u, d = C.c_uint64(), C.c_uint64()
        _check(lib().mb_basis_state(self._h, index, C.byref(u), C.byref(d)))
        return int(u.value), int(d.value)

    def index_of(self, up_bits: int, dn_bits: int) -> int:
        i = C.c_int64()
        _check(lib().mb_basis_index_of(self._h, up_bits, dn_bits, C.byref(i)))
        return int(i.value)

    def contains(self, up_bits: int, dn_bits: int) -> bool:
        try:
            self.index_of(up_bits, dn_bits)
            return True
        except IndexError:
            return False

    def product_index(self, first: int = 0, n: Optional[int] = None):
        import numpy as np
        n = self.size() - first if n is None else n
        out = np.zeros(n, np.int64)
        _check(lib().mb_basis_product_index(self._h, first, n, out.ctypes.data_as(C.c_void_p)))
        return out


# ------------------------------------------------------------------ models

class _HubbardBase:
    def _bonds_args(self):
        raise NotImplementedError

    def hamiltonian(self):
        """kinetic() + interaction() as a FermionExpression (hubbard_model.h)."""
        from .expression import hubbard_hamiltonian
        return hubbard_hamiltonian(self.t, self.u, self.sizes)

    def operator(self, basis: FermionBasis) -> "HubbardOperator":
        sizes = (C.c_int * len(self.sizes))(*self.sizes)
        h = C.c_void_p()
        _check(lib().mb_hubbard_create(basis._h, len(self.sizes), sizes, self.t, self.u, C.byref(h)))
        return HubbardOperator(h, basis, self)


class HubbardModel(_HubbardBase):
    """HubbardModel(t, u, size) — hubbard_model.h:10-59 (ring of `size` sites)."""

    def __init__(self, t: float, u: float, size: int):
        self.t, self.u, self.size = float(t), float(u), int(size)
        self.sizes = (self.size,)


class HubbardModel2D(_HubbardBase):
    """HubbardModel2D(t, u, size_x, size_y) — hubbard_model.h:61-129."""

    def __init__(self, t: float, u: float, size_x: int, size_y: int):
        self.t, self.u = float(t), float(u)
        self.size_x, self.size_y = int(size_x), int(size_y)
        self.sizes = (self.size_x, self.size_y)


class HubbardModel3D(_HubbardBase):
    """HubbardModel3D(t, u, size_x, size_y, size_z) — hubbard_model.h:131-212."""

    def __init__(self, t: float, u: float, size_x: int, size_y: int, size_z: int):
        self.t, self.u = float(t), float(u)
        self.sizes = (int(size_x), int(size_y), int(size_z))


def HubbardMomentumFactorizedFixedK(basis: FermionBasis, lattice_size, t: float, u: float,
                                    target_momentum) -> "HubbardOperator":
    """HubbardMomentumFactorizedFixedK (hubbard_momentum_factorized_fixed_k.h:41-90):
    H = t sum eps(k) n_k + (U/N) sum_q rho_{q,up} rho_{-q,dn} on a fixed-momentum
    basis.  Returns an operator with the same methods as the real-space one
    (FP64 H.v, orders, Lanczos, eigenpair, time evolution, energy/residual)."""
    sizes = tuple(int(v) for v in lattice_size)
    K = tuple(int(v) for v in target_momentum)
    h = C.c_void_p()
    _check(lib().mb_momentum_hubbard_create(basis._h, len(sizes), (C.c_int * len(sizes))(*sizes),
                                            float(t), float(u), (C.c_int * len(K))(*K), C.byref(h)))
    return HubbardOperator(h, basis, None)


def FermionExpressionOperator(basis: FermionBasis, expression) -> "HubbardOperator":
    """compute_matrix_elements(basis, A) (matrix_elements.h:74-105) of any
    FermionExpression A as a device operator: FP64 H.v, orders, fills, Lanczos,
    eigenpair, time evolution and energy/residual, on a fixed-N or
    fixed-momentum basis."""
    coeffs, offsets, ops = expression.arrays()
    h = C.c_void_p()
    _check(lib().mb_expression_create(basis._h, len(offsets) - 1,
                                      coeffs.ctypes.data_as(C.c_void_p),
                                      offsets.ctypes.data_as(C.c_void_p),
                                      ops.ctypes.data_as(C.c_void_p), C.byref(h)))
    op = HubbardOperator(h, basis, None)
    op.expression = expression
    return op


def compute_matrix_elements(basis_or_op, expression=None):
    """The sparse matrix <i|A|j> of compute_matrix_elements over the basis, in
    the reference's index order (built on the device): a SparseMatrix."""
    from .expression import SparseMatrix
    op = basis_or_op if expression is None else FermionExpressionOperator(basis_or_op, expression)
    h = C.c_void_p()
    _check(lib().mb_compute_matrix_elements(op._h, C.byref(h), _stream()))
    return SparseMatrix(h)


def _composite(fn, operands, *args):
    h = C.c_void_p()
    _check(fn(*[o._h for o in operands], *args, C.byref(h)))
    op = HubbardOperator(h, operands[0].basis, None)
    op._operands = operands  # the composite references its operands
    return op


def Negated(op: "HubbardOperator") -> "HubbardOperator":
    """Negated (linear_operator.h:93-105): -A."""
    return _composite(lib().mb_operator_negated, (op,))


def Scaled(op: "HubbardOperator", scale) -> "HubbardOperator":
    """Scaled (linear_operator.h:107-120): c A, c complex."""
    c = complex(scale)
    return _composite(lib().mb_operator_scaled, (op,), c.real, c.imag)


def Shifted(op: "HubbardOperator", shift) -> "HubbardOperator":
    """Shifted (linear_operator.h:122-135): A + c I."""
    c = complex(shift)
    return _composite(lib().mb_operator_shifted, (op,), c.real, c.imag)


def NegShift(op: "HubbardOperator", shift) -> "HubbardOperator":
    """NegShift (linear_operator.h:137-150): -A + c I."""
    c = complex(shift)
    return _composite(lib().mb_operator_neg_shifted, (op,), c.real, c.imag)


def Sum(a: "HubbardOperator", b: "HubbardOperator") -> "HubbardOperator":
    """Sum (linear_operator.h:152-172): A + B on bases of the same kind."""
    return _composite(lib().mb_operator_sum, (a, b))


def Difference(a: "HubbardOperator", b: "HubbardOperator") -> "HubbardOperator":
    """Difference (linear_operator.h:174-194): A - B."""
    return _composite(lib().mb_operator_difference, (a, b))


def Composed(a: "HubbardOperator", b: "HubbardOperator") -> "HubbardOperator":
    """Composed (linear_operator.h:196-216): A B, B applied first."""
    return _composite(lib().mb_operator_composed, (a, b))


def Identity(basis: FermionBasis) -> "HubbardOperator":
    """Identity (linear_operator.h:218-228) on the basis."""
    h = C.c_void_p()
    _check(lib().mb_operator_identity(basis._h, C.byref(h)))
    return HubbardOperator(h, basis, None)


def compute_rectangular_matrix_elements(row_basis: FermionBasis, col_basis: FermionBasis,
                                        expression):
    """compute_rectangular_matrix_elements (matrix_elements.h:110-167): the sparse
    matrix <i|A|j>, i in row_basis, j in col_basis (reference index order), built
    on the device.  A may change the particle numbers; the bases may be of any
    kinds with the same orbital count."""
    from .expression import SparseMatrix
    coeffs, offsets, ops = expression.arrays()
    h = C.c_void_p()
    _check(lib().mb_compute_rectangular_matrix_elements(
        row_basis._h, col_basis._h, len(offsets) - 1, coeffs.ctypes.data_as(C.c_void_p),
        offsets.ctypes.data_as(C.c_void_p), ops.ctypes.data_as(C.c_void_p), C.byref(h),
        _stream()))
    return SparseMatrix(h)


def compute_vector_elements(basis: FermionBasis, expression):
    """compute_vector_elements (matrix_elements.h:31-51): <i|A|0> for every state i
    of the basis (|0> the vacuum), a complex128 numpy vector in reference order."""
    import numpy as np
    coeffs, offsets, ops = expression.arrays()
    out = np.zeros(2 * max(basis.size(), 1), np.float64)
    _check(lib().mb_compute_vector_elements(
        basis._h, len(offsets) - 1, coeffs.ctypes.data_as(C.c_void_p),
        offsets.ctypes.data_as(C.c_void_p), ops.ctypes.data_as(C.c_void_p),
        out.ctypes.data_as(C.c_void_p), _stream()))
    return (out[0::2] + 1j * out[1::2])[:basis.size()]


class HubbardOperator:
    """The LinearOperator of a Hubbard model on a fixed-N basis (device tables)."""

    def __init__(self, handle, basis: FermionBasis, model):
        self._h = handle
        self.basis = basis  # keeps the basis alive (the model references it)
        self.model = model
        # a real-space Hubbard model on a fixed (N_up, N_dn) sector: the kind
        # with the FP32 H.v, FP32 Krylov and multi-vector KPM paths
        self.real_space = (model is not None and basis.n_up is not None
                           and not hasattr(basis, "momentum"))
        du, dd = C.c_int64(), C.c_int64()
        _check(lib().mb_hubbard_shape(self._h, C.byref(du), C.byref(dd)))
        self.d_up, self.d_dn = int(du.value), int(dd.value)

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.mb_hubbard_destroy(self._h)
            self._h = None

    # the reference's operator algebra (linear_operator.h:232-281): -A, A + B,
    # A - B, A * B (composition), scalar * A and A * scalar; numpy scalars defer
    # to these methods
    __array_ufunc__ = None

    def __neg__(self):
        return Negated(self)

    def __add__(self, other):
        if not isinstance(other, HubbardOperator):
            return NotImplemented  # A + c I is Shifted(A, c), as in the reference
        return Sum(self, other)

    def __sub__(self, other):
        if not isinstance(other, HubbardOperator):
            return NotImplemented
        return Difference(self, other)

    def __mul__(self, other):
        if isinstance(other, HubbardOperator):
            return Composed(self, other)
        if not isinstance(other, (int, float, complex)) and not hasattr(other, "__complex__"):
            return NotImplemented
        return Scaled(self, other)

    def __rmul__(self, other):
        if not isinstance(other, (int, float, complex)) and not hasattr(other, "__complex__"):
            return NotImplemented
        return Scaled(self, other)

    def dimension(self) -> int:
        return int(lib().mb_hubbard_dimension(self._h))

    # --- device order: planar real tensors of shape (2 * nvec, D)
    def empty(self, dtype=None, nvec: int = 1):
        """Uninitialised device-order storage for nvec complex vectors: a real
        tensor of shape (2 nvec, D), rows 2r and 2r+1 = re and im of vector r."""
        import torch
        dtype = dtype or torch.float64
        return torch.empty((2 * nvec, self.dimension()), dtype=dtype, device="cuda")

    def hv(self, x, y=None):
        """y = H x in device order: float64 (2, D) = one complex128 vector, float32
        (2 nvec, D) = nvec complex64 vectors."""
        import torch
        y = torch.empty_like(x) if y is None else y
        D = self.dimension()
        if x.numel() % (2 * D):
            raise ValueError("x must hold 2 * nvec planes of the operator's dimension")
        nvec = x.numel() // (2 * D)
        if x.dtype == torch.float64:
            if nvec != 1:
                raise ValueError("float64 H.v takes one complex vector")
            _check(lib().mb_hv_c128(self._h, _ptr(x), _ptr(y), _stream()))
        elif x.dtype == torch.float32:
            if nvec == 1:
                _check(lib().mb_hv_c64(self._h, _ptr(x), _ptr(y), _stream()))
            else:
                _check(lib().mb_hv_c64_multi(self._h, nvec, _ptr(x), _ptr(y), _stream()))
        else:
            raise ValueError("x must be float64 or float32 planes")
        return y

    def to_device_order(self, x_ref):
        """complex128 vector in reference index order -> device order (2, D)."""
        out = self.empty()
        _check(lib().mb_to_device_order_c128(self._h, _ptr(x_ref), _ptr(out), _stream()))
        return out

    def to_reference_order(self, x_dev):
        """device order (2, D) -> complex128 vector in reference index order."""
        import torch
        out = torch.empty(self.dimension(), dtype=torch.complex128, device="cuda")
        _check(lib().mb_to_reference_order_c128(self._h, _ptr(x_dev), _ptr(out), _stream()))
        return out

    def apply(self, x_ref):
        """LinearOperator::apply in reference order: returns a new complex128 vector."""
        import torch
        y = torch.empty_like(x_ref)
        work = self.empty(nvec=2)
        _check(lib().mb_apply_ref_c128(self._h, _ptr(x_ref), _ptr(y), _ptr(work), _stream()))
        return y

    def opposite_spin_correlation(self, x):
        """<x|G_updn(r)|x>/<x|x> for every displacement r (hubbard_model.h:40-53);
        x in device order.  Returns a numpy array of one value per site."""
        import numpy as np
        g = np.zeros(self.basis.orbitals)
        _check(lib().mb_opposite_spin_correlation(self._h, _ptr(x), g.ctypes.data_as(C.c_void_p),
                                                  _stream()))
        return g

    def energy_residual(self, x):
        """(<x|H|x>/<x|x>, |Hx - E x|/|x|) for x in device order."""
        e, r = C.c_double(), C.c_double()
        _check(lib().mb_energy_residual(self._h, _ptr(x), C.byref(e), C.byref(r), _stream()))
        return e.value, r.value

    def gaussian(self, seed: int):
        """Device-order complex vector (2, D), re/im ~ N(0,1) per reference index."""
        x = self.empty()
        _check(lib().mb_fill_gaussian_c128(self._h, seed, _ptr(x), _stream()))
        return x

    def random_phase(self, seed: int, nvec: int, fp32: Optional[bool] = None):
        """nvec random-phase vectors e^{i phi} in device order (2 nvec, D);
        fp32=None: FP32 for real-space Hubbard operators (the precision their
        KPM path runs in), FP64 otherwise."""
        import torch
        if fp32 is None:
            fp32 = self.real_space
        x = self.empty(torch.float32 if fp32 else torch.float64, nvec)
        if fp32:
            _check(lib().mb_fill_phase_c64_multi(self._h, seed, nvec, _ptr(x), _stream()))
        else:
            _check(lib().mb_fill_phase_c128_multi(self._h, seed, nvec, _ptr(x), _stream()))
        return x


def to_planar(z, device="cuda", dtype=None):
    """numpy complex array (D,) or (nvec, D) -> device-order tensor (2 nvec, D)."""
    import numpy as np
    import torch
    z = np.atleast_2d(np.asarray(z))
    planes = np.empty((2 * z.shape[0], z.shape[1]), dtype=np.float64)
    planes[0::2] = z.real
    planes[1::2] = z.imag
    t = torch.from_numpy(planes)
    if dtype is not None:
        t = t.to(dtype)
    return t.to(device)


def from_planar(t):
    """device-order tensor (2 nvec, D) -> numpy complex128 (D,) or (nvec, D)."""
    a = t.detach().cpu().numpy().astype("float64").reshape(-1, t.shape[-1])
    z = a[0::2] + 1j * a[1::2]
    return z[0] if z.shape[0] == 1 else z


# ------------------------------------------------------------------ Krylov

@dataclass
class LanczosDecomposition:
    alphas: List[float] = field(default_factory=list)
    betas: List[float] = field(default_factory=list)
    steps_taken: int = 0
    b_norm: float = 0.0


@dataclass
class EigenPair:
    value: float
    vector: object = None


@dataclass
class EvolutionOptions:
    krylov_steps: int = 30


LANCZOS_REORTHOGONALIZE = 1


def lanczos_pass_one(op: HubbardOperator, b, k: int,
                     reorthogonalize: bool = False) -> LanczosDecomposition:
    al = (C.c_double * max(k, 1))()
    be = (C.c_double * max(k, 1))()
    steps = C.c_int()
    bn = C.c_double()
    if reorthogonalize:
        _check(lib().mb_lanczos_pass_one_ex(op._h, _ptr(b), k, LANCZOS_REORTHOGONALIZE, al, be,
                                            C.byref(steps), C.byref(bn), _stream()))
    else:
        _check(lib().mb_lanczos_pass_one(op._h, _ptr(b), k, al, be, C.byref(steps), C.byref(bn),
                                         _stream()))
    m = steps.value
    return LanczosDecomposition(list(al[:m]), list(be[:max(m - 1, 0)]), m, bn.value)


def lanczos_pass_one_noalloc(op: HubbardOperator, b, k: int) -> LanczosDecomposition:
    """lanczos_pass_one with no device allocation (mb_lanczos_pass_one_noalloc):
    the work buffers come from the library's static per-GPU scratch array, so
    D is limited to about 1.04e6 (ValueError beyond); real-space Hubbard
    operators only."""
    al = (C.c_double * max(k, 1))()
    be = (C.c_double * max(k, 1))()
    steps = C.c_int()
    bn = C.c_double()
    _check(lib().mb_lanczos_pass_one_noalloc(op._h, _ptr(b), k, al, be, C.byref(steps),
                                             C.byref(bn), _stream()))
    m = steps.value
    return LanczosDecomposition(list(al[:m]), list(be[:max(m - 1, 0)]), m, bn.value)


def find_min_eigenpair(op: HubbardOperator, k: int, b, want_vector: bool = True,
                       reorthogonalize: bool = False) -> EigenPair:
    import torch
    val = C.c_double()
    x = op.empty() if want_vector else None
    xp = _ptr(x) if x is not None else None
    if reorthogonalize:
        _check(lib().mb_find_min_eigenpair_ex(op._h, _ptr(b), k, LANCZOS_REORTHOGONALIZE,
                                              C.byref(val), xp, _stream()))
    else:
        _check(lib().mb_find_min_eigenpair(op._h, _ptr(b), k, C.byref(val), xp, _stream()))
    return EigenPair(val.value, x)


EVOLVE_FP32 = 1


def time_evolve_state(op: HubbardOperator, psi, t: float,
                      opts: EvolutionOptions = EvolutionOptions(), fp32: bool = False):
    """exp(-iHt) psi (evolve_state.h:164-178).  fp32: Krylov vectors and H.v in
    FP32, state / alpha-beta reductions / exponential in FP64."""
    out = op.empty()
    _check(lib().mb_time_evolve_ex(op._h, _ptr(psi), t, opts.krylov_steps,
                                   EVOLVE_FP32 if fp32 else 0, _ptr(out), _stream()))
    return out


_STEP_CALLBACK = C.CFUNCTYPE(C.c_int, C.c_double, C.c_void_p, C.c_void_p)


def time_evolve_state_steps(op: HubbardOperator, psi, t0: float, t1: float, dt: float,
                            opts: EvolutionOptions = EvolutionOptions(), fp32: bool = False,
                            callback=None):
    """time_evolve_state_steps (evolve_state.h:199-231); callback(t, state) at t0
    and after every step, state = the device-order tensor being evolved (valid
    during the call; clone it to keep it)."""
    out = op.empty()
    if callback is None:
        _check(lib().mb_time_evolve_steps_ex(op._h, _ptr(psi), t0, t1, dt, opts.krylov_steps,
                                             EVOLVE_FP32 if fp32 else 0, _ptr(out), _stream()))
        return out

    def cb(t, _state, _user):
        try:
            callback(t, out)
            return 0
        except Exception:
            return 1

    f = _STEP_CALLBACK(cb)
    _check(lib().mb_time_evolve_steps_cb(op._h, _ptr(psi), t0, t1, dt, opts.krylov_steps,
                                         EVOLVE_FP32 if fp32 else 0, C.cast(f, C.c_void_p), None,
                                         _ptr(out), _stream()))
    return out


def imaginary_time_evolve_state(op: HubbardOperator, psi, tau: float,
                                opts: EvolutionOptions = EvolutionOptions()):
    """exp(-tau H) psi (imaginary_time_evolve_state, evolve_state.h:184-197), device order."""
    out = op.empty()
    _check(lib().mb_imaginary_time_evolve(op._h, _ptr(psi), tau, opts.krylov_steps, _ptr(out),
                                          _stream()))
    return out


class CurrentOperatorFixedK:
    """CurrentOperatorFixedK(source_basis, target_basis, lattice_size, t, Q, direction)
    (current_operator_fixed_k.h:28-69): J_d(Q) = sum_{k,s} v_d(k) c+_{k+Q,s} c_{k,s}
    with v_d(k) = 2 t (2 pi / L_d) sin(2 pi k_d / L_d), from the fixed-momentum
    basis K to K + Q.  Vectors are device-order planes (2, D) of the two sectors."""

    def __init__(self, source_basis: FermionBasis, target_basis: FermionBasis, lattice_size,
                 t: float, Q: int, direction: int):
        sizes = tuple(int(v) for v in lattice_size)
        h = C.c_void_p()
        _check(lib().mb_current_create(source_basis._h, target_basis._h, len(sizes),
                                       (C.c_int * len(sizes))(*sizes), float(t), int(Q),
                                       int(direction), C.byref(h)))
        self._h = h
        self.source_basis, self.target_basis = source_basis, target_basis

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.mb_current_destroy(self._h)
            self._h = None

    def source_dimension(self) -> int:
        return self.source_basis.size()

    def target_dimension(self) -> int:
        return self.target_basis.size()

    def apply(self, x):
        """J x: device order of sector K -> device order of sector K + Q."""
        import torch
        y = torch.empty((2, self.target_dimension()), dtype=torch.float64, device="cuda")
        _check(lib().mb_current_apply(self._h, _ptr(x), _ptr(y), _stream()))
        return y

    def adjoint_apply(self, y):
        """J^dagger y: sector K + Q -> sector K."""
        import torch
        x = torch.empty((2, self.source_dimension()), dtype=torch.float64, device="cuda")
        _check(lib().mb_current_adjoint_apply(self._h, _ptr(y), _ptr(x), _stream()))
        return x


def optical_correlator(h_k: HubbardOperator, h_kq: HubbardOperator, J: CurrentOperatorFixedK,
                       samples: int, beta: float, dt: float, steps: int, seed: int,
                       opts: EvolutionOptions = EvolutionOptions()):
    """Stochastic <J^dagger(t) J(0)> of hubbard_momentum_optical_conductivity_
    factorized.cpp:226-283 at t = i dt, i < steps: sample n starts from
    h_k.gaussian(seed + n), normalised and evolved by exp(-beta H_K / 2)."""
    import numpy as np
    c = np.zeros(2 * steps)
    _check(lib().mb_optical_correlator(h_k._h, h_kq._h, J._h, samples, beta, dt, steps,
                                       opts.krylov_steps, seed,
                                       c.ctypes.data_as(_dp), _stream()))
    return c[0::2] + 1j * c[1::2]


class OpticalConductivityResult:
    def __init__(self, frequencies, sigma, sum_rule):
        self.frequencies, self.sigma, self.sum_rule = frequencies, sigma, sum_rule


def compute_optical_conductivity(correlation, dt: float, beta: float,
                                 volume: float) -> OpticalConductivityResult:
    """compute_optical_conductivity (optical_conductivity.cpp:8-80)."""
    import numpy as np
    c = np.ascontiguousarray(correlation, dtype=np.complex128).view(np.float64)
    n = len(c) // 2
    nf = C.c_int()
    _check(lib().mb_optical_conductivity(c.ctypes.data_as(_dp), n, dt, beta, volume, None, None,
                                         0, C.byref(nf), None))
    om = np.zeros(nf.value)
    sg = np.zeros(2 * nf.value)
    sr = C.c_double()
    _check(lib().mb_optical_conductivity(c.ctypes.data_as(_dp), n, dt, beta, volume,
                                         om.ctypes.data_as(_dp), sg.ctypes.data_as(_dp),
                                         nf.value, C.byref(nf), C.byref(sr)))
    return OpticalConductivityResult(om, sg[0::2] + 1j * sg[1::2], sr.value)


_TRIDIAG_SOLVER = C.CFUNCTYPE(C.c_int, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double),
                              C.POINTER(C.c_double), C.c_void_p)


def solve(op: HubbardOperator, b, k: int, solver):
    """solve(op, b, k, solver) (lanczos.h:220-237): solver(alphas, betas) returns
    the m first-column coefficients y (real or complex); the result
    |b| sum_j y_j v_j in device order."""
    import numpy as np

    def cb(m, al, be, y, _user):
        try:
            a = np.ctypeslib.as_array(al, (m,)).copy()
            bb = np.ctypeslib.as_array(be, (m - 1,)).copy() if m > 1 else np.zeros(0)
            r = np.asarray(solver(a, bb), dtype=np.complex128).reshape(-1)
            if r.shape != (m,):
                return 1
            out = np.ctypeslib.as_array(y, (2 * m,))
            out[0::2] = r.real
            out[1::2] = r.imag
            return 0
        except Exception:
            return 1

    f = _TRIDIAG_SOLVER(cb)
    x = op.empty()
    _check(lib().mb_lanczos_solve(op._h, _ptr(b), int(k), C.cast(f, C.c_void_p), None, _ptr(x),
                                  _stream()))
    return x


def find_max_eigenpair(op: HubbardOperator, k: int, b, want_vector: bool = True) -> "EigenPair":
    """find_max_eigenpair(op, k) (lanczos.h:246-281) with an explicit start vector."""
    val = C.c_double()
    x = op.empty() if want_vector else None
    _check(lib().mb_find_max_eigenpair(op._h, _ptr(b), int(k), C.byref(val),
                                       _ptr(x) if x is not None else None, _stream()))
    return EigenPair(val.value, x)


class Exp:
    """Exp(op, ExpOptions{krylov_steps}) (exp_operator.h:50-89): apply(v) = exp(H) v
    in device order."""

    def __init__(self, op: HubbardOperator, krylov_steps: int = 30):
        self.op, self.krylov_steps = op, int(krylov_steps)

    def dimension(self) -> int:
        return self.op.dimension()

    def apply(self, v):
        out = self.op.empty()
        _check(lib().mb_exp_apply(self.op._h, _ptr(v), self.krylov_steps, _ptr(out), _stream()))
        return out


def power_method(op: HubbardOperator, seed, iterations: int) -> float:
    """power_method(op, v, iterations) (linear_operator_utils.h:18-50), seed in
    device order (not modified)."""
    v = C.c_double()
    _check(lib().mb_power_method(op._h, _ptr(seed), int(iterations), C.byref(v), _stream()))
    return v.value


@dataclass
class Bounds:
    alpha: float
    beta: float


def estimate_bounds(op: HubbardOperator, seed, power_iterations: int,
                    spectral_padding: float) -> Bounds:
    """estimate_bounds (linear_operator_utils.h:52-78) with an explicit seed."""
    a, b = C.c_double(), C.c_double()
    _check(lib().mb_estimate_bounds(op._h, _ptr(seed), int(power_iterations),
                                    float(spectral_padding), C.byref(a), C.byref(b), _stream()))
    return Bounds(a.value, b.value)


def kpm_dos_moments(op: HubbardOperator, a: float, b: float, M: int, nvec: int, seed: int,
                    fp32: Optional[bool] = None):
    """KPM density-of-states moments mu_0 .. mu_{M-1} of (H - b) / a with nvec
    random-phase vectors (phases keyed by reference index).  Real-space Hubbard
    operators run the multi-vector FP32 / FP64 path; every other operator
    (momentum sectors, fixed-N bases, expressions, composites) runs FP64 only,
    one vector at a time, with the same phases.  fp32=None picks FP32 H.v for
    real-space Hubbard operators and FP64 otherwise; fp32=True on another
    operator raises ValueError."""
    if fp32 is None:
        fp32 = op.real_space
    mu = (C.c_double * M)()
    _check(lib().mb_kpm_dos_moments(op._h, a, b, M, nvec, seed, int(fp32), mu, _stream()))
    return list(mu)


def tridiag_eig(alphas, betas, vectors: bool = True):
    import numpy as np
    a = np.ascontiguousarray(alphas, np.float64)
    m = len(a)
    b = np.ascontiguousarray(betas, np.float64) if m > 1 else np.zeros(1)
    w = np.zeros(max(m, 1))
    V = np.zeros((max(m, 1), max(m, 1)))
    _check(lib().mb_tridiag_eig(m, a.ctypes.data_as(C.c_void_p), b.ctypes.data_as(C.c_void_p),
                                w.ctypes.data_as(C.c_void_p),
                                V.ctypes.data_as(C.c_void_p) if vectors else None))
    # V is column-major: V_cm[i + m*j] = component i of vector j
    return w[:m], (V.reshape(max(m, 1), max(m, 1)).T[:m, :m] if vectors else None)


def projector_moments(M: int, epsilon: float):
    """kpm::projector_moments (kpm.h:46-59): M + 1 moments of theta(epsilon - x)."""
    mu = (C.c_double * (M + 1))()
    _check(lib().mb_projector_moments(int(M), float(epsilon), mu))
    return list(mu)


def lorentz_kernel(N: int, lam: float = 4.0):
    """kpm::lorentz_kernel (kpm.h:32-43): N + 1 damping factors."""
    g = (C.c_double * (N + 1))()
    _check(lib().mb_lorentz_kernel(int(N), float(lam), g))
    return list(g)


class SparseKPMProjector:
    """SparseKPMProjector(H, M, E_min, E_max, E_fermi, padding) (kpm.h:184-241) on a
    device operator: apply(v) = sum_m g_m mu_m T_m(H~) v (device order)."""

    def __init__(self, op: HubbardOperator, expansion_order: int, e_min: float, e_max: float,
                 e_fermi: float = 0.0, padding: float = 0.01):
        self.op, self.M = op, int(expansion_order)
        self.e_min, self.e_max, self.e_fermi, self.padding = (float(e_min), float(e_max),
                                                              float(e_fermi), float(padding))

    def dimension(self) -> int:
        return self.op.dimension()

    def apply(self, v):
        out = self.op.empty()
        _check(lib().mb_kpm_projector_apply(self.op._h, self.M, self.e_min, self.e_max,
                                            self.e_fermi, self.padding, _ptr(v), _ptr(out),
                                            _stream()))
        return out


class KPMProjector(SparseKPMProjector):
    """KPMProjector(H, M, E_fermi, padding) (kpm.h:67-180) on a device operator:
    the rescaling is estimate_rescaling(H, padding), whose extreme eigenvalues
    come from Lanczos (min(k, D) steps) where the reference diagonalises the
    dense H; apply is SparseKPMProjector's recurrence; build_matrix applies the
    projector to every unit vector (reference order, small operators)."""

    def __init__(self, op: HubbardOperator, expansion_order: int, e_fermi: float = 0.0,
                 padding: float = 0.01, k: int = 300, seed: int = 1):
        b = op.gaussian(seed=seed)
        kk = min(int(k), op.dimension())
        e_min = find_min_eigenpair(op, kk, b, want_vector=False).value
        e_max = find_max_eigenpair(op, kk, b, want_vector=False).value
        super().__init__(op, expansion_order, e_min, e_max, e_fermi, padding)
        self._rescaling = Rescaling.from_bounds(e_min, e_max, padding)
        eps = self._rescaling.rescale(self.e_fermi)
        if eps < -1.0 or eps > 1.0:
            raise RuntimeError("Fermi energy outside rescaled spectrum")

    def expansion_order(self) -> int:
        return self.M

    def rescaling(self) -> Rescaling:
        return self._rescaling

    def build_matrix(self):
        """The dense projector (kpm.h:144-170), columns P e_j in reference order."""
        import numpy as np
        import torch
        D = self.dimension()
        P = np.zeros((D, D), np.complex128)
        e = torch.zeros(D, dtype=torch.complex128, device="cuda")
        for j in range(D):
            e.zero_()
            e[j] = 1.0
            col = self.op.to_reference_order(self.apply(self.op.to_device_order(e)))
            P[:, j] = col.cpu().numpy()
        return P


def jackson_kernel(N: int):
    g = (C.c_double * (N + 1))()
    _check(lib().mb_jackson_kernel(N, g))
    return list(g)


def rescaling_from_bounds(e_min: float, e_max: float, padding: float = 0.01):
    a, b = C.c_double(), C.c_double()
    _check(lib().mb_rescaling_from_bounds(e_min, e_max, padding, C.byref(a), C.byref(b)))
    return a.value, b.value


class Rescaling:
    """rescaling::Rescaling (rescaling.h:9-19): H_scaled = (H - b) / a."""

    def __init__(self, a: float, b: float):
        self.a, self.b = float(a), float(b)

    def rescale(self, E: float) -> float:
        return (E - self.b) / self.a

    def inverse(self, E_scaled: float) -> float:
        return E_scaled * self.a + self.b

    @staticmethod
    def from_bounds(e_min: float, e_max: float, padding: float = 0.01) -> "Rescaling":
        """rescaling::from_bounds (rescaling.h:33-38)."""
        return Rescaling(*rescaling_from_bounds(e_min, e_max, padding))


def estimate_rescaling(op: "HubbardOperator", padding: float = 0.01, k: int = 300,
                       seed: int = 1) -> Rescaling:
    """rescaling::estimate_rescaling (rescaling.h:21-31) for a device operator: the
    extreme eigenvalues from Lanczos (min(k, D) steps from a seeded Gaussian
    start; exact up to rounding once k >= D) where the reference diagonalises a
    dense matrix."""
    b = op.gaussian(seed=seed)
    k = min(int(k), op.dimension())
    e_min = find_min_eigenpair(op, k, b, want_vector=False).value
    e_max = find_max_eigenpair(op, k, b, want_vector=False).value
    return Rescaling.from_bounds(e_min, e_max, padding)


def rescale_hamiltonian(op: "HubbardOperator", r: Rescaling) -> "HubbardOperator":
    """rescaling::rescale_hamiltonian (rescaling.h:40-48): (H - b I) / a as an
    operator (composite; it references op)."""
    return _composite(lib().mb_operator_rescaled, (op,), r.a, r.b)


# ------------------------------------------------------------------ batched

class Batch:
    """Many independent Lanczos solves in lockstep (one launch per phase of a
    step for the whole batch).  Keeps its operators (and their bases) alive."""

    def __init__(self, ops):
        self.ops = list(ops)
        arr = (C.c_void_p * len(self.ops))(*[op._h for op in self.ops])
        h = C.c_void_p()
        _check(lib().mb_batch_create(len(self.ops), arr, C.byref(h)))
        self._h = h

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.mb_batch_destroy(self._h)
            self._h = None

    def __len__(self):
        return len(self.ops)

    def footprint(self) -> int:
        return int(lib().mb_batch_footprint(self._h))

    def fill_gaussian(self, seeds):
        import numpy as np
        s = np.ascontiguousarray(seeds, np.uint64)
        _check(lib().mb_batch_fill_gaussian(self._h, s.ctypes.data_as(C.c_void_p), _stream()))

    def set_start(self, i: int, b):
        _check(lib().mb_batch_set_start(self._h, i, _ptr(b), _stream()))

    def set_mapping(self, tile: int, block: int):
        _check(lib().mb_batch_set_mapping(self._h, tile, block))

    def set_persistent(self, max_dim: int):
        """Instances up to max_dim run every step in the persistent per-instance
        kernel (0: all tiled over the per-phase kernels)."""
        _check(lib().mb_batch_set_persistent(self._h, int(max_dim)))

    def autotune(self, k_probe: int = 3):
        t, b = C.c_int(), C.c_int()
        _check(lib().mb_batch_autotune(self._h, k_probe, C.byref(t), C.byref(b), _stream()))
        return t.value, b.value

    def lanczos(self, k: int, use_graph: bool = True):
        import numpy as np
        n = len(self.ops)
        al = np.zeros((n, k))
        be = np.zeros((n, k))
        st = np.zeros(n, np.int32)
        bn = np.zeros(n)
        vp = lambda a: a.ctypes.data_as(C.c_void_p)
        _check(lib().mb_batch_lanczos(self._h, k, int(use_graph), vp(al), vp(be), vp(st), vp(bn),
                                      _stream()))
        return [LanczosDecomposition(list(al[i, :st[i]]), list(be[i, :max(st[i] - 1, 0)]),
                                     int(st[i]), float(bn[i])) for i in range(n)]

    def ground_energies(self, k: int, use_graph: bool = True):
        import numpy as np
        e0 = np.zeros(len(self.ops))
        _check(lib().mb_batch_ground_energies(self._h, k, int(use_graph),
                                              e0.ctypes.data_as(C.c_void_p), _stream()))
        return e0


# ------------------------------------------------------------------ serving

REQ_GROUND, REQ_KPM, REQ_EVOLVE = 0, 1, 2


class _Request(C.Structure):
    _fields_ = [("kind", C.c_int), ("ndim", C.c_int), ("sizes", C.c_int * 3),
                ("n_up", C.c_int), ("n_dn", C.c_int), ("t", C.c_double), ("u", C.c_double),
                ("steps", C.c_int), ("moments", C.c_int), ("nvec", C.c_int),
                ("t_final", C.c_double), ("dt", C.c_double), ("seed", C.c_uint64),
                ("arrival", C.c_double)]


class _Result(C.Structure):
    _fields_ = [("status", C.c_int), ("slot", C.c_int), ("dim", C.c_int64),
                ("footprint", C.c_int64), ("peak", C.c_int64), ("e0", C.c_double),
                ("e_min", C.c_double), ("e_max", C.c_double), ("overlap_re", C.c_double),
                ("overlap_im", C.c_double), ("norm", C.c_double), ("arrival", C.c_double),
                ("admitted", C.c_double), ("finished", C.c_double), ("n_moments", C.c_int),
                ("error", C.c_char * 256)]


@dataclass
class Request:
    """One serving request.  kind: REQ_GROUND (FP64 Lanczos, `steps` steps),
    REQ_KPM (`steps` FP64 Lanczos steps for the bounds, then `moments` FP32
    Chebyshev moments over `nvec` random-phase vectors) or REQ_EVOLVE
    (exp(-iHt) psi0 up to t_final in steps of dt, Krylov dimension `steps`).
    sizes: lattice sizes (1, 2 or 3 of them, periodic).  arrival: seconds after
    the server started (< 0: on submission)."""
    kind: int
    sizes: tuple
    n_up: int
    n_dn: int
    t: float = 1.0
    u: float = 0.0
    steps: int = 150
    moments: int = 256
    nvec: int = 4
    t_final: float = 0.0
    dt: float = 0.05
    seed: int = 0
    arrival: float = -1.0

    def _c(self) -> _Request:
        r = _Request()
        r.kind, r.ndim = self.kind, len(self.sizes)
        for i in range(3):
            r.sizes[i] = int(self.sizes[i]) if i < len(self.sizes) else 1
        r.n_up, r.n_dn, r.t, r.u = self.n_up, self.n_dn, self.t, self.u
        r.steps, r.moments, r.nvec = self.steps, self.moments, self.nvec
        r.t_final, r.dt, r.seed, r.arrival = self.t_final, self.dt, self.seed, self.arrival
        return r

    def footprint(self) -> int:
        r = self._c()
        return int(lib().mb_request_footprint(C.byref(r)))


class Server:
    """Serving loop: `slots` concurrent requests, device memory from one arena
    of `arena_bytes`; requests are admitted as they arrive and as slots and
    memory free up.  lane_bytes > 0 reserves part of the arena for the Lanczos
    lane, where small ground-state requests join a persistent batch at step
    boundaries."""

    def __init__(self, slots: int = 8, arena_bytes: int = 1 << 34, lane_bytes: int = 0):
        h = C.c_void_p()
        _check(lib().mb_server_create(slots, int(arena_bytes), int(lane_bytes), C.byref(h)))
        self._h = h

    def close(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.mb_server_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def submit(self, req: Request) -> int:
        r = req._c()
        i = C.c_int64()
        _check(lib().mb_server_submit(self._h, C.byref(r), C.byref(i)))
        return i.value

    def wait(self, rid: int, moments_cap: int = 4096) -> dict:
        import numpy as np
        res = _Result()
        mu = np.zeros(moments_cap)
        _check(lib().mb_server_wait(self._h, rid, C.byref(res), mu.ctypes.data_as(C.c_void_p),
                                    moments_cap))
        out = {f: getattr(res, f) for f, _ in _Result._fields_ if f != "error"}
        out["error"] = res.error.decode()
        out["moments"] = mu[:res.n_moments].copy()
        return out

    def drain(self):
        _check(lib().mb_server_drain(self._h))

    def now(self) -> float:
        return float(lib().mb_server_now(self._h))


class HubbardModelMomentum:
    """HubbardModelMomentum(t, u, lattice_size) (hubbard_model_momentum.h): its
    hamiltonian() expression in momentum space."""

    def __init__(self, t: float, u: float, lattice_size):
        self.t, self.u = float(t), float(u)
        self.sizes = tuple(int(v) for v in lattice_size)

    def hamiltonian(self):
        from .expression import hubbard_momentum_hamiltonian
        return hubbard_momentum_hamiltonian(self.t, self.u, self.sizes)

    def current(self, Q, direction: int):
        """current(Q, direction) (hubbard_model_momentum.h:52-91): J_d(Q) as an
        expression, for compute_rectangular_matrix_elements between sectors K and
        K + Q."""
        from .expression import hubbard_momentum_current
        return hubbard_momentum_current(self.t, self.sizes, Q, direction)


from .expression import (FermionExpression, SparseMatrix, Spin, annihilation,  # noqa: E402
                         creation, density, density_density, hopping, one_body, two_body)
