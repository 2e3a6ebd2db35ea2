// Observables of a device-order state (see observables.hip).
#pragma once

#include <vector>

#include "krylov.hpp"
#include "mb_internal.hpp"

namespace mb {

// <x|G_updn(r)|x> / <x|x> for every displacement r (flattened rx + Lx ry + Lx Ly rz).
std::vector<double> opposite_spin_correlation(const Hubbard &h, const void *x, hipStream_t s);
// E = <x|H|x> / <x|x>, residual = |Hx - E x| / |x|.
void energy_residual(const HvOp &op, const void *x, double *energy, double *residual,
                     hipStream_t s);

}  // namespace mb
