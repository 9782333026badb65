#pragma once
// Opaque kernel entry points that the printer routes calls to
// (to_amrex.py:56-63, 99-107).  The shallow-water and shallow-moment systems
// built here have analytic wave speeds and never call them; a model that does
// gets the reference's refusal contract: every entry is +inf
// (Source/UserFunctions.H:7-28), which the Rusanov speed and the CFL
// reduction turn into a visible non-finite result instead of a silent one.
#include <limits>
#include "AMReX_SmallMatrix.H"
#include "ZamrMath.H"

namespace amrex_user {
template <class... Args>
AMREX_GPU_HOST_DEVICE AMREX_FORCE_INLINE amrex::Real eigenvalues(Args...) {
    return std::numeric_limits<amrex::Real>::infinity();
}
template <class... Args>
AMREX_GPU_HOST_DEVICE AMREX_FORCE_INLINE amrex::Real eigensystem(Args...) {
    return std::numeric_limits<amrex::Real>::infinity();
}
template <class... Args>
AMREX_GPU_HOST_DEVICE AMREX_FORCE_INLINE amrex::Real solve(Args...) {
    return std::numeric_limits<amrex::Real>::infinity();
}
// Source/UserFunctions.H:223 — the derivative helper is a stub returning 0.
template <class... Args>
AMREX_GPU_HOST_DEVICE AMREX_FORCE_INLINE amrex::Real compute_derivative(Args...) {
    return amrex::Real(0);
}
}  // namespace amrex_user
