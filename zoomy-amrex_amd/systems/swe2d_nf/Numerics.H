#pragma once
// Riemann solver for the shallow-water system (per-cell Manning variant: the
// flux is that of systems/swe2d) in the interface zoomy_core's
// AmrexNumerics emits (called at Source/make_rhs.H:129-130 and
// Source/ZoomyAmr.cpp:483): a nonconservative Rusanov scheme.
//
//   numerical_flux         = (F(qL).n + F(qR).n)/2 - (s/2) (qR - qL)   rows h, hu, hv
//                            s = max(|lam|(qL), |lam|(qR)); the bed row b carries
//                            neither flux nor dissipation, so b never evolves
//   numerical_fluctuations = [D+ | D-], the straight-path integral of the bed
//                            coupling g h db split evenly, D+ = D- = g (hL+hR)/4 (bR-bL) n
//   local_max_abs_eigenvalue = |u.n| + sqrt(g max(h, 0))
//
// With this flux the order-1 fixtures of the reference (tests/refs/*.npz) are
// reproduced to round-off.
#include "Model.H"

struct Numerics {
    static constexpr bool fluctuations_are_zero = false;
    // the fluctuations are proportional to the bed jump qR(0) - qL(0), so a
    // well-balanced face (equal reconstructed beds) has none
    static constexpr bool zamr_fluct_from_bed_jump = true;

    static AMREX_GPU_HOST_DEVICE AMREX_FORCE_INLINE amrex::SmallMatrix<amrex::Real, 1, 1> local_max_abs_eigenvalue(
        Model::VQ const& Q, Model::VA const& Qaux, Model::VP const& p, Model::VN const& n)
    {
        amrex::SmallMatrix<amrex::Real, 1, 1> res{};
        const amrex::Real un = zamr_math::axis_dot(Q(2, 0), Q(3, 0), n(0, 0), n(1, 0)) * Qaux(6, 0);
        res(0, 0) = std::fabs(un) + zamr_math::sqrt(p(0, 0) * amrex::max(Q(1, 0), amrex::Real(0.0)));
        return res;
    }

    static AMREX_GPU_HOST_DEVICE AMREX_FORCE_INLINE Model::VQ numerical_flux(
        Model::VQ const& qL, Model::VQ const& qR, Model::VA const& aL, Model::VA const& aR,
        Model::VP const& p, Model::VN const& n)
    {
        Model::VQ res{};
        const amrex::Real g = p(0, 0);
        const amrex::Real nx = n(0, 0), ny = n(1, 0);
        const amrex::Real qnL = zamr_math::axis_dot(qL(2, 0), qL(3, 0), nx, ny);
        const amrex::Real qnR = zamr_math::axis_dot(qR(2, 0), qR(3, 0), nx, ny);
        const amrex::Real unL = qnL * aL(6, 0);
        const amrex::Real unR = qnR * aR(6, 0);
        const amrex::Real phL = 0.5 * g * qL(1, 0) * qL(1, 0);
        const amrex::Real phR = 0.5 * g * qR(1, 0) * qR(1, 0);
        const amrex::Real sL = std::fabs(unL) + zamr_math::sqrt(g * amrex::max(qL(1, 0), amrex::Real(0.0)));
        const amrex::Real sR = std::fabs(unR) + zamr_math::sqrt(g * amrex::max(qR(1, 0), amrex::Real(0.0)));
        const amrex::Real hs = 0.5 * amrex::max(sL, sR);
        res(0, 0) = 0.0;
        res(1, 0) = 0.5 * (qnL + qnR) - hs * (qR(1, 0) - qL(1, 0));
        res(2, 0) = 0.5 * ((qL(2, 0) * unL + zamr_math::axis_scale(phL, nx)) + (qR(2, 0) * unR + zamr_math::axis_scale(phR, nx))) -
                    hs * (qR(2, 0) - qL(2, 0));
        res(3, 0) = 0.5 * ((qL(3, 0) * unL + zamr_math::axis_scale(phL, ny)) + (qR(3, 0) * unR + zamr_math::axis_scale(phR, ny))) -
                    hs * (qR(3, 0) - qL(3, 0));
        return res;
    }

    static AMREX_GPU_HOST_DEVICE AMREX_FORCE_INLINE amrex::SmallMatrix<amrex::Real, 8, 1> numerical_fluctuations(
        Model::VQ const& qL, Model::VQ const& qR, Model::VA const& aL, Model::VA const& aR,
        Model::VP const& p, Model::VN const& n)
    {
        amrex::SmallMatrix<amrex::Real, 8, 1> res{};
        const amrex::Real w = 0.25 * p(0, 0) * (qL(1, 0) + qR(1, 0)) * (qR(0, 0) - qL(0, 0));
        res(2, 0) = zamr_math::axis_scale(w, n(0, 0));
        res(3, 0) = zamr_math::axis_scale(w, n(1, 0));
        res(4 + 2, 0) = zamr_math::axis_scale(w, n(0, 0));
        res(4 + 3, 0) = zamr_math::axis_scale(w, n(1, 0));
        return res;
    }
};
