#pragma once
// Nonconservative Rusanov solver for the one-direction shallow-water system,
// in the interface zoomy_core's AmrexNumerics emits (Source/make_rhs.H:129-130,
// Source/ZoomyAmr.cpp:483).  Same scheme as the two-direction system: central
// flux minus (s/2)(qR-qL) on h and hu, no flux or dissipation on b, and the
// bed coupling split evenly into D+ and D-.
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
        const amrex::Real un = Q(2, 0) * n(0, 0) * Qaux(3, 0);
        res(0, 0) = std::fabs(un) + zamr_math::sqrt(p(0, 0) * amrex::max(Q(1, 0), amrex::Real(0.0)));
        return res;
    }

    static AMREX_GPU_HOST_DEVICE AMREX_FORCE_INLINE Model::VQ numerical_flux(
        Model::VQ const& qL, Model::VQ const& qR, Model::VA const& aL, Model::VA const& aR,
        Model::VP const& p, Model::VN const& n)
    {
        Model::VQ res{};
        const amrex::Real g = p(0, 0);
        const amrex::Real nx = n(0, 0);
        const amrex::Real qnL = qL(2, 0) * nx;
        const amrex::Real qnR = qR(2, 0) * nx;
        const amrex::Real unL = qnL * aL(3, 0);
        const amrex::Real unR = qnR * aR(3, 0);
        const amrex::Real phL = 0.5 * g * qL(1, 0) * qL(1, 0);
        const amrex::Real phR = 0.5 * g * qR(1, 0) * qR(1, 0);
        const amrex::Real sL = std::fabs(unL) + zamr_math::sqrt(g * amrex::max(qL(1, 0), amrex::Real(0.0)));
        const amrex::Real sR = std::fabs(unR) + zamr_math::sqrt(g * amrex::max(qR(1, 0), amrex::Real(0.0)));
        const amrex::Real hs = 0.5 * amrex::max(sL, sR);
        res(0, 0) = 0.0;
        res(1, 0) = 0.5 * (qnL + qnR) - hs * (qR(1, 0) - qL(1, 0));
        res(2, 0) = 0.5 * ((qL(2, 0) * unL + phL * nx) + (qR(2, 0) * unR + phR * nx)) - hs * (qR(2, 0) - qL(2, 0));
        return res;
    }

    static AMREX_GPU_HOST_DEVICE AMREX_FORCE_INLINE amrex::SmallMatrix<amrex::Real, 6, 1> numerical_fluctuations(
        Model::VQ const& qL, Model::VQ const& qR, Model::VA const& aL, Model::VA const& aR,
        Model::VP const& p, Model::VN const& n)
    {
        amrex::SmallMatrix<amrex::Real, 6, 1> res{};
        const amrex::Real w = 0.25 * p(0, 0) * (qL(1, 0) + qR(1, 0)) * (qR(0, 0) - qL(0, 0));
        res(2, 0) = w * n(0, 0);
        res(3 + 2, 0) = w * n(0, 0);
        return res;
    }
};
