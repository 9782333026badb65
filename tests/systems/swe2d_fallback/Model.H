#pragma once
// TEST SYSTEM (built for the tests only): the swe2d system with a linear
// momentum source whose point-implicit matrix I - dt J can be made singular,
// to drive apply_implicit_source's explicit fallback (make_rhs.H:469-508)
// through the whole stage.
//
//   S = [0, 0, -kappa hu / h, -kappa hv / h]     (true divisions: NaN at h = 0)
//
// With kappa = -1/dt the momentum rows of I - dt J vanish where h = 1, the
// pivot search finds 0 and the stage applies q + dt S; where h = 0 the
// Jacobian and S are NaN, so the explicit value is not finite either and the
// cell keeps q_old; elsewhere the solve succeeds.  kappa is the parameter
// `kappa` (p(7)).
#define Model ZamrSwe2dBase
#include "../../../zoomy-amrex_amd/systems/swe2d/Model.H"
#undef Model

struct Model : ZamrSwe2dBase {
    static const std::vector<std::string> parameter_names() { return { "g", "ex", "ey", "ez", "rho", "wet_dry_eps", "nu", "kappa", "lambda_s" }; }

    static AMREX_GPU_HOST_DEVICE AMREX_FORCE_INLINE VQ source(
        VQ const& Q, VA const& Qaux, VP const& p, amrex::Real const time, VX const& X)
    {
        auto res = VQ{};
        res(2, 0) = -p(7, 0) * Q(2, 0) / Q(1, 0);
        res(3, 0) = -p(7, 0) * Q(3, 0) / Q(1, 0);
        return res;
    }

    static AMREX_GPU_HOST_DEVICE AMREX_FORCE_INLINE amrex::SmallMatrix<amrex::Real, 16, 1> source_jacobian_wrt_variables(
        VQ const& Q, VA const& Qaux, VP const& p)
    {
        auto res = amrex::SmallMatrix<amrex::Real, 16, 1>{};
        const amrex::Real k = p(7, 0), h = Q(1, 0);
        res(2 * 4 + 1, 0) = k * Q(2, 0) / (h * h);
        res(2 * 4 + 2, 0) = -k / h;
        res(3 * 4 + 1, 0) = k * Q(3, 0) / (h * h);
        res(3 * 4 + 3, 0) = -k / h;
        return res;
    }

    static AMREX_GPU_HOST_DEVICE AMREX_FORCE_INLINE amrex::SmallMatrix<amrex::Real, 28, 1> source_jacobian_wrt_aux_variables(
        VQ const& Q, VA const& Qaux, VP const& p)
    {
        return amrex::SmallMatrix<amrex::Real, 28, 1>{};
    }
};
