#pragma once
// Viscous shallow-water system, two horizontal directions: the swe2d system
// (../swe2d/Model.H: state [b, h, hu, hv], hinv in the last aux row, Rusanov
// flux, Manning friction) with a diffusive flux, in the printer's interface
// (zoomy_amrex/transformation/to_amrex.py:326-516).  has_diffusion = true
// switches on make_rhs.H's add_diffusive_flux (:166-229, gated by
// Model::has_diffusion, to_amrex.py:109-129,356).
//
//   diffusive_flux(Q, Qaux, gradQ, p)(v*dim + d) = nu * dQ_v/dx_d
//       for the rows h, hu, hv (v = 1, 2, 3); the bed row b does not diffuse.
//
// nu is the parameter `nu` (p(6), default 0: with nu = 0 the system is
// swe2d).  gradQ is laid out [dQ0_d0, dQ0_d1, dQ1_d0, ...] (n_dof_gradQ =
// n_dof_q * dimension), as make_rhs.H:196-200 fills it.
#define Model ZamrSwe2dInviscid
#include "../swe2d/Model.H"
#undef Model

struct Model : ZamrSwe2dInviscid {
    static constexpr bool has_diffusion = true;

    static AMREX_GPU_HOST_DEVICE AMREX_FORCE_INLINE amrex::SmallMatrix<amrex::Real, 8, 1> diffusive_flux(
        VQ const& Q, VA const& Qaux, amrex::SmallMatrix<amrex::Real, 8, 1> const& gradQ, VP const& p)
    {
        auto res = amrex::SmallMatrix<amrex::Real, 8, 1>{};
        const amrex::Real nu = p(6, 0);
        for (int v = 1; v < 4; ++v)
            for (int d = 0; d < 2; ++d) res(v * 2 + d, 0) = nu * gradQ(v * 2 + d, 0);
        return res;
    }
};
