#pragma once
// Shallow-water system, one horizontal direction: SME(level=0, dimension=2)
// with Newtonian, Manning and stress-free closures, in the interface
// zoomy-amrex's AmrexSystemModelPrinter emits
// (zoomy_amrex/transformation/to_amrex.py:326-516).  This is the model behind
// the reference's 1-D golden fixtures (tests/refs/{stoker_wet,ritter_dry,
// wb_lake,parallel_2rank,parity_numpy_n*}.npz), which it reproduces to round-off.
//
// State  Q    = [b, h, hu]
// Aux    Qaux = [a0, a1, a2 (identity rows), hinv]
// Params p    = [g, ex, ey, ez, rho, wet_dry_eps, nu, n_m, lambda_s]
// Formulas as in the two-direction system: KP-desingularised hinv, Rusanov
// flux [0, hu, hu u + g h^2/2], |lam| = |u| + sqrt(g max(h,0)), Manning friction.
#include <AMReX_Array4.H>
#include <AMReX_Vector.H>
#include <AMReX_SmallMatrix.H>
#include <vector>
#include <string>
#include <algorithm>
#include <cstring>

#if defined(__CUDACC__) || defined(__HIP__)
#define PORTABLE_FN __host__ __device__
#else
#define PORTABLE_FN
#endif

#include "UserFunctions.H"

struct Model {
    using T = amrex::Real;
    static constexpr int n_dof_q    = 3;
    static constexpr int n_dof_eq   = 3;
    static constexpr int n_dof_qaux = 4;
    static constexpr int n_parameters = 9;
    static constexpr int dimension  = 1;
    static constexpr int n_dof_gradQ = 3;
    static constexpr bool has_diffusion = false;
    static constexpr int n_boundary_tags = 3;
    static constexpr bool zamr_aux_is_local = true;
    static constexpr int zamr_n_aux_in = 0;          // aux rows read from storage (per-cell inputs)
    static constexpr int zamr_aux_in[1] = {0};
    static constexpr int zamr_idx_hinv = 3;
    // update_aux_variables(Q, 0) is zero but for hinv(Q(1), p(zamr_idx_eps)) in
    // row zamr_idx_hinv (a face's reconstructed states take their aux so)
    static constexpr bool zamr_aux_only_hinv = true;
    static constexpr int zamr_idx_eps = 5;
    static const std::vector<std::string> get_boundary_tags() { return { "East", "West", "wall" }; }
    static const std::vector<std::string> parameter_names() { return { "g", "ex", "ey", "ez", "rho", "wet_dry_eps", "nu", "n_m", "lambda_s" }; }
    static const std::vector<T> default_parameters() { return { 9.81, 0.0, 0.0, 1.0, 1.0, 0.0001, 0.0, 0.0, 0.0 }; }

    using VQ = amrex::SmallMatrix<amrex::Real, 3, 1>;
    using VA = amrex::SmallMatrix<amrex::Real, 4, 1>;
    using VP = amrex::SmallMatrix<amrex::Real, 9, 1>;
    using VN = amrex::SmallMatrix<amrex::Real, 1, 1>;
    using VX = amrex::SmallMatrix<amrex::Real, 3, 1>;
    // x^(-1/3) for x > 0 (zamr_math::rcbrt: the same steps on host and in
    // the bitwise device build, so the two agree bit for bit)
    static AMREX_GPU_HOST_DEVICE AMREX_FORCE_INLINE amrex::Real zamr_rcbrt(amrex::Real x)
    {
        return zamr_math::rcbrt(x);
    }

    static AMREX_GPU_HOST_DEVICE AMREX_FORCE_INLINE amrex::Real zamr_hinv(amrex::Real h, amrex::Real eps)
    {
#if defined(__HIP_DEVICE_COMPILE__) && !defined(ZAMR_BITWISE)
        // h >= eps gives h^4 >= eps^4 (rounding is monotone), where the
        // formula is 1/h: such a cell takes the reciprocal (one v_rcp and four
        // FMAs instead of a v_rsq and eleven).  Decided per cell, so a cell's
        // value never depends on the cells sharing its wave; a wave with no
        // dry cell skips the formula.
        if (h >= eps) return zamr_math::recip(h);
#endif
        const amrex::Real h4 = amrex::Math::powi<4>(h);
        const amrex::Real e4 = amrex::Math::powi<4>(eps);
        // h4 is +0, positive or NaN, so fmax is max(h4, e4) here in every case
        // (a NaN h gives a NaN hinv either way) in one device instruction
        return zamr_math::div_sqrt(1.4142135623730951 * h, h4 + std::fmax(h4, e4));
    }

    static AMREX_GPU_HOST_DEVICE AMREX_FORCE_INLINE VQ flux(
        VQ const& Q, VA const& Qaux, VP const& p)
    {
        auto res = VQ{};
        const amrex::Real u = Q(2, 0) * Qaux(3, 0);
        res(0, 0) = 0.0;
        res(1, 0) = Q(2, 0);
        res(2, 0) = Q(2, 0) * u + 0.5 * p(0, 0) * Q(1, 0) * Q(1, 0);
        return res;
    }

    static AMREX_GPU_HOST_DEVICE AMREX_FORCE_INLINE VQ dflux(VQ const& Q, VA const& Qaux, VP const& p)
    {
        return VQ{};
    }

    static AMREX_GPU_HOST_DEVICE AMREX_FORCE_INLINE VQ hydrostatic_pressure(VQ const& Q, VA const& Qaux, VP const& p)
    {
        return VQ{};
    }

    static AMREX_GPU_HOST_DEVICE AMREX_FORCE_INLINE amrex::SmallMatrix<amrex::Real, 9, 1> nonconservative_matrix(
        VQ const& Q, VA const& Qaux, VP const& p)
    {
        auto res = amrex::SmallMatrix<amrex::Real, 9, 1>{};
        res(2 * 3 + 0, 0) = p(0, 0) * Q(1, 0);
        return res;
    }

    static AMREX_GPU_HOST_DEVICE AMREX_FORCE_INLINE amrex::SmallMatrix<amrex::Real, 9, 1> quasilinear_matrix(
        VQ const& Q, VA const& Qaux, VP const& p)
    {
        auto res = amrex::SmallMatrix<amrex::Real, 9, 1>{};
        const amrex::Real u = Q(2, 0) * Qaux(3, 0);
        res(1 * 3 + 2, 0) = 1.0;
        res(2 * 3 + 0, 0) = p(0, 0) * Q(1, 0);
        res(2 * 3 + 1, 0) = p(0, 0) * Q(1, 0) - u * u;
        res(2 * 3 + 2, 0) = 2.0 * u;
        return res;
    }

    static AMREX_GPU_HOST_DEVICE AMREX_FORCE_INLINE VQ diffusive_flux(
        VQ const& Q, VA const& Qaux, amrex::SmallMatrix<amrex::Real, 3, 1> const& gradQ, VP const& p)
    {
        return VQ{};
    }

    static AMREX_GPU_HOST_DEVICE AMREX_FORCE_INLINE VQ eigenvalues(
        VQ const& Q, VA const& Qaux, VP const& p, VN const& n)
    {
        auto res = VQ{};
        const amrex::Real un = Q(2, 0) * n(0, 0) * Qaux(3, 0);
        const amrex::Real c = zamr_math::sqrt(p(0, 0) * amrex::max(Q(1, 0), amrex::Real(0.0)));
        res(0, 0) = 0.0;
        res(1, 0) = un - c;
        res(2, 0) = un + c;
        return res;
    }

    static AMREX_GPU_HOST_DEVICE AMREX_FORCE_INLINE amrex::SmallMatrix<amrex::Real, 9, 1> left_eigenvectors(
        VQ const& Q, VA const& Qaux, VP const& p, VN const& n)
    {
        return amrex::SmallMatrix<amrex::Real, 9, 1>{};
    }

    static AMREX_GPU_HOST_DEVICE AMREX_FORCE_INLINE amrex::SmallMatrix<amrex::Real, 9, 1> right_eigenvectors(
        VQ const& Q, VA const& Qaux, VP const& p, VN const& n)
    {
        return amrex::SmallMatrix<amrex::Real, 9, 1>{};
    }

    static AMREX_GPU_HOST_DEVICE AMREX_FORCE_INLINE VQ source(
        VQ const& Q, VA const& Qaux, VP const& p, amrex::Real const time, VX const& X)
    {
        auto res = VQ{};
        const amrex::Real u = Q(2, 0) * Qaux(3, 0);
        const amrex::Real speed = zamr_math::sqrt(u * u);
        const amrex::Real r = zamr_rcbrt(amrex::max(Q(1, 0), p(5, 0)));
        const amrex::Real k = p(0, 0) * p(7, 0) * p(7, 0) * speed * amrex::Math::powi<4>(r);
        res(0, 0) = 0.0;
        res(1, 0) = 0.0;
        res(2, 0) = -k * Q(2, 0);
        return res;
    }

    static AMREX_GPU_HOST_DEVICE AMREX_FORCE_INLINE amrex::SmallMatrix<amrex::Real, 9, 1> source_jacobian_wrt_variables(
        VQ const& Q, VA const& Qaux, VP const& p)
    {
        auto res = amrex::SmallMatrix<amrex::Real, 9, 1>{};
        const amrex::Real hinv = Qaux(3, 0);
        const amrex::Real u = Q(2, 0) * hinv;
        const amrex::Real speed = zamr_math::sqrt(u * u);
        const amrex::Real hf = amrex::max(Q(1, 0), p(5, 0));
        const amrex::Real r = zamr_rcbrt(hf);
        const amrex::Real c = p(0, 0) * p(7, 0) * p(7, 0) * amrex::Math::powi<4>(r);
        const amrex::Real k = c * speed;
        const amrex::Real dh = Q(1, 0) - p(5, 0);
        const amrex::Real heav = (dh > 0.0) ? amrex::Real(1.0) : ((dh < 0.0) ? amrex::Real(0.0) : amrex::Real(0.5));
        // 1/hf = r^3 and hinv/speed = hinv/sqrt(|u|^2): no division on the device
        const amrex::Real dk_dh = -(4.0 / 3.0) * k * (r * r * r) * heav;
        const amrex::Real ws = (speed > 0.0) ? zamr_math::div_sqrt(hinv, u * u) : amrex::Real(0.0);
        const amrex::Real dk_dhu = c * u * ws;
        res(2 * 3 + 1, 0) = -Q(2, 0) * dk_dh;
        res(2 * 3 + 2, 0) = -k - Q(2, 0) * dk_dhu;
        return res;
    }

    static AMREX_GPU_HOST_DEVICE AMREX_FORCE_INLINE amrex::SmallMatrix<amrex::Real, 12, 1> source_jacobian_wrt_aux_variables(
        VQ const& Q, VA const& Qaux, VP const& p)
    {
        auto res = amrex::SmallMatrix<amrex::Real, 12, 1>{};
        const amrex::Real m = zamr_math::sqrt(Q(2, 0) * Q(2, 0));
        const amrex::Real r = zamr_rcbrt(amrex::max(Q(1, 0), p(5, 0)));
        const amrex::Real c = p(0, 0) * p(7, 0) * p(7, 0) * amrex::Math::powi<4>(r);
        res(2 * 4 + 3, 0) = -c * m * Q(2, 0);
        return res;
    }

    static AMREX_GPU_HOST_DEVICE AMREX_FORCE_INLINE amrex::SmallMatrix<amrex::Real, 6, 1> project_2d_to_3d(
        VX const& X, VQ const& Q, VA const& Qaux, VP const& p)
    {
        return amrex::SmallMatrix<amrex::Real, 6, 1>{};
    }

    static AMREX_GPU_HOST_DEVICE AMREX_FORCE_INLINE VQ project_3d_to_2d(VQ const& Q, VA const& Qaux, VP const& p)
    {
        return VQ{};
    }

    static AMREX_GPU_HOST_DEVICE AMREX_FORCE_INLINE VQ residual(
        amrex::Real const time, VX const& X, amrex::Real const dX, VQ const& Q, VA const& Qaux, VP const& p)
    {
        return VQ{};
    }

    static AMREX_GPU_HOST_DEVICE AMREX_FORCE_INLINE VQ interpolate(VQ const& Q, VA const& Qaux, VP const& p)
    {
        return Q;
    }

    static AMREX_GPU_HOST_DEVICE AMREX_FORCE_INLINE VQ initial_condition(VX const& X, VP const& p)
    {
        return VQ{};
    }

    static AMREX_GPU_HOST_DEVICE AMREX_FORCE_INLINE VA initial_aux_condition(VX const& X, VP const& p)
    {
        return VA{};
    }

    static AMREX_GPU_HOST_DEVICE AMREX_FORCE_INLINE VQ update_variables(VQ const& Q, VA const& Qaux, VP const& p)
    {
        return Q;
    }

    static AMREX_GPU_HOST_DEVICE AMREX_FORCE_INLINE VA update_aux_variables(
        VQ const& Q, VA const& Qaux, VP const& p, amrex::Real const time, VX const& X)
    {
        VA res = Qaux;
        res(3, 0) = zamr_hinv(Q(1, 0), p(5, 0));
        return res;
    }

    static AMREX_GPU_HOST_DEVICE AMREX_FORCE_INLINE amrex::SmallMatrix<amrex::Real, 9, 1> update_variables_jacobian_wrt_variables(
        VQ const& Q, VA const& Qaux, VP const& p)
    {
        auto res = amrex::SmallMatrix<amrex::Real, 9, 1>{};
        for (int r = 0; r < 3; ++r) res(r * 3 + r, 0) = 1.0;
        return res;
    }

    static AMREX_GPU_HOST_DEVICE AMREX_FORCE_INLINE amrex::SmallMatrix<amrex::Real, 12, 1> update_aux_variables_jacobian_wrt_variables(
        VQ const& Q, VA const& Qaux, VP const& p, amrex::Real const time, VX const& X)
    {
        return amrex::SmallMatrix<amrex::Real, 12, 1>{};
    }

    // Tags sorted: East, West (extrapolation) and wall (reflective).
    static AMREX_GPU_HOST_DEVICE AMREX_FORCE_INLINE VQ boundary_conditions(
        const int bc_idx, VQ const& Q, VA const& Qaux, VN const& n, VX const& X,
        amrex::Real const time, amrex::Real const dX)
    {
        VQ res = Q;
        if (bc_idx == 2) {
            const amrex::Real qn = Q(2, 0) * n(0, 0);
            res(2, 0) = Q(2, 0) - 2.0 * qn * n(0, 0);
        }
        return res;
    }

    static AMREX_GPU_HOST_DEVICE AMREX_FORCE_INLINE VA aux_boundary_conditions(
        const int bc_idx, VQ const& Q, VA const& Qaux, VN const& n, VX const& X,
        amrex::Real const time, amrex::Real const dX)
    {
        return Qaux;
    }

    static constexpr int n_e2s = 3;
    static constexpr int e2s[3] = { 0, 1, 2 };
    static constexpr int n_deriv_aux = 0;
    static constexpr int deriv_aux[1][4] = {{0,0,0,0}};
    static constexpr int n_input_aux = 0;
    static constexpr int input_aux[1][4] = {{0,0,0,0}};
};
