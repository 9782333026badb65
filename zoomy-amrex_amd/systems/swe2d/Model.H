#pragma once
// Shallow-water system, two horizontal directions: SME(level=0) with
// Newtonian, Manning and stress-free closures, written out by hand in the
// exact interface that zoomy-amrex's AmrexSystemModelPrinter emits
// (zoomy_amrex/transformation/to_amrex.py:326-516): one `struct Model` of
// static host/device functions over amrex::SmallMatrix column vectors.
//
// State  Q    = [b, h, hu, hv]                      (Source/constants.H:5-7)
// Aux    Qaux = [a0 .. a5 (identity rows), hinv]     hinv in the LAST row
// Params p    = [g, ex, ey, ez, rho, wet_dry_eps, nu, n_m, lambda_s]
//               p(0) = g is read by make_rhs.H:287,395; p(5) = wet_dry_eps is
//               the only parameter the aux update reads (ZoomyAmr.cpp:836-840).
//
// Which formulas: the reference's golden fixtures (tests/refs/*.npz, order 1)
// are reproduced to round-off by
//   hinv  = sqrt(2) h / sqrt(h^4 + max(h^4, eps^4))      (Kurganov-Petrova)
//   flux  = [0, hu_n, hu u_n + g h^2 n_x / 2, hv u_n + g h^2 n_y / 2]
//   |lam| = |u_n| + sqrt(g max(h, 0))                    (guard_eigenvalue_powers)
// with u = hu * hinv.  Manning friction (n_m, default 0 as in the fixtures) is
// S_hu = -g n_m^2 |u| hu / max(h, eps)^(4/3), Jacobian written out with hinv
// held fixed (the printer emits dA/dQ == 0, to_amrex.py:478-480).
//
// Extra member `zamr_aux_is_local`: the live aux row is a per-cell function of
// (Q, p) and no operator reads the identity rows, so a device kernel may
// evaluate update_aux_variables on the fly instead of streaming Qaux.
// Extra member `zamr_source_rows`: the rows `source` can make nonzero (the
// global source solve carries only those).
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
    static constexpr int n_dof_q    = 4;
    static constexpr int n_dof_eq   = 4;
    static constexpr int n_dof_qaux = 7;
    static constexpr int n_parameters = 9;
    static constexpr int dimension  = 2;
    static constexpr int n_dof_gradQ = 8;
    static constexpr bool has_diffusion = false;
    static constexpr int n_boundary_tags = 5;
    static constexpr bool zamr_aux_is_local = true;
    static constexpr int zamr_source_rows[2] = {2, 3};  // Manning friction: momentum rows only
    static constexpr int zamr_n_aux_in = 0;          // aux rows read from storage (per-cell inputs)
    static constexpr int zamr_aux_in[1] = {0};
    static constexpr int zamr_idx_hinv = 6;
    // update_aux_variables(Q, 0) is zero but for hinv(Q(1), p(zamr_idx_eps)) in
    // row zamr_idx_hinv (a face's reconstructed states take their aux so)
    static constexpr bool zamr_aux_only_hinv = true;
    static constexpr int zamr_idx_eps = 5;
    static const std::vector<std::string> get_boundary_tags() { return { "East", "North", "South", "West", "wall" }; }
    static const std::vector<std::string> parameter_names() { return { "g", "ex", "ey", "ez", "rho", "wet_dry_eps", "nu", "n_m", "lambda_s" }; }
    static const std::vector<T> default_parameters() { return { 9.81, 0.0, 0.0, 1.0, 1.0, 0.0001, 0.0, 0.0, 0.0 }; }

    using VQ = amrex::SmallMatrix<amrex::Real, 4, 1>;
    using VA = amrex::SmallMatrix<amrex::Real, 7, 1>;
    using VP = amrex::SmallMatrix<amrex::Real, 9, 1>;
    using VN = amrex::SmallMatrix<amrex::Real, 2, 1>;
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

    static AMREX_GPU_HOST_DEVICE AMREX_FORCE_INLINE amrex::SmallMatrix<amrex::Real, 8, 1> flux(
        VQ const& Q, VA const& Qaux, VP const& p)
    {
        auto res = amrex::SmallMatrix<amrex::Real, 8, 1>{};
        const amrex::Real u = Q(2, 0) * Qaux(6, 0);
        const amrex::Real v = Q(3, 0) * Qaux(6, 0);
        const amrex::Real ph = 0.5 * p(0, 0) * Q(1, 0) * Q(1, 0);
        res(0, 0) = 0.0;
        res(1, 0) = Q(2, 0);
        res(2, 0) = Q(2, 0) * u + ph;
        res(3, 0) = Q(3, 0) * u;
        res(4, 0) = 0.0;
        res(5, 0) = Q(3, 0);
        res(6, 0) = Q(2, 0) * v;
        res(7, 0) = Q(3, 0) * v + ph;
        return res;
    }

    static AMREX_GPU_HOST_DEVICE AMREX_FORCE_INLINE amrex::SmallMatrix<amrex::Real, 8, 1> dflux(
        VQ const& Q, VA const& Qaux, VP const& p)
    {
        return amrex::SmallMatrix<amrex::Real, 8, 1>{};
    }

    static AMREX_GPU_HOST_DEVICE AMREX_FORCE_INLINE amrex::SmallMatrix<amrex::Real, 8, 1> hydrostatic_pressure(
        VQ const& Q, VA const& Qaux, VP const& p)
    {
        return amrex::SmallMatrix<amrex::Real, 8, 1>{};
    }

    // Dir-major flat (dim, n_eq, n_state): the bed couples into momentum as
    // g h db/dx (row hu, column b) and g h db/dy (row hv, column b).
    static AMREX_GPU_HOST_DEVICE AMREX_FORCE_INLINE amrex::SmallMatrix<amrex::Real, 32, 1> nonconservative_matrix(
        VQ const& Q, VA const& Qaux, VP const& p)
    {
        auto res = amrex::SmallMatrix<amrex::Real, 32, 1>{};
        res(2 * 4 + 0, 0) = p(0, 0) * Q(1, 0);
        res(16 + 3 * 4 + 0, 0) = p(0, 0) * Q(1, 0);
        return res;
    }

    static AMREX_GPU_HOST_DEVICE AMREX_FORCE_INLINE amrex::SmallMatrix<amrex::Real, 32, 1> quasilinear_matrix(
        VQ const& Q, VA const& Qaux, VP const& p)
    {
        auto res = amrex::SmallMatrix<amrex::Real, 32, 1>{};
        const amrex::Real g = p(0, 0);
        const amrex::Real h = Q(1, 0);
        const amrex::Real u = Q(2, 0) * Qaux(6, 0);
        const amrex::Real v = Q(3, 0) * Qaux(6, 0);
        // x direction
        res(1 * 4 + 2, 0) = 1.0;
        res(2 * 4 + 0, 0) = g * h;
        res(2 * 4 + 1, 0) = g * h - u * u;
        res(2 * 4 + 2, 0) = 2.0 * u;
        res(3 * 4 + 1, 0) = -u * v;
        res(3 * 4 + 2, 0) = v;
        res(3 * 4 + 3, 0) = u;
        // y direction
        res(16 + 1 * 4 + 3, 0) = 1.0;
        res(16 + 2 * 4 + 1, 0) = -u * v;
        res(16 + 2 * 4 + 2, 0) = v;
        res(16 + 2 * 4 + 3, 0) = u;
        res(16 + 3 * 4 + 0, 0) = g * h;
        res(16 + 3 * 4 + 1, 0) = g * h - v * v;
        res(16 + 3 * 4 + 3, 0) = 2.0 * v;
        return res;
    }

    static AMREX_GPU_HOST_DEVICE AMREX_FORCE_INLINE amrex::SmallMatrix<amrex::Real, 8, 1> diffusive_flux(
        VQ const& Q, VA const& Qaux, amrex::SmallMatrix<amrex::Real, 8, 1> const& gradQ, VP const& p)
    {
        return amrex::SmallMatrix<amrex::Real, 8, 1>{};
    }

    static AMREX_GPU_HOST_DEVICE AMREX_FORCE_INLINE VQ eigenvalues(
        VQ const& Q, VA const& Qaux, VP const& p, VN const& n)
    {
        auto res = VQ{};
        const amrex::Real un = (Q(2, 0) * n(0, 0) + Q(3, 0) * n(1, 0)) * Qaux(6, 0);
        const amrex::Real c = zamr_math::sqrt(p(0, 0) * amrex::max(Q(1, 0), amrex::Real(0.0)));
        res(0, 0) = 0.0;
        res(1, 0) = un - c;
        res(2, 0) = un;
        res(3, 0) = un + c;
        return res;
    }

    static AMREX_GPU_HOST_DEVICE AMREX_FORCE_INLINE amrex::SmallMatrix<amrex::Real, 16, 1> left_eigenvectors(
        VQ const& Q, VA const& Qaux, VP const& p, VN const& n)
    {
        return amrex::SmallMatrix<amrex::Real, 16, 1>{};
    }

    static AMREX_GPU_HOST_DEVICE AMREX_FORCE_INLINE amrex::SmallMatrix<amrex::Real, 16, 1> right_eigenvectors(
        VQ const& Q, VA const& Qaux, VP const& p, VN const& n)
    {
        return amrex::SmallMatrix<amrex::Real, 16, 1>{};
    }

    // Manning bottom friction, S_hu = -k hu, S_hv = -k hv with
    // k = g n_m^2 |u| / max(h, eps)^(4/3).
    static AMREX_GPU_HOST_DEVICE AMREX_FORCE_INLINE VQ source(
        VQ const& Q, VA const& Qaux, VP const& p, amrex::Real const time, VX const& X)
    {
        auto res = VQ{};
        const amrex::Real u = Q(2, 0) * Qaux(6, 0);
        const amrex::Real v = Q(3, 0) * Qaux(6, 0);
        const amrex::Real speed = zamr_math::sqrt(u * u + v * v);
        const amrex::Real r = zamr_rcbrt(amrex::max(Q(1, 0), p(5, 0)));
        // k = c |u| with c = g n^2 r^4, as the Jacobians form it (one evaluation)
        const amrex::Real c = p(0, 0) * p(7, 0) * p(7, 0) * amrex::Math::powi<4>(r);
        const amrex::Real k = c * speed;
        res(0, 0) = 0.0;
        res(1, 0) = 0.0;
        res(2, 0) = -k * Q(2, 0);
        res(3, 0) = -k * Q(3, 0);
        return res;
    }

    // Row-major (n_eq, n_state).  At zero speed the |u| derivative takes the
    // subgradient 0; the Max(h, eps) derivative is Heaviside(h - eps, 1/2).
    static AMREX_GPU_HOST_DEVICE AMREX_FORCE_INLINE amrex::SmallMatrix<amrex::Real, 16, 1> source_jacobian_wrt_variables(
        VQ const& Q, VA const& Qaux, VP const& p)
    {
        auto res = amrex::SmallMatrix<amrex::Real, 16, 1>{};
        const amrex::Real hinv = Qaux(6, 0);
        const amrex::Real u = Q(2, 0) * hinv;
        const amrex::Real v = Q(3, 0) * hinv;
        const amrex::Real speed = zamr_math::sqrt(u * u + v * v);
        const amrex::Real hf = amrex::max(Q(1, 0), p(5, 0));
        const amrex::Real r = zamr_rcbrt(hf);
        const amrex::Real c = p(0, 0) * p(7, 0) * p(7, 0) * amrex::Math::powi<4>(r);
        const amrex::Real k = c * speed;
        const amrex::Real dh = Q(1, 0) - p(5, 0);
        const amrex::Real heav = (dh > 0.0) ? amrex::Real(1.0) : ((dh < 0.0) ? amrex::Real(0.0) : amrex::Real(0.5));
        // 1/hf = r^3 and hinv/speed = hinv/sqrt(|u|^2): no division on the device
        const amrex::Real dk_dh = -(4.0 / 3.0) * k * (r * r * r) * heav;
        const amrex::Real ws = (speed > 0.0) ? zamr_math::div_sqrt(hinv, u * u + v * v) : amrex::Real(0.0);
        const amrex::Real dk_dhu = c * u * ws;
        const amrex::Real dk_dhv = c * v * ws;
        res(2 * 4 + 1, 0) = -Q(2, 0) * dk_dh;
        res(2 * 4 + 2, 0) = -k - Q(2, 0) * dk_dhu;
        res(2 * 4 + 3, 0) = -Q(2, 0) * dk_dhv;
        res(3 * 4 + 1, 0) = -Q(3, 0) * dk_dh;
        res(3 * 4 + 2, 0) = -Q(3, 0) * dk_dhu;
        res(3 * 4 + 3, 0) = -k - Q(3, 0) * dk_dhv;
        return res;
    }

    static AMREX_GPU_HOST_DEVICE AMREX_FORCE_INLINE amrex::SmallMatrix<amrex::Real, 28, 1> source_jacobian_wrt_aux_variables(
        VQ const& Q, VA const& Qaux, VP const& p)
    {
        auto res = amrex::SmallMatrix<amrex::Real, 28, 1>{};
        const amrex::Real m = zamr_math::sqrt(Q(2, 0) * Q(2, 0) + Q(3, 0) * Q(3, 0));
        const amrex::Real r = zamr_rcbrt(amrex::max(Q(1, 0), p(5, 0)));
        const amrex::Real c = p(0, 0) * p(7, 0) * p(7, 0) * amrex::Math::powi<4>(r);
        res(2 * 7 + 6, 0) = -c * m * Q(2, 0);
        res(3 * 7 + 6, 0) = -c * m * Q(3, 0);
        return res;
    }

    static AMREX_GPU_HOST_DEVICE AMREX_FORCE_INLINE amrex::SmallMatrix<amrex::Real, 6, 1> project_2d_to_3d(
        VX const& X, VQ const& Q, VA const& Qaux, VP const& p)
    {
        return amrex::SmallMatrix<amrex::Real, 6, 1>{};
    }

    static AMREX_GPU_HOST_DEVICE AMREX_FORCE_INLINE VQ project_3d_to_2d(
        VQ const& Q, VA const& Qaux, VP const& p)
    {
        return VQ{};
    }

    static AMREX_GPU_HOST_DEVICE AMREX_FORCE_INLINE VQ residual(
        amrex::Real const time, VX const& X, amrex::Real const dX, VQ const& Q, VA const& Qaux, VP const& p)
    {
        return VQ{};
    }

    static AMREX_GPU_HOST_DEVICE AMREX_FORCE_INLINE VQ interpolate(
        VQ const& Q, VA const& Qaux, VP const& p)
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

    static AMREX_GPU_HOST_DEVICE AMREX_FORCE_INLINE VQ update_variables(
        VQ const& Q, VA const& Qaux, VP const& p)
    {
        return Q;
    }

    static AMREX_GPU_HOST_DEVICE AMREX_FORCE_INLINE VA update_aux_variables(
        VQ const& Q, VA const& Qaux, VP const& p, amrex::Real const time, VX const& X)
    {
        VA res = Qaux;
        res(6, 0) = zamr_hinv(Q(1, 0), p(5, 0));
        return res;
    }

    static AMREX_GPU_HOST_DEVICE AMREX_FORCE_INLINE amrex::SmallMatrix<amrex::Real, 16, 1> update_variables_jacobian_wrt_variables(
        VQ const& Q, VA const& Qaux, VP const& p)
    {
        auto res = amrex::SmallMatrix<amrex::Real, 16, 1>{};
        for (int r = 0; r < 4; ++r) res(r * 4 + r, 0) = 1.0;
        return res;
    }

    static AMREX_GPU_HOST_DEVICE AMREX_FORCE_INLINE amrex::SmallMatrix<amrex::Real, 28, 1> update_aux_variables_jacobian_wrt_variables(
        VQ const& Q, VA const& Qaux, VP const& p, amrex::Real const time, VX const& X)
    {
        return amrex::SmallMatrix<amrex::Real, 28, 1>{};
    }

    // Tags sorted: East, North, South, West (extrapolation) and wall
    // (reflective: normal momentum flips sign).
    static AMREX_GPU_HOST_DEVICE AMREX_FORCE_INLINE VQ boundary_conditions(
        const int bc_idx, VQ const& Q, VA const& Qaux, VN const& n, VX const& X,
        amrex::Real const time, amrex::Real const dX)
    {
        VQ res = Q;
        if (bc_idx == 4) {
            const amrex::Real qn = Q(2, 0) * n(0, 0) + Q(3, 0) * n(1, 0);
            res(2, 0) = Q(2, 0) - 2.0 * qn * n(0, 0);
            res(3, 0) = Q(3, 0) - 2.0 * qn * n(1, 0);
        }
        return res;
    }

    static AMREX_GPU_HOST_DEVICE AMREX_FORCE_INLINE VA aux_boundary_conditions(
        const int bc_idx, VQ const& Q, VA const& Qaux, VN const& n, VX const& X,
        amrex::Real const time, amrex::Real const dX)
    {
        return Qaux;
    }

    static constexpr int n_e2s = 4;
    static constexpr int e2s[4] = { 0, 1, 2, 3 };
    static constexpr int n_deriv_aux = 0;
    static constexpr int deriv_aux[1][4] = {{0,0,0,0}};
    static constexpr int n_input_aux = 0;
    static constexpr int input_aux[1][4] = {{0,0,0,0}};
};
