#pragma once
// Shallow moment equations, level 2, two horizontal directions (the
// shallow-moment system of Kowalski & Torrilhon 2019, "Moment approximations
// and model cascades for shallow flow", extended to 2-D), written out by hand
// in the interface zoomy-amrex's AmrexSystemModelPrinter emits
// (zoomy_amrex/transformation/to_amrex.py:326-516).  The reference generates
// this model from zoomy_core, which is not available: the formulas below are
// the published level-2 system, not a copy of generated code.
//
// Velocity profile  u(zeta) = u_m + sum_j alpha_j phi_j(zeta),
//                   v(zeta) = v_m + sum_j beta_j  phi_j(zeta),   j = 1, 2,
// phi_j the shifted Legendre polynomials on [0, 1] with phi_j(0) = 1 at the bed:
//   phi_1 = 1 - 2 zeta,  phi_2 = 1 - 6 zeta + 6 zeta^2,  int phi_i phi_j = delta_ij / (2j+1).
//
// State  Q    = [b, h, h u_m, h alpha_1, h alpha_2, h v_m, h beta_1, h beta_2]
//               (b at row 0 as every reference layout, Source/constants.H:5-7)
// Aux    Qaux = [a0 .. a5 (identity rows), hinv]     hinv in the LAST row
// Params p    = [g, ex, ey, ez, rho, wet_dry_eps, nu, n_m, lambda_s]
//
// Conservative fluxes (x; y by symmetry), c_j = 1/(2j+1):
//   h:        h u_m
//   h u_m:    h (u_m^2 + sum c_j alpha_j^2) + g h^2/2
//   h alpha_i: h (2 u_m alpha_i + sum_jk A_ijk alpha_j alpha_k)
//   h v_m:    h (u_m v_m + sum c_j alpha_j beta_j)
//   h beta_i: h (u_m beta_i + v_m alpha_i + sum_jk A_ijk beta_j alpha_k)
// Non-conservative products, with D_j = d_x(h alpha_j) + d_y(h beta_j):
//   h alpha_i: - u_m D_i + sum_jk B_ijk alpha_k D_j   (on the left-hand side)
//   h beta_i:  - v_m D_i + sum_jk B_ijk beta_k  D_j
//   h u_m, h v_m: g h d_x b, g h d_y b (the bed coupling of the SWE systems)
// Friction (slip length lambda_s at the bed, Newtonian viscosity nu):
//   S_{h u_m}     = -(nu/lambda) u_b,   u_b = u_m + alpha_1 + alpha_2 (bed velocity)
//   S_{h alpha_i} = -(2i+1) (nu/lambda) u_b - (2i+1) (nu/h) sum_j C_ij alpha_j
// and the same for v, beta.  lambda_s = 0 switches the slip term off.
// Coefficients A_ijk = (2i+1) int phi_i phi_j phi_k, B_ijk = (2i+1) int phi_i'
// (int_0^zeta phi_j) phi_k, C_ij = int phi_i' phi_j' (exact rationals):
//   A_112 = A_121 = 2/5, A_211 = 2/3, A_222 = 2/7, others 0;
//   B_112 = 1/5, B_121 = -1/5, B_211 = -1, B_222 = -1/7, others 0;
//   C_11 = 4, C_22 = 12, C_12 = C_21 = 0.
// With u_m = hu * hinv etc. and hinv held fixed the friction is linear in Q,
// so the point-implicit step of make_rhs.H:429-509 is exact backward Euler.
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
    static constexpr int n_dof_q    = 8;
    static constexpr int n_dof_eq   = 8;
    static constexpr int n_dof_qaux = 7;
    static constexpr int n_parameters = 9;
    static constexpr int dimension  = 2;
    static constexpr int n_dof_gradQ = 16;
    static constexpr bool has_diffusion = false;
    static constexpr int n_boundary_tags = 5;
    static constexpr bool zamr_aux_is_local = true;
    static constexpr int zamr_n_aux_in = 0;
    static constexpr int zamr_aux_in[1] = {0};
    static constexpr int zamr_idx_hinv = 6;
    static const std::vector<std::string> get_boundary_tags() { return { "East", "North", "South", "West", "wall" }; }
    static const std::vector<std::string> parameter_names() { return { "g", "ex", "ey", "ez", "rho", "wet_dry_eps", "nu", "n_m", "lambda_s" }; }
    static const std::vector<T> default_parameters() { return { 9.81, 0.0, 0.0, 1.0, 1.0, 0.0001, 0.0, 0.0, 0.0 }; }

    using VQ = amrex::SmallMatrix<amrex::Real, 8, 1>;
    using VA = amrex::SmallMatrix<amrex::Real, 7, 1>;
    using VP = amrex::SmallMatrix<amrex::Real, 9, 1>;
    using VN = amrex::SmallMatrix<amrex::Real, 2, 1>;
    using VX = amrex::SmallMatrix<amrex::Real, 3, 1>;

    // Row indices of the momentum moments: m(d, k) is the row of the k-th
    // moment (0 = mean) of the velocity component along direction d.
    static AMREX_GPU_HOST_DEVICE AMREX_FORCE_INLINE constexpr int m(int d, int k) { return 2 + 3 * d + k; }

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

    // sum_jk A_ijk x_j y_k for i = 1, 2 (x, y the first and second moments)
    static AMREX_GPU_HOST_DEVICE AMREX_FORCE_INLINE amrex::Real zamr_A(int i, const amrex::Real* x, const amrex::Real* y)
    {
        using zamr_math::madd;
        using zamr_math::mul;
        return (i == 1) ? mul(0.4, madd(mul(x[0], y[1]), x[1], y[0]))
                        : madd(mul(mul(2.0 / 3.0, x[0]), y[0]), mul(2.0 / 7.0, x[1]), y[1]);
    }

    // Flux along direction d of the conservative part (rows 0..7).
    static AMREX_GPU_HOST_DEVICE AMREX_FORCE_INLINE VQ zamr_flux_dir(VQ const& Q, VA const& Qaux, VP const& p, int d)
    {
        VQ res{};
        const amrex::Real a = Qaux(6, 0);
        const amrex::Real h = Q(1, 0);
        const int e = 1 - d;  // the other direction
        const amrex::Real un = Q(m(d, 0), 0) * a, ut = Q(m(e, 0), 0) * a;
        const amrex::Real an[2] = {Q(m(d, 1), 0) * a, Q(m(d, 2), 0) * a};
        const amrex::Real at[2] = {Q(m(e, 1), 0) * a, Q(m(e, 2), 0) * a};
        res(0, 0) = 0.0;
        res(1, 0) = Q(m(d, 0), 0);
        // the Legendre weights 1/3, 1/5 as products (a division by a constant
        // is a full IEEE division on the device)
        res(m(d, 0), 0) = h * (un * un + an[0] * an[0] * (1.0 / 3.0) + an[1] * an[1] * (1.0 / 5.0)) + 0.5 * p(0, 0) * h * h;
        res(m(e, 0), 0) = h * (un * ut + an[0] * at[0] * (1.0 / 3.0) + an[1] * at[1] * (1.0 / 5.0));
        for (int i = 1; i <= 2; ++i) {
            res(m(d, i), 0) = h * (2.0 * un * an[i - 1] + zamr_A(i, an, an));
            res(m(e, i), 0) = h * (un * at[i - 1] + ut * an[i - 1] + zamr_A(i, at, an));
        }
        return res;
    }

    static AMREX_GPU_HOST_DEVICE AMREX_FORCE_INLINE amrex::SmallMatrix<amrex::Real, 16, 1> flux(
        VQ const& Q, VA const& Qaux, VP const& p)
    {
        auto res = amrex::SmallMatrix<amrex::Real, 16, 1>{};
        for (int d = 0; d < 2; ++d) {
            const VQ f = zamr_flux_dir(Q, Qaux, p, d);
            for (int r = 0; r < 8; ++r) res(d * 8 + r, 0) = f(r, 0);
        }
        return res;
    }

    static AMREX_GPU_HOST_DEVICE AMREX_FORCE_INLINE amrex::SmallMatrix<amrex::Real, 16, 1> dflux(
        VQ const& Q, VA const& Qaux, VP const& p)
    {
        return amrex::SmallMatrix<amrex::Real, 16, 1>{};
    }

    static AMREX_GPU_HOST_DEVICE AMREX_FORCE_INLINE amrex::SmallMatrix<amrex::Real, 16, 1> hydrostatic_pressure(
        VQ const& Q, VA const& Qaux, VP const& p)
    {
        return amrex::SmallMatrix<amrex::Real, 16, 1>{};
    }

    // Non-conservative matrix along direction d, (n_eq, n_state) row-major:
    // the bed coupling g h on the b column of the mean momentum, and the
    // moment products on the columns of the direction-d moments.
    static AMREX_GPU_HOST_DEVICE AMREX_FORCE_INLINE void zamr_ncp_dir(VQ const& Q, VA const& Qaux, VP const& p, int d,
                                                                      amrex::Real* B)
    {
        for (int k = 0; k < 64; ++k) B[k] = 0.0;
        const amrex::Real a = Qaux(6, 0);
        B[m(d, 0) * 8 + 0] = p(0, 0) * Q(1, 0);
        for (int c = 0; c < 2; ++c) {  // the alpha rows (c = 0) and the beta rows (c = 1)
            const amrex::Real mean = Q(m(c, 0), 0) * a;
            const amrex::Real mk[2] = {Q(m(c, 1), 0) * a, Q(m(c, 2), 0) * a};
            // B_ijk for (i, j, k): B_112 = 1/5, B_121 = -1/5, B_211 = -1, B_222 = -1/7
            const amrex::Real Bm[2][2] = {{0.2 * mk[1], -0.2 * mk[0]}, {-mk[0], -mk[1] * (1.0 / 7.0)}};
            for (int i = 1; i <= 2; ++i)
                for (int j = 1; j <= 2; ++j)
                    B[m(c, i) * 8 + m(d, j)] = zamr_math::add((i == j) ? -mean : 0.0, Bm[i - 1][j - 1]);
        }
    }

    static AMREX_GPU_HOST_DEVICE AMREX_FORCE_INLINE amrex::SmallMatrix<amrex::Real, 128, 1> nonconservative_matrix(
        VQ const& Q, VA const& Qaux, VP const& p)
    {
        auto res = amrex::SmallMatrix<amrex::Real, 128, 1>{};
        amrex::Real B[64];
        for (int d = 0; d < 2; ++d) {
            zamr_ncp_dir(Q, Qaux, p, d, B);
            for (int k = 0; k < 64; ++k) res(d * 64 + k, 0) = B[k];
        }
        return res;
    }

    // d(flux_d)/dQ + nonconservative_matrix_d, dir-major flat (dim, n_eq, n_state).
    static AMREX_GPU_HOST_DEVICE AMREX_FORCE_INLINE void zamr_quasilinear_dir(VQ const& Q, VA const& Qaux, VP const& p,
                                                                              int d, amrex::Real* A)
    {
        zamr_ncp_dir(Q, Qaux, p, d, A);
        const amrex::Real a = Qaux(6, 0);
        const amrex::Real h = Q(1, 0);
        const int e = 1 - d;
        const amrex::Real un = Q(m(d, 0), 0) * a, ut = Q(m(e, 0), 0) * a;
        const amrex::Real an[2] = {Q(m(d, 1), 0) * a, Q(m(d, 2), 0) * a};
        const amrex::Real at[2] = {Q(m(e, 1), 0) * a, Q(m(e, 2), 0) * a};
        const amrex::Real cj[2] = {1.0 / 3.0, 1.0 / 5.0};
        // h row
        A[1 * 8 + m(d, 0)] = zamr_math::add(A[1 * 8 + m(d, 0)], 1.0);
        // mean momentum along d
        {
            const int r = m(d, 0);
            A[r * 8 + 1] = zamr_math::add(A[r * 8 + 1], -un * un - cj[0] * an[0] * an[0] - cj[1] * an[1] * an[1] + p(0, 0) * h);
            A[r * 8 + m(d, 0)] = zamr_math::add(A[r * 8 + m(d, 0)], 2.0 * un);
            for (int j = 0; j < 2; ++j) A[r * 8 + m(d, j + 1)] = zamr_math::add(A[r * 8 + m(d, j + 1)], 2.0 * cj[j] * an[j]);
        }
        // mean momentum across d
        {
            const int r = m(e, 0);
            A[r * 8 + 1] = zamr_math::add(A[r * 8 + 1], -un * ut - cj[0] * an[0] * at[0] - cj[1] * an[1] * at[1]);
            A[r * 8 + m(d, 0)] = zamr_math::add(A[r * 8 + m(d, 0)], ut);
            A[r * 8 + m(e, 0)] = zamr_math::add(A[r * 8 + m(e, 0)], un);
            for (int j = 0; j < 2; ++j) {
                A[r * 8 + m(d, j + 1)] = zamr_math::add(A[r * 8 + m(d, j + 1)], cj[j] * at[j]);
                A[r * 8 + m(e, j + 1)] = zamr_math::add(A[r * 8 + m(e, j + 1)], cj[j] * an[j]);
            }
        }
        // moments along d: h (2 u_n a_i + A_i(a, a))
        for (int i = 1; i <= 2; ++i) {
            const int r = m(d, i);
            A[r * 8 + 1] = zamr_math::add(A[r * 8 + 1], -2.0 * un * an[i - 1] - zamr_A(i, an, an));
            A[r * 8 + m(d, 0)] = zamr_math::add(A[r * 8 + m(d, 0)], 2.0 * an[i - 1]);
            for (int q = 1; q <= 2; ++q) {
                amrex::Real ex[2] = {0.0, 0.0};
                ex[q - 1] = 1.0;
                // d/d(h a_q) of A_i(a, a) h = A_i(e_q, a) + A_i(a, e_q)
                A[r * 8 + m(d, q)] = zamr_math::add(A[r * 8 + m(d, q)], zamr_math::add(zamr_math::add((i == q) ? 2.0 * un : 0.0, zamr_A(i, ex, an)), zamr_A(i, an, ex)));
            }
        }
        // moments across d: h (u_n b_i + u_t a_i + A_i(b, a))
        for (int i = 1; i <= 2; ++i) {
            const int r = m(e, i);
            A[r * 8 + 1] = zamr_math::add(A[r * 8 + 1], -(un * at[i - 1] + ut * an[i - 1] + zamr_A(i, at, an)));
            A[r * 8 + m(d, 0)] = zamr_math::add(A[r * 8 + m(d, 0)], at[i - 1]);
            A[r * 8 + m(e, 0)] = zamr_math::add(A[r * 8 + m(e, 0)], an[i - 1]);
            for (int q = 1; q <= 2; ++q) {
                amrex::Real ex[2] = {0.0, 0.0};
                ex[q - 1] = 1.0;
                A[r * 8 + m(d, q)] = zamr_math::add(A[r * 8 + m(d, q)], zamr_math::add((i == q) ? ut : 0.0, zamr_A(i, at, ex)));
                A[r * 8 + m(e, q)] = zamr_math::add(A[r * 8 + m(e, q)], zamr_math::add((i == q) ? un : 0.0, zamr_A(i, ex, an)));
            }
        }
    }

    static AMREX_GPU_HOST_DEVICE AMREX_FORCE_INLINE amrex::SmallMatrix<amrex::Real, 128, 1> quasilinear_matrix(
        VQ const& Q, VA const& Qaux, VP const& p)
    {
        auto res = amrex::SmallMatrix<amrex::Real, 128, 1>{};
        amrex::Real A[64];
        for (int d = 0; d < 2; ++d) {
            zamr_quasilinear_dir(Q, Qaux, p, d, A);
            for (int k = 0; k < 64; ++k) res(d * 64 + k, 0) = A[k];
        }
        return res;
    }

    static AMREX_GPU_HOST_DEVICE AMREX_FORCE_INLINE amrex::SmallMatrix<amrex::Real, 16, 1> diffusive_flux(
        VQ const& Q, VA const& Qaux, amrex::SmallMatrix<amrex::Real, 16, 1> const& gradQ, VP const& p)
    {
        return amrex::SmallMatrix<amrex::Real, 16, 1>{};
    }

    // No closed-form spectrum at level 2: zeros, as the printer emits for a
    // system without eigenvalues (to_amrex.py:276-282); the numerics bound the
    // spectral radius instead.
    static AMREX_GPU_HOST_DEVICE AMREX_FORCE_INLINE VQ eigenvalues(VQ const& Q, VA const& Qaux, VP const& p, VN const& n)
    {
        return VQ{};
    }

    static AMREX_GPU_HOST_DEVICE AMREX_FORCE_INLINE amrex::SmallMatrix<amrex::Real, 64, 1> left_eigenvectors(
        VQ const& Q, VA const& Qaux, VP const& p, VN const& n)
    {
        return amrex::SmallMatrix<amrex::Real, 64, 1>{};
    }

    static AMREX_GPU_HOST_DEVICE AMREX_FORCE_INLINE amrex::SmallMatrix<amrex::Real, 64, 1> right_eigenvectors(
        VQ const& Q, VA const& Qaux, VP const& p, VN const& n)
    {
        return amrex::SmallMatrix<amrex::Real, 64, 1>{};
    }

    // Friction factors: slip = nu / lambda_s (0 when lambda_s = 0), newt = nu.
    static AMREX_GPU_HOST_DEVICE AMREX_FORCE_INLINE amrex::Real zamr_slip(VP const& p)
    {
        return (p(8, 0) > 0.0) ? p(6, 0) / p(8, 0) : 0.0;
    }

    static AMREX_GPU_HOST_DEVICE AMREX_FORCE_INLINE VQ source(
        VQ const& Q, VA const& Qaux, VP const& p, amrex::Real const time, VX const& X)
    {
        VQ res{};
        const amrex::Real a = Qaux(6, 0);
        const amrex::Real slip = zamr_slip(p), nu = p(6, 0);
        for (int d = 0; d < 2; ++d) {
            const amrex::Real ub = (Q(m(d, 0), 0) + Q(m(d, 1), 0) + Q(m(d, 2), 0)) * a;
            res(m(d, 0), 0) = -slip * ub;
            res(m(d, 1), 0) = -3.0 * slip * ub - 3.0 * nu * a * a * 4.0 * Q(m(d, 1), 0);
            res(m(d, 2), 0) = -5.0 * slip * ub - 5.0 * nu * a * a * 12.0 * Q(m(d, 2), 0);
        }
        return res;
    }

    // Row-major (n_eq, n_state), hinv held fixed: the friction is linear in Q.
    static AMREX_GPU_HOST_DEVICE AMREX_FORCE_INLINE amrex::SmallMatrix<amrex::Real, 64, 1> source_jacobian_wrt_variables(
        VQ const& Q, VA const& Qaux, VP const& p)
    {
        auto res = amrex::SmallMatrix<amrex::Real, 64, 1>{};
        const amrex::Real a = Qaux(6, 0);
        const amrex::Real slip = zamr_slip(p), nu = p(6, 0);
        const amrex::Real w[3] = {1.0, 3.0, 5.0};
        const amrex::Real cd[3] = {0.0, 4.0, 12.0};
        for (int d = 0; d < 2; ++d)
            for (int i = 0; i < 3; ++i) {
                for (int j = 0; j < 3; ++j) res(m(d, i) * 8 + m(d, j), 0) = -w[i] * slip * a;
                if (cd[i] != 0.0) res(m(d, i) * 8 + m(d, i), 0) += -w[i] * nu * a * a * cd[i];
            }
        return res;
    }

    static AMREX_GPU_HOST_DEVICE AMREX_FORCE_INLINE amrex::SmallMatrix<amrex::Real, 56, 1> source_jacobian_wrt_aux_variables(
        VQ const& Q, VA const& Qaux, VP const& p)
    {
        auto res = amrex::SmallMatrix<amrex::Real, 56, 1>{};
        const amrex::Real a = Qaux(6, 0);
        const amrex::Real slip = zamr_slip(p), nu = p(6, 0);
        const amrex::Real w[3] = {1.0, 3.0, 5.0};
        const amrex::Real cd[3] = {0.0, 4.0, 12.0};
        for (int d = 0; d < 2; ++d) {
            const amrex::Real sum = Q(m(d, 0), 0) + Q(m(d, 1), 0) + Q(m(d, 2), 0);
            for (int i = 0; i < 3; ++i)
                res(m(d, i) * 7 + 6, 0) = (cd[i] != 0.0) ? -w[i] * slip * sum - 2.0 * w[i] * nu * a * cd[i] * Q(m(d, i), 0)
                                                         : -w[i] * slip * sum;
        }
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
        res(6, 0) = zamr_hinv(Q(1, 0), p(5, 0));
        return res;
    }

    static AMREX_GPU_HOST_DEVICE AMREX_FORCE_INLINE amrex::SmallMatrix<amrex::Real, 64, 1> update_variables_jacobian_wrt_variables(
        VQ const& Q, VA const& Qaux, VP const& p)
    {
        auto res = amrex::SmallMatrix<amrex::Real, 64, 1>{};
        for (int r = 0; r < 8; ++r) res(r * 8 + r, 0) = 1.0;
        return res;
    }

    static AMREX_GPU_HOST_DEVICE AMREX_FORCE_INLINE amrex::SmallMatrix<amrex::Real, 56, 1> update_aux_variables_jacobian_wrt_variables(
        VQ const& Q, VA const& Qaux, VP const& p, amrex::Real const time, VX const& X)
    {
        return amrex::SmallMatrix<amrex::Real, 56, 1>{};
    }

    // Tags sorted: East, North, South, West (extrapolation) and wall
    // (reflective: every moment of the normal velocity flips sign).
    static AMREX_GPU_HOST_DEVICE AMREX_FORCE_INLINE VQ boundary_conditions(
        const int bc_idx, VQ const& Q, VA const& Qaux, VN const& n, VX const& X,
        amrex::Real const time, amrex::Real const dX)
    {
        VQ res = Q;
        if (bc_idx == 4) {
            for (int k = 0; k < 3; ++k) {
                const amrex::Real qn = Q(m(0, k), 0) * n(0, 0) + Q(m(1, k), 0) * n(1, 0);
                res(m(0, k), 0) = Q(m(0, k), 0) - 2.0 * qn * n(0, 0);
                res(m(1, k), 0) = Q(m(1, k), 0) - 2.0 * qn * n(1, 0);
            }
        }
        return res;
    }

    static AMREX_GPU_HOST_DEVICE AMREX_FORCE_INLINE VA aux_boundary_conditions(
        const int bc_idx, VQ const& Q, VA const& Qaux, VN const& n, VX const& X,
        amrex::Real const time, amrex::Real const dX)
    {
        return Qaux;
    }

    static constexpr int n_e2s = 8;
    static constexpr int e2s[8] = { 0, 1, 2, 3, 4, 5, 6, 7 };
    static constexpr int n_deriv_aux = 0;
    static constexpr int deriv_aux[1][4] = {{0,0,0,0}};
    static constexpr int n_input_aux = 0;
    static constexpr int input_aux[1][4] = {{0,0,0,0}};
};
