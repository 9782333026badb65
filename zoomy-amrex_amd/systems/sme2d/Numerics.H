#pragma once
// Riemann solver for the level-2 shallow moment equations in the interface
// zoomy_core's AmrexNumerics emits (called at Source/make_rhs.H:129-130 and
// Source/ZoomyAmr.cpp:483): a nonconservative Rusanov scheme.
//
//   numerical_flux         = (F(qL).n + F(qR).n)/2 - (s/2)(qR - qL) on rows h .. hv2,
//                            s = max(rho(qL), rho(qR)); the bed row carries nothing
//   numerical_fluctuations = [D+ | D-] with D+ = D- = (1/2) int_0^1 B(q(s)).n ds (qR - qL)
//                            along the straight path q(s) = qL + s (qR - qL), by
//                            3-point Gauss-Legendre (exact on the bed coupling g h db)
//   local_max_abs_eigenvalue = rho(q): a Gershgorin bound of the spectrum of the
//                            quasilinear matrix A.n (the reference's fallback when a
//                            system has no closed-form eigenvalues, README.md:20-22).
//
// The bound is taken on D^-1 (A.n) D with D = diag(1, c, .., c) over the rows
// and columns h .. hv2 (the bed row is zero, so the bed column does not move
// the spectrum), c = sqrt(g max(h, eps)): similar matrices share eigenvalues,
// and the scaling makes the bound |u.n| + sqrt(g h) when the moments vanish
// (the shallow-water speed) instead of the far looser plain row sum g h.
#include "Model.H"

struct Numerics {
    static constexpr bool fluctuations_are_zero = false;

    static AMREX_GPU_HOST_DEVICE AMREX_FORCE_INLINE amrex::SmallMatrix<amrex::Real, 1, 1> local_max_abs_eigenvalue(
        Model::VQ const& Q, Model::VA const& Qaux, Model::VP const& p, Model::VN const& n)
    {
        amrex::Real A[64], Ad[64];
        for (int k = 0; k < 64; ++k) A[k] = 0.0;
        for (int d = 0; d < 2; ++d) {
            if (n(d, 0) == 0.0) continue;
            Model::zamr_quasilinear_dir(Q, Qaux, p, d, Ad);
            for (int k = 0; k < 64; ++k) A[k] = zamr_math::madd(A[k], n(d, 0), Ad[k]);
        }
        const amrex::Real c2 = p(0, 0) * amrex::max(Q(1, 0), p(5, 0));
        const amrex::Real c = zamr_math::sqrt(c2);
        const amrex::Real rc = zamr_math::div_sqrt(1.0, c2);  // 1 / c
        amrex::Real rho = 0.0;
        for (int i = 1; i < 8; ++i) {
            amrex::Real s = 0.0;
            for (int j = 1; j < 8; ++j) {
                amrex::Real v = std::fabs(A[i * 8 + j]);
                if (i == 1 && j != 1) v = zamr_math::mul(v, c);
                if (i != 1 && j == 1) v = zamr_math::mul(v, rc);
                s = zamr_math::add(s, v);
            }
            rho = amrex::max(rho, s);
        }
        amrex::SmallMatrix<amrex::Real, 1, 1> res{};
        res(0, 0) = rho;
        return res;
    }

    static AMREX_GPU_HOST_DEVICE AMREX_FORCE_INLINE Model::VQ numerical_flux(
        Model::VQ const& qL, Model::VQ const& qR, Model::VA const& aL, Model::VA const& aR,
        Model::VP const& p, Model::VN const& n)
    {
        Model::VQ fL{}, fR{};
        for (int d = 0; d < 2; ++d) {
            if (n(d, 0) == 0.0) continue;
            const Model::VQ a = Model::zamr_flux_dir(qL, aL, p, d);
            const Model::VQ b = Model::zamr_flux_dir(qR, aR, p, d);
            for (int r = 0; r < 8; ++r) {
                fL(r, 0) = zamr_math::madd(fL(r, 0), n(d, 0), a(r, 0));
                fR(r, 0) = zamr_math::madd(fR(r, 0), n(d, 0), b(r, 0));
            }
        }
        amrex::Real sL = local_max_abs_eigenvalue(qL, aL, p, n)(0, 0);
        zamr_math::sequence_point(sL);  // one speed bound at a time (registers)
        const amrex::Real sR = local_max_abs_eigenvalue(qR, aR, p, n)(0, 0);
        const amrex::Real hs = 0.5 * amrex::max(sL, sR);
        Model::VQ res{};
        res(0, 0) = 0.0;
        for (int r = 1; r < 8; ++r) res(r, 0) = 0.5 * (fL(r, 0) + fR(r, 0)) - hs * (qR(r, 0) - qL(r, 0));
        return res;
    }

    static AMREX_GPU_HOST_DEVICE AMREX_FORCE_INLINE amrex::SmallMatrix<amrex::Real, 16, 1> numerical_fluctuations(
        Model::VQ const& qL, Model::VQ const& qR, Model::VA const& aL, Model::VA const& aR,
        Model::VP const& p, Model::VN const& n)
    {
        // Gauss-Legendre on [0, 1]: nodes 1/2 -+ sqrt(15)/10, 1/2; weights 5/18, 8/18, 5/18
        const amrex::Real s3[3] = {0.5 - 0.3872983346207417, 0.5, 0.5 + 0.3872983346207417};
        const amrex::Real w3[3] = {5.0 / 18.0, 8.0 / 18.0, 5.0 / 18.0};
        Model::VQ dq;
        for (int r = 0; r < 8; ++r) dq(r, 0) = qR(r, 0) - qL(r, 0);
        Model::VQ I{};
        amrex::Real B[64];
        for (int k = 0; k < 3; ++k) {
            Model::VQ q;
            for (int r = 0; r < 8; ++r) q(r, 0) = qL(r, 0) + s3[k] * dq(r, 0);
            Model::VA a = aL;
            a(6, 0) = Model::zamr_hinv(q(1, 0), p(5, 0));
            for (int d = 0; d < 2; ++d) {
                if (n(d, 0) == 0.0) continue;
                Model::zamr_ncp_dir(q, a, p, d, B);
                for (int i = 0; i < 8; ++i) {
                    amrex::Real acc = 0.0;
                    for (int j = 0; j < 8; ++j) acc = zamr_math::madd(acc, B[i * 8 + j], dq(j, 0));
                    I(i, 0) = zamr_math::madd(I(i, 0), zamr_math::mul(w3[k], n(d, 0)), acc);
                }
            }
        }
        amrex::SmallMatrix<amrex::Real, 16, 1> res{};
        for (int i = 0; i < 8; ++i) {
            res(i, 0) = zamr_math::mul(0.5, I(i, 0));
            res(8 + i, 0) = zamr_math::mul(0.5, I(i, 0));
        }
        return res;
    }
};
