#pragma once
// Hand-written reverse-mode (adjoint) functions of the per-cell-Manning
// shallow-water system in this directory (Model.H, Numerics.H).  They give,
// for one face or one cell, the transpose of the derivative of the forward
// map applied to an output adjoint.  The reference has no gradients; these
// are derived from the formulas of Model.H / Numerics.H and checked against
// finite differences of those same functions.
//
// Non-smooth points take one fixed one-sided derivative (a subgradient):
//   |x| and sqrt(x) at 0            -> 0
//   max(a, b)                       -> derivative of the argument max returns
//                                      ((a < b) ? b : a)
//   the Heaviside in the source Jacobian is piecewise constant -> 0.
#include "Numerics.H"

struct Adjoint {
    using VQ = Model::VQ;
    using VP = Model::VP;
    using VN = Model::VN;
    static constexpr int friction_row = 1;  // aux row holding the Manning field

    // d hinv / dh of Model::zamr_hinv: hinv = sqrt(2) h / sqrt(h^4 + max(h^4, eps^4)).
    static AMREX_GPU_HOST_DEVICE AMREX_FORCE_INLINE amrex::Real dhinv_dh(amrex::Real h, amrex::Real eps)
    {
        const amrex::Real h4 = amrex::Math::powi<4>(h);
        const amrex::Real e4 = amrex::Math::powi<4>(eps);
        const bool take_h = !(h4 < e4);
        const amrex::Real D = h4 + (take_h ? h4 : e4);
        const amrex::Real dD = (take_h ? 8.0 : 4.0) * h * h * h;
        return 1.4142135623730951 / std::sqrt(D) * (1.0 - 0.5 * h * dD / D);
    }

    // The point-implicit Manning step of make_rhs.H:429-509 for this model in
    // closed form.  With S = -k m, k = g n^2 |m| hinv r^4, r = max(h,eps)^(-1/3)
    // and hinv held at the stage-start value a, the linear system (I - dt J) dq
    // = dt S has dq_b = dq_h = 0 and, for the momenta m = (hu, hv),
    //   dq = -dt k m / (1 + dt k + dt q |m|^2),  q = g n^2 r^4 a^2 / (a|m|),
    // where dt q |m|^2 = dt k, so W = m (1 + tau) / (1 + 2 tau), tau = dt k.
    template <class T>
    static AMREX_GPU_HOST_DEVICE AMREX_FORCE_INLINE void manning_step(T h, T m0, T m1, T a, T nn, amrex::Real dt,
                                                                       VP const& p, T& w0, T& w1)
    {
        const T r = Model::zamr_rcbrt(amrex::max(h, T(p(5, 0))));
        const T r4 = amrex::Math::powi<4>(r);
        const T rho = std::sqrt(m0 * m0 + m1 * m1);
        const T tau = dt * p(0, 0) * nn * nn * r4 * a * rho;
        const T phi = (1.0 + tau) / (1.0 + 2.0 * tau);
        w0 = m0 * phi;
        w1 = m1 * phi;
    }

    // Reverse of W = implicit_source(V, aux(U), dt) at one cell, where the
    // source reads the stage-START aux (hinv of U_h, make_rhs.H:440) and the
    // stored Manning coefficient nn.  Given Wb, returns the adjoint of V
    // (Vb), the part that reaches U through hinv (Ub, h row only) and of nn.
    static AMREX_GPU_HOST_DEVICE AMREX_FORCE_INLINE void source_reverse(VQ const& U, VQ const& V, amrex::Real nn,
                                                                        amrex::Real dt, VP const& p, VQ const& Wb,
                                                                        VQ& Vb, VQ& Ub, amrex::Real& nb)
    {
        const amrex::Real eps = p(5, 0);
        const amrex::Real a = Model::zamr_hinv(U(1, 0), eps);
        const amrex::Real h = V(1, 0), m0 = V(2, 0), m1 = V(3, 0);
        const bool take_h = !(h < eps);
        const amrex::Real hf = take_h ? h : eps;
        const amrex::Real r = Model::zamr_rcbrt(hf);
        const amrex::Real r4 = amrex::Math::powi<4>(r);
        const amrex::Real rho = std::sqrt(m0 * m0 + m1 * m1);
        const amrex::Real C = dt * p(0, 0);
        const amrex::Real base = C * nn * nn * r4 * a;  // tau / rho
        const amrex::Real tau = base * rho;
        const amrex::Real den = 1.0 / (1.0 + 2.0 * tau);
        const amrex::Real phi = (1.0 + tau) * den;
        // phi'(tau) = -1 / (1 + 2 tau)^2
        const amrex::Real taub = -(den * den) * (Wb(2, 0) * m0 + Wb(3, 0) * m1);
        Vb(0, 0) = Wb(0, 0);
        Vb(2, 0) = Wb(2, 0) * phi;
        Vb(3, 0) = Wb(3, 0) * phi;
        // tau = C nn^2 r4 a rho
        const amrex::Real rhob = taub * base;
        if (rho > 0.0) {
            Vb(2, 0) += rhob * m0 / rho;
            Vb(3, 0) += rhob * m1 / rho;
        }
        const amrex::Real common = taub * C * r4 * rho;
        nb = common * 2.0 * nn * a;
        const amrex::Real ab = common * nn * nn;
        // d r4 / dh = -(4/3) r4 / hf on h >= eps
        const amrex::Real r4b = taub * C * nn * nn * a * rho;
        Vb(1, 0) = Wb(1, 0) + (take_h ? r4b * (-(4.0 / 3.0)) * r4 / hf : 0.0);
        Ub(0, 0) = 0.0;
        Ub(1, 0) = ab * dhinv_dh(U(1, 0), eps);
        Ub(2, 0) = 0.0;
        Ub(3, 0) = 0.0;
    }

    // Reverse of Numerics::numerical_flux and numerical_fluctuations on the
    // face states qL, qR with unit normal e_DIR, including the face aux
    // refresh (hinv of the face depths, make_rhs.H:114-115).  Fb, Dpb, Dmb are
    // the adjoints of flux, D+ and D-; qLb, qRb receive the adjoints of qL, qR.
    template <int DIR>
    static AMREX_GPU_HOST_DEVICE AMREX_FORCE_INLINE void face_reverse(VQ const& qL, VQ const& qR, VP const& p,
                                                                      VQ const& Fb, VQ const& Dpb, VQ const& Dmb,
                                                                      VQ& qLb, VQ& qRb)
    {
        constexpr int m = 2 + DIR;  // normal momentum row
        const amrex::Real g = p(0, 0), eps = p(5, 0);
        const amrex::Real hL = qL(1, 0), hR = qR(1, 0);
        const amrex::Real aL = Model::zamr_hinv(hL, eps), aR = Model::zamr_hinv(hR, eps);
        const amrex::Real qnL = qL(m, 0), qnR = qR(m, 0);
        const amrex::Real unL = qnL * aL, unR = qnR * aR;
        const amrex::Real cL = (hL > 0.0) ? std::sqrt(g * hL) : 0.0;
        const amrex::Real cR = (hR > 0.0) ? std::sqrt(g * hR) : 0.0;
        const amrex::Real sL = std::fabs(unL) + cL;
        const amrex::Real sR = std::fabs(unR) + cR;
        const bool take_R = sL < sR;
        const amrex::Real hs = 0.5 * (take_R ? sR : sL);

        amrex::Real hLb = 0.0, hRb = 0.0, unLb = 0.0, unRb = 0.0, hsb = 0.0;
        amrex::Real mLb[2] = {0.0, 0.0}, mRb[2] = {0.0, 0.0};
        // F_h = (qnL + qnR)/2 - hs (hR - hL)
        mLb[DIR] += 0.5 * Fb(1, 0);
        mRb[DIR] += 0.5 * Fb(1, 0);
        hsb -= Fb(1, 0) * (hR - hL);
        hLb += Fb(1, 0) * hs;
        hRb -= Fb(1, 0) * hs;
        // F_{m_d} = ((m_dL unL + phL n_d) + (m_dR unR + phR n_d))/2 - hs (m_dR - m_dL)
#pragma unroll
        for (int d = 0; d < 2; ++d) {
            const amrex::Real fb = Fb(2 + d, 0);
            const amrex::Real mL = qL(2 + d, 0), mR = qR(2 + d, 0);
            mLb[d] += 0.5 * fb * unL + fb * hs;
            mRb[d] += 0.5 * fb * unR - fb * hs;
            unLb += 0.5 * fb * mL;
            unRb += 0.5 * fb * mR;
            hsb -= fb * (mR - mL);
            if (d == DIR) {  // ph = g h^2 / 2
                hLb += 0.5 * fb * g * hL;
                hRb += 0.5 * fb * g * hR;
            }
        }
        // hs = max(sL, sR) / 2, s = |un| + sqrt(g max(h, 0))
        const amrex::Real sLb = take_R ? 0.0 : 0.5 * hsb;
        const amrex::Real sRb = take_R ? 0.5 * hsb : 0.0;
        unLb += (unL > 0.0) ? sLb : ((unL < 0.0) ? -sLb : 0.0);
        unRb += (unR > 0.0) ? sRb : ((unR < 0.0) ? -sRb : 0.0);
        if (hL > 0.0) hLb += sLb * 0.5 * g / cL;
        if (hR > 0.0) hRb += sRb * 0.5 * g / cR;
        // un = qn hinv(h)
        mLb[DIR] += unLb * aL;
        mRb[DIR] += unRb * aR;
        hLb += unLb * qnL * dhinv_dh(hL, eps);
        hRb += unRb * qnR * dhinv_dh(hR, eps);
        // D+ = D- = w e_DIR on the momentum rows, w = g (hL + hR)(bR - bL) / 4
        const amrex::Real wb = Dpb(m, 0) + Dmb(m, 0);
        const amrex::Real bL = qL(0, 0), bR = qR(0, 0);
        const amrex::Real wq = 0.25 * g * wb;
        hLb += wq * (bR - bL);
        hRb += wq * (bR - bL);
        qLb(0, 0) = -wq * (hL + hR);
        qRb(0, 0) = wq * (hL + hR);
        qLb(1, 0) = hLb;
        qRb(1, 0) = hRb;
        qLb(2, 0) = mLb[0];
        qLb(3, 0) = mLb[1];
        qRb(2, 0) = mRb[0];
        qRb(3, 0) = mRb[1];
    }

    // Reverse of the hydrostatic (Audusse) reconstruction of make_rhs.H:84-98
    // followed by the face functions on the reconstructed states:
    //   bf = max(bL, bR), h* = max(0, h + b - bf), momenta scaled by h*/h (by
    //   0 on a side with h <= 0), and b* = bf on both sides;
    // the reconstructed depths also feed the cells' balancing bed source
    // (make_rhs.H:160-162), g/2 h*^2 / dx on the normal momentum.  kL, kR
    // weigh that use: the adjoint of h*_L from it is kL h*_L (of h*_R,
    // kR h*_R).  qL, qR are the face states before the reconstruction; Fb,
    // Dpb, Dmb the adjoints of flux and fluctuations as for face_reverse.
    template <int DIR>
    static AMREX_GPU_HOST_DEVICE AMREX_FORCE_INLINE void wb_face_reverse(VQ const& qL, VQ const& qR, VP const& p,
                                                                         VQ const& Fb, VQ const& Dpb, VQ const& Dmb,
                                                                         amrex::Real kL, amrex::Real kR, VQ& qLb,
                                                                         VQ& qRb)
    {
        const amrex::Real bL = qL(0, 0), bR = qR(0, 0);
        const bool bf_from_R = bL < bR;  // amrex::max(bL, bR) = (bL < bR) ? bR : bL
        const amrex::Real bf = bf_from_R ? bR : bL;
        const amrex::Real hL = qL(1, 0), hR = qR(1, 0);
        const amrex::Real xL = hL + bL - bf, xR = hR + bR - bf;
        const bool posL = 0.0 < xL, posR = 0.0 < xR;  // max(0, x)
        const amrex::Real hLs = posL ? xL : amrex::Real(0.0), hRs = posR ? xR : amrex::Real(0.0);
        const bool wetL = hL > 0.0, wetR = hR > 0.0;
        const amrex::Real sL = wetL ? hLs / hL : amrex::Real(0.0), sR = wetR ? hRs / hR : amrex::Real(0.0);
        VQ L, R;
        L(0, 0) = bf;
        R(0, 0) = bf;
        L(1, 0) = hLs;
        R(1, 0) = hRs;
        for (int d = 0; d < 2; ++d) {
            L(2 + d, 0) = qL(2 + d, 0) * sL;
            R(2 + d, 0) = qR(2 + d, 0) * sR;
        }
        VQ lb, rb;
        face_reverse<DIR>(L, R, p, Fb, Dpb, Dmb, lb, rb);
        amrex::Real bfb = lb(0, 0) + rb(0, 0);
        amrex::Real hLsb = lb(1, 0) + kL * hLs, hRsb = rb(1, 0) + kR * hRs;
        amrex::Real sLb = 0.0, sRb = 0.0;
        for (int d = 0; d < 2; ++d) {
            qLb(2 + d, 0) = lb(2 + d, 0) * sL;
            qRb(2 + d, 0) = rb(2 + d, 0) * sR;
            sLb += lb(2 + d, 0) * qL(2 + d, 0);
            sRb += rb(2 + d, 0) * qR(2 + d, 0);
        }
        amrex::Real hLb = 0.0, hRb = 0.0, bLb = 0.0, bRb = 0.0;
        if (wetL) {  // s = h*/h
            hLsb += sLb / hL;
            hLb -= sLb * hLs / (hL * hL);
        }
        if (wetR) {
            hRsb += sRb / hR;
            hRb -= sRb * hRs / (hR * hR);
        }
        if (posL) {  // h* = h + b - bf
            hLb += hLsb;
            bLb += hLsb;
            bfb -= hLsb;
        }
        if (posR) {
            hRb += hRsb;
            bRb += hRsb;
            bfb -= hRsb;
        }
        if (bf_from_R) bRb += bfb;
        else bLb += bfb;
        qLb(0, 0) = bLb;
        qLb(1, 0) = hLb;
        qRb(0, 0) = bRb;
        qRb(1, 0) = hRb;
    }

    // Transpose of Model::boundary_conditions on a ghost adjoint: tags 0-3
    // (extrapolation) copy, tag 4 (wall) flips the momentum normal to side
    // direction dir.
    static AMREX_GPU_HOST_DEVICE AMREX_FORCE_INLINE VQ bc_reverse(int bc_idx, int dir, VQ v)
    {
        if (bc_idx == 4) v(2 + dir, 0) = -v(2 + dir, 0);
        return v;
    }
};
