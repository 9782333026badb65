// Host harness for the CPU tests of the hand-written reverse functions of
// systems/swe2d_nf/Adjoint.H: exposes the forward model functions (flux,
// fluctuations, the point-implicit source solved by the oracle's Gaussian
// elimination) and their reverse counterparts through a C ABI, so that the
// tests can compare the reverse functions with finite differences of the
// forward ones.  Built by the test itself with g++.
#include "Model.H"
#include "Numerics.H"
#include "Adjoint.H"
#include "zamr_oracle.hpp"

namespace {
using VQ = Model::VQ;
using VA = Model::VA;
using VP = Model::VP;
using VN = Model::VN;
using VX = Model::VX;

VQ vq(const double* a) {
    VQ v;
    for (int n = 0; n < 4; ++n) v(n, 0) = a[n];
    return v;
}
VP vp(const double* a) {
    VP v;
    for (int n = 0; n < Model::n_parameters; ++n) v(n, 0) = a[n];
    return v;
}
VA aux_of(const VQ& q, double nn, const VP& p) {
    VA a{};
    a(1, 0) = nn;
    return Model::update_aux_variables(q, a, p, 0.0, VX{});
}
}  // namespace

extern "C" {

// flux (4) then fluctuations D+ (4), D- (4) on the face with normal e_dir;
// face aux from update_aux_variables of each face state.
void hz_flux(const double* qL, const double* qR, int dir, const double* pp, double* out) {
    const VP p = vp(pp);
    const VQ L = vq(qL), R = vq(qR);
    const VA aL = aux_of(L, 0.0, p), aR = aux_of(R, 0.0, p);
    VN nh{};
    nh(dir, 0) = 1.0;
    const VQ F = Numerics::numerical_flux(L, R, aL, aR, p, nh);
    const auto D = Numerics::numerical_fluctuations(L, R, aL, aR, p, nh);
    for (int n = 0; n < 4; ++n) out[n] = F(n, 0);
    for (int n = 0; n < 8; ++n) out[4 + n] = D(n, 0);
}

void hz_face_reverse(const double* qL, const double* qR, int dir, const double* pp, const double* Fb,
                     const double* Dpb, const double* Dmb, double* qLb, double* qRb) {
    const VP p = vp(pp);
    VQ lb, rb;
    if (dir == 0) Adjoint::face_reverse<0>(vq(qL), vq(qR), p, vq(Fb), vq(Dpb), vq(Dmb), lb, rb);
    else Adjoint::face_reverse<1>(vq(qL), vq(qR), p, vq(Fb), vq(Dpb), vq(Dmb), lb, rb);
    for (int n = 0; n < 4; ++n) {
        qLb[n] = lb(n, 0);
        qRb[n] = rb(n, 0);
    }
}

// The oracle's face evaluation (make_rhs.H:31-131) with the hydrostatic
// reconstruction, order 1 (the face states are the two cells' states), on a
// 2x2 domain: flux (4), D+ (4), D- (4), then the reconstructed depths h*_L,
// h*_R (the balancing bed source's inputs).
void hz_wb_face(const double* qL, const double* qR, int dir, const double* pp, double* out) {
    zorc::Settings st;
    st.nx = 2;
    st.ny = 2;
    st.max_grid_size = 2;
    st.spatial_order = 1;
    st.well_balanced = true;
    zorc::Solver<Model, Numerics> o(st, pp);
    auto& fq = o.Q.fabs[0];
    const auto& fa = o.Qaux.fabs[0];
    const int iR = dir == 0 ? 1 : 0, jR = dir == 0 ? 0 : 1;
    for (int n = 0; n < 4; ++n) {
        fq(0, 0, n) = qL[n];
        fq(iR, jR, n) = qR[n];
    }
    VQ F;
    amrex::SmallMatrix<double, 8, 1> D;
    double hl = 0.0, hr = 0.0;
    o.eval_face(fq, fa, 0, 0, iR, jR, dir, 1, F, D, 0.0, true, &hl, &hr);
    for (int n = 0; n < 4; ++n) out[n] = F(n, 0);
    for (int n = 0; n < 8; ++n) out[4 + n] = D(n, 0);
    out[12] = hl;
    out[13] = hr;
}

void hz_wb_face_reverse(const double* qL, const double* qR, int dir, const double* pp, const double* Fb,
                        const double* Dpb, const double* Dmb, double kL, double kR, double* qLb, double* qRb) {
    const VP p = vp(pp);
    VQ lb, rb;
    if (dir == 0) Adjoint::wb_face_reverse<0>(vq(qL), vq(qR), p, vq(Fb), vq(Dpb), vq(Dmb), kL, kR, lb, rb);
    else Adjoint::wb_face_reverse<1>(vq(qL), vq(qR), p, vq(Fb), vq(Dpb), vq(Dmb), kL, kR, lb, rb);
    for (int n = 0; n < 4; ++n) {
        qLb[n] = lb(n, 0);
        qRb[n] = rb(n, 0);
    }
}

// apply_implicit_source (make_rhs.H:429-509) on V with the stage-start aux
// of U: Model source and Jacobians, I - dt J, the oracle's elimination.
int hz_implicit(const double* Vp, const double* Up, double nn, double dt, const double* pp, double* W) {
    const VP p = vp(pp);
    const VQ V = vq(Vp), U = vq(Up);
    const VA a = aux_of(U, nn, p);
    const VQ S = Model::source(V, a, p, 0.0, VX{});
    const auto J = Model::source_jacobian_wrt_variables(V, a, p);
    const auto JA = Model::source_jacobian_wrt_aux_variables(V, a, p);
    const auto AQ = Model::update_aux_variables_jacobian_wrt_variables(V, a, p, 0.0, VX{});
    amrex::SmallMatrix<double, 4, 4> A;
    amrex::SmallMatrix<double, 4, 1> rhs, dq;
    for (int r = 0; r < 4; ++r) {
        rhs(r, 0) = dt * S(r, 0);
        for (int c = 0; c < 4; ++c) {
            double chain = 0.0;
            for (int k = 0; k < Model::n_dof_qaux; ++k) chain += JA(r * Model::n_dof_qaux + k, 0) * AQ(k * 4 + c, 0);
            A(r, c) = (r == c ? 1.0 : 0.0) - dt * (J(r * 4 + c, 0) + chain);
        }
    }
    const bool ok = zorc::ge_solve<4>(A, rhs, dq);
    for (int n = 0; n < 4; ++n) W[n] = V(n, 0) + (ok ? dq(n, 0) : 0.0);
    return ok ? 1 : 0;
}

// the closed form the adjoint differentiates
void hz_manning(const double* Vp, const double* Up, double nn, double dt, const double* pp, double* W) {
    const VP p = vp(pp);
    const double a = Model::zamr_hinv(Up[1], p(5, 0));
    double w0, w1;
    Adjoint::manning_step<double>(Vp[1], Vp[2], Vp[3], a, nn, dt, p, w0, w1);
    W[0] = Vp[0];
    W[1] = Vp[1];
    W[2] = w0;
    W[3] = w1;
}

void hz_source_reverse(const double* Up, const double* Vp, double nn, double dt, const double* pp, const double* Wb,
                       double* Vb, double* Ub, double* nb) {
    VQ vb, ub;
    Adjoint::source_reverse(vq(Up), vq(Vp), nn, dt, vp(pp), vq(Wb), vb, ub, *nb);
    for (int n = 0; n < 4; ++n) {
        Vb[n] = vb(n, 0);
        Ub[n] = ub(n, 0);
    }
}

double hz_hinv(double h, double eps) { return Model::zamr_hinv(h, eps); }
double hz_dhinv(double h, double eps) { return Adjoint::dhinv_dh(h, eps); }

void hz_bc(int tag, int dir, double side, const double* q, double* out) {
    VN nh{};
    nh(dir, 0) = side;
    const VP p{};
    const VQ r = Model::boundary_conditions(tag, vq(q), VA{}, nh, VX{}, 0.0, 1.0);
    for (int n = 0; n < 4; ++n) out[n] = r(n, 0);
}
void hz_bc_reverse(int tag, int dir, const double* q, double* out) {
    const VQ r = Adjoint::bc_reverse(tag, dir, vq(q));
    for (int n = 0; n < 4; ++n) out[n] = r(n, 0);
}
}
