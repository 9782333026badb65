// Host harness for the CPU tests of the level-2 shallow-moment system
// (systems/sme2d): flux, quasilinear and non-conservative matrices, the
// spectral-radius bound and the path-integral fluctuations through a C ABI.
// Built by the tests themselves with g++.
#include "Model.H"
#include "Numerics.H"

namespace {
using VQ = Model::VQ;
using VA = Model::VA;
using VP = Model::VP;
using VN = Model::VN;
VQ vq(const double* a) {
    VQ v;
    for (int n = 0; n < 8; ++n) v(n, 0) = a[n];
    return v;
}
VP vp(const double* a) {
    VP v;
    for (int n = 0; n < 9; ++n) v(n, 0) = a[n];
    return v;
}
VA aux(const VQ& q, const VP& p) { return Model::update_aux_variables(q, VA{}, p, 0.0, Model::VX{}); }
}  // namespace

extern "C" {
void sz_flux(const double* q, int d, const double* p, double* out) {
    const VQ Q = vq(q);
    const VQ f = Model::zamr_flux_dir(Q, aux(Q, vp(p)), vp(p), d);
    for (int n = 0; n < 8; ++n) out[n] = f(n, 0);
}
void sz_quasi(const double* q, int d, const double* p, double* out) {
    const VQ Q = vq(q);
    Model::zamr_quasilinear_dir(Q, aux(Q, vp(p)), vp(p), d, out);
}
void sz_ncp(const double* q, int d, const double* p, double* out) {
    const VQ Q = vq(q);
    Model::zamr_ncp_dir(Q, aux(Q, vp(p)), vp(p), d, out);
}
double sz_speed(const double* q, double n0, double n1, const double* p) {
    const VQ Q = vq(q);
    VN n;
    n(0, 0) = n0;
    n(1, 0) = n1;
    return Numerics::local_max_abs_eigenvalue(Q, aux(Q, vp(p)), vp(p), n)(0, 0);
}
void sz_fluct(const double* qL, const double* qR, int d, const double* p, double* out) {
    const VQ L = vq(qL), R = vq(qR);
    VN n{};
    n(d, 0) = 1.0;
    const auto D = Numerics::numerical_fluctuations(L, R, aux(L, vp(p)), aux(R, vp(p)), vp(p), n);
    for (int k = 0; k < 16; ++k) out[k] = D(k, 0);
}
void sz_source(const double* q, const double* p, double* S, double* J) {
    const VQ Q = vq(q);
    const VA a = aux(Q, vp(p));
    const VQ s = Model::source(Q, a, vp(p), 0.0, Model::VX{});
    const auto j = Model::source_jacobian_wrt_variables(Q, a, vp(p));
    for (int n = 0; n < 8; ++n) S[n] = s(n, 0);
    for (int k = 0; k < 64; ++k) J[k] = j(k, 0);
}
}
