#pragma once
// sqrt(x), a / sqrt(x), 1/x and x^(-1/3) for the model headers.
//
// Host code and the bitwise build (ZAMR_BITWISE) use std::sqrt and a true
// division, correctly rounded, so that the CPU oracle and the bitwise GPU
// build agree bit for bit.
//
// Device code of the throughput build starts from the hardware estimate y of
// 1/sqrt(x) (v_rsq_f64, relative error near 2^-23):
//   sqrt:     one coupled Newton step on s = x y and h = y/2
//             (r = 1/2 - s h; s += s r; h += h r), then s += (x - s^2) h;
//             y is capped at 2^600 so that x = 0 gives 0;
//   a/sqrt:   two Newton steps y += y (1/2 - (x/2) y^2), then a y.
// No intermediate squares y^2, so every finite x >= 0, subnormals included,
// stays in range; the results land within an ulp or two of the correctly
// rounded values in under half the instructions.  Negative and NaN
// arguments give NaN as std::sqrt does; an infinite argument gives NaN where
// std::sqrt gives inf (a state that has blown up either way).
#include <cmath>
#include "AMReX_SmallMatrix.H"

namespace zamr_math {

AMREX_GPU_HOST_DEVICE AMREX_FORCE_INLINE double sqrt(double x) {
#if defined(__HIP_DEVICE_COMPILE__) && !defined(ZAMR_BITWISE)
    const double y = fmin(__builtin_amdgcn_rsq(x), 0x1p600);
    double s = x * y, h = 0.5 * y;
    const double r = fma(-s, h, 0.5);
    s = fma(s, r, s);
    h = fma(h, r, h);
    return fma(fma(-s, s, x), h, s);
#else
    return std::sqrt(x);
#endif
}

AMREX_GPU_HOST_DEVICE AMREX_FORCE_INLINE double div_sqrt(double a, double x) {
#if defined(__HIP_DEVICE_COMPILE__) && !defined(ZAMR_BITWISE)
    double y = __builtin_amdgcn_rsq(x);
    const double hx = 0.5 * x;
    y = fma(y, fma(-hx * y, y, 0.5), y);
    y = fma(y, fma(-hx * y, y, 0.5), y);
    return a * y;
#else
    return a / std::sqrt(x);
#endif
}

// 1/x, for x in the normal range: on the device of the throughput build the
// hardware estimate (v_rcp_f64) refined by two Newton steps (within an ulp
// or so of the correctly rounded quotient); elsewhere a true division.
AMREX_GPU_HOST_DEVICE AMREX_FORCE_INLINE double recip(double x) {
#if defined(__HIP_DEVICE_COMPILE__) && !defined(ZAMR_BITWISE)
    double r = __builtin_amdgcn_rcp(x);
    double e = fma(-x, r, 1.0);
    r = fma(r, e, r);
    e = fma(-x, r, 1.0);
    return fma(r, e, r);
#else
    return 1.0 / x;
#endif
}

// x^(-1/3) for x > 0 (normal range) from IEEE + - * only: the high word's
// exponent-and-mantissa bits divided by 3 and subtracted from a magic
// constant give a seed within 3.5% (the low word moves it by under 2^-20),
// then four Newton steps on y^-3 = x, y <- y (4 - x y^3) / 3.  The same
// steps on host and device, so the two agree bit for bit.  (A two-step
// series correction in e = 1 - x y^3 takes 14 operations instead of 20; it
// was measured slower in the stage kernels, whose register budget its four
// 64-bit constants strain.)
AMREX_GPU_HOST_DEVICE AMREX_FORCE_INLINE double rcbrt(double x) {
    unsigned long long bits;
    __builtin_memcpy(&bits, &x, sizeof(bits));
    const unsigned int hi = static_cast<unsigned int>(bits >> 32);
    bits = static_cast<unsigned long long>(0x553ef0ffu - hi / 3u) << 32;
    double y;
    __builtin_memcpy(&y, &bits, sizeof(y));
    const double third = 1.0 / 3.0;
    for (int it = 0; it < 4; ++it) {
        const double y3 = y * y * y;
        y = y * (4.0 - x * y3) * third;
    }
    return y;
}

// Arithmetic that leaves out terms with a structurally zero operand.  A value
// the compiler knows to be the constant zero after inlining (an entry of a
// model matrix that is never set, a component of an axis normal) makes a
// term that is an exact zero for finite operands, which IEEE rules forbid
// the compiler to drop (0 x is NaN for infinite x, 0 + x is +0 for x = -0):
// a printed model never contains such terms, and these helpers omit them.
// For finite values the results equal the plain expressions (up to the sign
// of a zero), on host and device alike.
AMREX_GPU_HOST_DEVICE AMREX_FORCE_INLINE bool known_zero(double v) { return __builtin_constant_p(v) && v == 0.0; }
// acc + x
AMREX_GPU_HOST_DEVICE AMREX_FORCE_INLINE double add(double acc, double x) {
    if (known_zero(x)) return acc;
    if (known_zero(acc)) return x;
    return acc + x;
}
// a b
AMREX_GPU_HOST_DEVICE AMREX_FORCE_INLINE double mul(double a, double b) {
    if (known_zero(a) || known_zero(b)) return 0.0;
    return a * b;
}
// acc + a b
AMREX_GPU_HOST_DEVICE AMREX_FORCE_INLINE double madd(double acc, double a, double b) {
    if (known_zero(a) || known_zero(b)) return acc;
    return add(acc, a * b);
}

// Projections on a face or boundary normal.  Every normal the driver hands
// the model (make_rhs.H's n_hat, the boundary normals of FillPhysicalBC) is a
// unit axis vector, so a component that is zero contributes nothing: it is
// skipped rather than multiplied, and where the normal is known at compile
// time (the stage kernels) no multiplication by zero is left.  For finite
// operands the value is a * nx + b * ny (up to the sign of a zero).
AMREX_GPU_HOST_DEVICE AMREX_FORCE_INLINE double axis_dot(double a, double b, double nx, double ny) {
    if (ny == 0.0) return a * nx;
    if (nx == 0.0) return b * ny;
    return a * nx + b * ny;
}
// x * n_d, zero where the component n_d is zero.
AMREX_GPU_HOST_DEVICE AMREX_FORCE_INLINE double axis_scale(double x, double nd) { return nd == 0.0 ? 0.0 : x * nd; }

// Keep the device compiler from moving work across this point (an empty asm
// statement that "redefines" x): used to stop it from interleaving two large
// independent evaluations, which costs registers a kernel at one wave per
// SIMD cannot spare.  No effect on values; nothing on the host.
AMREX_GPU_HOST_DEVICE AMREX_FORCE_INLINE void sequence_point(double& x) {
#if defined(__HIP_DEVICE_COMPILE__)
    asm volatile("" : "+v"(x));
#else
    (void)x;
#endif
}

}  // namespace zamr_math
