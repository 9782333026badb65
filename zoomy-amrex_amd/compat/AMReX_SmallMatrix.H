#pragma once
// Minimal stand-in for the AMReX names that the zoomy-amrex header printer
// emits into Model.H / Numerics.H (to_amrex.py:321-361): amrex::Real,
// amrex::SmallMatrix<T,N,M> with operator()(i,j) and value-initialisation,
// amrex::max/min, amrex::Math::powi<N>, and the AMREX_GPU_* macros.
//
// The same header serves the hipcc (device + host) build and the g++ CPU
// build, so a model header compiles unchanged for both.  Only the IEEE basic
// operations (+ - * / sqrt) are used by the helpers here, which keeps CPU and
// GPU results bit-identical when FMA contraction is off.

#include <cmath>
#include <cstddef>

#if defined(__HIPCC__) || defined(__HIP__)
#include <hip/hip_runtime.h>
#define AMREX_GPU_HOST_DEVICE __host__ __device__
#define AMREX_GPU_DEVICE __device__
#define AMREX_GPU_HOST __host__
#else
#define AMREX_GPU_HOST_DEVICE
#define AMREX_GPU_DEVICE
#define AMREX_GPU_HOST
#endif
#define AMREX_FORCE_INLINE inline __attribute__((always_inline))

namespace amrex {

using Real = double;

// Dense N x M matrix stored row-major in a plain array.  `SmallMatrix<...> m{}`
// zero-initialises (aggregate value-init), `SmallMatrix<...> m;` leaves it
// uninitialised, exactly like the AMReX type the printer targets.
template <class T, int N, int M>
struct SmallMatrix {
    T m_mat[N * M > 0 ? N * M : 1];

    AMREX_GPU_HOST_DEVICE AMREX_FORCE_INLINE T& operator()(int i, int j) noexcept {
        return m_mat[i * M + j];
    }
    AMREX_GPU_HOST_DEVICE AMREX_FORCE_INLINE const T& operator()(int i, int j) const noexcept {
        return m_mat[i * M + j];
    }
    AMREX_GPU_HOST_DEVICE AMREX_FORCE_INLINE T& operator[](int k) noexcept { return m_mat[k]; }
    AMREX_GPU_HOST_DEVICE AMREX_FORCE_INLINE const T& operator[](int k) const noexcept {
        return m_mat[k];
    }
    static constexpr int row_size = N;
    static constexpr int column_size = M;
};

// std::max / std::min semantics: (a < b) ? b : a and (b < a) ? b : a.
template <class T>
AMREX_GPU_HOST_DEVICE AMREX_FORCE_INLINE constexpr T max(T a, T b) noexcept {
    return (a < b) ? b : a;
}
template <class T>
AMREX_GPU_HOST_DEVICE AMREX_FORCE_INLINE constexpr T min(T a, T b) noexcept {
    return (b < a) ? b : a;
}

namespace Math {
// Integer power by repeated squaring, the same association AMReX uses:
// powi<4>(x) = (x*x)*(x*x), powi<3>(x) = x*(x*x).
template <int P, class T>
AMREX_GPU_HOST_DEVICE AMREX_FORCE_INLINE constexpr T powi(T x) noexcept {
    if constexpr (P < 0) {
        return T(1) / powi<-P>(x);
    } else if constexpr (P == 0) {
        return T(1);
    } else if constexpr (P == 1) {
        return x;
    } else if constexpr (P == 2) {
        return x * x;
    } else if constexpr (P % 2 == 0) {
        return powi<2>(powi<P / 2>(x));
    } else {
        return x * powi<P - 1>(x);
    }
}
template <class T>
AMREX_GPU_HOST_DEVICE AMREX_FORCE_INLINE T abs(T x) noexcept { return x < T(0) ? -x : x; }
}  // namespace Math

}  // namespace amrex
