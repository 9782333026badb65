// Bandwidth calibration for the stage kernels' roofline: the HBM rate their
// memory access pattern reaches with no arithmetic, beside a plain copy.
//
// pattern_kernel moves exactly the bytes of an order-2 stage over one n x n
// box of n_q SoA rows (2 ghost cells, Array4 layout as the solver stores it)
// in the stage kernel's decomposition: a wave per strip of 60 columns and 32
// rows, its 64 lanes on 64 consecutive columns (the strip and a 2-column halo
// each side), marching up the rows; per row it loads the 64 columns of every
// component of the stage input (the leading row, one row ahead, as the stage
// kernel prefetches it), with `heun` also the strip's 60 columns of Q^n, and
// stores the strip's 60 columns of every component.  Items are dealt to the
// XCDs as the stage kernel deals them.  copy_kernel copies one buffer to
// another with 32-byte accesses, grid-stride.  Neither computes anything:
// they bound what the stage kernels' traffic can reach on this device.
#include <hip/hip_runtime.h>

#include <cstddef>

namespace {

constexpr int kWave = 64, kHalo = 2, kStrip = kWave - 2 * kHalo, kRows = 32;

__device__ __forceinline__ int remap(int b, int nslots) {
    const int per = (nslots + 7) / 8;
    const int x = b & 7, k = b >> 3;
    const int it = x * per + k;
    return (it < nslots) ? it : -1;
}

template <int NQ>
__global__ void __launch_bounds__(256) pattern_kernel(const double* __restrict__ src, const double* __restrict__ qn,
                                                      double* __restrict__ dst, int n, int heun, int nslots,
                                                      int nstrip, int nitems) {
    const int slot = remap(blockIdx.x, nslots);
    if (slot < 0) return;
    const int it = slot * 4 + (threadIdx.x >> 6);
    if (it >= nitems) return;
    const int t = threadIdx.x & (kWave - 1);
    const int s = n + 2 * kHalo;  // row stride of the grown box
    const long long sc = (long long)s * s;
    const int i0 = (it % nstrip) * kStrip, j0 = (it / nstrip) * kRows;
    const int c = min(i0 - kHalo + t, n - 1 + kHalo);  // column, clamped into the grown box
    const bool own = t >= kHalo && t < kHalo + kStrip && i0 - kHalo + t <= n - 1;
    const int j1 = min(j0 + kRows, n);
    // the leading row of row j is j + 2 (clamped into the grown box)
    auto lead = [&](int j) { return (long long)(c + kHalo) + (long long)(min(j + 2, n - 1 + kHalo) + kHalo) * s; };
    double v[NQ];
    const long long l0 = lead(j0);
#pragma unroll
    for (int q = 0; q < NQ; ++q) v[q] = src[l0 + q * sc];
    for (int j = j0; j < j1; ++j) {
        double w[NQ];
        const long long l1 = lead(j + 1);
#pragma unroll
        for (int q = 0; q < NQ; ++q) w[q] = src[l1 + q * sc];  // next leading row
        const long long idx = (long long)(c + kHalo) + (long long)(j + kHalo) * s;
        if (own) {
#pragma unroll
            for (int q = 0; q < NQ; ++q) {
                double x = v[q];
                if (heun) x += qn[idx + q * sc];
                dst[idx + q * sc] = x;
            }
        }
#pragma unroll
        for (int q = 0; q < NQ; ++q) v[q] = w[q];
    }
}

__global__ void __launch_bounds__(256) copy_kernel(const double4* __restrict__ a, double4* __restrict__ b,
                                                   long long n4) {
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n4; i += (long long)gridDim.x * blockDim.x)
        b[i] = a[i];
}

int timed(hipStream_t st, int reps, double* ms, void (*launch)(hipStream_t, void*), void* arg) {
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = 1;
    if (hipEventCreate(&e0) == hipSuccess && hipEventCreate(&e1) == hipSuccess) {
        launch(st, arg);  // warm-up
        float f = 0.f;
        if (hipEventRecord(e0, st) == hipSuccess) {
            for (int r = 0; r < reps; ++r) launch(st, arg);
            if (hipEventRecord(e1, st) == hipSuccess && hipEventSynchronize(e1) == hipSuccess &&
                hipEventElapsedTime(&f, e0, e1) == hipSuccess && hipGetLastError() == hipSuccess) {
                *ms = f / reps;
                rc = 0;
            }
        }
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    return rc;
}

struct PatternArgs {
    const double *src, *qn;
    double* dst;
    int n, nq, heun, nslots, nstrip, nitems;
};
struct CopyArgs {
    const double4* a;
    double4* b;
    long long n4;
};

}  // namespace

extern "C" {

// Average milliseconds of one pattern launch over `reps` (after one warm-up)
// on an n x n box with nq = 4 or 8 rows; heun = 1 adds the stage-2 Q^n reads.
// Buffers are allocated and freed here.  Returns 0, or 1 on any HIP error or
// bad argument.
int zamr_probe_pattern(int n, int nq, int heun, int reps, double* ms) {
    if (n < 64 || (nq != 4 && nq != 8) || reps < 1 || !ms) return 1;
    const size_t s = (size_t)n + 2 * kHalo;
    const size_t bytes = s * s * (size_t)nq * sizeof(double);
    double *a = nullptr, *b = nullptr, *c = nullptr;
    int rc = 1;
    if (hipMalloc(&a, bytes) == hipSuccess && hipMalloc(&b, bytes) == hipSuccess && hipMalloc(&c, bytes) == hipSuccess &&
        hipMemset(a, 0, bytes) == hipSuccess && hipMemset(b, 0, bytes) == hipSuccess &&
        hipMemset(c, 0, bytes) == hipSuccess) {
        PatternArgs p;
        p.src = a;
        p.qn = b;
        p.dst = c;
        p.n = n;
        p.nq = nq;
        p.heun = heun;
        p.nstrip = (n + kStrip - 1) / kStrip;
        p.nitems = p.nstrip * ((n + kRows - 1) / kRows);
        p.nslots = (((p.nitems + 3) / 4 + 7) / 8) * 8;
        rc = timed(
            nullptr, reps, ms,
            [](hipStream_t st, void* v) {
                const PatternArgs& q = *static_cast<PatternArgs*>(v);
                if (q.nq == 4)
                    hipLaunchKernelGGL(pattern_kernel<4>, dim3(q.nslots), dim3(256), 0, st, q.src, q.qn, q.dst, q.n,
                                       q.heun, q.nslots, q.nstrip, q.nitems);
                else
                    hipLaunchKernelGGL(pattern_kernel<8>, dim3(q.nslots), dim3(256), 0, st, q.src, q.qn, q.dst, q.n,
                                       q.heun, q.nslots, q.nstrip, q.nitems);
            },
            &p);
    }
    if (a) (void)hipFree(a);
    if (b) (void)hipFree(b);
    if (c) (void)hipFree(c);
    return rc;
}

// Average milliseconds of one plain device copy of `bytes` (a multiple of 32).
int zamr_probe_copy(long long bytes, int reps, double* ms) {
    if (bytes < 32 || bytes % 32 || reps < 1 || !ms) return 1;
    double4 *a = nullptr, *b = nullptr;
    int rc = 1;
    if (hipMalloc(&a, bytes) == hipSuccess && hipMalloc(&b, bytes) == hipSuccess &&
        hipMemset(a, 0, bytes) == hipSuccess && hipMemset(b, 0, bytes) == hipSuccess) {
        CopyArgs p{a, b, bytes / 32};
        rc = timed(
            nullptr, reps, ms,
            [](hipStream_t st, void* v) {
                const CopyArgs& q = *static_cast<CopyArgs*>(v);
                hipLaunchKernelGGL(copy_kernel, dim3(256 * 64), dim3(256), 0, st, q.a, q.b, q.n4);
            },
            &p);
    }
    if (a) (void)hipFree(a);
    if (b) (void)hipFree(b);
    return rc;
}

}  // extern "C"
