// Shared device helpers and internal launch descriptors for the his_amd kernels (gfx950).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <type_traits>
#include "../../include/his_amd.h"

#define HIS_SQRT2 1.4142135623730951
#define HIS_TINY 1e-14

namespace his {

// error text for his_last_error()
void set_error(const char* fmt, ...);
int hip_fail(hipError_t e, const char* where);

#define HIS_HIP(call)                                   \
    do {                                                \
        hipError_t _e = (call);                         \
        if (_e != hipSuccess) return his::hip_fail(_e, #call); \
    } while (0)

// Reductions across one 64-lane wavefront.  The butterfly order is fixed, so a float64 sum
// over the same lane assignment is bit-reproducible from run to run.
template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}
template <typename T>
__device__ __forceinline__ T wave_max(T v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        T w = __shfl_xor(v, o, 64);
        v = w > v ? w : v;
    }
    return v;
}

template <typename T>
__device__ __forceinline__ T tmin(T a, T b) { return a < b ? a : b; }   // torch.minimum for ordered values
template <typename T>
__device__ __forceinline__ T tmax(T a, T b) { return a > b ? a : b; }
template <typename T>
__device__ __forceinline__ T relu(T a) { return a > T(0) ? a : T(0); }
template <typename T>
__device__ __forceinline__ T clamp0(T a) { return a < T(0) ? T(0) : a; }

// Per-member float64 series (his_series), evaluated with the host's arithmetic.
__device__ __forceinline__ double series_eval(const his_series& s, int b, double t) {
    switch (s.kind) {
    case 0:
        return s.vals ? s.vals[b] : 0.0;
    case 1: {
        const double* q = s.vals + (int64_t)b * s.K;
        double x = (t - s.t0) / s.step;
        if (x <= 0.0) return q[0];
        if (x >= (double)(s.K - 1)) return q[s.K - 1];
        int k = (int)x;
        double w = x - (double)k;
        return (1.0 - w) * q[k] + w * q[k + 1];
    }
    case 2: {
        int k = (int)floor((t - s.t0) / s.step + 1e-9);
        if (k < 0 || k >= s.K) return 0.0;
        return s.vals[(int64_t)b * s.K + k];
    }
    case 3: {
        double tt = t + s.shift[b];
        double eta = 0.0;
        for (int c = 0; c < s.K; ++c) eta = eta + s.amp[c] * cos(s.w[c] * tt - s.phase[c]);
        return s.base[b] + (s.on[b] ? eta : 0.0);
    }
    default:
        return 0.0;
    }
}

inline size_t align_up(size_t x, size_t a = 256) { return (x + a - 1) / a * a; }

}  // namespace his
