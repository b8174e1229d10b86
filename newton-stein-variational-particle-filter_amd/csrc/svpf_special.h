// svpf_special.h -- special functions shared by the kernels and the host layer.
#ifndef SVPF_SPECIAL_H
#define SVPF_SPECIAL_H

#include <hip/hip_runtime.h>
#include <math.h>

// psi(x) for x > 0: the recurrence psi(x) = psi(x + 1) - 1/x lifts x to >= 6 in
// at most six steps, then the asymptotic series
// ln x - 1/(2x) - sum_k B_2k / (2k x^2k), k = 1..5 (Abramowitz & Stegun 6.3.18).
// The step count is fixed, so a non-finite or non-positive argument (a diverged
// degrees-of-freedom parameter) returns NaN instead of looping.
__host__ __device__ inline float svpf_digamma(float x) {
    if (!(x > 0.0f) || !(x <= 3.0e38f)) return __builtin_nanf("");
    float r = 0.0f;
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        if (x < 6.0f) {
            r -= 1.0f / x;
            x += 1.0f;
        }
    }
    const float f = 1.0f / (x * x);
    const float tail = f * (1.0f / 12.0f - f * (1.0f / 120.0f - f * (1.0f / 252.0f - f * (1.0f / 240.0f - f / 132.0f))));
    return r + logf(x) - 0.5f / x - tail;
}

#endif
