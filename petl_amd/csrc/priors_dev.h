// Device functions of the relative-difference and p-Huber TV prior gradients, shared by the
// stand-alone gradient kernels (priors.hip) and the fused MLEM update (recon.hip).
//
// Reference: src/relative_differences.cu:21-28,107-173 and src/total_variation.cu:33-78,434-572.
// The 26 neighbours are clamped at the borders and summed in the reference's grouping (6 face, 12 edge,
// 8 corner neighbours, weighted 1, 1/sqrt2, 1/sqrt3) with contraction off.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace petl {

struct HuberParams {
    float p, delta, slope;
};

__device__ __forceinline__ float rd_clip(float v) { return v < 0.00000001f ? 0.0f : v; }

__device__ __forceinline__ float drd(float x, float y, float delta)
{
#pragma clang fp contract(off)
    if (x == 0.0f && y == 0.0f)
        return 0.0f;
    const float ad = fabsf(x - y);
    const float den = x + y + delta * ad;
    return (x - y) * (delta * ad + x + 3.0f * y) / (den * den);
}

__device__ __forceinline__ float dhuber(float x, const HuberParams& h)
{
#pragma clang fp contract(off)
    if (fabsf(x) <= h.delta)
        return x;
    return h.p * h.slope * powf(fabsf(x), h.p) / x;
}

// Gradient at (i,j,k) of an [N1][N2][N3] array.  kind 0: relative differences (with the f < 1e-8 clip of
// src/relative_differences.cu:107-115 applied to every value read); kind 1: p-Huber TV.
// `ld(index)` returns the value at a linear index (lets the image sit in any storage format).
template <int KIND, typename Load>
__device__ __forceinline__ float prior_grad_at_ld(const Load& ld, int i, int j, int k, int N1, int N2, int N3,
                                                  float delta, float beta, const HuberParams& h)
{
#pragma clang fp contract(off)
    const int im = max(0, i - 1), ip = min(N1 - 1, i + 1);
    const int jm = max(0, j - 1), jp = min(N2 - 1, j + 1);
    const int km = max(0, k - 1), kp = min(N3 - 1, k + 1);
    const float d1 = 1.0f * beta, d2 = 0.7071067811865475f * beta, d3 = 0.5773502691896258f * beta;
    auto F = [&](int a, int b, int c) {
        const float v = ld((int64_t(a) * N2 + b) * N3 + c);
        return KIND == 0 ? rd_clip(v) : v;
    };
    const float c0 = F(i, j, k);
    auto D = [&](int a, int b, int c) {
        return KIND == 0 ? drd(c0, F(a, b, c), delta) : dhuber(c0 - F(a, b, c), h);
    };
    if (KIND == 1 && N1 == 1) {
        return (D(i, jp, k) + D(i, jm, k) + D(i, j, kp) + D(i, j, km)) * d1 +
               (D(i, jp, kp) + D(i, jp, km) + D(i, jm, kp) + D(i, jm, km)) * d2;
    }
    const float s1 = D(ip, j, k) + D(im, j, k) + D(i, jp, k) + D(i, jm, k) + D(i, j, kp) + D(i, j, km);
    const float s2 = D(ip, jp, k) + D(ip, jm, k) + D(ip, j, kp) + D(ip, j, km) + D(im, jp, k) + D(im, jm, k) +
                     D(im, j, kp) + D(im, j, km) + D(i, jp, kp) + D(i, jp, km) + D(i, jm, kp) + D(i, jm, km);
    const float s3 = D(ip, jp, kp) + D(ip, jp, km) + D(ip, jm, kp) + D(ip, jm, km) + D(im, jp, kp) +
                     D(im, jp, km) + D(im, jm, kp) + D(im, jm, km);
    return s1 * d1 + s2 * d2 + s3 * d3;
}

template <int KIND>
__device__ __forceinline__ float prior_grad_at(const float* __restrict__ f, int i, int j, int k, int N1, int N2,
                                               int N3, float delta, float beta, const HuberParams& h)
{
    return prior_grad_at_ld<KIND>([f](int64_t x) { return f[x]; }, i, j, k, N1, N2, N3, delta, beta, h);
}

// TV_gradient's parameter rules (src/total_variation.cu:574-594): beta <= 0 -> 1, delta >= 1e-8,
// slope = delta^(2-p)/p in double.
inline HuberParams huber_params(float& delta, float& beta, float p)
{
    if (beta <= 0.0f)
        beta = 1.0f;
    if (delta < 1.0e-8f)
        delta = 1.0e-8f;
    HuberParams h;
    h.p = p;
    h.delta = delta;
    h.slope = float(pow(double(delta), 2.0 - double(p)) / double(p));
    return h;
}

}  // namespace petl
