// Per-voxel MLEM updates (src/petl_server.py:1586-1602), shared by the update kernels and the batched
// iterations, so that every launch form computes the same bits.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include "priors_dev.h"

namespace petl {

// f <- f * d / S, with d / S = 0 where S = 0 (dataCube DIV, src/data_cube.cpp:144-149)
__device__ __forceinline__ void mlem_update_at(float* __restrict__ f, const float* __restrict__ d,
                                               const float* __restrict__ S, int64_t i)
{
    const float s = S[i];
    const float q = (s == 0.0f) ? 0.0f : d[i] / s;
    f[i] = f[i] * q;
}

// one-step-late prior: fo <- f * d / max(0.1 S, S + grad R(f)) at voxel (i, j, k) of an [N1][N2][N3] image
template <int KIND>
__device__ __forceinline__ void mlem_update_prior_at(float* __restrict__ fo, const float* __restrict__ f,
                                                     const float* __restrict__ d, const float* __restrict__ S, int i,
                                                     int j, int k, int N1, int N2, int N3, float delta, float beta,
                                                     const HuberParams& h)
{
#pragma clang fp contract(off)
    const int64_t v = (int64_t(i) * N2 + j) * N3 + k;
    const float s = S[v];
    float G = prior_grad_at<KIND>(f, i, j, k, N1, N2, N3, delta, beta, h);
    G = G + s;
    G = fmaxf(0.1f * s, G);
    const float q = (G == 0.0f) ? 0.0f : d[v] / G;
    fo[v] = f[v] * q;
}

}  // namespace petl
