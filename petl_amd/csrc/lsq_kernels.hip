// Kernels of the least-squares reconstructions: Gaussian smoothing, TV cost / quadratic form, FP64 inner
// products and the conjugate-gradient vector updates.
//
// Reference: src/relative_differences.cu:314-420, 753-800 (Gaussian filter), src/total_variation.cu:21-78,
// 80-417, 662-800 (Huber, cost and quadratic-form terms and their sums), src/data_cube.cpp:249-254, 320-397
// (reciprocal, inner products) and the numpy vector arithmetic of src/petl_server.py:1691-1877.
//
// The reference smooths with one thread per output looping over the whole (2R+1)^d window; here the window
// is applied one axis at a time (the weight exp(-den (s_j^2 + s_k^2)) factors and the borders clamp per
// axis, so the passes give the same normalised sum), which reads 2R+1 instead of (2R+1)^d samples per
// output.  Reductions are two-pass with per-block FP64 partials summed in a fixed order, so every result is
// reproducible bit for bit.
#include <cmath>
#include "lsq_kernels.h"

namespace petl {

namespace {

constexpr int VT = 256;

inline int grid_for(int64_t n) { return int(std::min<int64_t>((n + VT - 1) / VT, 65536)); }

// ---------------------------------------------------------------------------------------------
// Gaussian smoothing along one axis of an [outer][n][inner] array
// ---------------------------------------------------------------------------------------------
template <bool WEIGHTS>
__global__ void __launch_bounds__(VT) gauss_axis_kernel(const float* __restrict__ in, float* __restrict__ out,
                                                        int64_t total, int n, int64_t inner, int R, float den)
{
#pragma clang fp contract(off)
    const int64_t slab = int64_t(n) * inner;
    for (int64_t e = int64_t(blockIdx.x) * VT + threadIdx.x; e < total; e += int64_t(gridDim.x) * VT) {
        const int64_t o = e / slab;
        const int64_t rem = e - o * slab;
        const int a = int(rem / inner);
        const int64_t base = o * slab + (rem - int64_t(a) * inner);
        float val = 0.0f, sum = 0.0f;
        for (int t = -R; t <= R; t++) {
            const int ai = max(0, min(a + t, n - 1));
            const int s = a - ai;
            const float w = expf(-den * float(s * s));
            val += w * in[base + int64_t(ai) * inner];
            sum += w;
        }
        float v = val / sum;
        if (WEIGHTS) {
            v = (v < 7.0f) ? 7.0f : v;  // np.clip(W, 7.0, None) keeps NaN
            v = (v == 0.0f) ? 0.0f : float(1.0 / double(v));
        }
        out[e] = v;
    }
}

void gauss_pass(const float* in, float* out, int64_t outer, int n, int64_t inner, const GaussParams& gp, bool weights,
                hipStream_t s)
{
    const int64_t total = outer * n * inner;
    if (weights)
        hipLaunchKernelGGL(gauss_axis_kernel<true>, dim3(grid_for(total)), dim3(VT), 0, s, in, out, total, n, inner,
                           gp.radius, gp.den);
    else
        hipLaunchKernelGGL(gauss_axis_kernel<false>, dim3(grid_for(total)), dim3(VT), 0, s, in, out, total, n, inner,
                           gp.radius, gp.den);
    PETL_LAUNCH_CHECK();
}

__global__ void __launch_bounds__(VT) copy_kernel(float* __restrict__ x, const float* __restrict__ y, int64_t n)
{
    for (int64_t i = int64_t(blockIdx.x) * VT + threadIdx.x; i < n; i += int64_t(gridDim.x) * VT)
        x[i] = y[i];
}

// ---------------------------------------------------------------------------------------------
// reductions
// ---------------------------------------------------------------------------------------------
__device__ __forceinline__ void block_sum_store(double acc, double* __restrict__ dst)
{
    __shared__ double red[VT];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int w = VT / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w)
            red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        *dst = red[0];
}

template <bool HAS_W>
__global__ void __launch_bounds__(VT) dot_partial_kernel(const float* __restrict__ x, const float* __restrict__ y,
                                                         const float* __restrict__ w, int64_t n,
                                                         double* __restrict__ part)
{
#pragma clang fp contract(off)
    double acc = 0.0;
    for (int64_t i = int64_t(blockIdx.x) * VT + threadIdx.x; i < n; i += int64_t(gridDim.x) * VT) {
        float p = x[i] * y[i];
        if (HAS_W)
            p = p * w[i];
        acc += double(p);
    }
    block_sum_store(acc, part + blockIdx.x);
}

__global__ void __launch_bounds__(VT) sum_final_kernel(const double* __restrict__ part, int nparts,
                                                       double* __restrict__ out)
{
    double acc = 0.0;
    for (int i = threadIdx.x; i < nparts; i += VT)
        acc += part[i];
    block_sum_store(acc, out);
}

// p-Huber of src/total_variation.cu:33-78 (USE_P_HUBER)
struct HuberFull {
    float p, delta, slope, shift;
};

__device__ __forceinline__ float huber_h(float x, const HuberFull& h)
{
#pragma clang fp contract(off)
    if (fabsf(x) <= h.delta)
        return 0.5f * x * x;
    return h.slope * powf(fabsf(x), h.p) + h.shift;
}

__device__ __forceinline__ float huber_dd(float x, const HuberFull& h)
{
#pragma clang fp contract(off)
    if (fabsf(x) <= h.delta)
        return 1.0f;
    return h.p * h.slope * powf(fabsf(x), h.p - 2.0f);
}

// cost term (QUAD false) or quadratic-form term (QUAD true) at (i, j, k): the neighbour sums of
// aTV_Huber_costTerm / aTV_Huber_quadFormTerm in their grouping (6 face, 12 edge, 8 corner neighbours)
template <bool QUAD>
__device__ __forceinline__ float tv_term_at(const float* __restrict__ f, const float* __restrict__ d, int i, int j,
                                            int k, int N1, int N2, int N3, float beta, const HuberFull& h)
{
#pragma clang fp contract(off)
    const int im = max(0, i - 1), ip = min(N1 - 1, i + 1);
    const int jm = max(0, j - 1), jp = min(N2 - 1, j + 1);
    const int km = max(0, k - 1), kp = min(N3 - 1, k + 1);
    const float d1 = 1.0f * beta, d2 = 0.7071067811865475f * beta, d3 = 0.5773502691896258f * beta;
    auto at = [&](int a, int b, int c) { return (int64_t(a) * N2 + b) * N3 + c; };
    const int64_t c0i = at(i, j, k);
    const float c0 = f[c0i];
    const float cd = QUAD ? d[c0i] : 0.0f;
    auto T = [&](int a, int b, int c) {
        const int64_t o = at(a, b, c);
        if (QUAD) {
            const float t = cd - d[o];
            return huber_dd(c0 - f[o], h) * (t * t);
        }
        return huber_h(c0 - f[o], h);
    };
    if (N1 == 1)
        return (T(i, jp, k) + T(i, jm, k) + T(i, j, kp) + T(i, j, km)) * d1 +
               (T(i, jp, kp) + T(i, jp, km) + T(i, jm, kp) + T(i, jm, km)) * d2;
    const float s1 = T(ip, j, k) + T(im, j, k) + T(i, jp, k) + T(i, jm, k) + T(i, j, kp) + T(i, j, km);
    const float s2 = T(ip, jp, k) + T(ip, jm, k) + T(ip, j, kp) + T(ip, j, km) + T(im, jp, k) + T(im, jm, k) +
                     T(im, j, kp) + T(im, j, km) + T(i, jp, kp) + T(i, jp, km) + T(i, jm, kp) + T(i, jm, km);
    const float s3 = T(ip, jp, kp) + T(ip, jp, km) + T(ip, jm, kp) + T(ip, jm, km) + T(im, jp, kp) +
                     T(im, jp, km) + T(im, jm, kp) + T(im, jm, km);
    return s1 * d1 + s2 * d2 + s3 * d3;
}

template <bool QUAD>
__global__ void __launch_bounds__(VT) tv_partial_kernel(const float* __restrict__ f, const float* __restrict__ d,
                                                        int N1, int N2, int N3, float beta, HuberFull h,
                                                        double* __restrict__ part)
{
    const int64_t n = int64_t(N1) * N2 * N3;
    const int64_t plane = int64_t(N2) * N3;
    double acc = 0.0;
    for (int64_t e = int64_t(blockIdx.x) * VT + threadIdx.x; e < n; e += int64_t(gridDim.x) * VT) {
        const int i = int(e / plane);
        const int64_t r = e - int64_t(i) * plane;
        const int j = int(r / N3);
        const int k = int(r - int64_t(j) * N3);
        acc += double(tv_term_at<QUAD>(f, d, i, j, k, N1, N2, N3, beta, h));
    }
    block_sum_store(acc, part + blockIdx.x);
}

// ---------------------------------------------------------------------------------------------
// vector updates
// ---------------------------------------------------------------------------------------------
#define PETL_EW_KERNEL(name, params, body)                                                               \
    __global__ void __launch_bounds__(VT) name params                                                   \
    {                                                                                                    \
        _Pragma("clang fp contract(off)") for (int64_t i = int64_t(blockIdx.x) * VT + threadIdx.x; i < n; \
                                               i += int64_t(gridDim.x) * VT) body                        \
    }

PETL_EW_KERNEL(sub_kernel, (float* __restrict__ x, const float* __restrict__ y, int64_t n), { x[i] = x[i] - y[i]; })
PETL_EW_KERNEL(add_kernel, (float* __restrict__ x, const float* __restrict__ y, int64_t n), { x[i] = x[i] + y[i]; })
PETL_EW_KERNEL(mul_to_kernel,
               (float* __restrict__ dst, const float* __restrict__ a, const float* __restrict__ b, int64_t n),
               { dst[i] = a[i] * b[i]; })
PETL_EW_KERNEL(xpay_kernel, (float* __restrict__ d, const float* __restrict__ u, float gamma, int64_t n), {
    const float t = gamma * d[i];
    d[i] = u[i] + t;
})
PETL_EW_KERNEL(step_kernel, (float* __restrict__ f, const float* __restrict__ d, float step, int clip, int64_t n), {
    const float t = step * d[i];
    float v = f[i] - t;
    if (clip && v < 0.0f)
        v = 0.0f;
    f[i] = v;
})
PETL_EW_KERNEL(resid_kernel,
               (float* __restrict__ r, const float* __restrict__ Pd, const float* __restrict__ g, float c, int64_t n),
               {
                   const float t = c * Pd[i];
                   const float v = r[i] + t;
                   r[i] = v - g[i];
               })
PETL_EW_KERNEL(recip_kernel, (float* __restrict__ x, float dz, int64_t n), {
    const float v = x[i];
    x[i] = (v == 0.0f) ? dz : float(1.0 / double(v));
})
PETL_EW_KERNEL(clip_neg_kernel, (float* __restrict__ x, int64_t n), {
    if (x[i] < 0.0f)
        x[i] = 0.0f;
})

}  // namespace

// -------------------------------------------------------------------------------------------------
GaussParams gauss_params(float FWHM)
{
    // formed in float as GaussianFilterKernel forms them
    const float sigma = FWHM / (2.0f * std::sqrt(2.0f * std::log(2.0f)));
    GaussParams gp;
    gp.radius = int(std::ceil(std::sqrt(2.0f * std::log(10.0f)) * sigma));
    gp.den = 1.0f / (2.0f * sigma);
    return gp;
}

void gaussian_filter_device(float* f, float* tmp, int N1, int N2, int N3, float FWHM, int numDims,
                            bool weights_epilogue, hipStream_t s)
{
    const GaussParams gp = gauss_params(FWHM);
    struct Pass {
        int64_t outer;
        int n;
        int64_t inner;
    };
    Pass passes[3];
    int np = 0;
    if (numDims != 1) {
        passes[np++] = {int64_t(N1) * N2, N3, 1};  // along the last axis
        passes[np++] = {int64_t(N1), N2, int64_t(N3)};
    }
    if (numDims != 2)
        passes[np++] = {1, N1, int64_t(N2) * N3};  // along the first axis
    float* src = f;
    float* dst = tmp;
    for (int q = 0; q < np; q++) {
        gauss_pass(src, dst, passes[q].outer, passes[q].n, passes[q].inner, gp, weights_epilogue && q == np - 1, s);
        std::swap(src, dst);
    }
    if (src != f) {
        const int64_t n = int64_t(N1) * N2 * N3;
        hipLaunchKernelGGL(copy_kernel, dim3(grid_for(n)), dim3(VT), 0, s, f, src, n);
        PETL_LAUNCH_CHECK();
    }
}

int64_t red_scratch_doubles(int64_t n) { return std::max<int64_t>(1, std::min<int64_t>((n + VT - 1) / VT, 4096)); }

void launch_dot(const float* x, const float* y, const float* w, int64_t n, double* scratch, double* out,
                hipStream_t s)
{
    const int nparts = int(red_scratch_doubles(n));
    if (w)
        hipLaunchKernelGGL(dot_partial_kernel<true>, dim3(nparts), dim3(VT), 0, s, x, y, w, n, scratch);
    else
        hipLaunchKernelGGL(dot_partial_kernel<false>, dim3(nparts), dim3(VT), 0, s, x, y, w, n, scratch);
    PETL_LAUNCH_CHECK();
    hipLaunchKernelGGL(sum_final_kernel, dim3(1), dim3(VT), 0, s, scratch, nparts, out);
    PETL_LAUNCH_CHECK();
}

void launch_tv_reduce(const float* f, const float* d, int N1, int N2, int N3, float delta, float beta, float p,
                      double* scratch, double* out, hipStream_t s)
{
    // parameter rules and constants of anisotropicTotalVariation_cost / _quadraticForm and
    // setConstantMemoryParameters (src/total_variation.cu:574-590, 662-680)
    if (beta <= 0.0f)
        beta = 1.0f;
    if (delta < 1.0e-8f)
        delta = 1.0e-8f;
    HuberFull h;
    h.p = p;
    h.delta = delta;
    h.slope = float(std::pow(double(delta), 2.0 - double(p)) / double(p));
    h.shift = float(double(delta * delta) * (0.5 - 1.0 / double(p)));
    const int64_t n = int64_t(N1) * N2 * N3;
    const int nparts = int(red_scratch_doubles(n));
    if (d)
        hipLaunchKernelGGL(tv_partial_kernel<true>, dim3(nparts), dim3(VT), 0, s, f, d, N1, N2, N3, beta, h, scratch);
    else
        hipLaunchKernelGGL(tv_partial_kernel<false>, dim3(nparts), dim3(VT), 0, s, f, d, N1, N2, N3, beta, h,
                           scratch);
    PETL_LAUNCH_CHECK();
    hipLaunchKernelGGL(sum_final_kernel, dim3(1), dim3(VT), 0, s, scratch, nparts, out);
    PETL_LAUNCH_CHECK();
}

#define PETL_EW_LAUNCH(kernel, n, ...)                                                      \
    do {                                                                                    \
        hipLaunchKernelGGL(kernel, dim3(grid_for(n)), dim3(VT), 0, s, __VA_ARGS__, n);      \
        PETL_LAUNCH_CHECK();                                                                \
    } while (0)

void launch_sub(float* x, const float* y, int64_t n, hipStream_t s) { PETL_EW_LAUNCH(sub_kernel, n, x, y); }
void launch_add(float* x, const float* y, int64_t n, hipStream_t s) { PETL_EW_LAUNCH(add_kernel, n, x, y); }
void launch_mul_to(float* dst, const float* a, const float* b, int64_t n, hipStream_t s)
{
    PETL_EW_LAUNCH(mul_to_kernel, n, dst, a, b);
}
void launch_xpay(float* d, const float* u, float gamma, int64_t n, hipStream_t s)
{
    PETL_EW_LAUNCH(xpay_kernel, n, d, u, gamma);
}
void launch_step(float* f, const float* d, float step, bool clip, int64_t n, hipStream_t s)
{
    PETL_EW_LAUNCH(step_kernel, n, f, d, step, clip ? 1 : 0);
}
void launch_resid_update(float* r, const float* Pd, const float* g, float c, int64_t n, hipStream_t s)
{
    PETL_EW_LAUNCH(resid_kernel, n, r, Pd, g, c);
}
void launch_reciprocal(float* x, float dz, int64_t n, hipStream_t s) { PETL_EW_LAUNCH(recip_kernel, n, x, dz); }
void launch_clip_neg(float* x, int64_t n, hipStream_t s) { PETL_EW_LAUNCH(clip_neg_kernel, n, x); }

}  // namespace petl
