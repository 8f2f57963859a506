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

inline bool aligned16(const void* p) { return (reinterpret_cast<uintptr_t>(p) & 15u) == 0; }

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

// per-block partial of sum double(x_i * y_i [* w_i]): float4 loads, DOT_U of them per array in flight per
// thread (n4 = 0 when a pointer is not 16-byte aligned), then the scalar tail from 4 * n4
constexpr int DOT_U = 4;

template <bool HAS_W>
__global__ void __launch_bounds__(VT) dot_partial_kernel(const float* __restrict__ x, const float* __restrict__ y,
                                                         const float* __restrict__ w, int64_t n, int64_t n4,
                                                         double* __restrict__ part)
{
#pragma clang fp contract(off)
    const float4* x4 = reinterpret_cast<const float4*>(x);
    const float4* y4 = reinterpret_cast<const float4*>(y);
    const float4* w4 = reinterpret_cast<const float4*>(w);
    const int64_t stride = int64_t(gridDim.x) * VT;
    double acc = 0.0;
    for (int64_t k = int64_t(blockIdx.x) * VT + threadIdx.x; k < n4; k += DOT_U * stride) {
        float4 vx[DOT_U], vy[DOT_U], vw[DOT_U];
#pragma unroll
        for (int u = 0; u < DOT_U; u++) {
            const int64_t kk = k + u * stride;
            if (kk < n4) {
                vx[u] = x4[kk];
                vy[u] = y4[kk];
                if (HAS_W)
                    vw[u] = w4[kk];
            }
        }
#pragma unroll
        for (int u = 0; u < DOT_U; u++) {
            const int64_t kk = k + u * stride;
            if (kk < n4) {
                float4 p;
                p.x = vx[u].x * vy[u].x;
                p.y = vx[u].y * vy[u].y;
                p.z = vx[u].z * vy[u].z;
                p.w = vx[u].w * vy[u].w;
                if (HAS_W) {
                    p.x = p.x * vw[u].x;
                    p.y = p.y * vw[u].y;
                    p.z = p.z * vw[u].z;
                    p.w = p.w * vw[u].w;
                }
                acc += double(p.x);
                acc += double(p.y);
                acc += double(p.z);
                acc += double(p.w);
            }
        }
    }
    for (int64_t i = 4 * n4 + int64_t(blockIdx.x) * VT + threadIdx.x; i < n; i += stride) {
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
//
// o[i] = op(o[i], a[i], b[i]) over n elements.  HBM needs many bytes in flight per wave to reach its
// bandwidth (a scalar-load grid-stride loop peaks near 1.1 TB/s here), so each thread issues EW_U float4
// loads per array before it computes and stores.  The scalar kernel covers unaligned pointers and the
// tail of fewer than four elements.
// ---------------------------------------------------------------------------------------------
constexpr int EW_U = 4;

template <bool READ_O, int NA, typename Op>
__global__ void __launch_bounds__(VT) ew_vec_kernel(float* __restrict__ o, const float* __restrict__ a,
                                                    const float* __restrict__ b, int64_t n4, Op op)
{
    float4* o4 = reinterpret_cast<float4*>(o);
    const float4* a4 = reinterpret_cast<const float4*>(a);
    const float4* b4 = reinterpret_cast<const float4*>(b);
    const int64_t stride = int64_t(gridDim.x) * VT;
    for (int64_t k = int64_t(blockIdx.x) * VT + threadIdx.x; k < n4; k += EW_U * stride) {
        float4 vo[EW_U], va[EW_U], vb[EW_U];
#pragma unroll
        for (int u = 0; u < EW_U; u++) {
            const int64_t kk = k + u * stride;
            if (kk < n4) {
                if (READ_O)
                    vo[u] = o4[kk];
                if (NA >= 1)
                    va[u] = a4[kk];
                if (NA >= 2)
                    vb[u] = b4[kk];
            }
        }
#pragma unroll
        for (int u = 0; u < EW_U; u++) {
            const int64_t kk = k + u * stride;
            if (kk < n4) {
                float4 r;
                r.x = op(READ_O ? vo[u].x : 0.0f, NA >= 1 ? va[u].x : 0.0f, NA >= 2 ? vb[u].x : 0.0f);
                r.y = op(READ_O ? vo[u].y : 0.0f, NA >= 1 ? va[u].y : 0.0f, NA >= 2 ? vb[u].y : 0.0f);
                r.z = op(READ_O ? vo[u].z : 0.0f, NA >= 1 ? va[u].z : 0.0f, NA >= 2 ? vb[u].z : 0.0f);
                r.w = op(READ_O ? vo[u].w : 0.0f, NA >= 1 ? va[u].w : 0.0f, NA >= 2 ? vb[u].w : 0.0f);
                o4[kk] = r;
            }
        }
    }
}

template <bool READ_O, int NA, typename Op>
__global__ void __launch_bounds__(VT) ew_scalar_kernel(float* __restrict__ o, const float* __restrict__ a,
                                                       const float* __restrict__ b, int64_t first, int64_t n,
                                                       Op op)
{
    for (int64_t i = first + int64_t(blockIdx.x) * VT + threadIdx.x; i < n; i += int64_t(gridDim.x) * VT)
        o[i] = op(READ_O ? o[i] : 0.0f, NA >= 1 ? a[i] : 0.0f, NA >= 2 ? b[i] : 0.0f);
}

template <bool READ_O, int NA, typename Op>
void ew_launch(float* o, const float* a, const float* b, int64_t n, Op op, hipStream_t s)
{
    if (n <= 0)
        return;
    const bool vec = aligned16(o) && (NA < 1 || aligned16(a)) && (NA < 2 || aligned16(b));
    const int64_t n4 = vec ? n / 4 : 0;
    if (n4 > 0) {
        const int blocks = int(std::min<int64_t>((n4 + int64_t(VT) * EW_U - 1) / (int64_t(VT) * EW_U), 65536));
        hipLaunchKernelGGL((ew_vec_kernel<READ_O, NA, Op>), dim3(blocks), dim3(VT), 0, s, o, a, b, n4, op);
        PETL_LAUNCH_CHECK();
    }
    if (4 * n4 < n) {
        hipLaunchKernelGGL((ew_scalar_kernel<READ_O, NA, Op>), dim3(grid_for(n - 4 * n4)), dim3(VT), 0, s, o, a, b,
                           4 * n4, n, op);
        PETL_LAUNCH_CHECK();
    }
}

// the operations (contraction off: every product rounds as numpy's float32 arithmetic does)
struct OpSub {
    __device__ float operator()(float o, float a, float) const { return o - a; }
};
struct OpAdd {
    __device__ float operator()(float o, float a, float) const { return o + a; }
};
struct OpMulTo {
    __device__ float operator()(float, float a, float b) const
    {
#pragma clang fp contract(off)
        return a * b;
    }
};
struct OpXpay {  // d = u + gamma d   (o = d, a = u)
    float gamma;
    __device__ float operator()(float o, float a, float) const
    {
#pragma clang fp contract(off)
        const float t = gamma * o;
        return a + t;
    }
};
struct OpStep {  // f = f - step d, clipped at 0   (o = f, a = d)
    float step;
    int clip;
    __device__ float operator()(float o, float a, float) const
    {
#pragma clang fp contract(off)
        const float t = step * a;
        const float v = o - t;
        return (clip && v < 0.0f) ? 0.0f : v;
    }
};
struct OpResid {  // r = (r + c Pd) - g   (o = r, a = Pd, b = g)
    float c;
    __device__ float operator()(float o, float a, float b) const
    {
#pragma clang fp contract(off)
        const float t = c * a;
        const float v = o + t;
        return v - b;
    }
};
struct OpRecip {
    float dz;
    __device__ float operator()(float o, float, float) const { return (o == 0.0f) ? dz : float(1.0 / double(o)); }
};
struct OpClipNeg {
    __device__ float operator()(float o, float, float) const { return o < 0.0f ? 0.0f : o; }
};
struct OpCopy {
    __device__ float operator()(float, float a, float) const { return a; }
};
struct OpFill {
    float v;
    __device__ float operator()(float, float, float) const { return v; }
};
struct OpMul {
    __device__ float operator()(float o, float a, float) const
    {
#pragma clang fp contract(off)
        return o * a;
    }
};

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
        launch_copy_vec(f, src, n, s);
    }
}

int64_t red_scratch_doubles(int64_t n) { return std::max<int64_t>(1, std::min<int64_t>((n + VT - 1) / VT, 4096)); }

void launch_dot(const float* x, const float* y, const float* w, int64_t n, double* scratch, double* out,
                hipStream_t s)
{
    const int nparts = int(red_scratch_doubles(n));
    const bool vec = aligned16(x) && aligned16(y) && (!w || aligned16(w));
    const int64_t n4 = vec ? n / 4 : 0;
    if (w)
        hipLaunchKernelGGL(dot_partial_kernel<true>, dim3(nparts), dim3(VT), 0, s, x, y, w, n, n4, scratch);
    else
        hipLaunchKernelGGL(dot_partial_kernel<false>, dim3(nparts), dim3(VT), 0, s, x, y, w, n, n4, scratch);
    PETL_LAUNCH_CHECK();
    hipLaunchKernelGGL(sum_final_kernel, dim3(1), dim3(VT), 0, s, scratch, nparts, out);
    PETL_LAUNCH_CHECK();
}

void launch_sum_parts(const double* part, int64_t n, double* out, hipStream_t s)
{
    if (n > INT32_MAX)
        throw hip_error("launch_sum_parts: too many partials");
    hipLaunchKernelGGL(sum_final_kernel, dim3(1), dim3(VT), 0, s, part, int(n), out);
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

void launch_sub(float* x, const float* y, int64_t n, hipStream_t s) { ew_launch<true, 1>(x, y, nullptr, n, OpSub{}, s); }
void launch_add(float* x, const float* y, int64_t n, hipStream_t s) { ew_launch<true, 1>(x, y, nullptr, n, OpAdd{}, s); }
void launch_mul_to(float* dst, const float* a, const float* b, int64_t n, hipStream_t s)
{
    ew_launch<false, 2>(dst, a, b, n, OpMulTo{}, s);
}
void launch_xpay(float* d, const float* u, float gamma, int64_t n, hipStream_t s)
{
    ew_launch<true, 1>(d, u, nullptr, n, OpXpay{gamma}, s);
}
void launch_step(float* f, const float* d, float step, bool clip, int64_t n, hipStream_t s)
{
    ew_launch<true, 1>(f, d, nullptr, n, OpStep{step, clip ? 1 : 0}, s);
}
void launch_resid_update(float* r, const float* Pd, const float* g, float c, int64_t n, hipStream_t s)
{
    ew_launch<true, 2>(r, Pd, g, n, OpResid{c}, s);
}
void launch_reciprocal(float* x, float dz, int64_t n, hipStream_t s)
{
    ew_launch<true, 0>(x, nullptr, nullptr, n, OpRecip{dz}, s);
}
void launch_clip_neg(float* x, int64_t n, hipStream_t s) { ew_launch<true, 0>(x, nullptr, nullptr, n, OpClipNeg{}, s); }
void launch_mul_inplace(float* x, const float* y, int64_t n, hipStream_t s)
{
    ew_launch<true, 1>(x, y, nullptr, n, OpMul{}, s);
}
void launch_copy_vec(float* dst, const float* src, int64_t n, hipStream_t s)
{
    ew_launch<false, 1>(dst, src, nullptr, n, OpCopy{}, s);
}
void launch_fill_vec(float* x, float v, int64_t n, hipStream_t s)
{
    ew_launch<false, 0>(x, nullptr, nullptr, n, OpFill{v}, s);
}

}  // namespace petl
