// Elementwise and reduction kernels of the device-resident reconstruction loop.
#include <cmath>
#include <cstring>
#include "petl_common.h"
#include "vec_kernels.h"
#include "priors_dev.h"
#include "mlem_update_dev.h"
#include "storage.h"

namespace petl {

namespace {

constexpr int VT = 256;

inline int grid_for(int64_t n) { return int(std::min<int64_t>((n + VT - 1) / VT, 65536)); }

__global__ void fill_kernel(float* __restrict__ x, float v, int64_t n)
{
    for (int64_t i = int64_t(blockIdx.x) * VT + threadIdx.x; i < n; i += int64_t(gridDim.x) * VT)
        x[i] = v;
}

__global__ void copy_kernel(float* __restrict__ x, const float* __restrict__ y, int64_t n)
{
    for (int64_t i = int64_t(blockIdx.x) * VT + threadIdx.x; i < n; i += int64_t(gridDim.x) * VT)
        x[i] = y[i];
}

__global__ void mul_kernel(float* __restrict__ x, const float* __restrict__ y, int64_t n)
{
    for (int64_t i = int64_t(blockIdx.x) * VT + threadIdx.x; i < n; i += int64_t(gridDim.x) * VT)
        x[i] = x[i] * y[i];
}

__global__ void sens_fix_kernel(float* __restrict__ S, int64_t n)
{
    for (int64_t i = int64_t(blockIdx.x) * VT + threadIdx.x; i < n; i += int64_t(gridDim.x) * VT)
        if (S[i] == 0.0f)
            S[i] = 1.0f;
}

__global__ void mlem_update_kernel(float* __restrict__ f, const float* __restrict__ d, const float* __restrict__ S,
                                   int64_t n)
{
    for (int64_t i = int64_t(blockIdx.x) * VT + threadIdx.x; i < n; i += int64_t(gridDim.x) * VT)
        mlem_update_at(f, d, S, i);
}

// index of the i-th bin of a row subset
__device__ __forceinline__ int64_t subset_index(int64_t i, int64_t row_elems, int first, int stride)
{
    const int64_t row = i / row_elems;
    return (first + row * stride) * row_elems + (i - row * row_elems);
}

__global__ void ratio_kernel(float* __restrict__ r, const float* __restrict__ y, const float* __restrict__ q,
                             const float* __restrict__ resp, int64_t n, int64_t row_elems, int first, int stride)
{
    for (int64_t ii = int64_t(blockIdx.x) * VT + threadIdx.x; ii < n; ii += int64_t(gridDim.x) * VT) {
        const int64_t i = subset_index(ii, row_elems, first, stride);
        const float qq = q[i];
        float out = (qq == 0.0f) ? 0.0f : y[i] / qq;
        if (resp)
            out = out * resp[i];
        r[i] = out;
    }
}

template <int KIND>
__global__ void __launch_bounds__(64) mlem_update_prior_kernel(float* __restrict__ fo, const float* __restrict__ f,
                                                               const float* __restrict__ d,
                                                               const float* __restrict__ S, int N1, int N2, int N3,
                                                               float delta, float beta, HuberParams h)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < N3)
        mlem_update_prior_at<KIND>(fo, f, d, S, blockIdx.z, blockIdx.y, k, N1, N2, N3, delta, beta, h);
}

// ---- storage-format forms (16-bit image / planograms, fp32 arithmetic) ----
template <int ST>
__global__ void fill_st_kernel(typename Stor<ST>::T* __restrict__ x, float v, float inv_scale, int64_t n)
{
    const typename Stor<ST>::T h = Stor<ST>::st(v, inv_scale);
    for (int64_t i = int64_t(blockIdx.x) * VT + threadIdx.x; i < n; i += int64_t(gridDim.x) * VT)
        x[i] = h;
}

template <int ST>
__global__ void mlem_update_st_kernel(typename Stor<ST>::T* __restrict__ f, const float* __restrict__ d,
                                      const float* __restrict__ S, int64_t n, float f_scale)
{
    const float inv = 1.0f / f_scale;
    for (int64_t i = int64_t(blockIdx.x) * VT + threadIdx.x; i < n; i += int64_t(gridDim.x) * VT) {
        const float s = S[i];
        const float q = (s == 0.0f) ? 0.0f : d[i] / s;
        f[i] = Stor<ST>::st(Stor<ST>::ld(f[i], f_scale) * q, inv);
    }
}

template <int KIND, int ST>
__global__ void __launch_bounds__(64) mlem_update_prior_st_kernel(typename Stor<ST>::T* __restrict__ fo,
                                                                  const typename Stor<ST>::T* __restrict__ f,
                                                                  const float* __restrict__ d,
                                                                  const float* __restrict__ S, int N1, int N2,
                                                                  int N3, float delta, float beta, HuberParams h,
                                                                  float f_scale)
{
#pragma clang fp contract(off)
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y;
    const int i = blockIdx.z;
    if (k >= N3)
        return;
    const int64_t v = (int64_t(i) * N2 + j) * N3 + k;
    const float s = S[v];
    auto ld = [f, f_scale](int64_t x) { return Stor<ST>::ld(f[x], f_scale); };
    float G = prior_grad_at_ld<KIND>(ld, i, j, k, N1, N2, N3, delta, beta, h);
    G = G + s;
    G = fmaxf(0.1f * s, G);
    const float q = (G == 0.0f) ? 0.0f : d[v] / G;
    fo[v] = Stor<ST>::st(ld(v) * q, 1.0f / f_scale);
}

template <int ST>
__global__ void convert_st_kernel(typename Stor<ST>::T* __restrict__ x, const float* __restrict__ y, int64_t n,
                                  float inv_scale)
{
    for (int64_t i = int64_t(blockIdx.x) * VT + threadIdx.x; i < n; i += int64_t(gridDim.x) * VT)
        x[i] = Stor<ST>::st(y[i], inv_scale);
}

template <int ST>
__global__ void copy_st_kernel(typename Stor<ST>::T* __restrict__ x, const typename Stor<ST>::T* __restrict__ y,
                               int64_t n)
{
    for (int64_t i = int64_t(blockIdx.x) * VT + threadIdx.x; i < n; i += int64_t(gridDim.x) * VT)
        x[i] = y[i];
}

// pass 1: per-block partial sums in FP64
__global__ void __launch_bounds__(VT) loglik_partial_kernel(const float* __restrict__ y, const float* __restrict__ q,
                                                            int64_t n, int64_t row_elems, int first, int stride,
                                                            double* __restrict__ part)
{
    __shared__ double red[VT];
    double acc = 0.0;
    for (int64_t ii = int64_t(blockIdx.x) * VT + threadIdx.x; ii < n; ii += int64_t(gridDim.x) * VT) {
        const int64_t i = subset_index(ii, row_elems, first, stride);
        const double qq = q[i];
        if (qq > 0.0)
            acc += double(y[i]) * log(qq) - qq;
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int w = VT / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w)
            red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        part[blockIdx.x] = red[0];
}

// pass 2: one block sums the partials in a fixed order
__global__ void __launch_bounds__(VT) loglik_final_kernel(const double* __restrict__ part, int nparts,
                                                          double* __restrict__ out, int accumulate)
{
    __shared__ double red[VT];
    double acc = 0.0;
    for (int i = threadIdx.x; i < nparts; i += VT)
        acc += part[i];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int w = VT / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w)
            red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        *out = accumulate ? (*out + red[0]) : red[0];
}

}  // namespace

void launch_fill_st(int st, void* x, float v, float scale, int64_t n, hipStream_t s)
{
    switch (st) {
    case STORE_F32: launch_fill(static_cast<float*>(x), v, n, s); return;
    case STORE_BF16:
        hipLaunchKernelGGL(fill_st_kernel<STORE_BF16>, dim3(grid_for(n)), dim3(VT), 0, s,
                           static_cast<uint16_t*>(x), v, 1.0f, n);
        break;
    case STORE_F16:
        hipLaunchKernelGGL(fill_st_kernel<STORE_F16>, dim3(grid_for(n)), dim3(VT), 0, s,
                           static_cast<_Float16*>(x), v, 1.0f / scale, n);
        break;
    default: throw hip_error("fill: unknown storage format");
    }
    PETL_LAUNCH_CHECK();
}

void launch_copy_st(int st, void* dst, const void* src, int64_t n, hipStream_t s)
{
    switch (st) {
    case STORE_F32: launch_copy(static_cast<float*>(dst), static_cast<const float*>(src), n, s); return;
    case STORE_BF16:
        hipLaunchKernelGGL(copy_st_kernel<STORE_BF16>, dim3(grid_for(n)), dim3(VT), 0, s,
                           static_cast<uint16_t*>(dst), static_cast<const uint16_t*>(src), n);
        break;
    case STORE_F16:
        hipLaunchKernelGGL(copy_st_kernel<STORE_F16>, dim3(grid_for(n)), dim3(VT), 0, s,
                           static_cast<_Float16*>(dst), static_cast<const _Float16*>(src), n);
        break;
    default: throw hip_error("copy: unknown storage format");
    }
    PETL_LAUNCH_CHECK();
}

void launch_convert_st(int st, void* dst, const float* src, int64_t n, float inv_scale, hipStream_t s)
{
    switch (st) {
    case STORE_F32: launch_copy(static_cast<float*>(dst), src, n, s); return;
    case STORE_BF16:
        hipLaunchKernelGGL(convert_st_kernel<STORE_BF16>, dim3(grid_for(n)), dim3(VT), 0, s,
                           static_cast<uint16_t*>(dst), src, n, 1.0f);
        break;
    case STORE_F16:
        hipLaunchKernelGGL(convert_st_kernel<STORE_F16>, dim3(grid_for(n)), dim3(VT), 0, s,
                           static_cast<_Float16*>(dst), src, n, inv_scale);
        break;
    default: throw hip_error("convert: unknown storage format");
    }
    PETL_LAUNCH_CHECK();
}

void launch_mlem_update_st(int st, void* f, const float* d, const float* S, int64_t n, float f_scale, hipStream_t s)
{
    switch (st) {
    case STORE_F32: launch_mlem_update(static_cast<float*>(f), d, S, n, s); return;
    case STORE_BF16:
        hipLaunchKernelGGL(mlem_update_st_kernel<STORE_BF16>, dim3(grid_for(n)), dim3(VT), 0, s,
                           static_cast<uint16_t*>(f), d, S, n, 1.0f);
        break;
    case STORE_F16:
        hipLaunchKernelGGL(mlem_update_st_kernel<STORE_F16>, dim3(grid_for(n)), dim3(VT), 0, s,
                           static_cast<_Float16*>(f), d, S, n, f_scale);
        break;
    default: throw hip_error("MLEM update: unknown storage format");
    }
    PETL_LAUNCH_CHECK();
}

template <int ST>
static void update_prior_st(void* fo, const void* f, const float* d, const float* S, int N1, int N2, int N3, int kind,
                            float delta, float beta, float p, float f_scale, hipStream_t s)
{
    using TS = typename Stor<ST>::T;
    const dim3 grid(ceil_div(N3, 64), N2, N1);
    if (kind == 0) {
        HuberParams h{};
        hipLaunchKernelGGL((mlem_update_prior_st_kernel<0, ST>), grid, dim3(64), 0, s, static_cast<TS*>(fo),
                           static_cast<const TS*>(f), d, S, N1, N2, N3, delta, beta, h, f_scale);
    } else {
        const HuberParams h = huber_params(delta, beta, p);
        hipLaunchKernelGGL((mlem_update_prior_st_kernel<1, ST>), grid, dim3(64), 0, s, static_cast<TS*>(fo),
                           static_cast<const TS*>(f), d, S, N1, N2, N3, delta, beta, h, f_scale);
    }
    PETL_LAUNCH_CHECK();
}

void launch_mlem_update_prior_st(int st, void* f_out, const void* f, const float* d, const float* S, int N1, int N2,
                                 int N3, int kind, float delta, float beta, float p, float f_scale, hipStream_t s)
{
    switch (st) {
    case STORE_F32:
        launch_mlem_update_prior(static_cast<float*>(f_out), static_cast<const float*>(f), d, S, N1, N2, N3, kind,
                                 delta, beta, p, s);
        return;
    case STORE_BF16: update_prior_st<STORE_BF16>(f_out, f, d, S, N1, N2, N3, kind, delta, beta, p, 1.0f, s); return;
    case STORE_F16: update_prior_st<STORE_F16>(f_out, f, d, S, N1, N2, N3, kind, delta, beta, p, f_scale, s); return;
    default: throw hip_error("MLEM prior update: unknown storage format");
    }
}

// ---- host conversions ----
void to_storage(int st, const float* x, void* out, int64_t n, float inv_scale)
{
    switch (st) {
    case STORE_F32: memcpy(out, x, size_t(n) * 4); return;
    case STORE_BF16: {
        uint16_t* o = static_cast<uint16_t*>(out);
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < n; i++)
            o[i] = Stor<STORE_BF16>::st(x[i], 1.0f);
        return;
    }
    case STORE_F16: {
        _Float16* o = static_cast<_Float16*>(out);
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < n; i++)
            o[i] = Stor<STORE_F16>::st(x[i], inv_scale);
        return;
    }
    default: throw hip_error("unknown storage format");
    }
}

void from_storage(int st, const void* x, float* out, int64_t n, float scale)
{
    switch (st) {
    case STORE_F32: memcpy(out, x, size_t(n) * 4); return;
    case STORE_BF16: {
        const uint16_t* a = static_cast<const uint16_t*>(x);
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < n; i++)
            out[i] = Stor<STORE_BF16>::ld(a[i], 1.0f);
        return;
    }
    case STORE_F16: {
        const _Float16* a = static_cast<const _Float16*>(x);
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < n; i++)
            out[i] = Stor<STORE_F16>::ld(a[i], scale);
        return;
    }
    default: throw hip_error("unknown storage format");
    }
}

float pow2_scale_for(float max_abs, float limit)
{
    if (!(max_abs > 0.0f) || !std::isfinite(max_abs))
        return 1.0f;
    float s = 1.0f;
    while (max_abs / s > limit)
        s *= 2.0f;
    while (max_abs / s <= 0.5f * limit)
        s *= 0.5f;
    return s;
}

void launch_fill(float* x, float v, int64_t n, hipStream_t s)
{
    hipLaunchKernelGGL(fill_kernel, dim3(grid_for(n)), dim3(VT), 0, s, x, v, n);
    PETL_LAUNCH_CHECK();
}

void launch_copy(float* dst, const float* src, int64_t n, hipStream_t s)
{
    hipLaunchKernelGGL(copy_kernel, dim3(grid_for(n)), dim3(VT), 0, s, dst, src, n);
    PETL_LAUNCH_CHECK();
}

void launch_mul(float* x, const float* y, int64_t n, hipStream_t s)
{
    hipLaunchKernelGGL(mul_kernel, dim3(grid_for(n)), dim3(VT), 0, s, x, y, n);
    PETL_LAUNCH_CHECK();
}

void launch_sens_fix(float* S, int64_t n, hipStream_t s)
{
    hipLaunchKernelGGL(sens_fix_kernel, dim3(grid_for(n)), dim3(VT), 0, s, S, n);
    PETL_LAUNCH_CHECK();
}

void launch_mlem_update(float* f, const float* d, const float* S, int64_t n, hipStream_t s)
{
    hipLaunchKernelGGL(mlem_update_kernel, dim3(grid_for(n)), dim3(VT), 0, s, f, d, S, n);
    PETL_LAUNCH_CHECK();
}

void launch_mlem_update_prior(float* f_out, const float* f, const float* d, const float* S, int N1, int N2, int N3,
                              int kind, float delta, float beta, float p, hipStream_t s)
{
    const dim3 grid(ceil_div(N3, 64), N2, N1);
    if (kind == 0) {
        HuberParams h{};
        hipLaunchKernelGGL(mlem_update_prior_kernel<0>, grid, dim3(64), 0, s, f_out, f, d, S, N1, N2, N3, delta, beta,
                           h);
    } else {
        const HuberParams h = huber_params(delta, beta, p);
        hipLaunchKernelGGL(mlem_update_prior_kernel<1>, grid, dim3(64), 0, s, f_out, f, d, S, N1, N2, N3, delta, beta,
                           h);
    }
    PETL_LAUNCH_CHECK();
}

static int64_t subset_bins(int64_t row_elems, int first, int stride, int nrows_total)
{
    return int64_t((nrows_total - first + stride - 1) / stride) * row_elems;
}

void launch_ratio(float* r, const float* y, const float* q, const float* resp, int64_t row_elems, int first,
                  int stride, int nrows_total, hipStream_t s)
{
    const int64_t n = subset_bins(row_elems, first, stride, nrows_total);
    hipLaunchKernelGGL(ratio_kernel, dim3(grid_for(n)), dim3(VT), 0, s, r, y, q, resp, n, row_elems, first, stride);
    PETL_LAUNCH_CHECK();
}

int64_t ll_scratch_doubles(int64_t n) { return std::min<int64_t>((n + VT - 1) / VT, 4096); }

void launch_loglik(const float* y, const float* q, int64_t row_elems, int first, int stride, int nrows_total,
                   double* scratch, double* out, bool accumulate, hipStream_t s)
{
    const int64_t n = subset_bins(row_elems, first, stride, nrows_total);
    const int nparts = int(ll_scratch_doubles(n));
    hipLaunchKernelGGL(loglik_partial_kernel, dim3(nparts), dim3(VT), 0, s, y, q, n, row_elems, first, stride,
                       scratch);
    PETL_LAUNCH_CHECK();
    hipLaunchKernelGGL(loglik_final_kernel, dim3(1), dim3(VT), 0, s, scratch, nparts, out, accumulate ? 1 : 0);
    PETL_LAUNCH_CHECK();
}

}  // namespace petl
