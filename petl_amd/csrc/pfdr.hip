// PFDR rebinning, ramp filter and FBP weights on the device (hipFFT / rocFFT for the transforms).
//
// Reference: src/pfdr.cu:16-315 (PFDR), src/ramp_filter.cu:20-349 (ramp filter), src/vector_ops.cpp:481-505
// (FFT sizes), src/planogram.cpp:252-330 (planogram weight).  Each kernel is a thread per output element
// with the reference's float / double arithmetic and contraction off; the transforms differ from cuFFT's
// only in float rounding.  Per v1 row PFDR transforms all N_u1 views of that row in one batched 2D
// real-to-complex transform and adds the row into the rebinned spectrum in place (each thread owns one
// rebinned coefficient, so no atomics); one batched inverse transform ends it.
#include <hipfft/hipfft.h>
#include <cmath>
#include <vector>
#include "pfdr.h"
#include "petl_model.h"

namespace petl {

namespace {

constexpr int PT = 256;

#define PETL_FFT_CHECK(expr)                                                                         \
    do {                                                                                             \
        hipfftResult _r = (expr);                                                                    \
        if (_r != HIPFFT_SUCCESS)                                                                    \
            throw hip_error(std::string(#expr) + " failed with hipFFT status " + std::to_string(int(_r))); \
    } while (0)

struct FftPlan {
    hipfftHandle h = 0;
    bool made = false;
    ~FftPlan()
    {
        if (made)
            (void)hipfftDestroy(h);
    }
};

inline unsigned blocks_for(int64_t n) { return unsigned((n + PT - 1) / PT); }

// copy_data_to_2d_fft_buffer: out[iu1][iv0][iu0] (zero-padded to NY x NX) = g[iv0][iu1][iu0] / sqrt(1 + v1^2 + v0^2)
__global__ void __launch_bounds__(PT) pfdr_load_kernel(float* __restrict__ out, const float* __restrict__ g, int B,
                                                       int NY, int NX, int Nv0, int Nu1, int Nu0, float Tv0,
                                                       float v0_0, float v1)
{
#pragma clang fp contract(off)
    const int64_t e = int64_t(blockIdx.x) * PT + threadIdx.x;
    if (e >= int64_t(B) * NY * NX)
        return;
    const int iu0 = int(e % NX);
    const int iv0 = int((e / NX) % NY);
    const int iu1 = int(e / (int64_t(NX) * NY));
    if (iv0 >= Nv0 || iu0 >= Nu0) {
        out[e] = 0.0f;
        return;
    }
    const float v0 = Tv0 * float(iv0) + v0_0;
    out[e] = g[(int64_t(iv0) * Nu1 + iu1) * Nu0 + iu0] * (1.0f / sqrtf(1.0f + v1 * v1 + v0 * v0));
}

struct RebinArgs {
    int Nu1, NY, NX, NXH;
    float T_U, ac, Tu1, u1_0, Tv1, v1;
};

__device__ __forceinline__ float freq(int i, int N, float T) { return i < N / 2 ? float(i) * T : float(i - N) * T; }

// rebin (src/pfdr.cu:74-117): one rebinned coefficient per thread, [iu1][iV0][iU0]
__global__ void __launch_bounds__(PT) pfdr_rebin_kernel(float2* __restrict__ Greb, const float2* __restrict__ G,
                                                        float* __restrict__ wts, RebinArgs A)
{
#pragma clang fp contract(off)
    const int64_t e = int64_t(blockIdx.x) * PT + threadIdx.x;
    if (e >= int64_t(A.Nu1) * A.NY * A.NXH)
        return;
    const int iU0 = int(e % A.NXH);
    const int iV0 = int((e / A.NXH) % A.NY);
    const int iu1 = int(e / (int64_t(A.NXH) * A.NY));
    const float U_0 = freq(iU0, A.NX, A.T_U);
    const float abs_v1 = fabsf(A.v1);
    const bool direct = abs_v1 <= 0.5f * A.Tv1;
    if (!direct && !(fabsf(U_0) > abs_v1 * A.ac))
        return;
    if (direct) {
        const float2 g = G[e];
        Greb[e].x = Greb[e].x + g.x;
        Greb[e].y = Greb[e].y + g.y;
        wts[e] = wts[e] + 1.0f;
        return;
    }
    // the oblique row is read at the unshifted u1: (u1 - u1_0) / T_u1 in float (src/pfdr.cu:97-99)
    const float u1 = float(iu1) * A.Tu1 + A.u1_0;
    const float iu1o = (u1 - A.u1_0) / A.Tu1;
    if (!(iu1o >= 0.0f && iu1o <= float(A.Nu1 - 1)))
        return;
    const int lo = int(iu1o);
    const int hi = min(A.Nu1 - 1, lo + 1);
    const float d = iu1o - float(lo);
    const int64_t rest = e - int64_t(iu1) * A.NY * A.NXH;
    const float2 glo = G[int64_t(lo) * A.NY * A.NXH + rest];
    const float2 ghi = G[int64_t(hi) * A.NY * A.NXH + rest];
    const double wlo = 1.0 - double(d);
    Greb[e].x = float(double(Greb[e].x) + (wlo * double(glo.x) + double(d * ghi.x)));
    Greb[e].y = float(double(Greb[e].y) + (wlo * double(glo.y) + double(d * ghi.y)));
    wts[e] = wts[e] + 1.0f;
}

// normalize (src/pfdr.cu:119-134)
__global__ void __launch_bounds__(PT) pfdr_normalize_kernel(float2* __restrict__ Greb, const float* __restrict__ wts,
                                                            int64_t n, int NXH, int NY, int NX)
{
#pragma clang fp contract(off)
    const int64_t e = int64_t(blockIdx.x) * PT + threadIdx.x;
    if (e >= n)
        return;
    double w = double(1.0f / float(NY * NX));
    const float c = wts[e];
    if (c > 0.0f && (e % NXH) != 0)
        w = w * double(1.0f / c);
    Greb[e].x = float(double(Greb[e].x) * w);
    Greb[e].y = float(double(Greb[e].y) * w);
}

// copy_data_from_2d_fft_buffer (src/pfdr.cu:136-150): reb[iv0][iu1][iu0] = real[iu1][iv0][iu0] sqrt(1 + v0^2)
__global__ void __launch_bounds__(PT) pfdr_store_kernel(float* __restrict__ reb, const float* __restrict__ real,
                                                        int Nv0, int Nu1, int Nu0, int NY, int NX, float Tv0,
                                                        float v0_0)
{
#pragma clang fp contract(off)
    const int64_t e = int64_t(blockIdx.x) * PT + threadIdx.x;
    if (e >= int64_t(Nv0) * Nu1 * Nu0)
        return;
    const int iu0 = int(e % Nu0);
    const int iu1 = int((e / Nu0) % Nu1);
    const int iv0 = int(e / (int64_t(Nu0) * Nu1));
    const float v0 = Tv0 * float(iv0) + v0_0;
    reb[e] = real[(int64_t(iu1) * NY + iv0) * NX + iu0] * sqrtf(1.0f + v0 * v0);
}

__global__ void __launch_bounds__(PT) zero_kernel(float* __restrict__ x, int64_t n)
{
    const int64_t e = int64_t(blockIdx.x) * PT + threadIdx.x;
    if (e < n)
        x[e] = 0.0f;
}

// setPaddedDataKernel (src/ramp_filter.cu:36-58): rows of Nu0 zero-padded to NH
__global__ void __launch_bounds__(PT) ramp_pad_kernel(float* __restrict__ pad, const float* __restrict__ x, int64_t rows,
                                                      int Nu0, int NH)
{
    const int64_t e = int64_t(blockIdx.x) * PT + threadIdx.x;
    if (e >= rows * NH)
        return;
    const int k = int(e % NH);
    const int64_t r = e / NH;
    pad[e] = k < Nu0 ? x[r * Nu0 + k] : 0.0f;
}

// multiplyRampFilterKernel (src/ramp_filter.cu:24-34)
__global__ void __launch_bounds__(PT) ramp_mul_kernel(float2* __restrict__ G, const float* __restrict__ H, int64_t n,
                                                      int NHH)
{
#pragma clang fp contract(off)
    const int64_t e = int64_t(blockIdx.x) * PT + threadIdx.x;
    if (e >= n)
        return;
    const float h = H[e % NHH];
    G[e].x = G[e].x * h;
    G[e].y = G[e].y * h;
}

// setFilteredDataKernel (src/ramp_filter.cu:60-72) followed by apply_planogram_weight(true) of the v1 = 0
// views (src/planogram.cpp:252-256, 304-330): x[iv0][iu1][iu0] = pad[row][iu0] / (1 / sqrt(1 + v0^2))
__global__ void __launch_bounds__(PT) ramp_store_kernel(float* __restrict__ x, const float* __restrict__ pad,
                                                        int64_t rows, int64_t row0, int Nu1, int Nu0, int NH,
                                                        float Tv0, float v0_0)
{
#pragma clang fp contract(off)
    const int64_t e = int64_t(blockIdx.x) * PT + threadIdx.x;
    if (e >= rows * Nu0)
        return;
    const int k = int(e % Nu0);
    const int64_t r = e / Nu0;
    const int iv0 = int((row0 + r) / Nu1);
    const float v0 = float(iv0) * Tv0 + v0_0;
    const float v1 = 0.0f;
    const float w = float(1.0 / sqrt(1.0 + double(v1 * v1) + double(v0 * v0)));
    const float c = float(1.0 / double(w));
    x[e] = pad[r * NH + k] * c;
}

}  // namespace

int optimal_fft_size(int N)
{
    // the smallest 2^(x+1) 3^y >= N found by the reference's own search
    if (N <= 2)
        return 2;
    const double c1 = std::log2(double(N) / 2.0) / std::log2(3);
    const double c2 = 1.0 / std::log2(3);
    const double xbar = std::log2(double(N) / 2.0);
    int best = int(std::pow(2, int(std::ceil(xbar)) + 1));
    for (int x = 0; x < int(std::ceil(xbar)); x++) {
        const int y = int(std::ceil(c1 - c2 * double(x)));
        const int v = int(std::pow(2, x + 1) * std::pow(3, y));
        if (v < best && y >= 0)
            best = v;
    }
    return best;
}

void pfdr_device(const PlanGeom& P, float v_m0, const float* d_g, float* d_reb, hipStream_t s)
{
    const int NX = optimal_fft_size(P.Nu0), NY = optimal_fft_size(P.Nv0), NXH = NX / 2 + 1, B = P.Nu1;
    const int64_t nreal = int64_t(B) * NY * NX, ncplx = int64_t(B) * NY * NXH;
    DevBuf real, freq, greb, wts;
    real.ensure(size_t(nreal) * sizeof(float));
    freq.ensure(size_t(ncplx) * sizeof(float2));
    greb.ensure(size_t(ncplx) * sizeof(float2));
    wts.ensure(size_t(ncplx) * sizeof(float));
    hipLaunchKernelGGL(zero_kernel, dim3(blocks_for(2 * ncplx)), dim3(PT), 0, s, greb.f(), 2 * ncplx);
    hipLaunchKernelGGL(zero_kernel, dim3(blocks_for(ncplx)), dim3(PT), 0, s, wts.f(), ncplx);
    PETL_LAUNCH_CHECK();

    FftPlan fwd, inv;
    int n[2] = {NY, NX}, inembed[2] = {NY, NX}, onembed[2] = {NY, NXH};
    PETL_FFT_CHECK(hipfftPlanMany(&fwd.h, 2, n, inembed, 1, NY * NX, onembed, 1, NY * NXH, HIPFFT_R2C, B));
    fwd.made = true;
    PETL_FFT_CHECK(hipfftPlanMany(&inv.h, 2, n, onembed, 1, NY * NXH, inembed, 1, NY * NX, HIPFFT_C2R, B));
    inv.made = true;
    PETL_FFT_CHECK(hipfftSetStream(fwd.h, s));
    PETL_FFT_CHECK(hipfftSetStream(inv.h, s));

    RebinArgs A;
    A.Nu1 = P.Nu1;
    A.NY = NY;
    A.NX = NX;
    A.NXH = NXH;
    A.T_U = float(1.0 / double(float(NX) * P.Tu0));
    A.ac = float(1.0 / (2.0 * double(P.Tu1) * 2.0 * double(v_m0)));
    A.Tu1 = P.Tu1;
    A.u1_0 = P.u1_0;
    A.Tv1 = P.Tv1;
    const int64_t view = int64_t(P.Nv0) * P.Nu1 * P.Nu0;
    for (int iv1 = 0; iv1 < P.Nv1; iv1++) {
        const float v1 = float(iv1) * P.Tv1 + P.v1_0;
        hipLaunchKernelGGL(pfdr_load_kernel, dim3(blocks_for(nreal)), dim3(PT), 0, s, real.f(), d_g + iv1 * view, B, NY,
                           NX, P.Nv0, P.Nu1, P.Nu0, P.Tv0, P.v0_0, v1);
        PETL_LAUNCH_CHECK();
        PETL_FFT_CHECK(hipfftExecR2C(fwd.h, real.f(), static_cast<hipfftComplex*>(freq.p)));
        A.v1 = v1;
        hipLaunchKernelGGL(pfdr_rebin_kernel, dim3(blocks_for(ncplx)), dim3(PT), 0, s, static_cast<float2*>(greb.p),
                           static_cast<const float2*>(freq.p), wts.f(), A);
        PETL_LAUNCH_CHECK();
    }
    hipLaunchKernelGGL(pfdr_normalize_kernel, dim3(blocks_for(ncplx)), dim3(PT), 0, s, static_cast<float2*>(greb.p),
                       wts.f(), ncplx, NXH, NY, NX);
    PETL_LAUNCH_CHECK();
    PETL_FFT_CHECK(hipfftExecC2R(inv.h, static_cast<hipfftComplex*>(greb.p), real.f()));
    hipLaunchKernelGGL(pfdr_store_kernel, dim3(blocks_for(view)), dim3(PT), 0, s, d_reb, real.f(), P.Nv0, P.Nu1,
                       P.Nu0, NY, NX, P.Tv0, P.v0_0);
    PETL_LAUNCH_CHECK();
    // the buffers are freed when this returns: finish the work that reads them first
    PETL_HIP_CHECK(hipStreamSynchronize(s));
}

void fbp_filter_device(float* d_reb, int N_v0, int N_u1, int N_u0, float T_u0, float T_v0, float v0_0, hipStream_t s)
{
    const int NH = optimal_fft_size(2 * N_u0), NHH = NH / 2 + 1;
    // H = Re(FFT(h)) / N_H with h[i] = 1 / (T pi (1/4 - s^2)) (src/ramp_filter.cu:74-155, 334-349); h is even,
    // so its transform is the cosine sum, formed here in double
    std::vector<float> h(NH), H(NHH);
    for (int i = 0; i < NH; i++) {
        const float sv = float(i < NH / 2 ? i : i - NH);
        h[i] = float(1.0 / (double(T_u0) * PETL_PI * (0.25 - double(sv * sv))));
    }
    for (int k = 0; k < NHH; k++) {
        double acc = 0.0;
        for (int i = 0; i < NH; i++)
            acc += double(h[i]) * std::cos(2.0 * PETL_PI * double((int64_t(k) * i) % NH) / double(NH));
        H[k] = float(acc) / float(NH);
    }
    const int64_t rows_total = int64_t(N_v0) * N_u1;
    const int64_t chunk = std::min<int64_t>(rows_total, int64_t(1) << 18);
    DevBuf pad, G, dH;
    pad.ensure(size_t(chunk) * NH * sizeof(float));
    G.ensure(size_t(chunk) * NHH * sizeof(float2));
    dH.ensure(size_t(NHH) * sizeof(float));
    PETL_HIP_CHECK(hipMemcpyAsync(dH.p, H.data(), size_t(NHH) * sizeof(float), hipMemcpyHostToDevice, s));
    for (int64_t r0 = 0; r0 < rows_total; r0 += chunk) {
        const int64_t rows = std::min(chunk, rows_total - r0);
        FftPlan fwd, inv;
        int n[1] = {NH}, emb_r[1] = {NH}, emb_c[1] = {NHH};
        PETL_FFT_CHECK(hipfftPlanMany(&fwd.h, 1, n, emb_r, 1, NH, emb_c, 1, NHH, HIPFFT_R2C, int(rows)));
        fwd.made = true;
        PETL_FFT_CHECK(hipfftPlanMany(&inv.h, 1, n, emb_c, 1, NHH, emb_r, 1, NH, HIPFFT_C2R, int(rows)));
        inv.made = true;
        PETL_FFT_CHECK(hipfftSetStream(fwd.h, s));
        PETL_FFT_CHECK(hipfftSetStream(inv.h, s));
        float* x = d_reb + r0 * N_u0;
        hipLaunchKernelGGL(ramp_pad_kernel, dim3(blocks_for(rows * NH)), dim3(PT), 0, s, pad.f(), x, rows, N_u0, NH);
        PETL_LAUNCH_CHECK();
        PETL_FFT_CHECK(hipfftExecR2C(fwd.h, pad.f(), static_cast<hipfftComplex*>(G.p)));
        hipLaunchKernelGGL(ramp_mul_kernel, dim3(blocks_for(rows * NHH)), dim3(PT), 0, s, static_cast<float2*>(G.p),
                           dH.f(), rows * NHH, NHH);
        PETL_LAUNCH_CHECK();
        PETL_FFT_CHECK(hipfftExecC2R(inv.h, static_cast<hipfftComplex*>(G.p), pad.f()));
        hipLaunchKernelGGL(ramp_store_kernel, dim3(blocks_for(rows * N_u0)), dim3(PT), 0, s, x, pad.f(), rows, r0,
                           N_u1, N_u0, NH, T_v0, v0_0);
        PETL_LAUNCH_CHECK();
        PETL_HIP_CHECK(hipStreamSynchronize(s));  // the plans and buffers go out of scope
    }
}

}  // namespace petl
