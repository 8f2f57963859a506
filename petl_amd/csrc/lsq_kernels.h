// Kernels of the least-squares reconstructions (RWLS / WLS / RLS / LS): Gaussian smoothing of the default
// weights, the TV cost and quadratic form of the step size, deterministic FP64 inner products and the
// conjugate-gradient vector updates.  All enqueue on the given stream; none allocates.
#pragma once
#include "petl_common.h"

namespace petl {

// ---- Gaussian smoothing (src/relative_differences.cu:314-420, 753-800) ----
struct GaussParams {
    int radius;  // ceil(sqrt(2 ln 10) sigma)
    float den;   // exponent factor 1 / (2 sigma): tap weight exp(-den * s^2)
};
GaussParams gauss_params(float FWHM);
// Smooths the [N1][N2][N3] array f in place along numDims axes (3: all, 2: the last two, 1: the first) as
// one pass per axis through tmp (same size).  weights_epilogue: the last pass writes 1 / max(x, 7) (the
// default RWLS weights, src/petl_server.py:1722-1726) instead of x.
void gaussian_filter_device(float* f, float* tmp, int N1, int N2, int N3, float FWHM, int numDims,
                            bool weights_epilogue, hipStream_t s);

// ---- deterministic reductions in FP64 ----
// scratch doubles needed by a reduction over n elements
int64_t red_scratch_doubles(int64_t n);
// *out = sum_i double(x_i * y_i [* w_i]) (w may be null); fixed reduction order
void launch_dot(const float* x, const float* y, const float* w, int64_t n, double* scratch, double* out,
                hipStream_t s);
// *out = sum of the n FP64 partials in index order (one workgroup)
void launch_sum_parts(const double* part, int64_t n, double* out, hipStream_t s);
// *out = p-Huber TV cost of f (src/total_variation.cu:80-216) or, with d != null, the quadratic form
// <d, R''(f) d> (src/total_variation.cu:218-417); beta/delta rules of TV_gradient apply.
void launch_tv_reduce(const float* f, const float* d, int N1, int N2, int N3, float delta, float beta, float p,
                      double* scratch, double* out, hipStream_t s);

// ---- conjugate-gradient vector updates (fp32, contraction off: every product rounds as numpy's; float4
//      loads, several per thread in flight, where the pointers are 16-byte aligned) ----
void launch_sub(float* x, const float* y, int64_t n, hipStream_t s);                  // x -= y
void launch_add(float* x, const float* y, int64_t n, hipStream_t s);                  // x += y
void launch_mul_to(float* dst, const float* a, const float* b, int64_t n, hipStream_t s);  // dst = a * b
void launch_xpay(float* d, const float* u, float gamma, int64_t n, hipStream_t s);    // d = u + gamma d
// f = f - step d; then f < 0 -> 0 when clip
void launch_step(float* f, const float* d, float step, bool clip, int64_t n, hipStream_t s);
// r = (r + c Pd) - g
void launch_resid_update(float* r, const float* Pd, const float* g, float c, int64_t n, hipStream_t s);
// x = (x == 0 ? dz : 1 / x), the quotient formed in double as dataCube's RECIPROCAL does
void launch_reciprocal(float* x, float dz, int64_t n, hipStream_t s);
void launch_clip_neg(float* x, int64_t n, hipStream_t s);  // x < 0 -> 0
void launch_mul_inplace(float* x, const float* y, int64_t n, hipStream_t s);  // x *= y
void launch_copy_vec(float* dst, const float* src, int64_t n, hipStream_t s);  // dst = src
void launch_fill_vec(float* x, float v, int64_t n, hipStream_t s);             // x = v

}  // namespace petl
