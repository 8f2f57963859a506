// Elementwise and reduction kernels used by the device-resident reconstruction.
#pragma once
#include "petl_common.h"

namespace petl {

void launch_fill(float* x, float v, int64_t n, hipStream_t s);
// dst = src (a kernel rather than a copy engine call, so it can sit in a captured graph)
void launch_copy(float* dst, const float* src, int64_t n, hipStream_t s);
// x *= y
void launch_mul(float* x, const float* y, int64_t n, hipStream_t s);
// S == 0 -> 1 (src/petl_server.py:1584)
void launch_sens_fix(float* S, int64_t n, hipStream_t s);
// f = f * (d / S), d/0 -> 0 (src/petl_server.py:1596 with src/data_cube.cpp:144-149)
void launch_mlem_update(float* f, const float* d, const float* S, int64_t n, hipStream_t s);
// OSL prior update (src/petl_server.py:1598-1602): G = grad R(f) + S; G = max(0.1 S, G); f_out = f * (d / G).
// kind 0: relative differences (delta, beta), kind 1: p-Huber TV (delta, beta, p).  f_out != f.
void launch_mlem_update_prior(float* f_out, const float* f, const float* d, const float* S, int N1, int N2, int N3,
                              int kind, float delta, float beta, float p, hipStream_t s);
// MLEM ratio from a stored projection: r = (q == 0 ? 0 : y / q), then r *= resp when resp != null
// (rdivide of src/data_cube.cpp:150-155 followed by the response weighting of the back projection).
// Bins are taken in rows of row_elems, rows first, first + stride, ... (< nrows_total): a v1 subset.
void launch_ratio(float* r, const float* y, const float* q, const float* resp, int64_t row_elems, int first,
                  int stride, int nrows_total, hipStream_t s);
// Poisson log-likelihood sum(y log(q) - q) over bins with q > 0, in FP64; q = projection (times response).
// Accumulates into *out (device double) via a two-pass deterministic reduction using `scratch`
// (at least ll_scratch_doubles(n) doubles).
// Storage-format forms (STORE_*; x = stored * scale for STORE_F16): fill, copy, and the image updates with
// the image f in storage format (d and S stay fp32).
void launch_fill_st(int st, void* x, float v, float scale, int64_t n, hipStream_t s);
void launch_copy_st(int st, void* dst, const void* src, int64_t n, hipStream_t s);
// dst (storage format st) = src (fp32) / scale, rounded to nearest even
void launch_convert_st(int st, void* dst, const float* src, int64_t n, float inv_scale, hipStream_t s);
void launch_mlem_update_st(int st, void* f, const float* d, const float* S, int64_t n, float f_scale, hipStream_t s);
void launch_mlem_update_prior_st(int st, void* f_out, const void* f, const float* d, const float* S, int N1, int N2,
                                 int N3, int kind, float delta, float beta, float p, float f_scale, hipStream_t s);
int64_t ll_scratch_doubles(int64_t n);
void launch_loglik(const float* y, const float* q, int64_t row_elems, int first, int stride, int nrows_total,
                   double* scratch, double* out, bool accumulate, hipStream_t s);

}  // namespace petl
