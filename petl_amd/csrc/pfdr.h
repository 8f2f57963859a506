// Analytic reconstruction: planogram frequency-distance rebinning (PFDR), the ramp filter and the pieces of
// FBP, on the device with hipFFT.
//
// Reference: src/pfdr.cu:1-315, src/ramp_filter.cu:1-349, src/vector_ops.cpp:481-505, src/petl.cpp:310-385,
// src/planogram.cpp:252-330.
#pragma once
#include "petl_common.h"

namespace petl {

// optimalFFTsize (src/vector_ops.cpp:481-505)
int optimal_fft_size(int N);

// PFDR of one planogram: d_g [N_v1][N_v0][N_u1][N_u0] -> d_reb [N_v0][N_u1][N_u0] (device arrays).
// v_m0: the planogram's v_m0 (sets the aperture constant).
void pfdr_device(const PlanGeom& P, float v_m0, const float* d_g, float* d_reb, hipStream_t s);

// In place: the ramp filter of every u0 row of d_reb ([rows][N_u0]) and the inverse planogram weight
// sqrt(1 + v0^2) of the rebinned (v1 = 0) views, rows ordered [N_v0][N_u1].
void fbp_filter_device(float* d_reb, int N_v0, int N_u1, int N_u0, float T_u0, float T_v0, float v0_0, hipStream_t s);

}  // namespace petl
