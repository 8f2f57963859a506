// Launchers for the projector, prior and vector kernels (all enqueue on the given stream).
#pragma once
#include "petl_common.h"

namespace petl {

// forward epilogues
enum Epilogue : int {
    EPI_NONE = 0,   // g = A f                      (times response when given)
    EPI_RATIO = 1,  // g = resp * y / (resp * A f), 0 where the projection is 0 (MLEM ratio)
    EPI_RESID = 2,  // g = (resp * A f - y) * w, the least-squares residual (times the weights w when given)
};

// ---- generic kernels: any psi, any pitch (proj_generic.hip) ----
// Views are restricted to v1 rows first, first+stride, ... (OSEM subsets); (0,1) = all.
void launch_fwd_generic(float* g, const float* f, const PlanGeom& P, const VolGeom& V, int epilogue,
                        const float* y, const float* resp, hipStream_t s, int v1_first = 0, int v1_stride = 1,
                        const float* wts = nullptr);
// f += A^T g (mode BACK_PETL or BACK_MATCHED)
void launch_back_generic(const float* g, float* f, const PlanGeom& P, const VolGeom& V, int mode, hipStream_t s,
                         int v1_first = 0, int v1_stride = 1);

// ---- priors (priors.hip) ----
void launch_rd_gradient(const float* f, float* Df, int N1, int N2, int N3, float delta, float beta, hipStream_t s);
void launch_tv_gradient(const float* f, float* Df, int N1, int N2, int N3, float delta, float beta, float p,
                        hipStream_t s);

}  // namespace petl
