// Forward epilogues and the prior launchers (all enqueue on the given stream).
#pragma once
#include "petl_common.h"

namespace petl {

// forward epilogues
enum Epilogue : int {
    EPI_NONE = 0,   // g = A f                      (times response when given)
    EPI_RATIO = 1,  // g = resp * y / (resp * A f), 0 where the projection is 0 (MLEM ratio)
    EPI_RESID = 2,  // g = (resp * A f - y) * w, the least-squares residual (times the weights w when given)
};

// ---- priors (priors.hip) ----
void launch_rd_gradient(const float* f, float* Df, int N1, int N2, int N3, float delta, float beta, hipStream_t s);
void launch_tv_gradient(const float* f, float* Df, int N1, int N2, int N3, float delta, float beta, float p,
                        hipStream_t s);

}  // namespace petl
