// Relative-difference and p-Huber TV prior gradients on a 3D array [N1][N2][N3] (N3 fastest).
//
// Reference: src/relative_differences.cu:107-173,481-532 and src/total_variation.cu:434-660.  Unlike
// the reference (threadIdx.x on the slowest axis and a separate clip pass on a device copy), threads run
// along the contiguous axis and the relative-difference clip is applied to each value as it is read, so
// the input is not modified and no copy is made.
#include "petl_common.h"
#include "proj_kernels.h"
#include "priors_dev.h"

namespace petl {

namespace {

template <int KIND>
__global__ void __launch_bounds__(64) prior_gradient_kernel(const float* __restrict__ f, float* __restrict__ Df,
                                                            int N1, int N2, int N3, float delta, float beta,
                                                            HuberParams h)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y;
    const int i = blockIdx.z;
    if (k >= N3)
        return;
    Df[(int64_t(i) * N2 + j) * N3 + k] = prior_grad_at<KIND>(f, i, j, k, N1, N2, N3, delta, beta, h);
}

}  // namespace

void launch_rd_gradient(const float* f, float* Df, int N1, int N2, int N3, float delta, float beta, hipStream_t s)
{
    const dim3 grid(ceil_div(N3, 64), N2, N1);
    HuberParams h{};
    hipLaunchKernelGGL(prior_gradient_kernel<0>, grid, dim3(64), 0, s, f, Df, N1, N2, N3, delta, beta, h);
    PETL_LAUNCH_CHECK();
}

void launch_tv_gradient(const float* f, float* Df, int N1, int N2, int N3, float delta, float beta, float p,
                        hipStream_t s)
{
    const HuberParams h = huber_params(delta, beta, p);
    const dim3 grid(ceil_div(N3, 64), N2, N1);
    hipLaunchKernelGGL(prior_gradient_kernel<1>, grid, dim3(64), 0, s, f, Df, N1, N2, N3, delta, beta, h);
    PETL_LAUNCH_CHECK();
}

}  // namespace petl
