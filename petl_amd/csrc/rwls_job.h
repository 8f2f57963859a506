// Device-resident RWLS (and its special cases WLS, RLS, LS): the preconditioned conjugate-gradient loop of
// PETLserver.RWLS (src/petl_server.py:1691-1877) with every array on the GPU.  The host keeps only the
// scalars that steer the loop (inner products, the step size), read back once or twice per iteration.
#pragma once
#include <vector>
#include "petl_model.h"

namespace petl {

enum RwlsWeights : int {
    RWLS_W_DEFAULT = 0,  // W = 1 / max(gaussian_filter(g, FWHM 3), 7) per planogram
    RWLS_W_ARRAYS = 1,   // caller's planogram weights
    RWLS_W_SCALAR = 2,   // a scalar: gradient unweighted, <Pd, Pd> scaled (the reference's handling)
};

struct RwlsParams {
    int numIter = 0;
    float delta = 1.0f, beta = 0.0f;
    int weights = RWLS_W_DEFAULT;
    float w_scalar = 1.0f;
    bool sqs = false;     // separable quadratic surrogate preconditioner
    bool nonneg = true;   // nonnegativity constraint
    bool response = false;  // projectors include the response (project -> R Pf, backproject(x) -> P*(R x))
    bool cost = false;      // the cost is read after every iteration (keeps the residual unweighted)
};

class RwlsJob {
public:
    RwlsJob(Model& m, const RwlsParams& prm);
    // uploads g (and W, response) and f0 (host arrays), forms the weights, the residual and the
    // preconditioner.  f0 follows the reference's start rules already (negatives clipped when nonneg).
    void setup(float* const* g, const float* f0, bool f0_nonzero, float* const* W, float* const* response,
               hipStream_t s);
    // one iteration; false when the loop ends at a step <= 0 (f unchanged by this call)
    bool iterate(int n, hipStream_t s);
    // 0.5 <r, r>_W + TV cost of the current image (the cost line the reference prints)
    double cost(hipStream_t s);
    void fetch(float* f, hipStream_t s);

private:
    void project(const float* x, std::vector<float*>& out, hipStream_t s);
    // r = R P x - g, times W when weigh (one forward with the residual epilogue)
    void project_residual(const float* x, bool weigh, hipStream_t s);
    void backproject(std::vector<float*>& y, bool y_is_scratch, float* out, hipStream_t s);
    float dot_vol(const float* x, const float* y, int slot, hipStream_t s);
    Model& m_;
    RwlsParams prm_;
    int64_t V_ = 0;
    size_t nP_ = 0;
    std::vector<int64_t> G_;
    std::vector<DevBuf> g_, r_, Pd_, W_, resp_;
    DevBuf f_, grad_, u_, d_, grad_old_, Q_, red_;
    std::vector<float*> gp_, rp_, Pdp_, Wp_, respp_;
    double* slots_ = nullptr;    // device result slots
    double* scratch_ = nullptr;  // reduction partials
    int64_t scratch_cap_ = 0;    // doubles at scratch_
    bool fuse_sumsq_ = false;    // <Pd, Pd>_W straight from the forward (Pd itself is not needed)
    double gg_old_ = 0.0;
    bool r_weighted_ = false;  // r already holds W r (weighting fused into the residual projection)
};

}  // namespace petl
