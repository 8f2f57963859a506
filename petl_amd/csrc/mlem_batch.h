// Batched MLEM iterations: one launch per phase of an iteration advances every request in the batch,
// whatever its panel and volume size (the serving path's batched form).
//
// An iteration is: the forward with the ratio epilogue of every (request, planogram); for planogram 0 then
// 1, the back projector's u0 gather and u1 gather; the image update.  Each phase is one launch (one more per
// group of requests beyond what one launch's kernel arguments hold) whose flat grid covers the work items of
// all requests: a workgroup finds its request from the prefix sums of the requests' item counts in the
// kernel arguments, where the requests' geometry and buffers travel too.  Every work item runs the same
// code as the per-request kernels (shared bodies), so a request's image does not depend on the batch it
// runs in.
#pragma once
#include <vector>
#include "proj_fast.h"

namespace petl {

constexpr int BATCH_MAX_REQUESTS = 32;  // requests per batched iteration
constexpr int BATCH_MAX_PLANOGRAMS = 2; // planograms per request in the batched form

// one request's iteration
struct BatchIterJob {
    int nplans = 0;
    const FastPlan* plan[BATCH_MAX_PLANOGRAMS] = {};
    float* ratio[BATCH_MAX_PLANOGRAMS] = {};       // forward ratio output = back input (planogram-sized)
    const float* data[BATCH_MAX_PLANOGRAMS] = {};  // measured data y
    const float* f_in = nullptr;                   // image the iteration starts from
    float* f_out = nullptr;                        // image it writes (= f_in without a prior)
    float* d = nullptr;                            // back projection (volume-sized)
    const float* S = nullptr;                      // sensitivity (zeros replaced by ones)
    float* h2 = nullptr;                           // back projector workspace of this request
    int prior = 0;                                 // 0 none, 1 relative differences, 2 p-Huber TV (OSL)
    float delta = 2.0f, beta = 0.0f, p = 1.0f;
    int Nx = 0, Ny = 0, Nz = 0;
};

class MlemBatch {
public:
    MlemBatch();
    ~MlemBatch();
    MlemBatch(const MlemBatch&) = delete;
    MlemBatch& operator=(const MlemBatch&) = delete;
    // whether a request whose planograms have these fast plans (null: not a fast geometry) can run batched:
    // at most BATCH_MAX_PLANOGRAMS planograms, each on the fast kernels with at most 255 u0 bins, in-plane
    // and slice axes of at most 128 voxels
    static bool eligible(const std::vector<const FastPlan*>& plans);
    // enqueues one iteration of every job on s (at most BATCH_MAX_REQUESTS)
    void launch(const std::vector<BatchIterJob>& jobs, hipStream_t s);

private:
    struct Impl;
    Impl* impl_;
};

}  // namespace petl
