// Device-resident MLEM / OSEM reconstruction of one model: the loop of src/petl_server.py:1534-1606
// (sensitivity src/petl_server.py:1257-1273) with every step on the GPU.  Shared by the synchronous
// MLEM_gpu export and the reconstruction service, so a request gives the same image either way.
#pragma once
#include <vector>
#include "petl_model.h"
#include "storage.h"
#include "mlem_batch.h"

namespace petl {

struct MlemParams {
    int numIter = 0;
    float delta = 2.0f, beta = 0.0f, p = 1.0f;
    int prior = 0;  // 0 none, 1 relative differences (OSL), 2 p-Huber TV (OSL)
    int nsub = 1;   // OSEM subsets over interleaved v1 rows (clamped to [1, min N_v1])
    bool loglik = false;
    bool mask = false, response = false;
    // storage of the planograms (data and ratio) and of the image (STORE_*): equal, or an fp32 image with
    // 16-bit planograms.  16-bit formats need the fast kernels and take no mask, response or
    // log-likelihood.  Sensitivities and back projections stay fp32.
    int storage = STORE_F32;
    int image_storage = STORE_F32;
};

// Initial estimate rules of src/petl_server.py:1571-1577 on a host volume: all zeros -> ones,
// otherwise negatives -> 0.
void mlem_initial_estimate(float* f, int64_t n);

class MlemJob {
public:
    MlemJob(Model& m, const MlemParams& prm);
    // device bytes the job holds (buffers + back projector workspace + kernel tables)
    static size_t device_bytes(Model& m, const MlemParams& prm);
    // allocates every buffer the job needs; no kernel runs
    void allocate();
    // enqueues on s: data, mask, response and the initial estimate f0 (host arrays; the caller keeps
    // them alive until s has passed this point), then the per-subset sensitivities
    void setup(float* const* g, const float* f0, float* const* mask, float* const* response, hipStream_t s);
    // enqueues one iteration (every subset) on s; it = log-likelihood slot when enabled.  The buffers
    // it touches are fixed, so the enqueued work can be captured once and replayed.
    void iterate(int it, hipStream_t s);
    // enqueues the copy of the image (and log-likelihoods) to host memory on s
    void fetch(float* f, double* loglik, hipStream_t s);
    const MlemParams& params() const { return prm_; }
    // the binary16 scales of the data and of the image chosen by the last setup(); a captured iteration
    // holds them as kernel arguments, so it can be replayed for another request only when they are equal
    float data_scale() const { return s_y_; }
    float image_scale() const { return s_f_; }
    // whether the job can run in batched iterations (MlemBatch): fp32 MLEM (one subset, no mask,
    // response or log-likelihood) in PETL mode on fast plans that MlemBatch takes; call after allocate()
    bool batch_eligible();
    // iteration `it` of the job as a batched-iteration entry (after setup()).  With a prior the image
    // alternates between two buffers (the update reads one and writes the other); batch_image(n) is the
    // buffer that holds the image after n batched iterations.
    BatchIterJob batch_job(int it);
    const float* batch_image(int n);
    // device bytes the job's buffers hold now (planogram and volume buffers, back projector workspace,
    // staging; not the kernel tables)
    size_t allocated_bytes() const;

private:
    void setup_storage(float* const* g, const float* f0, hipStream_t s);
    void iterate_storage(hipStream_t s);
    Model& m_;
    MlemParams prm_;
    int64_t V_ = 0;
    float s_y_ = 1.0f, s_f_ = 1.0f;  // binary16 scales of the data and of the image
    double* dll_ = nullptr;
    double* ll_scratch_ = nullptr;
};

// host <-> device helpers shared by the exports
void ensure_plan_bufs(Model& m, std::vector<DevBuf>& bufs, size_t elem_bytes = sizeof(float));
std::vector<float*> dev_ptrs(std::vector<DevBuf>& bufs);
void h2d(void* d, const void* h, size_t n, hipStream_t s);
void d2h(void* h, const void* d, size_t n, hipStream_t s);

}  // namespace petl
