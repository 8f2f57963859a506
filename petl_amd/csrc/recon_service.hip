// Reconstruction service: continuous batching of device-resident MLEM/OSEM requests on one GPU.
//
// Two host threads drive the device:
//   * the admission thread takes queued requests in arrival order, first fit against the free slots and
//     the HBM budget, builds each admitted request's model, buffers and stream, uploads its data, forms
//     its sensitivities, and captures one MLEM iteration of it as a hipGraph;
//   * the launch thread keeps up to two graph replays in flight per active request, round-robin over
//     the requests, and fetches the image when the last iteration has run.
// Requests therefore join and leave the device independently; their kernels overlap through their
// streams.  Finished requests return their memory to the budget, which admits the next ones.
//
// A request runs exactly the kernels MLEM_gpu runs for the same inputs (MlemJob), and nothing in
// them depends on other requests (the back projector gathers without atomics, every reduction has a
// fixed order), so its image is bit-identical alone and in any mix.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>
#include "mlem_job.h"
#include "../../include/petl_c_interface.h"

namespace petl {
namespace {

using Clock = std::chrono::steady_clock;

double seconds(Clock::time_point a, Clock::time_point b) { return std::chrono::duration<double>(b - a).count(); }

struct Request {
    int id = -1;
    petl_recon_request r{};
    std::vector<float*> y;
    size_t bytes = 0;
    Clock::time_point t_submit, t_admit, t_done;
    bool done = false, failed = false;
    std::vector<float> image;
};

struct Slot {
    std::shared_ptr<Request> req;
    std::unique_ptr<Model> model;
    std::unique_ptr<MlemJob> job;
    hipStream_t stream = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipEvent_t ev[2] = {nullptr, nullptr};  // completion of launches k with k % 2 == index
    float* pinned = nullptr;               // image staging (page-locked)
    int launched = 0, finished = 0;
    bool fetching = false;
    std::string sig;                       // geometry and MLEM parameters the resources were built for
    size_t bytes = 0;                      // device bytes the resources hold
    float cap_sy = 0.0f, cap_sf = 0.0f;    // binary16 data / image scales the captured iteration holds
    bool batched = false;                  // runs in the batched iterations, not its own graph
    hipEvent_t fetched = nullptr;          // batched: the image copy to `pinned` is done
};

// Captures one iteration of the slot's job as a graph (the binary16 scales of the request set up last
// become kernel arguments of the graph).
void capture_iteration(Slot& s)
{
    if (s.exec)
        PETL_HIP_CHECK(hipGraphExecDestroy(s.exec));
    if (s.graph)
        PETL_HIP_CHECK(hipGraphDestroy(s.graph));
    s.exec = nullptr;
    s.graph = nullptr;
    PETL_HIP_CHECK(hipStreamBeginCapture(s.stream, hipStreamCaptureModeThreadLocal));
    s.job->iterate(0, s.stream);
    PETL_HIP_CHECK(hipStreamEndCapture(s.stream, &s.graph));
    PETL_HIP_CHECK(hipGraphInstantiate(&s.exec, s.graph, nullptr, nullptr, 0));
    s.cap_sy = s.job->data_scale();
    s.cap_sf = s.job->image_scale();
}

// Requests with equal signatures can run on the same model, buffers and captured iteration.
std::string signature(const petl_recon_request& r)
{
    char buf[512];
    int n = std::snprintf(buf, sizeof(buf), "%d|%a|%a|%a|%a|%a|%a|%d|%d|%d|%a|%a|%d|%d|%a|%a|%a|%d|%d",
                          r.num_planograms, double(r.D), double(r.L), double(r.H), double(r.T), double(r.v_m0),
                          double(r.v_m1), r.numX, r.numY, r.numZ, double(r.voxelWidth), double(r.voxelHeight),
                          r.numSubsets, r.prior, double(r.delta), double(r.beta), double(r.p), r.storage,
                          r.image_storage);
    std::string sig(buf, size_t(std::max(0, n)));
    for (int i = 0; i < r.num_planograms && i < PETL_MAX_PLANOGRAMS; i++) {
        std::snprintf(buf, sizeof(buf), "|%a", double(r.psi[i]));
        sig += buf;
    }
    return sig;
}

std::unique_ptr<Model> make_model(const petl_recon_request& r)
{
    if (r.num_planograms < 1 || r.num_planograms > PETL_MAX_PLANOGRAMS || r.numIter < 0)
        return nullptr;
    auto m = std::make_unique<Model>();
    for (int i = 0; i < r.num_planograms; i++) {
        Planogram p;
        if (!p.init_panels(r.psi[i], r.D, r.L, r.H, r.v_m0, r.v_m1, r.T))
            return nullptr;
        m->plans.push_back(p);
    }
    m->numX = r.numX;
    m->numY = r.numY;
    m->numZ = r.numZ;
    m->voxelWidth = r.voxelWidth;
    m->voxelHeight = r.voxelHeight;
    if (!m->allDefined())
        return nullptr;
    return m;
}

MlemParams params_of(const petl_recon_request& r)
{
    MlemParams prm;
    prm.numIter = r.numIter;
    prm.nsub = r.numSubsets;
    prm.prior = r.prior;
    prm.delta = r.delta;
    prm.beta = r.beta;
    prm.p = r.p;
    prm.storage = r.storage;
    prm.image_storage = r.image_storage;
    return prm;
}

void destroy_slot(Slot& s)
{
    if (s.exec)
        (void)hipGraphExecDestroy(s.exec);
    if (s.graph)
        (void)hipGraphDestroy(s.graph);
    for (hipEvent_t e : s.ev)
        if (e)
            (void)hipEventDestroy(e);
    if (s.fetched)
        (void)hipEventDestroy(s.fetched);
    s.job.reset();
    s.model.reset();  // frees the request's device buffers and kernel tables
    if (s.stream)
        (void)hipStreamDestroy(s.stream);
    if (s.pinned)
        (void)hipHostFree(s.pinned);
    s = Slot();
}

class Service {
public:
    Service(double budget, int max_slots, bool batched) : max_slots_(std::max(1, max_slots))
    {
        if (batched) {
            batch_ = std::make_unique<MlemBatch>();
            PETL_HIP_CHECK(hipStreamCreateWithFlags(&bstream_, hipStreamNonBlocking));
            for (hipEvent_t& e : round_ev_)
                PETL_HIP_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        }
        if (budget <= 0.0) {
            size_t free_b = 0, total_b = 0;
            PETL_HIP_CHECK(hipMemGetInfo(&free_b, &total_b));
            budget = 0.8 * double(free_b);
        }
        budget_ = size_t(budget);
        admit_thread_ = std::thread([this] { admission_loop(); });
        launch_thread_ = std::thread([this] { launch_loop(); });
    }

    ~Service()
    {
        {
            std::lock_guard<std::mutex> lk(mu_);
            stop_ = true;
        }
        cv_admit_.notify_all();
        cv_launch_.notify_all();
        admit_thread_.join();
        launch_thread_.join();
        if (bstream_)
            (void)hipStreamSynchronize(bstream_);
        batch_.reset();
        for (hipEvent_t e : round_ev_)
            if (e)
                (void)hipEventDestroy(e);
        // the batch stream is left to process teardown, like the models' streams
    }

    int submit(const petl_recon_request& r)
    {
        auto model = make_model(r);
        if (!model || !r.y)
            return -1;
        auto req = std::make_shared<Request>();
        req->r = r;
        for (int i = 0; i < r.num_planograms; i++) {
            if (!r.y[i])
                return -1;
            req->y.push_back(r.y[i]);
        }
        req->bytes = MlemJob::device_bytes(*model, params_of(r));
        if (req->bytes > budget_)
            return -1;
        {
            std::lock_guard<std::mutex> lk(mu_);
            req->id = next_id_++;
            req->t_submit = Clock::now();
            all_[req->id] = req;
            pending_.push_back(req);
        }
        cv_admit_.notify_all();
        return req->id;
    }

    bool wait(int id, float* f, double* times)
    {
        std::shared_ptr<Request> req;
        std::unique_lock<std::mutex> lk(mu_);
        auto it = all_.find(id);
        if (it == all_.end())
            return false;
        req = it->second;
        cv_done_.wait(lk, [&] { return req->done; });
        all_.erase(id);
        lk.unlock();
        if (times) {
            times[0] = seconds(req->t_submit, req->t_admit);
            times[1] = seconds(req->t_admit, req->t_done);
            times[2] = seconds(req->t_submit, req->t_done);
        }
        if (req->failed)
            return false;
        if (f)
            std::memcpy(f, req->image.data(), req->image.size() * sizeof(float));
        return true;
    }

    void stats(double* out)
    {
        std::lock_guard<std::mutex> lk(mu_);
        out[0] = double(completed_);
        out[1] = double(failed_);
        out[2] = double(peak_active_);
        out[3] = double(peak_reserved_);
        out[4] = double(budget_);
        out[5] = double(pending_.size());
        out[6] = t_alloc_;
        out[7] = t_setup_;
        out[8] = t_capture_;
        out[9] = t_release_;
        out[10] = double(pool_hits_);
        out[11] = double(recaptures_);
        out[12] = double(batched_requests_);
        out[13] = double(batch_rounds_);
        out[14] = double(batch_peak_);
    }

private:
    // ---------------------------------------------------------------- admission thread
    void admission_loop()
    {
        for (;;) {
            std::shared_ptr<Request> req;
            std::vector<Slot> retire;
            {
                std::unique_lock<std::mutex> lk(mu_);
                cv_admit_.wait_for(lk, std::chrono::milliseconds(5), [&] {
                    return stop_ || !retired_.empty() || admissible_locked() >= 0;
                });
                retire.swap(retired_);
                const int k = admissible_locked();
                if (k >= 0) {
                    req = pending_[size_t(k)];
                    pending_.erase(pending_.begin() + k);
                    reserved_ += req->bytes;
                    active_++;
                    peak_active_ = std::max(peak_active_, active_);
                    peak_reserved_ = std::max(peak_reserved_, reserved_);
                } else if (stop_ && pending_.empty() && active_ == 0 && retire.empty()) {
                    std::vector<Slot> pool;
                    pool.swap(pool_);
                    pooled_ = 0;
                    lk.unlock();
                    for (Slot& p : pool)
                        destroy_slot(p);
                    return;
                }
            }
            for (Slot& s : retire)
                release(s);
            if (req)
                admit(req);
        }
    }

    // first queued request (arrival order) that fits the free slots and the budget; -1 if none
    int admissible_locked() const
    {
        if (active_ >= max_slots_)
            return -1;
        for (size_t i = 0; i < pending_.size(); i++)
            if (reserved_ + pending_[i]->bytes <= budget_)
                return int(i);
        return -1;
    }

    // A finished request's resources go to the pool (its memory stays counted in the budget as pooled),
    // unless they failed or the pool is full; freeing device memory waits for the whole device, so it is
    // done only when a new request needs the room.
    void release(Slot& s)
    {
        const size_t bytes = s.req->bytes;
        const bool keep = (s.exec != nullptr || s.batched) && !s.req->failed && int(pool_.size()) < max_slots_;
        const Clock::time_point t0 = Clock::now();
        if (!keep)
            destroy_slot(s);
        const double dt = seconds(t0, Clock::now());
        {
            std::lock_guard<std::mutex> lk(mu_);
            reserved_ -= bytes;
            active_--;
            t_release_ += dt;
            if (keep) {
                s.bytes = bytes;
                pooled_ += bytes;
                s.req.reset();
                pool_.push_back(std::move(s));
            }
        }
        cv_admit_.notify_all();
    }

    // frees pooled resources, oldest first, until `need` more bytes fit the budget beside the pool
    void evict_for(size_t need)
    {
        for (;;) {
            Slot victim;
            {
                std::lock_guard<std::mutex> lk(mu_);
                if (pool_.empty() || reserved_ + pooled_ <= budget_)
                    return;
                victim = std::move(pool_.front());
                pool_.erase(pool_.begin());
                pooled_ -= victim.bytes;
            }
            (void)need;
            const Clock::time_point t0 = Clock::now();
            destroy_slot(victim);
            std::lock_guard<std::mutex> lk(mu_);
            t_release_ += seconds(t0, Clock::now());
        }
    }

    // takes a pooled slot built for `sig`, if any
    bool take_pooled(const std::string& sig, Slot& out)
    {
        std::lock_guard<std::mutex> lk(mu_);
        for (size_t i = 0; i < pool_.size(); i++)
            if (pool_[i].sig == sig) {
                out = std::move(pool_[i]);
                pool_.erase(pool_.begin() + long(i));
                pooled_ -= out.bytes;
                pool_hits_++;
                return true;
            }
        return false;
    }

    void admit(const std::shared_ptr<Request>& req)
    {
        Slot s;
        req->t_admit = Clock::now();
        const std::string sig = signature(req->r);
        if (take_pooled(sig, s)) {
            s.req = req;
            s.launched = s.finished = 0;
            s.fetching = false;
            try {
                const Clock::time_point t1 = Clock::now();
                const int64_t V = s.model->vol_elems();
                std::fill(s.pinned, s.pinned + V, 1.0f);
                s.job->setup(req->y.data(), s.pinned, nullptr, nullptr, s.stream);
                PETL_HIP_CHECK(hipStreamSynchronize(s.stream));
                const Clock::time_point t2 = Clock::now();
                // binary16 requests pick their scales from their data: a graph captured with another
                // request's scales is captured again
                const bool recapture =
                    !s.batched && (s.job->data_scale() != s.cap_sy || s.job->image_scale() != s.cap_sf);
                if (recapture)
                    capture_iteration(s);
                std::lock_guard<std::mutex> lk(mu_);
                t_setup_ += seconds(t1, t2);
                t_capture_ += seconds(t2, Clock::now());
                recaptures_ += recapture ? 1 : 0;
                batched_requests_ += s.batched ? 1 : 0;
            } catch (const std::exception& e) {
                fprintf(stderr, "PETL service: request %d failed at admission: %s\n", req->id, e.what());
                (void)hipStreamSynchronize(s.stream);
                (void)hipGetLastError();
                finish(req, true);
                release(s);
                return;
            }
            {
                std::lock_guard<std::mutex> lk(mu_);
                ready_.push_back(std::move(s));
            }
            cv_launch_.notify_all();
            return;
        }
        evict_for(req->bytes);
        s.req = req;
        s.sig = sig;
        try {
            const petl_recon_request& r = req->r;
            const Clock::time_point t0 = Clock::now();
            s.model = make_model(r);
            if (!s.model)
                throw std::runtime_error("invalid request");
            s.job = std::make_unique<MlemJob>(*s.model, params_of(r));
            s.job->allocate();
            PETL_HIP_CHECK(hipStreamCreateWithFlags(&s.stream, hipStreamNonBlocking));
            for (hipEvent_t& e : s.ev)
                PETL_HIP_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
            const int64_t V = s.model->vol_elems();
            PETL_HIP_CHECK(hipHostMalloc(reinterpret_cast<void**>(&s.pinned), size_t(V) * sizeof(float)));
            std::fill(s.pinned, s.pinned + V, 1.0f);  // initial estimate: ones
            const Clock::time_point t1 = Clock::now();
            s.job->setup(req->y.data(), s.pinned, nullptr, nullptr, s.stream);
            PETL_HIP_CHECK(hipStreamSynchronize(s.stream));
            const Clock::time_point t2 = Clock::now();
            // requests the batched iterations take join them; the others replay a graph of their own
            s.batched = batch_ && s.job->batch_eligible();
            if (s.batched)
                PETL_HIP_CHECK(hipEventCreateWithFlags(&s.fetched, hipEventDisableTiming));
            else
                capture_iteration(s);
            const Clock::time_point t3 = Clock::now();
            std::lock_guard<std::mutex> lk(mu_);
            t_alloc_ += seconds(t0, t1);
            t_setup_ += seconds(t1, t2);
            t_capture_ += seconds(t2, t3);
            batched_requests_ += s.batched ? 1 : 0;
        } catch (const std::exception& e) {
            fprintf(stderr, "PETL service: request %d failed at admission: %s\n", req->id, e.what());
            if (s.stream) {
                hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
                if (hipStreamIsCapturing(s.stream, &st) == hipSuccess && st != hipStreamCaptureStatusNone) {
                    hipGraph_t g = nullptr;
                    (void)hipStreamEndCapture(s.stream, &g);
                    if (g)
                        (void)hipGraphDestroy(g);
                }
                (void)hipStreamSynchronize(s.stream);
            }
            (void)hipGetLastError();
            finish(req, true);
            release(s);
            return;
        }
        {
            std::lock_guard<std::mutex> lk(mu_);
            ready_.push_back(std::move(s));
        }
        cv_launch_.notify_all();
    }

    void finish(const std::shared_ptr<Request>& req, bool failed)
    {
        {
            std::lock_guard<std::mutex> lk(mu_);
            req->t_done = Clock::now();
            req->failed = failed;
            req->done = true;
            if (failed)
                failed_++;
            else
                completed_++;
        }
        cv_done_.notify_all();
    }

    // ---------------------------------------------------------------- launch thread
    // Batched slots: one round = one iteration of every batched slot that has iterations left (up to
    // BATCH_MAX_REQUESTS), enqueued on the batch stream; at most two rounds in flight.  A slot whose last
    // iteration is enqueued has its image copied after that round, on the same stream.  Sets retire[i] for
    // the batched slots that are finished.
    bool step_batched(std::vector<Slot>& active, std::vector<char>& retire)
    {
        bool progress = false;
        while (!rounds_.empty()) {
            const hipError_t q = hipEventQuery(round_ev_[rounds_.front() % ROUND_EVENTS]);
            if (q == hipErrorNotReady) {
                (void)hipGetLastError();
                break;
            }
            PETL_HIP_CHECK(q);
            rounds_.pop_front();
            progress = true;
        }
        for (size_t i = 0; i < active.size(); i++) {
            Slot& s = active[i];
            if (!s.batched || !s.fetching)
                continue;
            const hipError_t q = hipEventQuery(s.fetched);
            if (q == hipErrorNotReady) {
                (void)hipGetLastError();
                continue;
            }
            PETL_HIP_CHECK(q);
            s.req->image.assign(s.pinned, s.pinned + s.model->vol_elems());
            finish(s.req, false);
            retire[i] = 1;
            progress = true;
        }
        if (rounds_.size() >= 2)
            return progress;
        std::vector<BatchIterJob> jobs;
        std::vector<Slot*> in;
        for (Slot& s : active) {
            if (!s.batched || s.fetching)
                continue;
            if (s.launched < s.req->r.numIter && int(jobs.size()) < BATCH_MAX_REQUESTS) {
                jobs.push_back(s.job->batch_job(s.launched));
                in.push_back(&s);
            } else if (s.launched >= s.req->r.numIter) {
                in.push_back(&s);  // no iterations (left): only the image copy
            }
        }
        if (in.empty())
            return progress;
        if (!jobs.empty()) {
            batch_->launch(jobs, bstream_);
            const long k = round_count_++;
            PETL_HIP_CHECK(hipEventRecord(round_ev_[k % ROUND_EVENTS], bstream_));
            rounds_.push_back(k);
            std::lock_guard<std::mutex> lk(mu_);
            batch_rounds_++;
            batch_peak_ = std::max(batch_peak_, long(jobs.size()));
        }
        for (Slot* s : in) {
            const int n = s->req->r.numIter;
            if (s->launched < n)
                s->launched++;
            if (s->launched >= n) {
                d2h(s->pinned, s->job->batch_image(n), size_t(s->model->vol_elems()) * sizeof(float), bstream_);
                PETL_HIP_CHECK(hipEventRecord(s->fetched, bstream_));
                s->fetching = true;
            }
        }
        return true;
    }

    // returns true when the slot made progress; sets `retire` when it is finished
    bool step(Slot& s, bool& retire)
    {
        bool progress = false;
        const int n = s.req->r.numIter;
        while (s.finished < s.launched) {
            const hipError_t q = hipEventQuery(s.ev[s.finished % 2]);
            if (q == hipErrorNotReady) {
                (void)hipGetLastError();
                break;
            }
            PETL_HIP_CHECK(q);
            s.finished++;
            progress = true;
        }
        if (!s.fetching) {
            while (s.launched < n && s.launched - s.finished < 2) {
                PETL_HIP_CHECK(hipGraphLaunch(s.exec, s.stream));
                PETL_HIP_CHECK(hipEventRecord(s.ev[s.launched % 2], s.stream));
                s.launched++;
                progress = true;
            }
            if (s.finished == n) {
                s.job->fetch(s.pinned, nullptr, s.stream);
                PETL_HIP_CHECK(hipEventRecord(s.ev[0], s.stream));
                s.fetching = true;
                progress = true;
            }
        } else {
            const hipError_t q = hipEventQuery(s.ev[0]);
            if (q == hipErrorNotReady) {
                (void)hipGetLastError();
            } else {
                PETL_HIP_CHECK(q);
                const int64_t V = s.model->vol_elems();
                s.req->image.assign(s.pinned, s.pinned + V);
                finish(s.req, false);
                retire = true;
                progress = true;
            }
        }
        return progress;
    }

    void launch_loop()
    {
        std::vector<Slot> active;
        for (;;) {
            {
                std::unique_lock<std::mutex> lk(mu_);
                for (Slot& s : ready_)
                    active.push_back(std::move(s));
                ready_.clear();
                if (stop_ && active.empty() && pending_.empty() && active_ == 0)
                    return;
            }
            bool progress = false;
            std::vector<Slot> done;
            std::vector<char> bretire(active.size(), 0);
            if (batch_) {
                try {
                    progress |= step_batched(active, bretire);
                } catch (const std::exception& e) {
                    // a failed batched launch fails every batched request on the device
                    fprintf(stderr, "PETL service: batched iteration failed: %s\n", e.what());
                    (void)hipStreamSynchronize(bstream_);
                    (void)hipGetLastError();
                    rounds_.clear();
                    for (size_t i = 0; i < active.size(); i++)
                        if (active[i].batched && !bretire[i]) {
                            finish(active[i].req, true);
                            bretire[i] = 1;
                        }
                }
            }
            for (size_t i = 0, k = 0; i < active.size(); k++) {
                bool retire = false;
                if (active[i].batched) {
                    if (bretire[k]) {
                        done.push_back(std::move(active[i]));
                        active.erase(active.begin() + long(i));
                    } else {
                        i++;
                    }
                    continue;
                }
                try {
                    progress |= step(active[i], retire);
                } catch (const std::exception& e) {
                    fprintf(stderr, "PETL service: request %d failed: %s\n", active[i].req->id, e.what());
                    (void)hipStreamSynchronize(active[i].stream);
                    (void)hipGetLastError();
                    finish(active[i].req, true);
                    retire = true;
                }
                if (retire) {
                    done.push_back(std::move(active[i]));
                    active.erase(active.begin() + long(i));
                } else {
                    i++;
                }
            }
            if (!done.empty()) {
                {
                    std::lock_guard<std::mutex> lk(mu_);
                    for (Slot& s : done)
                        retired_.push_back(std::move(s));
                }
                cv_admit_.notify_all();
            }
            if (!progress) {
                std::unique_lock<std::mutex> lk(mu_);
                cv_launch_.wait_for(lk, std::chrono::microseconds(active.empty() ? 2000 : 50),
                                    [&] { return stop_ || !ready_.empty(); });
            }
        }
    }

    const int max_slots_;
    size_t budget_ = 0;
    std::mutex mu_;
    std::condition_variable cv_admit_, cv_launch_, cv_done_;
    bool stop_ = false;
    int next_id_ = 0;
    std::map<int, std::shared_ptr<Request>> all_;
    std::deque<std::shared_ptr<Request>> pending_;
    std::vector<Slot> ready_, retired_;
    size_t reserved_ = 0, peak_reserved_ = 0;
    int active_ = 0, peak_active_ = 0;
    long completed_ = 0, failed_ = 0;
    double t_alloc_ = 0, t_setup_ = 0, t_capture_ = 0, t_release_ = 0;  // host seconds spent per phase
    std::vector<Slot> pool_;  // resources of finished requests, oldest first
    size_t pooled_ = 0;       // device bytes they hold
    long pool_hits_ = 0;
    long recaptures_ = 0;  // pooled iterations captured again for other binary16 scales
    long batched_requests_ = 0, batch_rounds_ = 0, batch_peak_ = 0;
    // batched iterations (null when the service was created without them)
    static constexpr int ROUND_EVENTS = 4;
    std::unique_ptr<MlemBatch> batch_;
    hipStream_t bstream_ = nullptr;
    hipEvent_t round_ev_[ROUND_EVENTS] = {};
    std::deque<long> rounds_;  // batched rounds in flight (launch thread only)
    long round_count_ = 0;
    std::thread admit_thread_, launch_thread_;
};

template <typename Fn>
auto guarded_call(const char* what, Fn&& fn, decltype(fn()) fail) -> decltype(fn())
{
    try {
        return fn();
    } catch (const std::exception& e) {
        fprintf(stderr, "PETL %s failed: %s\n", what, e.what());
    } catch (...) {
        fprintf(stderr, "PETL %s failed\n", what);
    }
    return fail;
}

}  // namespace
}  // namespace petl

using petl::Service;

extern "C" {

void* petl_service_create(double hbm_budget_bytes, int max_slots)
{
    return petl::guarded_call("petl_service_create",
                              [&]() -> void* { return new Service(hbm_budget_bytes, max_slots, false); }, nullptr);
}

void* petl_service_create_ex(double hbm_budget_bytes, int max_slots, int flags)
{
    return petl::guarded_call(
        "petl_service_create_ex",
        [&]() -> void* { return new Service(hbm_budget_bytes, max_slots, (flags & PETL_SERVICE_BATCHED) != 0); },
        nullptr);
}

int petl_service_submit(void* service, const petl_recon_request* request)
{
    if (!service || !request)
        return -1;
    return petl::guarded_call("petl_service_submit",
                              [&] { return static_cast<Service*>(service)->submit(*request); }, -1);
}

double petl_service_request_bytes(const petl_recon_request* request)
{
    if (!request)
        return 0.0;
    return petl::guarded_call("petl_service_request_bytes", [&] {
        auto m = petl::make_model(*request);
        return m ? double(petl::MlemJob::device_bytes(*m, petl::params_of(*request))) : 0.0;
    }, 0.0);
}

bool petl_service_request_footprint(const petl_recon_request* request, double* out)
{
    if (!request || !out)
        return false;
    return petl::guarded_call("petl_service_request_footprint", [&] {
        auto m = petl::make_model(*request);
        if (!m)
            return false;
        petl::MlemJob job(*m, petl::params_of(*request));
        out[0] = double(petl::MlemJob::device_bytes(*m, petl::params_of(*request)));
        job.allocate();
        out[1] = double(job.allocated_bytes());
        return true;
    }, false);
}

bool petl_service_wait(void* service, int id, float* f, double* times)
{
    if (!service)
        return false;
    return petl::guarded_call("petl_service_wait",
                              [&] { return static_cast<Service*>(service)->wait(id, f, times); }, false);
}

bool petl_service_stats(void* service, double* stats)
{
    if (!service || !stats)
        return false;
    static_cast<Service*>(service)->stats(stats);
    return true;
}

void petl_service_destroy(void* service)
{
    delete static_cast<Service*>(service);
}

}  // extern "C"
