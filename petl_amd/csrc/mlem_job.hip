// Device-resident MLEM / OSEM (src/petl_server.py:1534-1606, 1257-1273).
#include <algorithm>
#include <cmath>
#include "mlem_job.h"
#include "proj_general.h"
#include "proj_kernels.h"
#include "vec_kernels.h"

namespace petl {

void ensure_plan_bufs(Model& m, std::vector<DevBuf>& bufs, size_t elem_bytes)
{
    bufs.resize(m.plans.size());
    for (size_t i = 0; i < m.plans.size(); i++)
        bufs[i].ensure(size_t(m.plans[i].elems()) * elem_bytes);
}

std::vector<float*> dev_ptrs(std::vector<DevBuf>& bufs)
{
    std::vector<float*> p(bufs.size());
    for (size_t i = 0; i < bufs.size(); i++)
        p[i] = bufs[i].f();
    return p;
}

void h2d(void* d, const void* h, size_t n, hipStream_t s)
{
    PETL_HIP_CHECK(hipMemcpyAsync(d, h, n, hipMemcpyHostToDevice, s));
}

void d2h(void* h, const void* d, size_t n, hipStream_t s)
{
    PETL_HIP_CHECK(hipMemcpyAsync(h, d, n, hipMemcpyDeviceToHost, s));
}

void mlem_initial_estimate(float* f, int64_t n)
{
    bool all_zero = true;
    for (int64_t i = 0; i < n; i++)
        if (f[i] != 0.0f) {
            all_zero = false;
            break;
        }
    if (all_zero)
        std::fill(f, f + n, 1.0f);
    else
        for (int64_t i = 0; i < n; i++)
            if (f[i] < 0.0f)
                f[i] = 0.0f;
}

static int clamp_subsets(const Model& m, int nsub)
{
    int nv1 = m.plans[0].N_v1;
    for (const auto& pl : m.plans)
        nv1 = std::min(nv1, pl.N_v1);
    return std::max(1, std::min(nsub, nv1));
}

MlemJob::MlemJob(Model& m, const MlemParams& prm) : m_(m), prm_(prm)
{
    prm_.nsub = clamp_subsets(m, prm.nsub);
    V_ = m.vol_elems();
    const bool st16 = prm_.storage != STORE_F32 || prm_.image_storage != STORE_F32;
    if (st16 && (prm_.mask || prm_.response || prm_.loglik))
        throw hip_error("16-bit storage takes no mask, response or log-likelihood");
    if (prm_.storage < STORE_F32 || prm_.storage > STORE_F16 ||
        !(prm_.image_storage == prm_.storage || prm_.image_storage == STORE_F32))
        throw hip_error("unsupported storage formats (planograms and image equal, or an fp32 image)");
}

// The staging buffer of 16-bit uploads: fp32 chunks of the data (2^24 values), then the initial image.
static size_t stage_bytes(int64_t V) { return size_t(std::max<int64_t>(int64_t(1) << 24, V)) * sizeof(float); }

size_t MlemJob::device_bytes(Model& m, const MlemParams& prm)
{
    const int nsub = clamp_subsets(m, prm.nsub);
    size_t G = 0, back_ws = 0;
    for (const auto& pl : m.plans)
        G += size_t(pl.elems());
    const VolGeom V = m.vgeom();
    for (const auto& pl : m.plans)  // the fast back projector's workspace, the largest of the planograms
        if (m.path == 0)
            back_ws = std::max(back_ws, fast_workspace_estimate(pl.geom(), V));
    // the general forward's copies of the volume: the transposed one (per-bin kernels) and the two padded
    // ones (separable kernels); which of them a model's planograms use is decided when their plans are built,
    // so both are counted
    // and the separable PETL back's partial volumes (the largest of the general planograms')
    size_t vol_t = 0, gen_ws = 0;
    for (const auto& pl : m.plans)
        if (m.path != 0 || fast_workspace_estimate(pl.geom(), V) == 0) {
            vol_t = (size_t(m.vol_elems()) + 2 * gen_padded_elems(V)) * sizeof(float);
            gen_ws = std::max(gen_ws, gen_back_workspace_estimate(pl.geom(), V));
        }
    vol_t += gen_ws;
    const size_t vol = size_t(m.vol_elems()) * sizeof(float);
    const size_t es = storage_bytes(prm.storage), ei = storage_bytes(prm.image_storage);
    const bool st16 = prm.storage != STORE_F32 || prm.image_storage != STORE_F32;
    // y and aux in storage format; response and the log-likelihood's projection fp32
    size_t plan_bytes = G * es * 2 + G * sizeof(float) * ((prm.response ? 1 : 0) + (prm.loglik ? 1 : 0));
    // image and its update buffer in storage format; back projection and sensitivities fp32; fp32
    // staging of 16-bit uploads; 1 MiB for the kernel tables and reduction scratch
    return plan_bytes + size_t(m.vol_elems()) * ei * 2 + vol * (1 + size_t(nsub)) + back_ws + vol_t +
           (st16 ? stage_bytes(m.vol_elems()) : 0) + (size_t(1) << 20);
}

size_t MlemJob::allocated_bytes() const
{
    const DeviceState& D = m_.dev;
    size_t n = D.f.bytes + D.f2.bytes + D.d.bytes + D.S.bytes + D.h2.bytes + D.stage.bytes + D.red.bytes + D.fT.bytes +
               D.fpad.bytes + D.genws.bytes;
    for (const auto* v : {&D.y, &D.aux, &D.resp, &D.g})
        for (const DevBuf& b : *v)
            n += b.bytes;
    return n;
}

void MlemJob::allocate()
{
    Model& m = m_;
    DeviceState& D = m.dev;
    m.prepare();
    m.ensure_workspace();
    m.ensure_general_buffers();  // a captured iteration must find them allocated
    if (prm_.storage != STORE_F32 || prm_.image_storage != STORE_F32)
        for (size_t i = 0; i < m.plans.size(); i++)
            if (!m.uses_fast(int(i)))
                throw hip_error("16-bit storage needs the fast kernels (psi 0 or 90 degrees, voxel pitch = crystal "
                                "pitch)");
    // planogram and image buffers hold storage-format elements
    ensure_plan_bufs(m, D.y, storage_bytes(prm_.storage));
    ensure_plan_bufs(m, D.aux, storage_bytes(prm_.storage));
    if (prm_.response)
        ensure_plan_bufs(m, D.resp);
    D.f.ensure(V_ * storage_bytes(prm_.image_storage));
    D.f2.ensure(V_ * storage_bytes(prm_.image_storage));
    if (prm_.storage != STORE_F32 || prm_.image_storage != STORE_F32)
        D.stage.ensure(stage_bytes(V_));
    D.d.ensure(V_ * sizeof(float));
    D.S.ensure(size_t(prm_.nsub) * V_ * sizeof(float));
    if (prm_.loglik) {
        int64_t max_el = 0;
        for (const auto& pl : m.plans)
            max_el = std::max<int64_t>(max_el, pl.elems());
        D.red.ensure(sizeof(double) * (size_t(prm_.numIter) + 1 + size_t(ll_scratch_doubles(max_el))));
        dll_ = static_cast<double*>(D.red.p);
        ll_scratch_ = dll_ + prm_.numIter + 1;
        ensure_plan_bufs(m, D.g);
    }
}

void MlemJob::setup(float* const* g, const float* f0, float* const* mask, float* const* response, hipStream_t s)
{
    if (prm_.storage != STORE_F32 || prm_.image_storage != STORE_F32) {
        setup_storage(g, f0, s);
        return;
    }
    Model& m = m_;
    DeviceState& D = m.dev;
    const size_t nP = m.plans.size();
    std::vector<float*> y = dev_ptrs(D.y), aux = dev_ptrs(D.aux);
    std::vector<float*> resp;
    if (prm_.response)
        resp = dev_ptrs(D.resp);
    for (size_t i = 0; i < nP; i++) {
        const size_t nb = size_t(m.plans[i].elems()) * sizeof(float);
        h2d(y[i], g[i], nb, s);
        if (prm_.response)
            h2d(resp[i], response[i], nb, s);
    }
    h2d(D.f.p, f0, V_ * sizeof(float), s);
    // sensitivity per subset: back(mask or ones, times response), zeros -> 1 (src/petl_server.py:1580-1585)
    for (int sub = 0; sub < prm_.nsub; sub++) {
        for (size_t i = 0; i < nP; i++) {
            const int64_t n = m.plans[i].elems();
            if (prm_.mask)
                h2d(aux[i], mask[i], n * sizeof(float), s);
            else
                launch_fill(aux[i], 1.0f, n, s);
            if (prm_.response)
                launch_mul(aux[i], resp[i], n, s);
        }
        float* Ss = D.S.f() + int64_t(sub) * V_;
        m.back(aux.data(), Ss, m.back_mode, s, sub, prm_.nsub);
        launch_sens_fix(Ss, V_, s);
    }
}

// 16-bit storage: data converted on the host (binary16: per-tensor power-of-two scale from the largest
// count), sensitivities as in fp32, then the initial image stored (binary16: with a scale that places
// 1024 x the expected mean image value sum(y) / sum(S), and the initial values, inside its range).
void MlemJob::setup_storage(float* const* g, const float* f0, hipStream_t s)
{
    Model& m = m_;
    DeviceState& D = m.dev;
    const int st = prm_.storage;
    const size_t nP = m.plans.size();
    std::vector<float*> aux = dev_ptrs(D.aux);
    const int sti = prm_.image_storage;
    // the data's largest count (binary16 data scale) and total (binary16 image scale)
    float ymax = 0.0f;
    double ysum = 0.0;
    if (st == STORE_F16 || sti == STORE_F16)
        for (size_t i = 0; i < nP; i++) {
            const float* gi = g[i];
            const int64_t n = m.plans[i].elems();
#pragma omp parallel for reduction(max : ymax) reduction(+ : ysum) schedule(static)
            for (int64_t k = 0; k < n; k++) {
                ymax = std::max(ymax, std::fabs(gi[k]));
                ysum += gi[k];
            }
        }
    s_y_ = (st == STORE_F16) ? pow2_scale_for(ymax, 16384.0f) : 1.0f;
    // data: fp32 chunks through a device staging buffer, converted on the device
    const int64_t chunk = int64_t(1) << 24;
    D.stage.ensure(stage_bytes(V_));
    for (size_t i = 0; i < nP; i++) {
        const int64_t n = m.plans[i].elems();
        for (int64_t o = 0; o < n; o += chunk) {
            const int64_t k = std::min(chunk, n - o);
            h2d(D.stage.p, g[i] + o, size_t(k) * sizeof(float), s);
            launch_convert_st(st, static_cast<char*>(D.y[i].p) + size_t(o) * storage_bytes(st), D.stage.f(), k,
                              1.0f / s_y_, s);
        }
    }
    for (int sub = 0; sub < prm_.nsub; sub++) {
        for (size_t i = 0; i < nP; i++)
            launch_fill_st(st, aux[i], 1.0f, 1.0f, m.plans[i].elems(), s);
        float* Ss = D.S.f() + int64_t(sub) * V_;
        m.back_st(st, reinterpret_cast<void* const*>(aux.data()), Ss, m.back_mode, s, sub, prm_.nsub, 1.0f);
        launch_sens_fix(Ss, V_, s);
    }
    float f0max = 0.0f;
    for (int64_t k = 0; k < V_; k++)
        f0max = std::max(f0max, std::fabs(f0[k]));
    s_f_ = 1.0f;
    if (sti == STORE_F16) {
        std::vector<float> S(V_);
        d2h(S.data(), D.S.f(), size_t(V_) * sizeof(float), s);
        PETL_HIP_CHECK(hipStreamSynchronize(s));
        double ssum = 0.0;
        for (float v : S)
            ssum += v;
        const double mean = ssum > 0.0 ? ysum / ssum : 1.0;
        s_f_ = pow2_scale_for(std::max(float(1024.0 * mean), f0max), 32768.0f);
    }
    // the initial image (volume-sized), through the same staging buffer once the data uploads are done
    PETL_HIP_CHECK(hipStreamSynchronize(s));
    h2d(D.stage.p, f0, size_t(V_) * sizeof(float), s);
    launch_convert_st(sti, D.f.p, D.stage.f(), V_, 1.0f / s_f_, s);
    PETL_HIP_CHECK(hipStreamSynchronize(s));
}

void MlemJob::iterate(int it, hipStream_t s)
{
    if (prm_.storage != STORE_F32 || prm_.image_storage != STORE_F32) {
        iterate_storage(s);
        return;
    }
    Model& m = m_;
    DeviceState& D = m.dev;
    const size_t nP = m.plans.size();
    const int nsub = prm_.nsub;
    std::vector<float*> y = dev_ptrs(D.y), aux = dev_ptrs(D.aux);
    std::vector<float*> resp;
    if (prm_.response)
        resp = dev_ptrs(D.resp);
    float* rp = nullptr;
    for (int sub = 0; sub < nsub; sub++) {
        if (prm_.loglik) {
            std::vector<float*> q = dev_ptrs(D.g);
            m.forward(q.data(), D.f.f(), EPI_NONE, nullptr, prm_.response ? resp.data() : nullptr, s, sub, nsub);
            for (size_t i = 0; i < nP; i++) {
                const Planogram& pl = m.plans[i];
                const int64_t row = int64_t(pl.N_v0) * pl.N_u1 * pl.N_u0;  // one v1 view row
                launch_loglik(y[i], q[i], row, sub, nsub, pl.N_v1, ll_scratch_, dll_ + it, !(i == 0 && sub == 0), s);
                rp = prm_.response ? resp[i] : nullptr;
                launch_ratio(aux[i], y[i], q[i], rp, row, sub, nsub, pl.N_v1, s);
            }
        } else {
            m.forward(aux.data(), D.f.f(), EPI_RATIO, y.data(), prm_.response ? resp.data() : nullptr, s, sub, nsub);
        }
        m.back(aux.data(), D.d.f(), m.back_mode, s, sub, nsub);
        const float* Ss = D.S.f() + int64_t(sub) * V_;
        if (prm_.beta <= 0.0f || prm_.prior == 0) {
            launch_mlem_update(D.f.f(), D.d.f(), Ss, V_, s);
        } else {
            launch_mlem_update_prior(D.f2.f(), D.f.f(), D.d.f(), Ss, m.numZ, m.numY, m.numX, prm_.prior == 2 ? 1 : 0,
                                     prm_.delta, prm_.beta, prm_.p, s);
            launch_copy(D.f.f(), D.f2.f(), V_, s);
        }
    }
}

void MlemJob::iterate_storage(hipStream_t s)
{
    Model& m = m_;
    DeviceState& D = m.dev;
    const int st = prm_.storage, sti = prm_.image_storage;
    const int nsub = prm_.nsub;
    std::vector<float*> y = dev_ptrs(D.y), aux = dev_ptrs(D.aux);
    void* const* yv = reinterpret_cast<void* const*>(y.data());
    void* const* av = reinterpret_cast<void* const*>(aux.data());
    for (int sub = 0; sub < nsub; sub++) {
        // ratio y / Af stored with scale 1 (it is near 1 where the model fits the data)
        m.forward_st(sti, st, av, D.f.p, EPI_RATIO, yv, nullptr, s, sub, nsub, s_f_, s_y_, 1.0f);
        m.back_st(st, av, D.d.f(), m.back_mode, s, sub, nsub, 1.0f);
        const float* Ss = D.S.f() + int64_t(sub) * V_;
        if (prm_.beta <= 0.0f || prm_.prior == 0) {
            launch_mlem_update_st(sti, D.f.p, D.d.f(), Ss, V_, s_f_, s);
        } else {
            launch_mlem_update_prior_st(sti, D.f2.p, D.f.p, D.d.f(), Ss, m.numZ, m.numY, m.numX,
                                        prm_.prior == 2 ? 1 : 0, prm_.delta, prm_.beta, prm_.p, s_f_, s);
            launch_copy_st(sti, D.f.p, D.f2.p, V_, s);
        }
    }
}

void MlemJob::fetch(float* f, double* loglik, hipStream_t s)
{
    if (prm_.image_storage != STORE_F32) {
        std::vector<char> tmp(size_t(V_) * storage_bytes(prm_.image_storage));
        d2h(tmp.data(), m_.dev.f.p, tmp.size(), s);
        PETL_HIP_CHECK(hipStreamSynchronize(s));
        from_storage(prm_.image_storage, tmp.data(), f, V_, s_f_);
        return;
    }
    d2h(f, m_.dev.f.p, V_ * sizeof(float), s);
    if (prm_.loglik && loglik && prm_.numIter > 0)
        d2h(loglik, dll_, sizeof(double) * prm_.numIter, s);
}

}  // namespace petl

namespace petl {

bool MlemJob::batch_eligible()
{
    if (prm_.storage != STORE_F32 || prm_.image_storage != STORE_F32 || prm_.nsub != 1 || prm_.mask ||
        prm_.response || prm_.loglik || m_.back_mode != BACK_PETL)
        return false;
    m_.prepare();
    std::vector<const FastPlan*> plans;
    for (size_t i = 0; i < m_.plans.size(); i++)
        plans.push_back(m_.dev.fast[i].get());
    return MlemBatch::eligible(plans);
}

static bool pingpong(const MlemParams& prm) { return prm.prior != 0 && prm.beta > 0.0f; }

BatchIterJob MlemJob::batch_job(int it)
{
    DeviceState& D = m_.dev;
    BatchIterJob j;
    j.nplans = int(m_.plans.size());
    for (int p = 0; p < j.nplans; p++) {
        j.plan[p] = D.fast[p].get();
        j.ratio[p] = D.aux[p].f();
        j.data[p] = D.y[p].f();
    }
    const bool pp = pingpong(prm_);
    j.f_in = (pp && it % 2 == 1) ? D.f2.f() : D.f.f();
    j.f_out = pp ? ((it % 2 == 0) ? D.f2.f() : D.f.f()) : D.f.f();
    j.d = D.d.f();
    j.S = D.S.f();
    j.h2 = D.h2.f();
    j.prior = prm_.prior;
    j.delta = prm_.delta;
    j.beta = prm_.beta;
    j.p = prm_.p;
    j.Nx = m_.numX;
    j.Ny = m_.numY;
    j.Nz = m_.numZ;
    return j;
}

const float* MlemJob::batch_image(int n)
{
    return (pingpong(prm_) && n % 2 == 1) ? m_.dev.f2.f() : m_.dev.f.f();
}

}  // namespace petl
