// Device-resident MLEM / OSEM (src/petl_server.py:1534-1606, 1257-1273).
#include <algorithm>
#include "mlem_job.h"
#include "proj_kernels.h"
#include "vec_kernels.h"

namespace petl {

void ensure_plan_bufs(Model& m, std::vector<DevBuf>& bufs)
{
    bufs.resize(m.plans.size());
    for (size_t i = 0; i < m.plans.size(); i++)
        bufs[i].ensure(size_t(m.plans[i].elems()) * sizeof(float));
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
}

size_t MlemJob::device_bytes(Model& m, const MlemParams& prm)
{
    const int nsub = clamp_subsets(m, prm.nsub);
    size_t G = 0, back_ws = 0;
    for (const auto& pl : m.plans)
        G += size_t(pl.elems());
    const VolGeom V = m.vgeom();
    for (const auto& pl : m.plans) {
        // the fast back projector's workspace, the largest of the planograms (see Model::prepare)
        if (m.path == 0) {
            const PlanGeom P = pl.geom();
            const bool sep = (V.Tx == P.Tu0 && V.Ty == P.Tu0 && V.Tz == P.Tu1);
            if (sep) {
                const size_t ns = size_t(std::max(V.Nx, V.Ny));
                back_ws = std::max(back_ws, size_t(P.Nv1) * P.Nu1 * ns * ns * sizeof(float));
            }
        }
    }
    const size_t vol = size_t(m.vol_elems()) * sizeof(float);
    size_t plan_bufs = 2 + (prm.response ? 1 : 0) + (prm.loglik ? 1 : 0);  // y, aux (+resp) (+q)
    return G * sizeof(float) * plan_bufs + vol * (4 + size_t(nsub)) + back_ws + (size_t(1) << 20);
}

void MlemJob::allocate()
{
    Model& m = m_;
    DeviceState& D = m.dev;
    m.prepare();
    ensure_plan_bufs(m, D.y);
    ensure_plan_bufs(m, D.aux);
    if (prm_.response)
        ensure_plan_bufs(m, D.resp);
    D.f.ensure(V_ * sizeof(float));
    D.f2.ensure(V_ * sizeof(float));
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

void MlemJob::iterate(int it, hipStream_t s)
{
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

void MlemJob::fetch(float* f, double* loglik, hipStream_t s)
{
    d2h(f, m_.dev.f.p, V_ * sizeof(float), s);
    if (prm_.loglik && loglik && prm_.numIter > 0)
        d2h(loglik, dll_, sizeof(double) * prm_.numIter, s);
}

}  // namespace petl
