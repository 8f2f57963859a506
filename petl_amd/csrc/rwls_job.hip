// Device-resident RWLS: src/petl_server.py:1691-1877 (RWLS, RWLSstepSize) with the helpers it calls --
// gaussian_filter (:2129-2155), clip and reciprocal (:1275-1281, 1488-1532), innerProd (:1143-1199),
// TVgradient / TVquadForm / TVcost (:1879-2026).
//
// Arrays stay on the GPU for the whole loop.  The scalars are formed as the reference forms them: every
// inner product of one array is rounded to float (inner_product3D/4D return float) and the per-planogram
// values are summed in double (the Python accumulator); gamma and the step are double quotients of those,
// rounded to float where they scale an fp32 array (numpy's float32 arithmetic with a Python scalar).
#include <cmath>
#include "rwls_job.h"
#include "lsq_kernels.h"
#include "mlem_job.h"
#include "proj_kernels.h"
#include "vec_kernels.h"

namespace petl {

namespace {
constexpr int NSLOTS = 64;
}

RwlsJob::RwlsJob(Model& m, const RwlsParams& prm) : m_(m), prm_(prm)
{
    if (prm_.numIter < 0)
        throw hip_error("RWLS: numIter must be >= 0");
    if (prm_.weights < RWLS_W_DEFAULT || prm_.weights > RWLS_W_SCALAR)
        throw hip_error("RWLS: unknown weight mode");
    nP_ = m.plans.size();
    if (nP_ + 4 > size_t(NSLOTS))
        throw hip_error("RWLS: too many planograms");
    V_ = m.vol_elems();
    G_.resize(nP_);
    for (size_t i = 0; i < nP_; i++)
        G_[i] = m.plans[i].elems();
}

void RwlsJob::project(const float* x, std::vector<float*>& out, hipStream_t s)
{
    m_.forward(out.data(), x, EPI_NONE, nullptr, prm_.response ? respp_.data() : nullptr, s);
}

void RwlsJob::project_residual(const float* x, bool weigh, hipStream_t s)
{
    m_.forward(rp_.data(), x, EPI_RESID, gp_.data(), prm_.response ? respp_.data() : nullptr, s, 0, 1,
               weigh ? Wp_.data() : nullptr);
    r_weighted_ = weigh;
}

// out = sum_i P_i^T (R_i y_i); the response product goes to Pd (or into y itself when y is scratch), as the
// reference multiplies a copy of the data (src/petl_server.py:743-746)
void RwlsJob::backproject(std::vector<float*>& y, bool y_is_scratch, float* out, hipStream_t s)
{
    if (!prm_.response) {
        m_.back(y.data(), out, m_.back_mode, s);
        return;
    }
    std::vector<float*>& t = y_is_scratch ? y : Pdp_;
    for (size_t i = 0; i < nP_; i++)
        launch_mul_to(t[i], y[i], respp_[i], G_[i], s);
    m_.back(t.data(), out, m_.back_mode, s);
}

void RwlsJob::setup(float* const* g, const float* f0, bool f0_nonzero, float* const* W, float* const* response,
                    hipStream_t s)
{
    Model& m = m_;
    m.prepare();
    const bool weighted = prm_.weights != RWLS_W_SCALAR;
    ensure_plan_bufs(m, g_);
    ensure_plan_bufs(m, r_);
    ensure_plan_bufs(m, Pd_);
    gp_ = dev_ptrs(g_);
    rp_ = dev_ptrs(r_);
    Pdp_ = dev_ptrs(Pd_);
    if (weighted) {
        ensure_plan_bufs(m, W_);
        Wp_ = dev_ptrs(W_);
    }
    if (prm_.response) {
        ensure_plan_bufs(m, resp_);
        respp_ = dev_ptrs(resp_);
    }
    const size_t vb = size_t(V_) * sizeof(float);
    f_.ensure(vb);
    grad_.ensure(vb);
    u_.ensure(vb);
    d_.ensure(vb);
    grad_old_.ensure(vb);
    if (prm_.sqs)
        Q_.ensure(vb);
    int64_t nmax = V_;
    for (int64_t e : G_)
        nmax = std::max(nmax, e);
    // with the non-negativity constraint Pd only feeds <Pd, Pd>_W, which the fast forward can sum without
    // writing Pd (the unconstrained update reads Pd itself)
    scratch_cap_ = red_scratch_doubles(nmax);
    fuse_sumsq_ = prm_.nonneg;
    for (size_t i = 0; i < nP_; i++) {
        const int64_t np = m.sumsq_parts(int(i));
        fuse_sumsq_ = fuse_sumsq_ && np > 0;
        scratch_cap_ = std::max(scratch_cap_, np);
    }
    red_.ensure(sizeof(double) * size_t(NSLOTS + scratch_cap_));
    slots_ = static_cast<double*>(red_.p);
    scratch_ = slots_ + NSLOTS;

    for (size_t i = 0; i < nP_; i++) {
        h2d(gp_[i], g[i], size_t(G_[i]) * sizeof(float), s);
        if (prm_.weights == RWLS_W_ARRAYS)
            h2d(Wp_[i], W[i], size_t(G_[i]) * sizeof(float), s);
        if (prm_.response)
            h2d(respp_[i], response[i], size_t(G_[i]) * sizeof(float), s);
    }
    h2d(f_.p, f0, vb, s);

    // default weights: W = copy(g); gaussian_filter(W, 3.0); clip(W, 7.0); reciprocal(W), each planogram
    // smoothed over (u1, u0) within its (v1, v0) rows (src/petl_server.py:2138-2142)
    if (prm_.weights == RWLS_W_DEFAULT)
        for (size_t i = 0; i < nP_; i++) {
            const Planogram& pl = m.plans[i];
            launch_copy_vec(Wp_[i], gp_[i], G_[i], s);
            gaussian_filter_device(Wp_[i], rp_[i], pl.N_v1 * pl.N_v0, pl.N_u1, pl.N_u0, 3.0f, 2, true, s);
        }

    // Pf - g, with Pf = 0 for an all-zero start
    if (f0_nonzero) {
        project_residual(f_.f(), false, s);
    } else {
        for (size_t i = 0; i < nP_; i++) {
            launch_fill_vec(rp_[i], 0.0f, G_[i], s);
            launch_sub(rp_[i], gp_[i], G_[i], s);
        }
        r_weighted_ = false;
    }

    // SQS: Q = 1 / P*(W P 1), 0 -> 1
    if (prm_.sqs) {
        launch_fill_vec(Q_.f(), 1.0f, V_, s);
        project(Q_.f(), Pdp_, s);
        if (weighted)
            for (size_t i = 0; i < nP_; i++)
                launch_mul_inplace(Pdp_[i], Wp_[i], G_[i], s);
        backproject(Pdp_, true, Q_.f(), s);
        launch_reciprocal(Q_.f(), 1.0f, V_, s);
    }
    launch_fill_vec(d_.f(), 0.0f, V_, s);
    launch_fill_vec(grad_old_.f(), 0.0f, V_, s);
    gg_old_ = 0.0;
}

// inner_product3D: the sum rounded to float
float RwlsJob::dot_vol(const float* x, const float* y, int slot, hipStream_t s)
{
    launch_dot(x, y, nullptr, V_, scratch_, slots_ + slot, s);
    double h = 0.0;
    d2h(&h, slots_ + slot, sizeof(double), s);
    PETL_HIP_CHECK(hipStreamSynchronize(s));
    return float(h);
}

bool RwlsJob::iterate(int n, hipStream_t s)
{
    const bool weighted = prm_.weights != RWLS_W_SCALAR;
    // WPf_minus_g aliases Pf_minus_g: the weighting is in place (src/petl_server.py:1790-1792); the
    // residual projection of the previous step may have applied it already
    if (weighted && !r_weighted_)
        for (size_t i = 0; i < nP_; i++)
            launch_mul_inplace(rp_[i], Wp_[i], G_[i], s);
    r_weighted_ = weighted;
    backproject(rp_, false, grad_.f(), s);
    if (prm_.beta > 0.0f) {
        launch_tv_gradient(f_.f(), u_.f(), m_.numZ, m_.numY, m_.numX, prm_.delta, prm_.beta, 1.0f, s);
        launch_add(grad_.f(), u_.f(), V_, s);
    }
    if (prm_.sqs)
        launch_mul_to(u_.f(), Q_.f(), grad_.f(), V_, s);
    else
        launch_copy_vec(u_.f(), grad_.f(), V_, s);

    float u_grad;
    if (n % 50 == 0) {
        launch_copy_vec(d_.f(), u_.f(), V_, s);
        u_grad = dot_vol(u_.f(), grad_.f(), 0, s);
    } else {
        launch_dot(u_.f(), grad_.f(), nullptr, V_, scratch_, slots_ + 0, s);
        launch_dot(u_.f(), grad_old_.f(), nullptr, V_, scratch_, slots_ + 1, s);
        double h[2];
        d2h(h, slots_, sizeof(h), s);
        PETL_HIP_CHECK(hipStreamSynchronize(s));
        u_grad = float(h[0]);
        if (gg_old_ == 0.0)
            throw hip_error("RWLS: float division by zero (the previous <u, grad> is 0)");
        const double gamma = (double(float(h[0])) - double(float(h[1]))) / gg_old_;
        launch_xpay(d_.f(), u_.f(), float(gamma), V_, s);
        if (dot_vol(d_.f(), grad_.f(), 2, s) <= 0.0f)
            launch_copy_vec(d_.f(), u_.f(), V_, s);
    }
    gg_old_ = double(u_grad);
    launch_copy_vec(grad_old_.f(), grad_.f(), V_, s);

    // num = <d, grad>; denominator <Pd, Pd>_W (+ <d, R''(f) d>), all read back at once
    if (fuse_sumsq_) {
        for (size_t i = 0; i < nP_; i++) {
            const int64_t np = m_.forward_sumsq_parts(int(i), d_.f(), prm_.response ? respp_[i] : nullptr,
                                                      weighted ? Wp_[i] : nullptr, scratch_, scratch_cap_, s);
            if (np == 0)
                throw hip_error("RWLS: the fused <Pd, Pd> forward is unavailable for this planogram");
            launch_sum_parts(scratch_, np, slots_ + 2 + i, s);
        }
    } else {
        project(d_.f(), Pdp_, s);
        for (size_t i = 0; i < nP_; i++)
            launch_dot(Pdp_[i], Pdp_[i], weighted ? Wp_[i] : nullptr, G_[i], scratch_, slots_ + 2 + i, s);
    }
    launch_dot(d_.f(), grad_.f(), nullptr, V_, scratch_, slots_ + 0, s);
    if (prm_.beta > 0.0f)
        launch_tv_reduce(f_.f(), d_.f(), m_.numZ, m_.numY, m_.numX, prm_.delta, prm_.beta, 1.0f, scratch_,
                         slots_ + 1, s);
    double h[NSLOTS];
    d2h(h, slots_, sizeof(double) * (2 + nP_), s);
    PETL_HIP_CHECK(hipStreamSynchronize(s));
    const double num = double(float(h[0]));
    double denA = 0.0;
    for (size_t i = 0; i < nP_; i++)
        denA += weighted ? double(float(h[2 + i])) : double(float(h[2 + i])) * double(prm_.w_scalar);
    const double denB = prm_.beta > 0.0f ? double(float(h[1])) : 0.0;
    const double den = denA + denB;
    const double step = std::fabs(den) > 1.0e-16 ? num / den : 0.0;
    if (step <= 0.0)
        return false;

    launch_step(f_.f(), d_.f(), float(step), prm_.nonneg, V_, s);
    if (prm_.nonneg) {
        // Pf - g and the next iteration's weighting in one forward (unless the cost needs r unweighted)
        project_residual(f_.f(), weighted && !prm_.cost, s);
    } else {
        // scalarAdd(Pf, -step, Pd) then subtract(Pf_minus_g, g) on the same (already weighted) array
        // (src/petl_server.py:1831-1834)
        for (size_t i = 0; i < nP_; i++)
            launch_resid_update(rp_[i], Pdp_[i], gp_[i], float(-step), G_[i], s);
        r_weighted_ = false;
    }
    return true;
}

double RwlsJob::cost(hipStream_t s)
{
    if (r_weighted_)
        throw hip_error("RWLS: the cost needs the unweighted residual (set RwlsParams::cost)");
    const bool weighted = prm_.weights != RWLS_W_SCALAR;
    for (size_t i = 0; i < nP_; i++)
        launch_dot(rp_[i], rp_[i], weighted ? Wp_[i] : nullptr, G_[i], scratch_, slots_ + 2 + i, s);
    if (prm_.beta > 0.0f)
        launch_tv_reduce(f_.f(), nullptr, m_.numZ, m_.numY, m_.numX, prm_.delta, prm_.beta, 1.0f, scratch_,
                         slots_ + 1, s);
    double h[NSLOTS];
    d2h(h, slots_, sizeof(double) * (2 + nP_), s);
    PETL_HIP_CHECK(hipStreamSynchronize(s));
    double fid = 0.0;
    for (size_t i = 0; i < nP_; i++)
        fid += weighted ? double(float(h[2 + i])) : double(float(h[2 + i])) * double(prm_.w_scalar);
    return 0.5 * fid + (prm_.beta > 0.0f ? double(float(h[1])) : 0.0);
}

void RwlsJob::fetch(float* f, hipStream_t s) { d2h(f, f_.p, size_t(V_) * sizeof(float), s); }

}  // namespace petl
