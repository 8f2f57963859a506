// extern "C" boundary of libpetl.so (declared in include/petl_c_interface.h).
//
// The reference keeps a process-global list of models and a global "current model" index that every
// Python call sets first (src/petl_c_interface.cpp:12-30, src/list_of_models.cpp:25-46); the same
// registry semantics are kept here.  Host-pointer entry points copy inputs to the device, run the HIP
// kernels on the model's stream and copy results back before returning, as the reference's do.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <vector>
#include "petl_common.h"
#include "petl_model.h"
#include "mlem_job.h"
#include "rwls_job.h"
#include "lsq_kernels.h"
#include "pfdr.h"
#include "binning.h"
#include "proj_kernels.h"
#include "vec_kernels.h"
#include "../../include/petl_c_interface.h"

#define PETL_VERSION "1.0-mi355x"

namespace petl {

namespace {

std::vector<std::unique_ptr<Model>> g_models;
int g_which = 0;

int append_model()
{
    g_models.push_back(std::make_unique<Model>());
    return int(g_models.size()) - 1;
}

}  // namespace

// src/list_of_models.cpp:37-46: list[i % size], creating one model on first use
Model* current_model()
{
    if (g_models.empty())
        append_model();
    const int n = int(g_models.size());
    int i = g_which % n;
    if (i < 0)
        i += n;
    return g_models[i].get();
}

namespace {

Model* M() { return current_model(); }

template <typename Fn>
bool guarded(const char* what, Fn&& fn)
{
    try {
        return fn();
    } catch (const std::exception& e) {
        fprintf(stderr, "PETL %s failed: %s\n", what, e.what());
    } catch (...) {
        fprintf(stderr, "PETL %s failed\n", what);
    }
    return false;
}

bool unsupported(const char* name)
{
    fprintf(stderr,
            "PETL: %s is outside the projector/MLEM path this library implements; it is not available.\n",
            name);
    return false;
}

// ---- host elementwise operations: src/data_cube.cpp:97-167, 320-397 ----
enum Op { OP_MULT, OP_DIV, OP_RDIV };

void binary_op(float* x, const float* y, int64_t n, Op op)
{
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++) {
        switch (op) {
        case OP_MULT:
            x[i] *= y[i];
            break;
        case OP_DIV:
            if (y[i] == 0.0f)
                x[i] = 0.0f;
            else
                x[i] /= y[i];
            break;
        case OP_RDIV:
            if (x[i] == 0.0f)
                x[i] = 0.0f;
            else
                x[i] = y[i] / x[i];
            break;
        }
    }
}

bool dims_ok(std::initializer_list<int> d)
{
    for (int v : d)
        if (v <= 0)
            return false;
    return true;
}

// double accumulation per N_1 slab then summed in order (src/data_cube.cpp:320-397)
float inner3(const float* x, const float* y, const float* w, int N1, int N2, int N3)
{
    std::vector<double> sums(N1);
    const int64_t slab = int64_t(N2) * N3;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < N1; i++) {
        double s = 0.0;
        const int64_t o = int64_t(i) * slab;
        for (int64_t k = 0; k < slab; k++)
            s += w ? double(x[o + k] * y[o + k] * w[o + k]) : double(x[o + k] * y[o + k]);
        sums[i] = s;
    }
    double acc = 0.0;
    for (int i = 0; i < N1; i++)
        acc += sums[i];
    return float(acc);
}

// ---- device-resident MLEM / OSEM (src/petl_server.py:1534-1606, 1257-1273) ----
bool run_mlem(Model& m, float** g, float* f, int numIter, float delta, float beta, int prior, float p, int nsub,
              float** mask, float** response, double* loglik, int storage = STORE_F32,
              int image_storage = STORE_F32)
{
    if (!g || !f || !m.allDefined() || numIter < 0)
        return false;
    for (size_t i = 0; i < m.plans.size(); i++)
        if (!g[i] || (mask && !mask[i]) || (response && !response[i]))
            return false;
    MlemParams prm;
    prm.numIter = numIter;
    prm.delta = delta;
    prm.beta = beta;
    prm.prior = prior;
    prm.p = p;
    prm.nsub = nsub;
    prm.loglik = loglik != nullptr;
    prm.mask = mask != nullptr;
    prm.response = response != nullptr;
    prm.storage = storage;
    prm.image_storage = image_storage;
    // initial estimate (src/petl_server.py:1571-1577), in place on the caller's array
    mlem_initial_estimate(f, m.vol_elems());
    hipStream_t s = m.stream();
    MlemJob job(m, prm);
    job.allocate();
    job.setup(g, f, mask, response, s);
    for (int it = 0; it < numIter; it++)
        job.iterate(it, s);
    job.fetch(f, loglik, s);
    PETL_HIP_CHECK(hipStreamSynchronize(s));
    return true;
}

// ---- list-mode binning (src/petl.cpp:65-79, src/binning.cpp) ----
// Adds n coincidences staged on the device to the model's device planograms; counts accepted events.
void bin_staged(Model& m, const float* d_blue, const float* d_pink, int64_t n, float* const* d_g, bool nearest,
                bool chest_wall, unsigned long long* d_count, hipStream_t s)
{
    for (size_t i = 0; i < m.plans.size(); i++)
        bin_events_device(bin_geom(m.plans[i], nearest, chest_wall), d_blue, d_pink, n, d_g[i], m.dev.bin,
                          d_count + i, s);
}

constexpr int64_t BIN_CHUNK = int64_t(1) << 22;  // coincidences staged per upload

// Source of coincidences for run_bin: a list-mode file or host arrays.
struct EventSource {
    ListModeReader* file = nullptr;
    const float* blue = nullptr;
    const float* pink = nullptr;
    int64_t n = 0, pos = 0;
    int64_t next(float* b, float* p, int64_t max)
    {
        if (file)
            return file->next(b, p, max);
        const int64_t k = std::min(max, n - pos);
        std::memcpy(b, blue + 3 * pos, size_t(k) * 12);
        std::memcpy(p, pink + 3 * pos, size_t(k) * 12);
        pos += k;
        return k;
    }
};

// Bins every coincidence of `src` into the host planograms g (zeroed first when `reset`, as PETL::bin
// does through planogram::set_constant(0) in src/binning.cpp:58).  accepted: optional host counts.
bool run_bin(Model& m, float** g, EventSource& src, bool reset, bool nearest, bool chest_wall, int64_t* accepted,
             int64_t* total_events)
{
    hipStream_t s = m.stream();
    DeviceState& D = m.dev;
    const size_t nP = m.plans.size();
    ensure_plan_bufs(m, D.g);
    std::vector<float*> dg = dev_ptrs(D.g);
    for (size_t i = 0; i < nP; i++) {
        const size_t nb = size_t(m.plans[i].elems()) * sizeof(float);
        if (reset)
            launch_fill(dg[i], 0.0f, m.plans[i].elems(), s);
        else
            h2d(dg[i], g[i], nb, s);
    }
    D.events.ensure(size_t(BIN_CHUNK) * 24);
    D.counter.ensure(nP * sizeof(unsigned long long));
    unsigned long long* d_count = static_cast<unsigned long long*>(D.counter.p);
    PETL_HIP_CHECK(hipMemsetAsync(d_count, 0, nP * sizeof(unsigned long long), s));
    float* hbuf = nullptr;
    PETL_HIP_CHECK(hipHostMalloc(reinterpret_cast<void**>(&hbuf), size_t(BIN_CHUNK) * 24));
    int64_t total = 0;
    try {
        float* d_blue = D.events.f();
        float* d_pink = d_blue + 3 * BIN_CHUNK;
        for (;;) {
            const int64_t n = src.next(hbuf, hbuf + 3 * BIN_CHUNK, BIN_CHUNK);
            if (n <= 0)
                break;
            h2d(d_blue, hbuf, size_t(n) * 12, s);
            h2d(d_pink, hbuf + 3 * BIN_CHUNK, size_t(n) * 12, s);
            bin_staged(m, d_blue, d_pink, n, dg.data(), nearest, chest_wall, d_count, s);
            PETL_HIP_CHECK(hipStreamSynchronize(s));  // the staging buffer is refilled next
            total += n;
        }
        std::vector<unsigned long long> cnt(nP);
        d2h(cnt.data(), d_count, nP * sizeof(unsigned long long), s);
        for (size_t i = 0; i < nP; i++)
            d2h(g[i], dg[i], size_t(m.plans[i].elems()) * sizeof(float), s);
        PETL_HIP_CHECK(hipStreamSynchronize(s));
        if (accepted)
            for (size_t i = 0; i < nP; i++)
                accepted[i] = int64_t(cnt[i]);
    } catch (...) {
        (void)hipHostFree(hbuf);
        throw;
    }
    (void)hipHostFree(hbuf);
    if (total_events)
        *total_events = total;
    return true;
}

// src/total_variation.cu:662-800: -1 for a null input; the sum of the per-voxel terms, as float
float tv_reduce_export(const char* what, bool quad, float* f, float* d, int N_1, int N_2, int N_3, float delta,
                       float beta, float p, bool data_on_cpu)
{
    if (!f || (quad && !d))
        return -1.0f;
    if (!dims_ok({N_1, N_2, N_3}))
        return -1.0f;
    float out = -1.0f;
    guarded(what, [&] {
        Model* m = M();
        hipStream_t s = m->stream();
        const int64_t n = int64_t(N_1) * N_2 * N_3;
        DevBuf a, b, red;
        const float* df = f;
        const float* dd = d;
        if (data_on_cpu) {
            a.ensure(n * sizeof(float));
            h2d(a.p, f, n * sizeof(float), s);
            df = a.f();
            if (d) {
                b.ensure(n * sizeof(float));
                h2d(b.p, d, n * sizeof(float), s);
                dd = b.f();
            }
        }
        red.ensure(sizeof(double) * size_t(1 + red_scratch_doubles(n)));
        double* slot = static_cast<double*>(red.p);
        launch_tv_reduce(df, dd, N_1, N_2, N_3, delta, beta, p, slot + 1, slot, s);
        double h = 0.0;
        d2h(&h, slot, sizeof(double), s);
        PETL_HIP_CHECK(hipStreamSynchronize(s));
        out = float(h);
        return true;
    });
    return out;
}

// RWLS on the GPU (src/petl_server.py:1691-1877); f is the caller's host image, updated in place.
bool run_rwls(Model& m, float** g, float* f, int numIter, float delta, float beta, int weight_mode, float** W,
              float w_scalar, int preconditioner, bool nonneg, float** response, double* cost, int* iterations)
{
    if (!g || !f || !m.allDefined() || numIter < 0)
        return false;
    if (weight_mode < RWLS_W_DEFAULT || weight_mode > RWLS_W_SCALAR || (weight_mode == RWLS_W_ARRAYS && !W))
        return false;
    for (size_t i = 0; i < m.plans.size(); i++)
        if (!g[i] || (weight_mode == RWLS_W_ARRAYS && !W[i]) || (response && !response[i]))
            return false;
    RwlsParams prm;
    prm.numIter = numIter;
    prm.delta = delta;
    prm.beta = beta;
    prm.weights = weight_mode;
    prm.w_scalar = w_scalar;
    prm.sqs = preconditioner == 1;
    prm.nonneg = nonneg;
    prm.response = response != nullptr;
    prm.cost = cost != nullptr;
    // start rules (src/petl_server.py:1743-1759): a nonzero image has its negatives clipped (in place) when
    // constrained and is projected; an all-zero image starts from Pf = 0
    const int64_t V = m.vol_elems();
    bool nonzero = false;
    for (int64_t i = 0; i < V && !nonzero; i++)
        nonzero = f[i] != 0.0f;
    if (nonzero && nonneg)
        for (int64_t i = 0; i < V; i++)
            if (f[i] < 0.0f)
                f[i] = 0.0f;
    hipStream_t s = m.stream();
    RwlsJob job(m, prm);
    job.setup(g, f, nonzero, W, response, s);
    int done = 0;
    for (int n = 0; n < numIter; n++) {
        if (!job.iterate(n, s))
            break;
        done = n + 1;
        if (cost)
            cost[n] = job.cost(s);
    }
    job.fetch(f, s);
    PETL_HIP_CHECK(hipStreamSynchronize(s));
    if (iterations)
        *iterations = done;
    return true;
}

// ray_trace / stopping_power on host planograms (src/petl.cpp:81-152): each planogram traced on the
// device; stopping_power with no modules sets the planograms to 1
// uploads the phantom's or the modules' solids to the model's device copy; returns their count
int upload_solids(Model& m, bool stopping, hipStream_t s)
{
    const std::vector<Solid>& solids = stopping ? m.modules : m.phantom;
    m.dev.solids.ensure(std::max<size_t>(1, solids.size()) * sizeof(Solid));
    if (!solids.empty()) {
        h2d(m.dev.solids.p, solids.data(), solids.size() * sizeof(Solid), s);
        PETL_HIP_CHECK(hipStreamSynchronize(s));  // the host vector may change after this call
    }
    return int(solids.size());
}

bool run_ray_trace(Model& m, float** g, int oversampling, bool stopping)
{
    for (size_t i = 0; i < m.plans.size(); i++)
        if (g[i] == nullptr)
            return false;
    if (stopping && m.modules.empty()) {
        for (size_t i = 0; i < m.plans.size(); i++)
            std::fill(g[i], g[i] + m.plans[i].elems(), 1.0f);
        return true;
    }
    hipStream_t s = m.stream();
    const int n = upload_solids(m, stopping, s);
    DevBuf d_g;
    for (size_t i = 0; i < m.plans.size(); i++) {
        const size_t bytes = size_t(m.plans[i].elems()) * sizeof(float);
        d_g.ensure(bytes);
        ray_trace_device(m.plans[i].geom(), static_cast<const Solid*>(m.dev.solids.p), n, oversampling, stopping,
                         d_g.f(), s);
        d2h(g[i], d_g.p, bytes, s);
        PETL_HIP_CHECK(hipStreamSynchronize(s));
    }
    return true;
}

// PETL::doPFDR (src/petl.cpp:357-385): every planogram rebinned to one v1 row (g_reb: [N_v0][N_u1][N_u0]);
// afterwards the model's planograms have N_v1 = 1 and v1_0 = 0, as the reference leaves them
bool run_pfdr(Model& m, float** g, float** g_reb)
{
    for (size_t i = 0; i < m.plans.size(); i++)
        if (g[i] == nullptr || g_reb[i] == nullptr)
            return false;
    hipStream_t s = m.stream();
    DevBuf dg, dr;
    for (size_t i = 0; i < m.plans.size(); i++) {
        const Planogram& p = m.plans[i];
        const size_t reb = size_t(p.N_v0) * p.N_u1 * p.N_u0;
        dg.ensure(size_t(p.elems()) * sizeof(float));
        dr.ensure(reb * sizeof(float));
        h2d(dg.p, g[i], size_t(p.elems()) * sizeof(float), s);
        pfdr_device(p.geom(), p.v_m0, dg.f(), dr.f(), s);
        d2h(g_reb[i], dr.p, reb * sizeof(float), s);
        PETL_HIP_CHECK(hipStreamSynchronize(s));
    }
    for (Planogram& p : m.plans) {
        p.N_v1 = 1;
        p.v1_0 = 0.0f;
    }
    m.touch();
    return true;
}

// PETL::doFBP (src/petl.cpp:310-355): PFDR when N_v1 > 1 (on a copy of the geometry), then the ramp filter
// and the inverse planogram weight of the rebinned views, and the back projection of the rebinned
// planograms into f.  Data that are already rebinned (N_v1 = 1) are filtered and weighted in place, as the
// reference does to the caller's arrays.
bool run_fbp(Model& m, float** g, float* f)
{
    for (size_t i = 0; i < m.plans.size(); i++)
        if (g[i] == nullptr)
            return false;
    Model reb;
    reb.copy_params_from(m);
    reb.back_mode = m.back_mode;
    reb.path = m.path;
    for (Planogram& p : reb.plans) {
        p.N_v1 = 1;
        p.v1_0 = 0.0f;
    }
    hipStream_t s = reb.stream();
    std::vector<DevBuf> dreb(m.plans.size());
    DevBuf dg;
    for (size_t i = 0; i < m.plans.size(); i++) {
        const Planogram& p = m.plans[i];
        const size_t n_reb = size_t(p.N_v0) * p.N_u1 * p.N_u0;
        dreb[i].ensure(n_reb * sizeof(float));
        if (p.N_v1 > 1) {
            dg.ensure(size_t(p.elems()) * sizeof(float));
            h2d(dg.p, g[i], size_t(p.elems()) * sizeof(float), s);
            pfdr_device(p.geom(), p.v_m0, dg.f(), dreb[i].f(), s);
        } else {
            h2d(dreb[i].p, g[i], n_reb * sizeof(float), s);
        }
        fbp_filter_device(dreb[i].f(), p.N_v0, p.N_u1, p.N_u0, p.T_u0, p.T_v0, p.v0_0, s);
        if (p.N_v1 == 1)
            d2h(g[i], dreb[i].p, n_reb * sizeof(float), s);
    }
    dg.release();
    DevBuf df;
    df.ensure(size_t(reb.vol_elems()) * sizeof(float));
    std::vector<float*> ptrs = dev_ptrs(dreb);
    reb.back(ptrs.data(), df.f(), reb.back_mode, s);
    d2h(f, df.p, size_t(reb.vol_elems()) * sizeof(float), s);
    PETL_HIP_CHECK(hipStreamSynchronize(s));
    return true;
}

}  // namespace
}  // namespace petl

using namespace petl;

// =================================================================================================
// Part 1: the reference ABI (src/petl_c_interface.cpp)
// =================================================================================================
extern "C" {

bool set_model(int i)
{
    g_which = i;
    return true;
}

int create_new_model()
{
    g_which = append_model();
    return g_which;
}

void about()
{
    printf("***********************************************************\n");
    printf("       Planogram Emission Tomography Library (PETL)\n");
    printf("           MI355X (gfx950) HIP implementation %s\n", PETL_VERSION);
    printf("***********************************************************\n");
}

bool copy_parameters(int param_id)
{
    if (0 <= param_id && param_id < int(g_models.size())) {
        Model* cur = M();
        Model* src = g_models[param_id].get();
        if (cur != src)
            cur->copy_params_from(*src);
        return true;
    }
    return false;
}

void clearAll()
{
    Model* m = M();
    m->plans.clear();
    m->numX = m->numY = m->numZ = 0;
    m->voxelWidth = m->voxelHeight = 0.0f;
    m->offsetX = m->offsetY = m->offsetZ = 0.0f;
    m->touch();
}

void print_parameters()
{
    Model* m = M();
    printf("\n======== PET Volume ========\n");
    printf("number of voxels (x, y, z): %d x %d x %d\n", m->numX, m->numY, m->numZ);
    printf("voxel size: %f mm x %f mm x %f mm\n", m->voxelWidth, m->voxelWidth, m->voxelHeight);
    if (m->offsetX != 0.0f || m->offsetY != 0.0f || m->offsetZ != 0.0f)
        printf("volume offset: %f mm, %f mm, %f mm\n", m->offsetX, m->offsetY, m->offsetZ);
    printf("\n======== PET Geometry ========\n");
    for (size_t i = 0; i < m->plans.size(); i++) {
        const Planogram& p = m->plans[i];
        printf("Planogram %d\n", int(i));
        printf("  angle = %f deg\n", p.psi * 180.0 / 3.1415926535897932385);
        printf("  number of samples: %d x %d x %d x %d\n", p.N_v1, p.N_v0, p.N_u1, p.N_u0);
        printf("  detector size: %f mm x %f mm\n", p.T_u1, p.T_u0);
        printf("  full detector length: %f mm\n", 2.0 * p.L);
        printf("  full detector height: %f mm\n", 2.0 * p.H);
        printf("  distance between panels: %f mm\n", 2.0 * p.R);
    }
    printf("\n");
}

bool volume_defined() { return M()->volumeDefined(); }
bool geometry_defined() { return M()->geometryDefined(); }

int get_numX() { return M()->numX; }
int get_numY() { return M()->numY; }
int get_numZ() { return M()->numZ; }
float get_voxelWidth() { return M()->voxelWidth; }
float get_voxelHeight() { return M()->voxelHeight; }
float get_offsetX() { return M()->offsetX; }
float get_offsetY() { return M()->offsetY; }
float get_offsetZ() { return M()->offsetZ; }

#define PETL_SETTER(name, field, type, cond) \
    bool name(type v)                       \
    {                                       \
        if (cond) {                         \
            M()->field = v;                 \
            M()->touch();                   \
            return true;                    \
        }                                   \
        return false;                       \
    }
PETL_SETTER(set_numX, numX, int, v > 0)
PETL_SETTER(set_numY, numY, int, v > 0)
PETL_SETTER(set_numZ, numZ, int, v > 0)
PETL_SETTER(set_voxelWidth, voxelWidth, float, v > 0.0f)
PETL_SETTER(set_voxelHeight, voxelHeight, float, v > 0.0f)
// src/petl_c_interface.cpp:154-195 accepts only positive offsets
PETL_SETTER(set_offsetX, offsetX, float, v > 0.0f)
PETL_SETTER(set_offsetY, offsetY, float, v > 0.0f)
PETL_SETTER(set_offsetZ, offsetZ, float, v > 0.0f)
#undef PETL_SETTER

int get_numPlanograms() { return int(M()->plans.size()); }

bool get_planogramSize(int which, int* shape)
{
    Model* m = M();
    if (!m->geometryDefined() || which < 0 || which >= int(m->plans.size()) || shape == nullptr)
        return false;
    const Planogram& p = m->plans[which];
    shape[0] = p.N_v1;
    shape[1] = p.N_v0;
    shape[2] = p.N_u1;
    shape[3] = p.N_u0;
    return true;
}

// src/petl_c_interface.cpp:216-219: degrees -> radians and full -> half sizes, in double
bool add_planogram(float psi, float R, float L, float H, float v_m0, float v_m1, float T)
{
    Planogram p;
    if (!p.init_panels(psi, R, L, H, v_m0, v_m1, T))
        return false;
    M()->plans.push_back(p);
    M()->touch();
    return true;
}

bool remove_planogram(int which)
{
    Model* m = M();
    if (which >= 0 && which < int(m->plans.size())) {
        m->plans.erase(m->plans.begin() + which);
        m->touch();
        return true;
    }
    return false;
}

bool keep_only_planogram(int which)
{
    Model* m = M();
    if (which >= 0 && which < int(m->plans.size())) {
        Planogram keep = m->plans[which];
        m->plans.clear();
        m->plans.push_back(keep);
        m->touch();
        return true;
    }
    return false;
}

// src/parameters.cpp:205-248
bool set_default_volume(float scale)
{
    Model* m = M();
    if (!m->geometryDefined())
        return false;
    m->offsetX = m->offsetY = m->offsetZ = 0.0f;
    float X = 0.0f, Y = 0.0f, Z = 0.0f;
    for (size_t i = 0; i < m->plans.size(); i++) {
        const Planogram& p = m->plans[i];
        const float W = p.L, Hh = p.R, psi = p.psi;
        const float X_new = std::fabs(W * std::cos(psi)) + std::fabs(Hh * std::sin(psi));
        const float Y_new = std::fabs(W * std::sin(psi)) + std::fabs(Hh * std::cos(psi));
        if (i == 0) {
            m->voxelWidth = p.T_u0;
            m->voxelHeight = p.T_u1;
            Z = p.H;
            X = X_new;
            Y = Y_new;
        } else {
            m->voxelWidth = std::min(m->voxelWidth, p.T_u0);
            m->voxelHeight = std::min(m->voxelHeight, p.T_u1);
            Z = std::min(Z, p.H);
            X = std::min(X, X_new);
            Y = std::min(Y, Y_new);
        }
    }
    m->voxelWidth *= scale;
    m->voxelHeight *= scale;
    m->numZ = 2 * int(Z / m->voxelHeight);
    m->numY = 2 * int(Y / m->voxelWidth);
    m->numX = 2 * int(X / m->voxelWidth);
    m->touch();
    return true;
}

// src/parameters.cpp:250-269
bool set_volume(int numX, int numY, int numZ, float voxelWidth, float voxelHeight, float offsetX, float offsetY,
                float offsetZ)
{
    if (numX > 0 && numY > 0 && numZ > 0 && voxelWidth > 0.0f && voxelHeight > 0.0f) {
        Model* m = M();
        m->numX = numX;
        m->numY = numY;
        m->numZ = numZ;
        m->voxelWidth = voxelWidth;
        m->voxelHeight = voxelHeight;
        m->offsetX = offsetX;
        m->offsetY = offsetY;
        m->offsetZ = offsetZ;
        m->touch();
        return true;
    }
    return false;
}

bool project(float** g, float* f)
{
    Model* m = M();
    if (g == nullptr || f == nullptr || !m->allDefined())
        return false;
    for (size_t i = 0; i < m->plans.size(); i++)
        if (!g[i])
            return false;
    return guarded("project", [&] {
        hipStream_t s = m->stream();
        m->prepare();
        const int64_t V = m->vol_elems();
        m->dev.f.ensure(V * sizeof(float));
        ensure_plan_bufs(*m, m->dev.g);
        std::vector<float*> dg = dev_ptrs(m->dev.g);
        h2d(m->dev.f.p, f, V * sizeof(float), s);
        m->forward(dg.data(), m->dev.f.f(), EPI_NONE, nullptr, nullptr, s);
        for (size_t i = 0; i < m->plans.size(); i++)
            d2h(g[i], dg[i], size_t(m->plans[i].elems()) * sizeof(float), s);
        PETL_HIP_CHECK(hipStreamSynchronize(s));
        return true;
    });
}

bool backproject(float** g, float* f)
{
    Model* m = M();
    if (g == nullptr || f == nullptr || !m->allDefined())
        return false;
    for (size_t i = 0; i < m->plans.size(); i++)
        if (!g[i])
            return false;
    return guarded("backproject", [&] {
        hipStream_t s = m->stream();
        m->prepare();
        const int64_t V = m->vol_elems();
        m->dev.f.ensure(V * sizeof(float));
        ensure_plan_bufs(*m, m->dev.g);
        std::vector<float*> dg = dev_ptrs(m->dev.g);
        for (size_t i = 0; i < m->plans.size(); i++)
            h2d(dg[i], g[i], size_t(m->plans[i].elems()) * sizeof(float), s);
        m->back(dg.data(), m->dev.f.f(), m->back_mode, s);
        d2h(f, m->dev.f.p, V * sizeof(float), s);
        PETL_HIP_CHECK(hipStreamSynchronize(s));
        return true;
    });
}

bool multiply3D(float* x, float* y, int N_1, int N_2, int N_3)
{
    if (!x || !y || !dims_ok({N_1, N_2, N_3}))
        return false;
    binary_op(x, y, int64_t(N_1) * N_2 * N_3, OP_MULT);
    return true;
}
bool multiply4D(float* x, float* y, int N_1, int N_2, int N_3, int N_4)
{
    if (!x || !y || !dims_ok({N_1, N_2, N_3, N_4}))
        return false;
    binary_op(x, y, int64_t(N_1) * N_2 * N_3 * N_4, OP_MULT);
    return true;
}
bool divide3D(float* x, float* y, int N_1, int N_2, int N_3)
{
    if (!x || !y || !dims_ok({N_1, N_2, N_3}))
        return false;
    binary_op(x, y, int64_t(N_1) * N_2 * N_3, OP_DIV);
    return true;
}
bool divide4D(float* x, float* y, int N_1, int N_2, int N_3, int N_4)
{
    if (!x || !y || !dims_ok({N_1, N_2, N_3, N_4}))
        return false;
    binary_op(x, y, int64_t(N_1) * N_2 * N_3 * N_4, OP_DIV);
    return true;
}
// denom = num / denom (denom == 0 -> 0)
bool rdivide3D(float* num, float* denom, int N_1, int N_2, int N_3)
{
    if (!num || !denom || !dims_ok({N_1, N_2, N_3}))
        return false;
    binary_op(denom, num, int64_t(N_1) * N_2 * N_3, OP_RDIV);
    return true;
}
bool rdivide4D(float* num, float* denom, int N_1, int N_2, int N_3, int N_4)
{
    if (!num || !denom || !dims_ok({N_1, N_2, N_3, N_4}))
        return false;
    binary_op(denom, num, int64_t(N_1) * N_2 * N_3 * N_4, OP_RDIV);
    return true;
}
static void reciprocal_n(float* x, int64_t n, float dz)
{
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++)
        x[i] = (x[i] == 0.0f) ? dz : float(1.0 / x[i]);
}
bool reciprocal3D(float* x, int N_1, int N_2, int N_3, float divide_by_zero_value)
{
    if (!x || !dims_ok({N_1, N_2, N_3}))
        return false;
    reciprocal_n(x, int64_t(N_1) * N_2 * N_3, divide_by_zero_value);
    return true;
}
bool reciprocal4D(float* x, int N_1, int N_2, int N_3, int N_4, float divide_by_zero_value)
{
    if (!x || !dims_ok({N_1, N_2, N_3, N_4}))
        return false;
    reciprocal_n(x, int64_t(N_1) * N_2 * N_3 * N_4, divide_by_zero_value);
    return true;
}
float inner_product3D(float* x, float* y, float* w, int N_1, int N_2, int N_3)
{
    if (!x || !y || !dims_ok({N_1, N_2, N_3}))
        return 0.0f;
    return inner3(x, y, w, N_1, N_2, N_3);
}
// 4D: float partials per N_1 slab accumulated in double (src/data_cube.cpp:339-356)
float inner_product4D(float* x, float* y, float* w, int N_1, int N_2, int N_3, int N_4)
{
    if (!x || !y || !dims_ok({N_1, N_2, N_3, N_4}))
        return 0.0f;
    const int64_t slab = int64_t(N_2) * N_3 * N_4;
    double acc = 0.0;
    for (int i = 0; i < N_1; i++)
        acc += double(inner3(x + i * slab, y + i * slab, w ? w + i * slab : nullptr, N_2, N_3, N_4));
    return float(acc);
}

bool TV_gradient(float* f, float* Df, int N_1, int N_2, int N_3, float delta, float beta, float p, bool data_on_cpu)
{
    if (!f || !Df || !dims_ok({N_1, N_2, N_3}))
        return false;
    return guarded("TV_gradient", [&] {
        Model* m = M();
        hipStream_t s = m->stream();
        const int64_t n = int64_t(N_1) * N_2 * N_3;
        if (data_on_cpu) {
            DevBuf a, b;
            a.ensure(n * sizeof(float));
            b.ensure(n * sizeof(float));
            h2d(a.p, f, n * sizeof(float), s);
            launch_tv_gradient(a.f(), b.f(), N_1, N_2, N_3, delta, beta, p, s);
            d2h(Df, b.p, n * sizeof(float), s);
            PETL_HIP_CHECK(hipStreamSynchronize(s));
        } else {
            launch_tv_gradient(f, Df, N_1, N_2, N_3, delta, beta, p, s);
            PETL_HIP_CHECK(hipStreamSynchronize(s));
        }
        return true;
    });
}

// src/relative_differences.cu:481-532 (Df == NULL writes into f)
bool relativeDifferences_gradient(float* f, float* Df, int N_1, int N_2, int N_3, float delta, float beta)
{
    if (f == nullptr || N_1 <= 0 || N_2 <= 0 || N_3 <= 0)
        return false;
    if (Df == nullptr)
        Df = f;
    return guarded("relativeDifferences_gradient", [&] {
        Model* m = M();
        hipStream_t s = m->stream();
        const int64_t n = int64_t(N_1) * N_2 * N_3;
        DevBuf a, b;
        a.ensure(n * sizeof(float));
        b.ensure(n * sizeof(float));
        h2d(a.p, f, n * sizeof(float), s);
        launch_rd_gradient(a.f(), b.f(), N_1, N_2, N_3, delta, beta, s);
        d2h(Df, b.p, n * sizeof(float), s);
        PETL_HIP_CHECK(hipStreamSynchronize(s));
        return true;
    });
}

// ---- TV cost / quadratic form and Gaussian smoothing (the RWLS step size and default weights) ----
float TV_cost(float* f, int N_1, int N_2, int N_3, float delta, float beta, float p, bool data_on_cpu)
{
    return tv_reduce_export("TV_cost", false, f, nullptr, N_1, N_2, N_3, delta, beta, p, data_on_cpu);
}

float TV_quadForm(float* f, float* d, int N_1, int N_2, int N_3, float delta, float beta, float p, bool data_on_cpu)
{
    return tv_reduce_export("TV_quadForm", true, f, d, N_1, N_2, N_3, delta, beta, p, data_on_cpu);
}

// src/relative_differences.cu:753-800 (host array, smoothed in place)
bool gaussian_filter(float* f, int N_1, int N_2, int N_3, float FWHM, int numDims)
{
    if (f == nullptr || N_1 <= 0 || N_2 <= 0 || N_3 <= 0)
        return false;
    return guarded("gaussian_filter", [&] {
        Model* m = M();
        hipStream_t s = m->stream();
        const int64_t n = int64_t(N_1) * N_2 * N_3;
        DevBuf a, b;
        a.ensure(n * sizeof(float));
        b.ensure(n * sizeof(float));
        h2d(a.p, f, n * sizeof(float), s);
        gaussian_filter_device(a.f(), b.f(), N_1, N_2, N_3, FWHM, numDims, false, s);
        d2h(f, a.p, n * sizeof(float), s);
        PETL_HIP_CHECK(hipStreamSynchronize(s));
        return true;
    });
}

// ---- response model: geometric solids, ray tracing, stopping power, solid angle ----
// src/petl_c_interface.cpp:271-310, src/petl.cpp:81-162
bool add_object(int type, float* c, float* r, float val, float* A, float* clip)
{
    Solid o;
    if (!solid_init(type, c, r, val, A, clip, o))
        return false;
    M()->phantom.push_back(o);
    return true;
}
bool add_module(int type, float* c, float* r, float val, float* A, float* clip)
{
    Solid o;
    if (!solid_init(type, c, r, val, A, clip, o))
        return false;
    M()->modules.push_back(o);
    return true;
}
bool clear_phantom()
{
    M()->phantom.clear();
    return true;
}
bool clear_modules()
{
    M()->modules.clear();
    return true;
}
bool scale_phantom(float scale_x, float scale_y, float scale_z)
{
    solids_scale(M()->phantom, scale_x, scale_y, scale_z);
    return true;
}
bool ray_trace(float** g, int oversampling)
{
    if (g == nullptr || !M()->geometryDefined())
        return false;
    return guarded("ray_trace", [&] { return run_ray_trace(*M(), g, oversampling, false); });
}
bool stopping_power(float** g, int oversampling)
{
    if (g == nullptr || !M()->geometryDefined())
        return false;
    return guarded("stopping_power", [&] { return run_ray_trace(*M(), g, oversampling, true); });
}
// planogram::apply_solid_angle_correction (src/planogram.cpp:257-301): each (v1, v0) view times
// c = 2 pi (1 + v0^2 + v1^2)^1.5 / (T_v1 T_v0), or times 1 / c
bool set_solid_angle_correction(float** g, bool do_inverse)
{
    if (g == nullptr)
        return false;
    Model* m = M();
    for (size_t i = 0; i < m->plans.size(); i++) {
        if (g[i] == nullptr)
            return false;
        const Planogram& pl = m->plans[i];
        const int64_t view = int64_t(pl.N_u1) * pl.N_u0;
#pragma omp parallel for collapse(2) schedule(static)
        for (int a = 0; a < pl.N_v1; a++)
            for (int b = 0; b < pl.N_v0; b++) {
                const float v1 = float(a) * pl.T_v1 + pl.v1_0;
                const float v0 = float(b) * pl.T_v0 + pl.v0_0;
                float c = float(2.0 * PETL_PI * std::pow(1.0 + v0 * v0 + v1 * v1, 1.5) / (pl.T_v1 * pl.T_v0));
                if (do_inverse)
                    c = float(1.0 / c);
                float* x = g[i] + (int64_t(a) * pl.N_v0 + b) * view;
                for (int64_t k = 0; k < view; k++)
                    x[k] *= c;
            }
    }
    return true;
}

// ---- outside the projector / MLEM path ----

bool simulate_scatter(float**, float*) { return unsupported("simulate_scatter"); }
// src/petl_c_interface.cpp:261-269, src/petl.cpp:310-385
bool FBP(float** g, float* f)
{
    Model* m = M();
    if (g == nullptr || f == nullptr || !m->allDefined())
        return false;
    return guarded("FBP", [&] { return run_fbp(*m, g, f); });
}
bool PFDR(float** g, float** g_reb)
{
    Model* m = M();
    if (g == nullptr || g_reb == nullptr || !m->geometryDefined())
        return false;
    return guarded("PFDR", [&] { return run_pfdr(*m, g, g_reb); });
}
bool voxelize(float*, int) { return unsupported("voxelize"); }
bool apply_corrections(float**, float**, float) { return unsupported("apply_corrections"); }
bool diffuse(float*, int, int, int, float, float, int, bool) { return unsupported("diffuse"); }
bool TV_denoise(float*, int, int, int, float, float, float, int, bool, bool) { return unsupported("TV_denoise"); }

// =================================================================================================
// Part 2: additions
// =================================================================================================
bool set_backprojector_mode(int mode)
{
    if (mode != BACK_PETL && mode != BACK_MATCHED)
        return false;
    M()->back_mode = mode;
    return true;
}
int get_backprojector_mode() { return M()->back_mode; }

bool set_projector_path(int path)
{
    if (path != 0 && path != 1)
        return false;
    M()->path = path;
    return true;
}
int get_projector_path() { return M()->path; }

int planogram_uses_fast_path(int which)
{
    Model* m = M();
    if (!m->allDefined())
        return 0;
    bool r = false;
    guarded("planogram_uses_fast_path", [&] {
        r = m->uses_fast(which);
        return true;
    });
    return r ? 1 : 0;
}

bool project_gpu(float** d_g, float* d_f, void* stream, bool sync)
{
    Model* m = M();
    if (!d_g || !d_f || !m->allDefined())
        return false;
    return guarded("project_gpu", [&] {
        hipStream_t s = static_cast<hipStream_t>(stream);  // NULL = the default stream
        m->forward(d_g, d_f, EPI_NONE, nullptr, nullptr, s);
        if (sync)
            PETL_HIP_CHECK(hipStreamSynchronize(s));
        return true;
    });
}

bool backproject_gpu(float** d_g, float* d_f, void* stream, bool sync)
{
    Model* m = M();
    if (!d_g || !d_f || !m->allDefined())
        return false;
    return guarded("backproject_gpu", [&] {
        hipStream_t s = static_cast<hipStream_t>(stream);  // NULL = the default stream
        m->back(d_g, d_f, m->back_mode, s);
        if (sync)
            PETL_HIP_CHECK(hipStreamSynchronize(s));
        return true;
    });
}

bool project_gpu_batch(float** d_g, float* d_f, int batch, void* stream, bool sync)
{
    Model* m = M();
    if (!d_g || !d_f || batch < 1 || !m->allDefined())
        return false;
    return guarded("project_gpu_batch", [&] {
        hipStream_t s = static_cast<hipStream_t>(stream);
        m->forward_batch(d_g, d_f, batch, s);
        if (sync)
            PETL_HIP_CHECK(hipStreamSynchronize(s));
        return true;
    });
}

bool backproject_gpu_batch(float** d_g, float* d_f, int batch, int mode, void* stream, bool sync)
{
    Model* m = M();
    if (!d_g || !d_f || batch < 1 || mode < -1 || mode > 1 || !m->allDefined())
        return false;
    return guarded("backproject_gpu_batch", [&] {
        hipStream_t s = static_cast<hipStream_t>(stream);
        m->back_batch(d_g, d_f, batch, mode < 0 ? m->back_mode : mode, s);
        if (sync)
            PETL_HIP_CHECK(hipStreamSynchronize(s));
        return true;
    });
}

bool project_gpu_ex(void** d_g, const void* d_f, int volume_storage, int planogram_storage, float f_scale,
                    float g_scale, void* stream, bool sync)
{
    Model* m = M();
    if (!d_g || !d_f || !m->allDefined())
        return false;
    return guarded("project_gpu_ex", [&] {
        hipStream_t s = static_cast<hipStream_t>(stream);
        m->forward_st(volume_storage, planogram_storage, d_g, d_f, EPI_NONE, nullptr, nullptr, s, 0, 1, f_scale, 1.0f,
                      g_scale);
        if (sync)
            PETL_HIP_CHECK(hipStreamSynchronize(s));
        return true;
    });
}

bool backproject_gpu_ex(void** d_g, float* d_f, int storage, float g_scale, void* stream, bool sync)
{
    Model* m = M();
    if (!d_g || !d_f || !m->allDefined() || storage < STORE_F32 || storage > STORE_F16)
        return false;
    return guarded("backproject_gpu_ex", [&] {
        hipStream_t s = static_cast<hipStream_t>(stream);
        m->back_st(storage, d_g, d_f, m->back_mode, s, 0, 1, g_scale);
        if (sync)
            PETL_HIP_CHECK(hipStreamSynchronize(s));
        return true;
    });
}

// src/petl.cpp:65-79: zero the planograms, then bin the file with the reference's defaults (millimetres,
// chest-wall origin, bilinear u bins since parameters::binningType defaults to PRIMARY)
bool bin(float** g, char* file_name)
{
    Model* m = M();
    if (g == nullptr || file_name == nullptr || file_name[0] == 0 || !m->geometryDefined())
        return false;
    for (size_t i = 0; i < m->plans.size(); i++)
        if (!g[i])
            return false;
    return guarded("bin", [&] {
        ListModeReader rd(file_name);
        if (!rd.ok()) {
            printf("PETL bin: list mode data file %s not found\n", file_name);
            return false;
        }
        EventSource src;
        src.file = &rd;
        int64_t total = 0;
        const bool ok = run_bin(*m, g, src, true, false, true, nullptr, &total);
        printf("PETL bin: binned %lld coincidences\n", (long long)total);
        return ok;
    });
}

bool bin_events(const float* blue, const float* pink, long long n, float** g, int nearest, int chest_wall,
                long long* accepted)
{
    Model* m = M();
    if (!blue || !pink || n < 0 || !g || !m->geometryDefined())
        return false;
    for (size_t i = 0; i < m->plans.size(); i++)
        if (!g[i])
            return false;
    return guarded("bin_events", [&] {
        EventSource src;
        src.blue = blue;
        src.pink = pink;
        src.n = n;
        std::vector<int64_t> acc(m->plans.size());
        const bool ok = run_bin(*m, g, src, false, nearest != 0, chest_wall != 0, acc.data(), nullptr);
        if (accepted)
            for (size_t i = 0; i < acc.size(); i++)
                accepted[i] = acc[i];
        return ok;
    });
}

bool bin_events_gpu(const float* d_blue, const float* d_pink, long long n, float** d_g, int nearest, int chest_wall,
                    void* stream, bool sync)
{
    Model* m = M();
    if (!d_blue || !d_pink || n < 0 || !d_g || !m->geometryDefined())
        return false;
    return guarded("bin_events_gpu", [&] {
        hipStream_t s = static_cast<hipStream_t>(stream);
        DeviceState& D = m->dev;
        D.counter.ensure(m->plans.size() * sizeof(unsigned long long));
        bin_staged(*m, d_blue, d_pink, n, d_g, nearest != 0, chest_wall != 0,
                   static_cast<unsigned long long*>(D.counter.p), s);
        if (sync)
            PETL_HIP_CHECK(hipStreamSynchronize(s));
        return true;
    });
}

bool MLEM_gpu_storage(float** g, float* f, int numIter, float delta, float beta, int prior, float p, int numSubsets,
                      int planogram_storage, int image_storage)
{
    Model* m = M();
    return guarded("MLEM_gpu_storage", [&] {
        return run_mlem(*m, g, f, numIter, delta, beta, prior, p, numSubsets, nullptr, nullptr, nullptr,
                        planogram_storage, image_storage);
    });
}

bool MLEM_gpu(float** g, float* f, int numIter, float delta, float beta, int prior, float p, int numSubsets,
              float** mask, float** response, double* loglik)
{
    Model* m = M();
    return guarded("MLEM_gpu",
                   [&] { return run_mlem(*m, g, f, numIter, delta, beta, prior, p, numSubsets, mask, response, loglik); });
}

bool ray_trace_gpu(float** d_g, int oversampling, int stopping, void* stream, bool sync)
{
    Model* m = M();
    if (!d_g || !m->geometryDefined())
        return false;
    for (size_t i = 0; i < m->plans.size(); i++)
        if (!d_g[i])
            return false;
    return guarded("ray_trace_gpu", [&] {
        hipStream_t s = static_cast<hipStream_t>(stream);
        if (stopping && m->modules.empty()) {
            for (size_t i = 0; i < m->plans.size(); i++)
                launch_fill(d_g[i], 1.0f, m->plans[i].elems(), s);
        } else {
            const int n = upload_solids(*m, stopping != 0, s);
            for (size_t i = 0; i < m->plans.size(); i++)
                ray_trace_device(m->plans[i].geom(), static_cast<const Solid*>(m->dev.solids.p), n, oversampling,
                                 stopping != 0, d_g[i], s);
        }
        if (sync)
            PETL_HIP_CHECK(hipStreamSynchronize(s));
        return true;
    });
}

bool RWLS_gpu(float** g, float* f, int numIter, float delta, float beta, int weight_mode, float** W, float w_scalar,
              int preconditioner, bool nonnegativityConstraint, float** response, double* cost, int* iterations)
{
    Model* m = M();
    return guarded("RWLS_gpu", [&] {
        return run_rwls(*m, g, f, numIter, delta, beta, weight_mode, W, w_scalar, preconditioner,
                        nonnegativityConstraint, response, cost, iterations);
    });
}

}  // extern "C"
