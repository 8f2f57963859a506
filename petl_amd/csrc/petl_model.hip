// Host model (geometry/volume parameters, registry) and device dispatch of the projectors.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include "petl_model.h"
#include "proj_kernels.h"
#include "vec_kernels.h"

namespace petl {

// ---------------------------------------------------------------------------------------------
// planogram geometry: src/planogram.cpp:143-186.  Float/double promotion as in the C++ source.
// ---------------------------------------------------------------------------------------------
bool Planogram::init(float psi_in, float R_in, float L_in, float H_in, float v_m0_in, float v_m1_in, float T)
{
    if (R_in <= 0.0f || L_in <= 0.0f || H_in <= 0.0f || v_m0_in <= 0.0f || v_m1_in < 0.0f || T <= 0.0f)
        return false;
    psi = psi_in;
    R = R_in;
    L = L_in;
    H = H_in;
    v_m0 = v_m0_in;
    const float h_over_r = H / R;
    v_m1 = (v_m1_in < h_over_r) ? v_m1_in : h_over_r;
    T_u0 = T;
    T_u1 = T;
    volatile float tv0 = T_u0 / R;
    volatile float tv1 = T_u1 / R;
    T_v0 = tv0;
    T_v1 = tv1;
    volatile float lr = R * v_m0;
    volatile float lsum = L + lr;
    N_u0 = 2 * int(lsum / T_u0);
    N_u1 = 2 * int(H / T_u1);
    N_v0 = 2 * int(v_m0 / T_v0) + 1;
    N_v1 = 2 * int(v_m1 / T_v1) + 1;
    T_v0 = float(double(v_m0) * 2.0 / double(float(N_v0)));
    volatile float a0 = -T_u0 * float(N_u0 - 1);
    volatile float a1 = -T_u1 * float(N_u1 - 1);
    volatile float a2 = -T_v0 * float(N_v0 - 1);
    volatile float a3 = -T_v1 * float(N_v1 - 1);
    u0_0 = float(double(a0) / 2.0);
    u1_0 = float(double(a1) / 2.0);
    v0_0 = float(double(a2) / 2.0);
    v1_0 = float(double(a3) / 2.0);
    cos_psi = float(std::cos(double(psi)));
    sin_psi = float(std::sin(double(psi)));
    return true;
}

bool Planogram::init_panels(float psi_deg, float D, float L, float H, float v_m0_in, float v_m1_in, float T)
{
    return init(float(double(psi_deg) * 3.1415926535897932385 / 180.0), float(0.5 * double(D)),
                float(0.5 * double(L)), float(0.5 * double(H)), v_m0_in, v_m1_in, T);
}

// src/planogram.cpp:91-97
bool Planogram::defined() const
{
    return R > 0.0f && L > 0.0f && H > 0.0f && v_m0 > 0.0f && v_m1 >= 0.0f && T_u0 > 0.0f && T_u1 > 0.0f &&
           T_v0 > 0.0f && T_v1 > 0.0f && N_u0 > 0 && N_u1 > 0 && N_v0 > 0 && N_v1 > 0;
}

PlanGeom Planogram::geom() const
{
    PlanGeom g;
    g.Nv1 = N_v1;
    g.Nv0 = N_v0;
    g.Nu1 = N_u1;
    g.Nu0 = N_u0;
    g.Tv1 = T_v1;
    g.Tv0 = T_v0;
    g.Tu1 = T_u1;
    g.Tu0 = T_u0;
    g.v1_0 = v1_0;
    g.v0_0 = v0_0;
    g.u1_0 = u1_0;
    g.u0_0 = u0_0;
    g.cos_psi = cos_psi;
    g.sin_psi = sin_psi;
    g.R = R;
    g.L = L;
    g.H = H;
    g.psi = psi;
    return g;
}

// src/parameters.cpp:175-203
bool Model::geometryDefined() const
{
    if (plans.empty())
        return false;
    for (const auto& p : plans)
        if (!p.defined())
            return false;
    return true;
}

bool Model::volumeDefined() const
{
    return !(numX <= 0 || numY <= 0 || numZ <= 0 || voxelWidth <= 0.0f || voxelHeight <= 0.0f);
}

// src/parameters.cpp:271-284
float Model::x_0() const { return float((0.0 - 0.5 * (numX - 1)) * double(voxelWidth) + double(offsetX)); }
float Model::y_0() const { return float((0.0 - 0.5 * (numY - 1)) * double(voxelWidth) + double(offsetY)); }
float Model::z_0() const { return float((0.0 - 0.5 * (numZ - 1)) * double(voxelHeight) + double(offsetZ)); }

VolGeom Model::vgeom() const
{
    VolGeom v;
    v.Nx = numX;
    v.Ny = numY;
    v.Nz = numZ;
    v.Tx = voxelWidth;
    v.Ty = voxelWidth;
    v.Tz = voxelHeight;
    v.x0 = x_0();
    v.y0 = y_0();
    v.z0 = z_0();
    return v;
}

void Model::copy_params_from(const Model& o)
{
    plans = o.plans;
    numX = o.numX;
    numY = o.numY;
    numZ = o.numZ;
    voxelWidth = o.voxelWidth;
    voxelHeight = o.voxelHeight;
    offsetX = o.offsetX;
    offsetY = o.offsetY;
    offsetZ = o.offsetZ;
    touch();
}

// ---------------------------------------------------------------------------------------------
// device state
// ---------------------------------------------------------------------------------------------
void DevBuf::ensure(size_t n)
{
    if (n <= bytes && p)
        return;
    release();
    PETL_HIP_CHECK(hipMalloc(&p, std::max<size_t>(n, 256)));
    bytes = n;
}

void DevBuf::release()
{
    if (p)
        (void)hipFree(p);
    p = nullptr;
    bytes = 0;
}

DeviceState::~DeviceState()
{
    // buffers free themselves; the stream is left to process teardown (destroying it after the HIP
    // runtime has shut down is not safe)
}

hipStream_t Model::stream()
{
    if (!dev.stream)
        PETL_HIP_CHECK(hipStreamCreateWithFlags(&dev.stream, hipStreamNonBlocking));
    return dev.stream;
}

void Model::prepare()
{
    if (prepared())
        return;
    dev.fast.clear();
    dev.gen.clear();
    const VolGeom V = vgeom();
    size_t ws = 0;
    for (const auto& p : plans) {
        std::unique_ptr<FastPlan> fp;
        if (path == 0)
            fp = make_fast_plan(p.geom(), V);
        if (fp)
            ws = std::max(ws, fast_back_workspace_bytes(*fp));
        dev.gen.push_back(fp ? nullptr : make_gen_plan(p.geom(), V));
        dev.fast.push_back(std::move(fp));
    }
    dev.h2_need = ws;  // allocated by the first back projection that needs it
    dev.built_version = version;
    dev.built_path = path;
}

bool Model::prepared() const
{
    return dev.built_version == version && dev.built_path == path && dev.fast.size() == plans.size();
}

void Model::ensure_workspace()
{
    if (dev.h2_need)
        dev.h2.ensure(dev.h2_need);
}

bool Model::uses_fast(int i)
{
    prepare();
    return i >= 0 && i < int(dev.fast.size()) && dev.fast[i] != nullptr;
}

void Model::gen_forward_one(size_t i, float* g, const float* f, GenCopies& made, int epilogue, const float* y,
                            const float* resp, const float* wts, hipStream_t s, int subset, int nsubsets)
{
    const GenPlan& gp = *dev.gen[i];
    const VolGeom V = vgeom();
    GenVolume vol;
    vol.f = f;
    if (gp.sep) {
        const size_t n = gen_padded_elems(V);
        if (!made.padded) {
            dev.fpad.ensure(2 * n * sizeof(float));
            launch_pad_volumes(f, dev.fpad.f(), dev.fpad.f() + n, V, s);
            made.padded = true;
        }
        vol.fpy = dev.fpad.f();
        vol.fpx = dev.fpad.f() + n;
    } else if (gp.any_xs) {
        if (!made.transposed) {
            dev.fT.ensure(size_t(vol_elems()) * sizeof(float));
            launch_transpose_xy(f, dev.fT.f(), V, s);
            made.transposed = true;
        }
        vol.fT = dev.fT.f();
    }
    gen_forward(gp, g, vol, epilogue, y, resp, wts, s, subset, nsubsets);
}

void Model::forward(float* const* g, const float* f, int epilogue, float* const* y, float* const* resp, hipStream_t s,
                    int subset, int nsubsets, float* const* wts)
{
    prepare();
    GenCopies made;
    for (size_t i = 0; i < plans.size(); i++) {
        const float* yi = y ? y[i] : nullptr;
        const float* ri = resp ? resp[i] : nullptr;
        const float* wi = wts ? wts[i] : nullptr;
        if (dev.fast[i])
            fast_forward(*dev.fast[i], g[i], f, epilogue, yi, ri, s, subset, nsubsets, wi);
        else
            gen_forward_one(i, g[i], f, made, epilogue, yi, ri, wi, s, subset, nsubsets);
    }
}

int64_t Model::sumsq_parts(int i)
{
    prepare();
    return dev.fast[i] ? fast_forward_groups(*dev.fast[i]) : 0;
}

int64_t Model::forward_sumsq_parts(int i, const float* f, const float* resp, const float* wts, double* part,
                                   int64_t cap, hipStream_t s)
{
    prepare();
    if (!dev.fast[i])
        return 0;
    const int64_t n = fast_forward_groups(*dev.fast[i]);
    if (n > cap)
        return 0;
    fast_forward_sumsq(*dev.fast[i], f, resp, wts, part, s);
    return n;
}

void Model::ensure_general_buffers()
{
    prepare();
    const VolGeom V = vgeom();
    for (size_t i = 0; i < plans.size(); i++) {
        if (dev.fast[i])
            continue;
        const GenPlan& gp = *dev.gen[i];
        if (gp.sep)
            dev.fpad.ensure(2 * gen_padded_elems(V) * sizeof(float));
        else if (gp.any_xs)
            dev.fT.ensure(size_t(vol_elems()) * sizeof(float));
        gen_back_workspace(i);
    }
}

float* Model::gen_back_workspace(size_t i)
{
    const size_t b = gen_back_workspace_bytes(*dev.gen[i]);
    if (!b)
        return nullptr;
    dev.genws.ensure(b);
    return dev.genws.f();
}

void Model::back(float* const* g, float* f, int mode, hipStream_t s, int subset, int nsubsets, float* workspace)
{
    prepare();
    // a caller's workspace means no allocation here: the general planograms then take the per-voxel PETL
    // kernel, which needs none
    const bool own = !workspace;
    if (own) {
        ensure_workspace();
        workspace = dev.h2.f();
    }
    const VolGeom V = vgeom();
    launch_fill(f, 0.0f, vol_elems(), s);
    for (size_t i = 0; i < plans.size(); i++) {
        if (dev.fast[i])
            fast_back(*dev.fast[i], g[i], f, mode, workspace, s, subset, nsubsets);
        else
            gen_back(*dev.gen[i], g[i], f, mode, s, subset, nsubsets, own ? gen_back_workspace(i) : nullptr);
    }
}

void Model::forward_batch(float* const* g, const float* f, int batch, hipStream_t s)
{
    prepare();
    const VolGeom V = vgeom();
    const int64_t nv = vol_elems();
    for (size_t i = 0; i < plans.size(); i++) {
        if (dev.fast[i]) {
            fast_forward_batch(*dev.fast[i], g[i], f, batch, s);
            continue;
        }
        const int64_t G = plans[i].elems();
        for (int b = 0; b < batch; b++) {
            GenCopies made;  // each batch element's own copies
            gen_forward_one(i, g[i] + b * G, f + b * nv, made, EPI_NONE, nullptr, nullptr, nullptr, s, 0, 1);
        }
    }
}

// the batched back projector's workspace holds this many bytes at most (whole batch elements, at least one)
constexpr size_t BATCH_WS_CAP = size_t(4) << 30;

void Model::back_batch(float* const* g, float* f, int batch, int mode, hipStream_t s)
{
    prepare();
    const VolGeom V = vgeom();
    const int64_t nv = vol_elems();
    launch_fill(f, 0.0f, nv * batch, s);
    for (size_t i = 0; i < plans.size(); i++) {
        const int64_t G = plans[i].elems();
        if (dev.fast[i]) {
            const size_t ws = fast_back_workspace_bytes(*dev.fast[i]);
            const int k = int(std::max<size_t>(1, std::min<size_t>(size_t(batch), BATCH_WS_CAP / ws)));
            dev.h2.ensure(ws * size_t(k));
            fast_back_batch(*dev.fast[i], g[i], f, mode, dev.h2.f(), k, batch, s);
            continue;
        }
        for (int b = 0; b < batch; b++)
            gen_back(*dev.gen[i], g[i] + b * G, f + b * nv, mode, s, 0, 1, gen_back_workspace(i));
    }
}

void Model::forward_st(int sf, int sg, void* const* g, const void* f, int epilogue, void* const* y, float* const* resp,
                       hipStream_t s, int subset, int nsubsets, float f_scale, float y_scale, float g_scale)
{
    prepare();
    GenCopies made;
    for (size_t i = 0; i < plans.size(); i++) {
        if (!dev.fast[i] && (sf != STORE_F32 || sg != STORE_F32))
            throw hip_error("16-bit storage needs the fast kernels (psi 0 or 90 degrees, voxel pitch = crystal pitch)");
        if (!dev.fast[i]) {
            gen_forward_one(i, static_cast<float*>(g[i]), static_cast<const float*>(f), made, epilogue,
                            y ? static_cast<const float*>(y[i]) : nullptr, resp ? resp[i] : nullptr, nullptr, s,
                            subset, nsubsets);
            continue;
        }
        fast_forward_st(*dev.fast[i], sf, sg, g[i], f, epilogue, y ? y[i] : nullptr, resp ? resp[i] : nullptr, s, subset,
                        nsubsets, f_scale, y_scale, g_scale);
    }
}

void Model::back_st(int st, void* const* g, float* f, int mode, hipStream_t s, int subset, int nsubsets, float g_scale)
{
    prepare();
    ensure_workspace();
    launch_fill(f, 0.0f, vol_elems(), s);
    for (size_t i = 0; i < plans.size(); i++) {
        if (!dev.fast[i] && st != STORE_F32)
            throw hip_error("16-bit storage needs the fast kernels (psi 0 or 90 degrees, voxel pitch = crystal pitch)");
        if (!dev.fast[i]) {
            gen_back(*dev.gen[i], static_cast<const float*>(g[i]), f, mode, s, subset, nsubsets,
                     gen_back_workspace(i));
            continue;
        }
        fast_back_st(*dev.fast[i], st, g[i], f, mode, dev.h2.f(), s, subset, nsubsets, g_scale);
    }
}

}  // namespace petl
