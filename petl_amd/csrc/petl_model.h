// Host model: planogram/volume parameters of one PETL "model" (src/parameters.h:20-115,
// src/planogram.h:11-67) plus the device state that keeps buffers, kernel tables and a stream alive
// between calls.
#pragma once
#include <memory>
#include <vector>
#include "petl_common.h"
#include "ray_trace.h"
#include "proj_fast.h"
#include "proj_general.h"

namespace petl {

struct Planogram {
    float psi = 0, R = 0, L = 0, H = 0, v_m0 = 0, v_m1 = 0;
    int N_u0 = 0, N_u1 = 0, N_v0 = 0, N_v1 = 0;
    float T_u0 = 0, T_u1 = 0, T_v0 = 0, T_v1 = 0;
    float u0_0 = 0, u1_0 = 0, v0_0 = 0, v1_0 = 0;
    float cos_psi = 1, sin_psi = 0;

    // src/planogram.cpp:143-186 (psi in radians, half sizes)
    bool init(float psi_in, float R_in, float L_in, float H_in, float v_m0_in, float v_m1_in, float T);
    // add_planogram's arguments (src/petl_c_interface.cpp:216-219): psi in degrees, full panel separation
    // D and full panel length/height, converted in double as the reference does
    bool init_panels(float psi_deg, float D, float L, float H, float v_m0_in, float v_m1_in, float T);
    bool defined() const;
    PlanGeom geom() const;
    int64_t elems() const { return int64_t(N_v1) * N_v0 * N_u1 * N_u0; }
};

// Owned device allocation that grows on demand.
struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    void ensure(size_t n);
    void release();
    float* f() const { return static_cast<float*>(p); }
    ~DevBuf() { release(); }
};

// device scratch of the list-mode binning (sort keys and weights, sort workspace), grown on demand
struct BinScratch {
    DevBuf keys[2], vals[2], tmp, counts;
};

struct DeviceState {
    hipStream_t stream = nullptr;
    uint64_t built_version = 0;                   // geometry version the fast plans were made for
    int built_path = -1;
    std::vector<std::unique_ptr<FastPlan>> fast;  // per planogram; null = general kernels
    std::vector<std::unique_ptr<GenPlan>> gen;    // per planogram without a fast plan
    DevBuf f, f2, S, d;                           // volume-sized
    std::vector<DevBuf> g, y, aux, resp;          // planogram-sized
    DevBuf h2;                                    // fast back projector workspace
    size_t h2_need = 0;                           // bytes it needs for one planogram (the largest)
    DevBuf fT;                                    // the volume with x and y swapped (general forward)
    DevBuf fpad;                                  // the separable general forward's two padded copies
    DevBuf genws;                                 // the separable general PETL back's partial volumes
    DevBuf red;                                   // reduction scratch
    BinScratch bin;                               // list-mode binning
    DevBuf events, counter;                       // staged coincidences, accepted-event counters
    DevBuf stage;                                 // fp32 upload staging for storage-format data
    DevBuf solids;                                // the ray tracer's solids
    ~DeviceState();
};

struct Model {
    std::vector<Planogram> plans;
    // geometric phantom (add_object) and scintillator modules (add_module) of the response model
    std::vector<Solid> phantom, modules;
    int numX = 0, numY = 0, numZ = 0;
    float voxelWidth = 0, voxelHeight = 0, offsetX = 0, offsetY = 0, offsetZ = 0;
    int back_mode = BACK_PETL;
    int path = 0;  // 0 automatic, 1 generic kernels only
    uint64_t version = 1;
    DeviceState dev;

    bool volumeDefined() const;
    bool geometryDefined() const;
    bool allDefined() const { return geometryDefined() && volumeDefined(); }
    float x_0() const;
    float y_0() const;
    float z_0() const;
    VolGeom vgeom() const;
    int64_t vol_elems() const { return int64_t(numX) * numY * numZ; }
    void touch() { version++; }
    void copy_params_from(const Model& o);

    // device helpers (petl_device.hip)
    hipStream_t stream();
    void prepare();  // (re)build per-planogram kernel plans when geometry changed
    // whether the kernel plans match the current geometry (no prepare() needed)
    bool prepared() const;
    // allocates the fast back projector's workspace (one planogram at a time) if it is not there yet
    void ensure_workspace();
    bool uses_fast(int i);
    // Which copies of the volume the general forwards of one call have made so far (dev.fT transposed,
    // dev.fpad padded): one of each serves every general planogram of the call.
    struct GenCopies {
        bool transposed = false, padded = false;
    };
    // the general forward of planogram i, making the copy of f its kernel reads first unless `made` has it
    void gen_forward_one(size_t i, float* g, const float* f, GenCopies& made, int epilogue, const float* y,
                         const float* resp, const float* wts, hipStream_t s, int subset, int nsubsets);
    // allocates now what the general planograms' projections would allocate on first use (the volume copies
    // of the forward, the back projectors' partial volumes), so that a later call, e.g. under graph
    // capture, allocates nothing
    void ensure_general_buffers();
    // the partial volumes of planogram i's column-driven general back projectors (allocated on first use),
    // or null when its kernels are the per-voxel ones, which need none
    float* gen_back_workspace(size_t i);
    // g_i = A_i f for all planograms (device pointers), with an optional epilogue (ratio, or residual with
    // the weights wts; the y, resp, wts lists may be null).  v1 subset (first, stride) restricts the views;
    // (0,1) = all.
    void forward(float* const* g, const float* f, int epilogue, float* const* y, float* const* resp, hipStream_t s,
                 int subset = 0, int nsubsets = 1, float* const* wts = nullptr);
    // Partials of sum_b w_b (resp_b (A_i f)_b)^2 for planogram i without writing the projection (fast
    // kernels only): returns the number of FP64 partials written to part (at most cap; summed in index
    // order they give the total), or 0 when planogram i takes the generic kernels or cap is too small.
    int64_t forward_sumsq_parts(int i, const float* f, const float* resp, const float* wts, double* part,
                                int64_t cap, hipStream_t s);
    // FP64 partials forward_sumsq_parts needs for planogram i (0 for the generic kernels)
    int64_t sumsq_parts(int i);
    // f = sum_i A_i^T g_i (overwrites f)
    // workspace: the fast back projector's workspace (at least dev.h2_need bytes), or null for the model's own
    void back(float* const* g, float* f, int mode, hipStream_t s, int subset = 0, int nsubsets = 1,
              float* workspace = nullptr);
    // Batches of `batch` volumes f (contiguous, vol_elems() each) and planograms g[i] (contiguous, batch x
    // planogram i): one launch per planogram on the fast kernels, a loop over the batch on the generic ones.
    // back_batch overwrites f with the sum over planograms, in `mode`.
    void forward_batch(float* const* g, const float* f, int batch, hipStream_t s);
    void back_batch(float* const* g, float* f, int batch, int mode, hipStream_t s);
    // The same with the volume in storage format sf and the planograms in sg (STORE_*; fast kernels only,
    // the generic kernels are fp32).  For STORE_F16, x = stored * scale with the scales given; back writes
    // an fp32 volume.
    void forward_st(int sf, int sg, void* const* g, const void* f, int epilogue, void* const* y, float* const* resp,
                    hipStream_t s, int subset, int nsubsets, float f_scale, float y_scale, float g_scale);
    void back_st(int st, void* const* g, float* f, int mode, hipStream_t s, int subset, int nsubsets, float g_scale);
};

Model* current_model();

}  // namespace petl
