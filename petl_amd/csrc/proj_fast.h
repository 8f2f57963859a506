// Fast separable SF projector for panels at psi = 0 or 90 degrees with voxel pitch equal to the crystal
// pitch (the geometry of every benchmark configuration).
//
// In that geometry each planogram view (v0,v1) is a parallel projection, and a voxel's footprint on
// the (u0,u1) bins is the product of two unit-width boxes whose centres move linearly with the slice
// index.  The footprint overlaps therefore reduce to a shifted 2-tap interpolation along u1 (per view
// row v1) followed by a shifted 2-tap interpolation along u0 (per v0), summed over slices:
//
//     g[v1][v0][u1][u0] = scale(v0,v1) * sum_s  L0(v0,s) * L1(v1,s) * f[:, s, :]
//
// The shifts and weights depend only on (view, slice); they are computed once per geometry in FP64 on
// the host (src/projectors.cu:100-177 evaluates the same overlaps per tap in fp32) and kept on the
// device in small tables.
#pragma once
#include <memory>
#include "petl_common.h"

namespace petl {

struct FastTables {
    int2* xtab = nullptr;     // [Nv0][Ns + pad]  {k, wq}: u0 interpolation of slice s in view v0
    int2* ftab = nullptr;     // forward: [Ns / SB][Nv0][SB] the xtab entries of one slice block and view
    float* fwp = nullptr;     // forward: [Ns / SB][Nv0][16]  SB x (1 - wq) formed in float, chunk mask, pad
    int4* ztab = nullptr;     // [Nv1][Ns]  {m1, wa, wb, 0}: u1 interpolation of slice s in view v1
    float* fscale = nullptr;  // [Nv1][Nv0] forward / matched scale  l_phi / sqrt(1+v0^2+v1^2)
    float* pscale = nullptr;  // [Nv1][Nv0] PETL back scale  T_x / max(|wy|,|wx|) / sqrt(1+v0^2+v1^2)
    int2* rows = nullptr;     // [2][Nv1]   active u1 rows {lo, hi}: [0] matched, [1] PETL mode
};

struct FastPlan {
    PlanGeom P;
    VolGeom V;
    int psi90 = 0;            // 0: slices along y, rows along x;  1: slices along x, rows along y
    int Ns = 0, Na = 0, Nz = 0;
    int64_t fs_s = 0, fs_a = 0, fs_z = 0;  // element strides of the slice / row / z axes in f[z][y][x]
    // forward kernel shape
    int J = 0, C = 0, R = 1, ngroups = 0, SB = 0, hpad = 0, hrow = 0, xrow = 0;
    // back kernel shape (a-groups of BACK_AG columns; the z chunk of back_z is chosen per launch)
    int nagroups = 0, nsgroups = 0, spad_lo = 0, srow = 0;
    FastTables dev;
    void* dev_block = nullptr;  // one allocation holding all tables
    ~FastPlan();
};

// Bytes of the back projector workspace a fast plan of this geometry holds (0 when the geometry is not
// one the fast kernels take); a host-only check that allocates nothing.
size_t fast_workspace_estimate(const PlanGeom& P, const VolGeom& V);
// Returns nullptr when the geometry is not eligible for the fast kernels.
std::unique_ptr<FastPlan> make_fast_plan(const PlanGeom& P, const VolGeom& V);

// g = A f (+ epilogue), one planogram.
// Views are restricted to v1 rows first, first+stride, ... (OSEM subsets); (0,1) = all.
// wts: weights of the EPI_RESID epilogue (nullable).
void fast_forward(const FastPlan& fp, float* g, const float* f, int epilogue, const float* y, const float* resp,
                  hipStream_t s, int v1_first = 0, int v1_stride = 1, const float* wts = nullptr);
// The same with the volume in storage format sf and the planogram and data y in sg (STORE_*; pairs: equal,
// or an fp32 volume with 16-bit planograms): x = stored * scale for STORE_F16 (f_scale, y_scale for the
// inputs, g_scale for the output); scales are ignored otherwise.
void fast_forward_st(const FastPlan& fp, int sf, int sg, void* g, const void* f, int epilogue, const void* y,
                     const float* resp, hipStream_t s, int v1_first, int v1_stride, float f_scale, float y_scale,
                     float g_scale);
// Weighted sum of squares of one projection, sum_b w_b (resp_b (A f)_b)^2 (w, resp nullable), without
// writing it: part[k] receives the FP64 partial of forward workgroup k; fast_forward_groups() of them.
// Summing the partials in index order gives a deterministic total.
int64_t fast_forward_groups(const FastPlan& fp, int v1_first = 0, int v1_stride = 1);
void fast_forward_sumsq(const FastPlan& fp, const float* f, const float* resp, const float* wts, double* part,
                        hipStream_t s, int v1_first = 0, int v1_stride = 1);
// g = A f for `batch` volumes f[b] (contiguous, V elements each) into planograms g[b] (contiguous, one
// planogram each): one launch (a batch axis in the grid) per 65535 / N_v1 elements.
void fast_forward_batch(const FastPlan& fp, float* g, const float* f, int batch, hipStream_t s);
// Bytes of scratch fast_back needs (the intermediate u0-gathered slabs).
size_t fast_back_workspace_bytes(const FastPlan& fp);
// f += A^T g (mode BACK_PETL or BACK_MATCHED), one planogram.
void fast_back(const FastPlan& fp, const float* g, float* f, int mode, float* workspace, hipStream_t s,
               int v1_first = 0, int v1_stride = 1);
// f[b] += A^T g[b] for `batch` planograms and volumes (contiguous), in launches of up to ws_elements
// elements (the workspace holds ws_elements x fast_back_workspace_bytes).
void fast_back_batch(const FastPlan& fp, const float* g, float* f, int mode, float* workspace, int ws_elements,
                     int batch, hipStream_t s);
// The same with the planogram in storage format st (g_scale as above); f stays fp32.
void fast_back_st(const FastPlan& fp, int st, const void* g, float* f, int mode, float* workspace, hipStream_t s,
                  int v1_first, int v1_stride, float g_scale);

}  // namespace petl
