// Generic SF planogram projectors for any panel angle psi and any voxel/crystal pitch.
//
// These are the reference-faithful kernels: one thread per planogram bin for the forward
// (src/projectors.cu:11-238) and one thread per voxel, gathering over bins, for both back projector
// modes.  Floating-point contraction is switched off so each weight rounds exactly as the reference's
// source expression (and the CPU oracle) does.  The fast separable kernels in proj_fast.hip cover the
// psi in {0, 90 deg} / equal-pitch case that the benchmarks run; everything else lands here.
#include "petl_common.h"
#include "proj_kernels.h"

namespace petl {

namespace {

struct ViewGeom {
    float v1, v0;
    float rx, ry, rz;
    bool x_slices;  // |r.x| >= |r.y|: loop over x slices (src/projectors.cu:77)
};

__device__ __forceinline__ float coordf(int i, float T, float s)
{
#pragma clang fp contract(off)
    return float(i) * T + s;
}

__device__ __forceinline__ ViewGeom make_view(const PlanGeom& P, int iv1, int iv0)
{
#pragma clang fp contract(off)
    ViewGeom w;
    w.v1 = coordf(iv1, P.Tv1, P.v1_0);
    w.v0 = coordf(iv0, P.Tv0, P.v0_0);
    w.rx = -(w.v0 * P.cos_psi + P.sin_psi);
    w.ry = -(w.v0 * P.sin_psi - P.cos_psi);
    w.rz = -w.v1;
    w.x_slices = fabsf(w.rx) >= fabsf(w.ry);
    return w;
}

// Forward masks (src/projectors.cu:30-43).
__device__ __forceinline__ bool u0_masked(const PlanGeom& P, float u0, float v0)
{
#pragma clang fp contract(off)
    return fabsf(u0) - 0.5f * P.Tu0 > P.L + P.R * fabsf(v0);
}
__device__ __forceinline__ bool u1_masked(const PlanGeom& P, float u1, float v1)
{
#pragma clang fp contract(off)
    return fabsf(u1) - 0.5f * P.Tu1 > P.H - P.R * fabsf(v1);
}

// ---------------------------------------------------------------------------------------------
// forward: one thread per bin of one (v1,v0) view.  grid = (ceil(Nu1*Nu0/256), Nv0, Nv1 slab)
// ---------------------------------------------------------------------------------------------
__global__ void __launch_bounds__(256) fwd_generic_kernel(float* __restrict__ g, const float* __restrict__ f,
                                                          PlanGeom P, VolGeom V, int v1_first, int v1_stride,
                                                          int v1_chunk, int epilogue,
                                                          const float* __restrict__ y,
                                                          const float* __restrict__ resp,
                                                          const float* __restrict__ wts)
{
#pragma clang fp contract(off)
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= P.Nu1 * P.Nu0)
        return;
    const int iu1 = b / P.Nu0, iu0 = b - iu1 * P.Nu0;
    const int iv0 = blockIdx.y;
    const int iv1 = v1_first + (v1_chunk + int(blockIdx.z)) * v1_stride;
    const int64_t gi = ((int64_t(iv1) * P.Nv0 + iv0) * P.Nu1 + iu1) * P.Nu0 + iu0;

    const ViewGeom w = make_view(P, iv1, iv0);
    const float u1 = coordf(iu1, P.Tu1, P.u1_0);
    const float u0 = coordf(iu0, P.Tu0, P.u0_0);
    float out = 0.0f;
    if (!u0_masked(P, u0, w.v0) && !u1_masked(P, u1, w.v1)) {
        const float iu0_min = iu0 - 0.5f, iu0_max = iu0 + 0.5f;
        const float iu1_min = iu1 - 0.5f, iu1_max = iu1 + 0.5f;
        const float px = u0 * P.cos_psi, py = u0 * P.sin_psi, pz = u1;
        const float Tx_inv = 1.0f / V.Tx, Ty_inv = 1.0f / V.Ty, Tz_inv = 1.0f / V.Tz;
        const float Tu1_inv = 1.0f / P.Tu1, Tu0_inv = 1.0f / P.Tu0;
        const float width_z = 0.5f * P.Tu1;
        const float Tz_over_Tu1 = V.Tz * Tu1_inv;
        const float z_fw = 0.5f * V.Tz * Tu1_inv;
        float val = 0.0f;
        if (w.x_slices) {
            const float width_y = 0.5f * P.Tu0 * fabsf(P.sin_psi);
            const float sin_inv = 1.0f / P.sin_psi;
            const float rx_inv = 1.0f / w.rx;
            const float l_phi = V.Tx * fabsf(sin_inv);
            const float y_fw = 0.5f * V.Ty * fabsf(sin_inv) * Tu0_inv;
            for (int ix = 0; ix < V.Nx; ix++) {
                const float x = ix * V.Tx + V.x0;
                const float t = (x - px) * rx_inv;
                const float y_c = py + t * w.ry;
                const float z_c = pz + t * w.rz;
                const float y_A = y_c - width_y, y_B = y_c + width_y;
                const int iy_min = max(0, int(ceilf((y_A - 0.5f * V.Ty - V.y0) * Ty_inv)));
                const int iy_max = min(V.Ny - 1, int(floorf((y_B + 0.5f * V.Ty - V.y0) * Ty_inv)));
                const float z_A = z_c - width_z, z_B = z_c + width_z;
                const int iz_min = max(0, int(ceilf((z_A - 0.5f * V.Tz - V.z0) * Tz_inv)));
                const int iz_max = min(V.Nz - 1, int(floorf((z_B + 0.5f * V.Tz - V.z0) * Tz_inv)));
                float yy = iy_min * V.Ty + V.y0 - t * w.ry;
                const float z_start = (iz_min * V.Tz + V.z0 - t * w.rz - P.u1_0) * Tu1_inv;
                for (int iy = iy_min; iy <= iy_max; iy++) {
                    const float c = (yy * sin_inv - P.u0_0) * Tu0_inv;
                    const float w0 = fminf(iu0_max, c + y_fw) - fmaxf(iu0_min, c - y_fw);
                    if (w0 > 0.0f) {
                        float z = z_start;
                        const float* frow = f + (int64_t(iz_min) * V.Ny + iy) * V.Nx + ix;
                        for (int iz = iz_min; iz <= iz_max; iz++) {
                            const float w1 = fmaxf(0.0f, fminf(iu1_max, z + z_fw) - fmaxf(iu1_min, z - z_fw));
                            if (w1 > 0.0f)
                                val += frow[0] * w0 * w1;
                            z += Tz_over_Tu1;
                            frow += int64_t(V.Ny) * V.Nx;
                        }
                    }
                    yy += V.Ty;
                }
            }
            val *= l_phi;
        } else {
            const float width_x = 0.5f * P.Tu0 * fabsf(P.cos_psi);
            const float cos_inv = 1.0f / P.cos_psi;
            const float ry_inv = 1.0f / w.ry;
            const float l_phi = V.Tx * fabsf(cos_inv);
            const float x_fw = 0.5f * V.Tx * fabsf(cos_inv) * Tu0_inv;
            for (int iy = 0; iy < V.Ny; iy++) {
                const float yv = iy * V.Ty + V.y0;
                const float t = (yv - py) * ry_inv;
                const float x_c = px + t * w.rx;
                const float z_c = pz + t * w.rz;
                const float x_A = x_c - width_x, x_B = x_c + width_x;
                const int ix_min = max(0, int(ceilf((x_A - 0.5f * V.Tx - V.x0) * Tx_inv)));
                const int ix_max = min(V.Nx - 1, int(floorf((x_B + 0.5f * V.Tx - V.x0) * Tx_inv)));
                const float z_A = z_c - width_z, z_B = z_c + width_z;
                const int iz_min = max(0, int(ceilf((z_A - 0.5f * V.Tz - V.z0) * Tz_inv)));
                const int iz_max = min(V.Nz - 1, int(floorf((z_B + 0.5f * V.Tz - V.z0) * Tz_inv)));
                float xx = ix_min * V.Tx + V.x0 - t * w.rx;
                const float z_start = (iz_min * V.Tz + V.z0 - t * w.rz - P.u1_0) * Tu1_inv;
                for (int ix = ix_min; ix <= ix_max; ix++) {
                    const float c = (xx * cos_inv - P.u0_0) * Tu0_inv;
                    const float w0 = fminf(iu0_max, c + x_fw) - fmaxf(iu0_min, c - x_fw);
                    if (w0 > 0.0f) {
                        float z = z_start;
                        const float* frow = f + (int64_t(iz_min) * V.Ny + iy) * V.Nx + ix;
                        for (int iz = iz_min; iz <= iz_max; iz++) {
                            const float w1 = fmaxf(0.0f, fminf(iu1_max, z + z_fw) - fmaxf(iu1_min, z - z_fw));
                            if (w1 > 0.0f)
                                val += frow[0] * w0 * w1;
                            z += Tz_over_Tu1;
                            frow += int64_t(V.Ny) * V.Nx;
                        }
                    }
                    xx += V.Tx;
                }
            }
            val *= l_phi;
        }
        out = val * (1.0f / sqrtf(1.0f + w.v0 * w.v0 + w.v1 * w.v1));
    }
    if (resp)
        out = out * resp[gi];
    if (epilogue == EPI_RATIO) {
        // rdivide (src/data_cube.cpp:150-155): r = y / Pf, Pf == 0 -> 0
        out = (out == 0.0f) ? 0.0f : y[gi] / out;
        if (resp)
            out = out * resp[gi];
    } else if (epilogue == EPI_RESID) {
        // Pf - g, then the in-place weighting of src/petl_server.py:1790-1792
        out = out - y[gi];
        if (wts)
            out = out * wts[gi];
    }
    g[gi] = out;
}

// ---------------------------------------------------------------------------------------------
// back projection, PETL mode: src/projectors.cu:240-377 with exact bilinear weights.
// One thread per voxel; all v1 slabs and all bins gathered in the kernel.  f += sum.
// ---------------------------------------------------------------------------------------------
__global__ void __launch_bounds__(256) back_petl_generic_kernel(const float* __restrict__ g, float* __restrict__ f,
                                                                PlanGeom P, VolGeom V, int v1_first, int v1_stride)
{
#pragma clang fp contract(off)
    const int64_t vox = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (vox >= vol_elems(V))
        return;
    const int ix = int(vox % V.Nx);
    const int iy = int((vox / V.Nx) % V.Ny);
    const int iz = int(vox / (int64_t(V.Nx) * V.Ny));
    const float x = ix * V.Tx + V.x0;
    const float y = iy * V.Ty + V.y0;
    const float z = iz * V.Tz + V.z0;
    const float cs = P.cos_psi, sn = P.sin_psi;
    const float t = y * cs - x * sn;
    const float xdt = x * cs + y * sn;
    const float Tu1_inv = 1.0f / P.Tu1, Tu0_inv = 1.0f / P.Tu0;
    const float xdt_dy = 0.5f * V.Ty * sn, t_dy = 0.5f * V.Ty * cs;
    const float xdt_dx = 0.5f * V.Tx * cs, t_dx = -0.5f * V.Tx * sn;
    const int64_t row = int64_t(P.Nu1) * P.Nu0;
    float fsum = 0.0f;
    for (int iv1 = v1_first; iv1 < P.Nv1; iv1 += v1_stride) {
        const float v1 = coordf(iv1, P.Tv1, P.v1_0);
        const float u1 = z + v1 * t;
        float val = 0.0f;
        if (fabsf(u1) - 0.5f * P.Tu1 <= P.H - P.R * fabsf(v1)) {
            const float iu1_mid = (u1 - P.u1_0) * Tu1_inv;
            const float iu1_A = iu1_mid - 0.5f * V.Tz * Tu1_inv;
            const float iu1_B = iu1_mid + 0.5f * V.Tz * Tu1_inv;
            const int iu1_lo = max(0, int(ceilf(iu1_A - 0.5f)));
            const int iu1_hi = min(P.Nu1 - 1, int(floorf(iu1_B + 0.5f)));
            for (int iv0 = 0; iv0 < P.Nv0; iv0++) {
                const float v0 = coordf(iv0, P.Tv0, P.v0_0);
                const float wy = v0 * cs + sn;
                const float wx = cs - v0 * sn;
                const float l_phi = 1.0f / fmaxf(fabsf(wy), fabsf(wx)) * (1.0f / sqrtf(1.0f + v0 * v0 + v1 * v1));
                float A, B;
                if (fabsf(wy) > fabsf(wx)) {
                    const float a = (xdt - xdt_dy + v0 * (t - t_dy) - P.u0_0) * Tu0_inv;
                    const float bb = (xdt + xdt_dy + v0 * (t + t_dy) - P.u0_0) * Tu0_inv;
                    A = (wy >= 0.0f) ? a : bb;
                    B = (wy >= 0.0f) ? bb : a;
                } else {
                    const float a = (xdt - xdt_dx + v0 * (t - t_dx) - P.u0_0) * Tu0_inv;
                    const float bb = (xdt + xdt_dx + v0 * (t + t_dx) - P.u0_0) * Tu0_inv;
                    A = (wx >= 0.0f) ? a : bb;
                    B = (wx >= 0.0f) ? bb : a;
                }
                const int iu0_lo = max(0, int(ceilf(A - 0.5f)));
                const int iu0_hi = min(P.Nu0 - 1, int(floorf(B + 0.5f)));
                const float* gv = g + (int64_t(iv1) * P.Nv0 + iv0) * row;
                for (int iu0 = iu0_lo; iu0 <= iu0_hi; iu0++) {
                    const float w0 = fmaxf(0.0f, fminf(float(iu0) + 0.5f, B) - fmaxf(float(iu0) - 0.5f, A));
                    if (!(w0 > 0.0f))
                        continue;
                    for (int iu1 = iu1_lo; iu1 <= iu1_hi; iu1++) {
                        const float w1 = fmaxf(0.0f, fminf(float(iu1) + 0.5f, iu1_B) - fmaxf(float(iu1) - 0.5f, iu1_A));
                        if (w1 > 0.0f)
                            val += gv[int64_t(iu1) * P.Nu0 + iu0] * l_phi * w0 * w1;
                    }
                }
            }
        }
        fsum += val * V.Tx;
    }
    f[vox] += fsum;
}

// Forward tap weight of voxel (ix,iy,iz) in bin (iu1,iu0) of view w, reproducing the forward's exact
// float sequence for the voxel's slice.  Returns false when the forward would not visit this tap.
__device__ __forceinline__ bool fwd_tap(const PlanGeom& P, const VolGeom& V, const ViewGeom& w, int iu1, int iu0,
                                        int ix, int iy, int iz, float& wt)
{
#pragma clang fp contract(off)
    const float u1 = coordf(iu1, P.Tu1, P.u1_0);
    const float u0 = coordf(iu0, P.Tu0, P.u0_0);
    const float iu0_min = iu0 - 0.5f, iu0_max = iu0 + 0.5f;
    const float iu1_min = iu1 - 0.5f, iu1_max = iu1 + 0.5f;
    const float px = u0 * P.cos_psi, py = u0 * P.sin_psi, pz = u1;
    const float Tx_inv = 1.0f / V.Tx, Ty_inv = 1.0f / V.Ty, Tz_inv = 1.0f / V.Tz;
    const float Tu1_inv = 1.0f / P.Tu1, Tu0_inv = 1.0f / P.Tu0;
    const float width_z = 0.5f * P.Tu1;
    const float Tz_over_Tu1 = V.Tz * Tu1_inv;
    const float z_fw = 0.5f * V.Tz * Tu1_inv;
    float t, w0, l_phi;
    int iz_min, iz_max;
    if (w.x_slices) {
        const float width_y = 0.5f * P.Tu0 * fabsf(P.sin_psi);
        const float sin_inv = 1.0f / P.sin_psi;
        const float rx_inv = 1.0f / w.rx;
        l_phi = V.Tx * fabsf(sin_inv);
        const float y_fw = 0.5f * V.Ty * fabsf(sin_inv) * Tu0_inv;
        const float x = ix * V.Tx + V.x0;
        t = (x - px) * rx_inv;
        const float y_c = py + t * w.ry;
        const float z_c = pz + t * w.rz;
        const float y_A = y_c - width_y, y_B = y_c + width_y;
        const int iy_min = max(0, int(ceilf((y_A - 0.5f * V.Ty - V.y0) * Ty_inv)));
        const int iy_max = min(V.Ny - 1, int(floorf((y_B + 0.5f * V.Ty - V.y0) * Ty_inv)));
        if (iy < iy_min || iy > iy_max)
            return false;
        const float z_A = z_c - width_z, z_B = z_c + width_z;
        iz_min = max(0, int(ceilf((z_A - 0.5f * V.Tz - V.z0) * Tz_inv)));
        iz_max = min(V.Nz - 1, int(floorf((z_B + 0.5f * V.Tz - V.z0) * Tz_inv)));
        float yy = iy_min * V.Ty + V.y0 - t * w.ry;
        for (int k = iy_min; k < iy; k++)
            yy += V.Ty;
        const float c = (yy * sin_inv - P.u0_0) * Tu0_inv;
        w0 = fminf(iu0_max, c + y_fw) - fmaxf(iu0_min, c - y_fw);
    } else {
        const float width_x = 0.5f * P.Tu0 * fabsf(P.cos_psi);
        const float cos_inv = 1.0f / P.cos_psi;
        const float ry_inv = 1.0f / w.ry;
        l_phi = V.Tx * fabsf(cos_inv);
        const float x_fw = 0.5f * V.Tx * fabsf(cos_inv) * Tu0_inv;
        const float yv = iy * V.Ty + V.y0;
        t = (yv - py) * ry_inv;
        const float x_c = px + t * w.rx;
        const float z_c = pz + t * w.rz;
        const float x_A = x_c - width_x, x_B = x_c + width_x;
        const int ix_min = max(0, int(ceilf((x_A - 0.5f * V.Tx - V.x0) * Tx_inv)));
        const int ix_max = min(V.Nx - 1, int(floorf((x_B + 0.5f * V.Tx - V.x0) * Tx_inv)));
        if (ix < ix_min || ix > ix_max)
            return false;
        const float z_A = z_c - width_z, z_B = z_c + width_z;
        iz_min = max(0, int(ceilf((z_A - 0.5f * V.Tz - V.z0) * Tz_inv)));
        iz_max = min(V.Nz - 1, int(floorf((z_B + 0.5f * V.Tz - V.z0) * Tz_inv)));
        float xx = ix_min * V.Tx + V.x0 - t * w.rx;
        for (int k = ix_min; k < ix; k++)
            xx += V.Tx;
        const float c = (xx * cos_inv - P.u0_0) * Tu0_inv;
        w0 = fminf(iu0_max, c + x_fw) - fmaxf(iu0_min, c - x_fw);
    }
    if (iz < iz_min || iz > iz_max || !(w0 > 0.0f))
        return false;
    float z = (iz_min * V.Tz + V.z0 - t * w.rz - P.u1_0) * Tu1_inv;
    for (int k = iz_min; k < iz; k++)
        z += Tz_over_Tu1;
    const float w1 = fmaxf(0.0f, fminf(iu1_max, z + z_fw) - fmaxf(iu1_min, z - z_fw));
    if (!(w1 > 0.0f))
        return false;
    wt = w0 * w1 * l_phi;
    return true;
}

// ---------------------------------------------------------------------------------------------
// back projection, matched mode: exact transpose of fwd_generic_kernel, as a per-voxel gather over
// the bins whose forward enumeration visits the voxel (no atomics).  f += sum.
// ---------------------------------------------------------------------------------------------
__global__ void __launch_bounds__(256) back_matched_generic_kernel(const float* __restrict__ g, float* __restrict__ f,
                                                                   PlanGeom P, VolGeom V, int v1_first, int v1_stride)
{
#pragma clang fp contract(off)
    const int64_t vox = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (vox >= vol_elems(V))
        return;
    const int ix = int(vox % V.Nx);
    const int iy = int((vox / V.Nx) % V.Ny);
    const int iz = int(vox / (int64_t(V.Nx) * V.Ny));
    const float x = ix * V.Tx + V.x0;
    const float y = iy * V.Ty + V.y0;
    const float z = iz * V.Tz + V.z0;
    const float cs = P.cos_psi, sn = P.sin_psi;
    const int64_t row = int64_t(P.Nu1) * P.Nu0;
    float fsum = 0.0f;
    for (int iv1 = v1_first; iv1 < P.Nv1; iv1 += v1_stride) {
        for (int iv0 = 0; iv0 < P.Nv0; iv0++) {
            const ViewGeom w = make_view(P, iv1, iv0);
            // candidate window: exact projection of the voxel centre (src/planogram.cpp:130-141) plus
            // the footprint half widths; the line spacing across slices is 1/|r| of the slice axis.
            const float tv = y * cs - x * sn;
            const float u0c = (x * cs + y * sn + w.v0 * tv - P.u0_0) / P.Tu0;
            const float hw0 = w.x_slices ? (P.Tu0 * fabsf(sn) + V.Ty) * fabsf(w.rx) / (2.0f * P.Tu0) + 1.5f
                                         : (P.Tu0 * fabsf(cs) + V.Tx) * fabsf(w.ry) / (2.0f * P.Tu0) + 1.5f;
            const int a0 = max(0, int(ceilf(u0c - hw0)));
            const int b0 = min(P.Nu0 - 1, int(floorf(u0c + hw0)));
            const float obl = 1.0f / sqrtf(1.0f + w.v0 * w.v0 + w.v1 * w.v1);
            const float* gv = g + (int64_t(iv1) * P.Nv0 + iv0) * row;
            float val = 0.0f;
            for (int iu0 = a0; iu0 <= b0; iu0++) {
                const float u0 = coordf(iu0, P.Tu0, P.u0_0);
                if (u0_masked(P, u0, w.v0))
                    continue;
                const float tb = w.x_slices ? (x - u0 * cs) / w.rx : (y - u0 * sn) / w.ry;
                const float u1c = (z - tb * w.rz - P.u1_0) / P.Tu1;
                const float hw1 = (P.Tu1 + V.Tz) / (2.0f * P.Tu1) + 1.5f;
                const int a1 = max(0, int(ceilf(u1c - hw1)));
                const int b1 = min(P.Nu1 - 1, int(floorf(u1c + hw1)));
                for (int iu1 = a1; iu1 <= b1; iu1++) {
                    const float u1 = coordf(iu1, P.Tu1, P.u1_0);
                    if (u1_masked(P, u1, w.v1))
                        continue;
                    float wt;
                    if (fwd_tap(P, V, w, iu1, iu0, ix, iy, iz, wt))
                        val += gv[int64_t(iu1) * P.Nu0 + iu0] * wt;
                }
            }
            fsum += val * obl;
        }
    }
    f[vox] += fsum;
}

}  // namespace

void launch_fwd_generic(float* g, const float* f, const PlanGeom& P, const VolGeom& V, int epilogue,
                        const float* y, const float* resp, hipStream_t s, int v1_first, int v1_stride,
                        const float* wts)
{
    const int nb = ceil_div(int64_t(P.Nu1) * P.Nu0, 256);
    const int nv1 = (P.Nv1 - v1_first + v1_stride - 1) / v1_stride;
    for (int c = 0; c < nv1; c += 65535) {
        const int n = min(65535, nv1 - c);
        hipLaunchKernelGGL(fwd_generic_kernel, dim3(nb, P.Nv0, n), dim3(256), 0, s, g, f, P, V, v1_first, v1_stride, c,
                           epilogue, y, resp, wts);
        PETL_LAUNCH_CHECK();
    }
}

void launch_back_generic(const float* g, float* f, const PlanGeom& P, const VolGeom& V, int mode, hipStream_t s,
                         int v1_first, int v1_stride)
{
    const int nb = ceil_div(vol_elems(V), 256);
    if (mode == BACK_MATCHED)
        hipLaunchKernelGGL(back_matched_generic_kernel, dim3(nb), dim3(256), 0, s, g, f, P, V, v1_first, v1_stride);
    else
        hipLaunchKernelGGL(back_petl_generic_kernel, dim3(nb), dim3(256), 0, s, g, f, P, V, v1_first, v1_stride);
    PETL_LAUNCH_CHECK();
}

}  // namespace petl
