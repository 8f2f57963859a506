// General SF projector (any psi, any pitch): per-view FP64-built crossing tables, interval-overlap
// weights shared by the forward and its exact transpose, and the reference's voxel-driven PETL-mode back
// projector.  See proj_general.h for the model.
//
// Kernels:
//   gen_fwd_sep_kernel      separable form (see proj_general.h): one wave per view (v1, v0), tile of 64
//                           u0 bins (lanes) and pass of GEN_RB u1 rows; each lane walks its bin column up
//                           the pass's rows, per slice reading GEN_RB + 1 z values of each of its two a
//                           cells from the padded copy and accumulating four taps per row;
//   gen_back_matched_sep_kernel  its exact transpose, gathered per voxel;
//   gen_fwd_kernel          a workgroup is one view (v1, v0) x a tile of 64 u0 bins (lanes) x 4 u1 rows
//                           (waves); each thread walks the slices for its bin, reading the volume through
//                           the layout whose footprint axis is contiguous (f for y planes, the transposed
//                           copy for x planes), so the lanes of a wave read neighbouring voxels;
//   gen_back_matched_kernel one thread per voxel gathers, per view, the bins whose footprint it meets
//                           (candidate bin windows from the inverted crossings, widened by one bin, every
//                           tap re-evaluated with the forward's own functions): no atomics;
//   gen_back_petl_kernel    one thread per (x, y) column and ZC heights: per view row v1 the voxel-centre
//                           u1 test and u1 interval per height, per v0 the u0 interval of the column
//                           (independent of the height), exact bilinear weights.
#include <algorithm>
#include <cmath>
#include <vector>
#include "proj_general.h"
#include "proj_kernels.h"

namespace petl {

namespace {

struct GenArgs {
    PlanGeom P;
    VolGeom V;
    int Nsmax;
    float Bz, hz;
    const GenView* views;
    const float* xi0;
    const float4* pcoef;
    int v1_first, v1_stride, v1_chunk;
};

__device__ __forceinline__ float coord(int i, float T, float s)
{
#pragma clang fp contract(off)
    return float(i) * T + s;
}

// Forward masks of the reference (src/projectors.cu:30-43), evaluated in float.
__device__ __forceinline__ bool mask_u0(const PlanGeom& P, float u0, float v0)
{
#pragma clang fp contract(off)
    return fabsf(u0) - 0.5f * P.Tu0 > P.L + P.R * fabsf(v0);
}
__device__ __forceinline__ bool mask_u1(const PlanGeom& P, float u1, float v1)
{
#pragma clang fp contract(off)
    return fabsf(u1) - 0.5f * P.Tu1 > P.H - P.R * fabsf(v1);
}

// overlap length of [c - h, c + h] with the unit cell [i - 1/2, i + 1/2] (negative when they are apart)
__device__ __forceinline__ float overlap(float c, float h, float i)
{
    return fminf(c + h, i + 0.5f) - fmaxf(c - h, i - 0.5f);
}

// The crossings of a bin's line with slice s, in voxel-index units.  The forward and the matched back
// projector call exactly these, so they compute identical weights.
__device__ __forceinline__ float cross_a(const GenView& w, float xs0, int iu0) { return fmaf(w.Au, float(iu0), xs0); }
__device__ __forceinline__ float cross_z_base(const GenView& w, int iu0, int s)
{
    return fmaf(w.Bu, float(iu0), fmaf(w.Bs, float(s), w.Z0));
}
__device__ __forceinline__ float cross_z(float Bz, float zb, int iu1) { return fmaf(Bz, float(iu1), zb); }

// ---------------------------------------------------------------------------------------------------------
// forward
// ---------------------------------------------------------------------------------------------------------
__global__ void __launch_bounds__(256) gen_fwd_kernel(float* __restrict__ g, const float* __restrict__ f,
                                                      const float* __restrict__ fT, GenArgs A, int epilogue,
                                                      const float* __restrict__ y, const float* __restrict__ resp,
                                                      const float* __restrict__ wts)
{
    const PlanGeom& P = A.P;
    const VolGeom& V = A.V;
    const int nt0 = (P.Nu0 + 63) >> 6;
    const int t1 = blockIdx.x / nt0, t0 = blockIdx.x - t1 * nt0;
    const int iu0 = t0 * 64 + (threadIdx.x & 63), iu1 = t1 * 4 + (threadIdx.x >> 6);
    const int iv0 = blockIdx.y, iv1 = A.v1_first + (A.v1_chunk + int(blockIdx.z)) * A.v1_stride;
    if (iu0 >= P.Nu0 || iu1 >= P.Nu1)
        return;  // no barrier follows
    const int64_t gi = ((int64_t(iv1) * P.Nv0 + iv0) * P.Nu1 + iu1) * P.Nu0 + iu0;
    const float v0 = coord(iv0, P.Tv0, P.v0_0), v1 = coord(iv1, P.Tv1, P.v1_0);
    float out = 0.0f;
    if (!mask_u0(P, coord(iu0, P.Tu0, P.u0_0), v0) && !mask_u1(P, coord(iu1, P.Tu1, P.u1_0), v1)) {
        const GenView w = A.views[int64_t(iv1) * P.Nv0 + iv0];
        const int Na = w.xs ? V.Ny : V.Nx, Ns = w.xs ? V.Nx : V.Ny;
        const float* vol = w.xs ? fT : f;
        const int64_t sk = int64_t(V.Nx) * V.Ny;
        const float* xr = A.xi0 + int64_t(iv0) * A.Nsmax;
        const float hb = w.hb, hz = A.hz;
        float acc = 0.0f;
        for (int s = 0; s < Ns; s++) {
            const float xi = cross_a(w, xr[s], iu0);
            const int ilo = max(0, int(floorf(xi - hb - 0.5f))), ihi = min(Na - 1, int(ceilf(xi + hb + 0.5f)));
            if (ilo > ihi)
                continue;
            const float ze = cross_z(A.Bz, cross_z_base(w, iu0, s), iu1);
            const int klo = max(0, int(floorf(ze - hz - 0.5f))), khi = min(V.Nz - 1, int(ceilf(ze + hz + 0.5f)));
            const float* vs = vol + int64_t(s) * Na;
            for (int k = klo; k <= khi; k++) {
                const float wz = overlap(ze, hz, float(k));
                if (!(wz > 0.0f))
                    continue;
                const float* vk = vs + int64_t(k) * sk;
                for (int i = ilo; i <= ihi; i++) {
                    const float wa = overlap(xi, hb, float(i));
                    if (wa > 0.0f)
                        acc += vk[i] * (wa * wz);
                }
            }
        }
        out = acc * w.scale;
    }
    {
#pragma clang fp contract(off)
        if (resp)
            out = out * resp[gi];
        if (epilogue == EPI_RATIO) {
            // dataCube RDIV (src/data_cube.cpp:150-155): y / Pf, 0 where Pf == 0
            out = (out == 0.0f) ? 0.0f : y[gi] / out;
            if (resp)
                out = out * resp[gi];
        } else if (epilogue == EPI_RESID) {
            // Pf - g, weighted in place (src/petl_server.py:1790-1792)
            out = out - y[gi];
            if (wts)
                out = out * wts[gi];
        }
    }
    g[gi] = out;
}

// ---------------------------------------------------------------------------------------------------------
// separable forward
// ---------------------------------------------------------------------------------------------------------
// word 3 of a raw buffer resource on gfx9-family GPUs (32-bit untyped dword accesses)
constexpr int GEN_BUFFER_FLAGS = 0x00020000;

// The taps of the separable form: cells c0 and c0 + 1 of the axis with the overlaps of [c - h, c + h]
// (h <= 1/2 meets no other cell; a negative overlap is a cell the interval misses).
__device__ __forceinline__ int sep_cell(float c, float h) { return int(floorf(c + (0.5f - h))); }
__device__ __forceinline__ float sep_weight(float c, float h, int cell) { return fmaxf(0.0f, overlap(c, h, float(cell))); }

template <int RB>
__global__ void __launch_bounds__(64) gen_fwd_sep_kernel(float* __restrict__ g, const float* __restrict__ fpy,
                                                         const float* __restrict__ fpx, GenArgs A, int epilogue,
                                                         const float* __restrict__ y,
                                                         const float* __restrict__ resp,
                                                         const float* __restrict__ wts)
{
    constexpr int PZ = GEN_PZ;
    static_assert(PZ >= RB + 2, "the padded copies must hold every z window of a pass");
    const PlanGeom& P = A.P;
    const VolGeom& V = A.V;
    const int lane = threadIdx.x;
    // blockIdx.x = pass * (u0 tiles) + tile: one pass of RB u1 rows per wave
    const int ntl = (P.Nu0 + 63) >> 6;
    const int pass = blockIdx.x / ntl;
    const int iu0 = (blockIdx.x - pass * ntl) * 64 + lane;
    const int iv0 = blockIdx.y, iv1 = A.v1_first + (A.v1_chunk + int(blockIdx.z)) * A.v1_stride;
    const GenView w = A.views[int64_t(iv1) * P.Nv0 + iv0];
    const int Na = w.xs ? V.Ny : V.Nx, Ns = w.xs ? V.Nx : V.Ny, NaP = Na + 2;
    const int plane = Ns * NaP;  // cells of one z plane of the padded copy (two floats each)
    // buffer resource over the copy this view reads (its size in bytes fits 32 bits: the plan checks the
    // padded element count); loads past it would return zero, and every load here stays inside it
    const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(
        const_cast<float*>(w.xs ? fpx : fpy), 0, 8 * (V.Nz + 2 * PZ) * plane, GEN_BUFFER_FLAGS);
    const float v0 = coord(iv0, P.Tv0, P.v0_0), v1 = coord(iv1, P.Tv1, P.v1_0);
    const bool lane_live = iu0 < P.Nu0 && !mask_u0(P, coord(iu0, P.Tu0, P.u0_0), v0);
    const int iu0c = min(iu0, P.Nu0 - 1);
    const float* xr = A.xi0 + int64_t(iv0) * A.Nsmax;
    const float hb = w.hb, hz = A.hz;
    const int64_t grow = (int64_t(iv1) * P.Nv0 + iv0) * P.Nu1 * int64_t(P.Nu0);
    {
        const int r0 = pass * RB;
        // bit r: row r0 + r exists and is not masked (one scalar word for the pass)
        unsigned rbits = 0;
#pragma unroll
        for (int r = 0; r < RB; r++)
            if (r0 + r < P.Nu1 && !mask_u1(P, coord(r0 + r, P.Tu1, P.u1_0), v1))
                rbits |= 1u << r;
        rbits = __builtin_amdgcn_readfirstlane(rbits);
        const bool any_row = rbits != 0;
        float acc[RB];
#pragma unroll
        for (int r = 0; r < RB; r++)
            acc[r] = 0.0f;
        if (any_row && __any(lane_live)) {
            float xs_next = xr[0];
            for (int s = 0; s < Ns; s++) {
                // the next slice's table entry is requested a slice ahead
                const float xs_cur = xs_next;
                xs_next = xr[min(s + 1, Ns - 1)];
                const float xi = cross_a(w, xs_cur, iu0c);
                const int i0 = sep_cell(xi, hb);
                const float zb = cross_z_base(w, iu0c, s);
                const int k0 = sep_cell(zb, hz);
                const int kw = k0 + r0;  // the pass reads z = kw ... kw + RB
                // a window that reaches a cell of the volume lies inside the padded copy; one that does not
                // contributes nothing and reads a clamped in-range column instead
                const bool live = lane_live && i0 >= -1 && i0 <= Na - 1 && kw >= -(RB + 1) && kw <= V.Nz;
                if (!__any(live))
                    continue;
                float wa0 = 0.0f, wa1 = 0.0f, wz0 = 0.0f, wz1 = 0.0f;
                if (live) {
                    wa0 = sep_weight(xi, hb, i0);
                    wa1 = sep_weight(xi, hb, i0 + 1);
                    wz0 = sep_weight(zb, hz, k0);
                    wz1 = sep_weight(zb, hz, k0 + 1);
                }
                const float W00 = wa0 * wz0, W01 = wa0 * wz1, W10 = wa1 * wz0, W11 = wa1 * wz1;
                // byte offset of the lane's cell (kw, s, i0) in the padded copy, whose cells hold the values of
                // two heights: one 16-byte load gives cells i0 and i0 + 1 at heights kw + t and kw + t + 1,
                // and plane t of the window adds the uniform t * plane cells as the load's scalar offset (9
                // loads per slice for the 17 heights of a pass instead of 17 of 8 bytes)
                const int voff = 8 * (((live ? kw : 0) + PZ) * plane + s * NaP + (live ? i0 : 0) + 1);
                float va[RB + 1], vb[RB + 1];
#pragma unroll
                for (int t = 0; t <= RB; t += 2) {
                    // elements taken out as integers first: __builtin_bit_cast of a subscript of this vector
                    // type compiles to a one-dword load here
                    const auto q = __builtin_amdgcn_raw_buffer_load_b128(rs, voff, 8 * t * plane, 0);
                    const unsigned q0 = q[0], q1 = q[1], q2 = q[2], q3 = q[3];
                    va[t] = __uint_as_float(q0);
                    vb[t] = __uint_as_float(q2);
                    if (t + 1 <= RB) {
                        va[t + 1] = __uint_as_float(q1);
                        vb[t + 1] = __uint_as_float(q3);
                    }
                }
#pragma unroll
                for (int r = 0; r < RB; r++) {
                    acc[r] = fmaf(va[r], W00, acc[r]);
                    acc[r] = fmaf(vb[r], W10, acc[r]);
                    acc[r] = fmaf(va[r + 1], W01, acc[r]);
                    acc[r] = fmaf(vb[r + 1], W11, acc[r]);
                }
            }
        }
        if (iu0 >= P.Nu0)
            return;  // no barrier follows
#pragma unroll
        for (int r = 0; r < RB; r++) {
            if (r0 + r >= P.Nu1)
                break;
            const int64_t gi = grow + int64_t(r0 + r) * P.Nu0 + iu0;
            float out = (lane_live && (rbits >> r & 1u)) ? acc[r] * w.scale : 0.0f;
            {
#pragma clang fp contract(off)
                if (resp)
                    out = out * resp[gi];
                if (epilogue == EPI_RATIO) {
                    // dataCube RDIV (src/data_cube.cpp:150-155): y / Pf, 0 where Pf == 0
                    out = (out == 0.0f) ? 0.0f : y[gi] / out;
                    if (resp)
                        out = out * resp[gi];
                } else if (epilogue == EPI_RESID) {
                    // Pf - g, weighted in place (src/petl_server.py:1790-1792)
                    out = out - y[gi];
                    if (wts)
                        out = out * wts[gi];
                }
            }
            g[gi] = out;
        }
    }
}

// the exact transpose of gen_fwd_sep_kernel, gathered per voxel: the bins whose two a cells include this
// voxel's column (candidates from the inverted crossing, each re-evaluated with the forward's functions),
// then the (at most two) u1 rows whose z window places this voxel at tap 0 or 1.  f += A^T g.
__global__ void __launch_bounds__(256) gen_back_matched_sep_kernel(const float* __restrict__ g,
                                                                   float* __restrict__ f, GenArgs A)
{
    const PlanGeom& P = A.P;
    const VolGeom& V = A.V;
    const int ix = blockIdx.x * 64 + (threadIdx.x & 63), iy = blockIdx.y * 4 + (threadIdx.x >> 6);
    const int iz = blockIdx.z;
    if (ix >= V.Nx || iy >= V.Ny)
        return;
    const float hz = A.hz;
    const int64_t row = int64_t(P.Nu1) * P.Nu0;
    float sum = 0.0f;
    for (int iv1 = A.v1_first; iv1 < P.Nv1; iv1 += A.v1_stride) {
        const float v1 = coord(iv1, P.Tv1, P.v1_0);
        for (int iv0 = 0; iv0 < P.Nv0; iv0++) {
            const GenView w = A.views[int64_t(iv1) * P.Nv0 + iv0];
            const int i = w.xs ? iy : ix, s = w.xs ? ix : iy;
            const float fi = float(i), hb = w.hb;
            const float xs0 = A.xi0[int64_t(iv0) * A.Nsmax + s];
            float lo = (fi - hb - 0.5f - xs0) / w.Au, hi = (fi + hb + 0.5f - xs0) / w.Au;
            if (lo > hi) {
                const float t = lo;
                lo = hi;
                hi = t;
            }
            const int b0 = max(0, int(floorf(lo)) - 1), b1 = min(P.Nu0 - 1, int(ceilf(hi)) + 1);
            if (b0 > b1)
                continue;
            const float v0 = coord(iv0, P.Tv0, P.v0_0);
            const float* gv = g + (int64_t(iv1) * P.Nv0 + iv0) * row;
            float acc = 0.0f;
            for (int iu0 = b0; iu0 <= b1; iu0++) {
                const float xi = cross_a(w, xs0, iu0);
                const int ta = i - sep_cell(xi, hb);
                if (ta < 0 || ta > 1)
                    continue;
                const float wa = sep_weight(xi, hb, i);
                if (!(wa > 0.0f) || mask_u0(P, coord(iu0, P.Tu0, P.u0_0), v0))
                    continue;
                const float zb = cross_z_base(w, iu0, s);
                const int k0 = sep_cell(zb, hz);
#pragma unroll
                for (int tz = 0; tz < 2; tz++) {
                    const int iu1 = iz - k0 - tz;
                    if (iu1 < 0 || iu1 >= P.Nu1)
                        continue;
                    const float wz = sep_weight(zb, hz, k0 + tz);
                    if (!(wz > 0.0f) || mask_u1(P, coord(iu1, P.Tu1, P.u1_0), v1))
                        continue;
                    acc += gv[int64_t(iu1) * P.Nu0 + iu0] * (wa * wz);
                }
            }
            sum += acc * w.scale;
        }
    }
    f[(int64_t(iz) * V.Ny + iy) * V.Nx + ix] += sum;
}

// ---------------------------------------------------------------------------------------------------------
// separable PETL-mode back projection.  One wave = 64 columns along the lane axis (x when |sin psi| <=
// |cos psi|, so that the u1 offset v1 * t_col barely changes across the lanes, else y) x ZB heights x one
// chunk of v1 rows; part[chunk] receives the chunk's sums (every element of it written), and
// sum_partials_kernel adds T_x * sum over chunks to f.
// ---------------------------------------------------------------------------------------------------------
constexpr int PETL_ZB = 4;  // heights per pass of the PETL column kernel: few accumulators, many waves in flight
// heights per pass of the matched column kernel, whose per-candidate work (crossings, weights, masks) is
// larger: 8 ran the PET/X preset's matched back projection 16 % faster than 4, 16 slower
constexpr int MATCHED_ZB = 8;

struct ColumnArgs {
    int lane_x;   // 1: lanes along x
    int nch;      // chunks of v1 rows
    int per;      // v1 rows (of the subset) per chunk
    int nv1;      // v1 rows of the subset
};

template <int NP, int ZB>
__global__ void __launch_bounds__(64) gen_back_petl_sep_kernel(const float* __restrict__ g, float* __restrict__ part,
                                                               GenArgs A, ColumnArgs S)
{
    const PlanGeom& P = A.P;
    const VolGeom& V = A.V;
    const int lane = threadIdx.x;
    const int La = S.lane_x ? V.Nx : V.Ny;
    const int nlc = (La + 63) >> 6;
    const int o = blockIdx.x / nlc, la = (blockIdx.x - o * nlc) * 64 + lane;
    const bool valid = la < La;
    const int lac = min(la, La - 1);
    const int ix = S.lane_x ? lac : o, iy = S.lane_x ? o : lac;
    const int zp = blockIdx.y * ZB;
    const int ch = blockIdx.z;
    float tcol;
    {
#pragma clang fp contract(off)
        // the voxel centre's signed distance along the panel normal direction (src/projectors.cu:284)
        tcol = coord(iy, V.Ty, V.y0) * P.cos_psi - coord(ix, V.Tx, V.x0) * P.sin_psi;
    }
    const float x = coord(ix, V.Tx, V.x0), yv = coord(iy, V.Ty, V.y0);
    const float hz1 = 0.5f * V.Tz / P.Tu1;
    const int64_t vplane = int64_t(P.Nu1) * P.Nu0;
    const int rowb = 4 * P.Nu0;  // bytes per u1 row
    float acc[ZB];
#pragma unroll
    for (int r = 0; r < ZB; r++)
        acc[r] = 0.0f;
    const int k1 = min(S.nv1, (ch + 1) * S.per);
    for (int k = ch * S.per; k < k1; k++) {
        const int iv1 = A.v1_first + k * A.v1_stride;
        const float v1 = coord(iv1, P.Tv1, P.v1_0);
        // voxel-centre u1 test per height (src/projectors.cu:284-289)
        unsigned lbits = 0;
#pragma unroll
        for (int r = 0; r < ZB; r++) {
            bool live;
            {
                // unfused, as the reference evaluates it: a voxel on the boundary keeps the reference's side
#pragma clang fp contract(off)
                const float u1 = coord(zp + r, V.Tz, V.z0) + v1 * tcol;
                live = valid && zp + r < V.Nz && !(fabsf(u1) - 0.5f * P.Tu1 > P.H - P.R * fabsf(v1));
            }
            lbits |= unsigned(live) << r;
        }
        if (!__any(lbits != 0))
            continue;
        // the pass's u1 rows qa, qa + 1 and weights, from its first height; height zp + r reads rows qa + r
        // and qa + r + 1
        float u1p;
        {
#pragma clang fp contract(off)
            u1p = coord(zp, V.Tz, V.z0) + v1 * tcol;
        }
        const float mid = (u1p - P.u1_0) / P.Tu1;
        const float A1 = mid - hz1, B1 = mid + hz1;
        const int qa = int(floorf(A1 + 0.5f));
        const float w1a = fmaxf(0.0f, fminf(float(qa) + 0.5f, B1) - fmaxf(float(qa) - 0.5f, A1));
        const float w1b = fmaxf(0.0f, fminf(float(qa) + 1.5f, B1) - fmaxf(float(qa) + 0.5f, A1));
        // every lane's row window inside the view plane: the offsets of the window's rows go in the loads'
        // scalar offset; otherwise each load reads its row clamped into the plane and rows outside it count
        // as zero
        const bool inside = __all(qa >= 0 && qa + ZB < P.Nu1);
        float a1[ZB];
#pragma unroll
        for (int r = 0; r < ZB; r++)
            a1[r] = 0.0f;
        const GenView* wr = A.views + int64_t(iv1) * P.Nv0;
        for (int iv0 = 0; iv0 < P.Nv0; iv0++) {
            const float4 pc = A.pcoef[iv0];
            const float l = wr[iv0].lpetl;
            const float c0 = fmaf(x, pc.x, fmaf(yv, pc.y, pc.z));
            const float A0 = c0 - pc.w, B0 = c0 + pc.w;
            const int pa = int(floorf(A0 + 0.5f));
            const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(
                const_cast<float*>(g + (int64_t(iv1) * P.Nv0 + iv0) * vplane), 0, int(4 * vplane), GEN_BUFFER_FLAGS);
            if (NP == 2 && inside && P.Nu0 >= 2) {
                // one 8-byte load per row gives two neighbouring cells pl, pl + 1 of the row, pl = pa clamped
                // to [0, Nu0 - 2]: the footprint (at most one cell wide) meets only cells pa and pa + 1, so a
                // cell of a clamped pair outside them has no overlap (up to rounding at an exact cell edge),
                // and where pl = pa the sums are those of the per-cell form below, in the same order
                const int pl = min(max(pa, 0), P.Nu0 - 2);
                float Wa[2], Wb[2];
#pragma unroll
                for (int pt = 0; pt < 2; pt++) {
                    const float p = float(pl + pt);
                    const float wl = fmaxf(0.0f, fminf(p + 0.5f, B0) - fmaxf(p - 0.5f, A0)) * l;
                    Wa[pt] = wl * w1a;
                    Wb[pt] = wl * w1b;
                }
                const int voff = qa * rowb + 4 * pl;
                float G0[ZB + 1], G1[ZB + 1];
#pragma unroll
                for (int t = 0; t <= ZB; t++) {
                    // elements taken out as integers (bit casts of this vector type's subscripts miscompile)
                    const auto q = __builtin_amdgcn_raw_buffer_load_b64(rs, voff, t * rowb, 0);
                    const unsigned e0 = q[0], e1 = q[1];
                    G0[t] = __uint_as_float(e0);
                    G1[t] = __uint_as_float(e1);
                }
#pragma unroll
                for (int r = 0; r < ZB; r++) {
                    a1[r] = fmaf(G0[r], Wa[0], a1[r]);
                    a1[r] = fmaf(G0[r + 1], Wb[0], a1[r]);
                }
#pragma unroll
                for (int r = 0; r < ZB; r++) {
                    a1[r] = fmaf(G1[r], Wa[1], a1[r]);
                    a1[r] = fmaf(G1[r + 1], Wb[1], a1[r]);
                }
                continue;
            }
#pragma unroll
            for (int pt = 0; pt < NP; pt++) {
                const int p = pa + pt;
                const float w0 = (p >= 0 && p < P.Nu0)
                                     ? fmaxf(0.0f, fminf(float(p) + 0.5f, B0) - fmaxf(float(p) - 0.5f, A0))
                                     : 0.0f;
                const float wl = w0 * l;
                const float Wa = wl * w1a, Wb = wl * w1b;
                const int pc4 = 4 * min(max(p, 0), P.Nu0 - 1);
                float G[ZB + 1];
                if (inside) {
                    const int voff = qa * rowb + pc4;
#pragma unroll
                    for (int t = 0; t <= ZB; t++)
                        G[t] = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rs, voff, t * rowb, 0));
                } else {
#pragma unroll
                    for (int t = 0; t <= ZB; t++) {
                        const int q = qa + t;
                        const float v = __builtin_bit_cast(
                            float, __builtin_amdgcn_raw_buffer_load_b32(rs, min(max(q, 0), P.Nu1 - 1) * rowb + pc4, 0, 0));
                        G[t] = (q >= 0 && q < P.Nu1) ? v : 0.0f;
                    }
                }
#pragma unroll
                for (int r = 0; r < ZB; r++) {
                    a1[r] = fmaf(G[r], Wa, a1[r]);
                    a1[r] = fmaf(G[r + 1], Wb, a1[r]);
                }
            }
        }
#pragma unroll
        for (int r = 0; r < ZB; r++)
            if (lbits >> r & 1u)
                acc[r] += a1[r];
    }
    if (!valid)
        return;
    float* pc = part + int64_t(ch) * V.Nx * V.Ny * V.Nz;
#pragma unroll
    for (int r = 0; r < ZB; r++)
        if (zp + r < V.Nz)
            pc[(int64_t(zp + r) * V.Ny + iy) * V.Nx + ix] = acc[r];
}

// ---------------------------------------------------------------------------------------------------------
// column-driven matched back projection (the exact transpose of gen_fwd_sep_kernel with the same weights).
// Lanes and chunks as gen_back_petl_sep_kernel.  Per view and column: the bins iu0 whose two cells include the
// column's in-plane index (candidates from the inverted crossing, each re-evaluated with the forward's
// functions), and per such bin the rows a height of the pass takes: height z reads row z - k0 (weight wz0)
// and row z - k0 - 1 (weight wz1), a window of ZB + 1 rows; rows the forward's u1 mask zeroes count as zero.
// ---------------------------------------------------------------------------------------------------------
template <int ZB>
__global__ void __launch_bounds__(64) gen_back_matched_col_kernel(const float* __restrict__ g,
                                                                  float* __restrict__ part, GenArgs A,
                                                                  const int2* __restrict__ mrows, ColumnArgs S)
{
    const PlanGeom& P = A.P;
    const VolGeom& V = A.V;
    const int lane = threadIdx.x;
    const int La = S.lane_x ? V.Nx : V.Ny;
    const int nlc = (La + 63) >> 6;
    const int o = blockIdx.x / nlc, la = (blockIdx.x - o * nlc) * 64 + lane;
    const bool valid = la < La;
    const int lac = min(la, La - 1);
    const int ix = S.lane_x ? lac : o, iy = S.lane_x ? o : lac;
    const int zp = blockIdx.y * ZB;
    const int ch = blockIdx.z;
    const float hz = A.hz;
    const int64_t vplane = int64_t(P.Nu1) * P.Nu0;
    const int rowb = 4 * P.Nu0;
    float acc[ZB];
#pragma unroll
    for (int r = 0; r < ZB; r++)
        acc[r] = 0.0f;
    const int k1 = min(S.nv1, (ch + 1) * S.per);
    for (int k = ch * S.per; k < k1; k++) {
        const int iv1 = A.v1_first + k * A.v1_stride;
        const int2 mr = mrows[iv1];  // unmasked rows of this view row
        if (mr.x > mr.y)
            continue;
        for (int iv0 = 0; iv0 < P.Nv0; iv0++) {
            const GenView w = A.views[int64_t(iv1) * P.Nv0 + iv0];
            const int i = w.xs ? iy : ix, s = w.xs ? ix : iy;
            const float fi = float(i), hb = w.hb;
            const float xs0 = A.xi0[int64_t(iv0) * A.Nsmax + s];
            // candidate bins: the column is one of a bin's two cells when its crossing lies in
            // [i - 3/2 + hb, i + 1/2 + hb); that interval inverted through an approximate reciprocal, widened
            // by 0.01 bin on each side (far above the reciprocal's and the crossing's rounding), and every
            // candidate re-evaluated below with the forward's own functions
            const float ra = __builtin_amdgcn_rcpf(w.Au);
            float lo = (fi - 1.5f + hb - xs0) * ra, hi = (fi + 0.5f + hb - xs0) * ra;
            if (lo > hi) {
                const float t = lo;
                lo = hi;
                hi = t;
            }
            const int b0 = max(0, int(ceilf(lo - 0.01f))), b1 = min(P.Nu0 - 1, int(floorf(hi + 0.01f)));
            const float v0 = coord(iv0, P.Tv0, P.v0_0);
            const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(
                const_cast<float*>(g + (int64_t(iv1) * P.Nv0 + iv0) * vplane), 0, int(4 * vplane), GEN_BUFFER_FLAGS);
            float av[ZB];
#pragma unroll
            for (int r = 0; r < ZB; r++)
                av[r] = 0.0f;
            // candidates in lockstep (a lane past its range carries zero weights), so that a candidate whose
            // row window lies inside the unmasked rows for every lane reads with the row step as the loads'
            // scalar offset
            for (int c = 0; __any(b0 + c <= b1); c++) {
                const int iu0 = min(b0 + c, P.Nu0 - 1);
                const float xi = cross_a(w, xs0, iu0);
                const int ta = i - sep_cell(xi, hb);
                const float wa = sep_weight(xi, hb, i);
                const bool ok = valid && b0 + c <= b1 && ta >= 0 && ta <= 1 && wa > 0.0f &&
                                !mask_u0(P, coord(iu0, P.Tu0, P.u0_0), v0);
                if (!__any(ok))
                    continue;
                const float zb = cross_z_base(w, iu0, s);
                const int k0 = sep_cell(zb, hz);
                const float W0 = ok ? wa * sep_weight(zb, hz, k0) : 0.0f;
                const float W1 = ok ? wa * sep_weight(zb, hz, k0 + 1) : 0.0f;
                // window rows qb ... qb + ZB: height zp + r takes G[r + 1] (weight W0) and G[r] (W1)
                const int qb = ok ? zp - k0 - 1 : mr.x;
                float G[ZB + 1];
                if (__all(qb >= mr.x && qb + ZB <= mr.y)) {
                    const int voff = qb * rowb + 4 * iu0;
#pragma unroll
                    for (int t = 0; t <= ZB; t++)
                        G[t] = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rs, voff, t * rowb, 0));
                } else {
#pragma unroll
                    for (int t = 0; t <= ZB; t++) {
                        const int q = qb + t;
                        const float v = __builtin_bit_cast(
                            float, __builtin_amdgcn_raw_buffer_load_b32(rs, min(max(q, 0), P.Nu1 - 1) * rowb + 4 * iu0, 0, 0));
                        G[t] = (q >= mr.x && q <= mr.y) ? v : 0.0f;
                    }
                }
#pragma unroll
                for (int r = 0; r < ZB; r++) {
                    av[r] += G[r + 1] * W0;
                    av[r] += G[r] * W1;
                }
            }
#pragma unroll
            for (int r = 0; r < ZB; r++)
                acc[r] += av[r] * w.scale;
        }
    }
    if (!valid)
        return;
    float* pc = part + int64_t(ch) * V.Nx * V.Ny * V.Nz;
#pragma unroll
    for (int r = 0; r < ZB; r++)
        if (zp + r < V.Nz)
            pc[(int64_t(zp + r) * V.Ny + iy) * V.Nx + ix] = acc[r];
}

// f += scale * sum over chunks of part[chunk], in chunk order
__global__ void __launch_bounds__(256) sum_partials_kernel(float* __restrict__ f, const float* __restrict__ part,
                                                           int nch, int64_t n, float scale)
{
    const int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x;
    if (i >= n)
        return;
    float t = 0.0f;
    for (int c = 0; c < nch; c++)
        t += part[int64_t(c) * n + i];
    f[i] += scale * t;
}

// fpy[z + PZ][y][x + 1] and fpx[z + PZ][x][y + 1] from f[z][y][x], each cell the pair (f at z, f at z + 1),
// zeros elsewhere: one thread per cell of the larger padded plane set, 32 x 32 tiles through LDS for the
// transposed copy
__global__ void __launch_bounds__(256) pad_volumes_kernel(const float* __restrict__ f, float2* __restrict__ fpy,
                                                          float2* __restrict__ fpx, int Nx, int Ny, int Nz)
{
    __shared__ float2 tile[32][33];
    const int bx = blockIdx.x * 32, by = blockIdx.y * 32;
    const int zp = blockIdx.z, z = zp - GEN_PZ;  // padded plane
    const bool zin0 = z >= 0 && z < Nz, zin1 = z + 1 >= 0 && z + 1 < Nz;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
    const int64_t zo = int64_t(z) * Nx * Ny;
    const float2 zero = make_float2(0.0f, 0.0f);
    // fpy: row y, column x + 1 (columns 0 and Nx + 1 are zero)
    for (int r = ty; r < 32; r += 8) {
        const int x = bx + tx, yy = by + r;
        float2 v = zero;
        if (x < Nx && yy < Ny) {
            const int64_t o = zo + int64_t(yy) * Nx + x;
            if (zin0)
                v.x = f[o];
            if (zin1)
                v.y = f[o + int64_t(Nx) * Ny];
        }
        tile[r][tx] = v;
        if (yy < Ny && x < Nx)
            fpy[(int64_t(zp) * Ny + yy) * (Nx + 2) + x + 1] = v;
        if (yy < Ny && x == 0)
            fpy[(int64_t(zp) * Ny + yy) * (Nx + 2)] = zero;
        if (yy < Ny && x == Nx - 1)
            fpy[(int64_t(zp) * Ny + yy) * (Nx + 2) + Nx + 1] = zero;
    }
    __syncthreads();
    // fpx: row x, column y + 1 (columns 0 and Ny + 1 are zero)
    for (int r = ty; r < 32; r += 8) {
        const int yy = by + tx, x = bx + r;
        if (x < Nx && yy < Ny)
            fpx[(int64_t(zp) * Nx + x) * (Ny + 2) + yy + 1] = tile[tx][r];
        if (x < Nx && yy == 0)
            fpx[(int64_t(zp) * Nx + x) * (Ny + 2)] = zero;
        if (x < Nx && yy == Ny - 1)
            fpx[(int64_t(zp) * Nx + x) * (Ny + 2) + Ny + 1] = zero;
    }
}

// ---------------------------------------------------------------------------------------------------------
// matched back projection: the exact transpose of gen_fwd_kernel, gathered per voxel.  f += A^T g.
// ---------------------------------------------------------------------------------------------------------
__global__ void __launch_bounds__(256) gen_back_matched_kernel(const float* __restrict__ g, float* __restrict__ f,
                                                               GenArgs A)
{
    const PlanGeom& P = A.P;
    const VolGeom& V = A.V;
    const int ix = blockIdx.x * 64 + (threadIdx.x & 63), iy = blockIdx.y * 4 + (threadIdx.x >> 6);
    const int iz = blockIdx.z;
    if (ix >= V.Nx || iy >= V.Ny)
        return;
    const float hz = A.hz, Bz = A.Bz, fz = float(iz);
    const int64_t row = int64_t(P.Nu1) * P.Nu0;
    float sum = 0.0f;
    for (int iv1 = A.v1_first; iv1 < P.Nv1; iv1 += A.v1_stride) {
        const float v1 = coord(iv1, P.Tv1, P.v1_0);
        for (int iv0 = 0; iv0 < P.Nv0; iv0++) {
            const GenView w = A.views[int64_t(iv1) * P.Nv0 + iv0];
            const int i = w.xs ? iy : ix, s = w.xs ? ix : iy;
            const float fi = float(i), hb = w.hb;
            const float xs0 = A.xi0[int64_t(iv0) * A.Nsmax + s];
            // bins whose a crossing can lie within hb + 1/2 of this voxel (inverted, one bin of margin)
            float lo = (fi - hb - 0.5f - xs0) / w.Au, hi = (fi + hb + 0.5f - xs0) / w.Au;
            if (lo > hi) {
                const float t = lo;
                lo = hi;
                hi = t;
            }
            const int b0 = max(0, int(floorf(lo)) - 1), b1 = min(P.Nu0 - 1, int(ceilf(hi)) + 1);
            if (b0 > b1)
                continue;
            const float v0 = coord(iv0, P.Tv0, P.v0_0);
            const float* gv = g + (int64_t(iv1) * P.Nv0 + iv0) * row;
            float acc = 0.0f;
            for (int iu0 = b0; iu0 <= b1; iu0++) {
                const float wa = overlap(cross_a(w, xs0, iu0), hb, fi);
                if (!(wa > 0.0f) || mask_u0(P, coord(iu0, P.Tu0, P.u0_0), v0))
                    continue;
                const float zb = cross_z_base(w, iu0, s);
                const int r0 = max(0, int(floorf((fz - hz - 0.5f - zb) / Bz)) - 1);
                const int r1 = min(P.Nu1 - 1, int(ceilf((fz + hz + 0.5f - zb) / Bz)) + 1);
                for (int iu1 = r0; iu1 <= r1; iu1++) {
                    const float wz = overlap(cross_z(Bz, zb, iu1), hz, fz);
                    if (!(wz > 0.0f) || mask_u1(P, coord(iu1, P.Tu1, P.u1_0), v1))
                        continue;
                    acc += gv[int64_t(iu1) * P.Nu0 + iu0] * (wa * wz);
                }
            }
            sum += acc * w.scale;
        }
    }
    f[(int64_t(iz) * V.Ny + iy) * V.Nx + ix] += sum;
}

// ---------------------------------------------------------------------------------------------------------
// PETL-mode back projection (src/projectors.cu:240-377 semantics, exact bilinear weights).  f += B g.
// ---------------------------------------------------------------------------------------------------------
template <int ZC>
__global__ void __launch_bounds__(256) gen_back_petl_kernel(const float* __restrict__ g, float* __restrict__ f,
                                                            GenArgs A)
{
    const PlanGeom& P = A.P;
    const VolGeom& V = A.V;
    const int ix = blockIdx.x * 64 + (threadIdx.x & 63), iy = blockIdx.y * 4 + (threadIdx.x >> 6);
    const int z0 = blockIdx.z * ZC;
    if (ix >= V.Nx || iy >= V.Ny)
        return;
    float tcol;
    {
#pragma clang fp contract(off)
        // the voxel centre's signed distance along the panel normal direction (src/projectors.cu:284)
        tcol = coord(iy, V.Ty, V.y0) * P.cos_psi - coord(ix, V.Tx, V.x0) * P.sin_psi;
    }
    const float x = coord(ix, V.Tx, V.x0), yv = coord(iy, V.Ty, V.y0);
    const float hz1 = 0.5f * V.Tz / P.Tu1;
    const int64_t row = int64_t(P.Nu1) * P.Nu0;
    float acc[ZC];
#pragma unroll
    for (int j = 0; j < ZC; j++)
        acc[j] = 0.0f;
    for (int iv1 = A.v1_first; iv1 < P.Nv1; iv1 += A.v1_stride) {
        const float v1 = coord(iv1, P.Tv1, P.v1_0);
        float A1[ZC], B1[ZC];
        int q0[ZC], q1[ZC];
        bool any = false;
#pragma unroll
        for (int j = 0; j < ZC; j++) {
            float u1;
            bool live;
            {
#pragma clang fp contract(off)
                u1 = coord(z0 + j, V.Tz, V.z0) + v1 * tcol;
                live = (z0 + j < V.Nz) && !(fabsf(u1) - 0.5f * P.Tu1 > P.H - P.R * fabsf(v1));
            }
            const float mid = (u1 - P.u1_0) / P.Tu1;
            A1[j] = mid - hz1;
            B1[j] = mid + hz1;
            q0[j] = max(0, int(ceilf(A1[j] - 0.5f)));
            q1[j] = live ? min(P.Nu1 - 1, int(floorf(B1[j] + 0.5f))) : -1;
            any = any || (live && q0[j] <= q1[j]);
        }
        if (!any)
            continue;
        const GenView* wr = A.views + int64_t(iv1) * P.Nv0;
        for (int iv0 = 0; iv0 < P.Nv0; iv0++) {
            const float4 pc = A.pcoef[iv0];
            const float l = wr[iv0].lpetl;
            const float c0 = fmaf(x, pc.x, fmaf(yv, pc.y, pc.z));
            const float A0 = c0 - pc.w, B0 = c0 + pc.w;
            const int p0 = max(0, int(ceilf(A0 - 0.5f))), p1 = min(P.Nu0 - 1, int(floorf(B0 + 0.5f)));
            const float* gv = g + (int64_t(iv1) * P.Nv0 + iv0) * row;
            for (int iu0 = p0; iu0 <= p1; iu0++) {
                const float w0 = fmaxf(0.0f, fminf(float(iu0) + 0.5f, B0) - fmaxf(float(iu0) - 0.5f, A0));
                if (!(w0 > 0.0f))
                    continue;
                const float wl = w0 * l;
                const float* gc = gv + iu0;
#pragma unroll
                for (int j = 0; j < ZC; j++) {
                    for (int iu1 = q0[j]; iu1 <= q1[j]; iu1++) {
                        const float w1 = fmaxf(0.0f, fminf(float(iu1) + 0.5f, B1[j]) - fmaxf(float(iu1) - 0.5f, A1[j]));
                        if (w1 > 0.0f)
                            acc[j] += gc[int64_t(iu1) * P.Nu0] * (wl * w1);
                    }
                }
            }
        }
    }
#pragma unroll
    for (int j = 0; j < ZC; j++)
        if (z0 + j < V.Nz)
            f[(int64_t(z0 + j) * V.Ny + iy) * V.Nx + ix] += V.Tx * acc[j];
}

// ---------------------------------------------------------------------------------------------------------
// fT[z][x][y] = f[z][y][x], 32 x 32 tiles through LDS
// ---------------------------------------------------------------------------------------------------------
__global__ void __launch_bounds__(256) transpose_xy_kernel(const float* __restrict__ f, float* __restrict__ fT,
                                                           int Nx, int Ny)
{
    __shared__ float tile[32][33];
    const int bx = blockIdx.x * 32, by = blockIdx.y * 32;
    const int64_t zo = int64_t(blockIdx.z) * Nx * Ny;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
    for (int r = ty; r < 32; r += 8) {
        const int x = bx + tx, y = by + r;
        if (x < Nx && y < Ny)
            tile[r][tx] = f[zo + int64_t(y) * Nx + x];
    }
    __syncthreads();
    for (int r = ty; r < 32; r += 8) {
        const int y = by + tx, x = bx + r;
        if (x < Nx && y < Ny)
            fT[zo + int64_t(x) * Ny + y] = tile[tx][r];
    }
}

inline float coord_host(int i, float T, float s)
{
    volatile float a = float(i) * T;
    return a + s;
}

GenArgs args_of(const GenPlan& gp, int v1_first, int v1_stride)
{
    GenArgs A;
    A.P = gp.P;
    A.V = gp.V;
    A.Nsmax = gp.Nsmax;
    A.Bz = gp.Bz;
    A.hz = gp.hz;
    A.views = gp.views;
    A.xi0 = gp.xi0;
    A.pcoef = gp.pcoef;
    A.v1_first = v1_first;
    A.v1_stride = v1_stride;
    A.v1_chunk = 0;
    return A;
}

}  // namespace

GenPlan::~GenPlan()
{
    if (block)
        (void)hipFree(block);
}

std::unique_ptr<GenPlan> make_gen_plan(const PlanGeom& P, const VolGeom& V)
{
    auto gp = std::make_unique<GenPlan>();
    gp->P = P;
    gp->V = V;
    gp->Nsmax = std::max(V.Nx, V.Ny);
    gp->Bz = float(double(P.Tu1) / double(V.Tz));
    gp->hz = float(0.5 * double(P.Tu1) / double(V.Tz));
    const double cs = P.cos_psi, sn = P.sin_psi;  // the float values the reference works with
    const double Tu0 = P.Tu0, Tu1 = P.Tu1, Tz = V.Tz;
    std::vector<GenView> views(size_t(P.Nv1) * P.Nv0);
    std::vector<float> xi0(size_t(P.Nv0) * gp->Nsmax, 0.0f);
    std::vector<float4> pcoef(P.Nv0);
    struct PerV0 {
        int xs;
        double pa, ps, rs, Ta, S0, Ts, dom_inv;
    };
    std::vector<PerV0> pv(P.Nv0);
    float max_hb = 0.0f, max_h0 = 0.0f;
    for (int iv0 = 0; iv0 < P.Nv0; iv0++) {
        const float v0f = coord_host(iv0, P.Tv0, P.v0_0);
        // slice axis: the reference's float decision (src/projectors.cu:77)
        volatile float t1 = v0f * P.cos_psi;
        volatile float t2 = v0f * P.sin_psi;
        const float rxf = -(t1 + P.sin_psi), ryf = -(t2 - P.cos_psi);
        const int xs = std::fabs(rxf) >= std::fabs(ryf) ? 1 : 0;
        const double v0 = v0f;
        const double rx = -(v0 * cs + sn), ry = -(v0 * sn - cs);
        PerV0& q = pv[iv0];
        q.xs = xs;
        double ra, a0, Na;
        if (!xs) {
            q.pa = cs, q.ps = sn, ra = rx, q.rs = ry, a0 = V.x0, q.Ta = V.Tx, q.S0 = V.y0, q.Ts = V.Ty, Na = V.Nx;
        } else {
            q.pa = sn, q.ps = cs, ra = ry, q.rs = rx, a0 = V.y0, q.Ta = V.Ty, q.S0 = V.x0, q.Ts = V.Tx, Na = V.Ny;
        }
        (void)Na;
        const int Ns = xs ? V.Nx : V.Ny;
        // a crossing of bin u0 with plane S:  a_c = u0 (p_a - p_s r_a / r_s) + S r_a / r_s
        const double ka = q.pa - q.ps * ra / q.rs, ks = ra / q.rs;
        for (int s = 0; s < Ns; s++) {
            const double S = q.S0 + double(s) * q.Ts;
            xi0[size_t(iv0) * gp->Nsmax + s] = float((ka * double(P.u0_0) + ks * S - a0) / q.Ta);
        }
        const float Au = float(ka * Tu0 / q.Ta), hb = float(0.5 * Tu0 * std::fabs(q.pa) / q.Ta);
        // PETL back: the column's u0 centre x w_x + y w_y and the half width of the voxel mid-line's
        // projection along the dominant of w_y, w_x (decided in float, src/projectors.cu:314-347)
        volatile float a1 = v0f * P.cos_psi;
        volatile float a2 = v0f * P.sin_psi;
        const float wyf = a1 + P.sin_psi, wxf = P.cos_psi - a2;
        const double wy = v0 * cs + sn, wx = cs - v0 * sn;
        const bool ydom = std::fabs(wyf) > std::fabs(wxf);
        const double h0 = ydom ? 0.5 * double(V.Ty) * std::fabs(wy) / Tu0 : 0.5 * double(V.Tx) * std::fabs(wx) / Tu0;
        pcoef[iv0] = make_float4(float(wx / Tu0), float(wy / Tu0), float(-double(P.u0_0) / Tu0), float(h0));
        q.dom_inv = 1.0 / std::max(std::fabs(wy), std::fabs(wx));
        max_h0 = std::max(max_h0, pcoef[iv0].w);
        for (int iv1 = 0; iv1 < P.Nv1; iv1++) {
            GenView& w = views[size_t(iv1) * P.Nv0 + iv0];
            w.xs = xs;
            w.Au = Au;
            w.hb = hb;
        }
        gp->any_xs = gp->any_xs || xs;
        max_hb = std::max(max_hb, hb);
    }
    gp->sep = gp->Bz == 1.0f && gp->hz <= 0.5f && max_hb <= 0.5f &&
              gen_padded_elems(V) < (size_t(1) << 29);  // byte offsets of the padded copies fit 31 bits
    // the separable PETL back: one row per voxel height, a u0 interval meeting at most NP cells, view planes
    // whose byte offsets fit 31 bits
    if (gp->Bz == 1.0f && size_t(P.Nu1) * P.Nu0 < (size_t(1) << 29))
        gp->petl_np = max_h0 <= 0.5f ? 2 : (max_h0 <= 1.0f ? 3 : 0);
    // the rows the forward's u1 mask leaves per view row, evaluated in float as mask_u1 does
    std::vector<int2> mrows(P.Nv1);
    for (int iv1 = 0; iv1 < P.Nv1; iv1++) {
        const float v1f = coord_host(iv1, P.Tv1, P.v1_0);
        volatile float rv = P.R * std::fabs(v1f);
        const float thr = P.H - rv;
        int lo = P.Nu1, hi = -1;
        for (int iu1 = 0; iu1 < P.Nu1; iu1++) {
            volatile float d = std::fabs(coord_host(iu1, P.Tu1, P.u1_0)) - 0.5f * P.Tu1;
            if (!(d > thr)) {
                lo = std::min(lo, iu1);
                hi = std::max(hi, iu1);
            }
        }
        mrows[iv1] = make_int2(lo, hi);
    }
    for (int iv1 = 0; iv1 < P.Nv1; iv1++) {
        const double v1 = coord_host(iv1, P.Tv1, P.v1_0);
        for (int iv0 = 0; iv0 < P.Nv0; iv0++) {
            const PerV0& q = pv[iv0];
            const double v0 = coord_host(iv0, P.Tv0, P.v0_0);
            GenView& w = views[size_t(iv1) * P.Nv0 + iv0];
            // z crossing: z_c = u1 + (r_z / r_s) (S - u0 p_s), r_z = -v1
            const double kz = -v1 / q.rs;
            w.Z0 = float((double(P.u1_0) - double(V.z0) + kz * q.S0 - kz * q.ps * double(P.u0_0)) / Tz);
            w.Bs = float(kz * q.Ts / Tz);
            w.Bu = float(-kz * q.ps * Tu0 / Tz);
            const double obl = 1.0 / std::sqrt(1.0 + v0 * v0 + v1 * v1);
            // slice length T_x / |p_a| (src/projectors.cu:84,138) over the footprint normalisations
            // T_u0 |p_a| / T_a and T_u1 / T_z, times the obliquity
            w.scale = float(double(V.Tx) * q.Ta * Tz * obl / (Tu0 * q.pa * q.pa * Tu1));
            w.lpetl = float(q.dom_inv * obl);
        }
    }
    const size_t nv = views.size() * sizeof(GenView), nx = xi0.size() * sizeof(float),
                 np = pcoef.size() * sizeof(float4), nr = mrows.size() * sizeof(int2);
    auto al = [](size_t n) { return (n + 255) & ~size_t(255); };
    char* blk = nullptr;
    PETL_HIP_CHECK(hipMalloc(&blk, al(nv) + al(nx) + al(np) + al(nr)));
    gp->block = blk;
    gp->views = reinterpret_cast<GenView*>(blk);
    gp->xi0 = reinterpret_cast<float*>(blk + al(nv));
    gp->pcoef = reinterpret_cast<float4*>(blk + al(nv) + al(nx));
    gp->rows = reinterpret_cast<int2*>(blk + al(nv) + al(nx) + al(np));
    PETL_HIP_CHECK(hipMemcpy(gp->views, views.data(), nv, hipMemcpyHostToDevice));
    PETL_HIP_CHECK(hipMemcpy(gp->xi0, xi0.data(), nx, hipMemcpyHostToDevice));
    PETL_HIP_CHECK(hipMemcpy(gp->pcoef, pcoef.data(), np, hipMemcpyHostToDevice));
    PETL_HIP_CHECK(hipMemcpy(gp->rows, mrows.data(), nr, hipMemcpyHostToDevice));
    return gp;
}

void launch_transpose_xy(const float* f, float* fT, const VolGeom& V, hipStream_t s)
{
    hipLaunchKernelGGL(transpose_xy_kernel, dim3((V.Nx + 31) / 32, (V.Ny + 31) / 32, V.Nz), dim3(256), 0, s, f, fT,
                       V.Nx, V.Ny);
    PETL_LAUNCH_CHECK();
}

size_t gen_padded_elems(const VolGeom& V)
{
    // floats of one copy: two per cell
    return 2 * std::max(size_t(V.Nz + 2 * GEN_PZ) * V.Ny * (V.Nx + 2), size_t(V.Nz + 2 * GEN_PZ) * V.Nx * (V.Ny + 2));
}

void launch_pad_volumes(const float* f, float* fpy, float* fpx, const VolGeom& V, hipStream_t s)
{
    hipLaunchKernelGGL(pad_volumes_kernel, dim3((V.Nx + 31) / 32, (V.Ny + 31) / 32, V.Nz + 2 * GEN_PZ), dim3(256), 0,
                       s, f, reinterpret_cast<float2*>(fpy), reinterpret_cast<float2*>(fpx), V.Nx, V.Ny, V.Nz);
    PETL_LAUNCH_CHECK();
}

void gen_forward(const GenPlan& gp, float* g, const GenVolume& vol, int epilogue, const float* y,
                 const float* resp, const float* wts, hipStream_t s, int v1_first, int v1_stride)
{
    GenArgs A = args_of(gp, v1_first, v1_stride);
    const PlanGeom& P = gp.P;
    const int nv1 = (P.Nv1 - v1_first + v1_stride - 1) / v1_stride;
    if (gp.sep) {
        if (!vol.fpy || !vol.fpx)
            throw hip_error("general forward: the padded volume copies are missing");
        for (int c = 0; c < nv1; c += 65535) {
            A.v1_chunk = c;
            hipLaunchKernelGGL(gen_fwd_sep_kernel<GEN_RB>,
                               dim3(((P.Nu0 + 63) / 64) * ((P.Nu1 + GEN_RB - 1) / GEN_RB), P.Nv0, std::min(65535, nv1 - c)),
                               dim3(64), 0, s, g, vol.fpy, vol.fpx, A, epilogue, y, resp, wts);
            PETL_LAUNCH_CHECK();
        }
        return;
    }
    if (gp.any_xs && !vol.fT)
        throw hip_error("general forward: the transposed volume is missing");
    const int tiles = ((P.Nu0 + 63) / 64) * ((P.Nu1 + 3) / 4);
    for (int c = 0; c < nv1; c += 65535) {
        A.v1_chunk = c;
        hipLaunchKernelGGL(gen_fwd_kernel, dim3(tiles, P.Nv0, std::min(65535, nv1 - c)), dim3(256), 0, s, g, vol.f,
                           vol.fT, A, epilogue, y, resp, wts);
        PETL_LAUNCH_CHECK();
    }
}

namespace {

// lanes along x when |sin psi| <= |cos psi|; chunks of v1 rows so that the launch has about 131072 waves
// (short waves balance better: 8192 ran the PET/X preset's back projection 1.6x slower), at most 1 GiB of
// partial volumes
ColumnArgs column_args(const PlanGeom& P, const VolGeom& V, int nv1)
{
    ColumnArgs S;
    S.lane_x = std::fabs(P.sin_psi) <= std::fabs(P.cos_psi) ? 1 : 0;
    const int La = S.lane_x ? V.Nx : V.Ny, Lo = S.lane_x ? V.Ny : V.Nx;
    const int64_t waves = int64_t(Lo) * ((La + 63) / 64) * ((V.Nz + PETL_ZB - 1) / PETL_ZB);
    const int64_t vbytes = int64_t(V.Nx) * V.Ny * V.Nz * 4;
    int nch = int(std::max<int64_t>(1, (131072 + waves - 1) / waves));
    nch = int(std::min<int64_t>({int64_t(nch), int64_t(std::max(1, nv1)), int64_t(65535),
                                 std::max<int64_t>(1, (int64_t(1) << 30) / vbytes)}));
    S.nv1 = nv1;
    S.per = (nv1 + nch - 1) / nch;
    S.nch = (nv1 + S.per - 1) / S.per;
    return S;
}

}  // namespace

size_t gen_back_workspace_bytes(const GenPlan& gp)
{
    if (!gp.petl_np && !gp.sep)
        return 0;
    return size_t(column_args(gp.P, gp.V, gp.P.Nv1).nch) * gp.V.Nx * gp.V.Ny * gp.V.Nz * sizeof(float);
}

size_t gen_back_workspace_estimate(const PlanGeom& P, const VolGeom& V)
{
    if (float(double(P.Tu1) / double(V.Tz)) != 1.0f)
        return 0;
    return size_t(column_args(P, V, P.Nv1).nch) * V.Nx * V.Ny * V.Nz * sizeof(float);
}

void gen_back(const GenPlan& gp, const float* g, float* f, int mode, hipStream_t s, int v1_first, int v1_stride,
              float* workspace)
{
    const GenArgs A = args_of(gp, v1_first, v1_stride);
    const VolGeom& V = gp.V;
    if (mode == BACK_MATCHED && gp.sep && workspace && gp.P.Nu1 * int64_t(gp.P.Nu0) < (int64_t(1) << 29)) {
        const int nv1 = (gp.P.Nv1 - v1_first + v1_stride - 1) / v1_stride;
        if (nv1 <= 0)
            return;
        const ColumnArgs S = column_args(gp.P, V, nv1);
        const int La = S.lane_x ? V.Nx : V.Ny, Lo = S.lane_x ? V.Ny : V.Nx;
        const dim3 grid(unsigned(Lo * ((La + 63) / 64)), unsigned((V.Nz + MATCHED_ZB - 1) / MATCHED_ZB), unsigned(S.nch));
        hipLaunchKernelGGL(gen_back_matched_col_kernel<MATCHED_ZB>, grid, dim3(64), 0, s, g, workspace, A, gp.rows, S);
        PETL_LAUNCH_CHECK();
        const int64_t n = int64_t(V.Nx) * V.Ny * V.Nz;
        hipLaunchKernelGGL(sum_partials_kernel, dim3(unsigned((n + 255) / 256)), dim3(256), 0, s, f, workspace, S.nch,
                           n, 1.0f);
        PETL_LAUNCH_CHECK();
        return;
    }
    if (mode != BACK_MATCHED && gp.petl_np && workspace) {
        const int nv1 = (gp.P.Nv1 - v1_first + v1_stride - 1) / v1_stride;
        if (nv1 <= 0)
            return;
        // a subset has no more v1 rows than the whole planogram, so it needs no more chunks than the
        // workspace holds
        const ColumnArgs S = column_args(gp.P, V, nv1);
        const int La = S.lane_x ? V.Nx : V.Ny, Lo = S.lane_x ? V.Ny : V.Nx;
        const dim3 grid(unsigned(Lo * ((La + 63) / 64)), unsigned((V.Nz + PETL_ZB - 1) / PETL_ZB), unsigned(S.nch));
        if (gp.petl_np == 2)
            hipLaunchKernelGGL((gen_back_petl_sep_kernel<2, PETL_ZB>), grid, dim3(64), 0, s, g, workspace, A, S);
        else
            hipLaunchKernelGGL((gen_back_petl_sep_kernel<3, PETL_ZB>), grid, dim3(64), 0, s, g, workspace, A, S);
        PETL_LAUNCH_CHECK();
        const int64_t n = int64_t(V.Nx) * V.Ny * V.Nz;
        hipLaunchKernelGGL(sum_partials_kernel, dim3(unsigned((n + 255) / 256)), dim3(256), 0, s, f, workspace, S.nch,
                           n, V.Tx);
        PETL_LAUNCH_CHECK();
        return;
    }
    if (mode == BACK_MATCHED && gp.sep) {
        hipLaunchKernelGGL(gen_back_matched_sep_kernel, dim3((V.Nx + 63) / 64, (V.Ny + 3) / 4, V.Nz), dim3(256), 0, s,
                           g, f, A);
    } else if (mode == BACK_MATCHED) {
        hipLaunchKernelGGL(gen_back_matched_kernel, dim3((V.Nx + 63) / 64, (V.Ny + 3) / 4, V.Nz), dim3(256), 0, s, g,
                           f, A);
    } else {
        constexpr int ZC = 4;
        hipLaunchKernelGGL(gen_back_petl_kernel<ZC>, dim3((V.Nx + 63) / 64, (V.Ny + 3) / 4, (V.Nz + ZC - 1) / ZC),
                           dim3(256), 0, s, g, f, A);
    }
    PETL_LAUNCH_CHECK();
}

}  // namespace petl
