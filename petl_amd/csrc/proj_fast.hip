// Fast separable SF projector and atomic-free gather back projector for gfx950 (see proj_fast.h).
//
// Axes.  For a psi = 0 panel the slices run along y and the u0 interpolation runs along x; for psi = 90
// deg the roles of x and y swap (src/projectors.cu:77 picks x slices when |r.x| >= |r.y|).  Below, "s"
// is the slice axis (Ns voxels), "a" the u0-interpolated axis (Na voxels) and z the u1-interpolated axis.
//
// Forward (fwd_fast_kernel), one workgroup per (u1 row, v1 view row, group of v0 views):
//   1. For a block of SB slices, interpolate the volume along z onto this u1 row,
//      H[s][a] = wa * f[iz_a][s][a] + wb * f[iz_a+1][s][a], into LDS (zero-padded rows; the v1 row's z
//      table sits in LDS, the next block's loads are in flight during the current block).
//   2. Each wave owns up to J v0 views; lane p of chunk c owns output bin u0 = 64c + p.  With the slice's
//      u0 interpolation out[p] = wq*H[p+k] + wp*H[p+k+1], every lane reads ONE value v = H[p+k] per
//      (view, slice) and accumulates Q[p] += wq*v, W[p] += wp*v; at the end out[p] = Q[p] + W[p+1].
//      wq, wp, k are wave-uniform (scalar loads from the FP64-built table, two views at a time), so the
//      inner step is one ds_read_b32 and two v_fmac_f32 per bin per slice.  Chunks that cannot touch the
//      volume's footprint for the whole slice block are skipped with a uniform branch.
//
// Back projection (two kernels, no atomics):
//   back_h2_kernel, one workgroup per (u1 row, v1 view row, group of a): the exact transpose of step 2.
//      Lane = slice; the workgroup streams the v0 rows g[v1][:, u1, :] through LDS (by LDS DMA, one block
//      of rows ahead) and each lane accumulates its slice's column H2[s][a] = sum_v0 wp*S[a+m] +
//      wq*S[a+m+1] in registers (sliding window: one LDS read and two FMAs per (view, voxel)).  The slab
//      is written transposed through LDS to workspace[v1][u1][s][a].
//   back_z_kernel, one wave per (a chunk, slice, z chunk): the transpose of step 1, gathering the two
//      u1 rows of every v1 for each voxel, f[iz][s][a] += sum_v1 wb*H2[iz+m1] + wa*H2[iz+m1+1].  In PETL
//      mode the reference's voxel-centre u1 test (src/projectors.cu:284-289) is applied per voxel here.
#include <cmath>
#include <vector>
#include <algorithm>
#include <cstring>
#include <cstdlib>
#include "petl_common.h"
#include "proj_kernels.h"
#include "proj_fast.h"
#include "storage.h"
#include "mlem_batch.h"
#include "mlem_update_dev.h"

namespace petl {

namespace {

// views per wave of the forward (compile-time tuning constant; 2 * J * C accumulators per lane)
#ifndef PETL_FWD_J
#define PETL_FWD_J 7
#endif
constexpr int FWD_J = PETL_FWD_J;
// waves per forward workgroup (compile-time tuning constant); each wave takes FWD_J views
#ifndef PETL_FWD_WAVES
#define PETL_FWD_WAVES 8
#endif
constexpr int FWD_WAVES = PETL_FWD_WAVES;
constexpr int FWD_THREADS = 64 * FWD_WAVES;
constexpr int FWD_WPROW = 16;  // floats per (block, view) row of the forward's wp table: SB wp, mask, pad
constexpr int FWD_SBMAX = 16; // largest slice block of the forward (table row padding)
constexpr int FWD_HP = 321;   // LDS row pitch (floats) of the forward's z-interpolated rows (Na <= 128)
constexpr int FWD_HP_WIDE = 449;  // the same for the wide instances (128 < Na <= 256)
// row pitch of the forward instance for in-plane axes of up to NAM voxels
constexpr int fwd_hp(int NAM) { return NAM <= 128 ? FWD_HP : FWD_HP_WIDE; }
constexpr int BACK_RB = 8;    // v0 rows per LDS block in the back projector
constexpr int BACK_AG = 128;  // a-values per lane in the back projector
constexpr int BACK_SG = 256;  // slices per back projector workgroup (one per lane)
constexpr int BACK_TW = 32;   // transpose tile width of the back epilogue
constexpr int BACK_STW = 256; // dwords per staged row of 16-bit planogram bins (Nu0 <= 511)


__device__ __forceinline__ float coordf(int i, float T, float s)
{
#pragma clang fp contract(off)
    return float(i) * T + s;
}
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

struct FwdArgs {
    PlanGeom P;
    int Ns, Na, Nz;
    int64_t fs_s, fs_a, fs_z;
    int hpad, hrow;
    int psi90;
    int v1_first, v1_stride;
    int xrow;  // padded row length of the u0 table
    float f_scale, y_scale, g_inv_scale;  // storage scales (16-bit formats; 1 for fp32)
    const float* wts;                     // EPI_RESID / sum-of-squares weights (nullable)
    double* part;                         // sum-of-squares instance: one partial per workgroup
    // batch: blockIdx.z = b * nv1 + k runs v1 row v1_first + k * v1_stride of element b, whose volume and
    // planogram start b * f_bstride and b * g_bstride elements after f and g (nv1 = gridDim.z, b = 0
    // without a batch)
    int nv1;
    int64_t f_bstride, g_bstride;
};

// compile-time epilogue of a forward instance: the runtime epilogue argument (none / ratio), the
// least-squares residual (g = (resp * Af - y) * w), or the weighted sum of squares sum_b w_b (resp * Af)_b^2
// with no store (per-workgroup FP64 partials in a fixed order)
constexpr int EK_RUNTIME = 0, EK_RESID = 1, EK_SUMSQ = 2;

// =============================================================================================
// forward
// =============================================================================================
// z interpolation of R u1 rows x SB slices, software-pipelined: the global loads of block b+1 are issued
// before the block-b accumulation and written to LDS after it.  LDS layout per buffer: [R][SB][HP].
// The loaded volume values stay in their storage format until zstore, so the loads remain in flight.
template <int MAXE, int ST>
struct ZLoads {
    typename Stor<ST>::T a[MAXE], b[MAXE];
    float wa[MAXE], wb[MAXE];
};

// The thread's elements of a block, the same for every block: slice sl within the block, the column's
// offset in the volume (ia * fs_a), the z-row base iu1_0 + r - 1 of its u1 row, whether that row is live,
// and the LDS destination (r * SB + sl) * HP + ia, or -1 past the block's R * SB * Na values (those
// elements load clamped addresses and are never stored).  Offsets are 32-bit: the fast path takes
// volumes of fewer than 2^31 voxels.
template <int MAXE>
struct ZMap {
    int sl[MAXE], aoff[MAXE], zr[MAXE], dst[MAXE];
    bool live[MAXE];
};

template <int SB, int R, int MAXE, int HP>
__device__ __forceinline__ void zmap(ZMap<MAXE>& M, const FwdArgs& A, int iu1_0, int nrows, int tid)
{
    const int per_row = SB * A.Na;
    const int n = R * per_row;
#pragma unroll
    for (int e = 0; e < MAXE; e++) {
        const int idx0 = tid + e * FWD_THREADS;
        const int idx = min(idx0, n - 1);
        const int r = idx / per_row;
        const int rem = idx - r * per_row;
        int sl, ia;
        if (A.psi90) {
            sl = rem % SB;
            ia = rem / SB;
        } else {
            sl = rem / A.Na;
            ia = rem - sl * A.Na;
        }
        M.sl[e] = sl;
        M.aoff[e] = ia * int(A.fs_a);
        M.zr[e] = iu1_0 + r - 1;
        M.live[e] = r < nrows;
        M.dst[e] = (idx0 < n) ? (r * SB + sl) * HP + ia : -1;
    }
}

// zs: this v1 row's z table (one entry per slice) in LDS, so that the volume addresses do not wait for a
// global load
template <int SB, int R, int MAXE, int ST>
__device__ __forceinline__ void zload(ZLoads<MAXE, ST>& L, const ZMap<MAXE>& M, const FwdArgs& A,
                                      const typename Stor<ST>::T* __restrict__ f, const int4* zs, int is0)
{
    // branch-free: every address is clamped into the volume and out-of-range taps get a zero weight, so
    // the loads stay in flight until zstore consumes them
    const int fs_s = int(A.fs_s), fs_z = int(A.fs_z);
#pragma unroll
    for (int e = 0; e < MAXE; e++) {
        const int is = is0 + M.sl[e];
        const bool live = is < A.Ns && M.live[e];
        const int isc = min(is, A.Ns - 1);
        const int4 z = zs[isc];
        const int iza = M.zr[e] - z.x;
        const int za = min(max(iza, 0), A.Nz - 1), zb = min(max(iza + 1, 0), A.Nz - 1);
        const int o = isc * fs_s + M.aoff[e];
        L.a[e] = f[o + za * fs_z];
        L.b[e] = f[o + zb * fs_z];
        L.wa[e] = (live && iza >= 0 && iza < A.Nz) ? __int_as_float(z.y) : 0.0f;
        L.wb[e] = (live && iza + 1 >= 0 && iza + 1 < A.Nz) ? __int_as_float(z.z) : 0.0f;
    }
}

template <int MAXE, int ST>
__device__ __forceinline__ void zstore(const ZLoads<MAXE, ST>& L, const ZMap<MAXE>& M, float* hb, float f_scale)
{
#pragma unroll
    for (int e = 0; e < MAXE; e++)
        if (M.dst[e] >= 0)
            hb[M.dst[e]] = L.wa[e] * Stor<ST>::ld(L.a[e], f_scale) + L.wb[e] * Stor<ST>::ld(L.b[e], f_scale);
}

// Inputs that the forward epilogue reads per bin of one view row: data y (ratio and residual), weights
// (residual) and response.  Every lane loads at its bin clamped to the row, with no per-bin branch, so the
// loads of a row issue back to back; the epilogue loads row j+1 before it stores row j, because loads and
// stores retire through one in-order counter and a load issued after a store waits for that store.
template <int C>
struct EpiIn {
    float y[C], w[C], rs[C];
};

template <int C, int SG, int EK>
__device__ __forceinline__ void epi_load(EpiIn<C>& in, const FwdArgs& A, int epilogue,
                                         const typename Stor<SG>::T* __restrict__ yv,
                                         const float* __restrict__ resp, int64_t row, int Nu0, int lane)
{
    int64_t gi[C];
#pragma unroll
    for (int c = 0; c < C; c++)
        gi[c] = row + min(64 * c + lane, Nu0 - 1);
    if (EK == EK_RESID || (EK == EK_RUNTIME && epilogue == EPI_RATIO)) {
#pragma unroll
        for (int c = 0; c < C; c++)
            in.y[c] = Stor<SG>::ld(yv[gi[c]], A.y_scale);
    }
    if (EK != EK_RUNTIME && A.wts) {
#pragma unroll
        for (int c = 0; c < C; c++)
            in.w[c] = A.wts[gi[c]];
    }
    if (resp) {
#pragma unroll
        for (int c = 0; c < C; c++)
            in.rs[c] = resp[gi[c]];
    }
}

// J views per wave, C output chunks of 64 bins, SB-slice LDS blocks, R u1 rows per workgroup, storage SF of
// the volume and SG of the planogram and the data y.  EK: the compile-time epilogue (EK_*); the residual
// and sum-of-squares forms are separate instances so that the projection kernels keep no weight pointer
// live.
// The work of one workgroup: views grp * 56 ..., u1 rows iu1_0 ... iu1_0 + R - 1, view row iv1 of the
// planogram g (already offset to its batch element); part_idx: the sum-of-squares partial it writes.
template <int J, int C, int SB, int R, int SF, int SG, int EK, int NAM>
__device__ __forceinline__ void fwd_body(typename Stor<SG>::T* __restrict__ g, const typename Stor<SF>::T* __restrict__ f,
                                         const FwdArgs& A, const int2* __restrict__ ftab, const float* __restrict__ fwp,
                                         const int4* __restrict__ ztab, const float* __restrict__ fscale, int epilogue,
                                         const typename Stor<SG>::T* __restrict__ yv, const float* __restrict__ resp,
                                         const int grp, const int iu1_0, const int iv1, const int64_t part_idx)
{
    using TS = typename Stor<SG>::T;
    constexpr bool RES = EK == EK_RESID, SQ = EK == EK_SUMSQ;
    constexpr int HP = fwd_hp(NAM);
    static_assert((2 * R * SB) % 4 == 0, "the z table after the row buffers must be 16-byte aligned");
    extern __shared__ __align__(16) float hbuf[];
    __shared__ int s_lo, s_hi;
    __shared__ double s_sq[FWD_WAVES];
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const PlanGeom& P = A.P;
    const int Nv0 = P.Nv0, Nu0 = P.Nu0, Nu1 = P.Nu1;
    const int Ns = A.Ns, Na = A.Na, Nz = A.Nz;
    const int nrows = min(R, Nu1 - iu1_0);
    const int iv0_base = grp * (FWD_WAVES * J) + wave;

    const float v1 = coordf(iv1, P.Tv1, P.v1_0);
    bool rmask[R];
    bool all_masked = true;
#pragma unroll
    for (int r = 0; r < R; r++) {
        rmask[r] = (r >= nrows) || u1_masked(P, coordf(iu1_0 + r, P.Tu1, P.u1_0), v1);
        all_masked = all_masked && rmask[r];
    }
    if (all_masked) {
        // every u1 row is masked (src/projectors.cu:38-43); a ratio of a zero projection is zero too
#pragma unroll
        for (int r = 0; r < R; r++) {
            if (r >= nrows)
                continue;
#pragma unroll
            for (int j = 0; j < J; j++) {
                const int iv0 = iv0_base + FWD_WAVES * j;
                if (iv0 < Nv0) {
                    const int64_t row = ((int64_t(iv1) * Nv0 + iv0) * Nu1 + iu1_0 + r) * Nu0;
                    if (RES) {
                        // residual of a zero projection: (resp * 0 - y) * w
                        EpiIn<C> in;
                        epi_load<C, SG, EK>(in, A, epilogue, yv, resp, row, Nu0, lane);
#pragma unroll
                        for (int c = 0; c < C; c++) {
#pragma clang fp contract(off)
                            const int p = 64 * c + lane;
                            float out = 0.0f;
                            if (resp)
                                out = out * in.rs[c];
                            out = out - in.y[c];
                            if (A.wts)
                                out = out * in.w[c];
                            if (p < Nu0)
                                g[row + p] = Stor<SG>::st(out, A.g_inv_scale);
                        }
                        continue;
                    }
                    if (!SQ)
                        for (int p = lane; p < Nu0; p += 64)
                            g[row + p] = Stor<SG>::st(0.0f, 1.0f);
                }
            }
        }
        if (SQ && tid == 0)
            A.part[part_idx] = 0.0;
        return;
    }

    // slices whose z-interpolated rows can be non-zero (union over the R rows)
    if (tid == 0) {
        s_lo = Ns;
        s_hi = -1;
    }
    const int hbuf_elems = 2 * R * SB * HP;
    // LDS after the two row buffers: this v1 row's z table (16-byte aligned: hbuf_elems is a multiple of 4)
    int4* s_z = reinterpret_cast<int4*>(hbuf + hbuf_elems);
    for (int i = tid; i < hbuf_elems; i += FWD_THREADS)
        hbuf[i] = 0.0f;
    __syncthreads();
    for (int is = tid; is < Ns; is += FWD_THREADS) {
        const int4 e = ztab[int64_t(iv1) * Ns + is];
        s_z[is] = e;
        bool nz = false;
#pragma unroll
        for (int r = 0; r < R; r++) {
            const int iza = iu1_0 + r - e.x - 1;
            nz = nz || (r < nrows && ((iza >= 0 && iza < Nz && __int_as_float(e.y) > 0.0f) ||
                                      (iza + 1 >= 0 && iza + 1 < Nz && __int_as_float(e.z) > 0.0f)));
        }
        if (nz) {
            atomicMin(&s_lo, is);
            atomicMax(&s_hi, is);
        }
    }
    __syncthreads();
    const int is_lo = __builtin_amdgcn_readfirstlane(s_lo), is_hi = __builtin_amdgcn_readfirstlane(s_hi);

    // Q[p] = sum wq*v, W[p] = sum wp*v over the values v = H[s][p+k] read by lane p;  out[p] = Q[p] + W[p+1].
    // Both sums hold non-negative terms for a non-negative volume, so a bin whose exact value is zero comes
    // out as exactly zero (a ratio y/q downstream depends on that).
    float Q[R][J][C], W[R][J][C];
#pragma unroll
    for (int r = 0; r < R; r++)
#pragma unroll
        for (int j = 0; j < J; j++)
#pragma unroll
            for (int c = 0; c < C; c++) {
                Q[r][j][c] = 0.0f;
                W[r][j][c] = 0.0f;
            }

    if (is_lo <= is_hi) {
        // blocks start at multiples of SB so the table can hold each slice's LDS row offset
        const int is_first = (is_lo / SB) * SB;
        const int nblk = (is_hi - is_first) / SB + 1;
        constexpr int MAXE = (R * SB * NAM + FWD_THREADS - 1) / FWD_THREADS;  // Na <= NAM
        ZMap<MAXE> M;
        zmap<SB, R, MAXE, HP>(M, A, iu1_0, nrows, tid);
        ZLoads<MAXE, SF> L;
        zload<SB, R>(L, M, A, f, s_z, is_first);
        zstore(L, M, hbuf + A.hpad, A.f_scale);
        __syncthreads();
        for (int blk = 0; blk < nblk; blk++) {
            const bool more = blk + 1 < nblk;
            if (more)
                zload<SB, R>(L, M, A, f, s_z, is_first + (blk + 1) * SB);
            const int is0 = is_first + blk * SB;
            const float* hb = hbuf + (blk & 1) * R * SB * HP + A.hpad + lane;
            // this block's table rows: [block][view] x {SB entries {(is % SB) * HP + k, wq}} and
            // {SB x wp, chunk mask, pad}; the wave's views sit at a fixed stride, so every view's scalar
            // loads are immediate offsets from these two addresses
            const int bi = is0 / SB;
            const int2* xt_b = ftab + (int64_t(bi) * Nv0 + iv0_base) * SB;
            const float* xp_b = fwp + (int64_t(bi) * Nv0 + iv0_base) * FWD_WPROW;
            // views in pairs: both views' table rows are requested together, so a pair costs one wait
            // for scalar loads instead of two (a wait on scalar loads is a full drain of the counter
            // they share with the LDS reads)
#pragma unroll
            for (int jp = 0; jp < J; jp += 2) {
                int off[2][SB];
                float wq[2][SB], wp[2][SB];
                int cmask[2];
#pragma unroll
                for (int q = 0; q < 2; q++) {
                    const int j = jp + q;
                    const int iv0 = iv0_base + FWD_WAVES * j;
                    cmask[q] = 0;
                    if (j < J && iv0 < Nv0) {
                        const int2* xt = xt_b + FWD_WAVES * SB * j;
                        const float* xp = xp_b + FWD_WAVES * FWD_WPROW * j;
#pragma unroll
                        for (int s = 0; s < SB; s++) {
                            const int2 e = xt[s];
                            off[q][s] = e.x;
                            wq[q][s] = __int_as_float(e.y);
                            wp[q][s] = xp[s];
                        }
                        // bit c: some slice of the block reaches output chunk c
                        cmask[q] = __float_as_int(xp[SB]);
                    }
                }
#pragma unroll
                for (int q = 0; q < 2; q++) {
                    const int j = jp + q;
                    if (j >= J)
                        continue;
                    const float* hp[SB];
#pragma unroll
                    for (int s = 0; s < SB; s++)
                        hp[s] = hb + off[q][s];
#pragma unroll
                    for (int c = 0; c < C; c++) {
                        if (cmask[q] & (1 << c)) {
#pragma unroll
                            for (int r = 0; r < R; r++) {
                                float v[SB];
#pragma unroll
                                for (int s = 0; s < SB; s++)
                                    v[s] = hp[s][r * SB * HP + 64 * c];
#pragma unroll
                                for (int s = 0; s < SB; s++) {
                                    W[r][j][c] += wp[q][s] * v[s];
                                    Q[r][j][c] += wq[q][s] * v[s];
                                }
                            }
                        }
                    }
                }
            }
            if (more)
                zstore(L, M, hbuf + ((blk + 1) & 1) * R * SB * HP + A.hpad, A.f_scale);
            __syncthreads();
        }
    }

    // epilogue: out[p] = (Q[p] + W[p+1]) * scale, masks, optional response / ratio / residual / square sum
    double sq = 0.0;
#pragma unroll
    for (int r = 0; r < R; r++) {
        if (r >= nrows)
            continue;
        const int64_t row_r = int64_t(iv1) * Nv0 * Nu1 + iu1_0 + r;
        EpiIn<C> in[2];
        epi_load<C, SG, EK>(in[0], A, epilogue, yv, resp, (row_r + int64_t(min(iv0_base, Nv0 - 1)) * Nu1) * Nu0,
                            Nu0, lane);
#pragma unroll
        for (int j = 0; j < J; j++) {
            if (j + 1 < J) {
                const int iv0n = min(iv0_base + FWD_WAVES * (j + 1), Nv0 - 1);
                epi_load<C, SG, EK>(in[(j + 1) & 1], A, epilogue, yv, resp, (row_r + int64_t(iv0n) * Nu1) * Nu0,
                                    Nu0, lane);
            }
            const EpiIn<C>& cur = in[j & 1];
            const int iv0 = iv0_base + FWD_WAVES * j;
            if (iv0 < Nv0) {
                const float sc = rmask[r] ? 0.0f : fscale[int64_t(iv1) * Nv0 + iv0];
                const float v0 = coordf(iv0, P.Tv0, P.v0_0);
                TS* gr = g + (row_r + int64_t(iv0) * Nu1) * Nu0;
#pragma unroll
                for (int c = 0; c < C; c++) {
#pragma clang fp contract(off)
                    float wn = __shfl_down(W[r][j][c], 1);
                    const int cn = (c + 1 < C) ? c + 1 : c;
                    const float w0 = (c + 1 < C) ? __shfl(W[r][j][cn], 0) : 0.0f;
                    if (lane == 63)
                        wn = w0;
                    const int p = 64 * c + lane;
                    float out = (Q[r][j][c] + wn) * sc;
                    if (u0_masked(P, coordf(p, P.Tu0, P.u0_0), v0))
                        out = 0.0f;
                    if (resp)
                        out = out * cur.rs[c];
                    if (SQ) {
                        // the weighted inner product's term: float(v * v) * w, summed in FP64
                        float t = out * out;
                        if (A.wts)
                            t = t * cur.w[c];
                        if (p < Nu0)
                            sq += double(t);
                        continue;
                    }
                    if (RES) {
                        out = out - cur.y[c];
                        if (A.wts)
                            out = out * cur.w[c];
                    } else if (epilogue == EPI_RATIO) {
                        out = (out == 0.0f) ? 0.0f : cur.y[c] / out;
                        if (resp)
                            out = out * cur.rs[c];
                    }
                    if (p < Nu0)
                        gr[p] = Stor<SG>::st(out, A.g_inv_scale);
                }
            }
        }
    }
    if (SQ) {
        // fixed-order workgroup sum: butterfly within each wave, then the waves in order
#pragma unroll
        for (int o = 32; o >= 1; o >>= 1)
            sq += __shfl_xor(sq, o);
        if (lane == 0)
            s_sq[wave] = sq;
        __syncthreads();
        if (tid == 0) {
            double t = 0.0;
            for (int w = 0; w < FWD_WAVES; w++)
                t += s_sq[w];
            A.part[part_idx] = t;
        }
    }
}

template <int J, int C, int SB, int R, int SF, int SG, int EK = EK_RUNTIME, int NAM = 128>
__global__ void __launch_bounds__(FWD_THREADS) fwd_fast_kernel(typename Stor<SG>::T* __restrict__ g,
                                                               const typename Stor<SF>::T* __restrict__ f,
                                                               FwdArgs A, const int2* __restrict__ ftab,
                                                               const float* __restrict__ fwp,
                                                               const int4* __restrict__ ztab,
                                                               const float* __restrict__ fscale, int epilogue,
                                                               const typename Stor<SG>::T* __restrict__ yv,
                                                               const float* __restrict__ resp)
{
    // one workgroup per (group of 56 views, u1 row, v1 row of a batch element); a renumbering that gives
    // each XCD a contiguous range of u1 rows (so resident workgroups share a band of the volume in L2)
    // measured 30 % slower
    const int bz = blockIdx.z, bel = bz / A.nv1;
    fwd_body<J, C, SB, R, SF, SG, EK, NAM>(g + bel * A.g_bstride, f + bel * A.f_bstride, A, ftab, fwp, ztab, fscale,
                                           epilogue, yv, resp, blockIdx.x, blockIdx.y * R,
                                           A.v1_first + (bz - bel * A.nv1) * A.v1_stride,
                                           (int64_t(blockIdx.z) * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x);
}

// =============================================================================================
// back projection, stage 1: u0 gather of every v0 row into per-slice columns
// =============================================================================================
struct BackArgs {
    PlanGeom P;
    VolGeom V;
    int Ns, Na, Nz;
    int64_t fs_s, fs_a, fs_z;
    int spad_lo, srow;
    int mode;
    int psi90;
    int v1_first, v1_stride;
    int sb;    // slice block of the forward table layout
    int hrow;  // the forward's LDS row pitch, which the table's offsets include
    float g_scale;  // storage scale of the planogram (16-bit formats; 1 for fp32)
    // batch: back_h2's blockIdx.z = b * nv1 + k (v1 row v1_first + k * v1_stride of element b) and
    // back_z's blockIdx.z = b * nzc + z chunk; element b's planogram, workspace and volume start b *
    // g_bstride, b * h_bstride and b * f_bstride elements after g, h2 and f
    int nv1, nzc;
    int64_t g_bstride, h_bstride, f_bstride;
    int nagroups;  // a groups per slice group (blockIdx.x of back_h2 = slice group x nagroups + a group)
};

// DMA: the next block of v0 rows goes from HBM straight into LDS (global_load_lds, one 64-dword row
// segment per wave instruction) while the current block accumulates, and each view's scale is folded
// into its two interpolation weights.  fp32 rows land in the other row buffer; 16-bit rows (an even
// number of bins) land as raw dwords in a staging array and one pass per block widens them into the row
// buffer.  Otherwise the block is read through registers and scaled (and, in matched mode, masked per
// bin) on its way into LDS.
// The work of one workgroup: a group ag of AG in-plane columns and slice group sgr (BACK_SG slices, one per
// lane) of u1 row iu1 and view row iv1; g and h2 already offset to their batch element.
template <int AG, int ST, bool DMA, bool MASKED>
__device__ __forceinline__ void back_h2_body(const typename Stor<ST>::T* __restrict__ g, float* __restrict__ h2,
                                             const BackArgs& A, const int2* __restrict__ xtab, int xrow,
                                             const float* __restrict__ scale, const int2* __restrict__ rows,
                                             const int ag, const int sgr, const int iu1, const int iv1)
{
    extern __shared__ float sbuf[];
    const int tid = threadIdx.x;
    const int nthr = blockDim.x;
    const PlanGeom& P = A.P;
    const int Nv0 = P.Nv0, Nu0 = P.Nu0, Nu1 = P.Nu1, Ns = A.Ns, Na = A.Na;
    const int2 rr = rows[iv1];
    if (iu1 < rr.x || iu1 > rr.y)
        return;
    const int a0 = ag * AG;
    const int s0 = sgr * BACK_SG, ns = min(BACK_SG, Ns - s0);  // this workgroup's slices [s0, s0 + ns)
    const int is = min(s0 + tid, Ns - 1);

    float acc[AG];
#pragma unroll
    for (int a = 0; a < AG; a++)
        acc[a] = 0.0f;

    const int buf_elems = 2 * BACK_RB * A.srow;
    for (int i = tid; i < buf_elems; i += nthr)
        sbuf[i] = 0.0f;
    __syncthreads();

    // DMA form: bins [0, Nu0) of each row come from the planogram; the LDS positions around them keep the
    // zeros written above (no lane writes there), and rows past Nv0 are never read.  In matched mode the
    // lanes of bins outside the view's u0 mask write a zero instead (they may hold a value from the view
    // this buffer row had two blocks earlier); the DMA and the zeros reach disjoint positions.
    const int nseg = (Nu0 + 63) / 64;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6), nwaves = nthr >> 6, lane = tid & 63;
    const int64_t g_row0 = (int64_t(iv1) * Nv0 * Nu1 + iu1) * Nu0;  // bin 0 of view 0 in this (v1, u1) row
    // rows go to waves round robin (no index division in the issue loop), each row in 64-bin segments
    auto dma = [&](int blk, int b) {
        for (int r = wave; r < BACK_RB; r += nwaves) {
            const int iv0 = blk * BACK_RB + r;
            if (iv0 >= Nv0)
                break;
            float* row = sbuf + b * BACK_RB * A.srow + A.spad_lo + r * A.srow;
            const typename Stor<ST>::T* src = g + g_row0 + int64_t(iv0) * Nu1 * Nu0;
            for (int seg = 0; seg < nseg; seg++) {
                const int p = seg * 64 + lane;
                if (p < Nu0) {
                    if (MASKED && u0_masked(P, coordf(p, P.Tu0, P.u0_0), coordf(iv0, P.Tv0, P.v0_0)))
                        row[p] = 0.0f;
                    else
                        __builtin_amdgcn_global_load_lds(reinterpret_cast<const void*>(src + p),
                                                         (__attribute__((address_space(3))) void*)(row + seg * 64),
                                                         4, 0, 0);
                }
            }
        }
    };

    // 16-bit DMA form: raw row dwords (two bins each) into the staging array, widened by convert()
    const int nd = Nu0 / 2;  // dwords per row (Nu0 is even on this path)
    auto dma16 = [&](int blk, uint32_t* stage) {
        for (int r = wave; r < BACK_RB; r += nwaves) {
            const int iv0 = blk * BACK_RB + r;
            if (iv0 >= Nv0)
                break;
            const uint32_t* src = reinterpret_cast<const uint32_t*>(g + g_row0 + int64_t(iv0) * Nu1 * Nu0);
            for (int seg = 0; seg * 64 < nd; seg++) {
                const int d = seg * 64 + lane;
                if (d < nd)
                    __builtin_amdgcn_global_load_lds(
                        reinterpret_cast<const void*>(src + d),
                        (__attribute__((address_space(3))) void*)(stage + r * BACK_STW + seg * 64), 4, 0, 0);
            }
        }
    };
    auto convert = [&](int blk, int b, const uint32_t* stage) {
        using TS16 = typename Stor<ST>::T;
        for (int r = 0; r < BACK_RB; r++) {
            const int iv0 = blk * BACK_RB + r;
            if (iv0 >= Nv0)
                break;
            float* row = sbuf + b * BACK_RB * A.srow + A.spad_lo + r * A.srow;
            for (int d = tid; d < nd; d += nthr) {
                const uint32_t w = stage[r * BACK_STW + d];
#pragma unroll
                for (int k = 0; k < 2; k++) {
                    const uint16_t bits = uint16_t(w >> (16 * k));
                    TS16 raw;
                    __builtin_memcpy(&raw, &bits, 2);
                    const int p = 2 * d + k;
                    float v = Stor<ST>::ld(raw, A.g_scale);
                    if (MASKED && u0_masked(P, coordf(p, P.Tu0, P.u0_0), coordf(iv0, P.Tv0, P.v0_0)))
                        v = 0.0f;
                    row[p] = v;
                }
            }
        }
    };

    // register form: elements idx = tid + k * nthr of the block's BACK_RB * Nu0 values, four at a time:
    // the four loads (clamped in bounds) issue together, then the scaled, masked values go to LDS.
    // Element -> (row, bin) by a float reciprocal: (idx + 0.5) / Nu0 lies at least 0.5 / Nu0 from an
    // integer, far above the rounding error of the product.
    const float inv_nu0 = 1.0f / float(Nu0);
    auto load = [&](int blk, int b) {
        float* sb = sbuf + b * BACK_RB * A.srow + A.spad_lo;
        const int n = BACK_RB * Nu0;
        for (int idx0 = tid; idx0 < n; idx0 += 4 * nthr) {
            typename Stor<ST>::T raw[4];
            float sc[4];
#pragma unroll
            for (int u = 0; u < 4; u++) {
                const int idx = min(idx0 + u * nthr, n - 1);
                const int r = int((float(idx) + 0.5f) * inv_nu0), p = idx - r * Nu0;
                const int iv0 = min(blk * BACK_RB + r, Nv0 - 1);
                raw[u] = g[g_row0 + int64_t(iv0) * Nu1 * Nu0 + p];
                sc[u] = scale[int64_t(iv1) * Nv0 + iv0];
            }
#pragma unroll
            for (int u = 0; u < 4; u++) {
                const int idx = idx0 + u * nthr;
                if (idx < n) {
                    const int r = int((float(idx) + 0.5f) * inv_nu0), p = idx - r * Nu0;
                    const int iv0 = blk * BACK_RB + r;
                    float val = 0.0f;
                    if (iv0 < Nv0) {
                        val = Stor<ST>::ld(raw[u], A.g_scale) * sc[u];
                        if (A.mode == BACK_MATCHED &&
                            u0_masked(P, coordf(p, P.Tu0, P.u0_0), coordf(iv0, P.Tv0, P.v0_0)))
                            val = 0.0f;
                    }
                    sb[r * A.srow + p] = val;
                }
            }
        }
    };

    const int nblk = (Nv0 + BACK_RB - 1) / BACK_RB;
    // staging of the 16-bit DMA form: its own array, so that the compiler sees that the DMA into it does
    // not alias the row buffer the accumulation reads
    constexpr bool STAGED = DMA && ST != STORE_F32;
    __shared__ __align__(16) uint32_t s_stage[STAGED ? BACK_RB * BACK_STW : 1];
    // Every wave reads rows that other waves' LDS DMA wrote: each wave waits for its own DMA (vmcnt(0))
    // before the barrier, so that after the barrier every wave's DMA has landed.
    if (STAGED) {
        dma16(0, s_stage);
        __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0)
        __syncthreads();
        convert(0, 0, s_stage);
    } else if (DMA) {
        dma(0, 0);
        __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0)
    } else {
        load(0, 0);
    }
    __syncthreads();
    for (int blk = 0; blk < nblk; blk++) {
        const float* sb = sbuf + (blk & 1) * BACK_RB * A.srow + A.spad_lo + a0;
        const int row_off = (is % A.sb) * A.hrow;
        if (DMA) {
            // this block's table entries first, waited for before the next block's DMA is issued: loads
            // and LDS DMA retire through one in-order counter, and the accumulation must not wait for
            // the DMA; it then runs on registers and LDS only
            int2 eb[BACK_RB];
            float scb[BACK_RB];  // the rows' view scales, requested together (one scalar-load wait a block)
#pragma unroll
            for (int r = 0; r < BACK_RB; r++) {
                eb[r] = xtab[int64_t(min(blk * BACK_RB + r, Nv0 - 1)) * xrow + is];
                scb[r] = scale[int64_t(iv1) * Nv0 + min(blk * BACK_RB + r, Nv0 - 1)];
            }
            __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0), expcnt and lgkmcnt left open (gfx9 encoding)
            if (blk + 1 < nblk) {
                if (STAGED)
                    dma16(blk + 1, s_stage);
                else
                    dma(blk + 1, (blk + 1) & 1);
            }
#pragma unroll
            for (int r = 0; r < BACK_RB; r++) {
                const int iv0 = blk * BACK_RB + r;
                if (iv0 < Nv0) {
                    const int m = -(eb[r].x - row_off) - 1;
                    const float sc = scb[r];
                    const float wq = __int_as_float(eb[r].y) * sc, wp = (1.0f - __int_as_float(eb[r].y)) * sc;
                    const float* s = sb + r * A.srow + m;
                    float xp = s[0];
#pragma unroll
                    for (int a = 0; a < AG; a++) {
                        const float x = s[a + 1];
                        acc[a] += wp * xp;
                        acc[a] += wq * x;
                        xp = x;
                    }
                }
            }
            // the next block's DMA (issued above) must have landed for every wave after the barrier
            __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0)
            __syncthreads();
            if (STAGED && blk + 1 < nblk) {
                convert(blk + 1, (blk + 1) & 1, s_stage);
                __syncthreads();
            }
            continue;
        }
        if (blk + 1 < nblk)
            load(blk + 1, (blk + 1) & 1);
        // the next view's table entry is loaded while the current view's window runs
        int2 e_next = xtab[int64_t(min(blk * BACK_RB, Nv0 - 1)) * xrow + is];
#pragma unroll 1
        for (int r = 0; r < BACK_RB; r++) {
            const int iv0 = blk * BACK_RB + r;
            const int2 e = e_next;
            if (r + 1 < BACK_RB)
                e_next = xtab[int64_t(min(iv0 + 1, Nv0 - 1)) * xrow + is];
            if (iv0 < Nv0) {
                const int m = -(e.x - row_off) - 1;  // the table's .x carries the forward's LDS row offset
                const float wq = __int_as_float(e.y), wp = 1.0f - wq;
                const float* s = sb + r * A.srow + m;
                float xp = s[0];
#pragma unroll
                for (int a = 0; a < AG; a++) {
                    const float x = s[a + 1];
                    acc[a] += wp * xp;
                    acc[a] += wq * x;
                    xp = x;
                }
            }
        }
        __syncthreads();
    }

    // transposed write: workspace[iv1][iu1][s][a]
    float* tile = sbuf;
    float* out = h2 + ((int64_t(iv1) * Nu1 + iu1) * Ns + s0) * int64_t(Na);
#pragma unroll
    for (int q = 0; q < AG / BACK_TW; q++) {
        if (tid < ns) {
#pragma unroll
            for (int t = 0; t < BACK_TW; t++)
                tile[tid * (BACK_TW + 1) + t] = acc[q * BACK_TW + t];
        }
        __syncthreads();
        for (int idx = tid; idx < ns * BACK_TW; idx += nthr) {
            const int s = idx / BACK_TW, t = idx - s * BACK_TW;
            const int a = a0 + q * BACK_TW + t;
            if (a < Na)
                out[int64_t(s) * Na + a] = tile[s * (BACK_TW + 1) + t];
        }
        __syncthreads();
    }
}

template <int AG, int ST, bool DMA = false, bool MASKED = false>
__global__ void __launch_bounds__(256) back_h2_kernel(const typename Stor<ST>::T* __restrict__ g,
                                                      float* __restrict__ h2, BackArgs A,
                                                      const int2* __restrict__ xtab, int xrow,
                                                      const float* __restrict__ scale,
                                                      const int2* __restrict__ rows)
{
    // blockIdx.x = slice group x a group; blockIdx.z = batch element x view row
    const int bz = blockIdx.z, bel = bz / A.nv1;
    const int sgr = blockIdx.x / A.nagroups, ag = blockIdx.x - sgr * A.nagroups;
    back_h2_body<AG, ST, DMA, MASKED>(g + bel * A.g_bstride, h2 + bel * A.h_bstride, A, xtab, xrow, scale, rows, ag,
                                      sgr, blockIdx.y, A.v1_first + (bz - bel * A.nv1) * A.v1_stride);
}

// =============================================================================================
// back projection, stage 2: u1 gather over all v1 rows, per voxel
// =============================================================================================
// The work of one wave: in-plane columns 64 * ac ... of slice is, heights z0 ... z0 + ZC - 1; h2 and f
// already offset to their batch element.  overwrite: f = sum instead of f += sum (the value is the same
// when f holds zeros: 0 + x = x for the non-negative sums of a back projection of non-negative data, and
// for any sum that is not -0).
template <int ZC>
__device__ __forceinline__ void back_z_body(const float* __restrict__ h2, float* __restrict__ f, const BackArgs& A,
                                            const int4* __restrict__ ztab, const int2* __restrict__ rows,
                                            const int lane, const int ac, const int is, const int z0,
                                            const bool overwrite)
{
#pragma clang fp contract(off)
    const int ia = ac * 64 + lane;
    const PlanGeom& P = A.P;
    const VolGeom& V = A.V;
    const int Ns = A.Ns, Na = A.Na, Nz = A.Nz, Nu1 = P.Nu1;
    const int iac = min(ia, Na - 1);
    const bool petl = (A.mode == BACK_PETL);

    // voxel-centre coordinates for the PETL-mode test
    const int ix = A.psi90 ? is : iac;
    const int iy = A.psi90 ? iac : is;
    const float xc = coordf(ix, V.Tx, V.x0), yc = coordf(iy, V.Ty, V.y0);
    const float tv = yc * P.cos_psi - xc * P.sin_psi;

    float acc[ZC];
#pragma unroll
    for (int i = 0; i < ZC; i++)
        acc[i] = 0.0f;

    const int64_t slab = int64_t(Ns) * Na;
    const float* hb = h2 + int64_t(is) * Na + iac;
    for (int iv1 = A.v1_first; iv1 < P.Nv1; iv1 += A.v1_stride) {
        const int4 e = ztab[int64_t(iv1) * Ns + is];
        const int2 rr = rows[iv1];
        const int r0 = z0 + e.x;
        const float wa = __int_as_float(e.y), wb = __int_as_float(e.z);
        const float* hv = hb + int64_t(iv1) * Nu1 * slab;
        float x[ZC + 1];
#pragma unroll
        for (int i = 0; i <= ZC; i++) {
            const int r = r0 + i;
            x[i] = (r >= rr.x && r <= rr.y) ? hv[int64_t(r) * slab] : 0.0f;
        }
        if (petl) {
            const float v1 = coordf(iv1, P.Tv1, P.v1_0);
            const float thr = P.H - P.R * fabsf(v1);
            const float v1t = v1 * tv;
#pragma unroll
            for (int i = 0; i < ZC; i++) {
                const float zc = coordf(z0 + i, V.Tz, V.z0);
                const float u1 = zc + v1t;
                const float c = wb * x[i] + wa * x[i + 1];
                acc[i] += (fabsf(u1) - 0.5f * P.Tu1 <= thr) ? c : 0.0f;
            }
        } else {
#pragma unroll
            for (int i = 0; i < ZC; i++)
                acc[i] += wb * x[i] + wa * x[i + 1];
        }
    }
    if (ia < Na) {
        float* fp = f + int64_t(is) * A.fs_s + int64_t(ia) * A.fs_a;
#pragma unroll
        for (int i = 0; i < ZC; i++)
            if (z0 + i < Nz) {
                if (overwrite)
                    fp[int64_t(z0 + i) * A.fs_z] = acc[i];
                else
                    fp[int64_t(z0 + i) * A.fs_z] += acc[i];
            }
    }
}

template <int ZC>
__global__ void __launch_bounds__(64) back_z_kernel(const float* __restrict__ h2, float* __restrict__ f, BackArgs A,
                                                    const int4* __restrict__ ztab, const int2* __restrict__ rows)
{
    const int bel = blockIdx.z / A.nzc;
    back_z_body<ZC>(h2 + bel * A.h_bstride, f + bel * A.f_bstride, A, ztab, rows, threadIdx.x, blockIdx.x, blockIdx.y,
                    (blockIdx.z - bel * A.nzc) * ZC, false);
}

// ---------------------------------------------------------------------------------------------
// host-side kernel selection
// ---------------------------------------------------------------------------------------------
// z rows per back_z wave: long chunks re-read less of h2 ((ZC+1)/ZC loads per voxel), but the grid
// has only ceil(Na/64)*Ns*ceil(Nz/ZC) waves, and each wave walks all v1 rows serially.  Chunks
// shrink until at least BACK_Z_WAVES waves are in flight (16 per CU).  Measured per launch on
// MI355X: 128^3 volume, 101 v1 rows: 245 / 174 / 155 / 156 us at ZC 32 / 16 / 8 / 4; 64^3, 63 rows:
// 128 / 76 / 53 / 41 us.
constexpr int BACK_Z_WAVES = 4096;
static int back_z_chunk(int Na, int Ns, int Nz)
{
    const int64_t cols = int64_t((Na + 63) / 64) * Ns;
    int zc = 32;
    while (zc > 4 && cols * ((Nz + zc - 1) / zc) < BACK_Z_WAVES)
        zc /= 2;
    return zc;
}

template <int SF, int SG>
using fwd_fn = void (*)(typename Stor<SG>::T*, const typename Stor<SF>::T*, FwdArgs, const int2*, const float*,
                        const int4*, const float*, int, const typename Stor<SG>::T*, const float*);

// FWD_J views per wave, one u1 row per workgroup, slice blocks of 8 (the fastest shape measured at 128^3;
// two rows per workgroup with 4 or 2 views per wave ran 17-37 % slower)
template <int SF, int SG, int NAM>
fwd_fn<SF, SG> pick_fwd_na(int C, int ek)
{
#define PETL_FWD_BY_C(EKV)                                                \
    switch (C) {                                                          \
    case 1: return fwd_fast_kernel<FWD_J, 1, 8, 1, SF, SG, EKV, NAM>;     \
    case 2: return fwd_fast_kernel<FWD_J, 2, 8, 1, SF, SG, EKV, NAM>;     \
    case 3: return fwd_fast_kernel<FWD_J, 3, 8, 1, SF, SG, EKV, NAM>;     \
    case 4: return fwd_fast_kernel<FWD_J, 4, 8, 1, SF, SG, EKV, NAM>;     \
    case 5: return fwd_fast_kernel<FWD_J, 5, 8, 1, SF, SG, EKV, NAM>;     \
    case 6: return fwd_fast_kernel<FWD_J, 6, 8, 1, SF, SG, EKV, NAM>;     \
    case 7: return fwd_fast_kernel<FWD_J, 7, 8, 1, SF, SG, EKV, NAM>;     \
    case 8: return fwd_fast_kernel<FWD_J, 8, 8, 1, SF, SG, EKV, NAM>;     \
    default: return nullptr;                                              \
    }
    if constexpr (SF == STORE_F32 && SG == STORE_F32) {
        if (ek == EK_RESID)
            PETL_FWD_BY_C(EK_RESID)
        if (ek == EK_SUMSQ)
            PETL_FWD_BY_C(EK_SUMSQ)
    } else if (ek != EK_RUNTIME) {
        return nullptr;  // the residual and sum-of-squares epilogues are fp32 only
    }
    PETL_FWD_BY_C(EK_RUNTIME)
#undef PETL_FWD_BY_C
}

// FWD_J views per wave, one u1 row per workgroup, slice blocks of 8 (the fastest shape measured at 128^3;
// two rows per workgroup with 4 or 2 views per wave ran 17-37 % slower); in-plane axes of up to 128
// voxels, or the wide instances (longer LDS rows, more z-stage elements per thread) up to 256
template <int SF, int SG>
fwd_fn<SF, SG> pick_fwd(int C, int ek, bool wide)
{
    return wide ? pick_fwd_na<SF, SG, 256>(C, ek) : pick_fwd_na<SF, SG, 128>(C, ek);
}

inline float fbits(int i)
{
    float f;
    memcpy(&f, &i, 4);
    return f;
}
inline int ibits(float f)
{
    int i;
    memcpy(&i, &f, 4);
    return i;
}

inline float coordh(int i, float T, float s)
{
    volatile float a = float(i) * T;
    return a + s;
}

}  // namespace

FastPlan::~FastPlan()
{
    if (dev_block)
        (void)hipFree(dev_block);
}

// Host check of the geometry conditions of the fast kernels (the LDS fits checked in make_fast_plan aside):
// psi = 0 or 90 degrees, voxel pitch = crystal pitch, every view on the same slice branch, at most 8 chunks
// of 64 u0 bins, in-plane axes of at most 256 and slice axes of at most 1024 voxels.  psi90 receives 1 for
// x slices.
static bool fast_geometry(const PlanGeom& P, const VolGeom& V, int& psi90)
{
    if (P.sin_psi == 0.0f && P.cos_psi == 1.0f)
        psi90 = 0;
    else if (P.sin_psi == 1.0f && std::fabs(P.cos_psi) < 1.0e-6f)
        psi90 = 1;
    else
        return false;
    if (!(V.Tx == P.Tu0 && V.Ty == P.Tu0 && V.Tz == P.Tu1))
        return false;
    // every view must take the same slice branch as this kernel assumes (src/projectors.cu:77)
    for (int iv0 = 0; iv0 < P.Nv0; iv0++) {
        const float v0 = coordh(iv0, P.Tv0, P.v0_0);
        volatile float t1 = v0 * P.cos_psi;
        volatile float t2 = v0 * P.sin_psi;
        const float rx = -(t1 + P.sin_psi), ry = -(t2 - P.cos_psi);
        const bool xs = std::fabs(rx) >= std::fabs(ry);
        if (xs != (psi90 == 1))
            return false;
    }
    const int Ns = psi90 ? V.Nx : V.Ny, Na = psi90 ? V.Ny : V.Nx;
    return (P.Nu0 + 1 + 63) / 64 <= 8 && Ns <= 1024 && Na <= 256 &&
           int64_t(V.Nx) * V.Ny * V.Nz < (int64_t(1) << 31);
}

size_t fast_workspace_estimate(const PlanGeom& P, const VolGeom& V)
{
    int psi90;
    if (!fast_geometry(P, V, psi90))
        return 0;
    return size_t(P.Nv1) * P.Nu1 * size_t(V.Nx) * V.Ny * sizeof(float);
}

std::unique_ptr<FastPlan> make_fast_plan(const PlanGeom& P, const VolGeom& V)
{
    int psi90;
    if (!fast_geometry(P, V, psi90))
        return nullptr;
    auto fp = std::make_unique<FastPlan>();
    fp->P = P;
    fp->V = V;
    fp->psi90 = psi90;
    fp->Nz = V.Nz;
    fp->fs_z = int64_t(V.Nx) * V.Ny;
    if (!psi90) {
        fp->Ns = V.Ny;
        fp->Na = V.Nx;
        fp->fs_s = V.Nx;
        fp->fs_a = 1;
    } else {
        fp->Ns = V.Nx;
        fp->Na = V.Ny;
        fp->fs_s = 1;
        fp->fs_a = V.Nx;
    }
    const int Ns = fp->Ns, Na = fp->Na;
    // forward shape: C chunks of 64 lanes cover the Nu0+1 positions the P/Q split needs
    fp->C = (P.Nu0 + 1 + 63) / 64;
    if (fp->C > 8 || Ns > 1024)
        return nullptr;
    // u1 rows per workgroup and views per wave: R rows share each (view, slice) table entry and LDS
    // address; each lane holds 2*R*J*C accumulators (56 at 228 u0 bins)
    fp->R = 1;
    fp->J = FWD_J;
    fp->ngroups = (P.Nv0 + FWD_WAVES * fp->J - 1) / (FWD_WAVES * fp->J);
    fp->SB = 8;
    fp->nagroups = (Na + BACK_AG - 1) / BACK_AG;
    fp->nsgroups = (Ns + BACK_SG - 1) / BACK_SG;
    if (Na > 256)  // the forward's z-interpolated rows hold the whole in-plane axis
        return nullptr;
    fp->hrow = (Na <= 128) ? FWD_HP : FWD_HP_WIDE;
    if (int64_t(V.Nx) * V.Ny * V.Nz >= (int64_t(1) << 31))  // the forward's 32-bit voxel offsets
        return nullptr;

    // ---- FP64 tables ----
    const double x0 = V.x0, y0 = V.y0, z0 = V.z0;
    const double cs = P.cos_psi;
    const int xrow = Ns + FWD_SBMAX;  // padded so a block of slice entries never leaves the row
    fp->xrow = xrow;
    std::vector<int2> xtab(size_t(P.Nv0) * xrow);
    std::vector<float> xwp(size_t(P.Nv0) * xrow);
    std::vector<int4> ztab(size_t(P.Nv1) * Ns);
    std::vector<float> fscale(size_t(P.Nv1) * P.Nv0), pscale(size_t(P.Nv1) * P.Nv0);
    std::vector<int2> rows(2 * size_t(P.Nv1));
    int kspread = 0, mmin = 1 << 30, mmax = -(1 << 30);
    for (int iv0 = 0; iv0 < P.Nv0; iv0++) {
        const double v0 = coordh(iv0, P.Tv0, P.v0_0);
        double alpha, beta;
        if (!psi90) {
            alpha = (x0 + y0 * v0 - double(P.u0_0)) / double(P.Tu0);
            beta = double(V.Ty) * v0 / double(P.Tu0);
        } else {
            const double kap = -(v0 - cs) / (1.0 + v0 * cs);
            alpha = (y0 + x0 * kap - double(P.u0_0)) / double(P.Tu0);
            beta = double(V.Tx) * kap / double(P.Tu0);
        }
        for (int is = 0; is < Ns; is++) {
            const double mu = alpha + double(is) * beta;
            const double m = std::floor(mu);
            const double phi = mu - m;
            const int mi = int(m);
            xtab[size_t(iv0) * xrow + is] = make_int2(-mi - 1, ibits(float(phi)));  // k for now
            mmin = std::min(mmin, mi);
            mmax = std::max(mmax, mi);
        }
        for (int is = Ns; is < xrow; is++)
            xtab[size_t(iv0) * xrow + is] = xtab[size_t(iv0) * xrow + Ns - 1];
        // wp = 1 - wq in float, the value the back projector forms from wq, so matched mode stays an exact
        // transpose of the forward
        for (int is = 0; is < xrow; is++)
            xwp[size_t(iv0) * xrow + is] = 1.0f - fbits(xtab[size_t(iv0) * xrow + is].y);
        for (int is = 0; is < xrow; is++) {
            const int lo = std::max(0, is - fp->SB + 1);
            const int kd = std::abs(xtab[size_t(iv0) * xrow + is].x - xtab[size_t(iv0) * xrow + lo].x);
            kspread = std::max(kspread, kd);
        }
        // the forward reads entry .x as the slice's LDS offset within an SB-aligned block
        for (int is = 0; is < xrow; is++)
            xtab[size_t(iv0) * xrow + is].x += (is % fp->SB) * fp->hrow;
    }
    for (int iv1 = 0; iv1 < P.Nv1; iv1++) {
        const double v1 = coordh(iv1, P.Tv1, P.v1_0);
        double gam, del;
        if (!psi90) {
            gam = (z0 + y0 * v1 - double(P.u1_0)) / double(P.Tu1);
            del = double(V.Ty) * v1 / double(P.Tu1);
        } else {
            gam = (z0 - x0 * v1 - double(P.u1_0)) / double(P.Tu1);
            del = -double(V.Tx) * v1 / double(P.Tu1);
        }
        for (int is = 0; is < Ns; is++) {
            const double nu = gam + double(is) * del;
            const double m = std::floor(nu);
            const double ps = nu - m;
            ztab[size_t(iv1) * Ns + is] = make_int4(int(m), ibits(float(ps)), ibits(float(1.0 - ps)), 0);
        }
        // active u1 rows.  matched: unmasked rows.  PETL: rows a voxel passing the centre test can touch.
        const float v1f = coordh(iv1, P.Tv1, P.v1_0);
        volatile float rv = P.R * std::fabs(v1f);
        const float thr = P.H - rv;
        int lo_m = P.Nu1, hi_m = -1, lo_p = P.Nu1, hi_p = -1;
        for (int iu1 = 0; iu1 < P.Nu1; iu1++) {
            const float u1 = coordh(iu1, P.Tu1, P.u1_0);
            volatile float d = std::fabs(u1) - 0.5f * P.Tu1;
            if (!(d > thr)) {
                lo_m = std::min(lo_m, iu1);
                hi_m = std::max(hi_m, iu1);
            }
            if (std::fabs(u1) <= thr + P.Tu1 + 0.5f * V.Tz + 0.01f * P.Tu1) {
                lo_p = std::min(lo_p, iu1);
                hi_p = std::max(hi_p, iu1);
            }
        }
        rows[iv1] = make_int2(lo_m, hi_m);
        rows[P.Nv1 + iv1] = make_int2(lo_p, hi_p);
        for (int iv0 = 0; iv0 < P.Nv0; iv0++) {
            const float v0 = coordh(iv0, P.Tv0, P.v0_0);
            volatile float q0 = v0 * v0;
            volatile float q1 = v1f * v1f;
            volatile float sum = 1.0f + q0;
            sum = sum + q1;
            const float obl = 1.0f / std::sqrt(float(sum));
            const float l_phi = psi90 ? V.Tx * std::fabs(1.0f / P.sin_psi) : V.Tx * std::fabs(1.0f / P.cos_psi);
            fscale[size_t(iv1) * P.Nv0 + iv0] = l_phi * obl;
            volatile float a1 = v0 * P.cos_psi;
            volatile float a2 = v0 * P.sin_psi;
            const float wy = a1 + P.sin_psi, wx = P.cos_psi - a2;
            const float lp = 1.0f / std::max(std::fabs(wy), std::fabs(wx)) * obl;
            pscale[size_t(iv1) * P.Nv0 + iv0] = lp * V.Tx;
        }
    }
    // forward tables in [slice block][view] order with each block's chunk mask
    const int SBk = fp->SB;
    const int nbt = (Ns + SBk - 1) / SBk;
    std::vector<int2> ftab(size_t(nbt) * P.Nv0 * SBk);
    std::vector<float> fwp(size_t(nbt) * P.Nv0 * FWD_WPROW, 0.0f);
    if (SBk + 1 > FWD_WPROW)
        return nullptr;
    for (int b = 0; b < nbt; b++)
        for (int iv0 = 0; iv0 < P.Nv0; iv0++) {
            const size_t row = size_t(b) * P.Nv0 + iv0;
            int mask = 0;
            for (int sl = 0; sl < SBk; sl++) {
                const int is = b * SBk + sl;  // < xrow: rows are padded past Ns
                const int2 e = xtab[size_t(iv0) * xrow + is];
                ftab[row * SBk + sl] = e;
                fwp[row * FWD_WPROW + sl] = xwp[size_t(iv0) * xrow + is];
                if (is < Ns) {
                    const int k = e.x - sl * fp->hrow;  // lane p reads H[s][p + 64 c + k]
                    for (int c = 0; c < fp->C; c++)
                        if (64 * c + k <= Na - 1 && 64 * c + 63 + k >= 0)
                            mask |= 1 << c;
                }
            }
            fwp[row * FWD_WPROW + SBk] = fbits(mask);
        }
    fp->hpad = 64 + kspread + 1;
    if (fp->hpad + Na + 63 + kspread >= fp->hrow)  // reads of an active chunk must stay inside the LDS row
        return nullptr;
    // back: S rows read at indices [m, m + NaPad] for m in [mmin, mmax]
    const int na_pad = fp->nagroups * BACK_AG;
    fp->spad_lo = std::max(0, -mmin);
    fp->srow = fp->spad_lo + std::max(P.Nu0, mmax + na_pad + 1) + 1;
    if ((fp->srow & 1) == 0)
        fp->srow += 1;
    if (size_t(2) * BACK_RB * fp->srow * 4 > 64 * 1024)
        return nullptr;

    // ---- upload ----
    const size_t nx = xtab.size() * sizeof(int2), nft = ftab.size() * sizeof(int2), nfw = fwp.size() * sizeof(float);
    const size_t nzb = ztab.size() * sizeof(int4);
    const size_t nf = fscale.size() * sizeof(float), nr = rows.size() * sizeof(int2);
    auto al = [](size_t n) { return (n + 255) & ~size_t(255); };
    const size_t total = al(nx) + al(nft) + al(nfw) + al(nzb) + 2 * al(nf) + al(nr);
    char* blk = nullptr;
    PETL_HIP_CHECK(hipMalloc(&blk, total));
    fp->dev_block = blk;
    size_t off = 0;
    fp->dev.xtab = reinterpret_cast<int2*>(blk + off);
    off += al(nx);
    fp->dev.ftab = reinterpret_cast<int2*>(blk + off);
    off += al(nft);
    fp->dev.fwp = reinterpret_cast<float*>(blk + off);
    off += al(nfw);
    fp->dev.ztab = reinterpret_cast<int4*>(blk + off);
    off += al(nzb);
    fp->dev.fscale = reinterpret_cast<float*>(blk + off);
    off += al(nf);
    fp->dev.pscale = reinterpret_cast<float*>(blk + off);
    off += al(nf);
    fp->dev.rows = reinterpret_cast<int2*>(blk + off);
    PETL_HIP_CHECK(hipMemcpy(fp->dev.xtab, xtab.data(), nx, hipMemcpyHostToDevice));
    PETL_HIP_CHECK(hipMemcpy(fp->dev.ftab, ftab.data(), nft, hipMemcpyHostToDevice));
    PETL_HIP_CHECK(hipMemcpy(fp->dev.fwp, fwp.data(), nfw, hipMemcpyHostToDevice));
    PETL_HIP_CHECK(hipMemcpy(fp->dev.ztab, ztab.data(), nzb, hipMemcpyHostToDevice));
    PETL_HIP_CHECK(hipMemcpy(fp->dev.fscale, fscale.data(), nf, hipMemcpyHostToDevice));
    PETL_HIP_CHECK(hipMemcpy(fp->dev.pscale, pscale.data(), nf, hipMemcpyHostToDevice));
    PETL_HIP_CHECK(hipMemcpy(fp->dev.rows, rows.data(), nr, hipMemcpyHostToDevice));
    return fp;
}

// kernel arguments of a plan (storage scales 1, no batch, no weights)
static FwdArgs fwd_args(const FastPlan& fp, int v1_first, int v1_stride)
{
    FwdArgs A{};
    A.P = fp.P;
    A.Ns = fp.Ns;
    A.Na = fp.Na;
    A.Nz = fp.Nz;
    A.fs_s = fp.fs_s;
    A.fs_a = fp.fs_a;
    A.fs_z = fp.fs_z;
    A.hpad = fp.hpad;
    A.hrow = fp.hrow;
    A.psi90 = fp.psi90;
    A.v1_first = v1_first;
    A.v1_stride = v1_stride;
    A.xrow = fp.xrow;
    A.f_scale = A.y_scale = A.g_inv_scale = 1.0f;
    A.wts = nullptr;
    A.part = nullptr;
    A.nv1 = (fp.P.Nv1 - v1_first + v1_stride - 1) / v1_stride;
    A.f_bstride = A.g_bstride = 0;
    return A;
}

static BackArgs back_args(const FastPlan& fp, int mode, int v1_first, int v1_stride)
{
    BackArgs A{};
    A.P = fp.P;
    A.V = fp.V;
    A.Ns = fp.Ns;
    A.Na = fp.Na;
    A.Nz = fp.Nz;
    A.fs_s = fp.fs_s;
    A.fs_a = fp.fs_a;
    A.fs_z = fp.fs_z;
    A.spad_lo = fp.spad_lo;
    A.srow = fp.srow;
    A.mode = mode;
    A.psi90 = fp.psi90;
    A.sb = fp.SB;
    A.hrow = fp.hrow;
    A.v1_first = v1_first;
    A.v1_stride = v1_stride;
    A.g_scale = 1.0f;
    A.nagroups = fp.nagroups;
    A.nv1 = (fp.P.Nv1 - v1_first + v1_stride - 1) / v1_stride;
    const int zc = back_z_chunk(fp.Na, fp.Ns, fp.Nz);
    A.nzc = (fp.Nz + zc - 1) / zc;
    A.g_bstride = plan_elems(fp.P);
    A.h_bstride = int64_t(fast_back_workspace_bytes(fp) / sizeof(float));
    A.f_bstride = int64_t(fp.V.Nx) * fp.V.Ny * fp.V.Nz;
    return A;
}

template <int SF, int SG>
static void fast_forward_t(const FastPlan& fp, void* g, const void* f, int epilogue, const void* y,
                           const float* resp, hipStream_t s, int v1_first, int v1_stride, float f_scale,
                           float y_scale, float g_scale, const float* wts = nullptr, int ek = -1,
                           double* part = nullptr, int batch = 1, int64_t f_bstride = 0, int64_t g_bstride = 0)
{
    using TG = typename Stor<SG>::T;
    using TF = typename Stor<SF>::T;
    FwdArgs A = fwd_args(fp, v1_first, v1_stride);
    A.f_scale = (SF == STORE_F16) ? f_scale : 1.0f;
    A.y_scale = (SG == STORE_F16) ? y_scale : 1.0f;
    A.g_inv_scale = (SG == STORE_F16) ? 1.0f / g_scale : 1.0f;
    A.wts = wts;
    A.part = part;
    if (ek < 0)
        ek = (epilogue == EPI_RESID) ? EK_RESID : EK_RUNTIME;
    const int nv1 = A.nv1;
    A.f_bstride = f_bstride;
    A.g_bstride = g_bstride;
    if (int64_t(nv1) * batch > 65535)
        throw hip_error("fast forward: batch x v1 rows exceeds the grid");
    fwd_fn<SF, SG> fn = pick_fwd<SF, SG>(fp.C, ek, fp.Na > 128);
    if (!fn || fp.J != FWD_J || fp.R != 1)
        throw hip_error("fast forward: no kernel instance for this shape");
    const size_t lds = size_t(2) * fp.R * fp.SB * fp.hrow * sizeof(float) + size_t(fp.Ns) * sizeof(int4);
    hipLaunchKernelGGL(fn, dim3(fp.ngroups, (fp.P.Nu1 + fp.R - 1) / fp.R, nv1 * batch), dim3(FWD_THREADS), lds, s,
                       static_cast<TG*>(g), static_cast<const TF*>(f), A, fp.dev.ftab, fp.dev.fwp, fp.dev.ztab,
                       fp.dev.fscale, epilogue, static_cast<const TG*>(y), resp);
    PETL_LAUNCH_CHECK();
}

void fast_forward(const FastPlan& fp, float* g, const float* f, int epilogue, const float* y, const float* resp,
                  hipStream_t s, int v1_first, int v1_stride, const float* wts)
{
    fast_forward_t<STORE_F32, STORE_F32>(fp, g, f, epilogue, y, resp, s, v1_first, v1_stride, 1.0f, 1.0f, 1.0f,
                                         wts);
}

void fast_forward_batch(const FastPlan& fp, float* g, const float* f, int batch, hipStream_t s)
{
    const int64_t V = int64_t(fp.V.Nx) * fp.V.Ny * fp.V.Nz, G = plan_elems(fp.P);
    const int nv1 = fp.P.Nv1;
    const int per = std::max(1, 65535 / nv1);  // batch elements per launch (grid z = elements x v1 rows)
    for (int b0 = 0; b0 < batch; b0 += per) {
        const int nb = std::min(per, batch - b0);
        fast_forward_t<STORE_F32, STORE_F32>(fp, g + b0 * G, f + b0 * V, EPI_NONE, nullptr, nullptr, s, 0, 1, 1.0f,
                                             1.0f, 1.0f, nullptr, -1, nullptr, nb, V, G);
    }
}

int64_t fast_forward_groups(const FastPlan& fp, int v1_first, int v1_stride)
{
    const int nv1 = (fp.P.Nv1 - v1_first + v1_stride - 1) / v1_stride;
    return int64_t(fp.ngroups) * ((fp.P.Nu1 + fp.R - 1) / fp.R) * nv1;
}

void fast_forward_sumsq(const FastPlan& fp, const float* f, const float* resp, const float* wts, double* part,
                        hipStream_t s, int v1_first, int v1_stride)
{
    fast_forward_t<STORE_F32, STORE_F32>(fp, nullptr, f, EPI_NONE, nullptr, resp, s, v1_first, v1_stride, 1.0f,
                                         1.0f, 1.0f, wts, EK_SUMSQ, part);
}

// (volume, planogram) formats: both fp32, both 16-bit of one kind, or an fp32 volume with 16-bit planograms
void fast_forward_st(const FastPlan& fp, int sf, int sg, void* g, const void* f, int epilogue, const void* y,
                     const float* resp, hipStream_t s, int v1_first, int v1_stride, float f_scale, float y_scale,
                     float g_scale)
{
#define PETL_FWD_CASE(A, B)                                                                                    \
    if (sf == A && sg == B) {                                                                                 \
        fast_forward_t<A, B>(fp, g, f, epilogue, y, resp, s, v1_first, v1_stride, f_scale, y_scale, g_scale); \
        return;                                                                                               \
    }
    PETL_FWD_CASE(STORE_F32, STORE_F32)
    PETL_FWD_CASE(STORE_BF16, STORE_BF16)
    PETL_FWD_CASE(STORE_F16, STORE_F16)
    PETL_FWD_CASE(STORE_F32, STORE_BF16)
    PETL_FWD_CASE(STORE_F32, STORE_F16)
#undef PETL_FWD_CASE
    throw hip_error("fast forward: unsupported pair of storage formats");
}

size_t fast_back_workspace_bytes(const FastPlan& fp)
{
    return size_t(fp.P.Nv1) * fp.P.Nu1 * size_t(fp.Ns) * fp.Na * sizeof(float);
}

template <int ST>
static void fast_back_t(const FastPlan& fp, const void* g, float* f, int mode, float* workspace, hipStream_t s,
                        int v1_first, int v1_stride, float g_scale, int batch = 1)
{
    using TS = typename Stor<ST>::T;
    BackArgs A = back_args(fp, mode, v1_first, v1_stride);
    A.g_scale = g_scale;
    const int nv1 = A.nv1, nzc = A.nzc, zc = back_z_chunk(fp.Na, fp.Ns, fp.Nz);
    if (int64_t(nv1) * batch > 65535 || int64_t(nzc) * batch > 65535)
        throw hip_error("fast back: batch x v1 rows exceeds the grid");
    const int2* rows = fp.dev.rows + (mode == BACK_PETL ? fp.P.Nv1 : 0);
    const float* scale = (mode == BACK_PETL) ? fp.dev.pscale : fp.dev.fscale;
    const int nthr = ((std::min(fp.Ns, BACK_SG) + 63) / 64) * 64;
    const size_t lds =
        std::max(size_t(2) * BACK_RB * fp.srow, size_t(std::min(fp.Ns, BACK_SG)) * (BACK_TW + 1)) * sizeof(float);
    const int gx = fp.nagroups * fp.nsgroups;
    // the 16-bit DMA form moves row dwords: rows of an even number of bins (and so 4-byte aligned); its
    // matched-mode instance needs registers for two waves per SIMD only, so that mode keeps the register form
    const bool dma = (ST == STORE_F32) ||
                     (mode == BACK_PETL && fp.P.Nu0 % 2 == 0 && fp.P.Nu0 <= 2 * BACK_STW);
    if (dma) {
        // MASKED: matched mode's per-bin u0 mask, a separate instance so that PETL mode carries none of it
        if (mode == BACK_MATCHED)
            hipLaunchKernelGGL((back_h2_kernel<BACK_AG, ST, true, true>), dim3(gx, fp.P.Nu1, nv1 * batch),
                               dim3(nthr), lds, s, static_cast<const TS*>(g), workspace, A, fp.dev.xtab, fp.xrow,
                               scale, rows);
        else
            hipLaunchKernelGGL((back_h2_kernel<BACK_AG, ST, true>), dim3(gx, fp.P.Nu1, nv1 * batch), dim3(nthr),
                               lds, s, static_cast<const TS*>(g), workspace, A, fp.dev.xtab, fp.xrow, scale, rows);
    } else {
        hipLaunchKernelGGL((back_h2_kernel<BACK_AG, ST>), dim3(gx, fp.P.Nu1, nv1 * batch), dim3(nthr), lds, s,
                           static_cast<const TS*>(g), workspace, A, fp.dev.xtab, fp.xrow, scale, rows);
    }
    PETL_LAUNCH_CHECK();
    const dim3 zgrid((fp.Na + 63) / 64, fp.Ns, nzc * batch);
    if (zc == 32)
        hipLaunchKernelGGL(back_z_kernel<32>, zgrid, dim3(64), 0, s, workspace, f, A, fp.dev.ztab, rows);
    else if (zc == 16)
        hipLaunchKernelGGL(back_z_kernel<16>, zgrid, dim3(64), 0, s, workspace, f, A, fp.dev.ztab, rows);
    else if (zc == 8)
        hipLaunchKernelGGL(back_z_kernel<8>, zgrid, dim3(64), 0, s, workspace, f, A, fp.dev.ztab, rows);
    else
        hipLaunchKernelGGL(back_z_kernel<4>, zgrid, dim3(64), 0, s, workspace, f, A, fp.dev.ztab, rows);
    PETL_LAUNCH_CHECK();
}

void fast_back(const FastPlan& fp, const float* g, float* f, int mode, float* workspace, hipStream_t s, int v1_first,
               int v1_stride)
{
    fast_back_t<STORE_F32>(fp, g, f, mode, workspace, s, v1_first, v1_stride, 1.0f);
}

void fast_back_batch(const FastPlan& fp, const float* g, float* f, int mode, float* workspace, int ws_elements,
                     int batch, hipStream_t s)
{
    const int64_t V = int64_t(fp.V.Nx) * fp.V.Ny * fp.V.Nz, G = plan_elems(fp.P);
    const int zc = back_z_chunk(fp.Na, fp.Ns, fp.Nz);
    const int nzc = (fp.Nz + zc - 1) / zc;
    const int per = std::max(1, std::min({ws_elements, 65535 / fp.P.Nv1, 65535 / nzc}));
    for (int b0 = 0; b0 < batch; b0 += per)
        fast_back_t<STORE_F32>(fp, g + b0 * G, f + b0 * V, mode, workspace, s, 0, 1, 1.0f, std::min(per, batch - b0));
}

void fast_back_st(const FastPlan& fp, int st, const void* g, float* f, int mode, float* workspace, hipStream_t s,
                  int v1_first, int v1_stride, float g_scale)
{
    switch (st) {
    case STORE_F32: fast_back_t<STORE_F32>(fp, g, f, mode, workspace, s, v1_first, v1_stride, 1.0f); break;
    case STORE_BF16: fast_back_t<STORE_BF16>(fp, g, f, mode, workspace, s, v1_first, v1_stride, 1.0f); break;
    case STORE_F16: fast_back_t<STORE_F16>(fp, g, f, mode, workspace, s, v1_first, v1_stride, g_scale); break;
    default: throw hip_error("fast back: unknown storage format");
    }
}


// =============================================================================================
// batched MLEM iterations (mlem_batch.h)
// =============================================================================================
namespace {

constexpr int BATCH_ZC = 8;            // z rows per back_z work item (the sums do not depend on it)
constexpr int BATCH_H2_THREADS = 128;  // back_h2 lanes = slices (the batched form takes Ns <= 128)

struct BFwd {
    FwdArgs A;
    float* g;
    const float* f;
    const int2* ftab;
    const float* fwp;
    const int4* ztab;
    const float* fscale;
    const float* y;
    int ngroups, nu1;
};

struct BBack {
    BackArgs A;
    const float* g;
    float* h2;
    float* f;
    const int2* xtab;
    const float* scale;
    const int2* rows;
    const int4* ztab;
    int xrow, gx, nu1, nac, nzc, overwrite;
};

struct BUpd {
    float* fo;
    const float* fi;
    const float* d;
    const float* S;
    int N1, N2, N3, prior, nk;  // nk = 64-voxel chunks per N3 row
    float delta, beta;
    HuberParams h;
};

// A phase's jobs travel as the launch's kernel argument (read with scalar loads, like any kernel
// argument; a table in global memory would cost the forward ~50 registers, since the compiler cannot
// prove that the work's stores leave it unchanged).  start[] holds the prefix sums of the jobs' work
// items; blockIdx.x is the flat item index.
constexpr int KARG_BYTES = 3800;  // below the 4 KB kernel-argument limit
template <typename Job>
struct PhaseArgs {
    static constexpr int MAXJ = (KARG_BYTES - 8) / int(sizeof(Job) + sizeof(int));
    int n;
    int start[MAXJ + 1];
    Job job[MAXJ];
};
static_assert(sizeof(PhaseArgs<BFwd>) <= 4096 && sizeof(PhaseArgs<BBack>) <= 4096 && sizeof(PhaseArgs<BUpd>) <= 4096,
              "kernel arguments of the batched phases");
static_assert(PhaseArgs<BFwd>::MAXJ >= 4 && PhaseArgs<BBack>::MAXJ >= 4, "jobs per batched launch");

// the job whose prefix range [start[j], start[j + 1]) holds item (start[0] = 0, item < start[n])
template <typename PA>
__device__ __forceinline__ int batch_job(const PA& A, int item)
{
    int j = 0;
    while (j + 1 < A.n && item >= A.start[j + 1])
        j++;
    return j;
}

__global__ void __launch_bounds__(FWD_THREADS) batch_fwd_kernel(const PhaseArgs<BFwd> P)
{
    const int item = blockIdx.x;
    const int j = batch_job(P, item);
    const BFwd& J = P.job[j];
    int l = item - P.start[j];
    const int grp = l % J.ngroups;
    l /= J.ngroups;
    const int iu1 = l % J.nu1, iv1 = l / J.nu1;
    fwd_body<FWD_J, 4, 8, 1, STORE_F32, STORE_F32, EK_RUNTIME, 128>(J.g, J.f, J.A, J.ftab, J.fwp, J.ztab, J.fscale,
                                                                    EPI_RATIO, J.y, nullptr, grp, iu1, iv1, 0);
}

__global__ void __launch_bounds__(BATCH_H2_THREADS) batch_h2_kernel(const PhaseArgs<BBack> P)
{
    const int item = blockIdx.x;
    const int j = batch_job(P, item);
    const BBack& J = P.job[j];
    int l = item - P.start[j];
    const int gxi = l % J.gx;
    l /= J.gx;
    const int iu1 = l % J.nu1, iv1 = l / J.nu1;
    const int sgr = gxi / J.A.nagroups, ag = gxi - sgr * J.A.nagroups;
    back_h2_body<BACK_AG, STORE_F32, true, false>(J.g, J.h2, J.A, J.xtab, J.xrow, J.scale, J.rows, ag, sgr, iu1, iv1);
}

// z items: start[] of these arguments holds the prefix sums of the back_z work (one wave each)
__global__ void __launch_bounds__(64) batch_z_kernel(const PhaseArgs<BBack> P)
{
    const int item = blockIdx.x;
    const int j = batch_job(P, item);
    const BBack& J = P.job[j];
    int l = item - P.start[j];
    const int ac = l % J.nac;
    l /= J.nac;
    const int is = l % J.A.Ns, zc = l / J.A.Ns;
    back_z_body<BATCH_ZC>(J.h2, J.f, J.A, J.ztab, J.rows, threadIdx.x, ac, is, zc * BATCH_ZC, J.overwrite != 0);
}

__global__ void __launch_bounds__(64) batch_update_kernel(const PhaseArgs<BUpd> P)
{
    const int item = blockIdx.x;
    const int j = batch_job(P, item);
    const BUpd& U = P.job[j];
    int l = item - P.start[j];
    const int kc = l % U.nk;
    l /= U.nk;
    const int jj = l % U.N2, ii = l / U.N2;
    const int k = kc * 64 + int(threadIdx.x);
    if (k >= U.N3)
        return;
    if (U.prior == 0)
        mlem_update_at(U.fo, U.d, U.S, (int64_t(ii) * U.N2 + jj) * U.N3 + k);
    else if (U.prior == 1)
        mlem_update_prior_at<0>(U.fo, U.fi, U.d, U.S, ii, jj, k, U.N1, U.N2, U.N3, U.delta, U.beta, U.h);
    else
        mlem_update_prior_at<1>(U.fo, U.fi, U.d, U.S, ii, jj, k, U.N1, U.N2, U.N3, U.delta, U.beta, U.h);
}

size_t batch_fwd_lds(const FastPlan& fp)
{
    return size_t(2) * fp.R * fp.SB * fp.hrow * sizeof(float) + size_t(fp.Ns) * sizeof(int4);
}

size_t batch_back_lds(const FastPlan& fp)
{
    return std::max(size_t(2) * BACK_RB * fp.srow, size_t(std::min(fp.Ns, BACK_SG)) * (BACK_TW + 1)) * sizeof(float);
}

// launches `kernel` over the jobs in groups of PhaseArgs::MAXJ; items(job) = its work items, lds(job) its LDS
template <typename Job, typename Kern, typename Items, typename Lds>
void launch_phase(Kern kernel, const std::vector<Job>& jobs, Items items, Lds lds, int threads, hipStream_t s)
{
    using PA = PhaseArgs<Job>;
    for (size_t b = 0; b < jobs.size(); b += PA::MAXJ) {
        PA A{};
        A.n = int(std::min<size_t>(PA::MAXJ, jobs.size() - b));
        A.start[0] = 0;
        size_t shared = 0;
        for (int k = 0; k < A.n; k++) {
            A.job[k] = jobs[b + k];
            A.start[k + 1] = A.start[k] + items(A.job[k]);
            shared = std::max(shared, lds(k));
        }
        if (A.start[A.n] == 0)
            continue;
        hipLaunchKernelGGL(kernel, dim3(A.start[A.n]), dim3(threads), shared, s, A);
        PETL_LAUNCH_CHECK();
    }
}

}  // namespace

struct MlemBatch::Impl {};

MlemBatch::MlemBatch() : impl_(nullptr) {}

MlemBatch::~MlemBatch() {}

bool MlemBatch::eligible(const std::vector<const FastPlan*>& plans)
{
    if (plans.empty() || plans.size() > size_t(BATCH_MAX_PLANOGRAMS))
        return false;
    for (const FastPlan* fp : plans)
        if (!fp || fp->C > 4 || fp->Na > 128 || fp->Ns > BATCH_H2_THREADS || fp->J != FWD_J || fp->R != 1 ||
            fp->SB != 8)
            return false;
    return true;
}

void MlemBatch::launch(const std::vector<BatchIterJob>& jobs, hipStream_t s)
{
    std::vector<BFwd> fw;
    std::vector<const FastPlan*> fplan;
    std::vector<BBack> bk[BATCH_MAX_PLANOGRAMS];
    std::vector<const FastPlan*> bplan[BATCH_MAX_PLANOGRAMS];
    std::vector<BUpd> up;
    for (const BatchIterJob& jb : jobs) {
        std::vector<const FastPlan*> pl(jb.plan, jb.plan + jb.nplans);
        if (!eligible(pl))
            throw hip_error("batched MLEM: a request outside the batched form");
        for (int p = 0; p < jb.nplans; p++) {
            const FastPlan& fp = *jb.plan[p];
            BFwd F{};
            F.A = fwd_args(fp, 0, 1);
            F.g = jb.ratio[p];
            F.f = jb.f_in;
            F.ftab = fp.dev.ftab;
            F.fwp = fp.dev.fwp;
            F.ztab = fp.dev.ztab;
            F.fscale = fp.dev.fscale;
            F.y = jb.data[p];
            F.ngroups = fp.ngroups;
            F.nu1 = fp.P.Nu1;
            fw.push_back(F);
            fplan.push_back(&fp);
            BBack B{};
            B.A = back_args(fp, BACK_PETL, 0, 1);
            B.g = jb.ratio[p];
            B.h2 = jb.h2;
            B.f = jb.d;
            B.xtab = fp.dev.xtab;
            B.scale = fp.dev.pscale;
            B.rows = fp.dev.rows + fp.P.Nv1;
            B.ztab = fp.dev.ztab;
            B.xrow = fp.xrow;
            B.gx = fp.nagroups * fp.nsgroups;
            B.nu1 = fp.P.Nu1;
            B.nac = (fp.Na + 63) / 64;
            B.nzc = (fp.Nz + BATCH_ZC - 1) / BATCH_ZC;
            B.overwrite = (p == 0) ? 1 : 0;  // d = the first planogram's sum, then += the next ones
            bk[p].push_back(B);
            bplan[p].push_back(&fp);
        }
        BUpd U{};
        U.fo = jb.f_out;
        U.fi = jb.f_in;
        U.d = jb.d;
        U.S = jb.S;
        U.N1 = jb.Nz;
        U.N2 = jb.Ny;
        U.N3 = jb.Nx;
        U.nk = (jb.Nx + 63) / 64;
        U.prior = (jb.beta > 0.0f) ? jb.prior : 0;
        U.delta = jb.delta;
        U.beta = jb.beta;
        if (U.prior == 2) {
            float dl = jb.delta, bt = jb.beta;
            U.h = huber_params(dl, bt, jb.p);
            U.delta = dl;
            U.beta = bt;
        }
        if (U.prior == 0 && U.fo != U.fi)
            throw hip_error("batched MLEM: the update without a prior is in place");
        up.push_back(U);
    }
    // forward with the ratio epilogue of every (request, planogram)
    size_t base = 0;
    for (size_t b = 0; b < fw.size(); b += PhaseArgs<BFwd>::MAXJ) {
        const size_t e = std::min(fw.size(), b + PhaseArgs<BFwd>::MAXJ);
        std::vector<BFwd> part(fw.begin() + long(b), fw.begin() + long(e));
        launch_phase(batch_fwd_kernel, part, [](const BFwd& F) { return F.ngroups * F.nu1 * F.A.P.Nv1; },
                     [&](int k) { return batch_fwd_lds(*fplan[b + size_t(k)]); }, FWD_THREADS, s);
    }
    (void)base;
    // back projection, planogram 0 then 1: the u0 gather, then the u1 gather into d
    for (int p = 0; p < BATCH_MAX_PLANOGRAMS; p++) {
        for (size_t b = 0; b < bk[p].size(); b += PhaseArgs<BBack>::MAXJ) {
            const size_t e = std::min(bk[p].size(), b + PhaseArgs<BBack>::MAXJ);
            std::vector<BBack> part(bk[p].begin() + long(b), bk[p].begin() + long(e));
            launch_phase(batch_h2_kernel, part, [](const BBack& B) { return B.gx * B.nu1 * B.A.P.Nv1; },
                         [&](int k) { return batch_back_lds(*bplan[p][b + size_t(k)]); }, BATCH_H2_THREADS, s);
            launch_phase(batch_z_kernel, part, [](const BBack& B) { return B.nac * B.A.Ns * B.nzc; },
                         [](int) { return size_t(0); }, 64, s);
        }
    }
    launch_phase(batch_update_kernel, up, [](const BUpd& U) { return U.nk * U.N2 * U.N1; }, [](int) { return size_t(0); },
                 64, s);
}

}  // namespace petl
