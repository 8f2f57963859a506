// General SF projector for any panel angle psi and any voxel / crystal pitch (the geometries the
// separable fast kernels do not take: oblique panels, voxels of another size than the crystals).
//
// Model.  A view (v1, v0) is a set of parallel lines of response.  The reference steps through the
// volume in planes ("slices") normal to the axis the line direction r is closest to
// (src/projectors.cu:77: x planes when |r_x| >= |r_y|, else y planes).  In slice s the line of bin
// (u0, u1) crosses the plane at in-plane coordinate a_c (x for y planes, y for x planes) and height z_c;
// the bin's footprint there is an interval of width T_u0 |p_a| along a (p_a = cos psi for y planes,
// sin psi for x planes) and T_u1 along z, and a voxel contributes the product of the overlap lengths of
// its box with those intervals (src/projectors.cu:100-177 evaluates the same overlaps as
// min(upper) - max(lower) in bin units).  Both crossings are affine in the bin indices and the slice
// index, so in voxel-index units
//
//     xi   = xi0[v0][s] + Au * iu0                      (a crossing; Au, xi0 depend on v0 only)
//     zeta = Z0 + Bs * s + Bu * iu0 + Bz * iu1          (z crossing; Z0, Bs, Bu per view)
//
// with every coefficient formed in FP64 on the host.  The bin value is scale * sum f * ov(xi, hb, i) *
// ov(zeta, hz, k), ov(c, h, i) the overlap of [c - h, c + h] with [i - 1/2, i + 1/2], and scale the
// slice length over the footprint normalisation and the obliquity 1 / sqrt(1 + v0^2 + v1^2).
//
// The forward and the matched back projector evaluate the crossings and overlaps with the same
// functions on the same table entries, and a tap counts exactly when both overlaps are positive, so the
// matched back projector is the exact transpose of the forward.
//
// With the separable forward, matched mode takes a column-driven form of its transpose too (per column and
// pass of heights, per view the bins whose two cells include the column, each with a window of rows), which
// writes per-chunk partial volumes like the PETL form.
//
// The separable PETL-mode back projector (petl_np > 0, voxel height = crystal pitch) uses that a voxel's u1
// interval moves by exactly one row per voxel height: per column it takes the pass's first height's two u1
// rows and weights for every height of the pass (the voxel-centre test stays per height) and reads a
// sliding window of planogram rows per u0 cell.
//
// Separable form.  When a u1 row steps the z crossing by exactly one voxel (Bz = 1: crystal pitch = voxel
// height) and a bin's footprint is at most one voxel wide on both axes (hb, hz <= 1/2), a bin meets at
// most two cells per axis in each slice: i0 = floor(xi + 1/2 - hb) and i0 + 1 along a, k0 + iu1 and
// k0 + iu1 + 1 along z with k0 = floor(zb + 1/2 - hz) taken from row 0's crossing zb = Z0 + Bs s + Bu iu0,
// and the z overlaps are row 0's for every row.  Those kernels (gen_sep = true) walk each bin column up
// the u1 rows with a sliding window of z values read from zero-padded volume copies, and the matched back
// projector applies the same rule per voxel, so the pair stays an exact transpose; other pitches use the
// per-bin kernels.  The PETL-mode back projector follows the
// reference's voxel-driven semantics instead (src/projectors.cu:240-377): per voxel and view row v1 the
// voxel-centre u1 test and u1 interval, per v0 the projection of the voxel's mid-line along the dominant
// axis onto u0, exact bilinear weights.
#pragma once
#include <memory>
#include "petl_common.h"

namespace petl {

constexpr int GEN_RB = 16;          // u1 rows per pass of the separable forward (register accumulators)
constexpr int GEN_PZ = GEN_RB + 2;  // zero planes below and above the volume in its padded copies

struct GenView {
    float Au;     // d xi / d iu0
    float hb;     // half width of a bin's footprint along a, in voxels
    float Z0, Bs, Bu;
    float scale;  // forward / matched per-bin factor
    float lpetl;  // PETL back: 1 / max(|w_y|, |w_x|) / sqrt(1 + v0^2 + v1^2)
    int xs;       // 1: x planes (a = y, slices along x); 0: y planes (a = x)
};

struct GenPlan {
    PlanGeom P;
    VolGeom V;
    int Nsmax = 0;         // max(Nx, Ny): row length of xi0
    float Bz = 0, hz = 0;  // d zeta / d iu1; half height of a bin's footprint in voxels
    bool any_xs = false;   // some view takes x planes (reads the transposed volume)
    bool sep = false;      // the separable kernels apply (Bz = 1, hb and hz at most 1/2)
    int petl_np = 0;       // u0 cells per column of the separable PETL back (2 or 3), 0: per-voxel kernel
    GenView* views = nullptr;   // [Nv1][Nv0]
    float* xi0 = nullptr;       // [Nv0][Nsmax]
    float4* pcoef = nullptr;    // [Nv0] PETL back: {Cx, Cy, C0, h0}: u0 centre = x Cx + y Cy + C0 (bins)
    int2* rows = nullptr;       // [Nv1] the u1 rows the forward's mask leaves, {first, last} (first > last: none)
    void* block = nullptr;
    ~GenPlan();
};

std::unique_ptr<GenPlan> make_gen_plan(const PlanGeom& P, const VolGeom& V);

// fT = f with x and y swapped ([z][x][y]); the per-bin forward reads x-plane views from it so that the
// footprint axis is the contiguous one.
void launch_transpose_xy(const float* f, float* fT, const VolGeom& V, hipStream_t s);

// The separable forward's zero-padded copies of f: fpy[z + GEN_PZ][y][x + 1] (y planes) and
// fpx[z + GEN_PZ][x][y + 1] (x planes), whose cells hold two floats, f at heights z and z + 1; each copy is
// gen_padded_elems(V) floats with GEN_PZ planes of zeros below and above the volume and one zero column on
// each side of the footprint axis (every element written, no separate clear).
size_t gen_padded_elems(const VolGeom& V);
void launch_pad_volumes(const float* f, float* fpy, float* fpx, const VolGeom& V, hipStream_t s);

// The volume forms a general forward reads: f itself, fT (launch_transpose_xy) for the per-bin kernels
// when gp.any_xs, the padded copies (launch_pad_volumes) for the separable ones.
struct GenVolume {
    const float* f = nullptr;
    const float* fT = nullptr;
    const float* fpy = nullptr;
    const float* fpx = nullptr;
};

// g = A f (+ epilogue: EPI_RATIO g = resp y / (resp A f), EPI_RESID g = (resp A f - y) w), v1 rows
// v1_first, v1_first + v1_stride, ...
void gen_forward(const GenPlan& gp, float* g, const GenVolume& vol, int epilogue, const float* y,
                 const float* resp, const float* wts, hipStream_t s, int v1_first = 0, int v1_stride = 1);
// f += A^T g in mode BACK_PETL or BACK_MATCHED (v1 rows as above).  workspace: at least
// gen_back_workspace_bytes(gp) bytes for the column-driven kernels (separable PETL mode, and matched mode
// with the separable forward), which write one partial volume per chunk of v1 rows there and sum them in
// chunk order; null takes the per-voxel kernels instead.
void gen_back(const GenPlan& gp, const float* g, float* f, int mode, hipStream_t s, int v1_first = 0,
              int v1_stride = 1, float* workspace = nullptr);
size_t gen_back_workspace_bytes(const GenPlan& gp);
// the same figure from the geometry alone (before a plan is built)
size_t gen_back_workspace_estimate(const PlanGeom& P, const VolGeom& V);

}  // namespace petl
