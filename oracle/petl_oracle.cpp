// =====================================================================================================
//  PETL hot-path CPU ORACLE  --  TEST INFRASTRUCTURE ONLY
//
//  This file is a plain C++17/OpenMP restatement of the reference algorithm for the separable-footprint
//  (SF) planogram projector path.  It exists to CHECK the HIP product path, and to be timed as the CPU
//  baseline.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it, and
//  only to check or time results.  Nothing in petl_amd/ links, loads or calls it.
//
//  Every routine is templated on the arithmetic type R:
//     R = float  : the fp32 restatement ("fp32 oracle").  Built with -ffp-contract=off so every
//                  multiply/add rounds exactly as written.
//     R = double : the FP64 oracle used for the precision study.  Geometry parameters stay the float
//                  values the reference computes; discrete decisions (masks, slice-branch choice) are
//                  evaluated in float exactly as in the fp32 path, so fp32 and fp64 differ only in
//                  continuous arithmetic.
//
//  Reference files followed (paths relative to the reference root):
//     planogram geometry          src/planogram.cpp:143-186, src/petl_c_interface.cpp:216-219
//     volume grid                 src/parameters.cpp:250-284
//     forward projector           src/projectors.cu:11-238 (project_planogram), :379-438 (project_SF)
//     back projector (PETL mode)  src/projectors.cu:240-377 (backproject_planogram), :440-497,
//                                 summed over planograms as in src/petl.cpp:226-277
//     relative-difference grad    src/relative_differences.cu:21-28, 107-173, 481-532
//     p-Huber TV gradient         src/total_variation.cu:21-78, 434-660
//     p-Huber TV cost, quad. form src/total_variation.cu:80-216, 218-417, 662-800
//     Gaussian smoothing          src/relative_differences.cu:314-420, 753-800
//     geometric solids            src/phantom.cpp:502-640 (init), 1120-1141 (scale)
//     ray tracing, stopping power src/analytic_ray_tracing_gpu.cu:48-862
//
//  Deliberate differences from the reference:
//     * texture fetches are replaced by exact array reads: the forward samples texel centres (exact in
//       CUDA as well); the backprojector uses exact bilinear weights instead of the hardware's 8-bit
//       fractional filter weights (parity for that rounding is unpinned),
//     * rsqrtf (built with --use_fast_math in the reference) is 1/sqrt,
//     * "matched" back projection mode = exact transpose of this forward (the reference has none).
// =====================================================================================================
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>
#include <algorithm>
#include <omp.h>

extern "C" {

typedef struct orc_plan {
    float psi, R, L, H, v_m0, v_m1;
    int N_u0, N_u1, N_v0, N_v1;
    float T_u0, T_u1, T_v0, T_v1;
    float u0_0, u1_0, v0_0, v1_0;
    float cos_psi, sin_psi;
} orc_plan;

typedef struct orc_vol {
    int numX, numY, numZ;
    float voxelWidth, voxelHeight;
    float offsetX, offsetY, offsetZ;
    float x0, y0, z0;
} orc_vol;

// one geometric solid of a phantom or of the scintillator module model (src/phantom.h:17-50)
typedef struct orc_solid {
    int type;  // 0 ellipsoid, 1 parallelepiped, 2-4 cylinder x/y/z, 5-7 cone x/y/z
    float c[3], r[3];
    float val;
    float A[9];
    float clip[6][4];
    int rotated;
    int nclip;
    float cone[2];
} orc_solid;

}  // extern "C"

namespace {

// ------------------------------------------------------------------------------------------------
// geometry
// ------------------------------------------------------------------------------------------------

// src/petl_c_interface.cpp:216-219 converts degrees and halves D/L/H in double, then
// src/planogram.cpp:143-186 derives the sampling.  Float/double promotion follows the C++ source.
int plan_init(float psi_deg, float D, float L, float H, float v_m0, float v_m1_in, float T, orc_plan* o)
{
    const float psi = (float)((double)psi_deg * 3.1415926535897932385 / 180.0);
    const float R = (float)(0.5 * (double)D);
    const float Lh = (float)(0.5 * (double)L);
    const float Hh = (float)(0.5 * (double)H);
    if (R <= 0.0f || Lh <= 0.0f || Hh <= 0.0f || v_m0 <= 0.0f || v_m1_in < 0.0f || T <= 0.0f)
        return 0;
    o->psi = psi;
    o->R = R;
    o->L = Lh;
    o->H = Hh;
    o->v_m0 = v_m0;
    const float h_over_r = Hh / R;
    o->v_m1 = (v_m1_in < h_over_r) ? v_m1_in : h_over_r;  // std::min(H/R, v_m1)
    o->T_u0 = T;
    o->T_u1 = T;
    const float Tv0 = T / R;
    const float Tv1 = T / R;
    o->N_u0 = 2 * (int)((Lh + R * v_m0) / T);
    o->N_u1 = 2 * (int)(Hh / T);
    o->N_v0 = 2 * (int)(v_m0 / Tv0) + 1;
    o->N_v1 = 2 * (int)(o->v_m1 / Tv1) + 1;
    o->T_v0 = (float)((double)v_m0 * 2.0 / (double)(float)o->N_v0);  // T_v0 recomputed, T_v1 is not
    o->T_v1 = Tv1;
    o->u0_0 = (float)((double)(-T * (float)(o->N_u0 - 1)) / 2.0);
    o->u1_0 = (float)((double)(-T * (float)(o->N_u1 - 1)) / 2.0);
    o->v0_0 = (float)((double)(-o->T_v0 * (float)(o->N_v0 - 1)) / 2.0);
    o->v1_0 = (float)((double)(-o->T_v1 * (float)(o->N_v1 - 1)) / 2.0);
    o->cos_psi = (float)std::cos((double)psi);
    o->sin_psi = (float)std::sin((double)psi);
    return 1;
}

// src/parameters.cpp:271-284: centre of voxel 0, computed in double and stored as float.
void vol_init(int nx, int ny, int nz, float vw, float vh, float ox, float oy, float oz, orc_vol* v)
{
    v->numX = nx;
    v->numY = ny;
    v->numZ = nz;
    v->voxelWidth = vw;
    v->voxelHeight = vh;
    v->offsetX = ox;
    v->offsetY = oy;
    v->offsetZ = oz;
    v->x0 = (float)((0.0 - 0.5 * (nx - 1)) * (double)vw + (double)ox);
    v->y0 = (float)((0.0 - 0.5 * (ny - 1)) * (double)vw + (double)oy);
    v->z0 = (float)((0.0 - 0.5 * (nz - 1)) * (double)vh + (double)oz);
}

inline float coord_f(int i, float T, float start) { return (float)i * T + start; }

// 1/sqrt(1+v0^2+v1^2) of a view (src/projectors.cu:237)
template <typename R>
inline R obliquity(const orc_plan& P, int iv1, int iv0)
{
    const R v1 = R(iv1) * R(P.T_v1) + R(P.v1_0);
    const R v0 = R(iv0) * R(P.T_v0) + R(P.v0_0);
    return R(1) / std::sqrt(R(1) + v0 * v0 + v1 * v1);
}

// Forward masks of src/projectors.cu:30-43, always evaluated in float.
inline bool bin_masked(const orc_plan& P, int iv1, int iv0, int iu1, int iu0)
{
    const float v1 = coord_f(iv1, P.T_v1, P.v1_0);
    const float v0 = coord_f(iv0, P.T_v0, P.v0_0);
    const float u1 = coord_f(iu1, P.T_u1, P.u1_0);
    const float u0 = coord_f(iu0, P.T_u0, P.u0_0);
    if (std::fabs(u0) - 0.5f * P.T_u0 > P.L + P.R * std::fabs(v0))
        return true;
    if (std::fabs(u1) - 0.5f * P.T_u1 > P.H - P.R * std::fabs(v1))
        return true;
    return false;
}

// ------------------------------------------------------------------------------------------------
// Forward projector: tap enumeration of one bin.  src/projectors.cu:46-237.
// The visitor receives (voxel linear index, weight w0*w1) for every tap with positive weights; the
// bin value is  l_phi * sum(f*w0*w1) / sqrt(1+v0^2+v1^2).  Returning the per-bin scale separately lets
// the forward and its exact transpose share one enumeration.  Returns l_phi; the caller applies
// 1/sqrt(1+v0^2+v1^2) (rsqrtf in the reference) after it, in the reference's order.
// ------------------------------------------------------------------------------------------------
template <typename R, typename Visit>
R enumerate_bin(const orc_plan& P, const orc_vol& V, int iv1, int iv0, int iu1, int iu0, Visit&& visit)
{
    const R v1 = R(iv1) * R(P.T_v1) + R(P.v1_0);
    const R v0 = R(iv0) * R(P.T_v0) + R(P.v0_0);
    const R u1 = R(iu1) * R(P.T_u1) + R(P.u1_0);
    const R u0 = R(iu0) * R(P.T_u0) + R(P.u0_0);
    const R cs = R(P.cos_psi), sn = R(P.sin_psi);
    const R half = R(0.5);

    const R iu0_min = R(iu0) - half, iu0_max = R(iu0) + half;
    const R iu1_min = R(iu1) - half, iu1_max = R(iu1) + half;

    // p + t*r (src/projectors.cu:52-53)
    const R px = u0 * cs, py = u0 * sn, pz = u1;
    const R rx = -(v0 * cs + sn), ry = -(v0 * sn - cs), rz = -v1;

    const R Tx = R(V.voxelWidth), Ty = R(V.voxelWidth), Tz = R(V.voxelHeight);
    const R x0 = R(V.x0), y0 = R(V.y0), z0 = R(V.z0);
    const R Tx_inv = R(1) / Tx, Ty_inv = R(1) / Ty, Tz_inv = R(1) / Tz;
    const R Tu1_inv = R(1) / R(P.T_u1), Tu0_inv = R(1) / R(P.T_u0);
    const R u0_0 = R(P.u0_0), u1_0 = R(P.u1_0);

    const R width_z = half * R(P.T_u1);
    const R Tz_over_Tu1 = Tz * Tu1_inv;
    const R z_fw = half * Tz * Tu1_inv;
    const int Nx = V.numX, Ny = V.numY, Nz = V.numZ;

    // branch decision in float (src/projectors.cu:77)
    const float v0f = coord_f(iv0, P.T_v0, P.v0_0);
    const float rxf = -(v0f * P.cos_psi + P.sin_psi), ryf = -(v0f * P.sin_psi - P.cos_psi);

    if (std::fabs(rxf) >= std::fabs(ryf)) {
        // x-slices (src/projectors.cu:79-129)
        const R width_y = half * R(P.T_u0) * std::fabs(sn);
        const R sin_inv = R(1) / sn;
        const R rx_inv = R(1) / rx;
        const R l_phi = Tx * std::fabs(sin_inv);
        const R y_fw = half * Ty * std::fabs(sin_inv) * Tu0_inv;
        for (int ix = 0; ix < Nx; ix++) {
            const R x = R(ix) * Tx + x0;
            const R t = (x - px) * rx_inv;
            const R y_c = py + t * ry;
            const R z_c = pz + t * rz;
            const R y_A = y_c - width_y, y_B = y_c + width_y;
            const int iy_min = std::max(0, (int)std::ceil((y_A - half * Ty - y0) * Ty_inv));
            const int iy_max = std::min(Ny - 1, (int)std::floor((y_B + half * Ty - y0) * Ty_inv));
            const R z_A = z_c - width_z, z_B = z_c + width_z;
            const int iz_min = std::max(0, (int)std::ceil((z_A - half * Tz - z0) * Tz_inv));
            const int iz_max = std::min(Nz - 1, (int)std::floor((z_B + half * Tz - z0) * Tz_inv));
            R y = R(iy_min) * Ty + y0 - t * ry;
            const R z_start = (R(iz_min) * Tz + z0 - t * rz - u1_0) * Tu1_inv;
            for (int iy = iy_min; iy <= iy_max; iy++) {
                const R c = (y * sin_inv - u0_0) * Tu0_inv;
                const R lo = c - y_fw, hi = c + y_fw;
                const R w0 = std::min(iu0_max, hi) - std::max(iu0_min, lo);
                if (w0 > R(0)) {
                    R z = z_start;
                    for (int iz = iz_min; iz <= iz_max; iz++) {
                        const R w1 = std::max(R(0), std::min(iu1_max, z + z_fw) - std::max(iu1_min, z - z_fw));
                        if (w1 > R(0))
                            visit((int64_t(iz) * Ny + iy) * Nx + ix, w0, w1);
                        z += Tz_over_Tu1;
                    }
                }
                y += Ty;
            }
        }
        return l_phi;
    } else {
        // y-slices (src/projectors.cu:133-185)
        const R width_x = half * R(P.T_u0) * std::fabs(cs);
        const R cos_inv = R(1) / cs;
        const R ry_inv = R(1) / ry;
        const R l_phi = Tx * std::fabs(cos_inv);
        const R x_fw = half * Tx * std::fabs(cos_inv) * Tu0_inv;
        for (int iy = 0; iy < Ny; iy++) {
            const R y = R(iy) * Ty + y0;
            const R t = (y - py) * ry_inv;
            const R x_c = px + t * rx;
            const R z_c = pz + t * rz;
            const R x_A = x_c - width_x, x_B = x_c + width_x;
            const int ix_min = std::max(0, (int)std::ceil((x_A - half * Tx - x0) * Tx_inv));
            const int ix_max = std::min(Nx - 1, (int)std::floor((x_B + half * Tx - x0) * Tx_inv));
            const R z_A = z_c - width_z, z_B = z_c + width_z;
            const int iz_min = std::max(0, (int)std::ceil((z_A - half * Tz - z0) * Tz_inv));
            const int iz_max = std::min(Nz - 1, (int)std::floor((z_B + half * Tz - z0) * Tz_inv));
            R x = R(ix_min) * Tx + x0 - t * rx;
            const R z_start = (R(iz_min) * Tz + z0 - t * rz - u1_0) * Tu1_inv;
            for (int ix = ix_min; ix <= ix_max; ix++) {
                const R c = (x * cos_inv - u0_0) * Tu0_inv;
                const R lo = c - x_fw, hi = c + x_fw;
                const R w0 = std::min(iu0_max, hi) - std::max(iu0_min, lo);
                if (w0 > R(0)) {
                    R z = z_start;
                    for (int iz = iz_min; iz <= iz_max; iz++) {
                        const R w1 = std::max(R(0), std::min(iu1_max, z + z_fw) - std::max(iu1_min, z - z_fw));
                        if (w1 > R(0))
                            visit((int64_t(iz) * Ny + iy) * Nx + ix, w0, w1);
                        z += Tz_over_Tu1;
                    }
                }
                x += Tx;
            }
        }
        return l_phi;
    }
}

template <typename R>
void project_plan(const orc_plan& P, const orc_vol& V, const R* f, R* g, int nthreads, int v1_begin = 0,
                  int v1_end = -1, int v1_stride = 1)
{
    if (v1_end < 0)
        v1_end = P.N_v1;
    const int nrows = (v1_end - v1_begin + v1_stride - 1) / v1_stride;
    const int NV = nrows * P.N_v0;
    const int64_t row = int64_t(P.N_u1) * P.N_u0;
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads)
    for (int vv = 0; vv < NV; vv++) {
        const int iv1 = v1_begin + (vv / P.N_v0) * v1_stride, iv0 = vv % P.N_v0;
        const int v = iv1 * P.N_v0 + iv0;
        R* gv = g + int64_t(v) * row;
        for (int iu1 = 0; iu1 < P.N_u1; iu1++)
            for (int iu0 = 0; iu0 < P.N_u0; iu0++) {
                R out = R(0);
                if (!bin_masked(P, iv1, iv0, iu1, iu0)) {
                    R val = R(0);
                    const R l_phi = enumerate_bin<R>(P, V, iv1, iv0, iu1, iu0,
                                                     [&](int64_t k, R w0, R w1) { val += f[k] * w0 * w1; });
                    out = (val * l_phi) * obliquity<R>(P, iv1, iv0);
                }
                gv[int64_t(iu1) * P.N_u0 + iu0] = out;
            }
    }
}

// Exact transpose of project_plan: scatter every tap with the same weight.  Accumulates into f.
template <typename R>
void backproject_matched_plan(const orc_plan& P, const orc_vol& V, const R* g, R* f, int nthreads)
{
    const int64_t nvox = int64_t(V.numX) * V.numY * V.numZ;
    const int NV = P.N_v1 * P.N_v0;
    const int64_t row = int64_t(P.N_u1) * P.N_u0;
    int nt = std::max(1, nthreads);
    std::vector<std::vector<R>> part(nt, std::vector<R>(nvox, R(0)));
#pragma omp parallel for schedule(static) num_threads(nt)
    for (int v = 0; v < NV; v++) {
        std::vector<R>& acc = part[omp_get_thread_num()];
        std::vector<int64_t> kv;
        std::vector<R> wv;
        const int iv1 = v / P.N_v0, iv0 = v % P.N_v0;
        const R* gv = g + int64_t(v) * row;
        for (int iu1 = 0; iu1 < P.N_u1; iu1++)
            for (int iu0 = 0; iu0 < P.N_u0; iu0++) {
                const R gb = gv[int64_t(iu1) * P.N_u0 + iu0];
                if (gb == R(0) || bin_masked(P, iv1, iv0, iu1, iu0))
                    continue;
                kv.clear();
                wv.clear();
                const R l_phi = enumerate_bin<R>(P, V, iv1, iv0, iu1, iu0, [&](int64_t k, R w0, R w1) {
                    kv.push_back(k);
                    wv.push_back(w0 * w1);
                });
                const R s = gb * (l_phi * obliquity<R>(P, iv1, iv0));
                for (size_t i = 0; i < kv.size(); i++)
                    acc[kv[i]] += s * wv[i];
            }
    }
    for (int t = 0; t < nt; t++)
#pragma omp parallel for schedule(static) num_threads(nt)
        for (int64_t i = 0; i < nvox; i++)
            f[i] += part[t][i];
}

// PETL-mode back projector: src/projectors.cu:240-377 with exact bilinear weights.  Accumulates into f.
template <typename R>
void backproject_petl_plan(const orc_plan& P, const orc_vol& V, const R* g, R* f, int nthreads, int v1_begin = 0,
                           int v1_end = -1, int v1_stride = 1)
{
    if (v1_end < 0)
        v1_end = P.N_v1;
    const int Nx = V.numX, Ny = V.numY, Nz = V.numZ;
    const R cs = R(P.cos_psi), sn = R(P.sin_psi), half = R(0.5);
    const R Tx = R(V.voxelWidth), Ty = R(V.voxelWidth), Tz = R(V.voxelHeight);
    const R Tu1_inv = R(1) / R(P.T_u1), Tu0_inv = R(1) / R(P.T_u0);
    const R u0_0 = R(P.u0_0), u1_0 = R(P.u1_0);
    const int64_t row = int64_t(P.N_u1) * P.N_u0;
    const int64_t slab = int64_t(P.N_v0) * row;
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads)
    for (int iz = 0; iz < Nz; iz++) {
        for (int iy = 0; iy < Ny; iy++)
            for (int ix = 0; ix < Nx; ix++) {
                const R x = R(ix) * Tx + R(V.x0);
                const R y = R(iy) * Ty + R(V.y0);
                const R z = R(iz) * Tz + R(V.z0);
                const float xf = coord_f(ix, V.voxelWidth, V.x0), yf = coord_f(iy, V.voxelWidth, V.y0);
                const float zf = coord_f(iz, V.voxelHeight, V.z0);
                R fsum = R(0);
                for (int iv1 = v1_begin; iv1 < v1_end; iv1 += v1_stride) {
                    R val = R(0);
                    // u1 test at the voxel centre (src/projectors.cu:284-289), evaluated in float
                    const float v1f = coord_f(iv1, P.T_v1, P.v1_0);
                    const float tf = yf * P.cos_psi - xf * P.sin_psi;
                    const float u1f = zf + v1f * tf;
                    if (std::fabs(u1f) - 0.5f * P.T_u1 <= P.H - P.R * std::fabs(v1f)) {
                        const R v1 = R(iv1) * R(P.T_v1) + R(P.v1_0);
                        const R t = y * cs - x * sn;
                        const R u1 = z + v1 * t;
                        const R xdt = x * cs + y * sn;
                        const R iu1_mid = (u1 - u1_0) * Tu1_inv;
                        const R iu1_A = iu1_mid - half * Tz * Tu1_inv;
                        const R iu1_B = iu1_mid + half * Tz * Tu1_inv;
                        const int iu1_lo = std::max(0, (int)std::ceil(iu1_A - half));
                        const int iu1_hi = std::min(P.N_u1 - 1, (int)std::floor(iu1_B + half));
                        const R xdt_dy = half * Ty * sn, t_dy = half * Ty * cs;
                        const R xdt_dx = half * Tx * cs, t_dx = -half * Tx * sn;
                        const R* gs = g + int64_t(iv1) * slab;
                        for (int iv0 = 0; iv0 < P.N_v0; iv0++) {
                            const R v0 = R(iv0) * R(P.T_v0) + R(P.v0_0);
                            const R wy = v0 * cs + sn;
                            const R wx = cs - v0 * sn;
                            const R l_phi = R(1) / std::max(std::fabs(wy), std::fabs(wx)) /
                                            std::sqrt(R(1) + v0 * v0 + v1 * v1);
                            R A, B;
                            if (std::fabs(wy) > std::fabs(wx)) {
                                const R a = (xdt - xdt_dy + v0 * (t - t_dy) - u0_0) * Tu0_inv;
                                const R b = (xdt + xdt_dy + v0 * (t + t_dy) - u0_0) * Tu0_inv;
                                if (wy >= R(0)) { A = a; B = b; } else { A = b; B = a; }
                            } else {
                                const R a = (xdt - xdt_dx + v0 * (t - t_dx) - u0_0) * Tu0_inv;
                                const R b = (xdt + xdt_dx + v0 * (t + t_dx) - u0_0) * Tu0_inv;
                                if (wx >= R(0)) { A = a; B = b; } else { A = b; B = a; }
                            }
                            const int iu0_lo = std::max(0, (int)std::ceil(A - half));
                            const int iu0_hi = std::min(P.N_u0 - 1, (int)std::floor(B + half));
                            const R* gv = gs + int64_t(iv0) * row;
                            for (int iu0 = iu0_lo; iu0 <= iu0_hi; iu0++) {
                                const R w0 = std::max(R(0), std::min(R(iu0) + half, B) - std::max(R(iu0) - half, A));
                                if (!(w0 > R(0)))
                                    continue;
                                for (int iu1 = iu1_lo; iu1 <= iu1_hi; iu1++) {
                                    const R w1 = std::max(R(0), std::min(R(iu1) + half, iu1_B) -
                                                                    std::max(R(iu1) - half, iu1_A));
                                    if (w1 > R(0))
                                        val += gv[int64_t(iu1) * P.N_u0 + iu0] * l_phi * w0 * w1;
                                }
                            }
                        }
                    }
                    fsum += val * Tx;
                }
                f[(int64_t(iz) * Ny + iy) * Nx + ix] += fsum;
            }
    }
}

// PETL-mode back projection as the reference evaluates it with hardware texture filtering
// (src/projectors.cu:296-377, texture set-up src/cuda_utils.cu:1020-1044): bins are taken in pairs (iu0,
// iu0 + 1) x (iu1, iu1 + 1) from the lower bound of each footprint range, and each pair is one bilinear
// texture fetch at (iu0 + w02 / (w01 + w02) + 1/2, iu1 + w12 / (w11 + w12) + 1/2) scaled by the weight sums;
// out-of-array bins read 0 (border addressing).  In exact arithmetic this equals backproject_petl_plan.
// The texture unit holds each interpolation fraction in 9-bit fixed point with 8 fractional bits (the
// public CUDA C Programming Guide, texture fetching appendix, linear filtering); quant: 0 exact fractions,
// 1 fractions rounded to the nearest multiple of 1/256, 2 fractions truncated to a multiple of 1/256 (the
// guide does not state the rounding).  Float arithmetic throughout, as the reference kernel.
static float tex_frac(float a, int quant)
{
    if (quant == 1)
        return std::nearbyint(a * 256.0f) / 256.0f;
    if (quant == 2)
        return std::floor(a * 256.0f) / 256.0f;
    return a;
}

void backproject_petl_tex(const orc_plan& P, const orc_vol& V, const float* g, float* f, int nthreads, int quant,
                          int v1_begin, int v1_end)
{
    const int Nx = V.numX, Ny = V.numY, Nz = V.numZ;
    const float cs = P.cos_psi, sn = P.sin_psi;
    const float Tx = V.voxelWidth, Ty = V.voxelWidth, Tz = V.voxelHeight;
    const float Tu1_inv = 1.0f / P.T_u1, Tu0_inv = 1.0f / P.T_u0;
    const int64_t row = int64_t(P.N_u1) * P.N_u0;
    const int64_t slab = int64_t(P.N_v0) * row;
    auto at = [&](const float* gv, int iu1, int iu0) -> float {
        if (iu0 < 0 || iu0 >= P.N_u0 || iu1 < 0 || iu1 >= P.N_u1)
            return 0.0f;
        return gv[int64_t(iu1) * P.N_u0 + iu0];
    };
    auto ovl = [](int i, float A, float B) {
        return std::max(0.0f, std::min(float(i) + 0.5f, B) - std::max(float(i) - 0.5f, A));
    };
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads)
    for (int iz = 0; iz < Nz; iz++)
        for (int iy = 0; iy < Ny; iy++)
            for (int ix = 0; ix < Nx; ix++) {
                const float x = coord_f(ix, V.voxelWidth, V.x0), y = coord_f(iy, V.voxelWidth, V.y0);
                const float z = coord_f(iz, V.voxelHeight, V.z0);
                const float t = y * cs - x * sn, xdt = x * cs + y * sn;
                float val = 0.0f;
                for (int iv1 = v1_begin; iv1 < v1_end; iv1++) {
                    const float v1 = coord_f(iv1, P.T_v1, P.v1_0);
                    const float u1 = z + v1 * t;
                    if (std::fabs(u1) - 0.5f * P.T_u1 > P.H - P.R * std::fabs(v1))
                        continue;
                    const float mid = (u1 - P.u1_0) * Tu1_inv;
                    const float A1 = mid - 0.5f * Tz * Tu1_inv, B1 = mid + 0.5f * Tz * Tu1_inv;
                    const int r0 = int(std::ceil(A1 - 0.5f)), r1 = int(std::floor(B1 + 0.5f));
                    const float xdy = 0.5f * Ty * sn, tdy = 0.5f * Ty * cs, xdx = 0.5f * Tx * cs, tdx = -0.5f * Tx * sn;
                    for (int iv0 = 0; iv0 < P.N_v0; iv0++) {
                        const float v0 = coord_f(iv0, P.T_v0, P.v0_0);
                        const float wy = v0 * cs + sn, wx = cs - v0 * sn;
                        const float l = 1.0f / std::max(std::fabs(wy), std::fabs(wx)) /
                                        std::sqrt(1.0f + v0 * v0 + v1 * v1);
                        float A0, B0;
                        if (std::fabs(wy) > std::fabs(wx)) {
                            const float a = (xdt - xdy + v0 * (t - tdy) - P.u0_0) * Tu0_inv;
                            const float b = (xdt + xdy + v0 * (t + tdy) - P.u0_0) * Tu0_inv;
                            A0 = wy >= 0.0f ? a : b;
                            B0 = wy >= 0.0f ? b : a;
                        } else {
                            const float a = (xdt - xdx + v0 * (t - tdx) - P.u0_0) * Tu0_inv;
                            const float b = (xdt + xdx + v0 * (t + tdx) - P.u0_0) * Tu0_inv;
                            A0 = wx >= 0.0f ? a : b;
                            B0 = wx >= 0.0f ? b : a;
                        }
                        const int q0 = int(std::ceil(A0 - 0.5f)), q1 = int(std::floor(B0 + 0.5f));
                        const float* gv = g + int64_t(iv1) * slab + int64_t(iv0) * row;
                        for (int iu0 = q0; iu0 <= q1; iu0 += 2) {
                            const float w02 = ovl(iu0 + 1, A0, B0), ws0 = w02 + ovl(iu0, A0, B0);
                            if (!(ws0 > 0.0f))
                                continue;
                            const float a0 = tex_frac(w02 / ws0, quant);
                            for (int iu1 = r0; iu1 <= r1; iu1 += 2) {
                                const float w12 = ovl(iu1 + 1, A1, B1), ws1 = w12 + ovl(iu1, A1, B1);
                                if (!(ws1 > 0.0f))
                                    continue;
                                const float a1 = tex_frac(w12 / ws1, quant);
                                const float tex = (1.0f - a0) * (1.0f - a1) * at(gv, iu1, iu0) +
                                                  a0 * (1.0f - a1) * at(gv, iu1, iu0 + 1) +
                                                  (1.0f - a0) * a1 * at(gv, iu1 + 1, iu0) + a0 * a1 * at(gv, iu1 + 1, iu0 + 1);
                                val += tex * l * ws0 * ws1;
                            }
                        }
                    }
                }
                f[(int64_t(iz) * Ny + iy) * Nx + ix] += val * Tx;
            }
}

// ------------------------------------------------------------------------------------------------
// Relative-difference prior gradient: src/relative_differences.cu:21-28, 107-173, 481-532.
// Arrays are [N1][N2][N3] (numpy (z,y,x)); neighbours are clamped.
// ------------------------------------------------------------------------------------------------
template <typename R>
inline R drd(R x, R y, R delta)
{
    if (x == R(0) && y == R(0))
        return R(0);
    const R ad = std::fabs(x - y);
    const R den = x + y + delta * ad;
    return (x - y) * (delta * ad + x + R(3) * y) / (den * den);
}

template <typename R>
void rd_gradient(const R* f_in, R* Df, int N1, int N2, int N3, R delta, R beta)
{
    const int64_t n = int64_t(N1) * N2 * N3;
    std::vector<R> f(f_in, f_in + n);
    for (int64_t i = 0; i < n; i++)  // clip kernel (src/relative_differences.cu:107-115)
        if (f[i] < R(1.0e-8f))
            f[i] = R(0);
    const R d1 = R(1.0f) * beta, d2 = R(0.7071067811865475f) * beta, d3 = R(0.5773502691896258f) * beta;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < N1; i++) {
        const int im = std::max(0, i - 1), ip = std::min(N1 - 1, i + 1);
        for (int j = 0; j < N2; j++) {
            const int jm = std::max(0, j - 1), jp = std::min(N2 - 1, j + 1);
            for (int k = 0; k < N3; k++) {
                const int km = std::max(0, k - 1), kp = std::min(N3 - 1, k + 1);
                auto F = [&](int a, int b, int c) { return f[(int64_t(a) * N2 + b) * N3 + c]; };
                const R c0 = F(i, j, k);
                auto D = [&](int a, int b, int c) { return drd<R>(c0, F(a, b, c), delta); };
                const R s1 = D(ip, j, k) + D(im, j, k) + D(i, jp, k) + D(i, jm, k) + D(i, j, kp) + D(i, j, km);
                const R s2 = D(ip, jp, k) + D(ip, jm, k) + D(ip, j, kp) + D(ip, j, km) + D(im, jp, k) +
                             D(im, jm, k) + D(im, j, kp) + D(im, j, km) + D(i, jp, kp) + D(i, jp, km) +
                             D(i, jm, kp) + D(i, jm, km);
                const R s3 = D(ip, jp, kp) + D(ip, jp, km) + D(ip, jm, kp) + D(ip, jm, km) + D(im, jp, kp) +
                             D(im, jp, km) + D(im, jm, kp) + D(im, jm, km);
                Df[(int64_t(i) * N2 + j) * N3 + k] = s1 * d1 + s2 * d2 + s3 * d3;
            }
        }
    }
}

// Relative-difference cost with correct indexing (the reference cost kernel at
// src/relative_differences.cu:77 drops N.y in one index); f is clipped first as in the gradient.
double rd_cost(const double* f_in, int N1, int N2, int N3, double delta, double beta)
{
    const int64_t n = int64_t(N1) * N2 * N3;
    std::vector<double> f(f_in, f_in + n);
    for (int64_t i = 0; i < n; i++)
        if (f[i] < 1.0e-8)
            f[i] = 0.0;
    auto phi = [&](double x, double y) {
        if (x == y)
            return 0.0;
        return (x - y) * (x - y) / ((x + y) + delta * std::fabs(x - y));
    };
    double acc = 0.0;
    for (int i = 0; i < N1; i++)
        for (int j = 0; j < N2; j++)
            for (int k = 0; k < N3; k++) {
                const double c0 = f[(int64_t(i) * N2 + j) * N3 + k];
                for (int di = -1; di <= 1; di++)
                    for (int dj = -1; dj <= 1; dj++)
                        for (int dk = -1; dk <= 1; dk++) {
                            const int m = std::abs(di) + std::abs(dj) + std::abs(dk);
                            if (m == 0)
                                continue;
                            const int a = std::min(N1 - 1, std::max(0, i + di));
                            const int b = std::min(N2 - 1, std::max(0, j + dj));
                            const int c = std::min(N3 - 1, std::max(0, k + dk));
                            const double w = (m == 1) ? 1.0 : (m == 2 ? (double)0.7071067811865475f
                                                                      : (double)0.5773502691896258f);
                            acc += w * phi(c0, f[(int64_t(a) * N2 + b) * N3 + c]);
                        }
            }
    return beta * acc;
}

// ------------------------------------------------------------------------------------------------
// p-Huber TV gradient: src/total_variation.cu:21-78 (USE_P_HUBER on), 434-572, 574-660.
// beta<=0 -> 1, delta floored at 1e-8, 26 neighbours (8 in-plane ones when N1 == 1).
// ------------------------------------------------------------------------------------------------
struct huber {
    float p, delta, slope, shift;
    huber(float delta_in, float p_in)
    {
        p = p_in;
        delta = delta_in;
        slope = (float)(std::pow((double)delta_in, 2.0 - (double)p_in) / (double)p_in);
        shift = (float)((double)delta_in * (double)delta_in * (0.5 - 1.0 / (double)p_in));
    }
    template <typename R>
    R d(R x) const
    {
        if (std::fabs(x) <= R(delta))
            return x;
        return R(p) * R(slope) * std::pow(std::fabs(x), R(p)) / x;
    }
    double h(double x) const
    {
        if (std::fabs(x) <= (double)delta)
            return 0.5 * x * x;
        return (double)slope * std::pow(std::fabs(x), (double)p) + (double)shift;
    }
    // curvature weight of the quadratic form (src/total_variation.cu:66-78): h'(t)/t, i.e. 1 inside delta
    double dd(double x) const
    {
        if (std::fabs(x) <= (double)delta)
            return 1.0;
        return (double)p * (double)slope * std::pow(std::fabs(x), (double)p - 2.0);
    }
};

template <typename R>
void tv_gradient(const R* f, R* Df, int N1, int N2, int N3, float delta, float beta, float p)
{
    if (beta <= 0.0f)
        beta = 1.0f;
    if (delta < 1.0e-8f)
        delta = 1.0e-8f;
    const huber hb(delta, p);
    const R d1 = R(1.0f * beta), d2 = R(0.7071067811865475f * beta), d3 = R(0.5773502691896258f * beta);
#pragma omp parallel for schedule(static)
    for (int i = 0; i < N1; i++) {
        const int im = std::max(0, i - 1), ip = std::min(N1 - 1, i + 1);
        for (int j = 0; j < N2; j++) {
            const int jm = std::max(0, j - 1), jp = std::min(N2 - 1, j + 1);
            for (int k = 0; k < N3; k++) {
                const int km = std::max(0, k - 1), kp = std::min(N3 - 1, k + 1);
                auto F = [&](int a, int b, int c) { return f[(int64_t(a) * N2 + b) * N3 + c]; };
                const R c0 = F(i, j, k);
                auto D = [&](int a, int b, int c) { return hb.d<R>(c0 - F(a, b, c)); };
                R out;
                if (N1 == 1) {
                    out = (D(i, jp, k) + D(i, jm, k) + D(i, j, kp) + D(i, j, km)) * d1 +
                          (D(i, jp, kp) + D(i, jp, km) + D(i, jm, kp) + D(i, jm, km)) * d2;
                } else {
                    const R s1 = D(ip, j, k) + D(im, j, k) + D(i, jp, k) + D(i, jm, k) + D(i, j, kp) + D(i, j, km);
                    const R s2 = D(ip, jp, k) + D(ip, jm, k) + D(ip, j, kp) + D(ip, j, km) + D(im, jp, k) +
                                 D(im, jm, k) + D(im, j, kp) + D(im, j, km) + D(i, jp, kp) + D(i, jp, km) +
                                 D(i, jm, kp) + D(i, jm, km);
                    const R s3 = D(ip, jp, kp) + D(ip, jp, km) + D(ip, jm, kp) + D(ip, jm, km) +
                                 D(im, jp, kp) + D(im, jp, km) + D(im, jm, kp) + D(im, jm, km);
                    out = s1 * d1 + s2 * d2 + s3 * d3;
                }
                Df[(int64_t(i) * N2 + j) * N3 + k] = out;
            }
        }
    }
}

// p-Huber TV cost: aTV_Huber_costTerm (src/total_variation.cu:80-216) summed over the array by
// anisotropicTotalVariation_cost (:746-800); h() above is the USE_P_HUBER Huber of :33-46.
double tv_cost(const double* f, int N1, int N2, int N3, float delta, float beta, float p)
{
    if (beta <= 0.0f)
        beta = 1.0f;
    if (delta < 1.0e-8f)
        delta = 1.0e-8f;
    const huber hb(delta, p);
    // neighbour weights formed in float as the kernels form them (dist_1..3 = w * beta)
    const double d1 = (double)(1.0f * beta), d2 = (double)(0.7071067811865475f * beta),
                 d3 = (double)(0.5773502691896258f * beta);
    double acc = 0.0;
    for (int i = 0; i < N1; i++)
        for (int j = 0; j < N2; j++)
            for (int k = 0; k < N3; k++) {
                const double c0 = f[(int64_t(i) * N2 + j) * N3 + k];
                for (int di = -1; di <= 1; di++)
                    for (int dj = -1; dj <= 1; dj++)
                        for (int dk = -1; dk <= 1; dk++) {
                            if (N1 == 1 && di != 0)
                                continue;
                            const int m = std::abs(di) + std::abs(dj) + std::abs(dk);
                            if (m == 0)
                                continue;
                            const int a = std::min(N1 - 1, std::max(0, i + di));
                            const int b = std::min(N2 - 1, std::max(0, j + dj));
                            const int c = std::min(N3 - 1, std::max(0, k + dk));
                            const double w = (m == 1) ? d1 : (m == 2 ? d2 : d3);
                            acc += w * hb.h(c0 - f[(int64_t(a) * N2 + b) * N3 + c]);
                        }
            }
    return acc;
}

// Quadratic form <d, R''(f) d> of the p-Huber TV prior: aTV_Huber_quadFormTerm (src/total_variation.cu:
// 218-417) summed over the array by anisotropicTotalVariation_quadraticForm (:662-744).  Each neighbour
// pair term is w * h''(f_i - f_j) * (d_i - d_j)^2 with the surrogate curvature dd() above, over the same
// 26 (8 in-plane when N1 == 1) border-clamped neighbours as the cost; beta <= 0 -> 1, delta >= 1e-8.
double tv_quadform(const double* f, const double* d, int N1, int N2, int N3, float delta, float beta, float p)
{
    if (beta <= 0.0f)
        beta = 1.0f;
    if (delta < 1.0e-8f)
        delta = 1.0e-8f;
    const huber hb(delta, p);
    // neighbour weights formed in float as the kernels form them (dist_1..3 = w * beta)
    const double d1 = (double)(1.0f * beta), d2 = (double)(0.7071067811865475f * beta),
                 d3 = (double)(0.5773502691896258f * beta);
    double acc = 0.0;
    for (int i = 0; i < N1; i++)
        for (int j = 0; j < N2; j++)
            for (int k = 0; k < N3; k++) {
                const int64_t c = (int64_t(i) * N2 + j) * N3 + k;
                for (int di = -1; di <= 1; di++)
                    for (int dj = -1; dj <= 1; dj++)
                        for (int dk = -1; dk <= 1; dk++) {
                            if (N1 == 1 && di != 0)
                                continue;
                            const int m = std::abs(di) + std::abs(dj) + std::abs(dk);
                            if (m == 0)
                                continue;
                            const int a = std::min(N1 - 1, std::max(0, i + di));
                            const int b = std::min(N2 - 1, std::max(0, j + dj));
                            const int e = std::min(N3 - 1, std::max(0, k + dk));
                            const int64_t o = (int64_t(a) * N2 + b) * N3 + e;
                            const double w = (m == 1) ? d1 : (m == 2 ? d2 : d3);
                            const double dt = d[c] - d[o];
                            acc += w * hb.dd(f[c] - f[o]) * dt * dt;
                        }
            }
    return acc;
}

// Gaussian smoothing of GaussianFilter (src/relative_differences.cu:753-800) with the kernels at
// :314-420.  numDims 3 smooths along all three axes, 2 along the last two, 1 along the first.  Parameters
// are formed in float as the kernels form them: sigma = FWHM / (2 sqrt(2 ln 2)), radius =
// ceil(sqrt(2 ln 10) sigma) and the exponent factor 1 / (2 sigma) (sigma, not sigma^2).  A tap at offset t
// reads the border-clamped sample and is weighted by exp(-factor * s^2), s = distance to that clamped
// sample (not t); the sum is divided by the sum of the weights.
void gaussian_filter(const double* f, double* out, int N1, int N2, int N3, float FWHM, int numDims)
{
    const float sigma = FWHM / (2.0f * std::sqrt(2.0f * std::log(2.0f)));
    const int R = int(std::ceil(std::sqrt(2.0f * std::log(10.0f)) * sigma));
    const double den = (double)(1.0f / (2.0f * sigma));
    const int Ri = (numDims == 2) ? 0 : R;
    const int Rj = (numDims == 1) ? 0 : R;
    const int Rk = (numDims == 1) ? 0 : R;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < N1; i++)
        for (int j = 0; j < N2; j++)
            for (int k = 0; k < N3; k++) {
                double val = 0.0, sum = 0.0;
                for (int di = -Ri; di <= Ri; di++) {
                    const int a = std::max(0, std::min(i + di, N1 - 1));
                    for (int dj = -Rj; dj <= Rj; dj++) {
                        const int b = std::max(0, std::min(j + dj, N2 - 1));
                        for (int dk = -Rk; dk <= Rk; dk++) {
                            const int c = std::max(0, std::min(k + dk, N3 - 1));
                            const double s2 = double((i - a) * (i - a) + (j - b) * (j - b) + (k - c) * (k - c));
                            const double w = std::exp(-den * s2);
                            val += w * f[(int64_t(a) * N2 + b) * N3 + c];
                            sum += w;
                        }
                    }
                }
                out[(int64_t(i) * N2 + j) * N3 + k] = val / sum;
            }
}

// ------------------------------------------------------------------------------------------------
// Analytic ray tracing through geometric solids: the line integrals of ray_trace and the stopping power of
// the scintillator modules (src/analytic_ray_tracing_gpu.cu:48-862, the GPU path the reference runs), with
// the solids built as geometricObject::init builds them (src/phantom.cpp:502-640).  The float / double mix
// of the kernels is kept: every quantity is formed in the type the reference forms it in.
// ------------------------------------------------------------------------------------------------
const float OOB = 1.0e12f;

bool solid_init(int type, const float* c, const float* r, float val, const float* A, const float* clip, orc_solid* o)
{
    std::memset(o, 0, sizeof(*o));
    o->type = -1;
    if (c == nullptr || r == nullptr || type < 0 || type > 7)
        return false;
    o->type = type;
    for (int i = 0; i < 3; i++) {
        o->c[i] = c[i];
        o->r[i] = r[i];
    }
    if (clip != nullptr) {
        o->nclip = 0;
        for (int i = 0; i < 3; i++)
            if (clip[i] != 0.0f) {
                o->clip[o->nclip][i] = -1.0f;
                o->clip[o->nclip][3] = -clip[i];
                o->nclip += 1;
            }
    }
    o->val = val;
    o->rotated = 0;
    if (A != nullptr)
        for (int i = 0; i < 9; i++) {
            o->A[i] = A[i];
            if (i == 0 || i == 4 || i == 8) {
                if (std::fabs(double(A[i]) - 1.0) > 1.0e-8)
                    o->rotated = 1;
            } else if (std::fabs(double(A[i])) > 1.0e-8) {
                o->rotated = 1;
            }
        }
    // cones: axis `ax` with end radii r_1, r_2 and half length l become the unit cone about the apex
    auto cone = [&](int ax, int i1, int i2, int il) {
        const double l = o->r[il], r_1 = o->r[i1], r_2 = o->r[i2];
        const int o1 = (ax + 1) % 3, o2 = (ax + 2) % 3;
        o->r[o1] = float(std::fabs(r_2 - r_1) / (2.0 * l));
        o->r[o2] = o->r[o1];
        o->r[ax] = 1.0f;
        o->c[ax] = float(o->c[ax] - l * (r_2 + r_1) / (r_2 - r_1));
        o->cone[0] = float(-l + l * (r_2 + r_1) / (r_2 - r_1));
        o->cone[1] = float(l + l * (r_2 + r_1) / (r_2 - r_1));
    };
    if (type == 5)
        cone(0, 1, 2, 0);  // CONE_X: radii = (l, r_1, r_2)
    else if (type == 6)
        cone(1, 0, 2, 1);  // CONE_Y: radii = (r_1, l, r_2)
    else if (type == 7)
        cone(2, 0, 1, 2);  // CONE_Z: radii = (r_1, r_2, l)
    return true;
}

struct f2 {
    float x, y;
};
struct d3 {
    double x, y, z;
};

bool p1d(f2& ts, float p, float r)
{
    if (std::fabs(r) < 1e-12f) {
        if (std::fabs(p) < 1.0f) {
            ts.x = -OOB;
            ts.y = OOB;
            return true;
        }
        return false;
    }
    if (r > 0.0f) {
        ts.x = (-1.0f - p) / r;
        ts.y = (1.0f - p) / r;
    } else {
        ts.y = (-1.0f - p) / r;
        ts.x = (1.0f - p) / r;
    }
    return true;
}

bool clip_planes(f2& ts, const float* p, const float* r, const orc_solid& s)
{
    for (int i = 0; i < s.nclip; i++) {
        const float p_dot_n = s.clip[i][0] * p[0] + s.clip[i][1] * p[1] + s.clip[i][2] * p[2];
        const float r_dot_n = s.clip[i][0] * r[0] + s.clip[i][1] * r[1] + s.clip[i][2] * r[2];
        if (std::fabs(r_dot_n) < 1.0e-12f) {
            if (p_dot_n < s.clip[i][3])
                return false;
        } else if (r_dot_n > 0.0f) {
            const float temp = (s.clip[i][3] - p_dot_n) / r_dot_n;
            if (ts.y < temp)
                return false;
            if (temp > ts.x)
                ts.x = temp;
        } else {
            const float temp = (s.clip[i][3] - p_dot_n) / r_dot_n;
            if (ts.x > temp)
                return false;
            if (ts.y > temp)
                ts.y = temp;
        }
    }
    return true;
}

// intersectionEndPoints_centeredAndNormalized (src/analytic_ray_tracing_gpu.cu:150-463)
bool centered(const d3& p, const d3& r, f2& ts, const orc_solid& s)
{
    ts.x = -OOB;
    ts.y = ts.x;
    const double r_dot_r = r.x * r.x + r.y * r.y + r.z * r.z;
    const double p_dot_r = r.x * p.x + r.y * p.y + r.z * p.z;
    if (s.type == 0) {
        double disc = p_dot_r * p_dot_r + r_dot_r * (1.0 - (p.x * p.x + p.y * p.y + p.z * p.z));
        if (!(disc > 0.0))
            return false;
        disc = std::sqrt(disc);
        ts.x = float((-p_dot_r - disc) / r_dot_r);
        ts.y = float((-p_dot_r + disc) / r_dot_r);
    } else if (s.type == 1) {
        f2 tx, ty, tz;
        if (!p1d(tx, float(p.x), float(r.x)) || !p1d(ty, float(p.y), float(r.y)) || !p1d(tz, float(p.z), float(r.z)))
            return false;
        ts.x = std::max(std::max(tx.x, ty.x), tz.x);
        ts.y = std::min(std::min(tx.y, ty.y), tz.y);
    } else if (s.type >= 2 && s.type <= 4) {
        // cylinders: ax = axis; the cross section is the unit disc in the other two coordinates
        const int ax = s.type - 2;  // 2: x, 3: y, 4: z
        const double pa = ax == 0 ? p.x : (ax == 1 ? p.y : p.z);
        const double ra = ax == 0 ? r.x : (ax == 1 ? r.y : r.z);
        double rr2, pr2, pp2, rz0, rz1;
        if (ax == 2) {
            rr2 = r_dot_r - r.z * r.z;
            pr2 = p_dot_r - p.z * r.z;
            pp2 = p.x * p.x + p.y * p.y;
            rz0 = r.x;
            rz1 = r.y;
        } else if (ax == 0) {
            rr2 = r_dot_r - r.x * r.x;
            pr2 = p.z * r.z + p.y * r.y;
            pp2 = p.z * p.z + p.y * p.y;
            rz0 = r.y;
            rz1 = r.z;
        } else {
            rr2 = r_dot_r - r.y * r.y;
            pr2 = p.x * r.x + p.z * r.z;
            pp2 = p.x * p.x + p.z * p.z;
            rz0 = r.x;
            rz1 = r.z;
        }
        double disc = pr2 * pr2 - rr2 * (pp2 - 1.0);
        if (disc > 0.0) {
            disc = std::sqrt(disc);
            const float tmin = float((-pr2 - disc) / rr2);
            const float tmax = float((-pr2 + disc) / rr2);
            f2 tz;
            if (!p1d(tz, float(pa), float(ra)))
                return false;
            ts.x = std::max(tmin, tz.x);
            ts.y = std::min(tmax, tz.y);
        } else if (rz0 == 0.0 && rz1 == 0.0 && pp2 <= 1.0) {
            return p1d(ts, float(pa), float(ra));  // returns without the final ordering test, as the reference
        } else {
            return false;
        }
    } else if (s.type == 5 || s.type == 7) {
        // CONE_X, CONE_Z: x'^2 + y'^2 = a'^2 about the axis, clipped along it by cone[]
        const bool zc = s.type == 7;
        const double pa = zc ? p.z : p.x, ra = zc ? r.z : r.x;
        const double pb = zc ? p.x : p.z, rb = zc ? r.x : r.z;
        double a = rb * rb + r.y * r.y - ra * ra;
        const double b_half = pb * rb + p.y * r.y - pa * ra;
        const double c = pb * pb + p.y * p.y - pa * pa;
        double disc = b_half * b_half - a * c;
        if (!(disc > 0.0))
            return false;
        disc = std::sqrt(disc);
        float tmin = float((-b_half - disc) / a);
        float tmax = float((-b_half + disc) / a);
        if (tmin > tmax) {
            a = tmin;
            tmin = tmax;
            tmax = float(a);
        }
        const double shift = 0.5 * (double(s.cone[1]) + double(s.cone[0]));
        const double scale = 0.5 * (double(s.cone[1]) - double(s.cone[0]));
        f2 tz;
        if (!p1d(tz, float((pa - shift) / scale), float(ra / scale)))
            return false;
        ts.x = std::max(tmin, tz.x);
        ts.y = std::min(tmax, tz.y);
    } else if (s.type == 6) {
        // CONE_Y, with the end-cap handling of src/analytic_ray_tracing_gpu.cu:353-451
        double a = r.x * r.x + r.z * r.z - r.y * r.y;
        const double b_half = p.x * r.x + p.z * r.z - p.y * r.y;
        const double c = p.x * p.x + p.z * p.z - p.y * p.y;
        double disc = b_half * b_half - a * c;
        if (disc > 0.0) {
            disc = std::sqrt(disc);
            double tmin = (-b_half - disc) / a;
            double tmax = (-b_half + disc) / a;
            if (tmin > tmax) {
                a = tmin;
                tmin = tmax;
                tmax = a;
            }
            const double shift = 0.5 * (double(s.cone[1]) + double(s.cone[0]));
            const double scale = 0.5 * (double(s.cone[1]) - double(s.cone[0]));
            f2 tz;
            if (p1d(tz, float((p.y - shift) / scale), float(r.y / scale))) {
                if (std::fabs(r.y) > 1.0e-12f) {
                    auto inside = [&](float t) {
                        double xv = p.x + t * r.x, yv = p.y + t * r.y, zv = p.z + t * r.z;
                        return xv * xv + zv * zv <= yv * yv;
                    };
                    const bool in0 = inside(tz.x), in1 = inside(tz.y);
                    if (in0) {
                        if (in1) {
                            ts.x = tz.x;
                            ts.y = tz.y;
                        } else if (tz.x <= tmin && tmin <= tz.y) {
                            ts.x = tz.x;
                            ts.y = float(tmin);
                        } else {
                            ts.x = tz.x;
                            ts.y = float(tmax);
                        }
                    } else {
                        if (in1) {
                            if (tz.x <= tmin && tmin <= tz.y) {
                                ts.x = float(tmin);
                                ts.y = tz.y;
                            } else {
                                ts.x = float(tmax);
                                ts.y = tz.y;
                            }
                        } else if (tz.x <= tmin && tmax <= tz.y) {
                            ts.x = float(tmin);
                            ts.y = float(tmax);
                        } else {
                            ts.x = -OOB;
                            ts.y = ts.x;
                            return false;
                        }
                    }
                } else {
                    ts.x = float(tmin);
                    ts.y = float(tmax);
                }
            }
        }
    } else {
        ts.x = -OOB;
        ts.y = -OOB;
        return false;
    }
    if (ts.x >= ts.y) {
        ts.x = -OOB;
        ts.y = ts.x;
        return false;
    }
    return true;
}

// intersectionEndPoints (src/analytic_ray_tracing_gpu.cu:465-518)
bool intersect(const float* p, const float* r, f2& ts, const orc_solid& s)
{
    d3 q, mr;
    if (!s.rotated) {
        q.x = (p[0] - s.c[0]) / s.r[0];
        q.y = (p[1] - s.c[1]) / s.r[1];
        q.z = (p[2] - s.c[2]) / s.r[2];
        mr.x = r[0] / s.r[0];
        mr.y = r[1] / s.r[1];
        mr.z = r[2] / s.r[2];
    } else {
        q.x = p[0] - s.c[0];
        q.y = p[1] - s.c[1];
        q.z = p[2] - s.c[2];
        const double tx = (q.x * s.A[0] + q.y * s.A[1] + q.z * s.A[2]) / s.r[0];
        const double ty = (q.x * s.A[3] + q.y * s.A[4] + q.z * s.A[5]) / s.r[1];
        const double tz = (q.x * s.A[6] + q.y * s.A[7] + q.z * s.A[8]) / s.r[2];
        q.x = tx;
        q.y = ty;
        q.z = tz;
        mr.x = (r[0] * s.A[0] + r[1] * s.A[1] + r[2] * s.A[2]) / s.r[0];
        mr.y = (r[0] * s.A[3] + r[1] * s.A[4] + r[2] * s.A[5]) / s.r[1];
        mr.z = (r[0] * s.A[6] + r[1] * s.A[7] + r[2] * s.A[8]) / s.r[2];
    }
    if (!centered(q, mr, ts, s))
        return false;
    if (!clip_planes(ts, p, r, s)) {
        ts.x = -OOB;
        ts.y = ts.x;
        return false;
    }
    return ts.y > ts.x;
}

// stoppingPower_geometricSolids / lineIntegral_geometricSolids (src/analytic_ray_tracing_gpu.cu:536-691):
// the union of the solids' chords cut at every end point; each piece belongs to the last-listed solid
// that contains its mid point
float ray_value(const float* p, const float* r, const orc_solid* s, int n, bool stopping)
{
    std::vector<float> ends;
    std::vector<float> t0(n), t1(n);
    std::vector<int> idx;
    for (int i = 0; i < n; i++) {
        f2 ts;
        if (intersect(p, r, ts, s[i])) {
            ends.push_back(ts.x);
            ends.push_back(ts.y);
            t0[i] = ts.x;
            t1[i] = ts.y;
            idx.push_back(i);
        } else {
            t0[i] = OOB;
            t1[i] = OOB;
        }
    }
    float fwd = 0.0f, bwd = 0.0f, all = 0.0f;
    const int count = int(idx.size());
    if (count > 0) {
        std::sort(ends.begin(), ends.end());
        for (int i = 0; i < 2 * count - 1; i++) {
            const float mid = (ends[i + 1] + ends[i]) * 0.5f;
            for (int ind = count - 1; ind >= 0; ind--) {
                const int j = idx[ind];
                if (t0[j] <= mid && mid <= t1[j]) {
                    const float piece = s[j].val * (ends[i + 1] - ends[i]);
                    all += piece;
                    if (mid >= 0.0f)
                        fwd += piece;
                    else
                        bwd += piece;
                    break;
                }
            }
        }
    }
    if (!stopping)
        return all;
    return (1.0f - std::exp(-fwd)) * (1.0f - std::exp(-bwd));
}

// rayTracingKernel / stoppingPowerKernel (src/analytic_ray_tracing_gpu.cu:693-862) for one planogram
void ray_trace_plan(const orc_plan& P, const orc_solid* s, int n, int oversampling, bool stopping, float* g,
                    int nthreads)
{
    oversampling = std::max(1, std::min(oversampling, 11));
    if (oversampling % 2 == 0)
        oversampling += 1;
    oversampling = std::max(1, std::min(oversampling, 11));
    const float cs = float(std::cos(double(P.psi))), sn = float(std::sin(double(P.psi)));
    const int64_t plane = int64_t(P.N_u1) * P.N_u0;
#pragma omp parallel for collapse(2) schedule(dynamic) num_threads(nthreads)
    for (int iv1 = 0; iv1 < P.N_v1; iv1++)
        for (int iv0 = 0; iv0 < P.N_v0; iv0++) {
            const float v1 = float(iv1) * P.T_v1 + P.v1_0;
            const float v0 = float(iv0) * P.T_v0 + P.v0_0;
            float* out = g + (int64_t(iv1) * P.N_v0 + iv0) * plane;
            // trajectory: r / rsqrt(r.r), i.e. r scaled by |r| (not normalised), as the kernel computes it
            float rv[3] = {-(v0 * cs + sn), -(v0 * sn - cs), -v1};
            const float mag = 1.0f / std::sqrt(rv[0] * rv[0] + rv[1] * rv[1] + rv[2] * rv[2]);
            rv[0] = rv[0] / mag;
            rv[1] = rv[1] / mag;
            rv[2] = rv[2] / mag;
            for (int iu1 = 0; iu1 < P.N_u1; iu1++)
                for (int iu0 = 0; iu0 < P.N_u0; iu0++) {
                    const float u1 = float(iu1) * P.T_u1 + P.u1_0;
                    const float u0 = float(iu0) * P.T_u0 + P.u0_0;
                    float& o = out[int64_t(iu1) * P.N_u0 + iu0];
                    // the u0 test uses T_v0 for the bin half width (the kernel reads T.x)
                    if (std::fabs(u1) - 0.5f * P.T_u1 > P.H - P.R * std::fabs(v1) ||
                        std::fabs(u0) - 0.5f * P.T_v0 > P.L + P.R * std::fabs(v0)) {
                        o = 0.0f;
                        continue;
                    }
                    if (oversampling <= 1) {
                        const float pos[3] = {u0 * cs, u0 * sn, u1};
                        o = ray_value(pos, rv, s, n, stopping);
                        continue;
                    }
                    const float T_v_os = P.T_u1 / float(oversampling + 1);
                    const float T_u_os = P.T_u0 / float(oversampling + 1);
                    const int rad = (oversampling - 1) / 2;
                    float acc = 0.0f;
                    for (int j = -rad; j <= rad; j++) {
                        const float du1 = j * T_v_os;
                        for (int k = -rad; k <= rad; k++) {
                            const float du0 = k * T_u_os;
                            const float pos[3] = {(u0 + du0) * cs, (u0 + du0) * sn, u1 + du1};
                            acc += ray_value(pos, rv, s, n, stopping);
                        }
                    }
                    o = acc / float(oversampling * oversampling);
                }
        }
}

}  // namespace

// ------------------------------------------------------------------------------------------------
// C entry points (loaded with ctypes by the tests)
// ------------------------------------------------------------------------------------------------
// Work measure of one projection: the number of (voxel, bin) pairs with a non-zero footprint weight over v1
// rows [v1_begin, v1_end), unmasked bins only (the visits of the forward loop above,
// src/projectors.cu:79-183, that add a non-zero term), with the weights formed in fp32 (f64 = 0, the GPU's
// arithmetic) or FP64 (f64 = 1; a sliver overlap that rounds to zero in fp32 still counts).
template <typename R>
static int64_t count_taps(const orc_plan& P, const orc_vol& V, int v1_begin, int v1_end, int nthreads)
{
    const int NV = (v1_end - v1_begin) * P.N_v0;
    int64_t total = 0;
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads) reduction(+ : total)
    for (int vv = 0; vv < NV; vv++) {
        const int iv1 = v1_begin + vv / P.N_v0, iv0 = vv % P.N_v0;
        int64_t n = 0;
        for (int iu1 = 0; iu1 < P.N_u1; iu1++)
            for (int iu0 = 0; iu0 < P.N_u0; iu0++)
                if (!bin_masked(P, iv1, iv0, iu1, iu0))
                    enumerate_bin<R>(P, V, iv1, iv0, iu1, iu0, [&](int64_t, R w0, R w1) {
                        n += (w0 * w1 != R(0)) ? 1 : 0;
                    });
        total += n;
    }
    return total;
}

extern "C" {

int orc_plan_init(float psi_deg, float D, float L, float H, float v_m0, float v_m1, float T, orc_plan* out)
{
    return plan_init(psi_deg, D, L, H, v_m0, v_m1, T, out);
}

void orc_vol_init(int nx, int ny, int nz, float vw, float vh, float ox, float oy, float oz, orc_vol* v)
{
    vol_init(nx, ny, nz, vw, vh, ox, oy, oz, v);
}

int orc_bin_masked(const orc_plan* P, int iv1, int iv0, int iu1, int iu0)
{
    return bin_masked(*P, iv1, iv0, iu1, iu0) ? 1 : 0;
}

void orc_project_f32(const orc_plan* P, const orc_vol* V, const float* f, float* g, int nthreads)
{
    project_plan<float>(*P, *V, f, g, nthreads);
}

void orc_project_f64(const orc_plan* P, const orc_vol* V, const double* f, double* g, int nthreads)
{
    project_plan<double>(*P, *V, f, g, nthreads);
}

// mode 0 = PETL (reference semantics), 1 = matched (exact transpose).  Accumulates into f.
void orc_backproject_f32(const orc_plan* P, const orc_vol* V, const float* g, float* f, int mode, int nthreads)
{
    if (mode == 1)
        backproject_matched_plan<float>(*P, *V, g, f, nthreads);
    else
        backproject_petl_plan<float>(*P, *V, g, f, nthreads);
}

void orc_backproject_f64(const orc_plan* P, const orc_vol* V, const double* g, double* f, int mode, int nthreads)
{
    if (mode == 1)
        backproject_matched_plan<double>(*P, *V, g, f, nthreads);
    else
        backproject_petl_plan<double>(*P, *V, g, f, nthreads);
}

// Taps of one bin (for brute-force / point-source checks).  Returns the tap count; writes at most cap.
int64_t orc_count_taps(const orc_plan* P, const orc_vol* V, int v1_begin, int v1_end, int f64, int nthreads)
{
    return f64 ? count_taps<double>(*P, *V, v1_begin, v1_end, nthreads)
               : count_taps<float>(*P, *V, v1_begin, v1_end, nthreads);
}

int orc_bin_taps(const orc_plan* P, const orc_vol* V, int iv1, int iv0, int iu1, int iu0, int64_t* idx,
                 double* w, int cap, double* scale)
{
    int n = 0;
    if (bin_masked(*P, iv1, iv0, iu1, iu0)) {
        *scale = 0.0;
        return 0;
    }
    const double l_phi = enumerate_bin<double>(*P, *V, iv1, iv0, iu1, iu0, [&](int64_t k, double w0, double w1) {
        if (n < cap) {
            idx[n] = k;
            w[n] = w0 * w1;
        }
        n++;
    });
    *scale = l_phi * obliquity<double>(*P, iv1, iv0);
    return n;
}

void orc_rd_gradient_f32(const float* f, float* Df, int N1, int N2, int N3, float delta, float beta)
{
    rd_gradient<float>(f, Df, N1, N2, N3, delta, beta);
}

void orc_rd_gradient_f64(const double* f, double* Df, int N1, int N2, int N3, double delta, double beta)
{
    rd_gradient<double>(f, Df, N1, N2, N3, delta, beta);
}

double orc_rd_cost_f64(const double* f, int N1, int N2, int N3, double delta, double beta)
{
    return rd_cost(f, N1, N2, N3, delta, beta);
}

void orc_tv_gradient_f32(const float* f, float* Df, int N1, int N2, int N3, float delta, float beta, float p)
{
    tv_gradient<float>(f, Df, N1, N2, N3, delta, beta, p);
}

void orc_tv_gradient_f64(const double* f, double* Df, int N1, int N2, int N3, float delta, float beta, float p)
{
    tv_gradient<double>(f, Df, N1, N2, N3, delta, beta, p);
}

double orc_tv_cost_f64(const double* f, int N1, int N2, int N3, float delta, float beta, float p)
{
    return tv_cost(f, N1, N2, N3, delta, beta, p);
}

double orc_tv_quadform_f64(const double* f, const double* d, int N1, int N2, int N3, float delta, float beta,
                           float p)
{
    return tv_quadform(f, d, N1, N2, N3, delta, beta, p);
}

void orc_gaussian_filter_f64(const double* f, double* out, int N1, int N2, int N3, float FWHM, int numDims)
{
    gaussian_filter(f, out, N1, N2, N3, FWHM, numDims);
}

int orc_solid_init(int type, const float* c, const float* r, float val, const float* A, const float* clip,
                   orc_solid* out)
{
    return solid_init(type, c, r, val, A, clip, out) ? 1 : 0;
}

// chord end points of one ray through one solid; returns 1 when it intersects
int orc_intersect(const orc_solid* s, const float* p, const float* r, float* ts)
{
    f2 t;
    const bool hit = intersect(p, r, t, *s);
    ts[0] = t.x;
    ts[1] = t.y;
    return hit ? 1 : 0;
}

float orc_ray_value(const orc_solid* s, int n, const float* p, const float* r, int stopping)
{
    return ray_value(p, r, s, n, stopping != 0);
}

void orc_ray_trace(const orc_plan* P, const orc_solid* s, int n, int oversampling, int stopping, float* g,
                   int nthreads)
{
    ray_trace_plan(*P, s, n, oversampling, stopping != 0, g, nthreads);
}

int orc_max_threads(void) { return omp_get_max_threads(); }

// Forward / PETL-mode back projection restricted to v1 rows [v1_begin, v1_end) (bounded timing samples).
void orc_project_rows_f32(const orc_plan* P, const orc_vol* V, const float* f, float* g, int v1_begin, int v1_end,
                          int nthreads)
{
    project_plan<float>(*P, *V, f, g, nthreads, v1_begin, v1_end);
}

// The reference's texture-fetch form of the PETL-mode back projection (quant: 0 exact fractions, 1 / 2
// fractions rounded / truncated to 8 fractional bits), v1 rows [v1_begin, v1_end), accumulated into f.
void orc_backproject_petl_tex_f32(const orc_plan* P, const orc_vol* V, const float* g, float* f, int quant,
                                  int v1_begin, int v1_end, int nthreads)
{
    backproject_petl_tex(*P, *V, g, f, nthreads, quant, v1_begin, v1_end);
}

void orc_backproject_petl_rows_f32(const orc_plan* P, const orc_vol* V, const float* g, float* f, int v1_begin,
                                   int v1_end, int nthreads)
{
    backproject_petl_plan<float>(*P, *V, g, f, nthreads, v1_begin, v1_end);
}

// OSEM subsets: v1 rows first, first + stride, ...
void orc_project_subset_f32(const orc_plan* P, const orc_vol* V, const float* f, float* g, int first, int stride,
                            int nthreads)
{
    project_plan<float>(*P, *V, f, g, nthreads, first, P->N_v1, stride);
}

void orc_backproject_petl_subset_f32(const orc_plan* P, const orc_vol* V, const float* g, float* f, int first,
                                     int stride, int nthreads)
{
    backproject_petl_plan<float>(*P, *V, g, f, nthreads, first, P->N_v1, stride);
}

// ------------------------------------------------------------------------------------------------
// List-mode binning.  Per coincidence (blue and pink end points, already in mm):
//   r = blue - pink; acceptance and (v1, v0) of src/planogram.cpp:113-128; (u1, u0) from the blue end
//   point, src/planogram.cpp:130-141; chest-wall origin: u1 -= H (src/binning.cpp:636-640).
// Then the flush of src/binning.cpp:662-750: nearest v bins by int(x + 0.5) (x + 0.5 in double,
// truncation toward zero), and in u either nearest (binning type 0) or the clamped bilinear split with
// weights formed in double and added to the float bins one at a time, in event order.
// Returns the number of accepted coincidences (src/binning.cpp:643 counts them).
// ------------------------------------------------------------------------------------------------
// (v1, v0, u1, u0) of one coincidence; false when the planogram does not accept it
static bool event_coords(const orc_plan& p, const float* b, const float* k, float* out)
{
    const float r0 = b[0] - k[0], r1 = b[1] - k[1], r2 = b[2] - k[2];
    const float c = -p.sin_psi * r0 + p.cos_psi * r1;
    if (c == 0.0f)
        return false;
    const float v0 = -(p.cos_psi * r0 + p.sin_psi * r1) / c;
    if (std::fabs(v0) > p.v_m0)
        return false;
    const float v1 = -r2 / c;
    if (std::fabs(v1) > p.v_m1)
        return false;
    const float t = b[1] * p.cos_psi - b[0] * p.sin_psi;
    out[0] = v1;
    out[1] = v0;
    out[2] = b[2] + v1 * t;
    out[3] = b[0] * p.cos_psi + b[1] * p.sin_psi + v0 * t;
    return true;
}

int orc_event_coords(const orc_plan* P, const float* blue, const float* pink, float* out)
{
    return event_coords(*P, blue, pink, out) ? 1 : 0;
}

int64_t orc_bin_events(const orc_plan* P, const float* blue, const float* pink, int64_t n, int nearest,
                       int chest_wall, float* g)
{
    const orc_plan& p = *P;
    const int64_t proj = int64_t(p.N_u1) * p.N_u0;
    const int64_t obl = int64_t(p.N_v0) * proj;
    int64_t accepted = 0;
    for (int64_t e = 0; e < n; e++) {
        float co[4];
        if (!event_coords(p, blue + 3 * e, pink + 3 * e, co))
            continue;
        const float v1 = co[0], v0 = co[1], u0 = co[3];
        float u1 = co[2];
        if (chest_wall)
            u1 -= p.H;
        accepted++;
        const float iv1_val = (v1 - p.v1_0) / p.T_v1;
        const int iv1 = int(double(iv1_val) + 0.5);
        if (iv1 < 0 || iv1 >= p.N_v1)
            continue;
        const int iv0 = int(double((v0 - p.v0_0) / p.T_v0) + 0.5);
        if (iv0 < 0 || iv0 >= p.N_v0)
            continue;
        float* pr = g + int64_t(iv1) * obl + int64_t(iv0) * proj;
        const float w = 1.0f;
        if (nearest) {
            const int a = int(double((u1 - p.u1_0) / p.T_u1) + 0.5);
            const int q = int(double((u0 - p.u0_0) / p.T_u0) + 0.5);
            if (a >= 0 && a < p.N_u1 && q >= 0 && q < p.N_u0)
                pr[int64_t(a) * p.N_u0 + q] += w;
            continue;
        }
        int lo1, hi1, lo0, hi0;
        float du1, du0;
        const float iu1 = (u1 - p.u1_0) / p.T_u1;
        if (iu1 <= 0.0f) { lo1 = hi1 = 0; du1 = 0.0f; }
        else if (iu1 >= float(p.N_u1 - 1)) { lo1 = hi1 = p.N_u1 - 1; du1 = 0.0f; }
        else { lo1 = int(iu1); hi1 = lo1 + 1; du1 = iu1 - float(lo1); }
        const float iu0 = (u0 - p.u0_0) / p.T_u0;
        if (iu0 <= 0.0f) { lo0 = hi0 = 0; du0 = 0.0f; }
        else if (iu0 >= float(p.N_u0 - 1)) { lo0 = hi0 = p.N_u0 - 1; du0 = 0.0f; }
        else { lo0 = int(iu0); hi0 = lo0 + 1; du0 = iu0 - float(lo0); }
        auto add = [&](int a, int q, double wt) {
            float& x = pr[int64_t(a) * p.N_u0 + q];
            x = float(double(x) + wt);
        };
        add(lo1, lo0, (1.0 - double(du1)) * (1.0 - double(du0)) * double(w));
        add(hi1, lo0, double(du1) * (1.0 - double(du0)) * double(w));
        add(lo1, hi0, (1.0 - double(du1)) * double(du0) * double(w));
        add(hi1, hi0, double(du1) * double(du0) * double(w));
    }
    return accepted;
}

}  // extern "C"
