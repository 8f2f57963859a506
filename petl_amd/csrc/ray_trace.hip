// Analytic ray tracing through geometric solids (line integrals and scintillator stopping power).
//
// Reference: src/phantom.cpp:502-640 (geometricObject::init), :1120-1141 (scale_phantom),
// src/analytic_ray_tracing_gpu.cu:48-862 (the kernels and their device functions).  The float / double
// mix of the reference's device code is kept (each quantity is formed in the type it is formed in there),
// with contraction off.
//
// Layout: one thread per bin, the solids read through uniform (scalar) loads; a solid whose bounding
// sphere the ray misses by a clear margin is skipped before the exact test (it could not report a chord).
// The reference gives every thread 4n floats and n ints of global scratch for its sorted end points; here
// a thread keeps the solids it hits (up to MAXHIT) in registers and sorts their end points there.  A ray that hits more solids walks
// the distinct end points in increasing order without storing them, recomputing the intersections it
// needs: the same pieces, in the same order, as sorting all end points.
#include <cmath>
#include "ray_trace.h"

namespace petl {

namespace {

constexpr float OOB = 1.0e12f;
constexpr int MAXHIT = 12;
constexpr int RT_THREADS = 128;

struct F2 {
    float x, y;
};
struct D3 {
    double x, y, z;
};

__device__ __forceinline__ bool p1d(F2& ts, float p, float r)
{
#pragma clang fp contract(off)
    if (fabsf(r) < 1e-12f) {
        if (fabsf(p) < 1.0f) {
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

__device__ __forceinline__ bool clip_planes(F2& ts, const float* p, const float* r, const Solid& s)
{
#pragma clang fp contract(off)
    for (int i = 0; i < s.nclip; i++) {
        const float p_dot_n = s.clip[i][0] * p[0] + s.clip[i][1] * p[1] + s.clip[i][2] * p[2];
        const float r_dot_n = s.clip[i][0] * r[0] + s.clip[i][1] * r[1] + s.clip[i][2] * r[2];
        if (fabsf(r_dot_n) < 1.0e-12f) {
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

// chord of the unit solid in its own frame (src/analytic_ray_tracing_gpu.cu:150-463)
__device__ bool centered(const D3& p, const D3& r, F2& ts, const Solid& s)
{
#pragma clang fp contract(off)
    ts.x = -OOB;
    ts.y = ts.x;
    const double r_dot_r = r.x * r.x + r.y * r.y + r.z * r.z;
    const double p_dot_r = r.x * p.x + r.y * p.y + r.z * p.z;
    if (s.type == 0) {
        double disc = p_dot_r * p_dot_r + r_dot_r * (1.0 - (p.x * p.x + p.y * p.y + p.z * p.z));
        if (!(disc > 0.0))
            return false;
        disc = sqrt(disc);
        ts.x = float((-p_dot_r - disc) / r_dot_r);
        ts.y = float((-p_dot_r + disc) / r_dot_r);
    } else if (s.type == 1) {
        F2 tx, ty, tz;
        if (!p1d(tx, float(p.x), float(r.x)) || !p1d(ty, float(p.y), float(r.y)) || !p1d(tz, float(p.z), float(r.z)))
            return false;
        ts.x = fmaxf(fmaxf(tx.x, ty.x), tz.x);
        ts.y = fminf(fminf(tx.y, ty.y), tz.y);
    } else if (s.type >= 2 && s.type <= 4) {
        // cylinders along x (2), y (3), z (4): unit disc across, the 1D slab along the axis
        const int ax = s.type - 2;
        const double pa = ax == 0 ? p.x : (ax == 1 ? p.y : p.z);
        const double ra = ax == 0 ? r.x : (ax == 1 ? r.y : r.z);
        double rr2, pr2, pp2, q0, q1;
        if (ax == 2) {
            rr2 = r_dot_r - r.z * r.z;
            pr2 = p_dot_r - p.z * r.z;
            pp2 = p.x * p.x + p.y * p.y;
            q0 = r.x;
            q1 = r.y;
        } else if (ax == 0) {
            rr2 = r_dot_r - r.x * r.x;
            pr2 = p.z * r.z + p.y * r.y;
            pp2 = p.z * p.z + p.y * p.y;
            q0 = r.y;
            q1 = r.z;
        } else {
            rr2 = r_dot_r - r.y * r.y;
            pr2 = p.x * r.x + p.z * r.z;
            pp2 = p.x * p.x + p.z * p.z;
            q0 = r.x;
            q1 = r.z;
        }
        double disc = pr2 * pr2 - rr2 * (pp2 - 1.0);
        if (disc > 0.0) {
            disc = sqrt(disc);
            const float tmin = float((-pr2 - disc) / rr2);
            const float tmax = float((-pr2 + disc) / rr2);
            F2 tz;
            if (!p1d(tz, float(pa), float(ra)))
                return false;
            ts.x = fmaxf(tmin, tz.x);
            ts.y = fminf(tmax, tz.y);
        } else if (q0 == 0.0 && q1 == 0.0 && pp2 <= 1.0) {
            return p1d(ts, float(pa), float(ra));  // the reference returns here, before the ordering test
        } else {
            return false;
        }
    } else if (s.type == 5 || s.type == 7) {
        // cones along x (5) and z (7): a'^2 + b'^2 = axial'^2, the axial range from cone[]
        const bool zc = s.type == 7;
        const double pa = zc ? p.z : p.x, ra = zc ? r.z : r.x;
        const double pb = zc ? p.x : p.z, rb = zc ? r.x : r.z;
        double a = rb * rb + r.y * r.y - ra * ra;
        const double b_half = pb * rb + p.y * r.y - pa * ra;
        const double c = pb * pb + p.y * p.y - pa * pa;
        double disc = b_half * b_half - a * c;
        if (!(disc > 0.0))
            return false;
        disc = sqrt(disc);
        float tmin = float((-b_half - disc) / a);
        float tmax = float((-b_half + disc) / a);
        if (tmin > tmax) {
            a = tmin;
            tmin = tmax;
            tmax = float(a);
        }
        const double shift = 0.5 * (double(s.cone[1]) + double(s.cone[0]));
        const double scale = 0.5 * (double(s.cone[1]) - double(s.cone[0]));
        F2 tz;
        if (!p1d(tz, float((pa - shift) / scale), float(ra / scale)))
            return false;
        ts.x = fmaxf(tmin, tz.x);
        ts.y = fminf(tmax, tz.y);
    } else if (s.type == 6) {
        // cone along y with its end-cap cases (src/analytic_ray_tracing_gpu.cu:353-451)
        double a = r.x * r.x + r.z * r.z - r.y * r.y;
        const double b_half = p.x * r.x + p.z * r.z - p.y * r.y;
        const double c = p.x * p.x + p.z * p.z - p.y * p.y;
        double disc = b_half * b_half - a * c;
        if (disc > 0.0) {
            disc = sqrt(disc);
            double tmin = (-b_half - disc) / a;
            double tmax = (-b_half + disc) / a;
            if (tmin > tmax) {
                a = tmin;
                tmin = tmax;
                tmax = a;
            }
            const double shift = 0.5 * (double(s.cone[1]) + double(s.cone[0]));
            const double scale = 0.5 * (double(s.cone[1]) - double(s.cone[0]));
            F2 tz;
            if (p1d(tz, float((p.y - shift) / scale), float(r.y / scale))) {
                if (fabs(r.y) > 1.0e-12f) {
                    double xv = p.x + tz.x * r.x, yv = p.y + tz.x * r.y, zv = p.z + tz.x * r.z;
                    const bool in0 = xv * xv + zv * zv <= yv * yv;
                    xv = p.x + tz.y * r.x;
                    yv = p.y + tz.y * r.y;
                    zv = p.z + tz.y * r.z;
                    const bool in1 = xv * xv + zv * zv <= yv * yv;
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
                    } else if (in1) {
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

// true when the line p + t r certainly misses the solid's bounding sphere: the distance from the centre
// to the line exceeds the radius by a margin far above the rounding of this test and of the exact one
__device__ __forceinline__ bool culled(const float* p, const float* r, const Solid& s)
{
    if (s.br < 0.0f)
        return false;
    const float wx = s.bc[0] - p[0], wy = s.bc[1] - p[1], wz = s.bc[2] - p[2];
    const float ww = wx * wx + wy * wy + wz * wz;
    const float wr = wx * r[0] + wy * r[1] + wz * r[2];
    const float rr = r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
    const float lim = s.br * 1.01f + 0.1f;
    return ww - wr * (wr / rr) > lim * lim;
}

// intersectionEndPoints (src/analytic_ray_tracing_gpu.cu:465-518); a culled solid is a miss, as the exact
// test finds it
__device__ bool intersect(const float* p, const float* r, F2& ts, const Solid& s)
{
#pragma clang fp contract(off)
    if (culled(p, r, s)) {
        ts.x = -OOB;
        ts.y = ts.x;
        return false;
    }
    D3 q, mr;
    if (!s.rotated) {
        q.x = (p[0] - s.c[0]) / s.r[0];
        q.y = (p[1] - s.c[1]) / s.r[1];
        q.z = (p[2] - s.c[2]) / s.r[2];
        mr.x = r[0] / s.r[0];
        mr.y = r[1] / s.r[1];
        mr.z = r[2] / s.r[2];
    } else {
        const double qx = p[0] - s.c[0], qy = p[1] - s.c[1], qz = p[2] - s.c[2];
        q.x = (qx * s.A[0] + qy * s.A[1] + qz * s.A[2]) / s.r[0];
        q.y = (qx * s.A[3] + qy * s.A[4] + qz * s.A[5]) / s.r[1];
        q.z = (qx * s.A[6] + qy * s.A[7] + qz * s.A[8]) / s.r[2];
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

struct Sums {
    float fwd, bwd, all;
};

__device__ __forceinline__ void add_piece(Sums& acc, float val, float lo, float hi, float mid)
{
#pragma clang fp contract(off)
    const float piece = val * (hi - lo);
    acc.all += piece;
    if (mid >= 0.0f)
        acc.fwd += piece;
    else
        acc.bwd += piece;
}

// storage-free form: the distinct end points in increasing order, each piece owned by the last-listed
// solid containing its mid point
__device__ Sums pieces_streaming(const float* p, const float* r, const Solid* __restrict__ S, int n)
{
#pragma clang fp contract(off)
    Sums acc{0.0f, 0.0f, 0.0f};
    float prev = 0.0f;
    bool have_prev = false;
    for (;;) {
        float nxt = 0.0f;
        bool found = false;
        for (int j = 0; j < n; j++) {
            F2 ts;
            if (!intersect(p, r, ts, S[j]))
                continue;
            if ((!have_prev || ts.x > prev) && (!found || ts.x < nxt)) {
                nxt = ts.x;
                found = true;
            }
            if ((!have_prev || ts.y > prev) && (!found || ts.y < nxt)) {
                nxt = ts.y;
                found = true;
            }
        }
        if (!found)
            break;
        if (have_prev) {
            const float mid = (nxt + prev) * 0.5f;
            for (int j = n - 1; j >= 0; j--) {
                F2 ts;
                if (intersect(p, r, ts, S[j]) && ts.x <= mid && mid <= ts.y) {
                    add_piece(acc, S[j].val, prev, nxt, mid);
                    break;
                }
            }
        }
        prev = nxt;
        have_prev = true;
    }
    return acc;
}

// stoppingPower_geometricSolids / lineIntegral_geometricSolids (src/analytic_ray_tracing_gpu.cu:536-691)
template <bool STOPPING>
__device__ float ray_value(const float* p, const float* r, const Solid* __restrict__ S, int n)
{
#pragma clang fp contract(off)
    float t0[MAXHIT], t1[MAXHIT];
    int id[MAXHIT];
    int cnt = 0;
    for (int i = 0; i < n; i++) {
        F2 ts;
        if (intersect(p, r, ts, S[i])) {
            if (cnt < MAXHIT) {
                t0[cnt] = ts.x;
                t1[cnt] = ts.y;
                id[cnt] = i;
            }
            cnt++;
        }
    }
    Sums acc{0.0f, 0.0f, 0.0f};
    if (cnt > MAXHIT) {
        acc = pieces_streaming(p, r, S, n);
    } else if (cnt > 0) {
        float e[2 * MAXHIT];
        for (int k = 0; k < cnt; k++) {
            e[2 * k] = t0[k];
            e[2 * k + 1] = t1[k];
        }
        for (int a = 1; a < 2 * cnt; a++) {  // insertion sort
            const float v = e[a];
            int b = a - 1;
            while (b >= 0 && e[b] > v) {
                e[b + 1] = e[b];
                b--;
            }
            e[b + 1] = v;
        }
        for (int i = 0; i < 2 * cnt - 1; i++) {
            const float mid = (e[i + 1] + e[i]) * 0.5f;
            for (int ind = cnt - 1; ind >= 0; ind--)
                if (t0[ind] <= mid && mid <= t1[ind]) {
                    add_piece(acc, S[id[ind]].val, e[i], e[i + 1], mid);
                    break;
                }
        }
    }
    if (!STOPPING)
        return acc.all;
    return (1.0f - expf(-acc.fwd)) * (1.0f - expf(-acc.bwd));
}

struct RtArgs {
    int Nv1, Nv0, Nu1, Nu0;
    float Tv1, Tv0, Tu1, Tu0, v1_0, v0_0, u1_0, u0_0;
    float cs, sn, R, L, H;
    int os;
};

// rayTracingKernel / stoppingPowerKernel (src/analytic_ray_tracing_gpu.cu:693-862), all views at once
template <bool STOPPING>
__global__ void __launch_bounds__(RT_THREADS) ray_trace_kernel(float* __restrict__ g, const Solid* __restrict__ S,
                                                               int n, RtArgs A, int64_t total)
{
#pragma clang fp contract(off)
    const int64_t b = int64_t(blockIdx.x) * RT_THREADS + threadIdx.x;
    if (b >= total)
        return;
    const int iu0 = int(b % A.Nu0);
    int64_t rest = b / A.Nu0;
    const int iu1 = int(rest % A.Nu1);
    rest /= A.Nu1;
    const int iv0 = int(rest % A.Nv0);
    const int iv1 = int(rest / A.Nv0);
    const float v1 = float(iv1) * A.Tv1 + A.v1_0;
    const float v0 = float(iv0) * A.Tv0 + A.v0_0;
    const float u1 = float(iu1) * A.Tu1 + A.u1_0;
    const float u0 = float(iu0) * A.Tu0 + A.u0_0;
    // the u0 half width is T_v0 in the reference's test (it reads T.x)
    if (fabsf(u1) - 0.5f * A.Tu1 > A.H - A.R * fabsf(v1) || fabsf(u0) - 0.5f * A.Tv0 > A.L + A.R * fabsf(v0)) {
        g[b] = 0.0f;
        return;
    }
    // trajectory divided by rsqrt(r.r): r scaled by |r| (the reference's setTrajectory)
    float r[3] = {-(v0 * A.cs + A.sn), -(v0 * A.sn - A.cs), -v1};
    const float mag = 1.0f / sqrtf(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
    r[0] = r[0] / mag;
    r[1] = r[1] / mag;
    r[2] = r[2] / mag;
    if (A.os <= 1) {
        const float p[3] = {u0 * A.cs, u0 * A.sn, u1};
        g[b] = ray_value<STOPPING>(p, r, S, n);
        return;
    }
    const float T_v_os = A.Tu1 / float(A.os + 1);
    const float T_u_os = A.Tu0 / float(A.os + 1);
    const int rad = (A.os - 1) / 2;
    float acc = 0.0f;
    for (int j = -rad; j <= rad; j++) {
        const float du1 = j * T_v_os;
        for (int k = -rad; k <= rad; k++) {
            const float du0 = k * T_u_os;
            const float p[3] = {(u0 + du0) * A.cs, (u0 + du0) * A.sn, u1 + du1};
            acc += ray_value<STOPPING>(p, r, S, n);
        }
    }
    g[b] = acc / float(A.os * A.os);
}

}  // namespace

bool solid_init(int type, const float* c, const float* r, float val, const float* A, const float* clip, Solid& o)
{
    o = Solid{};
    o.type = -1;
    if (c == nullptr || r == nullptr || type < 0 || type > 7)
        return false;
    o.type = type;
    for (int i = 0; i < 3; i++) {
        o.c[i] = c[i];
        o.r[i] = r[i];
    }
    if (clip != nullptr)
        for (int i = 0; i < 3; i++)
            if (clip[i] != 0.0f) {
                o.clip[o.nclip][i] = -1.0f;
                o.clip[o.nclip][3] = -clip[i];
                o.nclip++;
            }
    o.val = val;
    // ellipsoids, boxes and cylinders lie within |r| of their centre whatever their rotation or clipping;
    // cones are always tested (their scaling leaves the axial clip range unscaled)
    for (int i = 0; i < 3; i++)
        o.bc[i] = c[i];
    o.br = (type <= 4) ? float(std::sqrt(double(r[0]) * r[0] + double(r[1]) * r[1] + double(r[2]) * r[2])) : -1.0f;
    if (A != nullptr)
        for (int i = 0; i < 9; i++) {
            o.A[i] = A[i];
            const double dev = (i == 0 || i == 4 || i == 8) ? std::fabs(double(A[i]) - 1.0) : std::fabs(double(A[i]));
            if (dev > 1.0e-8)
                o.rotated = 1;
        }
    // a cone with end radii r_1, r_2 over half length l becomes the unit double cone about its apex,
    // clipped along the axis (src/phantom.cpp:591-637)
    auto cone = [&](int ax, int i1, int i2, int il) {
        const double l = o.r[il], r_1 = o.r[i1], r_2 = o.r[i2];
        const int a1 = (ax + 1) % 3, a2 = (ax + 2) % 3;
        o.r[a1] = float(std::fabs(r_2 - r_1) / (2.0 * l));
        o.r[a2] = o.r[a1];
        o.r[ax] = 1.0f;
        o.c[ax] = float(o.c[ax] - l * (r_2 + r_1) / (r_2 - r_1));
        o.cone[0] = float(-l + l * (r_2 + r_1) / (r_2 - r_1));
        o.cone[1] = float(l + l * (r_2 + r_1) / (r_2 - r_1));
    };
    if (type == 5)
        cone(0, 1, 2, 0);
    else if (type == 6)
        cone(1, 0, 2, 1);
    else if (type == 7)
        cone(2, 0, 1, 2);
    return true;
}

void solids_scale(std::vector<Solid>& s, float sx, float sy, float sz)
{
    for (Solid& o : s) {
        o.c[0] *= sx;
        o.c[1] *= sy;
        o.c[2] *= sz;
        o.r[0] *= sx;
        o.r[1] *= sy;
        o.r[2] *= sz;
        for (int k = 0; k < o.nclip; k++) {  // the reference scales these three entries so
            o.clip[k][1] *= sy;
            o.clip[k][2] *= sz;
            o.clip[k][3] *= sx;
        }
        o.bc[0] *= sx;
        o.bc[1] *= sy;
        o.bc[2] *= sz;
        if (o.br >= 0.0f)
            o.br *= std::max(std::fabs(sx), std::max(std::fabs(sy), std::fabs(sz)));
    }
}

void ray_trace_device(const PlanGeom& P, const Solid* d_solids, int n, int oversampling, bool stopping, float* g,
                      hipStream_t s)
{
    oversampling = std::max(1, std::min(oversampling, 11));
    if (oversampling % 2 == 0)
        oversampling += 1;
    oversampling = std::max(1, std::min(oversampling, 11));
    RtArgs A;
    A.Nv1 = P.Nv1;
    A.Nv0 = P.Nv0;
    A.Nu1 = P.Nu1;
    A.Nu0 = P.Nu0;
    A.Tv1 = P.Tv1;
    A.Tv0 = P.Tv0;
    A.Tu1 = P.Tu1;
    A.Tu0 = P.Tu0;
    A.v1_0 = P.v1_0;
    A.v0_0 = P.v0_0;
    A.u1_0 = P.u1_0;
    A.u0_0 = P.u0_0;
    // setConstantMemoryGeometryParameters: cos / sin of the stored float psi
    A.cs = float(std::cos(double(P.psi)));
    A.sn = float(std::sin(double(P.psi)));
    A.R = P.R;
    A.L = P.L;
    A.H = P.H;
    A.os = oversampling;
    const int64_t total = plan_elems(P);
    const int64_t blocks = (total + RT_THREADS - 1) / RT_THREADS;
    if (blocks > int64_t(0x7fffffff))
        throw hip_error("ray tracing: planogram too large for one launch");
    if (stopping)
        hipLaunchKernelGGL(ray_trace_kernel<true>, dim3(unsigned(blocks)), dim3(RT_THREADS), 0, s, g, d_solids, n, A,
                           total);
    else
        hipLaunchKernelGGL(ray_trace_kernel<false>, dim3(unsigned(blocks)), dim3(RT_THREADS), 0, s, g, d_solids, n, A,
                           total);
    PETL_LAUNCH_CHECK();
}

}  // namespace petl
