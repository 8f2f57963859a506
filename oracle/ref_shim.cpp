// TEST INFRASTRUCTURE ONLY.  Thin extern "C" shim over the reference's own geometry
// (src/planogram.cpp) and CPU elementwise (src/data_cube.cpp) classes, compiled from the reference
// sources where they lie.  It lets the tests pin the oracle's geometry and zero-division conventions
// against the reference itself.  The degree->radian conversion and half-size scaling restate the
// one line of src/petl_c_interface.cpp:218, which cannot be compiled here (it needs the CUDA build).
#include <vector>
#include "petl_defines.h"
#include "planogram.h"
#include "data_cube.h"
#include "vector_ops.h"

extern "C" {

// out = {N_v1, N_v0, N_u1, N_u0}; outf = {psi, R, L, H, v_m0, v_m1, T_u0, T_u1, T_v0, T_v1,
//        u0_0, u1_0, v0_0, v1_0, cos_psi, sin_psi}
int ref_plan_init(float psi, float D, float L, float H, float v_m0, float v_m1, float T, int* out, float* outf)
{
    planogram p;
    if (!p.init(psi * PI / 180.0, 0.5 * D, 0.5 * L, 0.5 * H, v_m0, v_m1, T))
        return 0;
    out[0] = p.N_v1; out[1] = p.N_v0; out[2] = p.N_u1; out[3] = p.N_u0;
    const float vals[16] = {p.psi, p.R, p.L, p.H, p.v_m0, p.v_m1, p.T_u0, p.T_u1, p.T_v0, p.T_v1,
                            p.u0_0, p.u1_0, p.v0_0, p.v1_0, p.cos_psi, p.sin_psi};
    for (int i = 0; i < 16; i++) outf[i] = vals[i];
    p.data = nullptr;
    return 1;
}

// The reference's own event geometry (src/planogram.cpp:113-141) for one coincidence: returns 1 when
// accepted, with out = {v1, v0, u1, u0} (u from the blue end point, no origin shift).
int ref_event_coords(float psi, float D, float L, float H, float v_m0, float v_m1, float T, const float* blue,
                     const float* pink, float* out)
{
    planogram p;
    if (!p.init(psi * PI / 180.0, 0.5 * D, 0.5 * L, 0.5 * H, v_m0, v_m1, T))
        return -1;
    p.data = nullptr;
    float r[3] = {blue[0] - pink[0], blue[1] - pink[1], blue[2] - pink[2]};
    float pos[3] = {blue[0], blue[1], blue[2]};
    float v1 = 0, v0 = 0, u1 = 0, u0 = 0;
    if (!p.is_within_acceptable_angle(r, v1, v0))
        return 0;
    p.get_u_coords(pos, v1, v0, u1, u0);
    out[0] = v1; out[1] = v0; out[2] = u1; out[3] = u0;
    return 1;
}

// which: 0 multiply, 1 divide, 2 rdivide (y = x / y semantics of rdivide3D), 3 add
int ref_binary3d(float* x, float* y, int N1, int N2, int N3, int which)
{
    dataCube X(x, N1, N2, N3), Y(y, N1, N2, N3);
    dataCube* r = nullptr;
    if (which == 0) r = X.multiply(&Y);
    else if (which == 1) r = X.divide(&Y);
    else if (which == 2) r = Y.rdivide(&X);
    else r = X.add(&Y);
    X.data = nullptr; Y.data = nullptr;
    return r != nullptr;
}

int ref_reciprocal3d(float* x, int N1, int N2, int N3, float dz)
{
    dataCube X(x, N1, N2, N3);
    X.reciprocal(dz);
    X.data = nullptr;
    return 1;
}

float ref_inner3d(float* x, float* y, float* w, int N1, int N2, int N3)
{
    dataCube X(x, N1, N2, N3), Y(y, N1, N2, N3);
    float r;
    if (w) { dataCube W(w, N1, N2, N3); r = X.innerProduct(&Y, &W); W.data = nullptr; }
    else r = X.innerProduct(&Y);
    X.data = nullptr; Y.data = nullptr;
    return r;
}

// The reference's own solid-angle factor (src/planogram.cpp:262-265) applied by
// planogram::apply_solid_angle_correction (:272-301) to a planogram of ones: out[iv1][iv0] = c or 1/c.
int ref_solid_angle(float psi, float D, float L, float H, float v_m0, float v_m1, float T, int do_inverse,
                    float* out)
{
    planogram p;
    if (!p.init(psi * PI / 180.0, 0.5 * D, 0.5 * L, 0.5 * H, v_m0, v_m1, T))
        return 0;
    const uint64 n = uint64(p.N_v1) * p.N_v0 * p.N_u1 * p.N_u0;
    std::vector<float> g(n, 1.0f);
    p.data = g.data();
    p.apply_solid_angle_correction(do_inverse != 0);
    for (int i = 0; i < p.N_v1; i++)
        for (int j = 0; j < p.N_v0; j++)
            out[i * p.N_v0 + j] = g[(uint64(i) * p.N_v0 + j) * p.N_u1 * p.N_u0];
    p.data = nullptr;
    return 1;
}

// The reference's own FFT size rule (src/vector_ops.cpp:481-505).
int ref_optimal_fft_size(int N) { return optimalFFTsize(N); }

}
