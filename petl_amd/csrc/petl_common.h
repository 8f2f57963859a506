// Common definitions for the MI355X PETL projector library: geometry records passed to kernels,
// HIP error handling, and small device helpers.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstdio>
#include <string>
#include <stdexcept>

namespace petl {

// One planogram's sampling, exactly the float fields the reference keeps on its planogram class
// (src/planogram.h:41-66) and packs into int4/float4 for its kernels (src/cuda_utils.cu:1363-1387).
struct PlanGeom {
    int Nv1, Nv0, Nu1, Nu0;
    float Tv1, Tv0, Tu1, Tu0;
    float v1_0, v0_0, u1_0, u0_0;
    float cos_psi, sin_psi;
    float R, L, H;   // half panel separation / half length / half height
    float psi;       // radians (float, as stored by the reference)
};

// Volume grid (src/cuda_utils.cu:1390-1404): f[z][y][x], x fastest, voxel-0 centre (x0,y0,z0).
struct VolGeom {
    int Nx, Ny, Nz;
    float Tx, Ty, Tz;
    float x0, y0, z0;
};

__host__ __device__ inline int64_t plan_elems(const PlanGeom& p) { return int64_t(p.Nv1) * p.Nv0 * p.Nu1 * p.Nu0; }
__host__ __device__ inline int64_t vol_elems(const VolGeom& v) { return int64_t(v.Nx) * v.Ny * v.Nz; }

// Backprojector semantics.
enum BackMode : int {
    BACK_PETL = 0,     // reference semantics: voxel-centre u1 test, no u0 mask (src/projectors.cu:240-377)
    BACK_MATCHED = 1,  // exact transpose of the forward projector, masks included
};

// Storage type of volumes/planograms on the device (arithmetic is always fp32 accumulate).
enum StorageType : int { STORE_F32 = 0, STORE_BF16 = 1, STORE_F16 = 2 };

struct hip_error : std::runtime_error {
    using std::runtime_error::runtime_error;
};

#define PETL_HIP_CHECK(expr)                                                                          \
    do {                                                                                              \
        hipError_t _e = (expr);                                                                       \
        if (_e != hipSuccess)                                                                         \
            throw ::petl::hip_error(std::string(#expr) + " failed: " + hipGetErrorString(_e) + " at " + \
                                    __FILE__ + ":" + std::to_string(__LINE__));                        \
    } while (0)

#define PETL_LAUNCH_CHECK() PETL_HIP_CHECK(hipGetLastError())

inline int ceil_div(int64_t a, int64_t b) { return int((a + b - 1) / b); }

constexpr double PETL_PI = 3.1415926535897932385;  // src/petl_defines.h:22

}  // namespace petl
