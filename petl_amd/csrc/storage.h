// Storage formats of volumes and planograms in HBM.  Arithmetic is always fp32; a 16-bit element is
// widened on load and rounded (to nearest, ties to even) on store.
//
//   STORE_F32   float
//   STORE_BF16  bfloat16 (8-bit exponent, 7-bit mantissa): the upper half of the fp32 bit pattern, no scale
//   STORE_F16   IEEE binary16 with a per-tensor power-of-two scale: stored = x / scale, x = stored * scale,
//               so a tensor's magnitude can be placed inside binary16's range (6.1e-5 .. 65504)
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstring>
#include "petl_common.h"

namespace petl {

template <int ST>
struct Stor;

template <>
struct Stor<STORE_F32> {
    using T = float;
    __host__ __device__ static float ld(T v, float) { return v; }
    __host__ __device__ static T st(float x, float) { return x; }
};

template <>
struct Stor<STORE_BF16> {
    using T = uint16_t;
    __host__ __device__ static float ld(T v, float)
    {
        const uint32_t u = uint32_t(v) << 16;
        float x;
        __builtin_memcpy(&x, &u, 4);
        return x;
    }
    __host__ __device__ static T st(float x, float)
    {
        uint32_t u;
        __builtin_memcpy(&u, &x, 4);
        if ((u & 0x7f800000u) == 0x7f800000u)  // inf / nan: keep, quieting a nan
            return T((u >> 16) | ((u & 0x007fffffu) ? 0x40u : 0u));
        u += 0x7fffu + ((u >> 16) & 1u);       // round to nearest, ties to even
        return T(u >> 16);
    }
};

template <>
struct Stor<STORE_F16> {
    using T = _Float16;
    __host__ __device__ static float ld(T v, float scale) { return float(v) * scale; }
    // saturating: a value beyond binary16's range is stored as +-65504, not as an infinity
    __host__ __device__ static T st(float x, float inv_scale)
    {
        return T(fminf(fmaxf(x * inv_scale, -65504.0f), 65504.0f));
    }
};

inline size_t storage_bytes(int st) { return st == STORE_F32 ? 4 : 2; }

// host conversions (uploads of data, downloads of images)
void to_storage(int st, const float* x, void* out, int64_t n, float inv_scale);
void from_storage(int st, const void* x, float* out, int64_t n, float scale);
// the power of two s with limit / 2 < max_abs / s <= limit (1 when max_abs is 0 or not finite)
float pow2_scale_for(float max_abs, float limit);

}  // namespace petl
