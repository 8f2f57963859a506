// List-mode binning on the GPU: coincidences (blue and pink end points) -> planogram counts.
//
// Reference: src/binning.cpp:558-660 (reading, acceptance, u coordinates, chest-wall origin) and
// :662-750 (the flush: nearest v bins by int(x + 0.5), nearest or clamped bilinear u bins, weights formed
// in double and added to the float bins one at a time), with the event geometry of
// src/planogram.cpp:113-141.
//
// The reference adds every contribution to its bin in event order.  Float addition is not associative,
// so to give the same bits this build keeps that order instead of using atomics: each coincidence
// writes its (bin, weight) pairs at position 4 * event + corner (bilinear; event for nearest), a stable
// radix sort groups them by bin keeping that order, and one thread per bin adds its run to the bin's value
// in sequence.
#include <hipcub/hipcub.hpp>
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>
#include "binning.h"

namespace petl {

namespace {

constexpr int BT = 256;

__device__ __forceinline__ bool event_coords(const BinGeom& B, const float* b, const float* k, float& v1, float& v0,
                                             float& u1, float& u0)
{
#pragma clang fp contract(off)
    const PlanGeom& P = B.P;
    const float r0 = b[0] - k[0], r1 = b[1] - k[1], r2 = b[2] - k[2];
    const float c = -P.sin_psi * r0 + P.cos_psi * r1;
    if (c == 0.0f)
        return false;
    v0 = -(P.cos_psi * r0 + P.sin_psi * r1) / c;
    if (fabsf(v0) > B.v_m0)
        return false;
    v1 = -r2 / c;
    if (fabsf(v1) > B.v_m1)
        return false;
    const float t = b[1] * P.cos_psi - b[0] * P.sin_psi;
    u1 = b[2] + v1 * t;
    u0 = b[0] * P.cos_psi + b[1] * P.sin_psi + v0 * t;
    return true;
}

// clamped linear position of src/binning.cpp:704-742
__device__ __forceinline__ void split(float x, int n, int& lo, int& hi, float& d)
{
    if (x <= 0.0f) {
        lo = hi = 0;
        d = 0.0f;
    } else if (x >= float(n - 1)) {
        lo = hi = n - 1;
        d = 0.0f;
    } else {
        lo = int(x);
        hi = lo + 1;
        d = x - float(lo);
    }
}

// thread per coincidence; SLOTS = 4 (bilinear: one slot per corner) or 1 (nearest)
template <int SLOTS>
__global__ void __launch_bounds__(BT) bin_contrib_kernel(const float* __restrict__ blue,
                                                         const float* __restrict__ pink, int64_t n, BinGeom B,
                                                         uint32_t no_bin, uint32_t* __restrict__ keys,
                                                         double* __restrict__ vals,
                                                         unsigned int* __restrict__ block_accepted)
{
#pragma clang fp contract(off)
    const int64_t e = int64_t(blockIdx.x) * BT + threadIdx.x;
    uint32_t key[4] = {no_bin, no_bin, no_bin, no_bin};
    double w[4] = {0.0, 0.0, 0.0, 0.0};
    bool ok = false;
    if (e < n) {
        const PlanGeom& P = B.P;
        float v1, v0, u1, u0;
        ok = event_coords(B, blue + 3 * e, pink + 3 * e, v1, v0, u1, u0);
        if (ok) {
            if (B.chest_wall)
                u1 = u1 - P.H;
            const int iv1 = int(double((v1 - P.v1_0) / P.Tv1) + 0.5);
            const int iv0 = int(double((v0 - P.v0_0) / P.Tv0) + 0.5);
            if (iv1 >= 0 && iv1 < P.Nv1 && iv0 >= 0 && iv0 < P.Nv0) {
                const uint32_t base = (uint32_t(iv1) * uint32_t(P.Nv0) + uint32_t(iv0)) * uint32_t(P.Nu1 * P.Nu0);
                if (SLOTS == 1) {
                    const int a = int(double((u1 - P.u1_0) / P.Tu1) + 0.5);
                    const int q = int(double((u0 - P.u0_0) / P.Tu0) + 0.5);
                    if (a >= 0 && a < P.Nu1 && q >= 0 && q < P.Nu0) {
                        key[0] = base + uint32_t(a * P.Nu0 + q);
                        w[0] = 1.0;
                    }
                } else {
                    int lo1, hi1, lo0, hi0;
                    float du1, du0;
                    split((u1 - P.u1_0) / P.Tu1, P.Nu1, lo1, hi1, du1);
                    split((u0 - P.u0_0) / P.Tu0, P.Nu0, lo0, hi0, du0);
                    const double a1 = 1.0 - double(du1), b1 = double(du1);
                    const double a0 = 1.0 - double(du0), b0 = double(du0);
                    key[0] = base + uint32_t(lo1 * P.Nu0 + lo0);
                    key[1] = base + uint32_t(hi1 * P.Nu0 + lo0);
                    key[2] = base + uint32_t(lo1 * P.Nu0 + hi0);
                    key[3] = base + uint32_t(hi1 * P.Nu0 + hi0);
                    w[0] = a1 * a0 * 1.0;
                    w[1] = b1 * a0 * 1.0;
                    w[2] = a1 * b0 * 1.0;
                    w[3] = b1 * b0 * 1.0;
                }
            }
        }
        if (SLOTS == 1) {
            keys[e] = key[0];
            vals[e] = w[0];
        } else {
            // 16-byte and two 16-byte stores per coincidence (slots 4e .. 4e+3 are contiguous and aligned)
            reinterpret_cast<uint4*>(keys)[e] = make_uint4(key[0], key[1], key[2], key[3]);
            reinterpret_cast<double2*>(vals)[2 * e] = make_double2(w[0], w[1]);
            reinterpret_cast<double2*>(vals)[2 * e + 1] = make_double2(w[2], w[3]);
        }
    }
    // accepted coincidences (src/binning.cpp:643), one count per block
    const int cnt = __syncthreads_count(ok);
    if (threadIdx.x == 0)
        block_accepted[blockIdx.x] = unsigned(cnt);
}

// adds the per-block counts to *accepted in a fixed order (one block)
__global__ void __launch_bounds__(BT) bin_count_kernel(const unsigned int* __restrict__ block_accepted, int nblocks,
                                                       unsigned long long* __restrict__ accepted)
{
    __shared__ unsigned long long red[BT];
    unsigned long long acc = 0;
    for (int i = threadIdx.x; i < nblocks; i += BT)
        acc += block_accepted[i];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int w = BT / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w)
            red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        *accepted += red[0];
}

// one thread per run of equal bins: add the run to the bin in order, each sum rounded to float
__global__ void __launch_bounds__(BT) bin_accumulate_kernel(const uint32_t* __restrict__ keys,
                                                            const double* __restrict__ vals, int64_t m,
                                                            uint32_t no_bin, float* __restrict__ g)
{
    const int64_t i = int64_t(blockIdx.x) * BT + threadIdx.x;
    if (i >= m)
        return;
    const uint32_t k = keys[i];
    if (k == no_bin || (i > 0 && keys[i - 1] == k))
        return;
    float acc = g[k];
    for (int64_t j = i; j < m && keys[j] == k; j++)
        acc = float(double(acc) + vals[j]);
    g[k] = acc;
}

int key_bits(uint64_t max_key)
{
    int b = 1;
    while (b < 32 && (uint64_t(1) << b) <= max_key)
        b++;
    return b;
}

}  // namespace

BinGeom bin_geom(const Planogram& pl, bool nearest, bool chest_wall)
{
    BinGeom B;
    B.P = pl.geom();
    B.v_m0 = pl.v_m0;
    B.v_m1 = pl.v_m1;
    B.nearest = nearest ? 1 : 0;
    B.chest_wall = chest_wall ? 1 : 0;
    return B;
}

void bin_events_device(const BinGeom& B, const float* blue, const float* pink, int64_t n, float* g, BinScratch& W,
                       unsigned long long* accepted, hipStream_t s)
{
    if (n <= 0)
        return;
    const uint64_t G = uint64_t(B.P.Nv1) * B.P.Nv0 * B.P.Nu1 * B.P.Nu0;
    if (G >= uint64_t(0xffffffffu))
        throw hip_error("binning: planogram too large for 32-bit bin keys");
    const uint32_t no_bin = uint32_t(G);  // sorts after every real bin
    const int slots = B.nearest ? 1 : 4;
    const int64_t m = slots * n;
    const int nblocks = ceil_div(n, BT);
    W.counts.ensure(size_t(nblocks) * sizeof(unsigned int));
    W.keys[0].ensure(size_t(m) * sizeof(uint32_t));
    W.keys[1].ensure(size_t(m) * sizeof(uint32_t));
    W.vals[0].ensure(size_t(m) * sizeof(double));
    W.vals[1].ensure(size_t(m) * sizeof(double));
    uint32_t* k0 = static_cast<uint32_t*>(W.keys[0].p);
    uint32_t* k1 = static_cast<uint32_t*>(W.keys[1].p);
    double* v0 = static_cast<double*>(W.vals[0].p);
    double* v1 = static_cast<double*>(W.vals[1].p);
    unsigned int* counts = static_cast<unsigned int*>(W.counts.p);
    if (slots == 1)
        hipLaunchKernelGGL(bin_contrib_kernel<1>, dim3(nblocks), dim3(BT), 0, s, blue, pink, n, B, no_bin, k0, v0,
                           counts);
    else
        hipLaunchKernelGGL(bin_contrib_kernel<4>, dim3(nblocks), dim3(BT), 0, s, blue, pink, n, B, no_bin, k0, v0,
                           counts);
    PETL_LAUNCH_CHECK();
    hipLaunchKernelGGL(bin_count_kernel, dim3(1), dim3(BT), 0, s, counts, nblocks, accepted);
    PETL_LAUNCH_CHECK();
    const int bits = key_bits(G);  // keys are <= G
    size_t tmp = 0;
    PETL_HIP_CHECK(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp, k0, k1, v0, v1, m, 0, bits, s));
    W.tmp.ensure(tmp);
    PETL_HIP_CHECK(hipcub::DeviceRadixSort::SortPairs(W.tmp.p, tmp, k0, k1, v0, v1, m, 0, bits, s));
    hipLaunchKernelGGL(bin_accumulate_kernel, dim3(ceil_div(m, BT)), dim3(BT), 0, s, k1, v1, m, no_bin, g);
    PETL_LAUNCH_CHECK();
}

// ---------------------------------------------------------------------------------------------
// list-mode files (src/binning.cpp:67-150 .lst, :558-660 acquisition layout), read in chunks
// ---------------------------------------------------------------------------------------------
ListModeReader::ListModeReader(const char* path) : path_(path ? path : "")
{
    fh_ = path ? std::fopen(path, "rb") : nullptr;
    const size_t L = path_.size();
    lst_ = L >= 4 && path_.compare(L - 4, 4, ".lst") == 0;
    if (fh_ && !lst_) {
        int64_t n = 0;
        if (std::fread(&n, sizeof(int64_t), 1, fh_) != 1)
            n = 0;
        remaining_ = n;
    }
}

ListModeReader::~ListModeReader()
{
    if (fh_)
        std::fclose(fh_);
}

int64_t ListModeReader::next(float* blue, float* pink, int64_t max_events)
{
    if (!fh_)
        return 0;
    int64_t got = 0;
    if (lst_) {
        // {float pos[4]; uint64 time} x 2 = 48 bytes per coincidence
        std::vector<unsigned char> buf(size_t(std::min<int64_t>(max_events, 1 << 16)) * 48);
        while (got < max_events) {
            const size_t want = size_t(std::min<int64_t>(max_events - got, int64_t(buf.size() / 48)));
            const size_t n = std::fread(buf.data(), 48, want, fh_);
            for (size_t i = 0; i < n; i++) {
                std::memcpy(blue + 3 * (got + int64_t(i)), buf.data() + 48 * i, 12);
                std::memcpy(pink + 3 * (got + int64_t(i)), buf.data() + 48 * i + 24, 12);
            }
            got += int64_t(n);
            if (n < want)
                break;
        }
        return got;
    }
    while (got < max_events && remaining_ > 0) {
        float* dst[2] = {blue + 3 * got, pink + 3 * got};
        bool ok = true;
        for (int side = 0; side < 2 && ok; side++) {
            int16_t nint = 0;
            float pos_e[4];
            ok = std::fread(&nint, sizeof(int16_t), 1, fh_) == 1 && std::fread(pos_e, sizeof(float), 4, fh_) == 4;
            if (ok && nint > 1)
                ok = std::fseek(fh_, long(18) * (nint - 1), SEEK_CUR) == 0;  // further interactions
            if (ok)
                std::memcpy(dst[side], pos_e, 12);
        }
        if (!ok) {
            remaining_ = 0;
            break;
        }
        remaining_--;
        got++;
    }
    return got;
}

}  // namespace petl
